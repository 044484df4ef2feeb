// TILD engine: set-up, channel layout, the step loop and the C entry points
// of the engine (mft_tild_create .. mft_tild_autotune).  Kernels live in
// tild_particles.hip, tild_grid.hip and tild_conv.hip.
#include "tild_impl.h"

namespace mft {

static int check_fft_size(int n, const char* what)
{
    MFT_REQUIRE(mft_fft_supported(n), MFT_ERR_UNSUPPORTED,
                "grid size %d along %s is not a supported FFT length (2,3,4,5,6,8,10,12,16,20,24,32,40,48,"
                "64,80,96,128,160,192,256)",
                n, what);
    return MFT_OK;
}

// host-side erf shape tables in FP64 (the reference's FP32 table cancels
// catastrophically at small k; this is the exact function, rounded once)
void erf_table(int kind, double sigma, double Rp, int nx, int ny, int nz, float lx, float ly, float lz,
               std::vector<float>& out)
{
    out.resize((size_t)nx * ny * nz);
    const double PI4 = 12.56637061435917295384;
    double sig2 = sigma * sigma;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                float kx = kcomp(ix, nx, lx), ky = kcomp(iy, ny, ly), kz = kcomp(iz, nz, lz);
                double k2 = (double)(float)((float)(kx * kx + ky * ky) + kz * kz);
                double k = std::sqrt(k2);
                double vol = PI4 * Rp * Rp * Rp / 3.0;
                double shape;
                if (k2 == 0.0) shape = vol;
                else {
                    double xx = Rp * k;
                    if (xx < 1e-2) {
                        double x2 = xx * xx;
                        shape = vol * (1.0 - x2 / 10.0 + x2 * x2 / 280.0);
                    } else
                        shape = PI4 * (std::sin(xx) - xx * std::cos(xx)) / (k2 * k);
                }
                double u = k2 == 0.0 ? (kind == MFT_POT_ERF2 ? vol * vol : vol)
                                     : std::exp(-k2 * sig2 * 0.5) * (kind == MFT_POT_ERF2 ? shape * shape : shape);
                out[ix + (size_t)nx * (iy + (size_t)ny * iz)] = (float)u;
            }
}

static int tild_setup(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    MFT_REQUIRE(d.nparticles > 0, MFT_ERR_INVALID, "Box created with no particles");
    MFT_REQUIRE(d.nparticles < (1 << kMetaIdBits), MFT_ERR_UNSUPPORTED, "at most %d particles per engine",
                (1 << kMetaIdBits) - 1);
    MFT_REQUIRE(d.nspecies >= 1 && d.nspecies <= kMaxSpecies, MFT_ERR_INVALID, "nspecies must be 1..%d",
                kMaxSpecies);
    MFT_REQUIRE(d.pmeorder >= 0 && d.pmeorder <= 4, MFT_ERR_INVALID, "pmeorder must be 0..4");
    MFT_REQUIRE(d.npotentials >= 0 && d.npotentials <= kMaxPots, MFT_ERR_INVALID, "too many potentials");
    MFT_REQUIRE(d.nbondtypes >= 0 && d.nbondtypes <= kMaxBondTypes, MFT_ERR_INVALID, "too many bond types");
    MFT_REQUIRE(d.maxbonds >= 0 && d.maxbonds <= 16, MFT_ERR_INVALID, "maxbonds must be 0..16");
    MFT_REQUIRE(d.lx > 0 && d.ly > 0 && d.lz > 0, MFT_ERR_INVALID, "box lengths must be positive");
    int rc;
    if ((rc = check_fft_size(d.nx, "x"))) return rc;
    if ((rc = check_fft_size(d.ny, "y"))) return rc;
    if ((rc = check_fft_size(d.nz, "z"))) return rc;
    MFT_REQUIRE(d.nx >= d.pmeorder + 1 && d.ny >= d.pmeorder + 1 && d.nz >= d.pmeorder + 1, MFT_ERR_INVALID,
                "grid smaller than the spline stencil");
    e->N = d.nparticles;
    e->M = d.nx * d.ny * d.nz;
    // PS_Box_initialization.cu:378 (double) and :727-731 (float copy)
    e->dx[0] = (float)((double)d.lx / (double)d.nx);
    e->dx[1] = (float)((double)d.ly / (double)d.ny);
    e->dx[2] = (float)((double)d.lz / (double)d.nz);
    e->gvol = e->dx[0] * e->dx[1] * e->dx[2];

    // potentials: one charges potential at most; the channel layout is set
    // once the charges are known (configure_channels)
    e->charge_pot = -1;
    for (int i = 0; i < d.npotentials; ++i) {
        const mft_tild_potential& P = e->pots[i];
        if (P.kind == MFT_POT_CHARGES) {
            MFT_REQUIRE(e->charge_pot < 0, MFT_ERR_UNSUPPORTED, "only one charges potential is supported");
            e->charge_pot = i;
        } else if (P.kind == MFT_POT_WALL) {
            MFT_REQUIRE(P.dir >= 0 && P.dir < 3 && (P.wall_sign == 1 || P.wall_sign == -1) && P.maskI != 0,
                        MFT_ERR_INVALID, "wall needs a group, an axis 0..2 and dirSign +1 or -1");
            MFT_REQUIRE(e->walls.n < kMaxWalls, MFT_ERR_UNSUPPORTED, "at most %d walls", kMaxWalls);
            const int w = e->walls.n++;
            e->walls.mask[w] = P.maskI;
            e->walls.axis[w] = P.dir;
            e->walls.sign[w] = P.wall_sign;
            e->walls.pos[w] = P.wall_pos;
            e->walls.k[w] = P.A;
            e->wall_pot.push_back(i);
        } else if (P.kind == MFT_POT_DPD) {
            const int rcd = tild_dpd_setup(e, i);
            if (rcd) return rcd;
        } else if (P.kind == MFT_POT_LANGEVIN) {
            MFT_REQUIRE(P.maskI != 0 && P.sigma > 0.0f, MFT_ERR_INVALID,
                        "ps-potentialLangevin.cu: did not find appropriate delt for the Langevin noise");
            MFT_REQUIRE(e->lang.n < kMaxWalls, MFT_ERR_UNSUPPORTED, "at most %d Langevin potentials", kMaxWalls);
            const int l = e->lang.n++;
            e->lang.mask[l] = P.maskI;
            e->lang.drag[l] = P.A;
            // noise_mag = sqrtf(2.0 * drag / delt) (ps_potentialLangevin.cu:64)
            e->lang.nmag[l] = sqrtf((float)(2.0 * (double)P.A / (double)P.sigma));
        } else if (P.kind == MFT_POT_BIAS) {
            MFT_REQUIRE(P.phase >= MFT_BIAS_LAM && P.phase <= MFT_BIAS_GYR, MFT_ERR_UNSUPPORTED,
                        "PHASE NOT FOUND in Box.cu");
            MFT_REQUIRE(P.dir >= 0 && P.dir < 3 && P.n_periods >= 1 && P.maskI != 0, MFT_ERR_INVALID,
                        "bias field needs a group, dir 0..2 and n_periods >= 1");
        } else {
            MFT_REQUIRE(P.kind >= 0 && P.kind <= MFT_POT_ERFG, MFT_ERR_UNSUPPORTED, "potential kind %d", P.kind);
        }
    }
    if (e->charge_pot >= 0)
        MFT_REQUIRE(d.has_charges, MFT_ERR_INVALID, "charges potential without charged particles");
    const int charge_pot = e->charge_pot;

    KParams& kp = e->kp;
    kp.npots = d.npotentials;
    kp.charge_pot = charge_pot;
    kp.energy = 0;
    e->tables.assign(d.npotentials, nullptr);
    for (int i = 0; i < d.npotentials; ++i) {
        const mft_tild_potential& P = e->pots[i];
        KPot& K = kp.pot[i];
        if (P.kind == MFT_POT_BIAS || P.kind == MFT_POT_WALL || P.kind == MFT_POT_LANGEVIN || P.kind == MFT_POT_DPD) {
            // not a pair potential: a no-op entry in the k-space pass (zero
            // amplitude, no species); a bias field is added to the force
            // fields after it, a wall to the particle forces in the integrate pass
            std::memset(&K, 0, sizeof(K));
            K.kind = MFT_POT_GAUSSIAN;
            e->pots[i].ramp = 0;
            if (P.kind == MFT_POT_BIAS && (rc = tild_bias_setup(e, i))) return rc;
            continue;
        }
        K.kind = P.kind;
        K.maskI = P.maskI;
        K.maskJ = P.maskJ;
        K.A = P.A;
        K.eA = P.A;
        K.sig2 = P.sigma * P.sigma;
        K.Rp = P.Rp;
        K.table = nullptr;
        if (P.kind == MFT_POT_ERF2 || P.kind == MFT_POT_ERFG) {
            std::vector<float> tab;
            erf_table(P.kind, P.sigma, P.Rp, d.nx, d.ny, d.nz, d.lx, d.ly, d.lz, tab);
            if ((rc = dalloc(&e->tables[i], tab.size()))) return rc;
            MFT_HIP(hipMemcpy(e->tables[i], tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
            K.table = e->tables[i];
        }
    }

    // bonds and integrator coefficients (GJF::finishInitialization,
    // ps_integratorGJF.cu:56-82, in float like the reference)
    std::memset(&e->bp, 0, sizeof(e->bp));
    e->bp.ntypes = d.nbondtypes;
    for (int i = 0; i < d.nbondtypes; ++i) {
        e->bp.k[i] = d.bond_k[i];
        e->bp.r0[i] = d.bond_r0[i];
        e->bp.style[i] = d.bond_style[i];
    }
    MFT_REQUIRE(d.integrator == MFT_INT_GJF || d.integrator == MFT_INT_EM || d.integrator == MFT_INT_NONE ||
                    d.integrator == MFT_INT_VV,
                MFT_ERR_INVALID, "unknown integrator %d", d.integrator);
    e->is_vv = d.integrator == MFT_INT_VV;
    std::memset(&e->gp, 0, sizeof(e->gp));
    for (int s = 0; s < d.nspecies; ++s) {
        float m = e->mass[s], D = e->mob[s], dt = d.delt;
        float b = m / (m + dt / (2.0f * D));
        float a = (1.0f - dt / (2.0f * D * m)) / (1.0f + dt / (2.0f * D * m));
        e->gp.a[s] = a;
        e->gp.b[s] = b;
        e->gp.nmag[s] = sqrtf(2.0f * dt) / sqrtf(D);
        e->gp.cf[s] = b * (dt * dt) / m;
        e->gp.cn[s] = b * dt / (2.0f * m);
        e->gp.inv_m[s] = 1.0f / m;
    }

    const int N = e->N, M = e->M;
    const size_t mb = d.maxbonds > 0 ? d.maxbonds : 1;
    for (int b = 0; b < 2; ++b) {
        if ((rc = dalloc(&e->X[b], (size_t)N))) return rc;
        if ((rc = dalloc(&e->XO[b], (size_t)N))) return rc;
        if ((rc = dalloc(&e->NZ[b], (size_t)N))) return rc;
        if ((rc = dalloc(&e->META[b], (size_t)N))) return rc;
        if ((rc = dalloc(&e->BS[b], (size_t)N * mb))) return rc;
        if ((rc = dalloc(&e->BT[b], (size_t)N * mb))) return rc;
    }
    if ((rc = dalloc(&e->XN, (size_t)N))) return rc;
    if ((rc = dalloc(&e->NN, (size_t)N))) return rc;
    if ((rc = dalloc(&e->fgrid, (size_t)N))) return rc;
    if ((rc = dalloc(&e->ftot, (size_t)N))) return rc;
    if ((rc = dalloc(&e->key, (size_t)N))) return rc;
    if ((rc = dalloc(&e->rank, (size_t)N))) return rc;
    if ((rc = dalloc(&e->map, (size_t)N))) return rc;
    if ((rc = dalloc(&e->slot_of, (size_t)N))) return rc;
    if ((rc = dalloc(&e->type, (size_t)N))) return rc;
    if (d.has_charges && (rc = dalloc(&e->charge, (size_t)N))) return rc;
    if ((rc = dalloc(&e->nBonds, (size_t)N))) return rc;
    if ((rc = dalloc(&e->bondedTo, (size_t)N * mb))) return rc;
    if ((rc = dalloc(&e->bondType, (size_t)N * mb))) return rc;
    // angles (PS_Box_initialization.cu:83-106, MAXANGLES :229-231)
    std::memset(&e->ap, 0, sizeof(e->ap));
    if (d.maxangles > 0) {
        MFT_REQUIRE(d.maxangles <= kMaxAngles, MFT_ERR_UNSUPPORTED, "MAXANGLES must be 1..%d", kMaxAngles);
        MFT_REQUIRE(d.nangletypes >= 1 && d.nangletypes <= kMaxAngleTypes && d.angle_k && d.angle_theta &&
                        d.angle_style,
                    MFT_ERR_INVALID, "angles need 1..%d angle types", kMaxAngleTypes);
        e->ap.ntypes = d.nangletypes;
        e->ap.maxangles = d.maxangles;
        for (int t = 0; t < d.nangletypes; ++t) {
            MFT_REQUIRE(d.angle_style[t] == 0 || d.angle_style[t] == 1, MFT_ERR_INVALID, "Angle style not supported!");
            e->ap.k[t] = d.angle_k[t];
            e->ap.theta[t] = d.angle_theta[t];
            e->ap.style[t] = d.angle_style[t];
        }
        const size_t ma = d.maxangles;
        for (int b = 0; b < 2; ++b) {
            if ((rc = dalloc(&e->NA[b], (size_t)N))) return rc;
            if ((rc = dalloc(&e->AT[b], (size_t)N * ma))) return rc;
            if ((rc = dalloc(&e->AG[b], (size_t)N * ma * 3))) return rc;
        }
        if ((rc = dalloc(&e->nAngles, (size_t)N))) return rc;
        if ((rc = dalloc(&e->angleType, (size_t)N * ma))) return rc;
        if ((rc = dalloc(&e->angleGroup, (size_t)N * ma * 3))) return rc;
        MFT_HIP(hipMemset(e->nAngles, 0, sizeof(int) * N));
        MFT_HIP(hipMemset(e->angleType, 0, sizeof(int) * N * ma));
        MFT_HIP(hipMemset(e->angleGroup, 0, sizeof(int) * N * ma * 3));
    }
    if ((rc = dalloc(&e->pio, 3 * (size_t)N))) return rc;
    if ((rc = dalloc(&e->pio2, 3 * (size_t)N))) return rc;
    if ((rc = dalloc(&e->pio3, 3 * (size_t)N))) return rc;
    if ((rc = dalloc(&e->force_slot, (size_t)d.nspecies))) return rc;
    const int nconv_blocks = tild_conv_blocks(d.nx, d.ny, d.nz);
    if ((rc = dalloc(&e->epart, (size_t)nconv_blocks * kMaxPots))) return rc;
    if ((rc = dalloc(&e->eout, (size_t)kMaxPots + 1))) return rc;
    if ((rc = dalloc(&e->d_step, (size_t)1))) return rc;
    if ((rc = dalloc(&e->bpart, (size_t)nblk(N)))) return rc;
    MFT_HIP(hipMemset(e->nBonds, 0, sizeof(int) * (size_t)N));
    MFT_HIP(hipMemset(e->bondedTo, 0, sizeof(int) * (size_t)N * mb));
    MFT_HIP(hipMemset(e->bondType, 0, sizeof(int) * (size_t)N * mb));
    MFT_HIP(hipMemset(e->type, 0, sizeof(int) * (size_t)N));
    MFT_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    return MFT_OK;
}


// Density channels and force fields once the charges are known.  When every
// species carries a single charge value (or there is no charges potential),
// the charge density is sum_s q_s rho_s, so electrostatics folds into the
// per-species k-space kernel and no separate charge channel, transform or
// field is needed; otherwise the charge density is a channel of its own.
static void build_groups(TildEngine* e);

static int configure_channels(TildEngine* e, const int* type, const float* charge)
{
    const mft_tild_desc& d = e->d;
    int rc;
    bool fold = true;
    std::vector<float> q(d.nspecies, 0.f);
    std::vector<int> seen(d.nspecies, 0);
    if (e->charge_pot >= 0) {
        for (size_t i = 0; i < (size_t)e->N; ++i) {
            int t = type[i];
            float c = charge ? charge[i] : 0.f;
            if (!seen[t]) {
                seen[t] = 1;
                q[t] = c;
            } else if (q[t] != c) {
                fold = false;
            }
        }
    }
    KParams& kp = e->kp;
    kp.fold = fold ? 1 : 0;
    for (int s = 0; s < kMaxSpecies; ++s) kp.qspec[s] = (fold && s < d.nspecies) ? q[s] : 0.f;
    e->nchan = d.nspecies;
    e->qchan = -1;
    if (e->charge_pot >= 0 && !fold) e->qchan = e->nchan++;
    MFT_REQUIRE(e->nchan <= kMaxChan, MFT_ERR_UNSUPPORTED, "at most %d density channels (species + charge)",
                kMaxChan);
    uint32_t forced = 0;
    for (int i = 0; i < d.npotentials; ++i) {
        const mft_tild_potential& P = e->pots[i];
        if (P.kind == MFT_POT_WALL || P.kind == MFT_POT_LANGEVIN || P.kind == MFT_POT_DPD)
            continue; // per-particle force
        if (P.kind != MFT_POT_CHARGES) forced |= P.maskI | (P.kind == MFT_POT_BIAS ? 0u : P.maskJ);
    }
    if (e->charge_pot >= 0 && fold)
        for (int s = 0; s < d.nspecies; ++s)
            if (q[s] != 0.f) forced |= 1u << s;
    e->force_slot_h.assign(d.nspecies, -1);
    e->nforce = 0;
    for (int s = 0; s < d.nspecies; ++s)
        if ((forced >> s) & 1u) {
            kp.force_species[e->nforce] = s;
            e->force_slot_h[s] = e->nforce++;
        }
    e->qslot = (e->charge_pot >= 0 && !fold) ? e->nforce : -1;
    e->nout4 = e->nforce + (e->qslot >= 0 ? 1 : 0);
    e->nout = 3 * e->nout4;
    e->npk = (e->nout + 1) / 2;
    kp.nchan = e->nchan;
    kp.qchan = e->qchan;
    kp.nforce = e->nforce;

    const size_t M = e->M;
    for (void* pbuf : {(void*)e->rho, (void*)e->S, (void*)e->Y, (void*)e->F4})
        if (pbuf) (void)hipFree(pbuf);
    e->rho = nullptr;
    e->S = nullptr;
    e->Y = nullptr;
    e->F4 = nullptr;
    if ((rc = dalloc(&e->rho, (size_t)e->nchan * M))) return rc;
    if ((rc = dalloc(&e->S, (size_t)e->nchan * M))) return rc;
    if ((rc = dalloc(&e->Y, (size_t)(e->npk > 0 ? e->npk : 1) * M))) return rc;
    if ((rc = dalloc(&e->F4, (size_t)(e->nout4 > 0 ? e->nout4 : 1) * M))) return rc;
    MFT_HIP(hipMemcpy(e->force_slot, e->force_slot_h.data(), sizeof(int) * d.nspecies, hipMemcpyHostToDevice));
    if ((rc = tild_tiles_configure(e))) return rc;
    build_groups(e);
    e->configured = 1;
    return MFT_OK;
}

// ---- launch helpers ----
struct TimedScope {
    TildEngine* e;
    int k;
    long slot = -1;
    TimedScope(TildEngine* e_, int k_) : e(e_), k(k_)
    {
        if (e->timing && ((e->tk_mask >> k) & 1u) && 2 * (e->evused + 1) <= e->evpool.size()) {
            slot = (long)e->evused++;
            e->evkind[slot] = k;
            (void)hipEventRecord(e->evpool[2 * slot], e->stream);
        }
    }
    ~TimedScope()
    {
        if (slot >= 0) (void)hipEventRecord(e->evpool[2 * slot + 1], e->stream);
    }
};

static void resolve_timing(TildEngine* e)
{
    if (e->evused == 0) return;
    (void)hipStreamSynchronize(e->stream);
    for (size_t i = 0; i < e->evused; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e->evpool[2 * i], e->evpool[2 * i + 1]) == hipSuccess) {
            e->tk_ms[e->evkind[i]] += ms;
            e->tk_n[e->evkind[i]] += 1;
        }
    }
    (void)hipGetLastError();
    e->evused = 0;
}

// Everything up to the per-particle grid force in fgrid.
static int grid_forces(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    int rc;
    {
        TimedScope ts(e, MFT_TK_DENSITY);
        if ((rc = tild_density(e))) return rc;
    }
    if (e->nout == 0) {
        MFT_HIP(hipMemsetAsync(e->fgrid, 0, sizeof(float4) * (size_t)e->N, e->stream));
        return MFT_OK;
    }
    const long M = e->M;
    AxisGeom gx{d.nx, d.ny, d.nz, 0}, gy{d.nx, d.ny, d.nz, 1};
    const TileGeom& tg = e->tg;
    const long cstride = (long)tg.ex[0] * tg.ex[1] * tg.ex[2];
    const bool plane = e->plane_passes && tild_plane_supported(d.nx, d.ny);
    if (plane) {
        // x and y of each z-plane in one pass (timed as the x pass)
        TimedScope ts(e, MFT_TK_FFT_X_FWD);
        if ((rc = tild_xy_fwd_tiles_plane(e->tiles, e->tile_tab, e->S, d.nx, d.ny, d.nz, e->nchan, cstride,
                                          e->stream)))
            return rc;
    } else {
        {
            TimedScope ts(e, MFT_TK_FFT_X_FWD);
            const bool staged = e->xfwd_staged && tild_x_fwd_staged_ok(d.nx, tg.ex[0], e->tile_maxcover);
            if ((rc = tild_x_fwd_tiles(e->tiles, e->tile_tab, e->S, d.nx, d.ny, d.nz, e->nchan, cstride, e->stream,
                                       tg.ex[0], staged, e->tile_maxcover <= 2)))
                return rc;
        }
        TimedScope ts(e, MFT_TK_FFT_Y_FWD);
        if ((rc = tild_axis(-1, 0, 0, e->S, e->S, gy, e->nchan, M, 1.0f, e->stream))) return rc;
    }
    {
        TimedScope ts(e, MFT_TK_CONV_Z);
        if ((rc = tild_conv(e))) return rc;
    }
    if (!plane) {
        TimedScope ts(e, MFT_TK_FFT_Y_INV);
        if ((rc = tild_axis(+1, 0, 0, e->Y, e->Y, gy, e->npk, M, 1.0f, e->stream))) return rc;
    }
    {
        TimedScope ts(e, MFT_TK_FFT_X_INV);
        if (plane) rc = tild_yx_inv_packed_plane(e->Y, e->F4, d.nx, d.ny, d.nz, e->nout4, e->stream);
        else rc = tild_x_inv_packed(e->Y, e->F4, d.nx, d.ny, d.nz, e->nout4, e->stream);
        if (rc) return rc;
        if (!e->bias.empty() && (rc = tild_bias_add(e))) return rc;
    }
    {
        TimedScope ts(e, MFT_TK_GATHER);
        if ((rc = tild_gather(e, e->gather_mode))) return rc;
    }
    return MFT_OK;
}

// Groups of potentials with one k-space shape and their coupling matrices
// (from the current, possibly ramped, amplitudes).
static void build_groups(TildEngine* e)
{
    KParams& kp = e->kp;
    kp.ngroups = 0;
    std::memset(kp.grp, 0, sizeof(kp.grp));
    auto find = [&](int kind, float sig2, const float* table) -> int {
        for (int g = 0; g < kp.ngroups; ++g)
            if (kp.grp[g].kind == kind && kp.grp[g].sig2 == sig2 && kp.grp[g].table == table) return g;
        if (kp.ngroups == kMaxGroups) return -1;
        KGroup& G = kp.grp[kp.ngroups];
        G.kind = kind;
        G.sig2 = sig2;
        G.table = table;
        return kp.ngroups++;
    };
    for (int i = 0; i < kp.npots; ++i) {
        const KPot& P = kp.pot[i];
        int g;
        if (P.A == 0.f && P.maskI == 0u && P.maskJ == 0u) continue; // no-op (bias fields)
        if (P.kind == MFT_POT_CHARGES) {
            if (!kp.fold) continue; // separate charge field
            g = find(MFT_POT_CHARGES, P.sig2, nullptr);
        } else if (P.kind == MFT_POT_GAUSSIAN) {
            g = find(MFT_POT_GAUSSIAN, P.sig2, nullptr);
        } else {
            g = find(P.kind, 0.f, P.table);
        }
        if (g < 0) {
            // more distinct shapes than kMaxGroups: the kernel evaluates each
            // potential on its own
            kp.ngroups = -1;
            return;
        }
        for (int f = 0; f < kp.nforce; ++f) {
            const int sp = kp.force_species[f];
            for (int c = 0; c < kp.nchan; ++c) {
                if (c == kp.qchan) continue;
                if (P.kind == MFT_POT_CHARGES) {
                    kp.grp[g].coef[f][c] += 12.566370614359172f * P.A * kp.qspec[sp] * kp.qspec[c];
                } else {
                    const float inI = (float)((P.maskI >> sp) & 1u), inJ = (float)((P.maskJ >> sp) & 1u);
                    kp.grp[g].coef[f][c] +=
                        P.A * (inI * (float)((P.maskJ >> c) & 1u) + inJ * (float)((P.maskI >> c) & 1u));
                }
            }
        }
    }
}

// PS_Potential::update_prefactor (ps_potential.cu:227-248): after step i of
// a run of nsteps the amplitude is Ao_i + (Ao_f - Ao_i)/nsteps * (i + 1).
static float ramp_amplitude(const mft_tild_potential& P, int64_t i, int64_t nsteps)
{
    return (float)((double)P.A + ((double)P.A_final - (double)P.A) / (double)nsteps * (double)(i + 1));
}

static void update_ramps(TildEngine* e, int64_t i, int64_t nsteps)
{
    for (int p = 0; p < e->d.npotentials; ++p) {
        const mft_tild_potential& P = e->pots[p];
        if (!P.ramp) continue;
        e->kp.pot[p].A = ramp_amplitude(P, i, nsteps);
        e->kp.pot[p].eA = e->kp.pot[p].A;
    }
    build_groups(e);
}

// advance = false: a force/energy evaluation at the current state (the step
// counter and the potential ramps stay where they are)
// energy_last: the k-space pass of the last step also sums the pair
// energies of that step's densities (the positions before the integrator
// moves them), with the amplitudes the ramps reach after the step -- what
// PS_Box::doTimeStep logs (PS_Box.cu:152-173: update_prefactor, then
// writeData -> computeThermoProps -> CalcEnergy on the step's d_rho).
static int do_steps(TildEngine* e, int nsteps, bool want_f, bool advance = true, bool energy_last = false)
{
    const mft_tild_desc& d = e->d;
    int rc;
    for (int i = 0; i < nsteps; ++i) {
        if (e->timing && 2 * (e->evused + MFT_TK_COUNT) > e->evpool.size()) resolve_timing(e);
        const bool elog = energy_last && i == nsteps - 1;
        if (elog) {
            e->kp.energy = 1;
            for (int p = 0; p < d.npotentials; ++p) {
                const mft_tild_potential& P = e->pots[p];
                if (!P.ramp || !advance) continue;
                e->kp.pot[p].eA = e->run_max > 0 ? ramp_amplitude(P, e->run_pos, e->run_max)
                                                 : ramp_amplitude(P, i, nsteps);
            }
        }
        rc = grid_forces(e);
        if (elog) {
            e->kp.energy = 0;
            for (int p = 0; p < d.npotentials; ++p) e->kp.pot[p].eA = e->kp.pot[p].A;
        }
        if (rc) return rc;
        if (e->dpd.on && (rc = tild_dpd_forces(e))) return rc;
        const bool last = i == nsteps - 1;
        {
            TimedScope ts(e, MFT_TK_INTEGRATE);
            if ((rc = tild_integrate(e, want_f && last))) return rc;
        }
        if (d.integrator != MFT_INT_NONE) {
            // the next positions into their key order
            {
                TimedScope ts(e, MFT_TK_SCAN);
                if ((rc = tild_scan(e, true))) return rc;
            }
            TimedScope ts(e, MFT_TK_SCATTER);
            if ((rc = tild_permute(e))) return rc;
        }
        if (!advance) continue;
        if (d.integrator == MFT_INT_VV) e->vv_state = 1;
        if (d.integrator == MFT_INT_GJF) e->nz_keyed = true;
        if (e->run_max > 0) update_ramps(e, e->run_pos++, e->run_max);
        else update_ramps(e, i, nsteps);
        e->step++;
    }
    return MFT_OK;
}

} // namespace mft

using namespace mft;

struct mft_tild : TildEngine {};

extern "C" {

int mft_tild_create(mft_tild** out, const mft_tild_desc* desc)
{
    MFT_REQUIRE(out && desc, MFT_ERR_INVALID, "null argument");
    *out = nullptr;
    mft_tild* e = new mft_tild();
    e->d = *desc;
    e->pots.assign(desc->potentials, desc->potentials + (desc->npotentials > 0 ? desc->npotentials : 0));
    e->d.potentials = nullptr;
    if (desc->nspecies > 0 && desc->nspecies <= kMaxSpecies) {
        e->mass.assign(desc->nspecies, 1.0f);
        e->mob.assign(desc->nspecies, 1.0f);
        for (int s = 0; s < desc->nspecies; ++s) {
            if (desc->mass) e->mass[s] = desc->mass[s];
            if (desc->mobility) e->mob[s] = desc->mobility[s];
        }
    }
    int rc = tild_setup(e);
    if (rc) {
        mft_tild_destroy(e);
        return rc;
    }
    *out = e;
    return MFT_OK;
}

int mft_tild_destroy(mft_tild* e)
{
    if (!e) return MFT_OK;
    void* ptrs[] = {e->X[0], e->X[1], e->XO[0], e->XO[1], e->NZ[0], e->NZ[1], e->META[0], e->META[1],
                    e->BS[0], e->BS[1], e->BT[0], e->BT[1], e->XN, e->NN, e->fgrid, e->ftot, e->key,
                    e->rank, e->map, e->slot_of, e->type, e->charge, e->nBonds, e->bondedTo, e->bondType, e->pio,
                    e->pio2, e->pio3, e->count, e->start, e->partial, e->rho, e->S, e->Y, e->F4, e->force_slot,
                    e->epart, e->eout, e->bpart, e->tiles, e->tile_tab, e->NA[0], e->NA[1], e->AT[0], e->AT[1],
                    e->AG[0], e->AG[1], e->nAngles, e->angleType, e->angleGroup, e->dcount, e->dstart,
                    e->dcursor, e->dcell, e->dslots, e->fdpd, e->d_step};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (float* t : e->tables)
        if (t) (void)hipFree(t);
    for (auto& b : e->bias) {
        (void)hipFree(b.F);
        (void)hipFree(b.u);
    }
    for (hipEvent_t ev : e->evpool) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return MFT_OK;
}

void* mft_tild_stream(mft_tild* e) { return e ? (void*)e->stream : nullptr; }

int mft_tild_set_angles(mft_tild* e, const int* nAngles, const int* angleType, const int* angleGroup)
{
    MFT_REQUIRE(e && nAngles && angleType && angleGroup, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(e->d.maxangles > 0, MFT_ERR_INVALID, "the engine was created without angles (maxangles 0)");
    MFT_REQUIRE(!e->have_state, MFT_ERR_INVALID, "set the angles before the state");
    const size_t N = e->N, ma = e->d.maxangles;
    for (size_t i = 0; i < N; ++i) {
        MFT_REQUIRE(nAngles[i] >= 0 && nAngles[i] <= (int)ma, MFT_ERR_INVALID, "particle %zu has %d angles", i,
                    nAngles[i]);
        for (int a = 0; a < nAngles[i]; ++a) {
            const int t = angleType[i * ma + a];
            MFT_REQUIRE(t >= 0 && t < e->ap.ntypes, MFT_ERR_INVALID, "angle type %d out of range", t);
            bool member = false;
            for (int c = 0; c < 3; ++c) {
                const int p = angleGroup[(i * ma + a) * 3 + c];
                MFT_REQUIRE(p >= 0 && (size_t)p < N, MFT_ERR_INVALID, "angle member %d out of range", p);
                member = member || (size_t)p == i;
            }
            MFT_REQUIRE(member, MFT_ERR_INVALID, "particle %zu is not a member of its angle %d", i, a);
        }
    }
    MFT_HIP(hipMemcpy(e->nAngles, nAngles, sizeof(int) * N, hipMemcpyHostToDevice));
    MFT_HIP(hipMemcpy(e->angleType, angleType, sizeof(int) * N * ma, hipMemcpyHostToDevice));
    MFT_HIP(hipMemcpy(e->angleGroup, angleGroup, sizeof(int) * N * ma * 3, hipMemcpyHostToDevice));
    return MFT_OK;
}

int mft_tild_angle_energy(mft_tild* e, double* energy)
{
    MFT_REQUIRE(e && energy && e->have_state, MFT_ERR_INVALID, "engine has no state");
    if (e->d.maxangles <= 0) {
        *energy = 0.0;
        return MFT_OK;
    }
    return tild_angle_energy(e, energy);
}

int mft_tild_set_topology(mft_tild* e, const int* type, const float* charge, const int* nBonds, const int* bondedTo,
                          const int* bondType)
{
    MFT_REQUIRE(e && type, MFT_ERR_INVALID, "null argument");
    const size_t N = e->N, mb = e->d.maxbonds > 0 ? e->d.maxbonds : 1;
    for (size_t i = 0; i < N; ++i)
        MFT_REQUIRE(type[i] >= 0 && type[i] < e->d.nspecies, MFT_ERR_INVALID, "particle %zu has type %d", i, type[i]);
    MFT_HIP(hipMemcpy(e->type, type, sizeof(int) * N, hipMemcpyHostToDevice));
    if (e->d.has_charges) {
        MFT_REQUIRE(charge, MFT_ERR_INVALID, "charges required");
        // PS_Box_initialization.cu:398-412: the box must be neutral
        double qt = 0.0;
        for (size_t i = 0; i < N; ++i) qt += charge[i];
        MFT_REQUIRE(std::fabs(qt) < 1e-3, MFT_ERR_INVALID, "Box net charge is not zero! qtotal: %g", qt);
        MFT_HIP(hipMemcpy(e->charge, charge, sizeof(float) * N, hipMemcpyHostToDevice));
    }
    int rc = configure_channels(e, type, e->d.has_charges ? charge : nullptr);
    if (rc) return rc;
    if (nBonds) {
        for (size_t i = 0; i < N; ++i) {
            MFT_REQUIRE(nBonds[i] >= 0 && nBonds[i] <= (int)mb, MFT_ERR_INVALID, "particle %zu has %d bonds", i,
                        nBonds[i]);
            for (int b = 0; b < nBonds[i]; ++b) {
                int id2 = bondedTo[i * mb + b], bt = bondType[i * mb + b];
                MFT_REQUIRE(id2 >= 0 && (size_t)id2 < N, MFT_ERR_INVALID, "bond partner %d out of range", id2);
                MFT_REQUIRE(bt >= 0 && bt < e->d.nbondtypes, MFT_ERR_INVALID, "bond type %d out of range", bt);
            }
        }
        MFT_HIP(hipMemcpy(e->nBonds, nBonds, sizeof(int) * N, hipMemcpyHostToDevice));
        MFT_HIP(hipMemcpy(e->bondedTo, bondedTo, sizeof(int) * N * mb, hipMemcpyHostToDevice));
        MFT_HIP(hipMemcpy(e->bondType, bondType, sizeof(int) * N * mb, hipMemcpyHostToDevice));
    }
    // the key-ordered state is built from the topology: set the state again
    e->have_state = 0;
    return MFT_OK;
}

int mft_tild_set_state(mft_tild* e, const float* x, const float* xo, const float* noise_old)
{
    MFT_REQUIRE(e && x, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(e->configured, MFT_ERR_INVALID, "set the topology before the state");
    const size_t n3 = 3 * (size_t)e->N;
    const float Ls[3] = {e->d.lx, e->d.ly, e->d.lz};
    for (size_t i = 0; i < n3; ++i)
        MFT_REQUIRE(std::isfinite(x[i]) && x[i] >= -Ls[i % 3] && x[i] <= 2.0f * Ls[i % 3], MFT_ERR_NONFINITE,
                    "position %zu is not finite or far outside the box: %g", i, x[i]);
    MFT_HIP(hipStreamSynchronize(e->stream));
    e->cur = 0;
    MFT_HIP(hipMemcpy(e->pio, x, sizeof(float) * n3, hipMemcpyHostToDevice));
    MFT_HIP(hipMemcpy(e->pio2, xo ? xo : x, sizeof(float) * n3, hipMemcpyHostToDevice));
    // velocity Verlet starts from rest with no stored force (the reference's
    // calloc'd velocities and freshly allocated force array)
    MFT_REQUIRE(!(noise_old && e->d.integrator == MFT_INT_VV), MFT_ERR_INVALID,
                "velocity Verlet starts from zero velocities: pass no noise_old");
    e->vv_state = 0;
    if (noise_old) MFT_HIP(hipMemcpy(e->pio3, noise_old, sizeof(float) * n3, hipMemcpyHostToDevice));
    else MFT_HIP(hipMemset(e->pio3, 0, sizeof(float) * n3));
    int rc = tild_sort_state(e);
    if (rc) return rc;
    MFT_HIP(hipStreamSynchronize(e->stream));
    e->have_state = 1;
    return MFT_OK;
}

int mft_tild_get_state(mft_tild* e, float* x, float* xo, float* noise_old)
{
    MFT_REQUIRE(e && e->have_state, MFT_ERR_INVALID, "engine has no state");
    const size_t n3 = 3 * (size_t)e->N;
    float* outs[3] = {x, xo, noise_old};
    for (int w = 0; w < 3; ++w) {
        if (!outs[w]) continue;
        int rc = tild_unsort(e, w);
        if (rc) return rc;
        MFT_HIP(hipStreamSynchronize(e->stream));
        MFT_HIP(hipMemcpy(outs[w], e->pio, sizeof(float) * n3, hipMemcpyDeviceToHost));
    }
    return MFT_OK;
}

int mft_tild_compute_forces(mft_tild* e, float* f)
{
    MFT_REQUIRE(e && e->have_state, MFT_ERR_INVALID, "engine has no state");
    int rc;
    int saved = e->d.integrator;
    e->d.integrator = MFT_INT_NONE;
    int64_t st = e->step;
    rc = do_steps(e, 1, true, false);
    e->d.integrator = saved;
    e->step = st;
    if (rc) return rc;
    if (f) {
        if ((rc = tild_unsort(e, 3))) return rc;
        MFT_HIP(hipStreamSynchronize(e->stream));
        MFT_HIP(hipMemcpy(f, e->pio, sizeof(float) * 3 * (size_t)e->N, hipMemcpyDeviceToHost));
    }
    return MFT_OK;
}

int mft_tild_get_densities(mft_tild* e, float* rho, float* rhoq)
{
    MFT_REQUIRE(e, MFT_ERR_INVALID, "null argument");
    if (!e->rho_valid) {
        int rc = tild_density_reduce(e);
        if (rc) return rc;
    }
    MFT_HIP(hipStreamSynchronize(e->stream));
    const size_t M = e->M;
    if (rho) MFT_HIP(hipMemcpy(rho, e->rho, sizeof(float) * M * e->d.nspecies, hipMemcpyDeviceToHost));
    if (rhoq) {
        MFT_REQUIRE(e->charge_pot >= 0, MFT_ERR_INVALID, "no charges potential");
        if (e->qchan >= 0) {
            MFT_HIP(hipMemcpy(rhoq, e->rho + (size_t)e->qchan * M, sizeof(float) * M, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> r(M * e->d.nspecies);
            MFT_HIP(hipMemcpy(r.data(), e->rho, sizeof(float) * r.size(), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < M; ++i) {
                double v = 0.0;
                for (int sp = 0; sp < e->d.nspecies; ++sp) v += (double)e->kp.qspec[sp] * r[(size_t)sp * M + i];
                rhoq[i] = (float)v;
            }
        }
    }
    return MFT_OK;
}

int mft_tild_get_grid_force(mft_tild* e, int s, float* F)
{
    MFT_REQUIRE(e && F && s >= 0 && s <= e->d.nspecies, MFT_ERR_INVALID, "bad argument");
    const size_t M = e->M;
    int slot = s == e->d.nspecies ? e->qslot : e->force_slot_h[s];
    if (slot < 0) {
        std::memset(F, 0, sizeof(float) * 3 * M);
        return MFT_OK;
    }
    // the field's first 3 M floats are its points' (fx, fy, fz)
    MFT_HIP(hipStreamSynchronize(e->stream));
    MFT_HIP(hipMemcpy(F, e->F4 + (size_t)slot * M, sizeof(float) * 3 * M, hipMemcpyDeviceToHost));
    return MFT_OK;
}

int mft_tild_step(mft_tild* e, int nsteps)
{
    MFT_REQUIRE(e && e->have_state, MFT_ERR_INVALID, "engine has no state");
    MFT_REQUIRE(nsteps >= 0, MFT_ERR_INVALID, "negative step count");
    return do_steps(e, nsteps, false);
}

int mft_tild_set_run(mft_tild* e, int64_t max_steps)
{
    MFT_REQUIRE(e && max_steps >= 0, MFT_ERR_INVALID, "bad argument");
    e->run_max = max_steps;
    e->run_pos = 0;
    return MFT_OK;
}

int mft_tild_synchronize(mft_tild* e)
{
    MFT_REQUIRE(e, MFT_ERR_INVALID, "null argument");
    MFT_HIP(hipStreamSynchronize(e->stream));
    return MFT_OK;
}

int mft_tild_bond_virial(mft_tild* e, float* vir)
{
    MFT_REQUIRE(e && vir, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(e->have_state, MFT_ERR_INVALID, "engine has no state");
    float* dv = nullptr;
    int rc = dalloc(&dv, (size_t)6 * e->N);
    if (rc) return rc;
    rc = tild_bond_virial(e, dv);
    if (!rc) {
        hipError_t err = hipStreamSynchronize(e->stream);
        if (err == hipSuccess) err = hipMemcpy(vir, dv, sizeof(float) * 6 * (size_t)e->N, hipMemcpyDeviceToHost);
        if (err != hipSuccess) {
            set_error(MFT_ERR_DEVICE, "bond virial readback: %s", hipGetErrorString(err));
            rc = MFT_ERR_DEVICE;
        }
    }
    (void)hipFree(dv);
    return rc;
}

int mft_tild_energies(mft_tild* e, double* pot_energy, double* bond_energy)
{
    MFT_REQUIRE(e && e->have_state, MFT_ERR_INVALID, "engine has no state");
    const mft_tild_desc& d = e->d;
    int rc;
    if (pot_energy && d.npotentials > 0) {
        // densities and spectra at the current positions
        int saved = d.integrator;
        e->d.integrator = MFT_INT_NONE;
        e->kp.energy = 1;
        int64_t st = e->step;
        rc = do_steps(e, 1, false, false);
        e->kp.energy = 0;
        e->d.integrator = saved;
        e->step = st;
        if (rc) return rc;
        MFT_HIP(hipStreamSynchronize(e->stream));
        MFT_HIP(hipMemcpy(pot_energy, e->eout, sizeof(double) * d.npotentials, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < e->bias.size(); ++k)
            if ((rc = tild_bias_energy(e, (int)k, &pot_energy[e->bias[k].pot]))) return rc;
        for (size_t w = 0; w < e->wall_pot.size(); ++w)
            if ((rc = tild_wall_energy(e, (int)w, &pot_energy[e->wall_pot[w]]))) return rc;
    }
    if (bond_energy) {
        if ((rc = tild_bond_energy(e))) return rc;
        MFT_HIP(hipStreamSynchronize(e->stream));
        double b;
        MFT_HIP(hipMemcpy(&b, e->eout + kMaxPots, sizeof(double), hipMemcpyDeviceToHost));
        *bond_energy = 0.5 * b;
    }
    return MFT_OK;
}

} // extern "C"

namespace mft {
// device-buffer energies: the pair-potential entries (zero for bias, wall,
// DPD and Langevin entries, which these do not cover) and the bond energy
// with the 1/2 of PS_Box::computeThermoProps (PS_Box.cu:292)
__global__ void k_finish_energies(double* __restrict__ out, const double* __restrict__ eout, int npot,
                                  unsigned pair_mask)
{
    const int p = threadIdx.x;
    if (p < npot) out[p] = ((pair_mask >> p) & 1u) ? eout[p] : 0.0;
    if (p == 0) out[npot] = 0.5 * eout[kMaxPots];
}
} // namespace mft

extern "C" {

int mft_tild_energies_device(mft_tild* e, double* d_out)
{
    MFT_REQUIRE(e && e->have_state && d_out, MFT_ERR_INVALID, "bad argument");
    const mft_tild_desc& d = e->d;
    int rc;
    unsigned pair_mask = 0;
    for (int p = 0; p < d.npotentials; ++p) {
        const int k = e->pots[p].kind;
        if (k == MFT_POT_GAUSSIAN || k == MFT_POT_ERF2 || k == MFT_POT_ERFG || k == MFT_POT_CHARGES)
            pair_mask |= 1u << p;
    }
    if (d.npotentials > 0) {
        int saved = d.integrator;
        e->d.integrator = MFT_INT_NONE;
        e->kp.energy = 1;
        int64_t st = e->step;
        rc = do_steps(e, 1, false, false);
        e->kp.energy = 0;
        e->d.integrator = saved;
        e->step = st;
        if (rc) return rc;
    }
    if ((rc = tild_bond_energy(e))) return rc;
    k_finish_energies<<<1, 64, 0, e->stream>>>(d_out, e->eout, d.npotentials, pair_mask);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// energies of every kind except the pair potentials, at the current positions
static int tail_energies(TildEngine* e, double* pot_energy, double* bond_energy)
{
    int rc;
    if (pot_energy) {
        for (size_t k = 0; k < e->bias.size(); ++k)
            if ((rc = tild_bias_energy(e, (int)k, &pot_energy[e->bias[k].pot]))) return rc;
        for (size_t w = 0; w < e->wall_pot.size(); ++w)
            if ((rc = tild_wall_energy(e, (int)w, &pot_energy[e->wall_pot[w]]))) return rc;
    }
    if (bond_energy) {
        if ((rc = tild_bond_energy(e))) return rc;
        MFT_HIP(hipStreamSynchronize(e->stream));
        double b;
        MFT_HIP(hipMemcpy(&b, e->eout + kMaxPots, sizeof(double), hipMemcpyDeviceToHost));
        *bond_energy = 0.5 * b;
    }
    return MFT_OK;
}

int mft_tild_step_energies(mft_tild* e, int nsteps, double* pot_energy, double* bond_energy)
{
    MFT_REQUIRE(e && e->have_state, MFT_ERR_INVALID, "engine has no state");
    MFT_REQUIRE(nsteps >= 1, MFT_ERR_INVALID, "step count must be at least 1");
    const mft_tild_desc& d = e->d;
    int rc;
    if (d.integrator == MFT_INT_VV) {
        // the reference's VV step computes its forces after the drift, at the
        // positions it ends with, which the fused pass reaches only in the
        // next step: evaluate there
        if ((rc = do_steps(e, nsteps, false))) return rc;
        return mft_tild_energies(e, pot_energy, bond_energy);
    }
    if ((rc = do_steps(e, nsteps, false, true, pot_energy && d.npotentials > 0))) return rc;
    if (pot_energy && d.npotentials > 0) {
        MFT_HIP(hipStreamSynchronize(e->stream));
        MFT_HIP(hipMemcpy(pot_energy, e->eout, sizeof(double) * d.npotentials, hipMemcpyDeviceToHost));
    }
    return tail_energies(e, pot_energy, bond_energy);
}

int mft_tild_timing(mft_tild* e, int enable, double* avg_ms, int64_t* launches)
{
    MFT_REQUIRE(e, MFT_ERR_INVALID, "null argument");
    if (enable && e->evpool.empty()) {
        const size_t pairs = 8192;
        e->evpool.resize(2 * pairs);
        e->evkind.assign(pairs, 0);
        for (auto& ev : e->evpool) MFT_HIP(hipEventCreate(&ev));
    }
    resolve_timing(e);
    if (enable && !e->timing) {
        for (int k = 0; k < MFT_TK_COUNT; ++k) {
            e->tk_ms[k] = 0;
            e->tk_n[k] = 0;
        }
    }
    e->timing = enable != 0;
    for (int k = 0; k < MFT_TK_COUNT; ++k) {
        if (avg_ms) avg_ms[k] = e->tk_n[k] ? e->tk_ms[k] / (double)e->tk_n[k] : 0.0;
        if (launches) launches[k] = e->tk_n[k];
    }
    return MFT_OK;
}

int mft_tild_timing_select(mft_tild* e, uint32_t mask)
{
    MFT_REQUIRE(e, MFT_ERR_INVALID, "null argument");
    e->tk_mask = mask;
    return MFT_OK;
}

int mft_tild_kernel_bytes(mft_tild* e, double* bytes)
{
    MFT_REQUIRE(e && bytes, MFT_ERR_INVALID, "null argument");
    const double N = e->N, M = e->M, C = e->nchan, O = e->npk, O4 = e->nout4;
    const bool q = e->d.has_charges;
    for (int k = 0; k < MFT_TK_COUNT; ++k) bytes[k] = 0;
    // minimal HBM traffic of each kernel's job (reads + writes once)
    const double mb = e->d.maxbonds > 0 ? e->d.maxbonds : 1;
    bytes[MFT_TK_BIN] = 0;                                    // sorting happens when the state is set
    bytes[MFT_TK_SCAN] = (double)e->nkeys * (4 + 4 + 4);      // count read, start write, count clear
    // permutation: key, rank, start, next x and noise, x, id/species, bond
    // types in; x, previous x, noise, id/species, bond types, map out; bond
    // slots remapped (no bond count: a slot's bonds are its partners >= 0)
    // under GJF the previous noise is regenerated, not carried (16 bytes in
    // and out of the permutation and of the integrator fewer)
    const double nz = e->d.integrator == MFT_INT_GJF ? 0.0 : 16.0;
    // bond types per slot only with more than one type (else none carried)
    const double bt = e->bp.ntypes > 1 ? 4 * mb : 0.0;
    bytes[MFT_TK_SCATTER] = N * ((4 + 4 + 4 + 12 + nz + 16 + 4 + bt) + (16 + 12 + nz + 4 + bt + 4) +
                                 (4 + 8 * mb + 4 * mb));
    bytes[MFT_TK_DENSITY] = N * 16 + M * 4 * C;               // key-ordered positions, rho
    const bool plane = e->plane_passes && tild_plane_supported(e->d.nx, e->d.ny);
    if (plane) {
        bytes[MFT_TK_FFT_X_FWD] = M * C * (4 + 8);          // densities in, x-y spectra out
        bytes[MFT_TK_FFT_X_INV] = M * O * 8 + M * O4 * 12;  // packed spectra in, force fields out
    } else {
        bytes[MFT_TK_FFT_X_FWD] = M * C * (4 + 8);          // densities in, spectra out (tiles summed on the fly)
        bytes[MFT_TK_FFT_Y_FWD] = M * C * 16;
        bytes[MFT_TK_FFT_Y_INV] = M * O * 16;
        bytes[MFT_TK_FFT_X_INV] = M * O * 8 + M * O4 * 12;   // force fields of 12-byte points
    }
    bytes[MFT_TK_CONV_Z] = M * 8 * (C + O);
    // per-slot grid forces and next positions as 12-byte records
    bytes[MFT_TK_GATHER] = N * (16 + 8 + 12) + M * 12 * O4;
    bytes[MFT_TK_INTEGRATE] = N * ((16 + 12 + nz + 4 + 4 * mb + bt + 12) + (12 + nz + 4 + 4));
    return MFT_OK;
}

int mft_tild_set_gather(mft_tild* e, int variant)
{
    MFT_REQUIRE(e, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(variant >= 0 && variant <= 3, MFT_ERR_INVALID, "gather variant %d", variant);
    e->gather_mode = variant;
    return MFT_OK;
}

int mft_tild_autotune_timings(mft_tild* e, double* ms, int n)
{
    MFT_REQUIRE(e && ms && n >= 0, MFT_ERR_INVALID, "bad argument");
    for (int i = 0; i < n; ++i) ms[i] = i < 8 ? e->tune_ms[i] : -1.0;
    return MFT_OK;
}

int mft_tild_set_grid_passes(mft_tild* e, int variant)
{
    MFT_REQUIRE(e, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(variant >= 0 && variant <= 2, MFT_ERR_INVALID, "grid pass variant %d", variant);
    MFT_REQUIRE(variant != 1 || tild_plane_supported(e->d.nx, e->d.ny), MFT_ERR_UNSUPPORTED,
                "plane passes need square x-y planes of 64, 96 or 128 points (grid %d x %d)", e->d.nx, e->d.ny);
    MFT_REQUIRE(variant != 2 || tild_x_fwd_staged_ok(e->d.nx, e->tg.ex[0], e->tile_maxcover), MFT_ERR_UNSUPPORTED,
                "the staged x pass does not fit this grid");
    e->plane_passes = variant == 1;
    e->xfwd_staged = variant == 2;
    return MFT_OK;
}

int mft_tild_autotune(mft_tild* e, int reps, int* choice)
{
    MFT_REQUIRE(e && e->have_state && reps >= 1, MFT_ERR_INVALID, "engine has no state");
    // one force evaluation fills the cell-ordered particles and force fields;
    // the gather only writes the per-particle grid force, so the variants
    // can be timed in place without changing the state
    int saved = e->d.integrator;
    int64_t st = e->step;
    e->d.integrator = MFT_INT_NONE;
    int rc = do_steps(e, 1, false, false);
    e->d.integrator = saved;
    e->step = st;
    if (rc) return rc;
    hipEvent_t a, b;
    MFT_HIP(hipEventCreate(&a));
    MFT_HIP(hipEventCreate(&b));
    float best = 1e30f;
    int best_mode = 0;
    for (int mode = 0; mode < 4; ++mode) {
        if (mode == 2 && !tild_gather_brick_fits(e)) continue;
        if ((rc = tild_gather(e, mode))) return rc;  // warm
        MFT_HIP(hipEventRecord(a, e->stream));
        for (int r = 0; r < reps; ++r)
            if ((rc = tild_gather(e, mode))) return rc;
        MFT_HIP(hipEventRecord(b, e->stream));
        MFT_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        MFT_HIP(hipEventElapsedTime(&ms, a, b));
        e->tune_ms[mode] = ms / reps;
        if (ms < best) {
            best = ms;
            best_mode = mode;
        }
    }
    e->gather_mode = best_mode;
    // line or plane x-y passes: the whole grid-force evaluation timed with
    // each (densities to per-particle grid forces; the state is unchanged)
    int best_plane = 0;
    if (tild_plane_supported(e->d.nx, e->d.ny)) {
        float bestg = 1e30f;
        for (int v = 0; v < 2; ++v) {
            e->plane_passes = v;
            if ((rc = grid_forces(e))) return rc;   // warm
            MFT_HIP(hipEventRecord(a, e->stream));
            for (int r = 0; r < reps; ++r)
                if ((rc = grid_forces(e))) return rc;
            MFT_HIP(hipEventRecord(b, e->stream));
            MFT_HIP(hipEventSynchronize(b));
            float ms = 0.f;
            MFT_HIP(hipEventElapsedTime(&ms, a, b));
            e->tune_ms[4 + v] = ms / reps;
            if (ms < bestg) {
                bestg = ms;
                best_plane = v;
            }
        }
    }
    e->plane_passes = best_plane;
    // forward x pass reading the tiles directly or through LDS-staged rows
    int best_staged = 0;
    if (!best_plane && tild_x_fwd_staged_ok(e->d.nx, e->tg.ex[0], e->tile_maxcover)) {
        float bests = 1e30f;
        for (int v = 0; v < 2; ++v) {
            e->xfwd_staged = v;
            if ((rc = grid_forces(e))) return rc;   // warm
            MFT_HIP(hipEventRecord(a, e->stream));
            for (int r = 0; r < reps; ++r)
                if ((rc = grid_forces(e))) return rc;
            MFT_HIP(hipEventRecord(b, e->stream));
            MFT_HIP(hipEventSynchronize(b));
            float ms = 0.f;
            MFT_HIP(hipEventElapsedTime(&ms, a, b));
            e->tune_ms[6 + v] = ms / reps;
            if (ms < bests) {
                bests = ms;
                best_staged = v;
            }
        }
    }
    e->xfwd_staged = best_staged;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (choice) *choice = best_mode | (best_plane << 2) | (best_staged << 3);
    return MFT_OK;
}

int mft_tild_num_outputs(mft_tild* e, int* nout)
{
    MFT_REQUIRE(e && nout, MFT_ERR_INVALID, "null argument");
    *nout = e->nout;
    return MFT_OK;
}

} // extern "C"

// ------------------------------------------------- replica groups ---
// Independent TILD systems advanced by one HIP graph: kGroupSteps steps of
// every engine are captured once, each engine's launches on its own stream
// between a fork from and a join back to the group's stream, so one graph
// launch runs those steps of all replicas with their kernels overlapping on
// the device.  The integrator reads the step index from device memory and the
// key scan advances it, so a replay keys its noise by the steps it runs; an
// even step count brings every engine's double buffers back to the captured
// parity.  Engines
// whose configuration a graph cannot replay (ramps, DPD, Langevin, velocity
// Verlet, timing) and GJF engines whose previous noise is not yet keyed (the
// first step after a state is set) take ordinary steps.
namespace mft {

constexpr int kGroupSteps = 8;   // steps per graph replay (even)

__global__ void k_set_step(uint32_t* __restrict__ d, uint32_t v) { *d = v; }

static bool graph_ready(const TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    if (d.integrator != MFT_INT_GJF && d.integrator != MFT_INT_EM) return false;
    if (e->dpd.on || e->lang.n || e->timing || e->nout == 0) return false;
    for (int p = 0; p < d.npotentials; ++p)
        if (e->pots[p].ramp) return false;
    if (d.integrator == MFT_INT_GJF && !e->nz_keyed) return false;
    return true;
}

} // namespace mft

struct mft_tild_group {
    std::vector<mft_tild*> eng;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> cur0, gather0, plane0;   // what the graph was captured with
    std::vector<int64_t> dstep_at;            // host step the device counter holds (-1: stale)
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join, pre, post;
};

static void group_drop_graph(mft_tild_group* g)
{
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->exec = nullptr;
    g->graph = nullptr;
}

// capture kGroupSteps steps of every engine into one graph (nothing runs;
// the host bookkeeping the capture advanced is restored)
static int group_capture(mft_tild_group* g)
{
    group_drop_graph(g);
    const size_t n = g->eng.size();
    std::vector<int64_t> step(n), run_pos(n);
    std::vector<int> cur(n);
    std::vector<bool> rho_valid(n), keyed(n);
    for (size_t i = 0; i < n; ++i) {
        TildEngine* e = g->eng[i];
        step[i] = e->step;
        run_pos[i] = e->run_pos;
        cur[i] = e->cur;
        rho_valid[i] = e->rho_valid;
        keyed[i] = e->nz_keyed;
        e->use_dstep = true;
    }
    int rc = MFT_OK;
    hipError_t err = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    if (err == hipSuccess) err = hipEventRecord(g->fork, g->stream);
    for (size_t i = 0; i < n && err == hipSuccess && rc == MFT_OK; ++i) {
        TildEngine* e = g->eng[i];
        err = hipStreamWaitEvent(e->stream, g->fork, 0);
        if (err != hipSuccess) break;
        rc = do_steps(e, kGroupSteps, false);
        if (rc) break;
        err = hipEventRecord(g->join[i], e->stream);
        if (err == hipSuccess) err = hipStreamWaitEvent(g->stream, g->join[i], 0);
    }
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(g->stream, &graph);
    for (size_t i = 0; i < n; ++i) {
        TildEngine* e = g->eng[i];
        e->use_dstep = false;
        e->step = step[i];
        e->run_pos = run_pos[i];
        e->cur = cur[i];
        e->rho_valid = rho_valid[i];
        e->nz_keyed = keyed[i];
    }
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (err == hipSuccess) err = end;
    if (err == hipSuccess) err = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        g->exec = nullptr;
        set_error(MFT_ERR_DEVICE, "capturing the replica group's steps: %s", hipGetErrorString(err));
        return MFT_ERR_DEVICE;
    }
    g->graph = graph;
    for (size_t i = 0; i < n; ++i) {
        g->cur0[i] = g->eng[i]->cur;
        g->gather0[i] = g->eng[i]->gather_mode;
        g->plane0[i] = g->eng[i]->plane_passes | (g->eng[i]->xfwd_staged << 1);
    }
    return MFT_OK;
}

extern "C" {

int mft_tild_group_create(mft_tild_group** out, mft_tild* const* engines, int n)
{
    MFT_REQUIRE(out && engines && n >= 1, MFT_ERR_INVALID, "bad argument");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        MFT_REQUIRE(engines[i], MFT_ERR_INVALID, "null engine %d", i);
        for (int j = 0; j < i; ++j) MFT_REQUIRE(engines[j] != engines[i], MFT_ERR_INVALID, "engine %d twice", i);
    }
    mft_tild_group* g = new mft_tild_group();
    g->eng.assign(engines, engines + n);
    g->cur0.assign(n, -1);
    g->gather0.assign(n, -1);
    g->plane0.assign(n, -1);
    g->dstep_at.assign(n, -1);
    g->join.assign(n, nullptr);
    g->pre.assign(n, nullptr);
    g->post.assign(n, nullptr);
    hipError_t err = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
    for (int i = 0; i < n && err == hipSuccess; ++i) {
        err = hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&g->pre[i], hipEventDisableTiming);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&g->post[i], hipEventDisableTiming);
    }
    if (err != hipSuccess) {
        mft_tild_group_destroy(g);
        set_error(MFT_ERR_DEVICE, "replica group set-up: %s", hipGetErrorString(err));
        return MFT_ERR_DEVICE;
    }
    *out = g;
    return MFT_OK;
}

int mft_tild_group_destroy(mft_tild_group* g)
{
    if (!g) return MFT_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    group_drop_graph(g);
    if (g->fork) (void)hipEventDestroy(g->fork);
    for (auto* v : {&g->join, &g->pre, &g->post})
        for (hipEvent_t ev : *v)
            if (ev) (void)hipEventDestroy(ev);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return MFT_OK;
}

int mft_tild_group_step(mft_tild_group* g, int nsteps)
{
    MFT_REQUIRE(g && nsteps >= 0, MFT_ERR_INVALID, "bad argument");
    const size_t n = g->eng.size();
    for (size_t i = 0; i < n; ++i) MFT_REQUIRE(g->eng[i]->have_state, MFT_ERR_INVALID, "engine %zu has no state", i);
    int rc;
    int done = 0;
    while (done < nsteps) {
        bool ready = nsteps - done >= kGroupSteps;
        for (size_t i = 0; i < n && ready; ++i) ready = graph_ready(g->eng[i]);
        if (!ready) {
            // one ordinary step of every engine, each on its own stream
            for (size_t i = 0; i < n; ++i) {
                if ((rc = do_steps(g->eng[i], 1, false))) return rc;
                g->dstep_at[i] = -1;
            }
            ++done;
            continue;
        }
        bool same = g->exec != nullptr;
        for (size_t i = 0; i < n && same; ++i) {
            const TildEngine* e = g->eng[i];
            same = g->gather0[i] == e->gather_mode && g->plane0[i] == (e->plane_passes | (e->xfwd_staged << 1));
        }
        if (!same && (rc = group_capture(g))) return rc;
        bool parity = true;
        for (size_t i = 0; i < n; ++i) parity = parity && g->eng[i]->cur == g->cur0[i];
        if (!parity) {
            // an odd number of ordinary steps since the capture: one more
            for (size_t i = 0; i < n; ++i) {
                if ((rc = do_steps(g->eng[i], 1, false))) return rc;
                g->dstep_at[i] = -1;
            }
            ++done;
            continue;
        }
        // each engine's earlier work and device step index before the replay
        for (size_t i = 0; i < n; ++i) {
            TildEngine* e = g->eng[i];
            if (g->dstep_at[i] != e->step) {
                k_set_step<<<1, 1, 0, e->stream>>>(e->d_step, (uint32_t)e->step);
                MFT_LAUNCH_CHECK();
            }
            MFT_HIP(hipEventRecord(g->pre[i], e->stream));
            MFT_HIP(hipStreamWaitEvent(g->stream, g->pre[i], 0));
        }
        MFT_HIP(hipGraphLaunch(g->exec, g->stream));
        for (size_t i = 0; i < n; ++i) {
            TildEngine* e = g->eng[i];
            MFT_HIP(hipEventRecord(g->post[i], g->stream));
            MFT_HIP(hipStreamWaitEvent(e->stream, g->post[i], 0));
            e->step += kGroupSteps;
            if (e->run_max > 0) e->run_pos += kGroupSteps;
            e->rho_valid = false;
            g->dstep_at[i] = e->step;
        }
        done += kGroupSteps;
    }
    return MFT_OK;
}

int mft_tild_group_synchronize(mft_tild_group* g)
{
    MFT_REQUIRE(g, MFT_ERR_INVALID, "null argument");
    MFT_HIP(hipStreamSynchronize(g->stream));
    for (mft_tild* e : g->eng) MFT_HIP(hipStreamSynchronize(e->stream));
    return MFT_OK;
}

} // extern "C"
