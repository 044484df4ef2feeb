// Per-kernel entry points on caller-owned device arrays in the reference's
// layouts, for callers that keep the reference's own loop and swap one
// kernel at a time: k-space potential tables, one potential's grid forces
// and energy, the GJF/EM position updates (TILD); species fields, one
// potential's field update, propagators, densities, Q and one potential's
// Hamiltonian term (CL-FTS).  Complex arrays are interleaved float (re, im).
// The engines (mft_tild_*, mft_fts_*) run the same arithmetic fused; these
// entry points trade speed for the reference's data flow.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "tild_impl.h"

namespace mft {

// ------------------------------------------------------------ helpers ---
// in-place 3D transform of one complex grid: forward scaled by 1/M (the
// reference's cufftWrapper*, Box.cu:58-127), inverse unscaled
static int fft3_complex(float2* a, int nx, int ny, int nz, int sign, hipStream_t st)
{
    const long M = (long)nx * ny * nz;
    const float s = sign < 0 ? (float)(1.0 / (double)M) : 1.0f;
    for (int ax = 0; ax < 3; ++ax) {
        AxisGeom g{nx, ny, nz, ax};
        const int rc = tild_axis(sign, 0, 0, a, a, g, 1, M, ax == 0 ? s : 1.0f, st);
        if (rc) return rc;
    }
    return MFT_OK;
}

static int check_grid(const int Nx[3])
{
    MFT_REQUIRE(Nx && Nx[0] > 0 && Nx[1] > 0 && Nx[2] > 0, MFT_ERR_INVALID, "grid sizes must be positive");
    for (int a = 0; a < 3; ++a)
        MFT_REQUIRE(mft_fft_supported(Nx[a]), MFT_ERR_UNSUPPORTED, "grid size %d is not a supported FFT length",
                    Nx[a]);
    return MFT_OK;
}

__global__ void k_real_to_complex(float2* __restrict__ out, const float* __restrict__ in, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) out[i] = make_float2(in[i], 0.f);
}

// out = fk[i*3 + j] * S[i]
__global__ void k_mul_component(float2* __restrict__ out, const float2* __restrict__ fk,
                                const float2* __restrict__ S, int j, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) out[i] = cmul(fk[i * 3 + j], S[i]);
}

// gridF[i*3 + j] += factor * Re a[i]  (accumulateGridForceComp, ps_groups.cu:164-178)
__global__ void k_accumulate_component(float* __restrict__ gridF, const float2* __restrict__ a, int j, float factor,
                                       long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) gridF[i * 3 + j] += factor * a[i].x;
}

__global__ void k_mul_inplace(float2* __restrict__ a, const float2* __restrict__ b, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) a[i] = cmul(a[i], b[i]);
}

// FP64 block partial sums of sum_i a[i] * Re b[i]
__global__ void k_dot_re(double* __restrict__ part, const float* __restrict__ a, const float2* __restrict__ b, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = i < M ? (double)a[i] * (double)b[i].x : 0.0;
    __shared__ double red[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// FP64 sums of device partials, read back in a fixed order
static int sum_partials(const double* dpart, int n, int ncomp, double* out, hipStream_t st)
{
    std::vector<double> h((size_t)n * ncomp);
    MFT_HIP(hipMemcpyAsync(h.data(), dpart, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    MFT_HIP(hipStreamSynchronize(st));
    for (int c = 0; c < ncomp; ++c) {
        double s = 0.0;
        for (int b = 0; b < n; ++b) s += h[(size_t)b * ncomp + c];
        out[c] = s;
    }
    return MFT_OK;
}

// ------------------------------------------------------------ TILD ---
// Potential tables of initializePotential (ps_potentialGaussian.cu:28-71,
// ps_potentialErf2.cu:31-67, ps_potentialErfG.cu:30-66,
// ps_potentialCharges.cu:30-80): u(k) and f_j(k) = -i k_j u(k).
static float table_value(int kind, float A, float sig2, float k2, float erf_shape)
{
    if (kind == MFT_POT_GAUSSIAN) return A * std::exp(-k2 * sig2 * 0.5f);
    if (kind == MFT_POT_CHARGES) return k2 > 0.f ? 12.566370614359172f * A / k2 * std::exp(-k2 * sig2) : 0.f;
    return A * erf_shape;
}

} // namespace mft

using namespace mft;

extern "C" {

int mft_tild_kspace_table(float* uk, float* fk, int kind, float A, float sigma, float Rp, const int Nx[3],
                          const float L[3], void* stream)
{
    int rc;
    if ((rc = check_grid(Nx))) return rc;
    MFT_REQUIRE(L && L[0] > 0.f && L[1] > 0.f && L[2] > 0.f, MFT_ERR_INVALID, "box lengths must be positive");
    MFT_REQUIRE(kind >= MFT_POT_GAUSSIAN && kind <= MFT_POT_CHARGES, MFT_ERR_INVALID, "potential kind %d", kind);
    const int nx = Nx[0], ny = Nx[1], nz = Nx[2];
    const size_t M = (size_t)nx * ny * nz;
    std::vector<float> shape;
    if (kind == MFT_POT_ERF2 || kind == MFT_POT_ERFG) erf_table(kind, sigma, Rp, nx, ny, nz, L[0], L[1], L[2], shape);
    std::vector<float2> u(uk ? M : 0), f(fk ? 3 * M : 0);
    const float sig2 = sigma * sigma;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const size_t i = ix + (size_t)nx * (iy + (size_t)ny * iz);
                const float k[3] = {kcomp(ix, nx, L[0]), kcomp(iy, ny, L[1]), kcomp(iz, nz, L[2])};
                const float k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                const float v = table_value(kind, A, sig2, k2, shape.empty() ? 0.f : shape[i]);
                if (uk) u[i] = make_float2(v, 0.f);
                if (fk)
                    for (int j = 0; j < 3; ++j) f[i * 3 + j] = make_float2(0.f, -k[j] * v);
            }
    hipStream_t st = (hipStream_t)stream;
    if (uk) MFT_HIP(hipMemcpyAsync(uk, u.data(), sizeof(float2) * M, hipMemcpyHostToDevice, st));
    if (fk) MFT_HIP(hipMemcpyAsync(fk, f.data(), sizeof(float2) * 3 * M, hipMemcpyHostToDevice, st));
    MFT_HIP(hipStreamSynchronize(st));
    return MFT_OK;
}

int mft_tild_pair_force(float* gridF, const float* rho, const float* fk, const int Nx[3], float factor, void* work,
                        void* stream)
{
    int rc;
    if ((rc = check_grid(Nx))) return rc;
    MFT_REQUIRE(gridF && rho && fk && work, MFT_ERR_INVALID, "null argument");
    const long M = (long)Nx[0] * Nx[1] * Nx[2];
    hipStream_t st = (hipStream_t)stream;
    float2* S = (float2*)work;
    float2* W = S + M;
    k_real_to_complex<<<nblk(M), 256, 0, st>>>(S, rho, M);
    if ((rc = fft3_complex(S, Nx[0], Nx[1], Nx[2], -1, st))) return rc;
    for (int j = 0; j < 3; ++j) {
        k_mul_component<<<nblk(M), 256, 0, st>>>(W, (const float2*)fk, S, j, M);
        if ((rc = fft3_complex(W, Nx[0], Nx[1], Nx[2], +1, st))) return rc;
        k_accumulate_component<<<nblk(M), 256, 0, st>>>(gridF, W, j, factor, M);
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int mft_tild_potential_energy(double* energy, const float* rhoI, const float* rhoJ, const float* uk,
                              const int Nx[3], double gvol, void* work, void* stream)
{
    int rc;
    if ((rc = check_grid(Nx))) return rc;
    MFT_REQUIRE(energy && rhoI && rhoJ && uk && work, MFT_ERR_INVALID, "null argument");
    const long M = (long)Nx[0] * Nx[1] * Nx[2];
    hipStream_t st = (hipStream_t)stream;
    float2* a = (float2*)work;
    double* part = (double*)(a + M);
    k_real_to_complex<<<nblk(M), 256, 0, st>>>(a, rhoJ, M);
    if ((rc = fft3_complex(a, Nx[0], Nx[1], Nx[2], -1, st))) return rc;
    k_mul_inplace<<<nblk(M), 256, 0, st>>>(a, (const float2*)uk, M);
    if ((rc = fft3_complex(a, Nx[0], Nx[1], Nx[2], +1, st))) return rc;
    k_dot_re<<<nblk(M), 256, 0, st>>>(part, rhoI, a, M);
    MFT_LAUNCH_CHECK();
    double e = 0.0;
    if ((rc = sum_partials(part, (int)nblk(M), 1, &e, st))) return rc;
    *energy = gvol * e;
    return MFT_OK;
}

} // extern "C"

namespace mft {

// Integrate_2 of GJF (ps_integratorGJF.cu:102-163) and EM
// (ps_integratorEM.cu:39-82) on AoS arrays, noise keyed by particle index and
// step like the engine's
__global__ void k_gjf_aos(float* __restrict__ x, float* __restrict__ xo, float* __restrict__ no,
                          const float* __restrict__ f, const int* __restrict__ type, GjfParams c, int n, float L0,
                          float L1, float L2, uint64_t seed, uint32_t step)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float z[4];
    normal4((uint32_t)i, step, TAG_TILD, 0u, seed, z);
    const int ty = min(max(type[i], 0), kMaxSpecies - 1);
    const float a = c.a[ty], b = c.b[ty], nm = c.nmag[ty], cf = c.cf[ty], cn = c.cn[ty];
    const float Ls[3] = {L0, L1, L2};
    for (int j = 0; j < 3; ++j) {
        const long k = (long)i * 3 + j;
        const float nnj = nm * z[j];
        const float xc = x[k];
        float xold = xo[k];
        if (xold - xc > Ls[j] / 2.0f) xold -= Ls[j];
        else if (xold - xc < -Ls[j] / 2.0f) xold += Ls[j];
        float v = 2.0f * b * xc - a * xold + cf * f[k] + cn * (nnj + no[k]);
        if (v > Ls[j]) v -= Ls[j];
        else if (v < 0.0f) v += Ls[j];
        x[k] = v;
        xo[k] = xc;
        no[k] = nnj;
    }
}

__global__ void k_em_aos(float* __restrict__ x, const float* __restrict__ f, int n, float delt, float L0, float L1,
                         float L2, uint64_t seed, uint32_t step)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float z[4];
    normal4((uint32_t)i, step, TAG_TILD, 0u, seed, z);
    const float Ls[3] = {L0, L1, L2};
    const float nm = sqrtf(2.0f * delt);
    for (int j = 0; j < 3; ++j) {
        const long k = (long)i * 3 + j;
        float v = x[k] + delt * f[k] + nm * z[j];
        if (v > Ls[j]) v -= Ls[j];
        else if (v < 0.0f) v += Ls[j];
        x[k] = v;
    }
}

} // namespace mft

extern "C" {

int mft_tild_gjf_step(float* x, float* xo, float* noise_old, const float* f, const int* type, int nparticles,
                      int nspecies, const float* mass, const float* mobility, float delt, const float L[3],
                      uint64_t seed, int64_t step, void* stream)
{
    MFT_REQUIRE(x && xo && noise_old && f && type && L && nparticles >= 0, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(nspecies >= 1 && nspecies <= kMaxSpecies, MFT_ERR_INVALID, "nspecies must be 1..%d", kMaxSpecies);
    GjfParams c;
    std::memset(&c, 0, sizeof(c));
    for (int s = 0; s < nspecies; ++s) {
        // GJF::finishInitialization (ps_integratorGJF.cu:56-82), in float
        const float m = mass ? mass[s] : 1.0f, D = mobility ? mobility[s] : 1.0f, dt = delt;
        const float b = m / (m + dt / (2.0f * D));
        c.a[s] = (1.0f - dt / (2.0f * D * m)) / (1.0f + dt / (2.0f * D * m));
        c.b[s] = b;
        c.nmag[s] = sqrtf(2.0f * dt) / sqrtf(D);
        c.cf[s] = b * (dt * dt) / m;
        c.cn[s] = b * dt / (2.0f * m);
    }
    if (nparticles == 0) return MFT_OK;
    k_gjf_aos<<<nblk(nparticles), 256, 0, (hipStream_t)stream>>>(x, xo, noise_old, f, type, c, nparticles, L[0],
                                                                        L[1], L[2], seed, (uint32_t)step);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int mft_tild_em_step(float* x, const float* f, int nparticles, float delt, const float L[3], uint64_t seed,
                     int64_t step, void* stream)
{
    MFT_REQUIRE(x && f && L && nparticles >= 0, MFT_ERR_INVALID, "null argument");
    if (nparticles == 0) return MFT_OK;
    k_em_aos<<<nblk(nparticles), 256, 0, (hipStream_t)stream>>>(x, f, nparticles, delt, L[0], L[1], L[2], seed,
                                                                       (uint32_t)step);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

} // extern "C"

// ----------------------------------------------- structure factor ---
namespace mft {

// density of a group of species (the sum of its species' densities)
__global__ void k_group_density(float2* __restrict__ out, const float* __restrict__ rho, uint32_t mask, int nspecies,
                                long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    float v = 0.f;
    for (int sp = 0; sp < nspecies; ++sp)
        if ((mask >> sp) & 1u) v += rho[(long)sp * M + i];
    out[i] = make_float2(v, 0.f);
}

// sk = |a|^2 (or +=): rho(k) rho(-k) of a real field (d_multiplyCpxByCpxConj)
__global__ void k_abs2(float* __restrict__ sk, const float2* __restrict__ a, long M, int accumulate)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const float2 v = a[i];
    const float s = v.x * v.x + v.y * v.y;
    sk[i] = accumulate ? sk[i] + s : s;
}

} // namespace mft

extern "C" int mft_tild_structure_factor(mft_tild* eng, uint32_t species_mask, float* sk, int accumulate, void* work)
{
    // mft_tild derives from TildEngine with nothing added (tild.hip)
    TildEngine* e = reinterpret_cast<TildEngine*>(eng);
    MFT_REQUIRE(e && sk && work, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(species_mask != 0 && (e->d.nspecies >= 32 || (species_mask >> e->d.nspecies) == 0), MFT_ERR_INVALID,
                "species mask 0x%x does not match %d species", species_mask, e->d.nspecies);
    int rc;
    if (!e->rho_valid && (rc = tild_density_reduce(e))) return rc;
    const long M = e->M;
    float2* a = (float2*)work;
    k_group_density<<<nblk(M), 256, 0, e->stream>>>(a, e->rho, species_mask, e->d.nspecies, M);
    if ((rc = fft3_complex(a, e->d.nx, e->d.ny, e->d.nz, -1, e->stream))) return rc;
    k_abs2<<<nblk(M), 256, 0, e->stream>>>(sk, a, M, accumulate);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// ------------------------------------------------------------- FTS ---
namespace mft {

constexpr int kApiMaxPots = 8;

struct SpeciesArgs {
    int kind[kApiMaxPots], I[kApiMaxPots], J[kApiMaxPots];
    const float2* wpl[kApiMaxPots];
    const float2* wmi[kApiMaxPots];
    int npots, nspecies;
    double invNr;
};

// FTS_Species::buildPotentialField (fts_species.cu:53-132)
__global__ void k_species_fields(float2* __restrict__ W, SpeciesArgs A, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    for (int s = 0; s < A.nspecies; ++s) {
        double wr = 0.0, wi = 0.0;
        for (int p = 0; p < A.npots; ++p) {
            const float2 w = A.wpl[p][i];
            if (A.kind[p] == MFT_FTS_FLORY) {
                if (s != A.I[p] && s != A.J[p]) continue;
                const float2 v = A.wmi[p][i];
                const double sg = s == A.I[p] ? -1.0 : 1.0;
                wr += sg * v.x - w.y;
                wi += sg * v.y + w.x;
            } else {
                wr -= w.y;
                wi += w.x;
            }
        }
        W[s * M + i] = make_float2((float)(wr * A.invNr), (float)(wi * A.invNr));
    }
}

struct UpdateArgs {
    int kind, I, J, noisy, col;
    double cpl, cmi, dt_pl, dt_mi, nmag_pl, nmag_mi, C, Nr;
    const float2* rho[kApiMaxPots];
    int nspecies;
    uint64_t seed;
    uint32_t iter;
};

// the potential's force (d_makeFloryForce fts_potential_flory.cu:279-304,
// d_makeHelfandForce fts_potential_helfand.cu:223-240, d_makeIncompressForce
// fts_potential_incompress.cu:202-218) and the EM step with real CL noise
// (fts_updateSchemes.cu:8-29, fts_make_CLnoise.cu:7-24)
__global__ void k_update_potential(float2* __restrict__ wpl, float2* __restrict__ wmi, UpdateArgs A, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    float z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (A.noisy) {
        normal4((uint32_t)i, A.iter, TAG_FTS, 0u, A.seed, z);
        normal4((uint32_t)i, A.iter, TAG_FTS, 1u, A.seed, z + 4);
    }
    const float2 wp = wpl[i];
    double wx = wp.x, wy = wp.y;
    if (A.kind == MFT_FTS_FLORY) {
        const float2 wm = wmi[i];
        double vx = wm.x, vy = wm.y;
        const float2 rI = A.rho[A.I][i], rJ = A.rho[A.J][i];
        const double fpx = A.cpl * wx - ((double)rI.y + rJ.y) / A.Nr;
        const double fpy = A.cpl * wy + ((double)rI.x + rJ.x) / A.Nr;
        const double fmx = A.cmi * vx + ((double)rJ.x - rI.x) / A.Nr;
        const double fmy = A.cmi * vy + ((double)rJ.y - rI.y) / A.Nr;
        wx -= A.dt_pl * fpx;
        wy -= A.dt_pl * fpy;
        vx -= A.dt_mi * fmx;
        vy -= A.dt_mi * fmy;
        if (A.noisy) {
            wx += A.nmag_pl * (double)z[A.col];
            vx += A.nmag_mi * (double)z[A.col + 1];
        }
        wmi[i] = make_float2((float)vx, (float)vy);
    } else {
        double rx = 0.0, ry = 0.0;
        for (int s = 0; s < A.nspecies; ++s) {
            rx += A.rho[s][i].x;
            ry += A.rho[s][i].y;
        }
        double fx, fy;
        if (A.kind == MFT_FTS_HELFAND) {
            fx = A.cpl * wx - ry / A.Nr;
            fy = A.cpl * wy + rx / A.Nr - A.C;
        } else {
            fx = -ry / A.Nr;
            fy = rx / A.Nr - A.C;
        }
        wx -= A.dt_pl * fx;
        wy -= A.dt_pl * fy;
        if (A.noisy && A.kind == MFT_FTS_HELFAND) wx += A.nmag_pl * (double)z[A.col];
    }
    wpl[i] = make_float2((float)wx, (float)wy);
}

// q = exp(-W)  and  q *= exp(-W)
__global__ void k_exp_neg(float2* __restrict__ q, const float2* __restrict__ W, long M, int mul)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const float2 w = W[i];
    const double a = std::exp(-(double)w.x);
    double s, c;
    sincos(-(double)w.y, &s, &c);
    const float2 e = make_float2((float)(a * c), (float)(a * s));
    q[i] = mul ? cmul(q[i], e) : e;
}

__global__ void k_copy2(float2* __restrict__ dst, const float2* __restrict__ src, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) dst[i] = src[i];
}

// rho_b = factor e^{+W_b} sum_{s in b} q_s qdag_{N-1-s}, added to the block's
// species density (LinearMolec::calcDensity, fts_molecule_linear.cu:278-363)
__global__ void k_block_density(float2* __restrict__ rho, const float2* __restrict__ q,
                                const float2* __restrict__ qd, const float2* __restrict__ Wb, int s0, int nb, int Ntot,
                                double fr, double fi, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    double ar = 0.0, ai = 0.0;
    for (int k = 0; k < nb; ++k) {
        const int s = s0 + k;
        const float2 a = q[(long)s * M + i], c = qd[(long)(Ntot - 1 - s) * M + i];
        ar += (double)a.x * c.x - (double)a.y * c.y;
        ai += (double)a.x * c.y + (double)a.y * c.x;
    }
    const float2 w = Wb[i];
    const double e = std::exp((double)w.x);
    double sn, cs;
    sincos((double)w.y, &sn, &cs);
    const double er = e * cs, ei = e * sn;
    const double tr = ar * er - ai * ei, ti = ar * ei + ai * er;
    const float2 o = rho[i];
    rho[i] = make_float2((float)((double)o.x + fr * tr - fi * ti), (float)((double)o.y + fr * ti + fi * tr));
}

// FP64 block partial sums of a complex field (and of its square)
__global__ void k_csum(double* __restrict__ part, const float2* __restrict__ a, long M, int squares)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0, 0, 0, 0};
    if (i < M) {
        const double x = a[i].x, y = a[i].y;
        v[0] = x;
        v[1] = y;
        if (squares) {
            v[2] = x * x - y * y;
            v[3] = 2.0 * x * y;
        }
    }
    __shared__ double red[4][4];
    for (int c = 0; c < 4; ++c) {
        const double s = wave_sum(v[c]);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] +
                                                               red[2][threadIdx.x] + red[3][threadIdx.x];
}

static int check_molecule(const mft_fts_molecule* mol)
{
    MFT_REQUIRE(mol && mol->nblocks >= 1 && mol->nblocks <= 16, MFT_ERR_INVALID, "molecule needs 1..16 blocks");
    for (int b = 0; b < mol->nblocks; ++b)
        MFT_REQUIRE(mol->N[b] >= 1 && mol->species[b] >= 0, MFT_ERR_INVALID, "block %d: N %d species %d", b,
                    mol->N[b], mol->species[b]);
    return MFT_OK;
}

static int csum(const float2* a, long M, bool squares, double out[4], double* dpart, hipStream_t st)
{
    k_csum<<<nblk(M), 256, 0, st>>>(dpart, a, M, squares ? 1 : 0);
    MFT_LAUNCH_CHECK();
    return sum_partials(dpart, (int)nblk(M), 4, out, st);
}

// ------------------------------------------- allocation-free Q variant ---
// Device-resident scratch of mft_fts_partition_function_noalloc: the four
// FP64 block partial sums of k_csum for every 256-point block of the largest
// accepted grid, followed by the four totals.  Every call writes all the
// partials it later reads, so nothing needs clearing between calls; calls that
// overlap on one device share this array and race.
constexpr long kNoAllocMaxPoints = 1L << 24;                     // 256^3 grid points
constexpr int kNoAllocMaxBlocks = (int)(kNoAllocMaxPoints / 256);
constexpr int kMaxDevicesNoAlloc = 64;
static __device__ __align__(256) double scratch[4 * kNoAllocMaxBlocks + 4];

static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevicesNoAlloc];   // device address of `scratch` per device

// Address of this device's copy of `scratch`, looked up once per device with
// that device current (the array's host-side name is not a device address).
static int scratch_address(double** out)
{
    int dev = 0;
    MFT_HIP(hipGetDevice(&dev));
    MFT_REQUIRE(dev >= 0 && dev < kMaxDevicesNoAlloc, MFT_ERR_UNSUPPORTED, "device %d beyond the %d supported",
                dev, kMaxDevicesNoAlloc);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(MFT_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                      hipGetErrorString(e));
            return MFT_ERR_DEVICE;
        }
        g_scratch_dev[dev] = (double*)p;
    }
    *out = g_scratch_dev[dev];
    return MFT_OK;
}

// out[c] = sum_b part[b*ncomp + c], one thread per component adding the
// partials in block order: the same order and rounding as sum_partials on
// the host
__global__ void k_sum_partials_ordered(double* __restrict__ out, const double* __restrict__ part, int n, int ncomp)
{
    const int c = threadIdx.x;
    if (c >= ncomp) return;
    double s = 0.0;
    for (int b = 0; b < n; ++b) s += part[(long)b * ncomp + c];
    out[c] = s;
}

} // namespace mft

extern "C" {

int mft_fts_species_fields(float* W, int nspecies, int npotentials, const mft_fts_potential* pots,
                           const float* const* wpl, const float* const* wmi, double Nr, long M, void* stream)
{
    MFT_REQUIRE(W && (npotentials == 0 || (pots && wpl)) && M > 0, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(npotentials >= 0 && npotentials <= kApiMaxPots, MFT_ERR_INVALID, "at most %d potentials",
                kApiMaxPots);
    MFT_REQUIRE(nspecies >= 1 && Nr > 0.0, MFT_ERR_INVALID, "bad species count or Nr");
    SpeciesArgs A;
    std::memset(&A, 0, sizeof(A));
    A.npots = npotentials;
    A.nspecies = nspecies;
    A.invNr = 1.0 / Nr;
    for (int p = 0; p < npotentials; ++p) {
        A.kind[p] = pots[p].kind;
        A.I[p] = pots[p].I;
        A.J[p] = pots[p].J;
        A.wpl[p] = (const float2*)wpl[p];
        A.wmi[p] = wmi ? (const float2*)wmi[p] : nullptr;
        MFT_REQUIRE(A.wpl[p], MFT_ERR_INVALID, "w+ of potential %d is null", p);
        MFT_REQUIRE(A.kind[p] != MFT_FTS_FLORY || A.wmi[p], MFT_ERR_INVALID, "Flory potential %d needs w-", p);
    }
    k_species_fields<<<nblk(M), 256, 0, (hipStream_t)stream>>>((float2*)W, A, M);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int mft_fts_update_fields(float* wpl, float* wmi, const mft_fts_potential* pot, const float* const* rho,
                          int nspecies, double C, double Nr, const int Nx[3], const float L[3], int cl,
                          uint64_t seed, int64_t iter, int noise_col, void* stream)
{
    int rc;
    if ((rc = check_grid(Nx))) return rc;
    MFT_REQUIRE(wpl && pot && rho && L && nspecies >= 1 && nspecies <= kApiMaxPots, MFT_ERR_INVALID,
                "bad argument");
    MFT_REQUIRE(pot->kind != MFT_FTS_FLORY || wmi, MFT_ERR_INVALID, "Flory needs w-");
    MFT_REQUIRE(pot->kind != MFT_FTS_FLORY ||
                    (pot->I >= 0 && pot->I < nspecies && pot->J >= 0 && pot->J < nspecies && pot->I != pot->J),
                MFT_ERR_INVALID, "Flory species %d %d out of range", pot->I, pot->J);
    for (int s = 0; s < nspecies; ++s) MFT_REQUIRE(rho[s], MFT_ERR_INVALID, "density %d is null", s);
    MFT_REQUIRE(noise_col >= 0 && noise_col + (pot->kind == MFT_FTS_FLORY ? 1 : 0) < 8, MFT_ERR_INVALID,
                "noise column %d", noise_col);
    const long M = (long)Nx[0] * Nx[1] * Nx[2];
    // FTS_Box.cu:466-476: gvol from the float box lengths in double
    const double gvol = ((double)L[0] / Nx[0]) * ((double)L[1] / Nx[1]) * ((double)L[2] / Nx[2]);
    UpdateArgs A;
    std::memset(&A, 0, sizeof(A));
    A.kind = pot->kind;
    A.I = pot->I;
    A.J = pot->J;
    A.col = noise_col;
    A.noisy = cl && pot->kind != MFT_FTS_INCOMPRESS;
    A.dt_pl = pot->delt_pl;
    A.dt_mi = pot->delt_mi;
    A.C = C;
    A.Nr = Nr;
    if (pot->kind == MFT_FTS_FLORY) {
        A.cpl = A.cmi = 2.0 * C / pot->chiN;
        A.nmag_mi = std::sqrt(2.0 * pot->delt_mi / gvol);
    } else if (pot->kind == MFT_FTS_HELFAND) {
        A.cpl = C / pot->kappaN;
    } else if (pot->kind == MFT_FTS_EDWARDS) {
        // d_makeEdwardsForce (fts_potential_edwards.cu:227-244): Helfand's form
        // with 1/B and no -iC term
        A.kind = MFT_FTS_HELFAND;
        A.cpl = 1.0 / pot->kappaN;
        A.C = 0.0;
    }
    A.nmag_pl = std::sqrt(2.0 * pot->delt_pl / gvol);
    for (int s = 0; s < nspecies; ++s) A.rho[s] = (const float2*)rho[s];
    A.nspecies = nspecies;
    A.seed = seed;
    A.iter = (uint32_t)iter;
    k_update_potential<<<nblk(M), 256, 0, (hipStream_t)stream>>>((float2*)wpl, (float2*)wmi, A, M);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int mft_fts_propagate(float* q, float* qdag, const float* W, const float* g, const mft_fts_molecule* mol,
                      const int Nx[3], void* stream)
{
    int rc;
    if ((rc = check_grid(Nx))) return rc;
    MFT_REQUIRE((q || qdag) && W && g, MFT_ERR_INVALID, "bad argument");
    if ((rc = check_molecule(mol))) return rc;
    const long M = (long)Nx[0] * Nx[1] * Nx[2];
    hipStream_t st = (hipStream_t)stream;
    const float2* W2 = (const float2*)W;
    // LinearMolec::calcPropagators (fts_molecule_linear.cu:144-269):
    // q_0 = exp(-W_b0), q_s = exp(-W_b(s)) IFFT[g FFT[q_(s-1)]]; the
    // complementary chain runs over the blocks in reverse
    for (int chain = 0; chain < 2; ++chain) {
        float2* out = (float2*)(chain == 0 ? q : qdag);
        if (!out) continue;
        int s = 0;
        for (int bb = 0; bb < mol->nblocks; ++bb) {
            const int b = chain == 0 ? bb : mol->nblocks - 1 - bb;
            const float2* Wb = W2 + (long)mol->species[b] * M;
            for (int k = 0; k < mol->N[b]; ++k, ++s) {
                float2* qs = out + (long)s * M;
                if (s == 0) {
                    k_exp_neg<<<nblk(M), 256, 0, st>>>(qs, Wb, M, 0);
                    continue;
                }
                k_copy2<<<nblk(M), 256, 0, st>>>(qs, qs - M, M);
                if ((rc = fft3_complex(qs, Nx[0], Nx[1], Nx[2], -1, st))) return rc;
                k_mul_inplace<<<nblk(M), 256, 0, st>>>(qs, (const float2*)g, M);
                if ((rc = fft3_complex(qs, Nx[0], Nx[1], Nx[2], +1, st))) return rc;
                k_exp_neg<<<nblk(M), 256, 0, st>>>(qs, Wb, M, 1);
            }
        }
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int mft_fts_density(float* rho, const float* q, const float* qdag, const float* W, const mft_fts_molecule* mol,
                    double factor_re, double factor_im, long M, void* stream)
{
    MFT_REQUIRE(rho && q && qdag && W && M > 0, MFT_ERR_INVALID, "bad argument");
    int rc = check_molecule(mol);
    if (rc) return rc;
    int Ntot = 0;
    for (int b = 0; b < mol->nblocks; ++b) Ntot += mol->N[b];
    hipStream_t st = (hipStream_t)stream;
    int s0 = 0;
    for (int b = 0; b < mol->nblocks; ++b) {
        const long sp = mol->species[b];
        k_block_density<<<nblk(M), 256, 0, st>>>((float2*)rho + sp * M, (const float2*)q, (const float2*)qdag,
                                                      (const float2*)W + sp * M, s0, mol->N[b], Ntot, factor_re,
                                                      factor_im, M);
        s0 += mol->N[b];
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int mft_fts_partition_function(double* Q, const float* q_last, double gvol, double V, long M, void* work,
                               void* stream)
{
    MFT_REQUIRE(Q && q_last && work && M > 0 && V > 0.0, MFT_ERR_INVALID, "bad argument");
    double s[4];
    int rc = csum((const float2*)q_last, M, false, s, (double*)work, (hipStream_t)stream);
    if (rc) return rc;
    // fts_molecule_linear.cu:265-266
    Q[0] = gvol / V * s[0];
    Q[1] = gvol / V * s[1];
    return MFT_OK;
}

int mft_fts_partition_function_noalloc(double* Q, const float* q_last, double gvol, double V, long M, void* stream)
{
    MFT_REQUIRE(Q && q_last && M > 0 && V > 0.0, MFT_ERR_INVALID, "bad argument");
    MFT_REQUIRE(M <= kNoAllocMaxPoints, MFT_ERR_INVALID,
                "grid of %ld points exceeds the allocation-free limit of %ld points (256^3)", M, kNoAllocMaxPoints);
    double* part = nullptr;
    int rc = scratch_address(&part);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int nb = (int)nblk(M);
    double* tot = part + 4L * kNoAllocMaxBlocks;
    k_csum<<<nb, 256, 0, st>>>(part, (const float2*)q_last, M, 0);
    k_sum_partials_ordered<<<1, 64, 0, st>>>(tot, part, nb, 4);
    MFT_LAUNCH_CHECK();
    double s[4];
    MFT_HIP(hipMemcpyAsync(s, tot, sizeof(s), hipMemcpyDeviceToHost, st));
    MFT_HIP(hipStreamSynchronize(st));
    // fts_molecule_linear.cu:265-266
    Q[0] = gvol / V * s[0];
    Q[1] = gvol / V * s[1];
    return MFT_OK;
}

int mft_fts_hamiltonian_term(double* H, const mft_fts_potential* pot, const float* wpl, const float* wmi, double C,
                             double gvol, long M, void* work, void* stream)
{
    MFT_REQUIRE(H && pot && wpl && work && M > 0, MFT_ERR_INVALID, "bad argument");
    MFT_REQUIRE(pot->kind != MFT_FTS_FLORY || wmi, MFT_ERR_INVALID, "Flory needs w-");
    hipStream_t st = (hipStream_t)stream;
    double a[4], b[4] = {0, 0, 0, 0};
    int rc = csum((const float2*)wpl, M, true, a, (double*)work, st);
    if (rc) return rc;
    if (pot->kind == MFT_FTS_FLORY && (rc = csum((const float2*)wmi, M, true, b, (double*)work, st))) return rc;
    if (pot->kind == MFT_FTS_FLORY) {
        // fts_potential_flory.cu:252-274: C/chiN gvol sum (w+^2 + w-^2)
        H[0] = C / pot->chiN * gvol * (a[2] + b[2]);
        H[1] = C / pot->chiN * gvol * (a[3] + b[3]);
    } else if (pot->kind == MFT_FTS_HELFAND) {
        // fts_potential_helfand.cu:254-275: C/(2 kappaN) gvol sum w^2 - i C gvol sum w
        H[0] = C / (2.0 * pot->kappaN) * gvol * a[2] + C * gvol * a[1];
        H[1] = C / (2.0 * pot->kappaN) * gvol * a[3] - C * gvol * a[0];
    } else if (pot->kind == MFT_FTS_EDWARDS) {
        // PotentialEdwards::calcHamiltonian (fts_potential_edwards.cu:256-270): gvol sum w^2 / (2B)
        H[0] = gvol * a[2] / (2.0 * pot->kappaN);
        H[1] = gvol * a[3] / (2.0 * pot->kappaN);
    } else {
        // fts_potential_incompress.cu:232-247: -i C gvol sum w
        H[0] = C * gvol * a[1];
        H[1] = -C * gvol * a[0];
    }
    return MFT_OK;
}

} // extern "C"
