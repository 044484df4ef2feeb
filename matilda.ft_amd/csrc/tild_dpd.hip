// DPD pair forces of the TILD step (DPD, ps_potentialDPD.cu, with the
// neighbour search of PS_NeighborList, ps_neighborList.cu).  The reference
// rebuilds a per-particle neighbour list each step (cells of width >= rcut,
// CUB sort, 3^3 cell search, no skin) and adds the pair forces of i < j with
// atomics.  Here the group's particles are binned into the same cells by a
// counting sort and each particle sums the forces of all its partners itself:
// the pair noise is one normal keyed by (lower id, higher id, step), so both
// members compute the same pair force with opposite sign and no atomics on
// forces are needed.  The result is added to the particle forces in the
// integrate pass.
#include <cmath>

#include "tild_impl.h"

namespace mft {

__device__ __forceinline__ int dpd_cell1(float x, float cw, int n)
{
    int c = (int)(x / cw);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// cell of every slot of the group (-1 outside it) and the cell counts
__global__ void k_dpd_count(const float4* __restrict__ X, const PMeta* __restrict__ META, DpdParams P, int ns,
                            int* __restrict__ count, int* __restrict__ cell)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    if (!((P.mask >> META[s].sp()) & 1u)) {
        cell[s] = -1;
        return;
    }
    const float4 x = X[s];
    const int c = dpd_cell1(x.x, P.cw[0], P.nc[0]) +
                  P.nc[0] * (dpd_cell1(x.y, P.cw[1], P.nc[1]) + P.nc[1] * dpd_cell1(x.z, P.cw[2], P.nc[2]));
    cell[s] = c;
    atomicAdd(&count[c], 1);
}

// exclusive scan of the n cell counts in one workgroup: start[0..n] (start[n]
// = total) and a copy in cursor[0..n-1] for the fill
__global__ void __launch_bounds__(1024) k_dpd_scan(const int* __restrict__ count, int* __restrict__ start,
                                                  int* __restrict__ cursor, int n)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int chunk = (n + 1023) / 1024;
    const int b = min(n, t * chunk), e = min(n, b + chunk);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += count[i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t ? part[t - 1] : 0;
    for (int i = b; i < e; ++i) {
        start[i] = run;
        cursor[i] = run;
        run += count[i];
    }
    if (t == 1023) start[n] = part[1023];
}

__global__ void k_dpd_fill(const int* __restrict__ cell, int ns, int* __restrict__ cursor, int* __restrict__ slots)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int c = cell[s];
    if (c < 0) return;
    slots[atomicAdd(&cursor[c], 1)] = s;
}

// d_DPD_forces (ps_potentialDPD.cu:63-137) for every pair of slot s within
// rcut, found in the 27 cells around its own; V null means zero velocities
__global__ void __launch_bounds__(256) k_dpd_force(const float4* __restrict__ X, const PMeta* __restrict__ META,
                                                   const float4* __restrict__ V, const int* __restrict__ cell,
                                                   const int* __restrict__ start, const int* __restrict__ slots,
                                                   DpdParams P, float L0, float L1, float L2, uint64_t seed,
                                                   uint32_t step, int ns, float4* __restrict__ F)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int c = cell[s];
    if (c < 0) {
        F[s] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float L[3] = {L0, L1, L2};
    const float Lh[3] = {0.5f * L0, 0.5f * L1, 0.5f * L2};
    const int id = META[s].id();
    const float4 x4 = X[s];
    const float xi[3] = {x4.x, x4.y, x4.z};
    float vi[3] = {0.f, 0.f, 0.f};
    if (V) {
        const float4 v4 = V[s];
        vi[0] = v4.x;
        vi[1] = v4.y;
        vi[2] = v4.z;
    }
    const int cx = c % P.nc[0], cy = (c / P.nc[0]) % P.nc[1], cz = c / (P.nc[0] * P.nc[1]);
    float f[3] = {0.f, 0.f, 0.f};
    for (int dz = -1; dz <= 1; ++dz) {
        const int ncz = (cz + dz + P.nc[2]) % P.nc[2];
        for (int dy = -1; dy <= 1; ++dy) {
            const int ncy = (cy + dy + P.nc[1]) % P.nc[1];
            for (int dx = -1; dx <= 1; ++dx) {
                const int ncx = (cx + dx + P.nc[0]) % P.nc[0];
                const int cc = ncx + P.nc[0] * (ncy + P.nc[1] * ncz);
                const int e = start[cc + 1];
                for (int k = start[cc]; k < e; ++k) {
                    const int t = slots[k];
                    if (t == s) continue;
                    const float4 xj4 = X[t];
                    const float xj[3] = {xj4.x, xj4.y, xj4.z};
                    float dr[3], r2 = 0.0f;
                    for (int d = 0; d < 3; ++d) {
                        float v = xi[d] - xj[d];
                        if (v > Lh[d]) v -= L[d];
                        else if (v < -Lh[d]) v += L[d];
                        dr[d] = v;
                        r2 += v * v;
                    }
                    if (r2 >= P.rcut2 || r2 == 0.0f) continue;
                    const float r = sqrtf(r2);
                    const float w = 1.0f - r / P.rcut;
                    const float winv = 1.0f / r;
                    float vj[3] = {0.f, 0.f, 0.f};
                    if (V) {
                        const float4 v4 = V[t];
                        vj[0] = v4.x;
                        vj[1] = v4.y;
                        vj[2] = v4.z;
                    }
                    float vdot = 0.0f;
                    for (int d = 0; d < 3; ++d) vdot += (vi[d] - vj[d]) * (dr[d] * winv);
                    const int jd = META[t].id();
                    float z[4];
                    normal4((uint32_t)min(id, jd), (uint32_t)max(id, jd), TAG_DPD, step, seed, z);
                    const float Fmag = -P.gamma * w * w * vdot + P.sigma * w * z[0] * P.inv_sqrt_delt;
                    for (int d = 0; d < 3; ++d) f[d] += Fmag * (dr[d] * winv);
                }
            }
        }
    }
    F[s] = make_float4(f[0], f[1], f[2], 0.f);
}

int tild_dpd_setup(TildEngine* e, int pot)
{
    const mft_tild_desc& d = e->d;
    const mft_tild_potential& Pt = e->pots[pot];
    DpdParams& P = e->dpd;
    MFT_REQUIRE(!P.on, MFT_ERR_UNSUPPORTED, "at most one dpd potential");
    MFT_REQUIRE(Pt.maskI != 0 && Pt.Rp > 0.0f, MFT_ERR_INVALID, "dpd needs a group and a neighbor list rcut > 0");
    MFT_REQUIRE(Pt.wall_pos > 0.0f, MFT_ERR_INVALID, "ps_potentialDPD: delt not set or invalid");
    const float L[3] = {d.lx, d.ly, d.lz};
    long ncell = 1;
    for (int j = 0; j < 3; ++j) {
        // PS_NeighborList::initializeNList (ps_neighborList.cu:150-175)
        int n = (int)(L[j] / Pt.Rp);
        if (n < 1) n = 1;
        MFT_REQUIRE(n >= 3, MFT_ERR_UNSUPPORTED,
                    "dpd: rcut %g must be at most a third of every box length (fewer cells repeat pairs)",
                    (double)Pt.Rp);
        P.nc[j] = n;
        P.cw[j] = L[j] / (float)n;
        ncell *= n;
    }
    MFT_REQUIRE(ncell <= (1L << 26), MFT_ERR_UNSUPPORTED, "dpd: %ld cells", ncell);
    P.on = 1;
    P.mask = Pt.maskI;
    P.gamma = Pt.A;
    P.sigma = sqrtf(2.0f * Pt.A * Pt.sigma);        // ps_potentialDPD.cu:12
    P.rcut = Pt.Rp;
    P.rcut2 = Pt.Rp * Pt.Rp;
    P.inv_sqrt_delt = 1.0f / sqrtf(Pt.wall_pos);    // ps_potentialDPD.cu:60
    int rc;
    if ((rc = dalloc(&e->dcount, (size_t)ncell))) return rc;
    if ((rc = dalloc(&e->dstart, (size_t)ncell + 1))) return rc;
    if ((rc = dalloc(&e->dcursor, (size_t)ncell))) return rc;
    if ((rc = dalloc(&e->dcell, (size_t)e->N))) return rc;
    if ((rc = dalloc(&e->dslots, (size_t)e->N))) return rc;
    if ((rc = dalloc(&e->fdpd, (size_t)e->N))) return rc;
    return MFT_OK;
}

int tild_dpd_forces(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    const DpdParams& P = e->dpd;
    const int c = e->cur, N = e->N;
    const int ncell = P.nc[0] * P.nc[1] * P.nc[2];
    MFT_HIP(hipMemsetAsync(e->dcount, 0, sizeof(int) * (size_t)ncell, e->stream));
    k_dpd_count<<<nblk(N), kBlock, 0, e->stream>>>(e->X[c], e->META[c], P, N, e->dcount, e->dcell);
    k_dpd_scan<<<1, 1024, 0, e->stream>>>(e->dcount, e->dstart, e->dcursor, ncell);
    k_dpd_fill<<<nblk(N), kBlock, 0, e->stream>>>(e->dcell, N, e->dcursor, e->dslots);
    // under VV these are the forces of the previous step: its velocities and noise
    const bool vv = e->is_vv;
    const uint32_t step = (uint32_t)e->step - ((vv && e->vv_state != 0) ? 1u : 0u);
    k_dpd_force<<<nblk(N), kBlock, 0, e->stream>>>(e->X[c], e->META[c], vv ? e->NZ[c] : nullptr, e->dcell,
                                                    e->dstart, e->dslots, P, d.lx, d.ly, d.lz, d.seed, step, N,
                                                    e->fdpd);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

} // namespace mft
