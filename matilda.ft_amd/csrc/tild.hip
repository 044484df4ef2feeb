// TILD particle-mesh step on gfx950.
//
// One step (PS_Box::doTimeStep, PS_Box.cu:73-207) runs as
//   bin (cell counting sort, fused into the previous step's integrator)
//   -> scan -> scatter into cell order
//   -> brick density: particles of a brick's halo region staged through LDS,
//      weights recomputed, int64 fixed-point LDS accumulation (deterministic,
//      no global atomics), one float per point and channel written
//   -> FFT x, FFT y (all density channels in one launch each)
//   -> fused z pass: forward z FFT, every potential applied in k-space as one
//      per-species kernel sum_c K_tc(k) rho_c(k) times -i k_j, inverse z FFT
//   -> inverse FFT y, inverse FFT x keeping the real part
//   -> gather in cell order (force on each particle from its species field and
//      q times the electric field)
//   -> bonds + GJF/EM per particle, writing the next positions and counting
//      the next step's cells.
// The reference instead materialises (p+1)^3 index/weight tables per particle,
// scatters with FP32 atomics per group, and runs 4 or 8 FFTs per potential.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "fft_core.h"
#include "spline.h"

namespace mft {

const float2* fft_twiddles_device(int n); // fft.hip

constexpr int kMaxSpecies = 16;
constexpr int kMaxPots = 32;
constexpr int kMaxBondTypes = 16;

struct KPot {
    int kind;
    uint32_t maskI, maskJ;
    float A;      // current amplitude (ramped)
    float sig2;
    float Rp;
    const float* table; // u(k)/A for erf kinds, else null
};

struct KParams {
    int npots;
    KPot pot[kMaxPots];
    int charge_pot;     // index of the charges potential or -1
    int nchan;          // species channels (+1 for charge density)
    int qchan;          // channel index of the charge density or -1
    int nforce;         // species receiving grid forces
    int force_species[kMaxSpecies];
    int fold;           // 1: every species has one charge; rho_q = sum_s q_s rho_s
    float qspec[kMaxSpecies];
    int energy;         // accumulate potential energies
};

struct BondParams {
    int ntypes;
    float k[kMaxBondTypes];
    float r0[kMaxBondTypes];
    int style[kMaxBondTypes];
};

struct GjfParams {
    float a[kMaxSpecies], b[kMaxSpecies], nmag[kMaxSpecies], cf[kMaxSpecies], cn[kMaxSpecies];
};

// ------------------------------------------------- grid weights (T3) ----
__global__ void k_grid_weights(float* __restrict__ W, int* __restrict__ inds, const float* __restrict__ x,
                               int ns, int nx, int ny, int nz, float dx0, float dx1, float dx2, int p)
{
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ns) return;
    float w[3][5];
    int g[3];
    float dxs[3] = {dx0, dx1, dx2};
    int n[3] = {nx, ny, nz};
    for (int j = 0; j < 3; ++j) g[j] = base_index_weights(x[id * 3 + j], dxs[j], p, w[j]);
    // the device kernel divides by the float product of the float spacings
    float gvol = dx0 * dx1 * dx2;
    int shift = p / 2;
    int G = (p + 1) * (p + 1) * (p + 1);
    int ct = 0;
    for (int ix = 0; ix <= p; ++ix) {
        int n0 = (g[0] + ix - shift + n[0]) % n[0];
        for (int iy = 0; iy <= p; ++iy) {
            int n1 = (g[1] + iy - shift + n[1]) % n[1];
            for (int iz = 0; iz <= p; ++iz) {
                int n2 = (g[2] + iz - shift + n[2]) % n[2];
                long o = (long)id * G + ct;
                inds[o] = n0 + (n1 + n2 * n[1]) * n[0];
                W[o] = w[0][ix] * w[1][iy] * w[2][iz] / gvol;
                ++ct;
            }
        }
    }
}

// Deterministic table-driven density (T4 drop-in): one thread per grid point
// would need an inverse map, so this entry point accumulates in int64 fixed
// point with global atomics; integer sums do not depend on the order.
__global__ void k_fill_density_tables(unsigned long long* __restrict__ acc, unsigned long long* __restrict__ accq,
                                      const float* __restrict__ q, const int* __restrict__ sites, int nsites,
                                      const int* __restrict__ inds, const float* __restrict__ W, int G, int M,
                                      int ns)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsites) return;
    int p = sites[s];
    if (p < 0 || p >= ns) return;
    float qp = q ? q[p] : 0.f;
    for (int i = 0; i < G; ++i) {
        long o = (long)p * G + i;
        int gi = inds[o];
        if (gi < 0 || gi >= M) continue;
        float w = W[o];
        atomicAdd(&acc[gi], (unsigned long long)__float2ll_rn(w * 4294967296.0f));
        if (accq) atomicAdd(&accq[gi], (unsigned long long)__float2ll_rn(w * qp * 4294967296.0f));
    }
}

__global__ void k_fixed_to_float_add(float* __restrict__ out, const unsigned long long* __restrict__ acc, int M)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    out[i] += (float)((double)(long long)acc[i] * (1.0 / 4294967296.0));
}

__global__ void k_map_forces_tables(float* __restrict__ f, const float* __restrict__ gridF,
                                    const float* __restrict__ q, const int* __restrict__ sites, int nsites,
                                    const int* __restrict__ inds, const float* __restrict__ W, int G, float gvol,
                                    int M, int ns)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsites) return;
    int p = sites[s];
    if (p < 0 || p >= ns) return;
    for (int j = 0; j < 3; ++j) {
        float acc = 0.f;
        for (int i = 0; i < G; ++i) {
            long o = (long)p * G + i;
            int gi = inds[o];
            if (gi < 0 || gi >= M) continue;
            acc += gridF[(long)gi * 3 + j] * W[o] * gvol;
        }
        f[p * 3 + j] += q ? acc * q[p] : acc;
    }
}

// ----------------------------------------------------------- bonds ------
__device__ __forceinline__ float min_image(float d, float L, float half)
{
    if (d > half) d -= L;
    else if (d < -half) d += L;
    return d;
}

// d_bonds (ps_deviceBonds.cu:7-79) for one particle.
__device__ __forceinline__ void bond_force(int i, int ns, const float* __restrict__ x,
                                           const int* __restrict__ nBonds, const int* __restrict__ bondedTo,
                                           const int* __restrict__ bondType, int maxbonds, const BondParams& bp, float L0, float L1, float L2,
                                           float x0, float x1, float x2, float* fo)
{
    int nb = nBonds[i];
    if (nb > maxbonds) nb = maxbonds;
    for (int b = 0; b < nb; ++b) {
        int id2 = bondedTo[i * maxbonds + b];
        int bt = bondType[i * maxbonds + b];
        if (id2 < 0 || id2 >= ns || bt < 0 || bt >= bp.ntypes) continue;
        float d0 = min_image(x0 - x[id2 * 3 + 0], L0, 0.5f * L0);
        float d1 = min_image(x1 - x[id2 * 3 + 1], L1, 0.5f * L1);
        float d2 = min_image(x2 - x[id2 * 3 + 2], L2, 0.5f * L2);
        float r2 = d0 * d0 + d1 * d1 + d2 * d2;
        float mf = 0.f;
        if (r2 > 1.0e-4f) {
            float r = sqrtf(r2);
            if (bp.style[bt] == 0) mf = 2.0f * bp.k[bt] * (r - bp.r0[bt]) / r;
            else if (bp.style[bt] == 1) {
                float u = r / bp.r0[bt];
                mf = 2.0f * bp.k[bt] * r / (1.0f - u * u) / r;
            }
        }
        fo[0] -= mf * d0;
        fo[1] -= mf * d1;
        fo[2] -= mf * d2;
    }
}

__global__ void k_bonds(float* __restrict__ f, const int* __restrict__ nBonds, const int* __restrict__ bondedTo,
                        const int* __restrict__ bondType, BondParams bp, const float* __restrict__ x, float L0,
                        float L1, float L2, int ns, int maxbonds)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    float fo[3] = {0.f, 0.f, 0.f};
    bond_force(i, ns, x, nBonds, bondedTo, bondType, maxbonds, bp, L0, L1, L2, x[i * 3], x[i * 3 + 1], x[i * 3 + 2],
               fo);
    f[i * 3 + 0] += fo[0];
    f[i * 3 + 1] += fo[1];
    f[i * 3 + 2] += fo[2];
}

// d_bondStressEnergy energy part (ps_deviceBonds.cu:85-171), FP64 block sums.
__global__ void k_bond_energy(double* __restrict__ partial, const int* __restrict__ nBonds,
                              const int* __restrict__ bondedTo, const int* __restrict__ bondType, BondParams bp,
                              const float* __restrict__ x, float L0, float L1, float L2, int ns, int maxbonds)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (i < ns) {
        float Ls[3] = {L0, L1, L2};
        for (int b = 0; b < nBonds[i]; ++b) {
            int id2 = bondedTo[i * maxbonds + b];
            int bt = bondType[i * maxbonds + b];
            float r2 = 0.f;
            for (int j = 0; j < 3; ++j) {
                // 0.5 * Lh as in ps_deviceBonds.cu:132-133
                float d = min_image(x[i * 3 + j] - x[id2 * 3 + j], Ls[j], 0.25f * Ls[j]);
                r2 += d * d;
            }
            if (r2 > 1.0e-4f) {
                float r = sqrtf(r2);
                if (bp.style[bt] == 0) {
                    float a = r - bp.r0[bt];
                    e += (double)(bp.k[bt] * a * a);
                } else if (bp.style[bt] == 1 && r / bp.r0[bt] < 1.f) {
                    float u = r / bp.r0[bt];
                    float arg = 1.f / (1.f - u * u);
                    e += (double)(2.0f * bp.k[bt] * arg * r);
                }
            }
        }
    }
    __shared__ double red[4];
    e = wave_sum(e);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_sum_partials(double* __restrict__ out, const double* __restrict__ partial, int n)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
    __shared__ double red[16];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += red[w];
        *out = t;
    }
}

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kScanThreads * kScanItems;

__device__ __forceinline__ int block_exclusive_scan(int v, int* sh, int* total)
{
    // wave inclusive scan then across waves
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) sh[wid] = inc;
    __syncthreads();
    int woff = 0, tot = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        if (w < wid) woff += sh[w];
        tot += sh[w];
    }
    __syncthreads();
    *total = tot;
    return woff + inc - v;
}

// ------------------------------------------------------------ binning ---
__device__ __forceinline__ int cell_of(float x0, float x1, float x2, float dx0, float dx1, float dx2, int p,
                                       int nx, int ny, int nz)
{
    int cx = wrap(base_index(x0, dx0, p), nx);
    int cy = wrap(base_index(x1, dx1, p), ny);
    int cz = wrap(base_index(x2, dx2, p), nz);
    return cx + nx * (cy + ny * cz);
}

__global__ void k_bin_count(const float* __restrict__ x, int ns, int* __restrict__ count, int* __restrict__ cell,
                            int* __restrict__ rank, float dx0, float dx1, float dx2, int p, int nx, int ny, int nz)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    int c = cell_of(x[i * 3], x[i * 3 + 1], x[i * 3 + 2], dx0, dx1, dx2, p, nx, ny, nz);
    cell[i] = c;
    rank[i] = atomicAdd(&count[c], 1);
}

__global__ void k_scan_reduce(const int* __restrict__ count, int M, int* __restrict__ partial)
{
    __shared__ int sh[16];
    long base = (long)blockIdx.x * kScanChunk;
    int s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        long i = base + (long)k * kScanThreads + threadIdx.x;
        if (i < M) s += count[i];
    }
    int tot;
    block_exclusive_scan(s, sh, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void k_scan_partials(int* __restrict__ partial, int n)
{
    __shared__ int sh[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += blockDim.x) {
        int i = base + threadIdx.x;
        int v = i < n ? partial[i] : 0;
        int tot;
        int ex = block_exclusive_scan(v, sh, &tot);
        if (i < n) partial[i] = ex + carry;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
}

// start[i] = exclusive prefix of count; count cleared for the next binning.
__global__ void k_scan_apply(int* __restrict__ count, int M, const int* __restrict__ partial,
                             int* __restrict__ start, int ns)
{
    __shared__ int sh[16];
    long base = (long)blockIdx.x * kScanChunk + (long)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        long i = base + k;
        v[k] = i < M ? count[i] : 0;
        s += v[k];
    }
    int tot;
    int ex = block_exclusive_scan(s, sh, &tot) + partial[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        long i = base + k;
        if (i < M) {
            start[i] = ex;
            count[i] = 0;
        }
        ex += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) start[M] = ns;
}

__global__ void k_bin_scatter(const float* __restrict__ x, const float* __restrict__ q, const int* __restrict__ type,
                              int ns, const int* __restrict__ cell, const int* __restrict__ rank,
                              const int* __restrict__ start, float4* __restrict__ spos, int2* __restrict__ smeta)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    int slot = start[cell[i]] + rank[i];
    spos[slot] = make_float4(x[i * 3], x[i * 3 + 1], x[i * 3 + 2], q ? q[i] : 0.f);
    smeta[slot] = make_int2(i, type[i]);
}

// ------------------------------------------------------- brick density ---
struct BrickGeom {
    int bx, by, bz;      // brick extent in grid points (powers of two)
    int sx, sy, sz;      // log2 of the extents
    int nbx, nby, nbz;   // bricks per axis
};

// Contiguous runs of cells along one axis covering the cells whose stencil
// reaches [g0, g0 + B): [g0 + shift - p, g0 + B - 1 + shift], wrapped, or the
// whole axis when that window is at least as long as the axis.
__device__ __forceinline__ int axis_window(int g0, int B, int p, int n, int* lo, int* len)
{
    int shift = p / 2;
    int w = B + p;
    if (w >= n) {
        lo[0] = 0;
        len[0] = n;
        return 1;
    }
    int a = wrap(g0 + shift - p, n);
    if (a + w <= n) {
        lo[0] = a;
        len[0] = w;
        return 1;
    }
    lo[0] = a;
    len[0] = n - a;
    lo[1] = 0;
    len[1] = w - (n - a);
    return 2;
}

constexpr int kDensThreads = 256;
constexpr int kMaxSegs = 2 * 40 * 40;

template <int P>
__global__ void __launch_bounds__(kDensThreads) k_density_brick(
    float* __restrict__ rho, int nchan, int qchan, const float4* __restrict__ spos, const int2* __restrict__ smeta,
    const int* __restrict__ start, int nx, int ny, int nz, float dx0, float dx1, float dx2, BrickGeom bg,
    float inv_gvol)
{
    constexpr int p = P;
    extern __shared__ unsigned long long acc[]; // [nchan][bx*by*bz]
    __shared__ int seg_begin[512];
    __shared__ int seg_pref[513];

    const int bid = blockIdx.x;
    const int ibx = bid % bg.nbx, iby = (bid / bg.nbx) % bg.nby, ibz = bid / (bg.nbx * bg.nby);
    const int gx0 = ibx * bg.bx, gy0 = iby * bg.by, gz0 = ibz * bg.bz;
    const int npts = bg.bx * bg.by * bg.bz;
    const float idx0 = 1.0f / dx0, idx1 = 1.0f / dx1, idx2 = 1.0f / dx2;

    for (int i = threadIdx.x; i < nchan * npts; i += blockDim.x) acc[i] = 0ull;

    // cell windows per axis
    int xlo[2], xlen[2], ylo[2], ylen[2], zlo[2], zlen[2];
    int nxs = axis_window(gx0, bg.bx, p, nx, xlo, xlen);
    int nys = axis_window(gy0, bg.by, p, ny, ylo, ylen);
    int nzs = axis_window(gz0, bg.bz, p, nz, zlo, zlen);
    int nyc = ylen[0] + (nys > 1 ? ylen[1] : 0);
    int nzc = zlen[0] + (nzs > 1 ? zlen[1] : 0);
    int nrows = nyc * nzc;

    // one segment per (row, x run): particles of cells [xlo, xlo+xlen) in row (cy, cz)
    __syncthreads();
    for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
        int iy = r % nyc, iz = r / nyc;
        int cy = iy < ylen[0] ? ylo[0] + iy : ylo[1] + (iy - ylen[0]);
        int cz = iz < zlen[0] ? zlo[0] + iz : zlo[1] + (iz - zlen[0]);
        for (int s = 0; s < nxs; ++s) {
            int c0 = xlo[s] + nx * (cy + ny * cz);
            int b = start[c0], e = start[c0 + xlen[s]];
            int k = r * nxs + s;
            seg_begin[k] = b;
            seg_pref[k] = e - b; // lengths for now
        }
    }
    __syncthreads();
    const int nseg = nrows * nxs;
    {
        // exclusive prefix of the segment lengths, two entries per thread
        __shared__ int sh[16];
        int k0 = 2 * threadIdx.x, k1 = k0 + 1;
        int l0 = k0 < nseg ? seg_pref[k0] : 0, l1 = k1 < nseg ? seg_pref[k1] : 0;
        int tot;
        int ex = block_exclusive_scan(l0 + l1, sh, &tot);
        __syncthreads();
        if (k0 < nseg) seg_pref[k0] = ex;
        if (k1 < nseg) seg_pref[k1] = ex + l0;
        if (threadIdx.x == 0) seg_pref[nseg] = tot;
    }
    __syncthreads();
    const int total = seg_pref[nseg];
    const int shift = p / 2;

    for (int it = threadIdx.x; it < total; it += blockDim.x) {
        // segment holding flattened index it
        int lo = 0, hi = nseg - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (seg_pref[mid] <= it) lo = mid;
            else hi = mid - 1;
        }
        int slot = seg_begin[lo] + (it - seg_pref[lo]);
        float4 pp = spos[slot];
        int typ = smeta[slot].y;
        float wx[5], wy[5], wz[5];
        const int g0 = base_index_weights_fast(pp.x, dx0, idx0, p, wx);
        const int g1 = base_index_weights_fast(pp.y, dx1, idx1, p, wy);
        const int g2 = base_index_weights_fast(pp.z, dx2, idx2, p, wz);
        // brick-local coordinate of the first stencil point, wrapped once
        const int bx = wrap(g0 - shift - gx0, nx), by = wrap(g1 - shift - gy0, ny), bz = wrap(g2 - shift - gz0, nz);
        const bool hasq = qchan >= 0 && pp.w != 0.f;
#pragma unroll
        for (int iz = 0; iz <= p; ++iz) {
            int lz = bz + iz;
            if (lz >= nz) lz -= nz;
            if (lz >= bg.bz) continue;
#pragma unroll
            for (int iy = 0; iy <= p; ++iy) {
                int ly = by + iy;
                if (ly >= ny) ly -= ny;
                if (ly >= bg.by) continue;
                float wyz = wy[iy] * wz[iz];
                const int rowi = ly + (lz << bg.sy);
#pragma unroll
                for (int ix = 0; ix <= p; ++ix) {
                    int lx = bx + ix;
                    if (lx >= nx) lx -= nx;
                    if (lx >= bg.bx) continue;
                    float w = wx[ix] * wyz;
                    int li = lx + (rowi << bg.sx);
                    atomicAdd(&acc[typ * npts + li], (unsigned long long)__float2ll_rn(w * 4294967296.0f));
                    if (hasq)
                        atomicAdd(&acc[qchan * npts + li],
                                  (unsigned long long)__float2ll_rn(w * pp.w * 4294967296.0f));
                }
            }
        }
    }
    __syncthreads();
    const long M = (long)nx * ny * nz;
    for (int i = threadIdx.x; i < nchan * npts; i += blockDim.x) {
        int c = i >> (bg.sx + bg.sy + bg.sz), li = i & (npts - 1);
        int lx = li & (bg.bx - 1), ly = (li >> bg.sx) & (bg.by - 1), lz = li >> (bg.sx + bg.sy);
        int gx = gx0 + lx, gy = gy0 + ly, gz = gz0 + lz;
        if (gx < nx && gy < ny && gz < nz) {
            double v = (double)(long long)acc[i] * (1.0 / 4294967296.0);
            rho[c * M + gx + (long)nx * (gy + (long)ny * gz)] = (float)(v * (double)inv_gvol);
        }
    }
}

// --------------------------------------------------------- FFT passes ---
// Line transform of nfield grids along one axis.  Input real (float) or
// complex, output complex or real part; scale applied on store.
template <int N, int SIGN, bool REAL_IN, bool REAL_OUT>
__global__ void __launch_bounds__(256) k_fft_axis(const void* __restrict__ in, void* __restrict__ out, AxisGeom geom,
                                                  long field_stride, const float2* __restrict__ twg, float scale)
{
    using P = FftPlan<N>;
    constexpr int T = P::T, E = P::E, B = 256 / T;
    __shared__ float2 lds[B * N];
    __shared__ float2 tw[N];
    for (int i = threadIdx.x; i < N; i += 256) tw[i] = twg[i];
    const bool inter = geom.axis != 0;
    const int lline = inter ? (threadIdx.x % B) : (threadIdx.x / T);
    const int t = inter ? (threadIdx.x / B) : (threadIdx.x % T);
    const int line = blockIdx.x * B + lline;
    const bool valid = line < geom.nlines();
    const int field = blockIdx.y;
    int ta, tb;
    const long base = valid ? geom.base(line, &ta, &tb) : 0;
    const long st = geom.stride();
    float2 v[E];
    if (REAL_IN) {
        const float* src = (const float*)in + field * field_stride + base;
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = valid ? make_float2(src[(t + T * s) * st], 0.f) : make_float2(0.f, 0.f);
    } else {
        const float2* src = (const float2*)in + field * field_stride + base;
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = valid ? src[(t + T * s) * st] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    LdsMap map{lline, B, inter, N};
    line_fft<N, SIGN>(v, lds, tw, t, map);
    if (!valid) return;
    if (REAL_OUT) {
        float* dst = (float*)out + field * field_stride + base;
#pragma unroll
        for (int s = 0; s < E; ++s) dst[(t + T * s) * st] = v[s].x * scale;
    } else {
        float2* dst = (float2*)out + field * field_stride + base;
#pragma unroll
        for (int s = 0; s < E; ++s) dst[(t + T * s) * st] = cscale(v[s], scale);
    }
}

// Potential shape at |k|^2 = k2 for grid point idx (float, reference forms).
__device__ __forceinline__ float pot_shape(const KPot& P, float k2, long idx)
{
    if (P.kind == MFT_POT_GAUSSIAN) return P.A * expf(-k2 * P.sig2 * 0.5f);
    return P.A * P.table[idx];
}

constexpr int kMaxChan = 8;

// Fused z pass: forward z FFT of every density channel (already transformed
// along x and y), k-space forces, inverse z FFT of each output component.
template <int N, int NCH>
__global__ void __launch_bounds__(256) k_conv_z(const float2* __restrict__ S, float2* __restrict__ Y, int nx, int ny,
                                                float Lx, float Ly, float Lz, KParams kp, const float2* __restrict__ twg,
                                                float invM, double* __restrict__ epart, float Mgvol)
{
    using P = FftPlan<N>;
    constexpr int T = P::T, E = P::E, B = 256 / T;
    __shared__ float2 lds[B * N];
    __shared__ float2 tw[N];
    __shared__ double ered[4 * (kMaxPots + 1)];
    for (int i = threadIdx.x; i < N; i += 256) tw[i] = twg[i];
    const int lline = threadIdx.x % B;
    const int t = threadIdx.x / B;
    const int line = blockIdx.x * B + lline;
    const long M = (long)nx * ny * N;
    const bool valid = line < nx * ny;
    const int ix = line % nx, iy = line / nx;
    const long st = (long)nx * ny;
    LdsMap map{lline, B, true, N};

    float2 spec[NCH][E];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        if (c >= kp.nchan) {
#pragma unroll
            for (int s = 0; s < E; ++s) spec[c][s] = make_float2(0.f, 0.f);
            continue;
        }
#pragma unroll
        for (int s = 0; s < E; ++s)
            spec[c][s] = valid ? S[c * M + line + (long)(t + T * s) * st] : make_float2(0.f, 0.f);
        line_fft<N, -1>(spec[c], lds, tw, t, map);
#pragma unroll
        for (int s = 0; s < E; ++s) spec[c][s] = cscale(spec[c][s], invM);
    }

    const float kx = kcomp(ix, nx, Lx), ky = kcomp(iy, ny, Ly);
    float kz[E], k2[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        kz[s] = kcomp(t + T * s, N, Lz);
        k2[s] = __fadd_rn(__fadd_rn(__fmul_rn(kx, kx), __fmul_rn(ky, ky)), __fmul_rn(kz[s], kz[s]));
    }

    int out = 0;
    for (int fi = 0; fi < kp.nforce + (kp.qchan >= 0 ? 1 : 0); ++fi) {
        const bool charge = fi == kp.nforce;
        float2 Z[E];
#pragma unroll
        for (int s = 0; s < E; ++s) Z[s] = make_float2(0.f, 0.f);
        if (charge) {
            // separate charge field (charges not uniform per species)
            const KPot& Pq = kp.pot[kp.charge_pot];
#pragma unroll
            for (int s = 0; s < E; ++s) {
                float u = k2[s] > 0.f ? 12.566370614359172f * Pq.A / k2[s] * expf(-k2[s] * Pq.sig2) : 0.f;
                float2 sq = make_float2(0.f, 0.f);
#pragma unroll
                for (int c = 0; c < NCH; ++c)
                    if (c == kp.qchan) sq = spec[c][s];
                Z[s] = cscale(sq, u);
            }
        } else {
            const int sp = kp.force_species[fi];
            for (int pi = 0; pi < kp.npots; ++pi) {
                const KPot& Pp = kp.pot[pi];
                if (Pp.kind == MFT_POT_CHARGES) {
                    // folded electrostatics: q_s u_q(k) sum_c q_c rho_c(k)
                    if (!kp.fold || kp.qspec[sp] == 0.f) continue;
#pragma unroll
                    for (int s = 0; s < E; ++s) {
                        float2 src = make_float2(0.f, 0.f);
#pragma unroll
                        for (int c = 0; c < NCH; ++c) {
                            if (c >= kp.nchan) continue;
                            src.x += kp.qspec[c] * spec[c][s].x;
                            src.y += kp.qspec[c] * spec[c][s].y;
                        }
                        float u = k2[s] > 0.f
                                      ? kp.qspec[sp] * 12.566370614359172f * Pp.A / k2[s] * expf(-k2[s] * Pp.sig2)
                                      : 0.f;
                        Z[s].x += u * src.x;
                        Z[s].y += u * src.y;
                    }
                    continue;
                }
                const bool inI = (Pp.maskI >> sp) & 1u, inJ = (Pp.maskJ >> sp) & 1u;
                if (!inI && !inJ) continue;
#pragma unroll
                for (int s = 0; s < E; ++s) {
                    float2 src = make_float2(0.f, 0.f);
#pragma unroll
                    for (int c = 0; c < NCH; ++c) {
                        if (c >= kp.nchan || c == kp.qchan) continue;
                        float m = (inI ? (float)((Pp.maskJ >> c) & 1u) : 0.f) +
                                  (inJ ? (float)((Pp.maskI >> c) & 1u) : 0.f);
                        src.x += m * spec[c][s].x;
                        src.y += m * spec[c][s].y;
                    }
                    long kidx = ix + (long)nx * (iy + (long)ny * (t + T * s));
                    float u = pot_shape(Pp, k2[s], kidx);
                    Z[s].x += u * src.x;
                    Z[s].y += u * src.y;
                }
            }
        }
        for (int j = 0; j < 3; ++j) {
            float2 v[E];
#pragma unroll
            for (int s = 0; s < E; ++s) {
                float kj = j == 0 ? kx : (j == 1 ? ky : kz[s]);
                // -i k_j Z
                v[s] = make_float2(kj * Z[s].y, -kj * Z[s].x);
            }
            line_fft<N, +1>(v, lds, tw, t, map);
            if (valid) {
#pragma unroll
                for (int s = 0; s < E; ++s) Y[out * M + line + (long)(t + T * s) * st] = v[s];
            }
            ++out;
        }
    }

    if (kp.energy) {
        // E_pot = gvol M sum_k Re[conj(rho_I) u rho_J]  (Parseval form of
        // ps_potential.cu:133-178); charges carry the 1/2 of
        // ps_potentialCharges.cu:115.
        for (int pi = 0; pi < kp.npots; ++pi) {
            const KPot& Pp = kp.pot[pi];
            double e = 0.0;
            if (valid) {
#pragma unroll
                for (int s = 0; s < E; ++s) {
                    float2 a = make_float2(0.f, 0.f), b = make_float2(0.f, 0.f);
                    float u;
                    if (Pp.kind == MFT_POT_CHARGES) {
#pragma unroll
                        for (int c = 0; c < NCH; ++c) {
                            if (c >= kp.nchan) continue;
                            if (kp.fold) {
                                a.x += kp.qspec[c] * spec[c][s].x;
                                a.y += kp.qspec[c] * spec[c][s].y;
                            } else if (c == kp.qchan) {
                                a = spec[c][s];
                            }
                        }
                        b = a;
                        u = k2[s] > 0.f ? 0.5f * 12.566370614359172f * Pp.A / k2[s] * expf(-k2[s] * Pp.sig2) : 0.f;
                    } else {
#pragma unroll
                        for (int c = 0; c < NCH; ++c) {
                            if (c >= kp.nchan || c == kp.qchan) continue;
                            if ((Pp.maskI >> c) & 1u) { a.x += spec[c][s].x; a.y += spec[c][s].y; }
                            if ((Pp.maskJ >> c) & 1u) { b.x += spec[c][s].x; b.y += spec[c][s].y; }
                        }
                        long kidx = ix + (long)nx * (iy + (long)ny * (t + T * s));
                        u = pot_shape(Pp, k2[s], kidx);
                    }
                    e += (double)u * ((double)a.x * b.x + (double)a.y * b.y);
                }
            }
            e = wave_sum(e);
            if ((threadIdx.x & 63) == 0) ered[(threadIdx.x >> 6) * (kMaxPots + 1) + pi] = e;
        }
        __syncthreads();
        if (threadIdx.x < kp.npots) {
            double s = 0.0;
            for (int w = 0; w < 4; ++w) s += ered[w * (kMaxPots + 1) + threadIdx.x];
            epart[(long)blockIdx.x * kMaxPots + threadIdx.x] = s * (double)Mgvol;
        }
    }
}

__global__ void k_energy_sum(double* __restrict__ out, const double* __restrict__ part, int nblocks, int npots)
{
    int p = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += part[(long)b * kMaxPots + p];
    __shared__ double red[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[p] = red[0] + red[1] + red[2] + red[3];
}

// ------------------------------------------------------------- gather ---
// Force on each particle from its species' total grid force (potentials and,
// with folded charges, q_s times the field) and, when charges are not uniform
// per species, q times the separate charge field (d_mapGridForcesToPartics and
// d_mapGridChargeForcesToPartics, device_utils.cu:358-461), in cell order.
template <int P>
__global__ void __launch_bounds__(256) k_gather(float* __restrict__ fgrid, const float4* __restrict__ F4,
                                                const int* __restrict__ force_slot, int qslot,
                                                const float4* __restrict__ spos, const int2* __restrict__ smeta,
                                                int ns, int nx, int ny, int nz, float dx0, float dx1, float dx2)
{
    constexpr int p = P;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    float4 pp = spos[i];
    int2 meta = smeta[i];
    const long M = (long)nx * ny * nz;
    int fs = force_slot[meta.y];
    bool charged = qslot >= 0 && pp.w != 0.f;
    float f[3] = {0.f, 0.f, 0.f};
    if (fs >= 0 || charged) {
        float wx[5], wy[5], wz[5];
        int g0 = base_index_weights_fast(pp.x, dx0, 1.0f / dx0, p, wx);
        int g1 = base_index_weights_fast(pp.y, dx1, 1.0f / dx1, p, wy);
        int g2 = base_index_weights_fast(pp.z, dx2, 1.0f / dx2, p, wz);
        const int shift = p / 2;
        const float4* Fs = F4 + (fs >= 0 ? fs : 0) * M;
        const float4* Fq = F4 + (qslot >= 0 ? qslot : 0) * M;
        float fq[3] = {0.f, 0.f, 0.f};
        int xs[p + 1], ys[p + 1], zs[p + 1];
        {
            int bx = wrap(g0 - shift, nx), by = wrap(g1 - shift, ny), bz = wrap(g2 - shift, nz);
#pragma unroll
            for (int i = 0; i <= p; ++i) {
                xs[i] = bx + i >= nx ? bx + i - nx : bx + i;
                ys[i] = by + i >= ny ? by + i - ny : by + i;
                zs[i] = bz + i >= nz ? bz + i - nz : bz + i;
            }
        }
#pragma unroll
        for (int iz = 0; iz <= p; ++iz) {
#pragma unroll
            for (int iy = 0; iy <= p; ++iy) {
                float wyz = wy[iy] * wz[iz];
                long row = (long)nx * (ys[iy] + (long)ny * zs[iz]);
#pragma unroll
                for (int ix = 0; ix <= p; ++ix) {
                    long g = row + xs[ix];
                    float w = wx[ix] * wyz;
                    if (fs >= 0) {
                        float4 v = Fs[g];
                        f[0] += v.x * w;
                        f[1] += v.y * w;
                        f[2] += v.z * w;
                    }
                    if (charged) {
                        float4 v = Fq[g];
                        fq[0] += v.x * w;
                        fq[1] += v.y * w;
                        fq[2] += v.z * w;
                    }
                }
            }
        }
        if (charged) {
            f[0] += fq[0] * pp.w;
            f[1] += fq[1] * pp.w;
            f[2] += fq[2] * pp.w;
        }
    }
    int pid = meta.x;
    fgrid[pid * 3 + 0] = f[0];
    fgrid[pid * 3 + 1] = f[1];
    fgrid[pid * 3 + 2] = f[2];
}

// Gather through LDS: a workgroup owns a brick of cells, stages the force
// fields on the brick's grid window (brick + p points per axis, periodic) in
// LDS once, then every particle whose base cell lies in the brick gathers its
// (p+1)^3 stencil from LDS.  Fields are staged in chunks that fit the LDS
// budget; each particle is handled in the chunk holding its field(s).
constexpr int kGatherThreads = 256;
constexpr int kGatherLdsFloats = 16384; // 64 KB

template <int P>
__global__ void __launch_bounds__(kGatherThreads) k_gather_brick(
    float* __restrict__ fgrid, const float4* __restrict__ F4, const int* __restrict__ force_slot, int qslot,
    int nout4, const float4* __restrict__ spos, const int2* __restrict__ smeta, const int* __restrict__ start, int nx,
    int ny, int nz, float dx0, float dx1, float dx2, BrickGeom bg, int fields_per_chunk)
{
    constexpr int p = P;
    extern __shared__ float tile[]; // [fields_per_chunk][3][window points]
    __shared__ int seg_begin[256];
    __shared__ int seg_pref[257];
    __shared__ long rowbase[256];   // grid offset of window row (ly, lz)
    __shared__ int xw[32];          // wrapped x of window column lx
    const int bid = blockIdx.x;
    const int ibx = bid % bg.nbx, iby = (bid / bg.nbx) % bg.nby, ibz = bid / (bg.nbx * bg.nby);
    const int cx0 = ibx * bg.bx, cy0 = iby * bg.by, cz0 = ibz * bg.bz;
    const int ex = min(bg.bx, nx - cx0), ey = min(bg.by, ny - cy0), ez = min(bg.bz, nz - cz0);
    const int shift = p / 2;
    // window of grid points: [c0 - shift, c0 + extent - 1 - shift + p]
    const int wx = ex + p, wy = ey + p, wz = ez + p;
    const int wpts = wx * wy * wz;
    const long M = (long)nx * ny * nz;
    const float idx0 = 1.0f / dx0, idx1 = 1.0f / dx1, idx2 = 1.0f / dx2;
    for (int r = threadIdx.x; r < wy * wz; r += blockDim.x) {
        int ly = r % wy, lz = r / wy;
        int gy = wrap(cy0 - shift + ly, ny), gz = wrap(cz0 - shift + lz, nz);
        rowbase[r] = (long)nx * (gy + (long)ny * gz);
    }
    if ((int)threadIdx.x < wx) xw[threadIdx.x] = wrap(cx0 - shift + (int)threadIdx.x, nx);

    // particles of the brick: one contiguous run per (cy, cz) row
    const int nrows = ey * ez;
    for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
        int cy = cy0 + r % ey, cz = cz0 + r / ey;
        int c0 = cx0 + nx * (cy + ny * cz);
        seg_begin[r] = start[c0];
        seg_pref[r] = start[c0 + ex] - start[c0];
    }
    __syncthreads();
    {
        __shared__ int sh[16];
        int l0 = (int)threadIdx.x < nrows ? seg_pref[threadIdx.x] : 0;
        int tot;
        int exs = block_exclusive_scan(l0, sh, &tot);
        __syncthreads();
        if ((int)threadIdx.x < nrows) seg_pref[threadIdx.x] = exs;
        if (threadIdx.x == 0) seg_pref[nrows] = tot;
    }
    __syncthreads();
    const int total = seg_pref[nrows];
    if (total == 0) return;

    for (int f0 = 0; f0 < nout4; f0 += fields_per_chunk) {
        const int nf = min(fields_per_chunk, nout4 - f0);
        __syncthreads();
        // stage fields f0 .. f0+nf-1 as [field][comp][point]
        // 16 lanes per window row (wx <= 12): lane column lx, row index step 16
        const int lx = threadIdx.x & 15;
        const int nrow = wy * wz;
        for (int f = 0; f < nf; ++f) {
            const float4* src = F4 + (long)(f0 + f) * M;
            float* tf = tile + f * 3 * wpts;
            if (lx < wx) {
                for (int r = threadIdx.x >> 4; r < nrow; r += blockDim.x >> 4) {
                    float4 v = src[rowbase[r] + xw[lx]];
                    int li = lx + wx * r;
                    tf[li] = v.x;
                    tf[wpts + li] = v.y;
                    tf[2 * wpts + li] = v.z;
                }
            }
        }
        __syncthreads();
        for (int it = threadIdx.x; it < total; it += blockDim.x) {
            int lo = 0, hi = nrows - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (seg_pref[mid] <= it) lo = mid;
                else hi = mid - 1;
            }
            int slot = seg_begin[lo] + (it - seg_pref[lo]);
            float4 pp = spos[slot];
            int2 meta = smeta[slot];
            int fs = force_slot[meta.y];
            bool charged = qslot >= 0 && pp.w != 0.f;
            // chunk indices of the particle's fields; the first of them writes,
            // a later one adds (chunks run in order inside this workgroup)
            const int k = f0 / fields_per_chunk;
            const int cf = fs >= 0 ? fs / fields_per_chunk : 0x7fffffff;
            const int cq = charged ? qslot / fields_per_chunk : 0x7fffffff;
            const int first = min(cf, cq);
            const bool hs = cf == k, hq = cq == k;
            const bool none = first == 0x7fffffff;
            if (!hs && !hq && !(none && k == 0)) continue;
            float a[3] = {0.f, 0.f, 0.f};
            if (hs || hq) {
                float w0[5], w1[5], w2[5];
                int g0 = base_index_weights_fast(pp.x, dx0, idx0, p, w0);
                int g1 = base_index_weights_fast(pp.y, dx1, idx1, p, w1);
                int g2 = base_index_weights_fast(pp.z, dx2, idx2, p, w2);
                // local window coordinates of the first stencil point
                int bx = wrap(g0 - cx0, nx), by = wrap(g1 - cy0, ny), bz = wrap(g2 - cz0, nz);
                float aq[3] = {0.f, 0.f, 0.f};
                const float* tfs = tile + (hs ? (fs - f0) : 0) * 3 * wpts;
                const float* tfq = tile + (hq ? (qslot - f0) : 0) * 3 * wpts;
#pragma unroll
                for (int iz = 0; iz <= p; ++iz) {
#pragma unroll
                    for (int iy = 0; iy <= p; ++iy) {
                        float wyz = w1[iy] * w2[iz];
                        int rowi = wx * ((by + iy) + wy * (bz + iz)) + bx;
#pragma unroll
                        for (int ix = 0; ix <= p; ++ix) {
                            float w = w0[ix] * wyz;
                            int li = rowi + ix;
                            if (hs) {
                                a[0] += tfs[li] * w;
                                a[1] += tfs[wpts + li] * w;
                                a[2] += tfs[2 * wpts + li] * w;
                            }
                            if (hq) {
                                aq[0] += tfq[li] * w;
                                aq[1] += tfq[wpts + li] * w;
                                aq[2] += tfq[2 * wpts + li] * w;
                            }
                        }
                    }
                }
                a[0] += aq[0] * pp.w;
                a[1] += aq[1] * pp.w;
                a[2] += aq[2] * pp.w;
            }
            const int pid = meta.x;
            if (none || k == first) {
                fgrid[pid * 3 + 0] = a[0];
                fgrid[pid * 3 + 1] = a[1];
                fgrid[pid * 3 + 2] = a[2];
            } else {
                fgrid[pid * 3 + 0] += a[0];
                fgrid[pid * 3 + 1] += a[1];
                fgrid[pid * 3 + 2] += a[2];
            }
        }
    }
}

// Last inverse pass along x for the three components of one force field,
// keeping real parts and writing the field as float4 (x, y, z, 0).
template <int N>
__global__ void __launch_bounds__(256) k_fft_x_inv_vec3(const float2* __restrict__ Y, float4* __restrict__ F4, int ny,
                                                        int nz, const float2* __restrict__ twg)
{
    using P = FftPlan<N>;
    constexpr int T = P::T, E = P::E, B = 256 / T;
    __shared__ float2 lds[B * N];
    __shared__ float2 tw[N];
    for (int i = threadIdx.x; i < N; i += 256) tw[i] = twg[i];
    const int lline = threadIdx.x / T;
    const int t = threadIdx.x % T;
    const int line = blockIdx.x * B + lline;
    const bool valid = line < ny * nz;
    const long M = (long)N * ny * nz;
    const long base = (long)line * N;
    const int f = blockIdx.y;
    float re[3][E];
    __syncthreads();
    LdsMap map{lline, B, false, N};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float2* src = Y + (3 * f + j) * M + base;
        float2 v[E];
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = valid ? src[t + T * s] : make_float2(0.f, 0.f);
        line_fft<N, +1>(v, lds, tw, t, map);
#pragma unroll
        for (int s = 0; s < E; ++s) re[j][s] = v[s].x;
    }
    if (!valid) return;
    float4* dst = F4 + f * M + base;
#pragma unroll
    for (int s = 0; s < E; ++s) dst[t + T * s] = make_float4(re[0][s], re[1][s], re[2][s], 0.f);
}

// ------------------------------------------------- bonds + integrator ---
// Bonds (ps_deviceBonds.cu:7-79) plus GJF (ps_integratorGJF.cu:102-163) or
// EM (ps_integratorEM.cu:39-82) on group "all".  Reads x (current), writes the
// next positions into xnext (the buffer that held x_old for GJF), counts the
// next step's cells.  integ = 2 only writes the total force.
__global__ void __launch_bounds__(256) k_bonds_integrate(
    const float* __restrict__ x, float* __restrict__ xnext, float* __restrict__ noise_old,
    const float* __restrict__ fgrid, float* __restrict__ fout, const int* __restrict__ type,
    const int* __restrict__ nBonds, const int* __restrict__ bondedTo, const int* __restrict__ bondType,
    int maxbonds, BondParams bp, GjfParams gp, int integ, float delt, float L0, float L1, float L2, int ns,
    uint64_t seed, uint32_t step, int* __restrict__ count, int* __restrict__ cell, int* __restrict__ rank,
    float dx0, float dx1, float dx2, int p, int nx, int ny, int nz)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    float xc[3] = {x[i * 3], x[i * 3 + 1], x[i * 3 + 2]};
    float fo[3] = {fgrid[i * 3], fgrid[i * 3 + 1], fgrid[i * 3 + 2]};
    bond_force(i, ns, x, nBonds, bondedTo, bondType, maxbonds, bp, L0, L1, L2, xc[0], xc[1], xc[2], fo);
    if (fout) {
        fout[i * 3 + 0] = fo[0];
        fout[i * 3 + 1] = fo[1];
        fout[i * 3 + 2] = fo[2];
    }
    if (integ == MFT_INT_NONE) return;
    float z[4];
    normal4((uint32_t)i, step, TAG_TILD, 0u, seed, z);
    const float Ls[3] = {L0, L1, L2};
    float xn[3];
    if (integ == MFT_INT_GJF) {
        int ty = type[i];
        float a = gp.a[ty], b = gp.b[ty], nm = gp.nmag[ty], cf = gp.cf[ty], cn = gp.cn[ty];
        for (int j = 0; j < 3; ++j) {
            float nn = nm * z[j];
            float xo = xnext[i * 3 + j];
            if (xo - xc[j] > Ls[j] / 2.0f) xo -= Ls[j];
            else if (xo - xc[j] < -Ls[j] / 2.0f) xo += Ls[j];
            float v = 2.0f * b * xc[j] - a * xo + cf * fo[j] + cn * (nn + noise_old[i * 3 + j]);
            if (v > Ls[j]) v -= Ls[j];
            else if (v < 0.0f) v += Ls[j];
            xn[j] = v;
            noise_old[i * 3 + j] = nn;
        }
    } else {
        float nm = sqrtf(2.0f * delt);
        for (int j = 0; j < 3; ++j) {
            float v = xc[j] + delt * fo[j] + nm * z[j];
            if (v > Ls[j]) v -= Ls[j];
            else if (v < 0.0f) v += Ls[j];
            xn[j] = v;
        }
    }
    xnext[i * 3 + 0] = xn[0];
    xnext[i * 3 + 1] = xn[1];
    xnext[i * 3 + 2] = xn[2];
    if (count) {
        int c = cell_of(xn[0], xn[1], xn[2], dx0, dx1, dx2, p, nx, ny, nz);
        cell[i] = c;
        rank[i] = atomicAdd(&count[c], 1);
    }
}

__global__ void k_copy_f(float* __restrict__ dst, const float* __restrict__ src, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ------------------------------------------------------------- engine ---
template <typename T>
static int dalloc(T** p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        set_error(MFT_ERR_OOM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
        return MFT_ERR_OOM;
    }
    return MFT_OK;
}

struct TildEngine {
    mft_tild_desc d;
    std::vector<mft_tild_potential> pots;
    std::vector<float> mass, mob;
    int N = 0, M = 0, nchan = 0, qchan = -1, nforce = 0, nout = 0, nout4 = 0, qslot = -1, charge_pot = -1;
    int configured = 0;
    float dx[3] = {0, 0, 0};
    float gvol = 0;
    hipStream_t stream = nullptr;
    // particle state
    float* x[2] = {nullptr, nullptr};
    int cur = 0;
    float *noise_old = nullptr, *fgrid = nullptr, *ftot = nullptr, *charge = nullptr;
    int *type = nullptr, *nBonds = nullptr, *bondedTo = nullptr, *bondType = nullptr;
    // binning
    int *count = nullptr, *cell = nullptr, *rank = nullptr, *start = nullptr, *partial = nullptr;
    int nscan_blocks = 0;
    float4* spos = nullptr;
    int2* smeta = nullptr;
    bool binned = false; // count/cell/rank hold the current positions
    // grids
    float* rho = nullptr;   // [nchan][M]
    float2* S = nullptr;    // [nchan][M]
    float2* Y = nullptr;    // [nout][M] force spectra (3 components per field)
    float4* F4 = nullptr;   // [nout4][M] force fields (x, y, z, 0)
    int* force_slot = nullptr; // [nspecies] -> force field slot or -1
    std::vector<int> force_slot_h;
    std::vector<float*> tables; // erf tables per potential (device)
    double* epart = nullptr;
    double* eout = nullptr;   // [kMaxPots + 1]
    double* bpart = nullptr;
    BondParams bp;
    GjfParams gp;
    KParams kp;
    BrickGeom bg;
    BrickGeom gbg;          // gather bricks
    int gather_fpc = 1;     // force fields staged per chunk
    int gather_mode = 0;    // 0 direct, 1 LDS-staged (mft_tild_autotune picks)
    size_t dens_lds = 0;
    int64_t step = 0;
    int have_state = 0;
    // timing: event pairs recorded around kernel launches on the engine
    // stream without host synchronisation, resolved when timing is read
    bool timing = false;
    std::vector<hipEvent_t> evpool;
    std::vector<int> evkind; // kernel id of pair i
    size_t evused = 0;
    double tk_ms[MFT_TK_COUNT];
    int64_t tk_n[MFT_TK_COUNT];
};

static const int kBlock = 256;
static inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

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
static void erf_table(int kind, double sigma, double Rp, const mft_tild_desc& d, std::vector<float>& out)
{
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    out.resize((size_t)nx * ny * nz);
    const double PI4 = 12.56637061435917295384;
    double sig2 = sigma * sigma;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                float kx = kcomp(ix, nx, d.lx), ky = kcomp(iy, ny, d.ly), kz = kcomp(iz, nz, d.lz);
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
        K.kind = P.kind;
        K.maskI = P.maskI;
        K.maskJ = P.maskJ;
        K.A = P.A;
        K.sig2 = P.sigma * P.sigma;
        K.Rp = P.Rp;
        K.table = nullptr;
        if (P.kind == MFT_POT_ERF2 || P.kind == MFT_POT_ERFG) {
            std::vector<float> tab;
            erf_table(P.kind, P.sigma, P.Rp, d, tab);
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
    }

    const int N = e->N, M = e->M;
    if ((rc = dalloc(&e->x[0], 3 * (size_t)N))) return rc;
    if ((rc = dalloc(&e->x[1], 3 * (size_t)N))) return rc;
    if ((rc = dalloc(&e->noise_old, 3 * (size_t)N))) return rc;
    if ((rc = dalloc(&e->fgrid, 3 * (size_t)N))) return rc;
    if ((rc = dalloc(&e->ftot, 3 * (size_t)N))) return rc;
    if ((rc = dalloc(&e->type, (size_t)N))) return rc;
    if (d.has_charges && (rc = dalloc(&e->charge, (size_t)N))) return rc;
    if ((rc = dalloc(&e->nBonds, (size_t)N))) return rc;
    if ((rc = dalloc(&e->bondedTo, (size_t)N * (d.maxbonds > 0 ? d.maxbonds : 1)))) return rc;
    if ((rc = dalloc(&e->bondType, (size_t)N * (d.maxbonds > 0 ? d.maxbonds : 1)))) return rc;
    if ((rc = dalloc(&e->count, (size_t)M))) return rc;
    if ((rc = dalloc(&e->cell, (size_t)N))) return rc;
    if ((rc = dalloc(&e->rank, (size_t)N))) return rc;
    if ((rc = dalloc(&e->start, (size_t)M + 1))) return rc;
    e->nscan_blocks = (M + kScanChunk - 1) / kScanChunk;
    if ((rc = dalloc(&e->partial, (size_t)e->nscan_blocks))) return rc;
    if ((rc = dalloc(&e->spos, (size_t)N))) return rc;
    if ((rc = dalloc(&e->smeta, (size_t)N))) return rc;
    if ((rc = dalloc(&e->force_slot, (size_t)d.nspecies))) return rc;
    int nconv_blocks = 0;
    {
        int T = 1;
#define MFT_T(n_) if (d.nz == n_) T = FftPlan<n_>::T;
        MFT_FOR_EACH_FFT_SIZE(MFT_T)
#undef MFT_T
        int B = 256 / T;
        nconv_blocks = (d.nx * d.ny + B - 1) / B;
    }
    if ((rc = dalloc(&e->epart, (size_t)nconv_blocks * kMaxPots))) return rc;
    if ((rc = dalloc(&e->eout, (size_t)kMaxPots + 1))) return rc;
    if ((rc = dalloc(&e->bpart, (size_t)nblk(N)))) return rc;
    MFT_HIP(hipMemset(e->count, 0, sizeof(int) * (size_t)M));
    MFT_HIP(hipMemset(e->noise_old, 0, sizeof(float) * 3 * (size_t)N));
    MFT_HIP(hipMemset(e->nBonds, 0, sizeof(int) * (size_t)N));
    MFT_HIP(hipMemset(e->type, 0, sizeof(int) * (size_t)N));
    MFT_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    return MFT_OK;
}


// Density channels and force fields once the charges are known.  When every
// species carries a single charge value (or there is no charges potential),
// the charge density is sum_s q_s rho_s, so electrostatics folds into the
// per-species k-space kernel and no separate charge channel, transform or
// field is needed; otherwise the charge density is a channel of its own.
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
        if (P.kind != MFT_POT_CHARGES) forced |= P.maskI | P.maskJ;
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
    kp.nchan = e->nchan;
    kp.qchan = e->qchan;
    kp.nforce = e->nforce;

    // brick geometry: 16x8x8 points for up to 4 channels, smaller beyond
    BrickGeom& bg = e->bg;
    auto pow2_floor = [](int v, int cap) {
        int b = 1;
        while (b * 2 <= v && b * 2 <= cap) b *= 2;
        return b;
    };
    auto log2i = [](int v) {
        int l = 0;
        while ((1 << l) < v) ++l;
        return l;
    };
    bg.bx = pow2_floor(d.nx, 16);
    bg.by = pow2_floor(d.ny, 8);
    bg.bz = pow2_floor(d.nz, e->nchan <= 4 ? 8 : 4);
    bg.sx = log2i(bg.bx);
    bg.sy = log2i(bg.by);
    bg.sz = log2i(bg.bz);
    bg.nbx = (d.nx + bg.bx - 1) / bg.bx;
    bg.nby = (d.ny + bg.by - 1) / bg.by;
    bg.nbz = (d.nz + bg.bz - 1) / bg.bz;
    e->dens_lds = (size_t)e->nchan * bg.bx * bg.by * bg.bz * sizeof(unsigned long long);
    {
        int wy = bg.by + d.pmeorder, wz = bg.bz + d.pmeorder;
        MFT_REQUIRE(2 * wy * wz <= 512, MFT_ERR_UNSUPPORTED, "brick halo too large");
    }
    // gather bricks: 8^3 cells, window (8+p)^3 points x 3 components per field
    BrickGeom& gg = e->gbg;
    gg.bx = pow2_floor(d.nx, 8);
    gg.by = pow2_floor(d.ny, 8);
    gg.bz = pow2_floor(d.nz, 8);
    gg.sx = log2i(gg.bx);
    gg.sy = log2i(gg.by);
    gg.sz = log2i(gg.bz);
    gg.nbx = (d.nx + gg.bx - 1) / gg.bx;
    gg.nby = (d.ny + gg.by - 1) / gg.by;
    gg.nbz = (d.nz + gg.bz - 1) / gg.bz;
    {
        int wpts = (gg.bx + d.pmeorder) * (gg.by + d.pmeorder) * (gg.bz + d.pmeorder);
        int fpc = kGatherLdsFloats / (3 * wpts);
        MFT_REQUIRE(fpc >= 1 && gg.by * gg.bz <= 256 && (gg.by + d.pmeorder) * (gg.bz + d.pmeorder) <= 256 &&
                        gg.bx + d.pmeorder <= 16,
                    MFT_ERR_UNSUPPORTED, "gather window too large");
        e->gather_fpc = std::max(1, std::min(fpc, std::max(1, e->nout4)));
    }
    const size_t M = e->M;
    for (void* pbuf : {(void*)e->rho, (void*)e->S, (void*)e->Y, (void*)e->F4})
        if (pbuf) (void)hipFree(pbuf);
    e->rho = nullptr;
    e->S = nullptr;
    e->Y = nullptr;
    e->F4 = nullptr;
    if ((rc = dalloc(&e->rho, (size_t)e->nchan * M))) return rc;
    if ((rc = dalloc(&e->S, (size_t)e->nchan * M))) return rc;
    if ((rc = dalloc(&e->Y, (size_t)(e->nout > 0 ? e->nout : 1) * M))) return rc;
    if ((rc = dalloc(&e->F4, (size_t)(e->nout4 > 0 ? e->nout4 : 1) * M))) return rc;
    MFT_HIP(hipMemcpy(e->force_slot, e->force_slot_h.data(), sizeof(int) * d.nspecies, hipMemcpyHostToDevice));
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
        if (e->timing && 2 * (e->evused + 1) <= e->evpool.size()) {
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

template <int N, int SIGN, bool RI, bool RO>
static void launch_axis_n(const void* in, void* out, AxisGeom g, int nf, long fs, float scale, hipStream_t st)
{
    constexpr int B = 256 / FftPlan<N>::T;
    dim3 grid((g.nlines() + B - 1) / B, nf);
    k_fft_axis<N, SIGN, RI, RO><<<grid, 256, 0, st>>>(in, out, g, fs, fft_twiddles_device(N), scale);
}

template <int SIGN, bool RI, bool RO>
static int launch_axis(const void* in, void* out, AxisGeom g, int nf, long fs, float scale, hipStream_t st)
{
    switch (g.length()) {
#define MFT_CASE(n_)                                                                                      \
    case n_: launch_axis_n<n_, SIGN, RI, RO>(in, out, g, nf, fs, scale, st); break;
        MFT_FOR_EACH_FFT_SIZE(MFT_CASE)
#undef MFT_CASE
    default:
        set_error(MFT_ERR_UNSUPPORTED, "unsupported FFT length %d", g.length());
        return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

static int launch_conv(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    float invM = 1.0f / (float)e->M;
    switch (d.nz) {
#define MFT_CASE(n_)                                                                                      \
    case n_: {                                                                                            \
        constexpr int B = 256 / FftPlan<n_>::T;                                                           \
        unsigned grid = (d.nx * d.ny + B - 1) / B;                                                        \
        float Mg = (float)((double)e->gvol * e->M);                                                       \
        const float2* tw = fft_twiddles_device(n_);                                                      \
        if (e->nchan <= 2)                                                                                \
            k_conv_z<n_, 2><<<grid, 256, 0, e->stream>>>(e->S, e->Y, d.nx, d.ny, d.lx, d.ly, d.lz, e->kp,  \
                                                         tw, invM, e->epart, Mg);                         \
        else if (e->nchan <= 4)                                                                           \
            k_conv_z<n_, 4><<<grid, 256, 0, e->stream>>>(e->S, e->Y, d.nx, d.ny, d.lx, d.ly, d.lz, e->kp,  \
                                                         tw, invM, e->epart, Mg);                         \
        else                                                                                              \
            k_conv_z<n_, 8><<<grid, 256, 0, e->stream>>>(e->S, e->Y, d.nx, d.ny, d.lx, d.ly, d.lz, e->kp,  \
                                                         tw, invM, e->epart, Mg);                         \
        if (e->kp.energy)                                                                                 \
            k_energy_sum<<<d.npotentials, 256, 0, e->stream>>>(e->eout, e->epart, grid, d.npotentials);   \
        break;                                                                                            \
    }
        MFT_FOR_EACH_FFT_SIZE(MFT_CASE)
#undef MFT_CASE
    default: return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}


static int launch_x_inv_vec3(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    switch (d.nx) {
#define MFT_CASE(n_)                                                                                      \
    case n_: {                                                                                            \
        constexpr int B = 256 / FftPlan<n_>::T;                                                           \
        dim3 grid((d.ny * d.nz + B - 1) / B, e->nout4);                                                   \
        k_fft_x_inv_vec3<n_><<<grid, 256, 0, e->stream>>>(e->Y, e->F4, d.ny, d.nz, fft_twiddles_device(n_)); \
        break;                                                                                            \
    }
        MFT_FOR_EACH_FFT_SIZE(MFT_CASE)
#undef MFT_CASE
    default: return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}


// gather variants: 0 = direct loads in cell order, 1 = LDS-staged bricks
static int launch_gather(TildEngine* e, int mode)
{
    const mft_tild_desc& d = e->d;
    if (mode == 0) {
#define MFT_G0(P_)                                                                                        \
    k_gather<P_><<<nblk(e->N), kBlock, 0, e->stream>>>(e->fgrid, e->F4, e->force_slot, e->qslot, e->spos, e->smeta, \
                                                        e->N, d.nx, d.ny, d.nz, e->dx[0], e->dx[1], e->dx[2])
        switch (d.pmeorder) {
        case 0: MFT_G0(0); break;
        case 1: MFT_G0(1); break;
        case 2: MFT_G0(2); break;
        case 3: MFT_G0(3); break;
        default: MFT_G0(4); break;
        }
#undef MFT_G0
    } else {
        const BrickGeom& gb = e->gbg;
        unsigned nb = gb.nbx * gb.nby * gb.nbz;
        size_t lds = sizeof(float) * 3 * (size_t)e->gather_fpc * (gb.bx + d.pmeorder) * (gb.by + d.pmeorder) *
                     (gb.bz + d.pmeorder);
#define MFT_G1(P_)                                                                                        \
    k_gather_brick<P_><<<nb, kGatherThreads, lds, e->stream>>>(e->fgrid, e->F4, e->force_slot, e->qslot, e->nout4, \
                                                               e->spos, e->smeta, e->start, d.nx, d.ny, d.nz,      \
                                                               e->dx[0], e->dx[1], e->dx[2], gb, e->gather_fpc)
        switch (d.pmeorder) {
        case 0: MFT_G1(0); break;
        case 1: MFT_G1(1); break;
        case 2: MFT_G1(2); break;
        case 3: MFT_G1(3); break;
        default: MFT_G1(4); break;
        }
#undef MFT_G1
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

static int bin_current(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    TimedScope ts(e, MFT_TK_BIN);
    k_bin_count<<<nblk(e->N), kBlock, 0, e->stream>>>(e->x[e->cur], e->N, e->count, e->cell, e->rank, e->dx[0],
                                                       e->dx[1], e->dx[2], d.pmeorder, d.nx, d.ny, d.nz);
    MFT_LAUNCH_CHECK();
    e->binned = true;
    return MFT_OK;
}

// Everything up to the per-particle grid force in fgrid.
static int grid_forces(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    int rc;
    if (!e->binned && (rc = bin_current(e))) return rc;
    {
        TimedScope ts(e, MFT_TK_SCAN);
        k_scan_reduce<<<e->nscan_blocks, kScanThreads, 0, e->stream>>>(e->count, e->M, e->partial);
        k_scan_partials<<<1, 1024, 0, e->stream>>>(e->partial, e->nscan_blocks);
        k_scan_apply<<<e->nscan_blocks, kScanThreads, 0, e->stream>>>(e->count, e->M, e->partial, e->start, e->N);
        MFT_LAUNCH_CHECK();
    }
    {
        TimedScope ts(e, MFT_TK_SCATTER);
        k_bin_scatter<<<nblk(e->N), kBlock, 0, e->stream>>>(e->x[e->cur], e->charge, e->type, e->N, e->cell, e->rank,
                                                             e->start, e->spos, e->smeta);
        MFT_LAUNCH_CHECK();
    }
    e->binned = false;
    {
        TimedScope ts(e, MFT_TK_DENSITY);
        unsigned nb = e->bg.nbx * e->bg.nby * e->bg.nbz;
#define MFT_DENS(P_)                                                                                      \
    k_density_brick<P_><<<nb, kDensThreads, e->dens_lds, e->stream>>>(e->rho, e->nchan, e->qchan, e->spos, e->smeta, \
                                                                       e->start, d.nx, d.ny, d.nz, e->dx[0],      \
                                                                       e->dx[1], e->dx[2], e->bg, 1.0f / e->gvol)
        switch (d.pmeorder) {
        case 0: MFT_DENS(0); break;
        case 1: MFT_DENS(1); break;
        case 2: MFT_DENS(2); break;
        case 3: MFT_DENS(3); break;
        default: MFT_DENS(4); break;
        }
#undef MFT_DENS
        MFT_LAUNCH_CHECK();
    }
    if (e->nout == 0) {
        MFT_HIP(hipMemsetAsync(e->fgrid, 0, sizeof(float) * 3 * (size_t)e->N, e->stream));
        return MFT_OK;
    }
    const long M = e->M;
    AxisGeom gx{d.nx, d.ny, d.nz, 0}, gy{d.nx, d.ny, d.nz, 1};
    {
        TimedScope ts(e, MFT_TK_FFT_X_FWD);
        if ((rc = launch_axis<-1, true, false>(e->rho, e->S, gx, e->nchan, M, 1.0f, e->stream))) return rc;
    }
    {
        TimedScope ts(e, MFT_TK_FFT_Y_FWD);
        if ((rc = launch_axis<-1, false, false>(e->S, e->S, gy, e->nchan, M, 1.0f, e->stream))) return rc;
    }
    {
        TimedScope ts(e, MFT_TK_CONV_Z);
        if ((rc = launch_conv(e))) return rc;
    }
    {
        TimedScope ts(e, MFT_TK_FFT_Y_INV);
        if ((rc = launch_axis<+1, false, false>(e->Y, e->Y, gy, e->nout, M, 1.0f, e->stream))) return rc;
    }
    {
        TimedScope ts(e, MFT_TK_FFT_X_INV);
        if ((rc = launch_x_inv_vec3(e))) return rc;
    }
    {
        TimedScope ts(e, MFT_TK_GATHER);
        if ((rc = launch_gather(e, e->gather_mode))) return rc;
        MFT_LAUNCH_CHECK();
    }
    return MFT_OK;
}

static void update_ramps(TildEngine* e, int64_t i, int64_t nsteps)
{
    // PS_Potential::update_prefactor (ps_potential.cu:227-248): after step
    // i the amplitude is Ao_i + (Ao_f - Ao_i)/maxsteps * (i + 1).
    for (int p = 0; p < e->d.npotentials; ++p) {
        const mft_tild_potential& P = e->pots[p];
        if (!P.ramp) continue;
        double A = (double)P.A + ((double)P.A_final - (double)P.A) / (double)nsteps * (double)(i + 1);
        e->kp.pot[p].A = (float)A;
    }
}

static int do_steps(TildEngine* e, int nsteps, bool want_f)
{
    const mft_tild_desc& d = e->d;
    int rc;
    for (int i = 0; i < nsteps; ++i) {
        if (e->timing && 2 * (e->evused + MFT_TK_COUNT) > e->evpool.size()) resolve_timing(e);
        if ((rc = grid_forces(e))) return rc;
        TimedScope ts(e, MFT_TK_INTEGRATE);
        int nxt = 1 - e->cur;
        const bool last = i == nsteps - 1;
        k_bonds_integrate<<<nblk(e->N), kBlock, 0, e->stream>>>(
            e->x[e->cur], e->x[nxt], e->noise_old, e->fgrid, (want_f && last) ? e->ftot : nullptr, e->type, e->nBonds,
            e->bondedTo, e->bondType, d.maxbonds, e->bp, e->gp, d.integrator, d.delt, d.lx, d.ly, d.lz, e->N, d.seed,
            (uint32_t)e->step, e->count, e->cell, e->rank, e->dx[0], e->dx[1], e->dx[2], d.pmeorder, d.nx, d.ny,
            d.nz);
        MFT_LAUNCH_CHECK();
        if (d.integrator != MFT_INT_NONE) {
            e->cur = nxt;
            e->binned = true;
        }
        update_ramps(e, i, nsteps);
        e->step++;
    }
    return MFT_OK;
}

} // namespace mft

using namespace mft;

struct mft_tild : TildEngine {};

extern "C" {

int mft_tild_grid_weights(float* gridW, int* gridInds, const float* x, int ns, const int Nx[3], const float dx[3],
                          int pmeorder, void* stream)
{
    MFT_REQUIRE(gridW && gridInds && x && Nx && dx && ns >= 0, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(pmeorder >= 0 && pmeorder <= 4, MFT_ERR_INVALID, "pmeorder must be 0..4");
    if (ns == 0) return MFT_OK;
    k_grid_weights<<<nblk(ns), kBlock, 0, (hipStream_t)stream>>>(gridW, gridInds, x, ns, Nx[0], Nx[1], Nx[2], dx[0],
                                                                   dx[1], dx[2], pmeorder);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int mft_tild_fill_density(float* rho, float* rhoq, const float* q, const int* sites, int nsites, int ns,
                          const int* gridInds, const float* gridW, int G, int M, void* stream)
{
    MFT_REQUIRE(rho && sites && gridInds && gridW && nsites >= 0 && G > 0 && M > 0, MFT_ERR_INVALID,
                "null argument");
    MFT_REQUIRE((rhoq == nullptr) == (q == nullptr), MFT_ERR_INVALID, "rhoq and q go together");
    hipStream_t st = (hipStream_t)stream;
    unsigned long long* acc = nullptr;
    int rc;
    if ((rc = dalloc(&acc, (size_t)M * (rhoq ? 2 : 1)))) return rc;
    MFT_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * (size_t)M * (rhoq ? 2 : 1), st));
    if (nsites > 0)
        k_fill_density_tables<<<nblk(nsites), kBlock, 0, st>>>(acc, rhoq ? acc + M : nullptr, q, sites, nsites,
                                                                gridInds, gridW, G, M, ns);
    k_fixed_to_float_add<<<nblk(M), kBlock, 0, st>>>(rho, acc, M);
    if (rhoq) k_fixed_to_float_add<<<nblk(M), kBlock, 0, st>>>(rhoq, acc + M, M);
    MFT_LAUNCH_CHECK();
    MFT_HIP(hipStreamSynchronize(st));
    MFT_HIP(hipFree(acc));
    return MFT_OK;
}

int mft_tild_map_forces(float* f, const float* gridF, const float* q, const int* sites, int nsites, int ns,
                        const int* gridInds, const float* gridW, int G, float gvol, int M, void* stream)
{
    MFT_REQUIRE(f && gridF && sites && gridInds && gridW && nsites >= 0 && G > 0, MFT_ERR_INVALID, "null argument");
    if (nsites == 0) return MFT_OK;
    k_map_forces_tables<<<nblk(nsites), kBlock, 0, (hipStream_t)stream>>>(f, gridF, q, sites, nsites, gridInds, gridW,
                                                                            G, gvol, M, ns);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int mft_tild_bonds(float* f, const int* nBonds, const int* bondedTo, const int* bondType, const float* bondReq,
                   const float* bondK, const int* bondStyle, int nbondtypes, const float* x, const float L[3], int ns,
                   int maxbonds, void* stream)
{
    MFT_REQUIRE(f && nBonds && x && L && ns >= 0, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(nbondtypes >= 0 && nbondtypes <= kMaxBondTypes, MFT_ERR_INVALID, "at most %d bond types",
                kMaxBondTypes);
    MFT_REQUIRE(nbondtypes == 0 || (bondReq && bondK && bondStyle), MFT_ERR_INVALID, "null bond tables");
    BondParams bp;
    std::memset(&bp, 0, sizeof(bp));
    bp.ntypes = nbondtypes;
    for (int i = 0; i < nbondtypes; ++i) {
        bp.k[i] = bondK[i];
        bp.r0[i] = bondReq[i];
        bp.style[i] = bondStyle[i];
    }
    if (ns == 0) return MFT_OK;
    k_bonds<<<nblk(ns), kBlock, 0, (hipStream_t)stream>>>(f, nBonds, bondedTo, bondType, bp, x, L[0], L[1], L[2], ns,
                                                           maxbonds);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

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
    void* ptrs[] = {e->x[0], e->x[1], e->noise_old, e->fgrid, e->ftot, e->charge, e->type, e->nBonds, e->bondedTo,
                    e->bondType, e->count, e->cell, e->rank, e->start, e->partial, e->spos, e->smeta, e->rho, e->S,
                    e->Y, e->F4, e->force_slot, e->epart, e->eout, e->bpart};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (float* t : e->tables)
        if (t) (void)hipFree(t);
    for (hipEvent_t ev : e->evpool) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return MFT_OK;
}

void* mft_tild_stream(mft_tild* e) { return e ? (void*)e->stream : nullptr; }

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
    MFT_HIP(hipMemcpy(e->x[0], x, sizeof(float) * n3, hipMemcpyHostToDevice));
    MFT_HIP(hipMemcpy(e->x[1], xo ? xo : x, sizeof(float) * n3, hipMemcpyHostToDevice));
    if (noise_old) MFT_HIP(hipMemcpy(e->noise_old, noise_old, sizeof(float) * n3, hipMemcpyHostToDevice));
    else MFT_HIP(hipMemset(e->noise_old, 0, sizeof(float) * n3));
    if (e->binned) {
        // counts of stale positions would be double-counted: clear them
        MFT_HIP(hipMemset(e->count, 0, sizeof(int) * (size_t)e->M));
        e->binned = false;
    }
    e->have_state = 1;
    return MFT_OK;
}

int mft_tild_get_state(mft_tild* e, float* x, float* xo, float* noise_old)
{
    MFT_REQUIRE(e, MFT_ERR_INVALID, "null argument");
    MFT_HIP(hipStreamSynchronize(e->stream));
    const size_t n3 = 3 * (size_t)e->N;
    if (x) MFT_HIP(hipMemcpy(x, e->x[e->cur], sizeof(float) * n3, hipMemcpyDeviceToHost));
    if (xo) MFT_HIP(hipMemcpy(xo, e->x[1 - e->cur], sizeof(float) * n3, hipMemcpyDeviceToHost));
    if (noise_old) MFT_HIP(hipMemcpy(noise_old, e->noise_old, sizeof(float) * n3, hipMemcpyDeviceToHost));
    return MFT_OK;
}

int mft_tild_compute_forces(mft_tild* e, float* f)
{
    MFT_REQUIRE(e && e->have_state, MFT_ERR_INVALID, "engine has no state");
    int rc;
    int saved = e->d.integrator;
    e->d.integrator = MFT_INT_NONE;
    int64_t st = e->step;
    rc = do_steps(e, 1, true);
    e->d.integrator = saved;
    e->step = st;
    if (rc) return rc;
    MFT_HIP(hipStreamSynchronize(e->stream));
    if (f) MFT_HIP(hipMemcpy(f, e->ftot, sizeof(float) * 3 * (size_t)e->N, hipMemcpyDeviceToHost));
    return MFT_OK;
}

int mft_tild_get_densities(mft_tild* e, float* rho, float* rhoq)
{
    MFT_REQUIRE(e, MFT_ERR_INVALID, "null argument");
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
    std::vector<float4> tmp(M);
    MFT_HIP(hipStreamSynchronize(e->stream));
    MFT_HIP(hipMemcpy(tmp.data(), e->F4 + (size_t)slot * M, sizeof(float4) * M, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < M; ++i) {
        F[3 * i + 0] = tmp[i].x;
        F[3 * i + 1] = tmp[i].y;
        F[3 * i + 2] = tmp[i].z;
    }
    return MFT_OK;
}

int mft_tild_step(mft_tild* e, int nsteps)
{
    MFT_REQUIRE(e && e->have_state, MFT_ERR_INVALID, "engine has no state");
    MFT_REQUIRE(nsteps >= 0, MFT_ERR_INVALID, "negative step count");
    return do_steps(e, nsteps, false);
}

int mft_tild_synchronize(mft_tild* e)
{
    MFT_REQUIRE(e, MFT_ERR_INVALID, "null argument");
    MFT_HIP(hipStreamSynchronize(e->stream));
    return MFT_OK;
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
        rc = do_steps(e, 1, false);
        e->kp.energy = 0;
        e->d.integrator = saved;
        e->step = st;
        if (rc) return rc;
        MFT_HIP(hipStreamSynchronize(e->stream));
        MFT_HIP(hipMemcpy(pot_energy, e->eout, sizeof(double) * d.npotentials, hipMemcpyDeviceToHost));
    }
    if (bond_energy) {
        unsigned nb = nblk(e->N);
        k_bond_energy<<<nb, kBlock, 0, e->stream>>>(e->bpart, e->nBonds, e->bondedTo, e->bondType, e->bp, e->x[e->cur],
                                                     d.lx, d.ly, d.lz, e->N, d.maxbonds);
        k_sum_partials<<<1, 1024, 0, e->stream>>>(e->eout + kMaxPots, e->bpart, nb);
        MFT_LAUNCH_CHECK();
        MFT_HIP(hipStreamSynchronize(e->stream));
        double b;
        MFT_HIP(hipMemcpy(&b, e->eout + kMaxPots, sizeof(double), hipMemcpyDeviceToHost));
        *bond_energy = 0.5 * b;
    }
    return MFT_OK;
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

int mft_tild_kernel_bytes(mft_tild* e, double* bytes)
{
    MFT_REQUIRE(e && bytes, MFT_ERR_INVALID, "null argument");
    const double N = e->N, M = e->M, C = e->nchan, O = e->nout, O4 = e->nout4;
    const bool q = e->d.has_charges;
    for (int k = 0; k < MFT_TK_COUNT; ++k) bytes[k] = 0;
    // minimal HBM traffic of each kernel's job (reads + writes once)
    bytes[MFT_TK_BIN] = N * (12 + 4 + 4 + 4);                 // x, cell, rank, count
    bytes[MFT_TK_SCAN] = M * (4 + 4 + 4);                     // count read, start write, count clear
    bytes[MFT_TK_SCATTER] = N * (12 + (q ? 4 : 0) + 4 + 4 + 4 + 4 + 16 + 8);
    bytes[MFT_TK_DENSITY] = N * 24 + M * 4 + M * 4 * C;       // sorted records, starts, rho
    bytes[MFT_TK_FFT_X_FWD] = M * C * (4 + 8);
    bytes[MFT_TK_FFT_Y_FWD] = M * C * 16;
    bytes[MFT_TK_CONV_Z] = M * 8 * (C + O);
    bytes[MFT_TK_FFT_Y_INV] = M * O * 16;
    bytes[MFT_TK_FFT_X_INV] = M * O * 8 + M * O4 * 16;
    bytes[MFT_TK_GATHER] = N * (24 + 12) + M * 16 * O4;
    bytes[MFT_TK_INTEGRATE] = N * (12 + 12 + 12 + 24 + 12 + 4 + 4 + 8 + 8 + 12) + N * 12;
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
    int rc = do_steps(e, 1, false);
    e->d.integrator = saved;
    e->step = st;
    if (rc) return rc;
    hipEvent_t a, b;
    MFT_HIP(hipEventCreate(&a));
    MFT_HIP(hipEventCreate(&b));
    float best = 1e30f;
    int best_mode = 0;
    for (int mode = 0; mode < 2; ++mode) {
        if ((rc = launch_gather(e, mode))) return rc;  // warm
        MFT_HIP(hipEventRecord(a, e->stream));
        for (int r = 0; r < reps; ++r)
            if ((rc = launch_gather(e, mode))) return rc;
        MFT_HIP(hipEventRecord(b, e->stream));
        MFT_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        MFT_HIP(hipEventElapsedTime(&ms, a, b));
        if (ms < best) {
            best = ms;
            best_mode = mode;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    e->gather_mode = best_mode;
    if (choice) *choice = best_mode;
    return MFT_OK;
}

int mft_tild_num_outputs(mft_tild* e, int* nout)
{
    MFT_REQUIRE(e && nout, MFT_ERR_INVALID, "null argument");
    *nout = e->nout;
    return MFT_OK;
}

} // extern "C"
