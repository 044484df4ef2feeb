// TILD particle kernels: per-kernel drop-in entry points (grid weights,
// table-driven density and force mapping, bonds), binning by brick-major
// key, tile density and reduction, gathers, and the fused bonds + integrator.
#include "tild_impl.h"

namespace mft {

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
// Base grid index of one coordinate wrapped into [0, n): the brick a particle
// is sorted into and the first point (+ p/2) of its spline stencil.
__device__ __forceinline__ int wrapped_base(float x, float dx, int p, int n) { return wrap(base_index(x, dx, p), n); }

__device__ __forceinline__ int key_of(float x0, float x1, float x2, int s, float dx0, float dx1, float dx2, int p,
                                      int nx, int ny, int nz, const TileGeom& tg)
{
    return tile_key(wrapped_base(x0, dx0, p, nx), wrapped_base(x1, dx1, p, ny), wrapped_base(x2, dx2, p, nz), s, tg);
}

// Rank of each lane's key among equal keys, one atomic per run of equal keys
// in consecutive lanes (particles in key order mostly keep their key, so a
// wave's lanes fall into a few runs).  Every lane of the wave calls this;
// inactive lanes do not count.
__device__ __forceinline__ int wave_rank(int* __restrict__ count, int key, bool active)
{
    const int lane = threadIdx.x & 63;
    const int k = active ? key : -1 - lane; // unique, never counted
    const int prev = __shfl_up(k, 1, 64);
    const bool head = lane == 0 || k != prev;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int hl = 63 - __clzll(heads & upto);
    const unsigned long long above = heads & ~upto;
    const int next = above ? __ffsll((long long)above) - 1 : 64;
    int base = 0;
    if (head && active) base = atomicAdd(&count[k], next - lane);
    base = __shfl(base, hl, 64);
    return base + (lane - hl);
}

// Initial sort from particle-id order: key and rank of every particle.
__global__ void k_sort_count(const float* __restrict__ x, const int* __restrict__ type, int ns,
                             int* __restrict__ count, int* __restrict__ key, int* __restrict__ rank, float dx0,
                             float dx1, float dx2, int p, int nx, int ny, int nz, TileGeom tg)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    int c = key_of(x[i * 3], x[i * 3 + 1], x[i * 3 + 2], type[i], dx0, dx1, dx2, p, nx, ny, nz, tg);
    key[i] = c;
    rank[i] = atomicAdd(&count[c], 1);
}

// Initial sort: every particle's state into its slot.
__global__ void k_sort_place(const float* __restrict__ x, const float* __restrict__ xo,
                             const float* __restrict__ noise, const float* __restrict__ q,
                             const int* __restrict__ type, const int* __restrict__ bondType, int maxbonds, int ns, const int* __restrict__ key,
                             const int* __restrict__ rank, const int* __restrict__ start, int* __restrict__ slot_of,
                             float4* __restrict__ X, float4* __restrict__ XO, float4* __restrict__ NZ,
                             PMeta* __restrict__ META, int* __restrict__ BT)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int s = start[key[i]] + rank[i];
    slot_of[i] = s;
    X[s] = make_float4(x[i * 3], x[i * 3 + 1], x[i * 3 + 2], q ? q[i] : 0.f);
    ff_store(XO, s, xo[i * 3], xo[i * 3 + 1], xo[i * 3 + 2]);
    NZ[s] = make_float4(noise[i * 3], noise[i * 3 + 1], noise[i * 3 + 2], 0.f);
    META[s] = make_meta(i, type[i]);
    for (int b = 0; b < maxbonds; ++b) BT[s * maxbonds + b] = bondType[i * maxbonds + b];
}

// Initial sort: bond partners as slots, -1 past the particle's bonds (the
// key-ordered state keeps no bond count: a slot's bonds are its entries >= 0).
__global__ void k_sort_bonds(const int* __restrict__ bondedTo, const int* __restrict__ nBonds, int maxbonds, int ns,
                             const int* __restrict__ slot_of, int* __restrict__ BS)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int s = slot_of[i];
    for (int b = 0; b < maxbonds; ++b) BS[s * maxbonds + b] = b < nBonds[i] ? slot_of[bondedTo[i * maxbonds + b]] : -1;
}

// Initial sort of the angles: counts, types and members as slots.
__global__ void k_sort_angles(const int* __restrict__ nAngles, const int* __restrict__ angleType,
                              const int* __restrict__ angleGroup, int ma, int ns, const int* __restrict__ slot_of,
                              int* __restrict__ NA, int* __restrict__ AT, int* __restrict__ AG)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int s = slot_of[i], na = nAngles[i];
    NA[s] = na;
    for (int a = 0; a < ma; ++a) {
        AT[s * ma + a] = angleType[i * ma + a];
        for (int c = 0; c < 3; ++c)
            AG[(s * ma + a) * 3 + c] = a < na ? slot_of[angleGroup[(i * ma + a) * 3 + c]] : 0;
    }
}

// key order -> particle-id order: out[3 id + j] = src[slot].j (packed: src
// holds 12-byte records, ff_load)
__global__ void k_unsort(float* __restrict__ out, const float4* __restrict__ src, const PMeta* __restrict__ META,
                         int ns, int packed)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    float4 v;
    if (packed) {
        const float3 p = ff_load(src, s);
        v = make_float4(p.x, p.y, p.z, 0.f);
    } else {
        v = src[s];
    }
    const int id = META[s].id();
    out[id * 3 + 0] = v.x;
    out[id * 3 + 1] = v.y;
    out[id * 3 + 2] = v.z;
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

// start[i] = exclusive prefix of count; count cleared for the next binning.
// Each block adds the totals of the blocks before it (at most a few thousand
// values, read from L2), so the scan takes two launches.
__global__ void k_scan_apply(int* __restrict__ count, int M, const int* __restrict__ partial,
                             int* __restrict__ start, int ns, uint32_t* __restrict__ dstep)
{
    __shared__ int sh[16];
    int c = 0;
    for (int j = threadIdx.x; j < (int)blockIdx.x; j += blockDim.x) c += partial[j];
    int carry;
    block_exclusive_scan(c, sh, &carry);
    // M is a multiple of 64 (64 sub-cell keys per brick and species), so a
    // thread's 16 keys are all in range or all past it: int4 loads and stores
    static_assert(kScanItems == 16, "four int4 per thread");
    const long base = (long)blockIdx.x * kScanChunk + (long)threadIdx.x * kScanItems;
    const bool in = base < M;
    int4 v4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        v4[k] = in ? reinterpret_cast<const int4*>(count + base)[k] : make_int4(0, 0, 0, 0);
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) s += v4[k].x + v4[k].y + v4[k].z + v4[k].w;
    int tot;
    int ex = block_exclusive_scan(s, sh, &tot) + carry;
    if (in) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int4 o;
            o.x = ex;
            o.y = o.x + v4[k].x;
            o.z = o.y + v4[k].y;
            o.w = o.z + v4[k].z;
            ex = o.w + v4[k].w;
            reinterpret_cast<int4*>(start + base)[k] = o;
            reinterpret_cast<int4*>(count + base)[k] = make_int4(0, 0, 0, 0);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        start[M] = ns;
        // the step's last reader of the device step index ran before this
        // scan (the integrator): advance it for the next step
        if (dstep) *dstep += 1u;
    }
}

// Particle state into the next step's key order: the slot's next position
// and noise (from the integrator), its current position as the next previous
// position, identity and bonds; map[slot] = new slot.  BT/BT2 null: a single
// bond type, whose slots' types are not carried (every bond is type 0).
__global__ void k_permute(int ns, int maxbonds, const int* __restrict__ key, const int* __restrict__ rank,
                          const int* __restrict__ start, int* __restrict__ map, const float4* __restrict__ XN,
                          const float4* __restrict__ NN, const float4* __restrict__ X, const PMeta* __restrict__ META,
                          const int* __restrict__ BT, float4* __restrict__ X2, float4* __restrict__ XO2,
                          float4* __restrict__ NZ2, PMeta* __restrict__ META2, int* __restrict__ BT2)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int n = start[key[s]] + rank[s];
    map[s] = n;
    const float4 xc = X[s];
    const float3 xn = ff_load(XN, s);
    X2[n] = make_float4(xn.x, xn.y, xn.z, xc.w);
    ff_store(XO2, n, xc.x, xc.y, xc.z);
    if (NN) NZ2[n] = NN[s];
    META2[n] = META[s];
    if (BT2)
        for (int b = 0; b < maxbonds; ++b) BT2[n * maxbonds + b] = BT[s * maxbonds + b];
}

// bond partner slots into the new order (after every map entry is written)
__global__ void k_remap_bonds(int ns, int maxbonds, const int* __restrict__ map, const int* __restrict__ BS,
                              int* __restrict__ BS2)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int n = map[s];
    if (maxbonds == 2) {
        // both partners' map entries loaded together
        const int2 b = reinterpret_cast<const int2*>(BS)[s];
        const int m0 = b.x >= 0 ? map[b.x] : -1;
        const int m1 = b.y >= 0 ? map[b.y] : -1;
        reinterpret_cast<int2*>(BS2)[n] = make_int2(m0, m1);
        return;
    }
    for (int b = 0; b < maxbonds; ++b) {
        const int p = BS[s * maxbonds + b];
        BS2[n * maxbonds + b] = p >= 0 ? map[p] : -1;
    }
}

// angles into the new order (after every map entry is written)
__global__ void k_permute_angles(int ns, int ma, const int* __restrict__ map, const int* __restrict__ NA,
                                 const int* __restrict__ AT, const int* __restrict__ AG, int* __restrict__ NA2,
                                 int* __restrict__ AT2, int* __restrict__ AG2)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int n = map[s], na = NA[s];
    NA2[n] = na;
    for (int a = 0; a < ma; ++a) {
        AT2[n * ma + a] = AT[s * ma + a];
        for (int c = 0; c < 3; ++c) AG2[(n * ma + a) * 3 + c] = a < na ? map[AG[(s * ma + a) * 3 + c]] : 0;
    }
}

// --------------------------------------------------------- tile density ---
constexpr int kTileThreads = 256;

// Spline weights of one particle and its brick-local stencil origin; false
// when the particle is not in the brick (cannot happen for keys computed from
// the same positions; the check keeps LDS accesses in bounds regardless).
template <int P>
__device__ __forceinline__ bool stencil_in_brick(const float4& pp, float dx0, float dx1, float dx2, float i0,
                                                 float i1, float i2, int nx, int ny, int nz, int o0, int o1, int o2,
                                                 const TileGeom& tg, float* wx, float* wy, float* wz, int* l)
{
    int g0 = wrap(base_index_weights_fast(pp.x, dx0, i0, P, wx), nx);
    int g1 = wrap(base_index_weights_fast(pp.y, dx1, i1, P, wy), ny);
    int g2 = wrap(base_index_weights_fast(pp.z, dx2, i2, P, wz), nz);
    l[0] = g0 - o0;
    l[1] = g1 - o1;
    l[2] = g2 - o2;
    return (unsigned)l[0] < (unsigned)tg.b[0] && (unsigned)l[1] < (unsigned)tg.b[1] &&
           (unsigned)l[2] < (unsigned)tg.b[2];
}

// Upper bound of one (p+1)^3 spline weight w_x w_y w_z.
__host__ __device__ constexpr float stencil_wmax(int p)
{
    return p <= 1 ? 1.0f : (p == 2 ? 0.43f : (p == 3 ? 0.30f : 0.22f));
}

// Passes of the deposit whose particles the caller loads up front (before the
// tile is cleared), so their loads overlap the clearing and each other.
constexpr int kDepositPre = 2;

template <int P, int BM, typename ACC>
__device__ __forceinline__ void deposit_one(ACC* acc, int qchan, const float4& pp, int typ, float sc, float dx0,
                                            float dx1, float dx2, float i0, float i1, float i2, int nx, int ny,
                                            int nz, int o0, int o1, int o2, const TileGeom& tg)
{
    constexpr int TP = BM + P, T3 = TP * TP * TP;
    float wx[P + 1], wy[P + 1], wz[P + 1];
    int l[3];
    if (!stencil_in_brick<P>(pp, dx0, dx1, dx2, i0, i1, i2, nx, ny, nz, o0, o1, o2, tg, wx, wy, wz, l)) return;
    ACC* a = acc + typ * T3 + l[0] + TP * (l[1] + TP * l[2]);
#pragma unroll
    for (int iz = 0; iz <= P; ++iz) {
#pragma unroll
        for (int iy = 0; iy <= P; ++iy) {
            const float wyz = wy[iy] * wz[iz] * sc;
#pragma unroll
            for (int ix = 0; ix <= P; ++ix)
                atomicAdd(a + ix + TP * (iy + TP * iz), (ACC)__float2uint_rz(fmaf(wx[ix], wyz, 0.5f)));
        }
    }
    if (sizeof(ACC) == 8 && qchan >= 0 && pp.w != 0.f) {
        ACC* aq = acc + qchan * T3 + l[0] + TP * (l[1] + TP * l[2]);
#pragma unroll
        for (int iz = 0; iz <= P; ++iz) {
#pragma unroll
            for (int iy = 0; iy <= P; ++iy) {
                const float wyz = wy[iy] * wz[iz] * pp.w * sc;
#pragma unroll
                for (int ix = 0; ix <= P; ++ix)
                    atomicAdd(aq + ix + TP * (iy + TP * iz), (ACC)__float2ll_rn(wx[ix] * wyz));
            }
        }
    }
}

// Particle r of the brick taken by pass-j thread: particles are in cell
// order, so neighbouring ones share stencil points; the 64 lanes of a wave
// take particles n/64 apart (a 64 x S transpose) so one atomic instruction
// rarely hits the same LDS word twice.  -1: none.
__device__ __forceinline__ int deposit_rank(int j, int n)
{
    const int S = (n + 63) >> 6;
    const int r = (j & 63) * S + (j >> 6);
    return (j < 64 * S && r < n) ? r : -1;
}

// Species of the brick's particle at slot it: the key order puts a brick's
// particles species by species, so the species follows from the segment
// starts (uniform over the workgroup: scalar loads), with no per-particle
// record read.  key0 = the brick's first key.
__device__ __forceinline__ int species_in_brick(const int* __restrict__ start, int key0, int nspec, int it)
{
    int t = 0;
    for (int sp = 1; sp < nspec; ++sp) t += it >= start[key0 + (sp << 6)] ? 1 : 0;
    return t;
}

// Deposit the brick's particles into the LDS tile.  ACC is the fixed-point
// accumulator word: 32-bit with 2^sc scaling when n w_max 2^sc < 2^32 bounds
// every sum, else 64-bit with 2^31 (and signed charges).  Integer sums do not
// depend on the order of the atomics.  pp/typ: this thread's particles of the
// first kDepositPre passes, loaded by the caller.
template <int P, int BM, typename ACC>
__device__ __forceinline__ void deposit_tile(ACC* acc, int qchan, const float4* __restrict__ spos,
                                             const int* __restrict__ start, int key0, int pb, int n, float sc,
                                             float dx0, float dx1, float dx2, int nx, int ny, int nz, int o0, int o1,
                                             int o2, const TileGeom& tg, const float4 (&pp)[kDepositPre],
                                             const int (&typ)[kDepositPre])
{
    const float i0 = 1.0f / dx0, i1 = 1.0f / dx1, i2 = 1.0f / dx2;
#pragma unroll
    for (int q = 0; q < kDepositPre; ++q)
        if (deposit_rank(threadIdx.x + q * kTileThreads, n) >= 0)
            deposit_one<P, BM, ACC>(acc, qchan, pp[q], typ[q], sc, dx0, dx1, dx2, i0, i1, i2, nx, ny, nz, o0, o1,
                                    o2, tg);
    const int S = (n + 63) >> 6;
    for (int j = threadIdx.x + kDepositPre * kTileThreads; j < 64 * S; j += kTileThreads) {
        const int r = deposit_rank(j, n);
        if (r < 0) continue;
        deposit_one<P, BM, ACC>(acc, qchan, spos[pb + r], species_in_brick(start, key0, tg.nspec, pb + r), sc, dx0,
                                dx1, dx2, i0, i1, i2, nx, ny, nz, o0, o1, o2, tg);
    }
}

// Walk of one thread over a brick's tile entries i, i + kTileThreads, ...
// (i = c T3 + t0 + TP (t1 + TP t2)) with the entry's offset in the
// halo-expanded tile array: the tile digits (t0, t1, t2) and the offset
// advance by the mixed-radix digits of kTileThreads with carries, so no
// entry needs an integer division or a 64-bit multiply.
template <int TP>
struct TileWalk {
    static constexpr int T3 = TP * TP * TP;
    static constexpr int DC = kTileThreads / T3, R = kTileThreads % T3;
    static constexpr int D2 = R / (TP * TP), D1 = (R / TP) % TP, D0 = R % TP;
    int t0, t1, t2;
    long off;
    long step;     // offset change with no carry
    int j0, j1;    // extra offset of a carry out of t0, out of t1
    long j2;       // ... out of t2 (into the channel)
    __device__ TileWalk(int i, int ex0, int ex1, long cstride)
    {
        const int c = i / T3, t = i - c * T3;
        t0 = t % TP;
        t1 = (t / TP) % TP;
        t2 = t / (TP * TP);
        const int e01 = ex0 * ex1;
        off = c * cstride + t0 + (long)ex0 * t1 + (long)e01 * t2;
        step = D0 + (long)D1 * ex0 + (long)D2 * e01 + DC * cstride;
        j0 = ex0 - TP;
        j1 = e01 - TP * ex0;
        j2 = cstride - (long)TP * e01;
    }
    __device__ __forceinline__ void advance()
    {
        t0 += D0;
        const bool c0 = t0 >= TP;
        t0 -= c0 ? TP : 0;
        t1 += D1 + (c0 ? 1 : 0);
        const bool c1 = t1 >= TP;
        t1 -= c1 ? TP : 0;
        t2 += D2 + (c1 ? 1 : 0);
        const bool c2 = t2 >= TP;
        t2 -= c2 ? TP : 0;
        off += step + (c0 ? j0 : 0) + (c1 ? j1 : 0) + (c2 ? j2 : 0);
    }
};

// One workgroup per brick: the brick's particles (contiguous in key order)
// spread their (p+1)^3 weights over the brick's tile in LDS in fixed point;
// the tile is written as float times 1/gvol.
template <int P, int BM>
__global__ void __launch_bounds__(kTileThreads) k_density_tile(
    float* __restrict__ tiles, int nchan, int qchan, const float4* __restrict__ spos,
    const int* __restrict__ start, TileGeom tg, int nx, int ny, int nz, float dx0, float dx1, float dx2,
    double inv_gvol)
{
    constexpr int TP = BM + P, T3 = TP * TP * TP;
    extern __shared__ unsigned long long acc64[]; // [nchan][T3] (32-bit words in the narrow case)
    unsigned int* acc32 = reinterpret_cast<unsigned int*>(acc64);
    const int brick = brick_of_block(blockIdx.x, tg.nbricks);
    const int k0 = brick % tg.nb[0], k1 = (brick / tg.nb[0]) % tg.nb[1], k2 = brick / (tg.nb[0] * tg.nb[1]);
    const int o0 = k0 << tg.sb[0], o1 = k1 << tg.sb[1], o2 = k2 << tg.sb[2];
    const int key0 = (brick * tg.nspec) << 6;
    const int pb = start[key0], pe = start[((brick + 1) * tg.nspec) << 6];
    const int n = pe - pb;
    // the first passes' particles are loaded before the tile is cleared, so
    // their latency overlaps the clearing and the barrier (which waits for LDS
    // only)
    float4 pp0[kDepositPre];
    int typ0[kDepositPre];
#pragma unroll
    for (int q = 0; q < kDepositPre; ++q) {
        const int r = deposit_rank(threadIdx.x + q * kTileThreads, n);
        pp0[q] = r >= 0 ? spos[pb + r] : make_float4(0.f, 0.f, 0.f, 0.f);
        typ0[q] = r >= 0 ? species_in_brick(start, key0, tg.nspec, pb + r) : 0;
    }
    // largest power-of-two scale 2^sc with n (w_max 2^sc + 1/2) < 2^32, the
    // bound of one point's sum of rounded deposits (checked for sc >= 16)
    const float bound = (float)n * (stencil_wmax(P) + 1.0f / 65536.0f) + 1.0f;
    int need = 1;
    while (need < 31 && (float)(1u << need) <= bound) ++need;
    const int sc = 32 - need;
    const bool narrow = qchan < 0 && sc >= 16;
    // tile rows go to the halo-expanded grid: brick k at k * TP per axis
    const long cstride = (long)tg.ex[0] * tg.ex[1] * tg.ex[2];
    float* dst = tiles + (long)k0 * TP + (long)tg.ex[0] * (k1 * TP + (long)tg.ex[1] * (k2 * TP));
    const int nt = nchan * T3;
    if (narrow) {
        for (int i = threadIdx.x; i < nchan * T3; i += kTileThreads) acc32[i] = 0u;
        __syncthreads();
        deposit_tile<P, BM, unsigned int>(acc32, qchan, spos, start, key0, pb, n, (float)(1u << sc), dx0, dx1, dx2,
                                          nx, ny, nz, o0, o1, o2, tg, pp0, typ0);
        __syncthreads();
        const double scale = inv_gvol / (double)(1u << sc);
        TileWalk<TP> w(threadIdx.x, tg.ex[0], tg.ex[1], cstride);
        for (int i = threadIdx.x; i < nt; i += kTileThreads, w.advance())
            dst[w.off] = (float)((double)acc32[i] * scale);
    } else {
        for (int i = threadIdx.x; i < nchan * T3; i += kTileThreads) acc64[i] = 0ull;
        __syncthreads();
        deposit_tile<P, BM, unsigned long long>(acc64, qchan, spos, start, key0, pb, n, 2147483648.0f, dx0, dx1,
                                                dx2, nx, ny, nz, o0, o1, o2, tg, pp0, typ0);
        __syncthreads();
        const double scale = inv_gvol / 2147483648.0;
        TileWalk<TP> w(threadIdx.x, tg.ex[0], tg.ex[1], cstride);
        for (int i = threadIdx.x; i < nt; i += kTileThreads, w.advance())
            dst[w.off] = (float)((double)(long long)acc64[i] * scale);
    }
}

// rho at each grid point: the tile entries of every brick whose tile covers
// it, summed in a fixed order.  tab holds, per axis and coordinate, the count
// and the (brick, tile index) offsets of the covering tiles along that axis.
// One 64-lane workgroup per (x segment, y, z): the y and z entries are uniform.
__global__ void __launch_bounds__(64) k_tile_reduce(float* __restrict__ rho, const float* __restrict__ tiles,
                                                    const int* __restrict__ tab, int nx, int ny, int nz, int nchan,
                                                    long cstride)
{
    const long M = (long)nx * ny * nz;
    const int gx = blockIdx.x * 64 + threadIdx.x;
    const int gy = blockIdx.y, gz = blockIdx.z;
    if (gx >= nx) return;
    const int* tx = tab + 5 * gx;
    const int* ty = tab + 5 * (nx + gy);
    const int* tz = tab + 5 * (nx + ny + gz);
    const long id = gx + (long)nx * (gy + (long)ny * gz);
    for (int ch = 0; ch < nchan; ++ch) rho[ch * M + id] = tile_sum(tiles + ch * cstride, tz, ty, tx);
}

// ------------------------------------------------------------- gather ---
// Force on each particle from its species' total grid force (potentials and,
// with folded charges, q_s times the field) and, when charges are not uniform
// per species, q times the separate charge field (d_mapGridForcesToPartics and
// d_mapGridChargeForcesToPartics, device_utils.cu:358-461), written per slot
// of the key order.  Direct variant: one thread per particle, stencil loaded
// from the global fields.
template <int P>
__global__ void __launch_bounds__(256) k_gather(float4* __restrict__ fgrid, const float4* __restrict__ F4,
                                                const int* __restrict__ force_slot, int qslot,
                                                const float4* __restrict__ spos, const PMeta* __restrict__ smeta,
                                                int ns, int nx, int ny, int nz, float dx0, float dx1, float dx2)
{
    constexpr int p = P;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    float4 pp = spos[i];
    PMeta meta = smeta[i];
    const long M = (long)nx * ny * nz;
    int fs = force_slot[meta.sp()];
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
                        const float3 v = ff_load(Fs, g);
                        f[0] += v.x * w;
                        f[1] += v.y * w;
                        f[2] += v.z * w;
                    }
                    if (charged) {
                        const float3 v = ff_load(Fq, g);
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
    ff_store(fgrid, i, f[0], f[1], f[2]);
}

// Tile variant.  The force field window of a brick ((b+p)^3 points from
// origin - p/2, periodic) is staged in LDS by 16 lanes per window row, every
// load of a lane issued before its first LDS store.  The window keeps the
// three force components only (12 bytes a point): the stencil sums read a
// quarter fewer LDS bytes than with the padded float4.
struct Win3 {
    float x, y, z;
};

template <int P, int BM>
__device__ __forceinline__ void stage_window(Win3* ft, int* xg, int* yo, int* zo, unsigned char (*rows)[2],
                                             const float4* __restrict__ F, int o0, int o1, int o2, int e0, int e1,
                                             int e2, int nx, int ny, int nz)
{
    constexpr int TP = BM + P, shift = P / 2;
    // a full window that does not wrap (most bricks): its points split by the
    // compile-time TP and address the field directly, with no tables
    const int g0 = o0 - shift, g1 = o1 - shift, g2 = o2 - shift;
    const bool direct = e0 == TP && e1 == TP && e2 == TP && g0 >= 0 && g1 >= 0 && g2 >= 0 && g0 + TP <= nx &&
                        g1 + TP <= ny && g2 + TP <= nz;
    const int nrows = e1 * e2;
    if (!direct) {
        if ((int)threadIdx.x < TP) {
            const int t = threadIdx.x;
            xg[t] = wrap(o0 - shift + t, nx);
            yo[t] = nx * wrap(o1 - shift + t, ny);
            zo[t] = nx * ny * wrap(o2 - shift + t, nz);
        }
        for (int r = threadIdx.x; r < nrows; r += kTileThreads) {
            rows[r][0] = (unsigned char)(r % e1);
            rows[r][1] = (unsigned char)(r / e1);
        }
        __syncthreads();
    }
    // the window's points dealt to every lane in turn (x fastest, so a wave
    // reads a few whole rows); points past the window re-load the last point
    // (in bounds) and skip the store, so the arrays stay in registers.  Row r
    // and column x of point q: r = floor((q + 1/2) / e0) in float, exact for
    // q < 2^20 with e0 <= TP
    constexpr int KMAX = (TP * TP * TP + kTileThreads - 1) / kTileThreads;
    const int npts = nrows * e0;
    const float ie0 = 1.0f / (float)e0;
    float3 v[KMAX];
    int dofs[KMAX];
    if (direct) {
        const int base = g0 + nx * (g1 + ny * g2);
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            const int q = min((int)threadIdx.x + kTileThreads * k, npts - 1);
            const int x = q % TP, t1 = (q / TP) % TP, t2 = q / (TP * TP);
            v[k] = ff_load(F, base + x + nx * (t1 + ny * t2));
            dofs[k] = q;
        }
    } else {
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            const int q = min((int)threadIdx.x + kTileThreads * k, npts - 1);
            const int r = (int)(((float)q + 0.5f) * ie0);
            const int x = q - r * e0;
            const int t1 = rows[r][0], t2 = rows[r][1];
            v[k] = ff_load(F, xg[x] + yo[t1] + zo[t2]);
            dofs[k] = x + TP * (t1 + TP * t2);
        }
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
        if ((int)threadIdx.x + kTileThreads * k < npts) ft[dofs[k]] = Win3{v[k].x, v[k].y, v[k].z};
    __syncthreads();
}

// sum_w w F over the particle's stencil, from the staged window
template <int P, int BM>
__device__ __forceinline__ float3 stencil_sum(const Win3* ft, const int* l, const float* wx, const float* wy,
                                              const float* wz)
{
    constexpr int TP = BM + P;
    const Win3* t = ft + l[0] + TP * (l[1] + TP * l[2]);
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
    for (int iz = 0; iz <= P; ++iz) {
#pragma unroll
        for (int iy = 0; iy <= P; ++iy) {
            const float wyz = wy[iy] * wz[iz];
#pragma unroll
            for (int ix = 0; ix <= P; ++ix) {
                const float w = wx[ix] * wyz;
                const Win3 v = t[ix + TP * (iy + TP * iz)];
                a0 += v.x * w;
                a1 += v.y * w;
                a2 += v.z * w;
            }
        }
    }
    return make_float3(a0, a1, a2);
}

struct GatherMap {
    int slot[kMaxSpecies]; // force field of each species, or -1
};

// One workgroup per (brick, species): the species' particles of the brick are
// contiguous in key order; the workgroup stages that species' force field
// window and writes the particles' grid forces (zero for a species no
// potential acts on).
template <int P, int BM>
__global__ void __launch_bounds__(kTileThreads) k_gather_species(
    float4* __restrict__ fgrid, const float4* __restrict__ F4, GatherMap gm, const float4* __restrict__ spos,
    const PMeta* __restrict__ smeta, const int* __restrict__ start, TileGeom tg, int nx, int ny, int nz, float dx0,
    float dx1, float dx2)
{
    constexpr int TP = BM + P, T3 = TP * TP * TP;
    __shared__ Win3 ft[T3];
    __shared__ int xg[TP], yo[TP], zo[TP];
    __shared__ unsigned char rows[TP * TP][2];
    const int brick = brick_of_block(blockIdx.x, tg.nbricks);
    const int s = blockIdx.y;
    const int key0 = (brick * tg.nspec + s) << 6;
    const int pb = start[key0], pe = start[key0 + 64];
    if (pb == pe) return;
    const int f = gm.slot[s];
    if (f < 0) {
        for (int it = pb + (int)threadIdx.x; it < pe; it += kTileThreads)
            ff_store(fgrid, it, 0.f, 0.f, 0.f);
        return;
    }
    const int k0 = brick % tg.nb[0], k1 = (brick / tg.nb[0]) % tg.nb[1], k2 = brick / (tg.nb[0] * tg.nb[1]);
    const int o0 = k0 << tg.sb[0], o1 = k1 << tg.sb[1], o2 = k2 << tg.sb[2];
    const int e0 = min(tg.b[0], nx - o0) + P, e1 = min(tg.b[1], ny - o1) + P, e2 = min(tg.b[2], nz - o2) + P;
    const long M = (long)nx * ny * nz;
    // the first pass's particle is loaded before the window is staged, so its
    // latency overlaps the window loads (the barriers wait for LDS only)
    const int it0 = pb + (int)threadIdx.x;
    const float4 pp0 = it0 < pe ? spos[it0] : make_float4(0.f, 0.f, 0.f, 0.f);
    stage_window<P, BM>(ft, xg, yo, zo, rows, F4 + f * M, o0, o1, o2, e0, e1, e2, nx, ny, nz);
    const float i0 = 1.0f / dx0, i1 = 1.0f / dx1, i2 = 1.0f / dx2;
    for (int it = it0; it < pe; it += kTileThreads) {
        const float4 pp = it == it0 ? pp0 : spos[it];
        float wx[P + 1], wy[P + 1], wz[P + 1];
        int l[3];
        float3 a = make_float3(0.f, 0.f, 0.f);
        if (stencil_in_brick<P>(pp, dx0, dx1, dx2, i0, i1, i2, nx, ny, nz, o0, o1, o2, tg, wx, wy, wz, l))
            a = stencil_sum<P, BM>(ft, l, wx, wy, wz);
        ff_store(fgrid, it, a.x, a.y, a.z);
    }
}

// Pipelined variant: each workgroup walks a contiguous run of (brick,
// species) tasks and loads the next task's window into registers while it
// sums the current one from LDS, so the window loads overlap the stencil
// sums.  A lane owns window column lane16 and rows row16 + 16 j; their
// periodic grid offsets are computed in registers (no LDS tables).
template <int P, int BM>
__device__ __forceinline__ void window_loads(float3 (&v)[((BM + P) * (BM + P) + 15) / 16], const float4* __restrict__ F,
                                             int o0, int o1, int o2, int e0, int e1, int e2, int nx, int ny, int nz)
{
    constexpr int TP = BM + P, shift = P / 2, RMAX = (TP * TP + 15) / 16;
    const int lane16 = threadIdx.x & 15, row16 = threadIdx.x >> 4;
    const int nrows = e1 * e2;
    const int gx = wrap(o0 - shift + min(lane16, e0 - 1), nx);
#pragma unroll
    for (int j = 0; j < RMAX; ++j) {
        const int r = min(row16 + 16 * j, nrows - 1);
        const int t1 = r % e1, t2 = r / e1;
        const long off = gx + (long)nx * (wrap(o1 - shift + t1, ny) + (long)ny * wrap(o2 - shift + t2, nz));
        v[j] = ff_load(F, off);
    }
}

template <int P, int BM>
__global__ void __launch_bounds__(kTileThreads) k_gather_species_pipe(
    float4* __restrict__ fgrid, const float4* __restrict__ F4, GatherMap gm, const float4* __restrict__ spos,
    const PMeta* __restrict__ smeta, const int* __restrict__ start, TileGeom tg, int nx, int ny, int nz, float dx0,
    float dx1, float dx2, int ntasks, int per)
{
    constexpr int TP = BM + P, T3 = TP * TP * TP, RMAX = (TP * TP + 15) / 16;
    __shared__ Win3 ft[T3];
    const int chunk = brick_of_block(blockIdx.x, gridDim.x);
    const int t0 = chunk * per, t1 = min(t0 + per, ntasks);
    const long M = (long)nx * ny * nz;
    const int lane16 = threadIdx.x & 15, row16 = threadIdx.x >> 4;
    const float i0 = 1.0f / dx0, i1 = 1.0f / dx1, i2 = 1.0f / dx2;
    // the next task that needs a window (has particles and a force field)
    auto next_window = [&](int t) {
        for (; t < t1; ++t) {
            const int brick = t / tg.nspec, sp = t - brick * tg.nspec;
            const int key0 = (brick * tg.nspec + sp) << 6;
            if (start[key0] != start[key0 + 64] && gm.slot[sp] >= 0) return t;
        }
        return t1;
    };
    float3 v[RMAX];
    int tw = next_window(t0);
    auto geom = [&](int t, int* o, int* e) {
        const int brick = t / tg.nspec;
        const int k0 = brick % tg.nb[0], k1 = (brick / tg.nb[0]) % tg.nb[1], k2 = brick / (tg.nb[0] * tg.nb[1]);
        o[0] = k0 << tg.sb[0];
        o[1] = k1 << tg.sb[1];
        o[2] = k2 << tg.sb[2];
        e[0] = min(tg.b[0], nx - o[0]) + P;
        e[1] = min(tg.b[1], ny - o[1]) + P;
        e[2] = min(tg.b[2], nz - o[2]) + P;
    };
    if (tw < t1) {
        int o[3], e[3];
        geom(tw, o, e);
        window_loads<P, BM>(v, F4 + gm.slot[tw % tg.nspec] * M, o[0], o[1], o[2], e[0], e[1], e[2], nx, ny, nz);
    }
    for (int t = t0; t < t1; ++t) {
        const int brick = t / tg.nspec, sp = t - brick * tg.nspec;
        const int key0 = (brick * tg.nspec + sp) << 6;
        const int pb = start[key0], pe = start[key0 + 64];
        if (pb == pe) continue;
        if (gm.slot[sp] < 0) {
            for (int it = pb + (int)threadIdx.x; it < pe; it += kTileThreads)
                ff_store(fgrid, it, 0.f, 0.f, 0.f);
            continue;
        }
        // t == tw: its window is in v
        int o[3], e[3];
        geom(t, o, e);
        __syncthreads();   // the previous window's sums are done
        if (lane16 < e[0]) {
            const int nrows = e[1] * e[2];
#pragma unroll
            for (int j = 0; j < RMAX; ++j) {
                const int r = row16 + 16 * j;
                if (r < nrows) {
                    const int t1r = r % e[1], t2r = r / e[1];
                    ft[lane16 + TP * (t1r + TP * t2r)] = Win3{v[j].x, v[j].y, v[j].z};
                }
            }
        }
        __syncthreads();
        tw = next_window(t + 1);
        if (tw < t1) {
            int on[3], en[3];
            geom(tw, on, en);
            window_loads<P, BM>(v, F4 + gm.slot[tw % tg.nspec] * M, on[0], on[1], on[2], en[0], en[1], en[2], nx,
                                ny, nz);
        }
        for (int it = pb + (int)threadIdx.x; it < pe; it += kTileThreads) {
            const float4 pp = spos[it];
            float wx[P + 1], wy[P + 1], wz[P + 1];
            int l[3];
            float3 a = make_float3(0.f, 0.f, 0.f);
            if (stencil_in_brick<P>(pp, dx0, dx1, dx2, i0, i1, i2, nx, ny, nz, o[0], o[1], o[2], tg, wx, wy, wz, l))
                a = stencil_sum<P, BM>(ft, l, wx, wy, wz);
            ff_store(fgrid, it, a.x, a.y, a.z);
        }
    }
}

// Brick variant: one workgroup per brick stages the force windows of every
// species with a force field (the window tables built once) and writes the
// grid forces of all the brick's particles, whatever their species; fewer,
// fuller workgroups than one per (brick, species).  LDS: nwin windows of
// (b+p)^3 12-byte points (dynamic).
template <int P, int BM>
__global__ void __launch_bounds__(kTileThreads) k_gather_brick(
    float4* __restrict__ fgrid, const float4* __restrict__ F4, GatherMap gm, const float4* __restrict__ spos,
    const PMeta* __restrict__ smeta, const int* __restrict__ start, TileGeom tg, int nx, int ny, int nz, float dx0,
    float dx1, float dx2)
{
    constexpr int TP = BM + P, T3 = TP * TP * TP, shift = P / 2;
    extern __shared__ Win3 ftw[];   // [window][T3]
    __shared__ int xg[TP], yo[TP], zo[TP];
    __shared__ unsigned char rows[TP * TP][2];
    const int brick = brick_of_block(blockIdx.x, tg.nbricks);
    const int pb = start[(brick * tg.nspec) << 6], pe = start[((brick + 1) * tg.nspec) << 6];
    if (pb == pe) return;
    const int k0 = brick % tg.nb[0], k1 = (brick / tg.nb[0]) % tg.nb[1], k2 = brick / (tg.nb[0] * tg.nb[1]);
    const int o0 = k0 << tg.sb[0], o1 = k1 << tg.sb[1], o2 = k2 << tg.sb[2];
    const int e0 = min(tg.b[0], nx - o0) + P, e1 = min(tg.b[1], ny - o1) + P, e2 = min(tg.b[2], nz - o2) + P;
    const long M = (long)nx * ny * nz;
    if ((int)threadIdx.x < TP) {
        const int t = threadIdx.x;
        xg[t] = wrap(o0 - shift + t, nx);
        yo[t] = nx * wrap(o1 - shift + t, ny);
        zo[t] = nx * ny * wrap(o2 - shift + t, nz);
    }
    const int nrows = e1 * e2;
    for (int r = threadIdx.x; r < nrows; r += kTileThreads) {
        rows[r][0] = (unsigned char)(r % e1);
        rows[r][1] = (unsigned char)(r / e1);
    }
    __syncthreads();
    // windows of the species that have force fields, in species order
    constexpr int RMAX = (TP * TP + 15) / 16;
    const int lane16 = threadIdx.x & 15, row16 = threadIdx.x >> 4;
    int nwin = 0;
    for (int sp = 0; sp < tg.nspec; ++sp) {
        const int f = gm.slot[sp];
        if (f < 0) continue;
        Win3* ft = ftw + nwin * T3;
        ++nwin;
        if (lane16 < e0) {
            const float4* src = F4 + f * M;
            const int x0 = xg[lane16];
            Win3* dst = ft + lane16;
            float3 v[RMAX];
            int dofs[RMAX];
#pragma unroll
            for (int j = 0; j < RMAX; ++j) {
                const int r = min(row16 + 16 * j, nrows - 1);
                const int t1 = rows[r][0], t2 = rows[r][1];
                v[j] = ff_load(src, x0 + yo[t1] + zo[t2]);
                dofs[j] = TP * (t1 + TP * t2);
            }
#pragma unroll
            for (int j = 0; j < RMAX; ++j)
                if (row16 + 16 * j < nrows) dst[dofs[j]] = Win3{v[j].x, v[j].y, v[j].z};
        }
    }
    __syncthreads();
    const float i0 = 1.0f / dx0, i1 = 1.0f / dx1, i2 = 1.0f / dx2;
    for (int it = pb + (int)threadIdx.x; it < pe; it += kTileThreads) {
        const float4 pp = spos[it];
        const int sp = smeta[it].sp();
        float3 a = make_float3(0.f, 0.f, 0.f);
        if (gm.slot[sp] >= 0) {
            int w = 0;
            for (int q = 0; q < sp; ++q) w += gm.slot[q] >= 0 ? 1 : 0;
            float wx[P + 1], wy[P + 1], wz[P + 1];
            int l[3];
            if (stencil_in_brick<P>(pp, dx0, dx1, dx2, i0, i1, i2, nx, ny, nz, o0, o1, o2, tg, wx, wy, wz, l))
                a = stencil_sum<P, BM>(ftw + w * T3, l, wx, wy, wz);
        }
        ff_store(fgrid, it, a.x, a.y, a.z);
    }
}

// Separate charge field (charges not uniform per species): one workgroup per
// brick adds q times the field to every charged particle's grid force, after
// the species pass.
template <int P, int BM>
__global__ void __launch_bounds__(kTileThreads) k_gather_charge(
    float4* __restrict__ fgrid, const float4* __restrict__ Fq, const float4* __restrict__ spos,
    const PMeta* __restrict__ smeta, const int* __restrict__ start, TileGeom tg, int nx, int ny, int nz, float dx0,
    float dx1, float dx2)
{
    constexpr int TP = BM + P, T3 = TP * TP * TP;
    __shared__ Win3 ft[T3];
    __shared__ int xg[TP], yo[TP], zo[TP];
    __shared__ unsigned char rows[TP * TP][2];
    const int brick = brick_of_block(blockIdx.x, tg.nbricks);
    const int pb = start[(brick * tg.nspec) << 6], pe = start[((brick + 1) * tg.nspec) << 6];
    if (pb == pe) return;
    const int k0 = brick % tg.nb[0], k1 = (brick / tg.nb[0]) % tg.nb[1], k2 = brick / (tg.nb[0] * tg.nb[1]);
    const int o0 = k0 << tg.sb[0], o1 = k1 << tg.sb[1], o2 = k2 << tg.sb[2];
    const int e0 = min(tg.b[0], nx - o0) + P, e1 = min(tg.b[1], ny - o1) + P, e2 = min(tg.b[2], nz - o2) + P;
    stage_window<P, BM>(ft, xg, yo, zo, rows, Fq, o0, o1, o2, e0, e1, e2, nx, ny, nz);
    const float i0 = 1.0f / dx0, i1 = 1.0f / dx1, i2 = 1.0f / dx2;
    for (int it = pb + (int)threadIdx.x; it < pe; it += kTileThreads) {
        const float4 pp = spos[it];
        if (pp.w == 0.f) continue;
        const int pid = it;
        float wx[P + 1], wy[P + 1], wz[P + 1];
        int l[3];
        if (!stencil_in_brick<P>(pp, dx0, dx1, dx2, i0, i1, i2, nx, ny, nz, o0, o1, o2, tg, wx, wy, wz, l)) continue;
        const float3 a = stencil_sum<P, BM>(ft, l, wx, wy, wz);
        const float3 o = ff_load(fgrid, pid);
        ff_store(fgrid, pid, o.x + pp.w * a.x, o.y + pp.w * a.y, o.z + pp.w * a.z);
    }
}

// ------------------------------------------------- bonds + integrator ---
// Bonds (ps_deviceBonds.cu:7-79) plus GJF (ps_integratorGJF.cu:102-163) or
// EM (ps_integratorEM.cu:39-82) on group "all", one thread per slot of the
// key order.  Writes the next position and noise of the slot and counts the
// next key; integ = MFT_INT_NONE only writes the total force.
__device__ __forceinline__ void bond_force_slots(int s, int ns, const float4* __restrict__ X,
                                                 const int* __restrict__ BS,
                                                 const int* __restrict__ BT, int maxbonds, const BondParams& bp,
                                                 float L0, float L1, float L2, float x0, float x1, float x2,
                                                 float* fo)
{
    for (int b = 0; b < maxbonds; ++b) {
        const int s2 = BS[s * maxbonds + b];
        const int bt = BT ? BT[s * maxbonds + b] : 0;
        if (s2 < 0 || s2 >= ns || bt < 0 || bt >= bp.ntypes) continue;
        const float4 y = X[s2];
        const float d0 = min_image(x0 - y.x, L0, 0.5f * L0);
        const float d1 = min_image(x1 - y.y, L1, 0.5f * L1);
        const float d2 = min_image(x2 - y.z, L2, 0.5f * L2);
        const float r2 = d0 * d0 + d1 * d1 + d2 * d2;
        float mf = 0.f;
        if (r2 > 1.0e-4f) {
            const float r = sqrtf(r2);
            if (bp.style[bt] == 0) mf = 2.0f * bp.k[bt] * (r - bp.r0[bt]) / r;
            else if (bp.style[bt] == 1) {
                const float u = r / bp.r0[bt];
                mf = 2.0f * bp.k[bt] * r / (1.0f - u * u) / r;
            }
        }
        fo[0] -= mf * d0;
        fo[1] -= mf * d1;
        fo[2] -= mf * d2;
    }
}

// One bond of bond_force_slots with the partner's position given
// one bond's force with its type's parameters (style, k, r0) already loaded
__device__ __forceinline__ void bond_force_params(int style, float kb, float r0, float L0, float L1, float L2,
                                                  float4 x, float4 y, float* fo)
{
    const float d0 = min_image(x.x - y.x, L0, 0.5f * L0);
    const float d1 = min_image(x.y - y.y, L1, 0.5f * L1);
    const float d2 = min_image(x.z - y.z, L2, 0.5f * L2);
    const float r2 = d0 * d0 + d1 * d1 + d2 * d2;
    float mf = 0.f;
    if (r2 > 1.0e-4f) {
        const float r = sqrtf(r2);
        if (style == 0) mf = 2.0f * kb * (r - r0) / r;
        else if (style == 1) {
            const float u = r / r0;
            mf = 2.0f * kb * r / (1.0f - u * u) / r;
        }
    }
    fo[0] -= mf * d0;
    fo[1] -= mf * d1;
    fo[2] -= mf * d2;
}

__device__ __forceinline__ void bond_force_pair(const BondParams& bp, int bt, float L0, float L1, float L2,
                                                float4 x, float4 y, float* fo)
{
    bond_force_params(bp.style[bt], bp.k[bt], bp.r0[bt], L0, L1, L2, x, y, fo);
}

// One angle of slot s (d_angles, ps_deviceAngles.cu:8-133): the force on
// whichever member s is, and the angle's energy (d_anglesStressEnergy,
// ps_deviceAngles.cu:138-290); false when a bond vector is too short.
__device__ __forceinline__ bool angle_terms(int s, const float4* __restrict__ X, int ai, int aj, int ak, int t,
                                            const AngleParams& ap, float L0, float L1, float L2, float* fs,
                                            float* energy)
{
    const float Ls[3] = {L0, L1, L2};
    const float4 xi = X[ai], xj = X[aj], xk = X[ak];
    const float pi[3] = {xi.x, xi.y, xi.z}, pj[3] = {xj.x, xj.y, xj.z}, pk[3] = {xk.x, xk.y, xk.z};
    float rij[3], rkj[3], mrij2 = 0.0f, mrkj2 = 0.0f;
    for (int j = 0; j < 3; ++j) {
        rij[j] = min_image(pi[j] - pj[j], Ls[j], 0.5f * Ls[j]);
        mrij2 += rij[j] * rij[j];
    }
    if (mrij2 < 1.0E-4f) return false;
    const float mrij = sqrtf(mrij2);
    for (int j = 0; j < 3; ++j) {
        rkj[j] = min_image(pk[j] - pj[j], Ls[j], 0.5f * Ls[j]);
        mrkj2 += rkj[j] * rkj[j];
    }
    if (mrkj2 < 1.0E-4f) return false;
    const float mrkj = sqrtf(mrkj2);
    float dot = 0.0f;
    for (int j = 0; j < 3; ++j) dot += rij[j] * rkj[j];
    const float cos_th = dot / mrij / mrkj;
    float fi[3], fk[3];
    if (ap.style[t] == 0) {
        // WLC: k (1 + cos theta)
        const float kang = ap.k[t];
        for (int j = 0; j < 3; ++j) {
            fi[j] = -kang * (rkj[j] / mrij / mrkj - cos_th * rij[j] / mrij2);
            fk[j] = -kang * (rij[j] / mrkj / mrij - cos_th * rkj[j] / mrkj2);
        }
        *energy = (float)((double)kang * (1.0 + (double)cos_th));
    } else {
        // harmonic: k (theta - theta_eq)^2
        if (cos_th < -1.f || cos_th > 1.f) return false;
        const float theta = acosf(cos_th);
        const float denom = sqrtf(1.f - cos_th * cos_th);
        const float iSinTheta = denom > 1.0E-4f ? (float)(1.0 / (double)denom) : 0.f;
        const float dtheta = theta - ap.theta[t];
        const float du_mag = 2.f * ap.k[t] * dtheta;
        for (int j = 0; j < 3; ++j) {
            fi[j] = du_mag * iSinTheta * (rkj[j] / mrij / mrkj - cos_th * rij[j] / mrij2);
            fk[j] = du_mag * iSinTheta * (rij[j] / mrij / mrkj - cos_th * rkj[j] / mrkj2);
        }
        *energy = ap.k[t] * dtheta * dtheta;
    }
    for (int j = 0; j < 3; ++j) {
        if (s == ai) fs[j] = fi[j];
        else if (s == aj) fs[j] = -(fi[j] + fk[j]);
        else if (s == ak) fs[j] = fk[j];
        else fs[j] = 0.f;
    }
    return true;
}

__device__ __forceinline__ void angle_force_slots(int s, const float4* __restrict__ X, const int* __restrict__ NA,
                                                  const int* __restrict__ AT, const int* __restrict__ AG,
                                                  const AngleParams& ap, float L0, float L1, float L2, float* fo)
{
    const int ma = ap.maxangles, na = min(NA[s], ma);
    float lf[3] = {0.f, 0.f, 0.f};
    for (int a = 0; a < na; ++a) {
        const int* g = AG + (s * ma + a) * 3;
        float fs[3], en;
        if (!angle_terms(s, X, g[0], g[1], g[2], AT[s * ma + a], ap, L0, L1, L2, fs, &en)) continue;
        lf[0] += fs[0];
        lf[1] += fs[1];
        lf[2] += fs[2];
    }
    fo[0] += lf[0];
    fo[1] += lf[1];
    fo[2] += lf[2];
}

struct StepArgs {
    const float4* XO;
    const float4* NZ;
    float4* XN;
    float4* NN;
    GjfParams gp;
    int integ;
    float delt;
    float L[3];
    uint64_t seed;
    uint32_t step;
    float dx[3];
    int p, nx, ny, nz;
    WallParams walls;
    const int *NA, *AT, *AG;   // angles (null without)
    AngleParams ap;
    int vv;                    // velocity Verlet: 0 before the first step (TildEngine::vv_state)
    int use_v;                 // the noise slot holds velocities (VV)
    const float4* FD;          // DPD pair forces per slot (null without)
    LangevinParams lang;
    int nz_keyed;              // GJF: previous noise regenerated from (particle, step - 1), NZ not read
    int write_nn;              // store the new noise in NN (off under GJF: the next step regenerates it)
};

// Wall forces of ps_potentialWall.cu:92-108 on a particle of species ty
__device__ __forceinline__ void wall_forces(const WallParams& W, int ty, float4 xc, float* fo)
{
    for (int w = 0; w < W.n; ++w) {
        if (!((W.mask[w] >> ty) & 1u)) continue;
        const int a = W.axis[w];
        const float disp = W.pos[w] - (a == 0 ? xc.x : (a == 1 ? xc.y : xc.z));
        if ((float)W.sign[w] * disp > 0.0f) fo[a] += W.k[w] * disp;
    }
}

// d_add_Langevin_forces (ps_potentialLangevin.cu:101-133) on particle id of
// species ty with velocity v: one normal per axis, keyed by (particle, step,
// 'LANG', potential + 1)
__device__ __forceinline__ void langevin_forces(const LangevinParams& P, int id, int ty, const float* v,
                                                uint64_t seed, uint32_t step, float* fo)
{
    for (int l = 0; l < P.n; ++l) {
        if (!((P.mask[l] >> ty) & 1u)) continue;
        float z[4];
        normal4((uint32_t)id, step, TAG_LANGEVIN, (uint32_t)(l + 1), seed, z);
        for (int j = 0; j < 3; ++j) fo[j] += -v[j] * P.drag[l] + P.nmag[l] * z[j];
    }
}

// GJF (ps_integratorGJF.cu:102-163) or EM (ps_integratorEM.cu:39-82) update
// of slot s from its total force fo: writes the next position and noise and
// returns the next key.
// Per-species integrator coefficients of one particle, read (per lane, so
// from memory) before the integrator needs them: GJF a, b, noise magnitude,
// force and noise factors; VV's 1/m in a.
struct SpeciesCoef {
    float a, b, nm, cf, cn;
};

__device__ __forceinline__ SpeciesCoef species_coef(const StepArgs& A, int ty)
{
    // every table read unconditionally (no branch between the loads and
    // their use), the integrator's own picked after
    const float a = A.gp.a[ty], b = A.gp.b[ty], nm = A.gp.nmag[ty], cf = A.gp.cf[ty], cn = A.gp.cn[ty];
    const float im = A.gp.inv_m[ty];
    const bool vv = A.integ == MFT_INT_VV;
    return SpeciesCoef{vv ? im : a, b, nm, cf, cn};
}

__device__ __forceinline__ int integrate_slot(int s, float4 xc, PMeta meta, const float* fo, const StepArgs& A,
                                              uint32_t step, const TileGeom& tg, float4 xo4, float4 nz4,
                                              const SpeciesCoef& sc)
{
    float z[4];
    normal4((uint32_t)meta.id(), step, TAG_TILD, 0u, A.seed, z);
    const float xcv[3] = {xc.x, xc.y, xc.z};
    float xn[3], nn[3] = {0.f, 0.f, 0.f};
    if (A.integ == MFT_INT_GJF) {
        const float a = sc.a, b = sc.b, nm = sc.nm, cf = sc.cf, cn = sc.cn;
        float nov[3];
        if (A.nz_keyed) {
            // the previous step's nm * normal, the same bits as when it was drawn
            float zo[4];
            normal4((uint32_t)meta.id(), step - 1u, TAG_TILD, 0u, A.seed, zo);
            for (int j = 0; j < 3; ++j) nov[j] = nm * zo[j];
        } else {
            nov[0] = nz4.x;
            nov[1] = nz4.y;
            nov[2] = nz4.z;
        }
        const float xov[3] = {xo4.x, xo4.y, xo4.z};
        for (int j = 0; j < 3; ++j) {
            const float nnj = nm * z[j];
            float xo = xov[j];
            if (xo - xcv[j] > A.L[j] / 2.0f) xo -= A.L[j];
            else if (xo - xcv[j] < -A.L[j] / 2.0f) xo += A.L[j];
            float v = 2.0f * b * xcv[j] - a * xo + cf * fo[j] + cn * (nnj + nov[j]);
            if (v > A.L[j]) v -= A.L[j];
            else if (v < 0.0f) v += A.L[j];
            xn[j] = v;
            nn[j] = nnj;
        }
    } else if (A.integ == MFT_INT_VV) {
        // The reference's step k is Integrate_1 (half kick with the stored
        // force, drift), the forces F_k at the new positions, Integrate_2
        // (half kick with F_k; ps_integratorVV.cu:31-94).  Step k of this pass
        // has the forces at its current positions, which are F_(k-1): it
        // applies step k-1's Integrate_2, then step k's Integrate_1, and keeps
        // the half-kicked velocity.  Step 0 starts from rest with the zero force
        // array of a fresh allocation: neither kick nor drift changes anything.
        const float inv_m = sc.a;
        const float hd = A.delt * 0.5f;
        const float vv0[3] = {nz4.x, nz4.y, nz4.z};
        for (int j = 0; j < 3; ++j) {
            float v = vv0[j];
            float x = xcv[j];
            if (A.vv != 0) {
                v += hd * inv_m * fo[j];
                v += hd * inv_m * fo[j];
                x = xcv[j] + (hd * 2.0f) * v;
                if (x >= A.L[j]) x -= A.L[j];
                else if (x < 0.0f) x += A.L[j];
            }
            xn[j] = x;
            nn[j] = v;
        }
    } else {
        const float nm = sqrtf(2.0f * A.delt);
        for (int j = 0; j < 3; ++j) {
            float v = xcv[j] + A.delt * fo[j] + nm * z[j];
            if (v > A.L[j]) v -= A.L[j];
            else if (v < 0.0f) v += A.L[j];
            xn[j] = v;
        }
    }
    ff_store(A.XN, s, xn[0], xn[1], xn[2]);
    if (A.write_nn) A.NN[s] = make_float4(nn[0], nn[1], nn[2], 0.f);
    return key_of(xn[0], xn[1], xn[2], meta.sp(), A.dx[0], A.dx[1], A.dx[2], A.p, A.nx, A.ny, A.nz, tg);
}

// Bonds (ps_deviceBonds.cu:7-79) plus the integrator on group "all", one
// thread per slot of the key order.  Writes the next position and noise of
// the slot and counts the next key; integ = MFT_INT_NONE only writes the
// total force.  TWO: at most two bonds per particle (the slots' bond lists
// are loaded with the slot's own records, so a slot waits for two rounds of
// memory latency: its records, then its partners' positions).  dstep: the
// step index on the device (graph replays), else null.
template <bool TWO>
__global__ void __launch_bounds__(256) k_integrate(
    const float4* __restrict__ X, const PMeta* __restrict__ META, const int* __restrict__ BS,
    const int* __restrict__ BT, const float4* __restrict__ fgrid,
    float4* __restrict__ ftot, int maxbonds, BondParams bp, StepArgs A, int ns, int* __restrict__ count,
    int* __restrict__ key, int* __restrict__ rank, TileGeom tg, const uint32_t* __restrict__ dstep)
{
    // contiguous slot ranges per XCD (brick_of_block): bond partners, mostly
    // a few slots away in key order, are read from the L2 that holds the
    // block's own slots rather than from another XCD's lines
    const int s = brick_of_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool active = s < ns;
    int k = 0;
    if (active) {
        // every load that does not depend on another issued first, the bond
        // partners' positions next: two rounds of memory latency per slot
        const float4 xc = X[s];
        const float3 fg = ff_load(fgrid, s);
        const PMeta meta = META[s];
        const bool gjf = A.integ == MFT_INT_GJF;
        const bool need_nz = (gjf && !A.nz_keyed) || A.integ == MFT_INT_VV || A.use_v;
        const float3 xo3 = gjf ? ff_load(A.XO, s) : make_float3(0.f, 0.f, 0.f);
        const float4 xo4 = make_float4(xo3.x, xo3.y, xo3.z, 0.f);
        const float4 nz4 = need_nz ? A.NZ[s] : make_float4(0.f, 0.f, 0.f, 0.f);
        int2 bs = make_int2(-1, -1), bt = make_int2(0, 0);
        if constexpr (TWO) {
            bs = reinterpret_cast<const int2*>(BS)[s];
            if (BT) bt = reinterpret_cast<const int2*>(BT)[s];
        }
        const uint32_t step = dstep ? *dstep : A.step;
        float fo[3] = {fg.x, fg.y, fg.z};
        SpeciesCoef sco;
        if constexpr (TWO) {
            const bool v0 = (bs.x >= 0) & (bs.x < ns) & (bt.x >= 0) & (bt.x < bp.ntypes);
            const bool v1 = (bs.y >= 0) & (bs.y < ns) & (bt.y >= 0) & (bt.y < bp.ntypes);
            // the bond types' parameters (indexed per lane, so read from
            // memory) are loaded with the partners' positions, not after them
            const int t0 = v0 ? bt.x : 0, t1 = v1 ? bt.y : 0;
            const int st0 = bp.style[t0], st1 = bp.style[t1];
            const float k0 = bp.k[t0], k1 = bp.k[t1], q0 = bp.r0[t0], q1 = bp.r0[t1];
            const float4 y0 = X[v0 ? bs.x : s];
            const float4 y1 = X[v1 ? bs.y : s];
            // the species' coefficients in the same round of loads
            sco = species_coef(A, meta.sp());
            if (v0) bond_force_params(st0, k0, q0, A.L[0], A.L[1], A.L[2], xc, y0, fo);
            if (v1) bond_force_params(st1, k1, q1, A.L[0], A.L[1], A.L[2], xc, y1, fo);
        } else {
            sco = species_coef(A, meta.sp());
            bond_force_slots(s, ns, X, BS, BT, maxbonds, bp, A.L[0], A.L[1], A.L[2], xc.x, xc.y, xc.z, fo);
        }
        if (A.NA) angle_force_slots(s, X, A.NA, A.AT, A.AG, A.ap, A.L[0], A.L[1], A.L[2], fo);
        if (A.walls.n) wall_forces(A.walls, meta.sp(), xc, fo);
        if (A.FD) {
            const float4 fd = A.FD[s];
            fo[0] += fd.x;
            fo[1] += fd.y;
            fo[2] += fd.z;
        }
        if (A.lang.n) {
            // the velocities the forces see: VV's (the half-kicked ones the
            // reference holds at force time), zero under GJF and EM
            float v[3] = {0.f, 0.f, 0.f};
            if (A.use_v) {
                v[0] = nz4.x;
                v[1] = nz4.y;
                v[2] = nz4.z;
            }
            // under VV these are the forces of the previous step (its noise)
            const uint32_t ls = (A.use_v && A.vv != 0) ? step - 1u : step;
            langevin_forces(A.lang, meta.id(), meta.sp(), v, A.seed, ls, fo);
        }
        if (ftot) ftot[s] = make_float4(fo[0], fo[1], fo[2], 0.f);
        if (A.integ != MFT_INT_NONE) k = integrate_slot(s, xc, meta, fo, A, step, tg, xo4, nz4, sco);
    }
    if (A.integ == MFT_INT_NONE) return; // uniform over the grid
    const int r = wave_rank(count, k, active);
    if (active) {
        key[s] = k;
        rank[s] = r;
    }
}

// d_bondStressEnergy virial part (ps_deviceBonds.cu:85-171): per particle the
// six terms -mf dr_a dr_b (xx, yy, zz, xy, xz, yz) summed over its bonds in
// float, written in particle-id order (the layout of the reference's
// d_bondVirScratch); dr with the 0.5*Lh minimum image (:132-133); no term for
// r^2 <= 1e-4 or for FENE at or beyond r0, where the reference leaves mf unset.
__global__ void k_bond_virial_slots(float* __restrict__ vir, const float4* __restrict__ X,
                                    const PMeta* __restrict__ META, const int* __restrict__ BS,
                                    const int* __restrict__ BT, BondParams bp, float L0,
                                    float L1, float L2, int ns, int maxbonds)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const float Ls[3] = {L0, L1, L2};
    const float4 x = X[s];
    const float xv[3] = {x.x, x.y, x.z};
    float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int b = 0; b < maxbonds; ++b) {
        const int s2 = BS[s * maxbonds + b];
        if (s2 < 0) continue;
        const int bt = BT ? BT[s * maxbonds + b] : 0;
        const float4 y = X[s2];
        const float yv[3] = {y.x, y.y, y.z};
        float d[3], r2 = 0.f;
        for (int j = 0; j < 3; ++j) {
            d[j] = min_image(xv[j] - yv[j], Ls[j], 0.25f * Ls[j]);
            r2 += d[j] * d[j];
        }
        float mf = 0.f;
        if (r2 > 1.0e-4f) {
            const float r = sqrtf(r2);
            if (bp.style[bt] == 0) mf = 2.0f * bp.k[bt] * (r - bp.r0[bt]) / r;
            else if (bp.style[bt] == 1 && r / bp.r0[bt] < 1.f) {
                const float u = r / bp.r0[bt];
                mf = 2.0f * bp.k[bt] * (1.f / (1.f - u * u));
            }
        }
        v[0] += -mf * d[0] * d[0];
        v[1] += -mf * d[1] * d[1];
        v[2] += -mf * d[2] * d[2];
        v[3] += -mf * d[0] * d[1];
        v[4] += -mf * d[0] * d[2];
        v[5] += -mf * d[1] * d[2];
    }
    const long id = META[s].id();
    for (int c = 0; c < 6; ++c) vir[id * 6 + c] = v[c];
}

int tild_bond_virial(TildEngine* e, float* vir)
{
    const int c = e->cur;
    k_bond_virial_slots<<<nblk(e->N), kBlock, 0, e->stream>>>(vir, e->X[c], e->META[c], e->BS[c],
                                                               slot_bond_types(e, c), e->bp, e->d.lx, e->d.ly, e->d.lz, e->N,
                                                               std::max(1, e->d.maxbonds));
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// d_bondStressEnergy energy part (ps_deviceBonds.cu:85-171) over slots, FP64
// block sums, with the reference's 0.5*Lh minimum image (:132-133).
__global__ void k_bond_energy_slots(double* __restrict__ partial, const float4* __restrict__ X,
                                    const int* __restrict__ BS,
                                    const int* __restrict__ BT, BondParams bp, float L0, float L1, float L2, int ns,
                                    int maxbonds)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (s < ns) {
        const float Ls[3] = {L0, L1, L2};
        const float4 x = X[s];
        const float xv[3] = {x.x, x.y, x.z};
        for (int b = 0; b < maxbonds; ++b) {
            const int s2 = BS[s * maxbonds + b];
            if (s2 < 0) continue;
            const int bt = BT ? BT[s * maxbonds + b] : 0;
            const float4 y = X[s2];
            const float yv[3] = {y.x, y.y, y.z};
            float r2 = 0.f;
            for (int j = 0; j < 3; ++j) {
                const float d = min_image(xv[j] - yv[j], Ls[j], 0.25f * Ls[j]);
                r2 += d * d;
            }
            if (r2 > 1.0e-4f) {
                const float r = sqrtf(r2);
                if (bp.style[bt] == 0) {
                    const float a = r - bp.r0[bt];
                    e += (double)(bp.k[bt] * a * a);
                } else if (bp.style[bt] == 1 && r / bp.r0[bt] < 1.f) {
                    const float u = r / bp.r0[bt];
                    const float arg = 1.f / (1.f - u * u);
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

// ----------------------------------------------------------- launchers ---
static int pow2_floor(int v, int cap)
{
    int b = 1;
    while (b * 2 <= v && b * 2 <= cap) b *= 2;
    return b;
}

static int log2i(int v)
{
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

// Brick geometry, key space, LDS budgets and the per-axis tables of tile
// offsets covering each grid coordinate.  Bricks are 8^3 points for up to 4
// density channels, 4^3 beyond.
int tild_tiles_configure(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    const int p = d.pmeorder, shift = p / 2;
    const int BM = e->nchan <= 4 ? 8 : 4;
    TileGeom& tg = e->tg;
    const int n[3] = {d.nx, d.ny, d.nz};
    for (int a = 0; a < 3; ++a) {
        tg.b[a] = pow2_floor(n[a], BM);
        tg.sb[a] = log2i(tg.b[a]);
        tg.nb[a] = (n[a] + tg.b[a] - 1) / tg.b[a];
    }
    tg.lbmax = log2i(BM);
    tg.pitch = BM + p;
    tg.nbricks = tg.nb[0] * tg.nb[1] * tg.nb[2];
    tg.nspec = d.nspecies;
    for (int a = 0; a < 3; ++a) tg.ex[a] = tg.nb[a] * tg.pitch;
    const long nkeys = ((long)tg.nbricks * tg.nspec) << 6;
    const long T3 = (long)tg.pitch * tg.pitch * tg.pitch;
    MFT_REQUIRE(nkeys < (1L << 30) && (long)tg.nbricks * e->nchan * T3 < (1L << 31), MFT_ERR_UNSUPPORTED,
                "grid too large for the brick decomposition");
    e->nkeys = (int)nkeys;
    e->dens_lds = sizeof(unsigned long long) * (size_t)e->nchan * T3;
    MFT_REQUIRE(e->dens_lds <= 64 * 1024, MFT_ERR_UNSUPPORTED, "density tile exceeds LDS");

    // per-axis covering tiles: brick k covers coordinates k b - shift + t,
    // t in [0, len_k + p), periodically; its entry t sits at k pitch + t of
    // the halo-expanded grid along that axis
    std::vector<int> tab(5 * (size_t)(n[0] + n[1] + n[2]), 0);
    const long estride[3] = {1, tg.ex[0], (long)tg.ex[0] * tg.ex[1]};
    int base = 0, maxcover = 0;
    for (int a = 0; a < 3; ++a) {
        for (int X = 0; X < n[a]; ++X) {
            int* row = &tab[5 * (size_t)(base + X)];
            int cnt = 0;
            for (int k = 0; k < tg.nb[a]; ++k) {
                const int len = std::min(tg.b[a], n[a] - k * tg.b[a]) + p;
                int t0 = ((X - k * tg.b[a] + shift) % n[a] + n[a]) % n[a];
                for (int t = t0; t < len; t += n[a]) {
                    MFT_REQUIRE(cnt < 4, MFT_ERR_UNSUPPORTED, "tile cover table overflow");
                    row[1 + cnt++] = (int)(((long)k * tg.pitch + t) * estride[a]);
                }
            }
            row[0] = cnt;
            maxcover = std::max(maxcover, cnt);
        }
        base += n[a];
    }
    e->tile_maxcover = maxcover;
    if (e->tile_tab) (void)hipFree(e->tile_tab);
    e->tile_tab = nullptr;
    int rc;
    if ((rc = dalloc(&e->tile_tab, tab.size()))) return rc;
    MFT_HIP(hipMemcpy(e->tile_tab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
    if (e->tiles) (void)hipFree(e->tiles);
    e->tiles = nullptr;
    if ((rc = dalloc(&e->tiles, (size_t)tg.nbricks * e->nchan * T3))) return rc;
    // key-space arrays
    for (void* pbuf : {(void*)e->count, (void*)e->start, (void*)e->partial})
        if (pbuf) (void)hipFree(pbuf);
    e->count = e->start = e->partial = nullptr;
    if ((rc = dalloc(&e->count, (size_t)e->nkeys))) return rc;
    if ((rc = dalloc(&e->start, (size_t)e->nkeys + 1))) return rc;
    e->nscan_blocks = (e->nkeys + kScanChunk - 1) / kScanChunk;
    if ((rc = dalloc(&e->partial, (size_t)e->nscan_blocks))) return rc;
    MFT_HIP(hipMemset(e->count, 0, sizeof(int) * (size_t)e->nkeys));
    e->have_state = 0; // the key order changed: the state is sorted again when set
    return MFT_OK;
}

// particle-id-order staging (pio: positions, pio2: previous positions, pio3:
// previous noise) -> key order, with counts, starts and bond slots
int tild_sort_state(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    const int N = e->N, mb = std::max(1, d.maxbonds), c = e->cur;
    e->nz_keyed = false; // NZ holds the given previous noise
    MFT_HIP(hipMemsetAsync(e->count, 0, sizeof(int) * (size_t)e->nkeys, e->stream));
    k_sort_count<<<nblk(N), kBlock, 0, e->stream>>>(e->pio, e->type, N, e->count, e->key, e->rank, e->dx[0],
                                                     e->dx[1], e->dx[2], d.pmeorder, d.nx, d.ny, d.nz, e->tg);
    MFT_LAUNCH_CHECK();
    int rc;
    if ((rc = tild_scan(e))) return rc;
    k_sort_place<<<nblk(N), kBlock, 0, e->stream>>>(e->pio, e->pio2, e->pio3, e->charge, e->type,
                                                     e->bondType, mb, N, e->key, e->rank, e->start, e->slot_of,
                                                     e->X[c], e->XO[c], e->NZ[c], e->META[c], e->BT[c]);
    k_sort_bonds<<<nblk(N), kBlock, 0, e->stream>>>(e->bondedTo, e->nBonds, mb, N, e->slot_of, e->BS[c]);
    if (d.maxangles > 0)
        k_sort_angles<<<nblk(N), kBlock, 0, e->stream>>>(e->nAngles, e->angleType, e->angleGroup, d.maxangles, N,
                                                         e->slot_of, e->NA[c], e->AT[c], e->AG[c]);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// previous GJF noise in particle-id order, regenerated from (particle,
// step - 1): out[3 id + j] = nmag(type) * normal_j
__global__ void k_unsort_keyed_noise(float* __restrict__ out, const PMeta* __restrict__ META, GjfParams gp, int ns,
                                     uint64_t seed, uint32_t prev_step)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const PMeta m = META[s];
    const int id = m.id();
    float z[4];
    normal4((uint32_t)id, prev_step, TAG_TILD, 0u, seed, z);
    const float nm = gp.nmag[m.sp()];
    out[id * 3 + 0] = nm * z[0];
    out[id * 3 + 1] = nm * z[1];
    out[id * 3 + 2] = nm * z[2];
}

// key order -> particle-id order into pio: what 0 positions, 1 previous
// positions, 2 previous noise, 3 total force of the last evaluation
int tild_unsort(TildEngine* e, int what)
{
    const int c = e->cur;
    if (what == 2 && e->nz_keyed) {
        k_unsort_keyed_noise<<<nblk(e->N), kBlock, 0, e->stream>>>(e->pio, e->META[c], e->gp, e->N, e->d.seed,
                                                                   (uint32_t)e->step - 1u);
        MFT_LAUNCH_CHECK();
        return MFT_OK;
    }
    const float4* src = what == 0 ? e->X[c] : what == 1 ? e->XO[c] : what == 2 ? e->NZ[c] : e->ftot;
    k_unsort<<<nblk(e->N), kBlock, 0, e->stream>>>(e->pio, src, e->META[c], e->N, what == 1 ? 1 : 0);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int tild_scan(TildEngine* e, bool advance_step)
{
    uint32_t* ds = (advance_step && e->use_dstep) ? e->d_step : nullptr;
    k_scan_reduce<<<e->nscan_blocks, kScanThreads, 0, e->stream>>>(e->count, e->nkeys, e->partial);
    k_scan_apply<<<e->nscan_blocks, kScanThreads, 0, e->stream>>>(e->count, e->nkeys, e->partial, e->start, e->N,
                                                                    ds);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// next positions into the next key order (after the integrator and the scan)
int tild_permute(TildEngine* e)
{
    const int N = e->N, mb = std::max(1, e->d.maxbonds), c = e->cur, n = 1 - c;
    // under GJF the next step regenerates the previous noise: NZ stays behind
    const float4* nn = e->d.integrator == MFT_INT_GJF ? nullptr : e->NN;
    k_permute<<<nblk(N), kBlock, 0, e->stream>>>(N, mb, e->key, e->rank, e->start, e->map, e->XN, nn, e->X[c],
                                                  e->META[c], slot_bond_types(e, c), e->X[n], e->XO[n], e->NZ[n],
                                                  e->META[n], slot_bond_types(e, n));
    k_remap_bonds<<<nblk(N), kBlock, 0, e->stream>>>(N, mb, e->map, e->BS[c], e->BS[n]);
    if (e->d.maxangles > 0)
        k_permute_angles<<<nblk(N), kBlock, 0, e->stream>>>(N, e->d.maxangles, e->map, e->NA[c], e->AT[c], e->AG[c],
                                                            e->NA[n], e->AT[n], e->AG[n]);
    MFT_LAUNCH_CHECK();
    e->cur = n;
    return MFT_OK;
}

int tild_density(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    const TileGeom& tg = e->tg;
    const int BM = 1 << tg.lbmax;
    const double inv_gvol = 1.0 / (double)e->gvol;
#define MFT_DENS(P_, B_)                                                                                  \
    k_density_tile<P_, B_><<<tg.nbricks, kTileThreads, e->dens_lds, e->stream>>>(                         \
        e->tiles, e->nchan, e->qchan, e->X[e->cur], e->start, tg, d.nx, d.ny, d.nz, e->dx[0], e->dx[1], \
        e->dx[2], inv_gvol)
#define MFT_DENS_B(P_)                                                                                    \
    if (BM == 8) MFT_DENS(P_, 8);                                                                         \
    else MFT_DENS(P_, 4);
    switch (d.pmeorder) {
    case 0: MFT_DENS_B(0); break;
    case 1: MFT_DENS_B(1); break;
    case 2: MFT_DENS_B(2); break;
    case 3: MFT_DENS_B(3); break;
    default: MFT_DENS_B(4); break;
    }
#undef MFT_DENS_B
#undef MFT_DENS
    MFT_LAUNCH_CHECK();
    e->rho_valid = false;
    return MFT_OK;
}

int tild_density_reduce(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    const TileGeom& tg = e->tg;
    const long cstride = (long)tg.ex[0] * tg.ex[1] * tg.ex[2];
    dim3 rgrid((d.nx + 63) / 64, d.ny, d.nz);
    k_tile_reduce<<<rgrid, 64, 0, e->stream>>>(e->rho, e->tiles, e->tile_tab, d.nx, d.ny, d.nz, e->nchan, cstride);
    MFT_LAUNCH_CHECK();
    e->rho_valid = true;
    return MFT_OK;
}

// LDS bytes of the brick gather's windows (every species with a force field)
static size_t gather_brick_lds(const TildEngine* e)
{
    const TileGeom& tg = e->tg;
    const int TP = (1 << tg.lbmax) + e->d.pmeorder;
    int nwin = 0;
    for (int sp = 0; sp < e->d.nspecies; ++sp) nwin += e->force_slot_h[sp] >= 0 ? 1 : 0;
    return sizeof(Win3) * (size_t)nwin * TP * TP * TP;
}

bool tild_gather_brick_fits(const TildEngine* e) { return gather_brick_lds(e) <= 64 * 1024; }

// gather variants: 0 = direct loads in key order, 1 = LDS windows per (brick,
// species), 2 = LDS windows of every species per brick, 3 = as 1 with each
// workgroup walking a run of (brick, species) tasks, the next window's loads
// in flight during the current sums
int tild_gather(TildEngine* e, int mode)
{
    const mft_tild_desc& d = e->d;
    if (mode == 2 && !tild_gather_brick_fits(e)) mode = 1;
    if (mode == 3) {
        const TileGeom& tg = e->tg;
        const int BM = 1 << tg.lbmax;
        GatherMap gm;
        for (int sp = 0; sp < kMaxSpecies; ++sp) gm.slot[sp] = sp < d.nspecies ? e->force_slot_h[sp] : -1;
        const int ntasks = tg.nbricks * tg.nspec;
        const int nwg = std::min(ntasks, 256 * 8);
        const int per = (ntasks + nwg - 1) / nwg;
        const int grid = (ntasks + per - 1) / per;
        const float4* Fq = e->F4 + (size_t)(e->qslot >= 0 ? e->qslot : 0) * e->M;
#define MFT_G3(P_, B_)                                                                                    \
    do {                                                                                                  \
        k_gather_species_pipe<P_, B_><<<grid, kTileThreads, 0, e->stream>>>(                              \
            e->fgrid, e->F4, gm, e->X[e->cur], e->META[e->cur], e->start, tg, d.nx, d.ny, d.nz, e->dx[0], e->dx[1], \
            e->dx[2], ntasks, per);                                                                       \
        if (e->qslot >= 0)                                                                                \
            k_gather_charge<P_, B_><<<tg.nbricks, kTileThreads, 0, e->stream>>>(                          \
                e->fgrid, Fq, e->X[e->cur], e->META[e->cur], e->start, tg, d.nx, d.ny, d.nz, e->dx[0], e->dx[1], e->dx[2]); \
    } while (0)
#define MFT_G3_B(P_)                                                                                      \
    if (BM == 8) MFT_G3(P_, 8);                                                                           \
    else MFT_G3(P_, 4);
        switch (d.pmeorder) {
        case 0: MFT_G3_B(0); break;
        case 1: MFT_G3_B(1); break;
        case 2: MFT_G3_B(2); break;
        case 3: MFT_G3_B(3); break;
        default: MFT_G3_B(4); break;
        }
#undef MFT_G3_B
#undef MFT_G3
        MFT_LAUNCH_CHECK();
        return MFT_OK;
    }
    if (mode == 2) {
        const TileGeom& tg = e->tg;
        const int BM = 1 << tg.lbmax;
        GatherMap gm;
        for (int sp = 0; sp < kMaxSpecies; ++sp) gm.slot[sp] = sp < d.nspecies ? e->force_slot_h[sp] : -1;
        const size_t lds = gather_brick_lds(e);
        const float4* Fq = e->F4 + (size_t)(e->qslot >= 0 ? e->qslot : 0) * e->M;
#define MFT_G2(P_, B_)                                                                                    \
    do {                                                                                                  \
        k_gather_brick<P_, B_><<<tg.nbricks, kTileThreads, lds, e->stream>>>(e->fgrid, e->F4, gm, e->X[e->cur], \
                                                                             e->META[e->cur], e->start, tg, d.nx, \
                                                                             d.ny, d.nz, e->dx[0], e->dx[1], \
                                                                             e->dx[2]);                   \
        if (e->qslot >= 0)                                                                                \
            k_gather_charge<P_, B_><<<tg.nbricks, kTileThreads, 0, e->stream>>>(                          \
                e->fgrid, Fq, e->X[e->cur], e->META[e->cur], e->start, tg, d.nx, d.ny, d.nz, e->dx[0], e->dx[1], e->dx[2]); \
    } while (0)
#define MFT_G2_B(P_)                                                                                      \
    if (BM == 8) MFT_G2(P_, 8);                                                                           \
    else MFT_G2(P_, 4);
        switch (d.pmeorder) {
        case 0: MFT_G2_B(0); break;
        case 1: MFT_G2_B(1); break;
        case 2: MFT_G2_B(2); break;
        case 3: MFT_G2_B(3); break;
        default: MFT_G2_B(4); break;
        }
#undef MFT_G2_B
#undef MFT_G2
        MFT_LAUNCH_CHECK();
        return MFT_OK;
    }
    if (mode == 0) {
#define MFT_G0(P_)                                                                                        \
    k_gather<P_><<<nblk(e->N), kBlock, 0, e->stream>>>(e->fgrid, e->F4, e->force_slot, e->qslot, e->X[e->cur], e->META[e->cur], \
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
        const TileGeom& tg = e->tg;
        const int BM = 1 << tg.lbmax;
        GatherMap gm;
        for (int sp = 0; sp < kMaxSpecies; ++sp) gm.slot[sp] = sp < d.nspecies ? e->force_slot_h[sp] : -1;
        dim3 grid(tg.nbricks, d.nspecies);
        const float4* Fq = e->F4 + (size_t)(e->qslot >= 0 ? e->qslot : 0) * e->M;
#define MFT_G1(P_, B_)                                                                                    \
    do {                                                                                                  \
        k_gather_species<P_, B_><<<grid, kTileThreads, 0, e->stream>>>(e->fgrid, e->F4, gm, e->X[e->cur], e->META[e->cur], \
                                                                         e->start, tg, d.nx, d.ny, d.nz, e->dx[0], \
                                                                         e->dx[1], e->dx[2]);             \
        if (e->qslot >= 0)                                                                                \
            k_gather_charge<P_, B_><<<tg.nbricks, kTileThreads, 0, e->stream>>>(                          \
                e->fgrid, Fq, e->X[e->cur], e->META[e->cur], e->start, tg, d.nx, d.ny, d.nz, e->dx[0], e->dx[1], e->dx[2]); \
    } while (0)
#define MFT_G1_B(P_)                                                                                      \
    if (BM == 8) MFT_G1(P_, 8);                                                                           \
    else MFT_G1(P_, 4);
        switch (d.pmeorder) {
        case 0: MFT_G1_B(0); break;
        case 1: MFT_G1_B(1); break;
        case 2: MFT_G1_B(2); break;
        case 3: MFT_G1_B(3); break;
        default: MFT_G1_B(4); break;
        }
#undef MFT_G1_B
#undef MFT_G1
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

static StepArgs step_args(const TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    const int c = e->cur;
    StepArgs A;
    A.XO = e->XO[c];
    A.NZ = e->NZ[c];
    A.XN = e->XN;
    A.NN = e->NN;
    A.gp = e->gp;
    A.integ = d.integrator;
    A.delt = d.delt;
    A.L[0] = d.lx;
    A.L[1] = d.ly;
    A.L[2] = d.lz;
    A.seed = d.seed;
    A.step = (uint32_t)e->step;
    for (int j = 0; j < 3; ++j) A.dx[j] = e->dx[j];
    A.p = d.pmeorder;
    A.nx = d.nx;
    A.ny = d.ny;
    A.nz = d.nz;
    A.walls = e->walls;
    const bool ang = d.maxangles > 0;
    A.NA = ang ? e->NA[c] : nullptr;
    A.AT = ang ? e->AT[c] : nullptr;
    A.AG = ang ? e->AG[c] : nullptr;
    A.ap = e->ap;
    A.vv = e->vv_state;
    A.lang = e->lang;
    A.use_v = e->is_vv ? 1 : 0;
    A.FD = e->dpd.on ? e->fdpd : nullptr;
    A.nz_keyed = (d.integrator == MFT_INT_GJF && e->nz_keyed) ? 1 : 0;
    A.write_nn = d.integrator == MFT_INT_GJF ? 0 : 1;
    return A;
}

int tild_integrate(TildEngine* e, bool want_f)
{
    const mft_tild_desc& d = e->d;
    const int c = e->cur;
    const uint32_t* ds = e->use_dstep ? e->d_step : nullptr;
    if (d.maxbonds == 2)
        k_integrate<true><<<nblk(e->N), kBlock, 0, e->stream>>>(e->X[c], e->META[c], e->BS[c],
                                                              slot_bond_types(e, c), e->fgrid,
                                                              want_f ? e->ftot : nullptr, 2, e->bp,
                                                              step_args(e), e->N, e->count, e->key, e->rank, e->tg,
                                                              ds);
    else
        k_integrate<false><<<nblk(e->N), kBlock, 0, e->stream>>>(e->X[c], e->META[c], e->BS[c],
                                                               slot_bond_types(e, c), e->fgrid,
                                                               want_f ? e->ftot : nullptr,
                                                               std::max(1, d.maxbonds), e->bp, step_args(e), e->N,
                                                               e->count, e->key, e->rank, e->tg, ds);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// d_wallHarmonicEnergy (ps_potentialWall.cu:129-147), FP64 block sums
__global__ void k_wall_energy(double* __restrict__ partial, const float4* __restrict__ X,
                              const PMeta* __restrict__ META, WallParams W, int w, int ns)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (s < ns && ((W.mask[w] >> META[s].sp()) & 1u)) {
        const float4 xc = X[s];
        const int a = W.axis[w];
        const float disp = W.pos[w] - (a == 0 ? xc.x : (a == 1 ? xc.y : xc.z));
        if ((float)W.sign[w] * disp > 0.0f) v = 0.5f * W.k[w] * disp * disp;
    }
    v = wave_sum(v);
    __shared__ double red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) t += red[k];
        partial[blockIdx.x] = t;
    }
}

int tild_wall_energy(TildEngine* e, int w, double* energy)
{
    const int c = e->cur;
    const unsigned nb = nblk(e->N);
    k_wall_energy<<<nb, kBlock, 0, e->stream>>>(e->bpart, e->X[c], e->META[c], e->walls, w, e->N);
    MFT_LAUNCH_CHECK();
    std::vector<double> h(nb);
    MFT_HIP(hipMemcpyAsync(h.data(), e->bpart, sizeof(double) * nb, hipMemcpyDeviceToHost, e->stream));
    MFT_HIP(hipStreamSynchronize(e->stream));
    double t = 0.0;
    for (double v : h) t += v;
    *energy = t;
    return MFT_OK;
}

// d_anglesStressEnergy summed over particles / 3 (PS_Box.cu:296-305), FP64
__global__ void k_angle_energy_slots(double* __restrict__ partial, const float4* __restrict__ X,
                                     const int* __restrict__ NA, const int* __restrict__ AT,
                                     const int* __restrict__ AG, AngleParams ap, float L0, float L1, float L2, int ns)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (s < ns) {
        const int ma = ap.maxangles, na = min(NA[s], ma);
        float es = 0.0f;
        for (int a = 0; a < na; ++a) {
            const int* g = AG + (s * ma + a) * 3;
            float fs[3], en;
            if (angle_terms(s, X, g[0], g[1], g[2], AT[s * ma + a], ap, L0, L1, L2, fs, &en)) es += en;
        }
        v = es;
    }
    v = wave_sum(v);
    __shared__ double red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) t += red[k];
        partial[blockIdx.x] = t;
    }
}

int tild_angle_energy(TildEngine* e, double* energy)
{
    const mft_tild_desc& d = e->d;
    const int c = e->cur;
    const unsigned nb = nblk(e->N);
    k_angle_energy_slots<<<nb, kBlock, 0, e->stream>>>(e->bpart, e->X[c], e->NA[c], e->AT[c], e->AG[c], e->ap, d.lx,
                                                        d.ly, d.lz, e->N);
    MFT_LAUNCH_CHECK();
    std::vector<double> h(nb);
    MFT_HIP(hipMemcpyAsync(h.data(), e->bpart, sizeof(double) * nb, hipMemcpyDeviceToHost, e->stream));
    MFT_HIP(hipStreamSynchronize(e->stream));
    double t = 0.0;
    for (double v : h) t += v;
    *energy = t / 3.0;
    return MFT_OK;
}

int tild_bond_energy(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    const int c = e->cur;
    unsigned nb = nblk(e->N);
    k_bond_energy_slots<<<nb, kBlock, 0, e->stream>>>(e->bpart, e->X[c], e->BS[c], slot_bond_types(e, c), e->bp, d.lx,
                                                       d.ly, d.lz, e->N, std::max(1, d.maxbonds));
    k_sum_partials<<<1, 1024, 0, e->stream>>>(e->eout + kMaxPots, e->bpart, nb);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

} // namespace mft

using namespace mft;

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

} // extern "C"
