// Line-FFT passes of the TILD grid pipeline: x and y transforms of all
// channels in one launch, and the last inverse x pass that writes the float4
// force fields.
#include "tild_impl.h"

namespace mft {

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

// Last inverse pass along x.  The force components are packed two per
// complex field (component r of field f = 3f + j is the real part of packed
// field r/2 for even r, the imaginary part for odd r); one workgroup row
// transforms the three packed lines of fields 2g and 2g+1 and writes them as
// float4 (x, y, z, 0).
template <int N>
__global__ void __launch_bounds__(256) k_fft_x_inv_packed(const float2* __restrict__ Y, float4* __restrict__ F4,
                                                          int ny, int nz, int nfield, int npk,
                                                          const float2* __restrict__ twg)
{
    using P = FftPlan<N>;
    constexpr int T = P::T, E = P::E, B = 256 / T;
    constexpr int NP = swz_pitch(N);
    __shared__ float2 lds[B * NP];
    __shared__ float2 tw[N];
    for (int i = threadIdx.x; i < N; i += 256) tw[i] = twg[i];
    const int lline = threadIdx.x / T;
    const int t = threadIdx.x % T;
    const int line = blockIdx.x * B + lline;
    const bool valid = line < ny * nz;
    const long M = (long)N * ny * nz;
    const long base = (long)line * N;
    const int g = blockIdx.y;
    float2 pk[3][E];
    // the three packed lines' loads in flight before the first transform
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int q = 3 * g + j;
#pragma unroll
        for (int s = 0; s < E; ++s)
            pk[j][s] = (valid && q < npk) ? Y[q * M + base + t + T * s] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    // a line's T threads are consecutive threads of one wavefront
    LdsMap map{lline, B, false, NP, true};
#pragma unroll
    for (int j = 0; j < 3; ++j) line_fft<N, +1, true>(pk[j], lds, tw, t, map);
    if (!valid) return;
    // fields 2g (components in pk0.re, pk0.im, pk1.re) and 2g+1 (pk1.im,
    // pk2.re, pk2.im)
    float4* d0 = F4 + (2 * g) * M;
#pragma unroll
    for (int s = 0; s < E; ++s) ff_store(d0, base + t + T * s, pk[0][s].x, pk[0][s].y, pk[1][s].x);
    if (2 * g + 1 < nfield) {
        float4* d1 = F4 + (2 * g + 1) * M;
#pragma unroll
        for (int s = 0; s < E; ++s) ff_store(d1, base + t + T * s, pk[1][s].y, pk[2][s].x, pk[2][s].y);
    }
}

// First forward pass along x, reading the density tiles: each grid value is
// the sum of the tile entries that cover it, in the fixed order of
// k_tile_reduce (z, then y, then x contributions), so the densities are the
// same bits as the reduced grid.  tab holds per axis and coordinate the count
// and the offsets of the covering tile entries.  The block's B lines are
// summed into LDS first (consecutive lanes on consecutive x, so the tile rows
// are read coalesced), then transformed from there.
template <int N>
__global__ void __launch_bounds__(256) k_fft_x_fwd_tiles(const float* __restrict__ tiles, const int* __restrict__ tab,
                                                         float2* __restrict__ out, int ny, int nz, long cstride,
                                                         const float2* __restrict__ twg, int cover2)
{
    using P = FftPlan<N>;
    constexpr int T = P::T, E = P::E, B = 256 / T;
    constexpr int NP = swz_pitch(N);
    __shared__ float2 lds[B * NP];
    __shared__ float2 tw[N];
    __shared__ int tabx[5 * N];
    for (int i = threadIdx.x; i < N; i += 256) tw[i] = twg[i];
    for (int i = threadIdx.x; i < 5 * N; i += 256) tabx[i] = tab[i];
    __syncthreads();
    const int c = blockIdx.y;
    const int nlines = ny * nz;
    const float* tc = tiles + c * cstride;
    constexpr int PER = (B * N + 255) / 256;
    float sums[PER];
    bool lane_cover = false;
    if constexpr (N % 64 == 0) lane_cover = cover2 && cstride < (1L << 30);
    if (lane_cover) {
        // at most two covering tiles per axis, and a wavefront's 64 lanes on
        // one line (N a multiple of 64): the line's z and y cover rows, and
        // which of its (z, y) entries exist, are uniform over the wave (scalar
        // loads and branches); only the x cover is per lane.  Entries a line
        // lacks are not loaded; a lane without a second x entry re-reads its
        // first.  Byte offsets within a channel fit 32 bits (cstride < 2^30),
        // so every load is a base plus a 32-bit lane offset.
        const char* tb = reinterpret_cast<const char*>(tc);
        int cz[PER], cy[PER], oz[PER][2], oy[PER][2], cx[PER], bx[PER][2];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = threadIdx.x + 256 * k;
            const int l = __builtin_amdgcn_readfirstlane(i / N);
            const int x = i - l * N;
            const int line = blockIdx.x * B + l;
            const bool valid = l < B && line < nlines;
            const int gy = valid ? line % ny : 0, gz = valid ? line / ny : 0;
            const int* tz = tab + 5 * (N + ny + gz);
            const int* ty = tab + 5 * (N + gy);
            const int* tx = tabx + 5 * x;
            cz[k] = valid ? tz[0] : 0;
            cy[k] = ty[0];
            oz[k][0] = tz[1];
            oz[k][1] = tz[2];
            oy[k][0] = ty[1];
            oy[k][1] = ty[2];
            cx[k] = tx[0];
            bx[k][0] = 4 * tx[1];
            bx[k][1] = 4 * (tx[0] > 1 ? tx[2] : tx[1]);
        }
        float v[PER][8];
#pragma unroll
        for (int k = 0; k < PER; ++k)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    float v0 = 0.f, v1 = 0.f;
                    if (a < cz[k] && b < cy[k]) {
                        const unsigned o = 4u * (unsigned)(oz[k][a] + oy[k][b]);
                        v0 = *reinterpret_cast<const float*>(tb + (o + (unsigned)bx[k][0]));
                        v1 = *reinterpret_cast<const float*>(tb + (o + (unsigned)bx[k][1]));
                    }
                    v[k][4 * a + 2 * b] = v0;
                    v[k][4 * a + 2 * b + 1] = v1;
                }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            float sum = 0.f;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int d = 0; d < 2; ++d) {
                        const bool in = (a < cz[k]) & (b < cy[k]) & (d < cx[k]);
                        sum = in ? sum + v[k][4 * a + 2 * b + d] : sum;
                    }
            sums[k] = sum;
        }
    } else if (cover2) {
        // at most two covering tiles per axis: every point's cover rows are
        // read first, then all its (up to eight) tile entries are loaded
        // together, branch-free, and summed in k_tile_reduce's order
        int cnt[PER][3], oz[PER][2], oy[PER][2], ox[PER][2];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = threadIdx.x + 256 * k;
            const int l = i / N, x = i - l * N;
            const int line = blockIdx.x * B + l;
            const bool valid = i < B * N && line < nlines;
            const int gy = valid ? line % ny : 0, gz = valid ? line / ny : 0;
            const int* tz = tab + 5 * (N + ny + gz);
            const int* ty = tab + 5 * (N + gy);
            const int* tx = tabx + 5 * (valid ? x : 0);
            cnt[k][0] = valid ? tz[0] : 0;
            cnt[k][1] = ty[0];
            cnt[k][2] = tx[0];
            oz[k][0] = tz[1];
            oz[k][1] = tz[2];
            oy[k][0] = ty[1];
            oy[k][1] = ty[2];
            ox[k][0] = tx[1];
            ox[k][1] = tx[2];
        }
        float v[PER][8];
#pragma unroll
        for (int k = 0; k < PER; ++k)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int d = 0; d < 2; ++d) {
                        // entries a point lacks read its first entry instead (an
                        // offset in range, on a line the wave reads anyway), so
                        // every load is unconditional and none waits on a branch
                        const bool in = (a < cnt[k][0]) & (b < cnt[k][1]) & (d < cnt[k][2]);
                        const int off = in ? oz[k][a] + oy[k][b] + ox[k][d] : oz[k][0] + oy[k][0] + ox[k][0];
                        const float t = tc[off];
                        v[k][4 * a + 2 * b + d] = in ? t : 0.f;
                    }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            float sum = 0.f;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int d = 0; d < 2; ++d)
                    {
                        const bool in = (a < cnt[k][0]) & (b < cnt[k][1]) & (d < cnt[k][2]);
                        sum = in ? sum + v[k][4 * a + 2 * b + d] : sum;
                    }
            sums[k] = sum;
        }
    } else {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = threadIdx.x + 256 * k;
            const int l = i / N, x = i - l * N;
            const int line = blockIdx.x * B + l;
            float sum = 0.f;
            if (i < B * N && line < nlines) {
                const int gy = line % ny, gz = line / ny;
                sum = tile_sum(tc, tab + 5 * (N + ny + gz), tab + 5 * (N + gy), tabx + 5 * x);
            }
            sums[k] = sum;
        }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = threadIdx.x + 256 * k;
        if (i < B * N) lds[(i / N) * NP + i % N] = make_float2(sums[k], 0.f);
    }
    __syncthreads();
    const int lline = threadIdx.x / T;
    const int t = threadIdx.x % T;
    const int line = blockIdx.x * B + lline;
    float2 v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = lds[lline * NP + t + T * s];
    // lines NP apart: each line's exchanges stay inside its own region
    LdsMap map{lline, B, false, NP, true};
    line_fft<N, -1, true>(v, lds, tw, t, map);
    if (line >= nlines) return;
    float2* dst = out + c * (long)N * nlines + (long)line * N;
#pragma unroll
    for (int s = 0; s < E; ++s) dst[t + T * s] = v[s];
}

// ------------------------------------------------------- plane passes ---
// The x and y transforms of one z-plane of an N x N grid in one workgroup:
// the plane sits in LDS with a row pitch of N + 1 (a column's elements fall in
// different banks) and its lines are transformed in place, T consecutive
// threads per line (so a line's exchanges stay inside one wavefront), LINES
// lines at a time.  The arithmetic is that of the line passes (same plans,
// twiddles and summation order; the compiler may still contract different
// multiply-adds in the two kernels, so results can differ by rounding); the
// plane passes save one HBM round trip of the spectra per direction.
template <int N>
struct PlaneShape {
    using P = FftPlan<N>;
    static constexpr int T = P::T, E = P::E;
    static constexpr int NT = T * N >= 1024 ? 1024 : T * N;
    static constexpr int LINES = NT / T;
    static constexpr int ROUNDS = N / LINES;
    static constexpr int PITCH = N + 1;
    static constexpr size_t LDS = sizeof(float2) * N * PITCH;
    static_assert(N % LINES == 0, "whole rounds of lines");
};

// Forward x then y transforms of channel blockIdx.y, z-plane blockIdx.x, from
// the density tiles (each grid value summed from the covering tile entries in
// k_tile_reduce's order), written as spectra x-fastest.
template <int N>
__global__ void __launch_bounds__(PlaneShape<N>::NT) k_fft_xy_fwd_tiles_plane(
    const float* __restrict__ tiles, const int* __restrict__ tab, float2* __restrict__ out, int nz, long cstride,
    const float2* __restrict__ twg)
{
    using S = PlaneShape<N>;
    constexpr int T = S::T, E = S::E, NT = S::NT, LINES = S::LINES, PITCH = S::PITCH;
    extern __shared__ float2 plane[];
    __shared__ float2 tw[N];
    __shared__ int tabxy[5 * 2 * N];
    for (int i = threadIdx.x; i < N; i += NT) tw[i] = twg[i];
    for (int i = threadIdx.x; i < 5 * 2 * N; i += NT) tabxy[i] = tab[i];
    __syncthreads();
    const int iz = blockIdx.x, c = blockIdx.y;
    const float* tc = tiles + c * cstride;
    const int* tz = tab + 5 * (2 * N + iz);
    for (int i = threadIdx.x; i < N * N; i += NT) {
        const int y = i / N, x = i - y * N;
        plane[y * PITCH + x] = make_float2(tile_sum(tc, tz, tabxy + 5 * (N + y), tabxy + 5 * x), 0.f);
    }
    __syncthreads();
    const int lt = threadIdx.x / T, t = threadIdx.x % T;
#pragma unroll
    for (int r = 0; r < S::ROUNDS; ++r) {
        LdsMap map{r * LINES + lt, 0, false, PITCH};   // row y
        float2 v[E];
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = plane[map(t + T * s)];
        line_fft<N, -1, true>(v, plane, tw, t, map);
#pragma unroll
        for (int s = 0; s < E; ++s) plane[map(t + T * s)] = v[s];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < S::ROUNDS; ++r) {
        LdsMap map{r * LINES + lt, PITCH, true, N};    // column x
        float2 v[E];
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = plane[map(t + T * s)];
        line_fft<N, -1, true>(v, plane, tw, t, map);
#pragma unroll
        for (int s = 0; s < E; ++s) plane[map(t + T * s)] = v[s];
    }
    __syncthreads();
    float2* dst = out + (long)c * N * N * nz + (long)iz * N * N;
    for (int i = threadIdx.x; i < N * N; i += NT) {
        const int y = i / N, x = i - y * N;
        dst[i] = plane[y * PITCH + x];
    }
}

// Inverse y then x transforms of the packed force spectra of z-plane
// blockIdx.x, written as the float4 force fields 2g and 2g + 1 (g =
// blockIdx.y) like k_fft_x_inv_packed: the three packed planes 3g .. 3g + 2
// are transformed one after the other, the first one's values (and the
// second one's imaginary parts) kept in registers until the fields they
// complete are written.
template <int N>
__global__ void __launch_bounds__(PlaneShape<N>::NT) k_fft_yx_inv_packed_plane(
    const float2* __restrict__ Y, float4* __restrict__ F4, int nz, int nfield, int npk,
    const float2* __restrict__ twg)
{
    using S = PlaneShape<N>;
    constexpr int T = S::T, E = S::E, NT = S::NT, LINES = S::LINES, PITCH = S::PITCH, R = S::ROUNDS;
    extern __shared__ float2 plane[];
    __shared__ float2 tw[N];
    for (int i = threadIdx.x; i < N; i += NT) tw[i] = twg[i];
    const int iz = blockIdx.x, g = blockIdx.y;
    const long M = (long)N * N * nz, pbase = (long)iz * N * N;
    const int lt = threadIdx.x / T, t = threadIdx.x % T;
    float2 keep[R][E];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int s = 0; s < E; ++s) keep[r][s] = make_float2(0.f, 0.f);
    float4* d0 = F4 + (long)(2 * g) * M;
    float4* d1 = F4 + (long)(2 * g + 1) * M;
    const bool second = 2 * g + 1 < nfield;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int q = 3 * g + j;
        const bool have = q < npk;   // a missing packed plane is zero
        if (have) {
            const float2* src = Y + (long)q * M + pbase;
            for (int i = threadIdx.x; i < N * N; i += NT) {
                const int y = i / N, x = i - y * N;
                plane[y * PITCH + x] = src[i];
            }
        }
        __syncthreads();
        if (have) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                LdsMap map{r * LINES + lt, PITCH, true, N};    // column x
                float2 v[E];
#pragma unroll
                for (int s = 0; s < E; ++s) v[s] = plane[map(t + T * s)];
                line_fft<N, +1, true>(v, plane, tw, t, map);
#pragma unroll
                for (int s = 0; s < E; ++s) plane[map(t + T * s)] = v[s];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = r * LINES + lt;
            float2 v[E];
            if (have) {
                LdsMap map{row, 0, false, PITCH};          // row y
#pragma unroll
                for (int s = 0; s < E; ++s) v[s] = plane[map(t + T * s)];
                line_fft<N, +1, true>(v, plane, tw, t, map);
            } else {
#pragma unroll
                for (int s = 0; s < E; ++s) v[s] = make_float2(0.f, 0.f);
            }
            const long o = (long)row * N + t;
            if (j == 1) {
                // field 2g: packed 3g (re, im) and 3g + 1 (re)
#pragma unroll
                for (int s = 0; s < E; ++s) ff_store(d0, pbase + o + T * s, keep[r][s].x, keep[r][s].y, v[s].x);
            } else if (j == 2 && second) {
                // field 2g + 1: packed 3g + 1 (im) and 3g + 2 (re, im)
#pragma unroll
                for (int s = 0; s < E; ++s) ff_store(d1, pbase + o + T * s, keep[r][s].y, v[s].x, v[s].y);
            }
#pragma unroll
            for (int s = 0; s < E; ++s) keep[r][s] = v[s];
        }
        __syncthreads();   // the plane is reloaded next
    }
}

// grids whose z-planes the plane passes handle (square planes that fit LDS)
bool tild_plane_supported(int nx, int ny)
{
    return nx == ny && (nx == 64 || nx == 96 || nx == 128);
}

// the plane kernels use more dynamic LDS than the default limit: raise it
// once per kernel (setting it again is harmless)
static int allow_plane_lds(const void* fn, size_t bytes, bool* done)
{
    if (*done) return MFT_OK;
    MFT_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    *done = true;
    return MFT_OK;
}

int tild_xy_fwd_tiles_plane(const float* tiles, const int* tab, float2* S, int nx, int ny, int nz, int nchan,
                            long cstride, hipStream_t st)
{
    if (!tild_plane_supported(nx, ny)) return MFT_ERR_UNSUPPORTED;
    switch (nx) {
#define MFT_CASE(n_)                                                                                      \
    case n_: {                                                                                            \
        using PS = PlaneShape<n_>;                                                                        \
        static bool attr = false;                                                                         \
        int rc = allow_plane_lds((const void*)k_fft_xy_fwd_tiles_plane<n_>, PS::LDS, &attr);              \
        if (rc) return rc;                                                                                \
        k_fft_xy_fwd_tiles_plane<n_><<<dim3(nz, nchan), PS::NT, PS::LDS, st>>>(tiles, tab, S, nz, cstride, \
                                                                            fft_twiddles_device(n_));    \
        break;                                                                                            \
    }
        MFT_CASE(64)
        MFT_CASE(96)
        MFT_CASE(128)
#undef MFT_CASE
    default: return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int tild_yx_inv_packed_plane(const float2* Y, float4* F4, int nx, int ny, int nz, int nfield, hipStream_t st)
{
    if (!tild_plane_supported(nx, ny)) return MFT_ERR_UNSUPPORTED;
    const int npk = (3 * nfield + 1) / 2;
    switch (nx) {
#define MFT_CASE(n_)                                                                                      \
    case n_: {                                                                                            \
        using PS = PlaneShape<n_>;                                                                        \
        static bool attr = false;                                                                         \
        int rc = allow_plane_lds((const void*)k_fft_yx_inv_packed_plane<n_>, PS::LDS, &attr);             \
        if (rc) return rc;                                                                                \
        k_fft_yx_inv_packed_plane<n_><<<dim3(nz, (nfield + 1) / 2), PS::NT, PS::LDS, st>>>(              \
            Y, F4, nz, nfield, npk, fft_twiddles_device(n_));                                             \
        break;                                                                                            \
    }
        MFT_CASE(64)
        MFT_CASE(96)
        MFT_CASE(128)
#undef MFT_CASE
    default: return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// The same pass with the tile rows staged in LDS first: the (z cover, line,
// y cover) rows of the halo-expanded tile grid that the block's B lines read
// (at most two covers per axis) are copied with float4 loads, then each point
// sums its entries from LDS in k_tile_reduce's order (the same bits), and the
// staging buffer is reused for the transform's exchanges.  Needs the
// halo-expanded row length ex0 to be a multiple of 4.
template <int N>
__global__ void __launch_bounds__(256) k_fft_x_fwd_tiles_staged(const float* __restrict__ tiles,
                                                                const int* __restrict__ tab,
                                                                float2* __restrict__ out, int ny, int nz,
                                                                long cstride, int ex0,
                                                                const float2* __restrict__ twg)
{
    using P = FftPlan<N>;
    constexpr int T = P::T, E = P::E, B = 256 / T;
    extern __shared__ float4 stage4[];   // [2 z covers][B lines][2 y covers][ex0 / 4]
    float* stage = reinterpret_cast<float*>(stage4);
    float2* lds = reinterpret_cast<float2*>(stage4);
    __shared__ float2 tw[N];
    __shared__ int tabx[5 * N];
    __shared__ int rowoff[2 * B * 2];
    __shared__ int ncz[B], ncy[B];
    for (int i = threadIdx.x; i < N; i += 256) tw[i] = twg[i];
    for (int i = threadIdx.x; i < 5 * N; i += 256) tabx[i] = tab[i];
    const int c = blockIdx.y;
    const int nlines = ny * nz;
    if ((int)threadIdx.x < B) {
        const int l = threadIdx.x, line = blockIdx.x * B + l;
        int cz = 0, cy = 0;
        const int* tz = nullptr;
        const int* ty = nullptr;
        if (line < nlines) {
            tz = tab + 5 * (N + ny + line / ny);
            ty = tab + 5 * (N + line % ny);
            cz = tz[0];
            cy = ty[0];
        }
        ncz[l] = cz;
        ncy[l] = cy;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                rowoff[(a * B + l) * 2 + b] = (a < cz && b < cy) ? tz[1 + a] + ty[1 + b] : -1;
    }
    __syncthreads();
    const int q = ex0 >> 2;
    const float4* tc4 = reinterpret_cast<const float4*>(tiles + c * cstride);
    for (int i = threadIdx.x; i < 4 * B * q; i += 256) {
        const int r = i / q, j = i - r * q;
        const int off = rowoff[r];
        if (off >= 0) stage4[(long)r * q + j] = tc4[(off >> 2) + j];
    }
    __syncthreads();
    constexpr int PER = (B * N + 255) / 256;
    float sums[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = threadIdx.x + 256 * k;
        float s = 0.f;
        if (i < B * N) {
            const int l = i / N, x = i - l * N;
            const int* tx = tabx + 5 * x;
            const int cz = ncz[l], cy = ncy[l], cx = tx[0];
            for (int a = 0; a < cz; ++a)
                for (int b = 0; b < cy; ++b) {
                    const float* row = stage + (long)((a * B + l) * 2 + b) * ex0;
                    for (int d = 0; d < cx; ++d) s += row[tx[1 + d]];
                }
        }
        sums[k] = s;
    }
    __syncthreads();   // the staged rows are read; the buffer takes the transform
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = threadIdx.x + 256 * k;
        if (i < B * N) lds[i] = make_float2(sums[k], 0.f);
    }
    __syncthreads();
    const int lline = threadIdx.x / T;
    const int t = threadIdx.x % T;
    const int line = blockIdx.x * B + lline;
    float2 v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = lds[lline * N + t + T * s];
    LdsMap map{lline, B, false, N};
    line_fft<N, -1, true>(v, lds, tw, t, map);
    if (line >= nlines) return;
    float2* dst = out + c * (long)N * nlines + (long)line * N;
#pragma unroll
    for (int s = 0; s < E; ++s) dst[t + T * s] = v[s];
}

// staged x pass possible: row length a multiple of 4 floats, at most two
// covering tiles per coordinate (every axis), the staging buffer in 64 KB
bool tild_x_fwd_staged_ok(int nx, int ex0, int maxcover)
{
    int T = 0;
#define MFT_T(n_) if (nx == n_) T = FftPlan<n_>::T;
    MFT_FOR_EACH_FFT_SIZE(MFT_T)
#undef MFT_T
    if (T == 0 || ex0 % 4 != 0 || maxcover > 2) return false;
    const size_t B = 256 / T;
    return 16 * B * (size_t)ex0 <= 64 * 1024 && B * (size_t)nx * 8 <= 16 * B * (size_t)ex0;
}

int tild_x_fwd_tiles(const float* tiles, const int* tab, float2* S, int nx, int ny, int nz, int nchan, long cstride,
                     hipStream_t st, int ex0, bool staged, bool cover2)
{
    if (staged) {
        switch (nx) {
#define MFT_CASE(n_)                                                                                      \
    case n_: {                                                                                            \
        constexpr int B = 256 / FftPlan<n_>::T;                                                           \
        dim3 grid((ny * nz + B - 1) / B, nchan);                                                          \
        k_fft_x_fwd_tiles_staged<n_><<<grid, 256, 16 * B * (size_t)ex0, st>>>(tiles, tab, S, ny, nz, cstride, ex0, \
                                                                            fft_twiddles_device(n_));    \
        break;                                                                                            \
    }
            MFT_FOR_EACH_FFT_SIZE(MFT_CASE)
#undef MFT_CASE
        default: return MFT_ERR_UNSUPPORTED;
        }
        MFT_LAUNCH_CHECK();
        return MFT_OK;
    }
    switch (nx) {
#define MFT_CASE(n_)                                                                                      \
    case n_: {                                                                                            \
        constexpr int B = 256 / FftPlan<n_>::T;                                                           \
        dim3 grid((ny * nz + B - 1) / B, nchan);                                                          \
        k_fft_x_fwd_tiles<n_><<<grid, 256, 0, st>>>(tiles, tab, S, ny, nz, cstride, fft_twiddles_device(n_), \
                                                    cover2 ? 1 : 0);                                      \
        break;                                                                                            \
    }
        MFT_FOR_EACH_FFT_SIZE(MFT_CASE)
#undef MFT_CASE
    default: return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
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

// the three variants the engine uses: forward real -> complex, forward and
// inverse complex -> complex
int tild_axis(int sign, int real_in, int real_out, const void* in, void* out, AxisGeom g, int nf, long fs,
              float scale, hipStream_t st)
{
    if (sign < 0 && real_in && !real_out) return launch_axis<-1, true, false>(in, out, g, nf, fs, scale, st);
    if (sign < 0 && !real_in && !real_out) return launch_axis<-1, false, false>(in, out, g, nf, fs, scale, st);
    if (sign > 0 && !real_in && !real_out) return launch_axis<+1, false, false>(in, out, g, nf, fs, scale, st);
    set_error(MFT_ERR_UNSUPPORTED, "axis pass variant not built");
    return MFT_ERR_UNSUPPORTED;
}

int tild_x_inv_packed(const float2* Y, float4* F4, int nx, int ny, int nz, int nfield, hipStream_t st)
{
    const int npk = (3 * nfield + 1) / 2;
    switch (nx) {
#define MFT_CASE(n_)                                                                                      \
    case n_: {                                                                                            \
        constexpr int B = 256 / FftPlan<n_>::T;                                                           \
        dim3 grid((ny * nz + B - 1) / B, (nfield + 1) / 2);                                               \
        k_fft_x_inv_packed<n_><<<grid, 256, 0, st>>>(Y, F4, ny, nz, nfield, npk, fft_twiddles_device(n_)); \
        break;                                                                                            \
    }
        MFT_FOR_EACH_FFT_SIZE(MFT_CASE)
#undef MFT_CASE
    default: return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

} // namespace mft
