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
    __shared__ float2 lds[B * N];
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
    LdsMap map{lline, B, false, N};
#pragma unroll
    for (int j = 0; j < 3; ++j) line_fft<N, +1, true>(pk[j], lds, tw, t, map);
    if (!valid) return;
    // fields 2g (components in pk0.re, pk0.im, pk1.re) and 2g+1 (pk1.im,
    // pk2.re, pk2.im)
    float4* d0 = F4 + (2 * g) * M + base;
#pragma unroll
    for (int s = 0; s < E; ++s) d0[t + T * s] = make_float4(pk[0][s].x, pk[0][s].y, pk[1][s].x, 0.f);
    if (2 * g + 1 < nfield) {
        float4* d1 = F4 + (2 * g + 1) * M + base;
#pragma unroll
        for (int s = 0; s < E; ++s) d1[t + T * s] = make_float4(pk[1][s].y, pk[2][s].x, pk[2][s].y, 0.f);
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
                                                         const float2* __restrict__ twg)
{
    using P = FftPlan<N>;
    constexpr int T = P::T, E = P::E, B = 256 / T;
    __shared__ float2 lds[B * N];
    __shared__ float2 tw[N];
    __shared__ int tabx[5 * N];
    for (int i = threadIdx.x; i < N; i += 256) tw[i] = twg[i];
    for (int i = threadIdx.x; i < 5 * N; i += 256) tabx[i] = tab[i];
    __syncthreads();
    const int c = blockIdx.y;
    const int nlines = ny * nz;
    const float* tc = tiles + c * cstride;
    for (int i = threadIdx.x; i < B * N; i += 256) {
        const int l = i / N, x = i - l * N;
        const int line = blockIdx.x * B + l;
        float sum = 0.f;
        if (line < nlines) {
            const int gy = line % ny, gz = line / ny;
            sum = tile_sum(tc, tab + 5 * (N + ny + gz), tab + 5 * (N + gy), tabx + 5 * x);
        }
        lds[i] = make_float2(sum, 0.f);
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

int tild_x_fwd_tiles(const float* tiles, const int* tab, float2* S, int nx, int ny, int nz, int nchan, long cstride,
                     hipStream_t st)
{
    switch (nx) {
#define MFT_CASE(n_)                                                                                      \
    case n_: {                                                                                            \
        constexpr int B = 256 / FftPlan<n_>::T;                                                           \
        dim3 grid((ny * nz + B - 1) / B, nchan);                                                          \
        k_fft_x_fwd_tiles<n_><<<grid, 256, 0, st>>>(tiles, tab, S, ny, nz, cstride, fft_twiddles_device(n_)); \
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
