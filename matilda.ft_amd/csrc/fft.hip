// Batched 3D complex FFT (plan/exec) over the line kernels of fft_core.h and
// the twiddle tables every FFT-using kernel stages into LDS.
#include <cmath>
#include <mutex>
#include <vector>

#include "common.h"
#include "fft_core.h"

namespace mft {

// exp(-2 pi i m / n), m < n, rounded once from double precision; one table
// per length and device, created on first use (never inside graph capture:
// engines create theirs at setup).
const float2* fft_twiddles_device(int n)
{
    static std::mutex mu;
    static std::vector<std::pair<long, float2*>> cache; // key = device * 4096 + n
    int dev = 0;
    (void)hipGetDevice(&dev);
    long key = (long)dev * 4096 + n;
    std::lock_guard<std::mutex> lk(mu);
    for (auto& kv : cache)
        if (kv.first == key) return kv.second;
    std::vector<float2> h(n);
    for (int m = 0; m < n; ++m) {
        double a = -2.0 * M_PI * (double)m / (double)n;
        h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    float2* d = nullptr;
    if (hipMalloc((void**)&d, sizeof(float2) * n) != hipSuccess) return nullptr;
    (void)hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice);
    cache.push_back({key, d});
    return d;
}

template <int N, int SIGN>
__global__ void __launch_bounds__(256) k_fft_lines(const float2* __restrict__ in, float2* __restrict__ out,
                                                   AxisGeom geom, long field_stride, const float2* __restrict__ twg,
                                                   float scale)
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
    int ta, tb;
    const long base = valid ? geom.base(line, &ta, &tb) : 0;
    const long st = geom.stride();
    const float2* src = in + blockIdx.y * field_stride + base;
    float2 v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = valid ? src[(t + T * s) * st] : make_float2(0.f, 0.f);
    __syncthreads();
    LdsMap map{lline, B, inter, N};
    line_fft<N, SIGN>(v, lds, tw, t, map);
    if (!valid) return;
    float2* dst = out + blockIdx.y * field_stride + base;
#pragma unroll
    for (int s = 0; s < E; ++s) dst[(t + T * s) * st] = cscale(v[s], scale);
}

template <int SIGN>
static int launch_lines(const float2* in, float2* out, AxisGeom g, int batch, long fs, float scale, hipStream_t st)
{
    switch (g.length()) {
#define MFT_CASE(n_)                                                                                       \
    case n_: {                                                                                             \
        constexpr int B = 256 / FftPlan<n_>::T;                                                            \
        dim3 grid((g.nlines() + B - 1) / B, batch);                                                        \
        k_fft_lines<n_, SIGN><<<grid, 256, 0, st>>>(in, out, g, fs, fft_twiddles_device(n_), scale);       \
        break;                                                                                             \
    }
        MFT_FOR_EACH_FFT_SIZE(MFT_CASE)
#undef MFT_CASE
    default:
        set_error(MFT_ERR_UNSUPPORTED, "unsupported FFT length %d", g.length());
        return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

} // namespace mft

using namespace mft;

struct mft_fft_plan {
    int n[3];
    int batch;
};

extern "C" {

int mft_fft_supported(int n)
{
    switch (n) {
#define MFT_CASE(n_) case n_:
        MFT_FOR_EACH_FFT_SIZE(MFT_CASE)
#undef MFT_CASE
        return 1;
    default: return 0;
    }
}

int mft_fft_plan_create(mft_fft_plan** plan, int nx, int ny, int nz, int batch)
{
    MFT_REQUIRE(plan, MFT_ERR_INVALID, "null plan");
    *plan = nullptr;
    MFT_REQUIRE(batch >= 1, MFT_ERR_INVALID, "batch must be positive");
    int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        MFT_REQUIRE(n[a] >= 1, MFT_ERR_INVALID, "grid size must be positive");
        MFT_REQUIRE(n[a] == 1 || mft_fft_supported(n[a]), MFT_ERR_UNSUPPORTED,
                    "FFT length %d is not supported (2,3,4,5,6,8,10,12,16,20,24,32,40,48,64,80,96,128,160,192,256)", n[a]);
        if (n[a] > 1) MFT_REQUIRE(fft_twiddles_device(n[a]) != nullptr, MFT_ERR_OOM, "twiddle table");
    }
    mft_fft_plan* p = new mft_fft_plan;
    p->n[0] = nx;
    p->n[1] = ny;
    p->n[2] = nz;
    p->batch = batch;
    *plan = p;
    return MFT_OK;
}

int mft_fft_exec(mft_fft_plan* p, const void* in, void* out, int dir, void* stream)
{
    MFT_REQUIRE(p && in && out, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(dir == 1 || dir == -1, MFT_ERR_INVALID, "dir must be 1 (forward) or -1 (inverse)");
    hipStream_t st = (hipStream_t)stream;
    const long M = (long)p->n[0] * p->n[1] * p->n[2];
    const float scale_fwd = (float)(1.0 / (double)M);
    const float2* src = (const float2*)in;
    float2* dst = (float2*)out;
    int last = -1;
    for (int a = 0; a < 3; ++a)
        if (p->n[a] > 1) last = a;
    if (last < 0) {
        if (in != out) MFT_HIP(hipMemcpyAsync(out, in, sizeof(float2) * M * p->batch, hipMemcpyDeviceToDevice, st));
        return MFT_OK;
    }
    int rc;
    for (int a = 0; a < 3; ++a) {
        if (p->n[a] <= 1) continue;
        AxisGeom g{p->n[0], p->n[1], p->n[2], a};
        float scale = (dir == 1 && a == last) ? scale_fwd : 1.0f;
        rc = dir == 1 ? launch_lines<-1>(src, dst, g, p->batch, M, scale, st)
                      : launch_lines<+1>(src, dst, g, p->batch, M, scale, st);
        if (rc) return rc;
        src = dst;
    }
    return MFT_OK;
}

int mft_fft_plan_destroy(mft_fft_plan* p)
{
    delete p;
    return MFT_OK;
}

} // extern "C"
