// Batched line FFTs in registers + LDS for gfx950.
//
// A line of length N is owned by T threads; thread t holds the E = N/T
// elements t, t+T, t+2T, ... ("cyclic" distribution).  Each Stockham pass of
// radix R gives every thread E/R butterflies whose R inputs are exactly its own
// registers b + r*(E/R); between passes the outputs travel through LDS to
// their self-sorted positions and come back in cyclic order.  The last pass
// leaves frequency j + r*N/R in register b + r*(E/R), i.e. the cyclic
// distribution again, so a forward transform, a pointwise k-space operation and
// an inverse transform chain without touching memory.
//
// Twiddles are exp(-2 pi i m / N) rounded once from double precision (host
// table, staged in LDS).  Radices 2, 3, 4, 5 and 8 cover every grid size in
// the supported table below.
#pragma once
#include <hip/hip_runtime.h>

namespace mft {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// i*z (SIGN=+1) or -i*z (SIGN=-1)
template <int SIGN>
__device__ __forceinline__ float2 rot90(float2 z)
{
    return SIGN > 0 ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
}

template <int R, int SIGN>
struct Dft;

template <int SIGN>
struct Dft<2, SIGN> {
    __device__ __forceinline__ static void run(float2* a)
    {
        float2 t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    }
};

template <int SIGN>
struct Dft<4, SIGN> {
    __device__ __forceinline__ static void run(float2* a)
    {
        float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        float2 s13 = cadd(a[1], a[3]), d13 = rot90<SIGN>(csub(a[1], a[3]));
        a[0] = cadd(s02, s13);
        a[2] = csub(s02, s13);
        a[1] = cadd(d02, d13);
        a[3] = csub(d02, d13);
    }
};

template <int SIGN>
struct Dft<8, SIGN> {
    __device__ __forceinline__ static void run(float2* a)
    {
        float2 e[4] = {a[0], a[2], a[4], a[6]};
        float2 o[4] = {a[1], a[3], a[5], a[7]};
        Dft<4, SIGN>::run(e);
        Dft<4, SIGN>::run(o);
        const float c = 0.70710678118654752f;
        // w8^k for k = 1, 2, 3 with w8 = exp(SIGN * 2 pi i / 8)
        float2 o1 = make_float2(c * (o[1].x - SIGN * o[1].y), c * (o[1].y + SIGN * o[1].x));
        float2 o2 = rot90<SIGN>(o[2]);
        float2 o3 = make_float2(-c * (o[3].x + SIGN * o[3].y), c * (SIGN * o[3].x - o[3].y));
        a[0] = cadd(e[0], o[0]);
        a[4] = csub(e[0], o[0]);
        a[1] = cadd(e[1], o1);
        a[5] = csub(e[1], o1);
        a[2] = cadd(e[2], o2);
        a[6] = csub(e[2], o2);
        a[3] = cadd(e[3], o3);
        a[7] = csub(e[3], o3);
    }
};

template <int SIGN>
struct Dft<3, SIGN> {
    __device__ __forceinline__ static void run(float2* a)
    {
        const float h = 0.86602540378443865f; // sqrt(3)/2
        float2 t1 = cadd(a[1], a[2]);
        float2 t2 = csub(a[1], a[2]);
        float2 m = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
        float2 n = cscale(rot90<SIGN>(t2), h);
        a[0] = cadd(a[0], t1);
        a[1] = cadd(m, n);
        a[2] = csub(m, n);
    }
};

template <int SIGN>
struct Dft<5, SIGN> {
    __device__ __forceinline__ static void run(float2* a)
    {
        const float c1 = 0.30901699437494742f;  // cos(2 pi/5)
        const float c2 = -0.80901699437494742f; // cos(4 pi/5)
        const float s1 = 0.95105651629515357f;  // sin(2 pi/5)
        const float s2 = 0.58778525229247313f;  // sin(4 pi/5)
        float2 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
        float2 t3 = csub(a[1], a[4]), t4 = csub(a[2], a[3]);
        float2 b1 = make_float2(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
        float2 b2 = make_float2(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
        float2 d1 = rot90<SIGN>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
        float2 d2 = rot90<SIGN>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
        a[0] = cadd(a[0], cadd(t1, t2));
        a[1] = cadd(b1, d1);
        a[4] = csub(b1, d1);
        a[2] = cadd(b2, d2);
        a[3] = csub(b2, d2);
    }
};

// ------------------------------------------------------------ plans ------
template <int... Rs>
struct Radices {};

template <int N>
struct FftPlan;
#define MFT_PLAN(N_, T_, ...)                                                   \
    template <>                                                                 \
    struct FftPlan<N_> {                                                        \
        static constexpr int N = N_;                                            \
        static constexpr int T = T_;                                            \
        static constexpr int E = N_ / T_;                                       \
        using R = Radices<__VA_ARGS__>;                                         \
    };
MFT_PLAN(2, 1, 2)
MFT_PLAN(3, 1, 3)
MFT_PLAN(4, 1, 4)
MFT_PLAN(5, 1, 5)
MFT_PLAN(6, 1, 2, 3)
MFT_PLAN(8, 1, 8)
MFT_PLAN(10, 1, 2, 5)
MFT_PLAN(12, 1, 4, 3)
MFT_PLAN(16, 2, 8, 2)
MFT_PLAN(20, 1, 4, 5)
MFT_PLAN(24, 2, 4, 3, 2)
MFT_PLAN(32, 4, 8, 4)
MFT_PLAN(40, 2, 4, 2, 5)
MFT_PLAN(48, 4, 4, 4, 3)
MFT_PLAN(64, 8, 8, 8)
MFT_PLAN(80, 4, 4, 4, 5)
MFT_PLAN(96, 8, 4, 4, 3, 2)
MFT_PLAN(128, 16, 8, 8, 2)
MFT_PLAN(160, 8, 4, 4, 2, 5)
MFT_PLAN(192, 16, 4, 4, 4, 3)
MFT_PLAN(256, 32, 8, 8, 4)
#undef MFT_PLAN

// Plans with four elements per thread for kernels that keep several lines
// in registers at once (the fused k-space pass): more threads per line,
// fewer registers per thread, one more LDS pass.  Sizes without one use the
// default plan.
template <int N>
struct FftPlanLow : FftPlan<N> {};
#define MFT_PLAN_LOW(N_, T_, ...)                                               \
    template <>                                                                 \
    struct FftPlanLow<N_> {                                                     \
        static constexpr int N = N_;                                            \
        static constexpr int T = T_;                                            \
        static constexpr int E = N_ / T_;                                       \
        using R = Radices<__VA_ARGS__>;                                         \
    };
MFT_PLAN_LOW(32, 8, 4, 4, 2)
MFT_PLAN_LOW(64, 16, 4, 4, 4)
MFT_PLAN_LOW(128, 32, 4, 4, 4, 2)
MFT_PLAN_LOW(256, 64, 4, 4, 4, 4)
#undef MFT_PLAN_LOW

// Expands F(N) for every supported size; used by host dispatch switches.
#define MFT_FOR_EACH_FFT_SIZE(F)                                                 \
    F(2) F(3) F(4) F(5) F(6) F(8) F(10) F(12) F(16) F(20) F(24) F(32) F(40)    \
        F(48) F(64) F(80) F(96) F(128) F(160) F(192) F(256)

// LDS addressing of element e of local line l: lines contiguous (x axis) or
// interleaved so that neighbouring lanes hit neighbouring words (y, z axes).
// swz (contiguous lines): element e sits at e + e/8 within a line pitch of at
// least swz_pitch(N).  A Stockham exchange writes element R j + r from thread
// j, so without the skew the lanes of one write land R elements (2 R words)
// apart and all lines start on the same bank: a few banks take every word.
// One skipped word every eight elements spreads those lanes over the banks.
struct LdsMap {
    int line;
    int nlines; // lines per workgroup (B)
    bool interleaved;
    int n;
    bool swz = false;
    __device__ __forceinline__ int operator()(int e) const
    {
        return interleaved ? e * nlines + line : line * n + (swz ? e + (e >> 3) : e);
    }
};

// line pitch that holds a skewed line of N elements
__host__ __device__ constexpr int swz_pitch(int N) { return N + (N + 7) / 8; }

// Exchange ordering.  When every thread of a line sits in one wavefront
// (T consecutive threads, T <= 64, not interleaved) the exchange only orders
// the wave's own LDS accesses, which execute in issue order: a compiler-level
// barrier suffices.  Otherwise the workgroup synchronises.
template <bool WAVE>
__device__ __forceinline__ void exchange_barrier()
{
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

template <int N, int T, int R, int NS, int SIGN, bool LAST, bool WAVE = false>
__device__ __forceinline__ void fft_pass(float2 (&v)[N / T], float2* lds, const float2* tw, int t,
                                         const LdsMap& map)
{
    constexpr int E = N / T;
    constexpr int NB = E / R;
    static_assert(E % R == 0, "radix must divide the elements per thread");
    float2 out[E];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        float2 a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = v[b + r * NB];
        const int j = t + T * b;
        const int k = j % NS;
        if (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[r * k * (N / (NS * R))];
                if (SIGN > 0) w.y = -w.y;
                a[r] = cmul(a[r], w);
            }
        }
        Dft<R, SIGN>::run(a);
#pragma unroll
        for (int r = 0; r < R; ++r) out[b + r * NB] = a[r];
    }
    if (LAST) {
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = out[s];
    } else {
        exchange_barrier<WAVE>(); // every thread finished reading the previous layout
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + T * b;
            const int k = j % NS;
            const int d = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[map(d + r * NS)] = out[b + r * NB];
        }
        exchange_barrier<WAVE>();
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = lds[map(t + T * s)];
    }
}

template <int N, int T, int SIGN, bool WAVE, int NS, int R, int... Rest>
__device__ __forceinline__ void fft_passes(float2 (&v)[N / T], float2* lds, const float2* tw, int t,
                                           const LdsMap& map)
{
    constexpr bool LAST = sizeof...(Rest) == 0;
    fft_pass<N, T, R, NS, SIGN, LAST, WAVE>(v, lds, tw, t, map);
    if constexpr (!LAST) fft_passes<N, T, SIGN, WAVE, NS * R, Rest...>(v, lds, tw, t, map);
}

template <int N, int T, int SIGN, bool WAVE = false, int... Rs>
__device__ __forceinline__ void fft_run(float2 (&v)[N / T], float2* lds, const float2* tw, int t,
                                        const LdsMap& map, Radices<Rs...>)
{
    fft_passes<N, T, SIGN, WAVE, 1, Rs...>(v, lds, tw, t, map);
}

// Unnormalised line transform; SIGN = -1 forward, +1 inverse.  WAVE: the
// line's T threads are consecutive threads of one wavefront and the LDS map
// is not interleaved (exchanges need no workgroup barrier).
template <int N, int SIGN, bool WAVE = false>
__device__ __forceinline__ void line_fft(float2 (&v)[N / FftPlan<N>::T], float2* lds, const float2* tw,
                                         int t, const LdsMap& map)
{
    using P = FftPlan<N>;
    static_assert(!WAVE || (P::T <= 64 && 64 % P::T == 0), "a line must fit in one wavefront");
    fft_run<N, P::T, SIGN, WAVE>(v, lds, tw, t, map, typename P::R{});
}

// The same transform with an explicit plan.
template <typename Plan, int SIGN>
__device__ __forceinline__ void line_fft_plan(float2 (&v)[Plan::E], float2* lds, const float2* tw, int t,
                                              const LdsMap& map)
{
    fft_run<Plan::N, Plan::T, SIGN>(v, lds, tw, t, map, typename Plan::R{});
}

// Geometry of the lines along one axis of an x-fastest (nx, ny, nz) grid.
struct AxisGeom {
    int nx, ny, nz;
    int axis; // 0 = x, 1 = y, 2 = z
    __host__ __device__ int nlines() const
    {
        return axis == 0 ? ny * nz : (axis == 1 ? nx * nz : nx * ny);
    }
    __host__ __device__ int length() const { return axis == 0 ? nx : (axis == 1 ? ny : nz); }
    __host__ __device__ long stride() const { return axis == 0 ? 1 : (axis == 1 ? nx : (long)nx * ny); }
    // offset of element 0 of line l; also returns the two transverse indices
    __host__ __device__ long base(int l, int* a, int* b) const
    {
        if (axis == 0) {
            *a = l % ny;
            *b = l / ny;
            return (long)l * nx;
        }
        if (axis == 1) {
            *a = l % nx;
            *b = l / nx;
            return (long)*a + (long)nx * ny * *b;
        }
        *a = l % nx;
        *b = l / nx;
        return l;
    }
};

} // namespace mft
