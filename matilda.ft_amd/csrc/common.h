// Shared device/host helpers for the MI355X engine: status handling,
// counter-based RNG, reference k-vector arithmetic, reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mft.h"

namespace mft {

// Block index -> work item such that each of the 8 XCDs (blocks are
// dispatched to them round-robin) takes one contiguous slab of the n items:
// neighbouring items share that XCD's L2 and address translations.
__device__ __forceinline__ int xcd_slab(int blk, int n)
{
    const int per = (n + 7) >> 3;
    const int x = blk & 7, j = blk >> 3;
    const int full = n - 8 * (per - 1); // XCDs with `per` items
    const int base = x < full ? x * per : full * per + (x - full) * (per - 1);
    return base + j;
}

// ----------------------------------------------------------- status ------
void set_error(int code, const char* fmt, ...);
int last_code();

#define MFT_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::mft::set_error(MFT_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,            \
                             hipGetErrorString(e_), __FILE__, __LINE__);                \
            return MFT_ERR_DEVICE;                                                      \
        }                                                                               \
    } while (0)

#define MFT_LAUNCH_CHECK()                                                              \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) {                                                         \
            ::mft::set_error(MFT_ERR_DEVICE, "kernel launch failed: %s (%s:%d)",        \
                             hipGetErrorString(e_), __FILE__, __LINE__);                \
            return MFT_ERR_DEVICE;                                                      \
        }                                                                               \
    } while (0)

#define MFT_REQUIRE(cond, code, ...)                                                    \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::mft::set_error(code, __VA_ARGS__);                                        \
            return code;                                                                \
        }                                                                               \
    } while (0)

// ---------------------------------------------------------------- RNG ----
// Philox4x32-10 (Salmon et al., SC'11) and the Box-Muller mapping shared with
// the CPU oracle: u1 = ((r0 >> 8) + 1) 2^-24, u2 = (r1 >> 8) 2^-24.
constexpr uint32_t TAG_TILD = 0x54494C44u; // 'TILD'
constexpr uint32_t TAG_LANGEVIN = 0x4C414E47u; // 'LANG'
constexpr uint32_t TAG_DPD = 0x44504450u;      // 'DPDP'
constexpr uint32_t TAG_FTS = 0x4654534Eu;  // 'FTSN'

__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ inline void normal4(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed,
                               float z[4])
{
    uint32_t c[4] = {c0, c1, c2, c3};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float u1 = (float)((c[2 * h] >> 8) + 1u) * 5.9604644775390625e-8f;
        float u2 = (float)(c[2 * h + 1] >> 8) * 5.9604644775390625e-8f;
        float rad = sqrtf(-2.0f * logf(u1));
        float s, co;
        sincospif(2.0f * u2, &s, &co);
        z[2 * h] = rad * co;
        z[2 * h + 1] = rad * s;
    }
}

// ------------------------------------------------------------ k grid ----
// Box::get_kD (Box.cu:157-203): k_i = PI2*n/L for n < N/2, else PI2*(n-N)/L,
// with PI2 = 6.2831853071f (constants.h:9), all in float, so the Nyquist
// index maps to the negative frequency.
__host__ __device__ inline float kcomp(int n, int N, float L)
{
    const float PI2 = 6.2831853071f;
    float t = (2 * n < N) ? PI2 * (float)n : PI2 * (float)(n - N);
    return t / L;
}

// ------------------------------------------------------- reductions ----
template <typename T>
__device__ inline T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

} // namespace mft
