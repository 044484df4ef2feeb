// Status handling, device helpers and RNG entry points of the C ABI.
#include <cstdarg>
#include <cstring>

#include "common.h"

namespace mft {

static thread_local char g_msg[1024] = "";
static thread_local int g_code = MFT_OK;

void set_error(int code, const char* fmt, ...)
{
    g_code = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
}

int last_code() { return g_code; }

__global__ void k_philox(uint32_t* __restrict__ raw, float* __restrict__ nrm, int64_t nblocks, uint64_t seed,
                         uint32_t c1, uint32_t c2, uint32_t c3)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nblocks) return;
    if (raw) {
        uint32_t c[4] = {(uint32_t)i, c1, c2, c3};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        raw[4 * i + 0] = c[0];
        raw[4 * i + 1] = c[1];
        raw[4 * i + 2] = c[2];
        raw[4 * i + 3] = c[3];
    } else {
        float z[4];
        normal4((uint32_t)i, c1, c2, c3, seed, z);
        nrm[4 * i + 0] = z[0];
        nrm[4 * i + 1] = z[1];
        nrm[4 * i + 2] = z[2];
        nrm[4 * i + 3] = z[3];
    }
}

} // namespace mft

using namespace mft;

extern "C" {

const char* mft_last_error(void) { return g_msg; }

int mft_version(void) { return 1; }

int mft_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    MFT_REQUIRE(free_bytes && total_bytes, MFT_ERR_INVALID, "null argument");
    MFT_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return MFT_OK;
}

int mft_device_count(int* n)
{
    MFT_REQUIRE(n, MFT_ERR_INVALID, "null argument");
    *n = 0;
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) {
        *n = 0;
        (void)hipGetLastError();
    }
    return MFT_OK;
}

int mft_malloc(void** dptr, size_t bytes)
{
    MFT_REQUIRE(dptr, MFT_ERR_INVALID, "null argument");
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        set_error(MFT_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return MFT_ERR_OOM;
    }
    return MFT_OK;
}

int mft_free(void* dptr)
{
    if (dptr) MFT_HIP(hipFree(dptr));
    return MFT_OK;
}

int mft_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    MFT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MFT_OK;
}

int mft_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    MFT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MFT_OK;
}

int mft_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    MFT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return MFT_OK;
}

int mft_memset(void* dst, int value, size_t bytes)
{
    MFT_HIP(hipMemset(dst, value, bytes));
    return MFT_OK;
}

int mft_synchronize(void)
{
    MFT_HIP(hipDeviceSynchronize());
    return MFT_OK;
}

int mft_stream_create(void** stream)
{
    MFT_REQUIRE(stream, MFT_ERR_INVALID, "null argument");
    hipStream_t s;
    MFT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return MFT_OK;
}

int mft_stream_destroy(void* stream)
{
    if (stream) MFT_HIP(hipStreamDestroy((hipStream_t)stream));
    return MFT_OK;
}

int mft_stream_synchronize(void* stream)
{
    MFT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return MFT_OK;
}

int mft_philox_raw(uint32_t* out, int64_t nblocks, uint64_t seed, uint32_t c1, uint32_t c2, uint32_t c3,
                   void* stream)
{
    MFT_REQUIRE(out && nblocks >= 0, MFT_ERR_INVALID, "bad argument");
    if (nblocks == 0) return MFT_OK;
    k_philox<<<(unsigned)((nblocks + 255) / 256), 256, 0, (hipStream_t)stream>>>(out, nullptr, nblocks, seed, c1, c2,
                                                                                  c3);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int mft_philox_normal(float* out, int64_t nblocks, uint64_t seed, uint32_t c1, uint32_t c2, uint32_t c3,
                      void* stream)
{
    MFT_REQUIRE(out && nblocks >= 0, MFT_ERR_INVALID, "bad argument");
    if (nblocks == 0) return MFT_OK;
    k_philox<<<(unsigned)((nblocks + 255) / 256), 256, 0, (hipStream_t)stream>>>(nullptr, out, nblocks, seed, c1, c2,
                                                                                  c3);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

} // extern "C"
