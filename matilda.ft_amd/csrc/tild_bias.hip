// Static bias fields of the TILD step (BiasField, ps_potentialBias.cu): an
// external field A u(r) seeding a target phase on the particles of one group.
// The field is built on the host once (Box::make_bias_field, Box.cu:308-458),
// its force -grad u is taken spectrally once on the device (the reference's
// initializePotential, ps_potentialBias.cu:86-185), and every step adds that
// force field to the force fields of the group's species before the gather.
#include <cmath>
#include <vector>

#include "tild_impl.h"

namespace mft {

// A u(r) at grid point (ix, iy, iz), r = n dx (Box::get_r, Box.cu:206-218),
// the reference's float arguments, PI = 3.141592654f (constants.h)
static double bias_value(const mft_tild_potential& P, const float r[3], const float L[3])
{
    const float PI = 3.141592654f;
    const double A = P.A;
    const float n = (float)P.n_periods;
    if (P.phase == MFT_BIAS_LAM) {
        const float arg = 2.0f * PI * n / L[P.dir];
        return A * std::sin((double)(arg * r[P.dir]));
    }
    if (P.phase == MFT_BIAS_HEX) {
        // cylinders along dir: the two other axes carry the cosines
        int d1 = 0, d2 = 1;
        if (P.dir == d1) d1 = 2;
        else if (P.dir == d2) d2 = 2;
        const float a1 = 2.0f * PI * n / L[d1], a2 = 2.0f * PI * n / L[d2];
        return A * std::cos((double)(a1 * r[d1])) * std::cos((double)(a2 * r[d2]));
    }
    // gyroid: the reference keeps the sines, cosines and their sum in float
    float c[3], s[3];
    for (int j = 0; j < 3; ++j) {
        const float a = 2.0f * PI * n / L[j];
        c[j] = (float)std::cos((double)(a * r[j]));
        s[j] = (float)std::sin((double)(a * r[j]));
    }
    const float e = s[0] * c[1] + s[1] * c[2] + s[2] * c[0];
    return A * ((double)(e * e) - 1.44);
}

__global__ void k_bias_to_complex(float2* __restrict__ a, const float* __restrict__ u, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) a[i] = make_float2(u[i], 0.f);
}

// out = -i k_j S (fk = -I kv[j] uk, ps_potentialBias.cu:119-124), k of
// Box::get_kD with the Nyquist index negative and kept
__global__ void k_bias_grad(float2* __restrict__ out, const float2* __restrict__ S, int j, int nx, int ny, int nz,
                            float Lx, float Ly, float Lz)
{
    const long M = (long)nx * ny * nz;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int ix = (int)(i % nx), iy = (int)((i / nx) % ny), iz = (int)(i / ((long)nx * ny));
    const float k = j == 0 ? kcomp(ix, nx, Lx) : (j == 1 ? kcomp(iy, ny, Ly) : kcomp(iz, nz, Lz));
    const float2 v = S[i];
    out[i] = make_float2(k * v.y, -k * v.x);
}

__global__ void k_bias_store(float4* __restrict__ F, const float2* __restrict__ a, int j, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    float4 f = F[i];
    if (j == 0) f.x = a[i].x;
    else if (j == 1) f.y = a[i].x;
    else f.z = a[i].x;
    F[i] = f;
}

static int fft3_inplace(float2* a, const mft_tild_desc& d, int sign)
{
    const long M = (long)d.nx * d.ny * d.nz;
    const float s = sign < 0 ? (float)(1.0 / (double)M) : 1.0f;
    for (int ax = 0; ax < 3; ++ax) {
        AxisGeom g{d.nx, d.ny, d.nz, ax};
        const int rc = tild_axis(sign, 0, 0, a, a, g, 1, M, ax == 0 ? s : 1.0f, nullptr);
        if (rc) return rc;
    }
    return MFT_OK;
}

int tild_bias_setup(TildEngine* e, int pot)
{
    const mft_tild_desc& d = e->d;
    const mft_tild_potential& P = e->pots[pot];
    const long M = (long)d.nx * d.ny * d.nz;
    const float L[3] = {d.lx, d.ly, d.lz};
    std::vector<float> u(M);
    for (long i = 0; i < M; ++i) {
        const int n[3] = {(int)(i % d.nx), (int)((i / d.nx) % d.ny), (int)(i / ((long)d.nx * d.ny))};
        const float r[3] = {(float)n[0] * e->dx[0], (float)n[1] * e->dx[1], (float)n[2] * e->dx[2]};
        u[i] = (float)bias_value(P, r, L); // ur[i] = urDD[i] (ps_potentialBias.cu:96-98)
    }
    TildEngine::Bias b{pot, P.maskI, nullptr, nullptr};
    int rc;
    if ((rc = dalloc(&b.F, (size_t)M))) return rc;
    if ((rc = dalloc(&b.u, (size_t)M))) return rc;
    e->bias.push_back(b);
    MFT_HIP(hipMemcpy(b.u, u.data(), sizeof(float) * M, hipMemcpyHostToDevice));
    MFT_HIP(hipMemset(b.F, 0, sizeof(float4) * M));
    float2 *S = nullptr, *W = nullptr;
    if ((rc = dalloc(&S, (size_t)M))) return rc;
    if ((rc = dalloc(&W, (size_t)M))) {
        (void)hipFree(S);
        return rc;
    }
    k_bias_to_complex<<<nblk(M), kBlock>>>(S, b.u, M);
    rc = fft3_inplace(S, d, -1);
    for (int j = 0; j < 3 && !rc; ++j) {
        k_bias_grad<<<nblk(M), kBlock>>>(W, S, j, d.nx, d.ny, d.nz, d.lx, d.ly, d.lz);
        if ((rc = fft3_inplace(W, d, +1))) break;
        k_bias_store<<<nblk(M), kBlock>>>(b.F, W, j, M);
    }
    const hipError_t err = hipDeviceSynchronize();
    (void)hipFree(S);
    (void)hipFree(W);
    if (rc) return rc;
    MFT_REQUIRE(err == hipSuccess, MFT_ERR_DEVICE, "bias field set-up: %s", hipGetErrorString(err));
    return MFT_OK;
}

// F: an engine force field (12 bytes a point); B: the bias force (float4)
__global__ void k_bias_add(float4* __restrict__ F, const float4* __restrict__ B, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const float3 f = ff_load(F, i);
    const float4 b = B[i];
    ff_store(F, i, f.x + b.x, f.y + b.y, f.z + b.z);
}

// accumulateGridForceComp of the bias on its group (ps_potentialBias.cu:200-211)
int tild_bias_add(TildEngine* e)
{
    const long M = e->M;
    for (const auto& b : e->bias)
        for (int s = 0; s < e->d.nspecies; ++s) {
            if (!((b.mask >> s) & 1u)) continue;
            const int f = e->force_slot_h[s];
            MFT_REQUIRE(f >= 0, MFT_ERR_INVALID, "biased species %d has no force field", s);
            k_bias_add<<<nblk(M), kBlock, 0, e->stream>>>(e->F4 + (long)f * M, b.F, M);
        }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// FP64 block sums of sum_{s in mask} rho_s u
__global__ void k_bias_energy(double* __restrict__ part, const float* __restrict__ rho, const float* __restrict__ u,
                              uint32_t mask, int nspecies, long M)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < M) {
        double r = 0.0;
        for (int s = 0; s < nspecies; ++s)
            if ((mask >> s) & 1u) r += (double)rho[(long)s * M + i];
        v = r * (double)u[i];
    }
    v = wave_sum(v);
    __shared__ double red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

// BiasField::CalcEnergy (ps_potentialBias.cu:215-228): gvol sum rho_I A u
int tild_bias_energy(TildEngine* e, int k, double* energy)
{
    int rc;
    if (!e->rho_valid && (rc = tild_density_reduce(e))) return rc;
    const long M = e->M;
    const unsigned nb = nblk(M);
    double* part = nullptr;
    if ((rc = dalloc(&part, nb))) return rc;
    const auto& b = e->bias[k];
    k_bias_energy<<<nb, kBlock, 0, e->stream>>>(part, e->rho, b.u, b.mask, e->d.nspecies, M);
    std::vector<double> h(nb);
    hipError_t err = hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    (void)hipFree(part);
    MFT_REQUIRE(err == hipSuccess, MFT_ERR_DEVICE, "bias energy: %s", hipGetErrorString(err));
    double s = 0.0;
    for (double v : h) s += v;
    *energy = (double)e->gvol * s;
    return MFT_OK;
}

} // namespace mft

extern "C" int mft_bias_field(double* u, const int Nx[3], const float L[3], int phase, double A, int dir,
                              int n_periods)
{
    MFT_REQUIRE(u && Nx && L, MFT_ERR_INVALID, "null argument");
    MFT_REQUIRE(phase >= MFT_BIAS_LAM && phase <= MFT_BIAS_GYR, MFT_ERR_INVALID,
                "bias phase %d is not lamellar, cylinder or gyroid", phase);
    MFT_REQUIRE(dir >= 0 && dir < 3 && n_periods > 0 && Nx[0] > 0 && Nx[1] > 0 && Nx[2] > 0, MFT_ERR_INVALID,
                "bad bias field arguments");
    mft_tild_potential P{};
    P.phase = phase;
    P.A = 1.0f;
    P.dir = dir;
    P.n_periods = n_periods;
    const float dx[3] = {L[0] / (float)Nx[0], L[1] / (float)Nx[1], L[2] / (float)Nx[2]};
    const long M = (long)Nx[0] * Nx[1] * Nx[2];
    for (long i = 0; i < M; ++i) {
        const int n[3] = {(int)(i % Nx[0]), (int)((i / Nx[0]) % Nx[1]), (int)(i / ((long)Nx[0] * Nx[1]))};
        const float r[3] = {(float)n[0] * dx[0], (float)n[1] * dx[1], (float)n[2] * dx[2]};
        // the prefactor enters in double, as Box::make_bias_field's Ao
        u[i] = A * mft::bias_value(P, r, L);
    }
    return MFT_OK;
}
