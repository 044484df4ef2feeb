// Fused k-space pass of the TILD pipeline: forward z FFT of every density
// channel, all potentials applied as one per-species kernel, inverse z FFT of
// each force component; optional potential energies by Parseval.
#include "tild_impl.h"

namespace mft {


// Potential shape at |k|^2 = k2 for grid point idx (float, reference forms).
__device__ __forceinline__ float pot_shape(const KPot& P, float k2, long idx)
{
    if (P.kind == MFT_POT_GAUSSIAN) return P.A * expf(-k2 * P.sig2 * 0.5f);
    return P.A * P.table[idx];
}


// Fused z pass: forward z FFT of every density channel (already transformed
// along x and y), k-space forces, inverse z FFT of each output component.
template <int N, int NCH>
__device__ __forceinline__ void conv_z_body(const float2* __restrict__ S, float2* __restrict__ Y, int nx, int ny,
                                            float Lx, float Ly, float Lz, const KParams& kp,
                                            const float2* __restrict__ twg, float invM, double* __restrict__ epart,
                                            float Mgvol)
{
    using P = FftPlanLow<N>;
    constexpr int T = P::T, E = P::E, B = 256 / T;
    __shared__ float2 lds[B * N];
    __shared__ float2 tw[N];
    __shared__ double ered[4 * (kMaxPots + 1)];
    for (int i = threadIdx.x; i < N; i += 256) tw[i] = twg[i];
    const int lline = threadIdx.x % B;
    const int t = threadIdx.x / B;
    // consecutive line groups on one XCD: a group's B lines are half of a
    // 128-byte row of x, and the group beside it reads the other half, so both
    // halves are fetched into the same L2 (brick_of_block's mapping)
    const int line = brick_of_block(blockIdx.x, gridDim.x) * B + lline;
    const long M = (long)nx * ny * N;
    const bool valid = line < nx * ny;
    const int ix = line % nx, iy = line / nx;
    const long st = (long)nx * ny;
    LdsMap map{lline, B, true, N};

    float2 spec[NCH][E];
    // every channel's loads in flight before the first transform
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
        for (int s = 0; s < E; ++s)
            spec[c][s] = (valid && c < kp.nchan) ? S[c * M + line + (long)(t + T * s) * st] : make_float2(0.f, 0.f);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        if (c >= kp.nchan) continue;
        line_fft_plan<P, -1>(spec[c], lds, tw, t, map);
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

    int out = 0;        // packed output field
    float2 pend[E];     // even component waiting for its partner
    bool pending = false;
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
        } else if (kp.ngroups >= 0) {
            for (int g = 0; g < kp.ngroups; ++g) {
                const KGroup& G = kp.grp[g];
#pragma unroll
                for (int s = 0; s < E; ++s) {
                    float2 src = make_float2(0.f, 0.f);
#pragma unroll
                    for (int c = 0; c < NCH; ++c) {
                        if (c >= kp.nchan) continue;
                        const float m = G.coef[fi][c];
                        src.x += m * spec[c][s].x;
                        src.y += m * spec[c][s].y;
                    }
                    float u;
                    if (G.kind == MFT_POT_GAUSSIAN) u = expf(-k2[s] * G.sig2 * 0.5f);
                    else if (G.kind == MFT_POT_CHARGES) u = k2[s] > 0.f ? expf(-k2[s] * G.sig2) / k2[s] : 0.f;
                    else u = G.table[ix + (long)nx * (iy + (long)ny * (t + T * s))];
                    Z[s].x += u * src.x;
                    Z[s].y += u * src.y;
                }
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
            // V = -i k_j Z, with the Nyquist plane of axis j dropped: there
            // V is anti-Hermitian and adds nothing to the real part the force
            // field keeps, so V is the spectrum of a real field and two such
            // components share one complex inverse transform (re, im)
            float2 v[E];
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const bool nyq = j == 0 ? 2 * ix == nx : (j == 1 ? 2 * iy == ny : 2 * (t + T * s) == N);
                const float kj = nyq ? 0.f : (j == 0 ? kx : (j == 1 ? ky : kz[s]));
                v[s] = make_float2(kj * Z[s].y, -kj * Z[s].x);
            }
            if (!pending) {
#pragma unroll
                for (int s = 0; s < E; ++s) pend[s] = v[s];
                pending = true;
                continue;
            }
#pragma unroll
            for (int s = 0; s < E; ++s) v[s] = make_float2(pend[s].x - v[s].y, pend[s].y + v[s].x);
            pending = false;
            line_fft_plan<P, +1>(v, lds, tw, t, map);
            if (valid) {
#pragma unroll
                for (int s = 0; s < E; ++s) Y[out * M + line + (long)(t + T * s) * st] = v[s];
            }
            ++out;
        }
    }
    if (pending) {
        line_fft_plan<P, +1>(pend, lds, tw, t, map);
        if (valid) {
#pragma unroll
            for (int s = 0; s < E; ++s) Y[out * M + line + (long)(t + T * s) * st] = pend[s];
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
                        KPot Pe = Pp;
                        Pe.A = Pp.eA;
                        u = pot_shape(Pe, k2[s], kidx);
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

// Up to three channels the pass fits 128 registers (four waves per SIMD);
// more channels keep the register count they need.
template <int N, int NCH>
__global__ void __launch_bounds__(256, 4) k_conv_z(const float2* __restrict__ S, float2* __restrict__ Y, int nx, int ny,
                                                float Lx, float Ly, float Lz, KParams kp, const float2* __restrict__ twg,
                                                float invM, double* __restrict__ epart, float Mgvol)
{
    conv_z_body<N, NCH>(S, Y, nx, ny, Lx, Ly, Lz, kp, twg, invM, epart, Mgvol);
}

template <int N, int NCH>
__global__ void __launch_bounds__(256) k_conv_z_wide(const float2* __restrict__ S, float2* __restrict__ Y, int nx,
                                                     int ny, float Lx, float Ly, float Lz, KParams kp,
                                                     const float2* __restrict__ twg, float invM,
                                                     double* __restrict__ epart, float Mgvol)
{
    conv_z_body<N, NCH>(S, Y, nx, ny, Lx, Ly, Lz, kp, twg, invM, epart, Mgvol);
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

int tild_conv_blocks(int nx, int ny, int nz)
{
    int T = 1;
#define MFT_T(n_) if (nz == n_) T = FftPlanLow<n_>::T;
    MFT_FOR_EACH_FFT_SIZE(MFT_T)
#undef MFT_T
    int B = 256 / T;
    return (nx * ny + B - 1) / B;
}

// lengths whose low-register plan keeps the two- and three-channel pass in
// 128 registers without spilling
template <int N>
constexpr bool conv_four_waves()
{
    return N == 32 || N == 64 || N == 128;
}

template <int N>
static void launch_conv(TildEngine* e, float invM)
{
    const mft_tild_desc& d = e->d;
    constexpr int B = 256 / FftPlanLow<N>::T;
    const unsigned grid = (d.nx * d.ny + B - 1) / B;
    const float Mg = (float)((double)e->gvol * e->M);
    const float2* tw = fft_twiddles_device(N);
    bool done = false;
    if constexpr (conv_four_waves<N>()) {
        if (e->nchan <= 2) {
            k_conv_z<N, 2><<<grid, 256, 0, e->stream>>>(e->S, e->Y, d.nx, d.ny, d.lx, d.ly, d.lz, e->kp, tw, invM,
                                                        e->epart, Mg);
            done = true;
        } else if (e->nchan == 3) {
            k_conv_z<N, 3><<<grid, 256, 0, e->stream>>>(e->S, e->Y, d.nx, d.ny, d.lx, d.ly, d.lz, e->kp, tw, invM,
                                                        e->epart, Mg);
            done = true;
        }
    }
    if (done)
        ;
    else if (e->nchan <= 2)
        k_conv_z_wide<N, 2><<<grid, 256, 0, e->stream>>>(e->S, e->Y, d.nx, d.ny, d.lx, d.ly, d.lz, e->kp, tw, invM,
                                                         e->epart, Mg);
    else if (e->nchan == 3)
        k_conv_z_wide<N, 3><<<grid, 256, 0, e->stream>>>(e->S, e->Y, d.nx, d.ny, d.lx, d.ly, d.lz, e->kp, tw, invM,
                                                         e->epart, Mg);
    else if (e->nchan <= 4)
        k_conv_z_wide<N, 4><<<grid, 256, 0, e->stream>>>(e->S, e->Y, d.nx, d.ny, d.lx, d.ly, d.lz, e->kp, tw, invM,
                                                         e->epart, Mg);
    else
        k_conv_z_wide<N, 8><<<grid, 256, 0, e->stream>>>(e->S, e->Y, d.nx, d.ny, d.lx, d.ly, d.lz, e->kp, tw, invM,
                                                         e->epart, Mg);
    if (e->kp.energy) k_energy_sum<<<d.npotentials, 256, 0, e->stream>>>(e->eout, e->epart, grid, d.npotentials);
}

int tild_conv(TildEngine* e)
{
    const mft_tild_desc& d = e->d;
    float invM = 1.0f / (float)e->M;
    switch (d.nz) {
#define MFT_CASE(n_)                                                                                      \
    case n_: launch_conv<n_>(e, invM); break;
        MFT_FOR_EACH_FFT_SIZE(MFT_CASE)
#undef MFT_CASE
    default: return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

} // namespace mft
