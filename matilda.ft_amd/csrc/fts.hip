// Complex-Langevin FTS step on gfx950 (FTS_Box::doTimeStep, FTS_Box.cu:19-123).
//
// Per iteration and replica:
//   update: every potential's EM step with real CL noise from the previous
//     densities (fts_potential_*.cu updateFields), species fields
//     W_s (fts_species.cu:53-132) and E_s = exp(-W_s), the first propagator
//     slice of every chain, and FP64 partial sums of the Hamiltonian field
//     terms -- one elementwise pass;
//   propagators as the reference's q_s = e^{-W_b(s)} F^-1[g F q_{s-1}]
//     (fts_molecule_linear.cu:144-269), D = F^-1 g F applied axis by axis
//     because g(k) = exp(-k^2/(Nr-1)) factorises over kx, ky, kz: the x-y pass
//     reads q_{s-1} alone and the z pass multiplies by E_b(s), so a bead costs
//     two reads and two writes of a slice plus one read of E; the forward and
//     complementary chains of every replica share each launch;
//   density: rho_b ~ e^{W_b} sum_s q_s qdag_{N-1-s} (fts_molecule_linear.cu:
//     298-327) and Q = gvol/V sum q_{N-1}, FP64 accumulation; the n/(QV)
//     factor is applied by the next update.
// Replicas with equal grids form one group; every launch covers a group.
#include <cmath>
#include <algorithm>
#include <array>
#include <complex>
#include <cstring>
#include <cstdlib>
#include <deque>
#include <map>
#include <vector>

#include "common.h"
#include "fft_core.h"

namespace mft {

const float2* fft_twiddles_device(int n); // fft.hip

constexpr int kFtsMaxSpecies = 8;
constexpr int kFtsMaxPots = 8;
constexpr int kFtsMaxMol = 4;
constexpr int kFtsMaxBlocks = 16;
constexpr int kFtsThreads = 256;
constexpr int kFtsMaxSnapshots = 8;

struct FtsPotDev {
    int kind, I, J;
    int scheme;               // MFT_FTS_EM or MFT_FTS_EMPC
    double cpl, cmi;          // force coefficients of w+ and w-
    double cterm;             // constant imaginary force: C (Helfand, Incompress), 0 (Edwards)
    double dt_pl, dt_mi;
    double nmag_pl, nmag_mi;  // sqrt(2 dt / gvol) or 0 (no noise)
    int col_pl, col_mi;       // noise normal index or -1
    float2 *wpl, *wmi;
    // EMPC predictor data: fields before the predictor and its forces
    float2 *wpl_o, *wmi_o;
    double2 *fpl_o, *fmi_o;
    // 1S: the forces, for the k-space step that follows the pass
    float2 *f1pl, *f1mi;
};

struct FtsUpdDev {
    int M;
    int nspecies, npots, nblk;     // nblk = blocks over all molecules
    FtsPotDev pot[kFtsMaxPots];
    int blk_species[kFtsMaxMol * kFtsMaxBlocks];
    int blk_mol[kFtsMaxMol * kFtsMaxBlocks];
    float2* rhot;                  // [nblk][M] unscaled block densities
    const double2* Q;              // [nmol]
    double nmolecs[kFtsMaxMol];
    double activity[kFtsMaxMol];   // > 0: grand canonical, density factor = activity
    double V, C, Nr;
    float2* W;                     // [nspecies][M]
    float2* E;                     // [nspecies][M] exp(-W)
    int nchain;
    float2* chain_p0[2 * kFtsMaxMol];
    int chain_species0[2 * kFtsMaxMol];
    double* Hpart;                 // [gridDim.x][npots*6]
    uint64_t seed;
    uint32_t stream;
    const int64_t* iter;           // device counter (CL iteration)
    int cl;
    int first;                     // initialisation pass: no field update
    int corrector;                 // EMPC corrector pass (fts_potential_flory.cu:198-236)
    int have_rho;                  // densities exist (false before the first density pass)
};

// Register-array helpers with compile-time indices.  The selection is a
// multiply by 0 or 1 rather than a branch: with a wave-uniform index the
// compiler turns a select chain back into an indexed access to scratch
// memory.  (Adding 0 leaves a finite value unchanged.)
template <int N>
__device__ __forceinline__ float pick(const float (&a)[N], int k)
{
    float v = 0.f;
#pragma unroll
    for (int j = 0; j < N; ++j) v = fmaf(j == k ? 1.f : 0.f, a[j], v);
    return v;
}

template <int N>
__device__ __forceinline__ double2 pick(const double2 (&a)[N], int k)
{
    double2 v = make_double2(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const double m = j == k ? 1.0 : 0.0;
        v.x = fma(m, a[j].x, v.x);
        v.y = fma(m, a[j].y, v.y);
    }
    return v;
}

template <int N>
__device__ __forceinline__ float2 pick(const float2 (&a)[N], int k)
{
    float2 v = make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const float m = j == k ? 1.f : 0.f;
        v.x = fmaf(m, a[j].x, v.x);
        v.y = fmaf(m, a[j].y, v.y);
    }
    return v;
}

template <int N>
__device__ __forceinline__ void add_at(double2 (&a)[N], int k, double x, double y)
{
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const double m = j == k ? 1.0 : 0.0;
        a[j].x = fma(m, x, a[j].x);
        a[j].y = fma(m, y, a[j].y);
    }
}

// v times the per-axis diffusion factor g = hi + lo (exact in double) with a
// single rounding to float.  Rounding v*hi first and adding v*lo after rounds
// twice, and when the slice mean sits near a power of two (weak fields keep
// it near 1) the first rounding pulls the mean the same way every bead: log Q
// drifted by ~1e-6 over 20 beads that way.
__device__ __forceinline__ float2 scale_hilo(float2 v, float2 g)
{
    const double gd = (double)g.x + (double)g.y;
    return make_float2((float)((double)v.x * gd), (float)((double)v.y * gd));
}

// Elementwise update of one replica (blockIdx.y selects the replica).  NS
// and NP bound the species and potentials of every replica of the launch, so
// the per-thread arrays stay in registers.
template <int NS, int NP>
__global__ void __launch_bounds__(kFtsThreads) k_fts_update(const FtsUpdDev* __restrict__ descs)
{
    const FtsUpdDev& D = descs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < D.M;
    double hs[NP * 6];
#pragma unroll
    for (int k = 0; k < NP * 6; ++k) hs[k] = 0.0;
    if (valid) {
        // species densities of the previous iteration
        double2 rho[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) rho[s] = make_double2(0.0, 0.0);
        if (D.have_rho) {
            for (int b = 0; b < D.nblk; ++b) {
                int m = D.blk_mol[b];
                double2 Q = D.Q[m];
                // factor = n / (Q V)
                double den = (Q.x * Q.x + Q.y * Q.y) * D.V;
                double fr = D.nmolecs[m] * Q.x / den, fi = -D.nmolecs[m] * Q.y / den;
                if (D.activity[m] > 0.0) {
                    // fts_molecule_linear.cu:289-291
                    fr = D.activity[m];
                    fi = 0.0;
                }
                float2 r = D.rhot[(long)b * D.M + i];
                add_at(rho, D.blk_species[b], fr * r.x - fi * r.y, fr * r.y + fi * r.x);
            }
        }
        double2 rt = make_double2(0.0, 0.0);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            rt.x += rho[s].x;
            rt.y += rho[s].y;
        }
        float z[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) z[k] = 0.f;
        if (D.cl && !D.first) {
            int64_t it = *D.iter;
            normal4((uint32_t)i, (uint32_t)it, TAG_FTS, D.stream * 16u, D.seed, z);
            normal4((uint32_t)i, (uint32_t)it, TAG_FTS, D.stream * 16u + 1u, D.seed, z + 4);
        }
        double2 W[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) W[s] = make_double2(0.0, 0.0);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            if (p >= D.npots) break;
            const FtsPotDev& P = D.pot[p];
            float2 wp = P.wpl[i];
            double2 w = make_double2(wp.x, wp.y);
            if (P.kind == MFT_FTS_FLORY) {
                float2 wm = P.wmi[i];
                double2 v = make_double2(wm.x, wm.y);
                if (!D.first && !D.corrector && P.scheme == MFT_FTS_1S) {
                    // forces for the k-space step (fts_potential_flory.cu:172-193)
                    double2 rI = pick(rho, P.I), rJ = pick(rho, P.J);
                    P.f1pl[i] = make_float2((float)(P.cpl * w.x - (rI.y + rJ.y) / D.Nr),
                                            (float)(P.cpl * w.y + (rI.x + rJ.x) / D.Nr));
                    P.f1mi[i] = make_float2((float)(P.cmi * v.x + (rJ.x - rI.x) / D.Nr),
                                            (float)(P.cmi * v.y + (rJ.y - rI.y) / D.Nr));
                } else if (!D.first && (!D.corrector || P.scheme == MFT_FTS_EMPC)) {
                    // d_makeFloryForce (fts_potential_flory.cu:279-304)
                    double2 rI = pick(rho, P.I), rJ = pick(rho, P.J);
                    double fpx = P.cpl * w.x - (rI.y + rJ.y) / D.Nr;
                    double fpy = P.cpl * w.y + (rI.x + rJ.x) / D.Nr;
                    double fmx = P.cmi * v.x + (rJ.x - rI.x) / D.Nr;
                    double fmy = P.cmi * v.y + (rJ.y - rI.y) / D.Nr;
                    const double npl = P.col_pl >= 0 ? P.nmag_pl * (double)pick(z, P.col_pl) : 0.0;
                    const double nmi = P.col_mi >= 0 ? P.nmag_mi * (double)pick(z, P.col_mi) : 0.0;
                    if (!D.corrector) {
                        if (P.scheme == MFT_FTS_EMPC) {
                            // storePredictorData (fts_potential_flory.cu:243-249)
                            P.wpl_o[i] = wp;
                            P.wmi_o[i] = wm;
                            P.fpl_o[i] = make_double2(fpx, fpy);
                            P.fmi_o[i] = make_double2(fmx, fmy);
                        }
                        // d_fts_updateEM (fts_updateSchemes.cu:8-29), real noise
                        w.x -= P.dt_pl * fpx;
                        w.y -= P.dt_pl * fpy;
                        v.x -= P.dt_mi * fmx;
                        v.y -= P.dt_mi * fmy;
                        w.x += npl;
                        v.x += nmi;
                    } else {
                        // d_fts_updateEMPC (fts_updateSchemes.cu:56-81): the
                        // predictor's fields, forces and noise
                        const float2 wo = P.wpl_o[i], vo = P.wmi_o[i];
                        const double2 fo = P.fpl_o[i], go = P.fmi_o[i];
                        const double hp = 0.5 * P.dt_pl, hm = 0.5 * P.dt_mi;
                        w.x = (double)wo.x - hp * (fpx + fo.x) + npl;
                        w.y = (double)wo.y - hp * (fpy + fo.y);
                        v.x = (double)vo.x - hm * (fmx + go.x) + nmi;
                        v.y = (double)vo.y - hm * (fmy + go.y);
                    }
                    float2 wf = make_float2((float)w.x, (float)w.y), vf = make_float2((float)v.x, (float)v.y);
                    P.wpl[i] = wf;
                    P.wmi[i] = vf;
                    w = make_double2(wf.x, wf.y);
                    v = make_double2(vf.x, vf.y);
                }
                // species fields: -w- on I, +w- on J, i w+ on both
                add_at(W, P.I, -v.x - w.y, -v.y + w.x);
                add_at(W, P.J, v.x - w.y, v.y + w.x);
                hs[p * 6 + 0] += w.x * w.x - w.y * w.y;
                hs[p * 6 + 1] += 2.0 * w.x * w.y;
                hs[p * 6 + 4] += v.x * v.x - v.y * v.y;
                hs[p * 6 + 5] += 2.0 * v.x * v.y;
            } else {
                if (!D.first && (!D.corrector || P.scheme == MFT_FTS_EMPC)) {
                    double fx, fy;
                    if (P.kind == MFT_FTS_HELFAND) {
                        // d_makeHelfandForce (fts_potential_helfand.cu:223-240)
                        fx = P.cpl * w.x - rt.y / D.Nr;
                        fy = P.cpl * w.y + rt.x / D.Nr - P.cterm;
                    } else {
                        // d_makeIncompressForce (fts_potential_incompress.cu:202-218)
                        fx = -rt.y / D.Nr;
                        fy = rt.x / D.Nr - D.C;
                    }
                    if (P.scheme == MFT_FTS_1S) {
                        // the field changes in the k-space step that follows
                        P.f1pl[i] = make_float2((float)fx, (float)fy);
                    } else {
                        const double npl = P.col_pl >= 0 ? P.nmag_pl * (double)pick(z, P.col_pl) : 0.0;
                        if (!D.corrector) {
                            if (P.scheme == MFT_FTS_EMPC) {
                                P.wpl_o[i] = wp;
                                P.fpl_o[i] = make_double2(fx, fy);
                            }
                            w.x -= P.dt_pl * fx;
                            w.y -= P.dt_pl * fy;
                            w.x += npl;
                        } else {
                            const float2 wo = P.wpl_o[i];
                            const double2 fo = P.fpl_o[i];
                            const double hp = 0.5 * P.dt_pl;
                            w.x = (double)wo.x - hp * (fx + fo.x) + npl;
                            w.y = (double)wo.y - hp * (fy + fo.y);
                        }
                        float2 wf = make_float2((float)w.x, (float)w.y);
                        P.wpl[i] = wf;
                        w = make_double2(wf.x, wf.y);
                    }
                }
                // i w+ on every species
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    W[s].x -= w.y;
                    W[s].y += w.x;
                }
                hs[p * 6 + 0] += w.x * w.x - w.y * w.y;
                hs[p * 6 + 1] += 2.0 * w.x * w.y;
                hs[p * 6 + 2] += w.x;
                hs[p * 6 + 3] += w.y;
            }
        }
        double inv = 1.0 / D.Nr;
        float2 Es[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            Es[s] = make_float2(0.f, 0.f);
            if (s >= D.nspecies) continue;
            double wr = W[s].x * inv, wi = W[s].y * inv;
            D.W[(long)s * D.M + i] = make_float2((float)wr, (float)wi);
            // exp(-W): the bead weight of the q-form (fts_molecule_linear.cu:153-199)
            double a = exp(-wr), sn, cs;
            sincos(-wi, &sn, &cs);
            Es[s] = make_float2((float)(a * cs), (float)(a * sn));
            D.E[(long)s * D.M + i] = Es[s];
        }
        for (int c = 0; c < D.nchain; ++c) D.chain_p0[c][i] = pick(Es, D.chain_species0[c]);
    }
    // Hamiltonian partial sums (fields after the update), FP64, fixed order
    __shared__ double red[(kFtsThreads / 64) * NP * 6];
#pragma unroll
    for (int k = 0; k < NP * 6; ++k) {
        if (k >= D.npots * 6) break;
        double v = wave_sum(hs[k]);
        if ((threadIdx.x & 63) == 0) red[(threadIdx.x >> 6) * NP * 6 + k] = v;
    }
    __syncthreads();
    const int nh = D.npots * 6;
    if ((int)threadIdx.x < nh) {
        double s = 0.0;
        for (int w = 0; w < kFtsThreads / 64; ++w) s += red[w * NP * 6 + threadIdx.x];
        D.Hpart[(long)blockIdx.x * kFtsMaxPots * 6 + threadIdx.x] = s;
    }
}

// One propagation slice along one axis for a list of chains.
struct ZmDesc {
    float2* w;
    const double* Hpart;   // the update pass's block sums (sum w at [p*6+2], [p*6+3])
    int nHpart, p, M;
};

struct PropDesc {
    const float2* src;
    float2* dst;
    const float2* pre;   // multiply before, or null
    const float2* post;  // multiply after (E_b(s) on the last axis), or null
    const float2* g;     // exp(-k_a^2/(Nr-1)) / N_a along this axis as hi + lo
};

// Line plans of the single-axis pass.  The default 24- and 96-point plans
// hold twelve elements per thread, which with the line's exp(-W) factors and
// addresses takes ~200 VGPRs (two waves per SIMD); six elements per thread
// (radix-2 and radix-3 passes) take 120-141 (24-point z pass 13.4 -> 9.5 us
// in the sweep, 96^3 iteration 2.33 -> 2.24 ms).  At 48 points the five-pass
// six-element plan measured slower (27.8 -> 30.1 us) and the default stays.
template <int N>
struct FtsAxisPlan : FftPlan<N> {};
#define MFT_AXIS_PLAN(N_, T_, ...)                                              \
    template <>                                                                 \
    struct FtsAxisPlan<N_> {                                                    \
        static constexpr int N = N_;                                            \
        static constexpr int T = T_;                                            \
        static constexpr int E = N_ / T_;                                       \
        using R = Radices<__VA_ARGS__>;                                         \
    };
MFT_AXIS_PLAN(24, 4, 2, 2, 2, 3)
MFT_AXIS_PLAN(96, 16, 2, 2, 2, 2, 2, 3)
#undef MFT_AXIS_PLAN

template <int N>
__global__ void __launch_bounds__(256) k_fts_prop_axis(const PropDesc* __restrict__ descs, int nx, int ny, int nz,
                                                       int axis, const float2* __restrict__ twg)
{
    using P = FtsAxisPlan<N>;
    constexpr int T = P::T, E = P::E, B = 256 / T;
    __shared__ float2 lds[B * N];
    __shared__ float2 tw[N];
    __shared__ float2 gk[N];
    const PropDesc& D = descs[blockIdx.y];
    for (int i = threadIdx.x; i < N; i += 256) {
        tw[i] = twg[i];
        gk[i] = D.g[i];
    }
    AxisGeom geom{nx, ny, nz, axis};
    const bool inter = axis != 0;
    const int lline = inter ? (threadIdx.x % B) : (threadIdx.x / T);
    const int t = inter ? (threadIdx.x / B) : (threadIdx.x % T);
    const int line = blockIdx.x * B + lline;
    const bool valid = line < geom.nlines();
    int ta, tb;
    const long base = valid ? geom.base(line, &ta, &tb) : 0;
    const long st = geom.stride();
    float2 v[E], pv[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        long o = base + (long)(t + T * s) * st;
        float2 a = valid ? D.src[o] : make_float2(0.f, 0.f);
        if (D.pre && valid) a = cmul(a, D.pre[o]);
        v[s] = a;
    }
    // the post factors (E = exp(-W) on the last axis) are loaded with the
    // line, so their latency overlaps the transforms instead of following them
    if (D.post) {
#pragma unroll
        for (int s = 0; s < E; ++s)
            pv[s] = valid ? D.post[base + (long)(t + T * s) * st] : make_float2(1.f, 0.f);
    }
    __syncthreads();
    LdsMap map{lline, B, inter, N};
    line_fft_plan<P, -1>(v, lds, tw, t, map);
#pragma unroll
    for (int s = 0; s < E; ++s) {
        // two-float factor: the per-axis 1/N of a non-power-of-two length is not
        // a float; applied with one rounding (scale_hilo)
        const float2 g = gk[t + T * s];
        v[s] = scale_hilo(v[s], g);
    }
    line_fft_plan<P, +1>(v, lds, tw, t, map);
    if (!valid) return;
#pragma unroll
    for (int s = 0; s < E; ++s) {
        long o = base + (long)(t + T * s) * st;
        float2 a = v[s];
        if (D.post) a = cmul(a, pv[s]);
        D.dst[o] = a;
    }
}

// x and y diffusion of one z-plane in one workgroup: the plane (times pre,
// when given) is staged in LDS, every row is transformed along x, multiplied
// by g_x and transformed back in place, then every column along y; the x and y
// passes of k_fts_prop_axis with the plane kept in LDS between them (the same
// arithmetic, one global read and write instead of two).  Rows are padded by
// PAD values so the rows a half-wave touches start on different LDS banks;
// lines past the plane use a scratch region for their exchanges.
template <int N, int NT>
__global__ void __launch_bounds__(NT) k_fts_prop_xy(const PropDesc* __restrict__ dxs,
                                                    const PropDesc* __restrict__ dys,
                                                    const float2* __restrict__ twg)
{
    using P = FftPlan<N>;
    constexpr int T = P::T, E = P::E, B = NT / T;
    constexpr int NL = (N + B - 1) / B;
    // row stride in LDS: room for the skewed row exchanges (swz_pitch) when N
    // is a multiple of 16
    constexpr int PAD = (N % 16 == 0) ? (N / 8 > 8 ? N / 8 : 8) : 1, RS = N + PAD;
    constexpr int NV = N * N / 2;                            // float4 (two values) per plane
    constexpr int PER = (NV + NT - 1) / NT;
    __shared__ float2 plane[N * RS];
    __shared__ float2 scratch[(N % B) ? B * N : 1];
    __shared__ float2 tw[N], gx[N], gy[N];
    const PropDesc& DX = dxs[blockIdx.y];
    const PropDesc& DY = dys[blockIdx.y];
    const long off = (long)blockIdx.x * N * N;
    {
        // every load of the plane in flight before the first LDS store
        const float4* src = reinterpret_cast<const float4*>(DX.src + off);
        const float4* pre = DX.pre ? reinterpret_cast<const float4*>(DX.pre + off) : nullptr;
        float4 a[PER], e[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int v = threadIdx.x + NT * k;
            const int vc = v < NV ? v : NV - 1;
            a[k] = src[vc];
            e[k] = pre ? pre[vc] : make_float4(1.f, 0.f, 1.f, 0.f);
        }
        for (int i = threadIdx.x; i < N; i += NT) {
            tw[i] = twg[i];
            gx[i] = DX.g[i];
            gy[i] = DY.g[i];
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int v = threadIdx.x + NT * k;
            if (v >= NV) continue;
            float2 lo = make_float2(a[k].x, a[k].y), hi = make_float2(a[k].z, a[k].w);
            if (pre) {
                lo = cmul(lo, make_float2(e[k].x, e[k].y));
                hi = cmul(hi, make_float2(e[k].z, e[k].w));
            }
            const int i0 = 2 * v, r = i0 / N, c = i0 - r * N; // N even: both values on one row
            plane[r * RS + c] = lo;
            plane[r * RS + c + 1] = hi;
        }
    }
    __syncthreads();
    // rows (x): T consecutive threads of one wavefront per row, each row's
    // exchanges in its own LDS row: no workgroup barrier until the columns
    {
        const int lx = threadIdx.x / T, t = threadIdx.x % T;
        for (int round = 0; round < NL; ++round) {
            const int row = round * B + lx;
            const bool valid = row < N;
            float2 v[E];
#pragma unroll
            for (int s = 0; s < E; ++s) v[s] = valid ? plane[row * RS + t + T * s] : make_float2(0.f, 0.f);
            float2* base = valid ? plane : scratch;
            // skewed exchanges within the padded row (RS >= swz_pitch(N) when
            // PAD is 8, N <= 64)
            LdsMap map{valid ? row : lx, N, false, valid ? RS : N, valid && RS >= swz_pitch(N)};
            line_fft<N, -1, true>(v, base, tw, t, map);
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const float2 g = gx[t + T * s];
                v[s] = scale_hilo(v[s], g);
            }
            line_fft<N, +1, true>(v, base, tw, t, map);
            exchange_barrier<true>();
            if (valid) {
#pragma unroll
                for (int s = 0; s < E; ++s) plane[row * RS + t + T * s] = v[s];
            }
            exchange_barrier<true>();
        }
    }
    __syncthreads();
    // columns (y): B consecutive threads cover B columns, element stride RS
    {
        const int lc = threadIdx.x % B, t = threadIdx.x / B;
        for (int round = 0; round < NL; ++round) {
            const int col = round * B + lc;
            const bool valid = col < N;
            float2 v[E];
#pragma unroll
            for (int s = 0; s < E; ++s) v[s] = valid ? plane[(t + T * s) * RS + col] : make_float2(0.f, 0.f);
            float2* base = valid ? plane : scratch;
            LdsMap map{valid ? col : lc, valid ? RS : B, true, N};
            line_fft<N, -1>(v, base, tw, t, map);
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const float2 g = gy[t + T * s];
                v[s] = scale_hilo(v[s], g);
            }
            line_fft<N, +1>(v, base, tw, t, map);
            __syncthreads();
            if (valid) {
#pragma unroll
                for (int s = 0; s < E; ++s) plane[(t + T * s) * RS + col] = v[s];
            }
            __syncthreads();
        }
    }
    float4* dst = reinterpret_cast<float4*>(DY.dst + off);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int v = threadIdx.x + NT * k;
        if (v >= NV) continue;
        const int i0 = 2 * v, r = i0 / N, c = i0 - r * N;
        const float2 lo = plane[r * RS + c], hi = plane[r * RS + c + 1];
        dst[v] = make_float4(lo.x, lo.y, hi.x, hi.y);
    }
}

struct DensDesc {
    const float2* p;       // forward chain [Ntot][M]
    const float2* pd;      // complementary chain (== p when symmetric)
    const float2* W;       // species fields [nspecies][M]
    float2* rhot;          // [nblocks][M]
    double* Qpart;         // [gridDim.x]
    int M, Ntot, nblocks;
    int N[kFtsMaxBlocks];
    int species[kFtsMaxBlocks];
};

// Points per thread of k_fts_density: two (float4 loads of the propagators)
// when the grid has an even number of points, so every pair is aligned.
__host__ __device__ constexpr int fts_density_points(long M) { return (M & 1) == 0 ? 2 : 1; }

template <int PTS>
__device__ __forceinline__ void load_pts(const float2* __restrict__ p, float2 (&v)[PTS])
{
    if constexpr (PTS == 2) {
        const float4 q = *reinterpret_cast<const float4*>(p);
        v[0] = make_float2(q.x, q.y);
        v[1] = make_float2(q.z, q.w);
    } else {
        v[0] = *p;
    }
}

// rho_b(r) = e^{+W_b(r)} sum_{s in b} q_s(r) qdag_{Ntot-1-s}(r)
// (fts_molecule_linear.cu:298-327) and the partial sums of Q = sum q_{Ntot-1}
// (fts_molecule_linear.cu:265-266).
// PTS consecutive points per thread; the beads of a block are taken four at
// a time with all eight propagator loads issued before the sums, which run in
// bead order for each point as before.
template <int PTS>
__global__ void __launch_bounds__(kFtsThreads) k_fts_density(const DensDesc* __restrict__ descs)
{
    const int bx = blockIdx.x;
    const DensDesc& D = descs[blockIdx.y];
    const int i = (bx * blockDim.x + threadIdx.x) * PTS;
    double qr = 0.0, qi = 0.0;
    if (i < D.M) {
        const long M = D.M;
        int s = 0;
        for (int b = 0; b < D.nblocks; ++b) {
            double ar[PTS], ai[PTS];
#pragma unroll
            for (int j = 0; j < PTS; ++j) ar[j] = ai[j] = 0.0;
            const int nb = D.N[b];
            int k = 0;
            for (; k + 4 <= nb; k += 4, s += 4) {
                float2 a[4][PTS], c[4][PTS];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    load_pts<PTS>(D.p + (long)(s + u) * M + i, a[u]);
                    load_pts<PTS>(D.pd + (long)(D.Ntot - 1 - s - u) * M + i, c[u]);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int j = 0; j < PTS; ++j) {
                        ar[j] += (double)a[u][j].x * c[u][j].x - (double)a[u][j].y * c[u][j].y;
                        ai[j] += (double)a[u][j].x * c[u][j].y + (double)a[u][j].y * c[u][j].x;
                    }
            }
            for (; k < nb; ++k, ++s) {
                float2 a[PTS], c[PTS];
                load_pts<PTS>(D.p + (long)s * M + i, a);
                load_pts<PTS>(D.pd + (long)(D.Ntot - 1 - s) * M + i, c);
#pragma unroll
                for (int j = 0; j < PTS; ++j) {
                    ar[j] += (double)a[j].x * c[j].x - (double)a[j].y * c[j].y;
                    ai[j] += (double)a[j].x * c[j].y + (double)a[j].y * c[j].x;
                }
            }
            float2 w[PTS];
            load_pts<PTS>(D.W + (long)D.species[b] * M + i, w);
#pragma unroll
            for (int j = 0; j < PTS; ++j) {
                double sn, cs;
                sincos((double)w[j].y, &sn, &cs);
                const double e = exp((double)w[j].x), er = e * cs, ei = e * sn;
                D.rhot[(long)b * M + i + j] =
                    make_float2((float)(ar[j] * er - ai[j] * ei), (float)(ar[j] * ei + ai[j] * er));
            }
        }
        float2 pl[PTS];
        load_pts<PTS>(D.p + (long)(D.Ntot - 1) * M + i, pl);
#pragma unroll
        for (int j = 0; j < PTS; ++j) {
            qr += pl[j].x;
            qi += pl[j].y;
        }
    }
    __shared__ double red[2 * (kFtsThreads / 64)];
    qr = wave_sum(qr);
    qi = wave_sum(qi);
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = qr;
        red[2 * (threadIdx.x >> 6) + 1] = qi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kFtsThreads / 64; ++w) {
            a += red[2 * w];
            b += red[2 * w + 1];
        }
        D.Qpart[2 * bx] = a;
        D.Qpart[2 * bx + 1] = b;
    }
}

struct QDesc {
    const double* Qpart;
    int nparts;
    double scale;    // gvol / V
    double2* Q;
};

__global__ void k_fts_qreduce(const QDesc* __restrict__ descs)
{
    const QDesc& D = descs[blockIdx.x];
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < D.nparts; k += blockDim.x) {
        a += D.Qpart[2 * k];
        b += D.Qpart[2 * k + 1];
    }
    __shared__ double red[2 * 16];
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = a;
        red[2 * (threadIdx.x >> 6) + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) {
            x += red[2 * w];
            y += red[2 * w + 1];
        }
        *D.Q = make_double2(x * D.scale, y * D.scale);
    }
}

__global__ void k_fts_iter_incr(int64_t* it, int n)
{
    if (threadIdx.x < n) it[threadIdx.x] += 1;
}

// Species densities rho_s = sum_b factor_m rho~_b (output only).
__global__ void k_fts_species_rho(const FtsUpdDev* __restrict__ descs, float2* __restrict__ out)
{
    const FtsUpdDev& D = descs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.M) return;
    for (int s = 0; s < D.nspecies; ++s) {
        double rx = 0.0, ry = 0.0;
        for (int b = 0; b < D.nblk; ++b) {
            if (D.blk_species[b] != s) continue;
            int m = D.blk_mol[b];
            double2 Q = D.Q[m];
            double den = (Q.x * Q.x + Q.y * Q.y) * D.V;
            double fr = D.nmolecs[m] * Q.x / den, fi = -D.nmolecs[m] * Q.y / den;
            if (D.activity[m] > 0.0) {
                fr = D.activity[m];
                fi = 0.0;
            }
            float2 r = D.rhot[(long)b * D.M + i];
            rx += fr * r.x - fi * r.y;
            ry += fr * r.y + fi * r.x;
        }
        out[(long)s * D.M + i] = make_float2((float)rx, (float)ry);
    }
}

// ----------------------------------------------------------- host side --
struct FtsRep {
    mft_fts_desc d;
    int id = 0;
    int M = 0;
    double gvol = 0, V = 0;
    int Ntot[kFtsMaxMol] = {0};
    bool sym[kFtsMaxMol] = {false};
    double nmolecs[kFtsMaxMol] = {0};
    // activity the device descriptors use; a change takes effect after the
    // next density pass (pending until then)
    double act_dev[kFtsMaxMol] = {0};
    bool act_pending = false;
    int nblk = 0;
    float2 *wpl[kFtsMaxPots] = {nullptr}, *wmi[kFtsMaxPots] = {nullptr};
    // EMPC predictor data per potential (fields before the predictor, forces)
    float2 *wpl_o[kFtsMaxPots] = {nullptr}, *wmi_o[kFtsMaxPots] = {nullptr};
    double2 *fpl_o[kFtsMaxPots] = {nullptr}, *fmi_o[kFtsMaxPots] = {nullptr};
    bool pc = false;              // some potential uses the predictor-corrector
    // 1S: forces and the real linear coefficients A(k) of each field
    float2 *f1pl[kFtsMaxPots] = {nullptr}, *f1mi[kFtsMaxPots] = {nullptr};
    double *Akpl[kFtsMaxPots] = {nullptr}, *Akmi[kFtsMaxPots] = {nullptr};
    bool s1 = false;              // some potential uses the one-step k-space scheme
    bool zm = false;              // some potential keeps a zero-mean w+ (modify zeroMean)
    float2 *W = nullptr, *E = nullptr, *rhot = nullptr;
    float2 *p[kFtsMaxMol] = {nullptr}, *pd[kFtsMaxMol] = {nullptr};
    float2* tmp = nullptr; // [2*nmol][M]
    float2* g[3] = {nullptr, nullptr, nullptr};
    double *Hpart = nullptr, *Qpart = nullptr;
    double2* Q = nullptr;
    int nHpart = 0, nQpart = 0;
    int slot = 0;                 // iteration counter slot
    // device-resident copies of the potential fields (all w+, then every
    // Flory w-) and the activities, kept for a caller that returns to an
    // earlier state of a sweep
    float2* snap[kFtsMaxSnapshots] = {nullptr};
    double snap_act[kFtsMaxSnapshots][kFtsMaxMol] = {{0}};
    bool alive = true;
    std::vector<void*> allocs;    // device buffers owned by the replica
};

struct FtsGroup {
    int n[3];
    bool pc = false;                 // predictor-corrector replicas (two passes per iteration)
    bool s1 = false;                 // replicas with a one-step k-space update
    bool zm = false;                 // replicas with zero-mean fields
    ZmDesc* d_zm = nullptr;          // one per zero-mean field of the group, the EMPC ones first
    int nzm = 0, nzm_pc = 0;
    std::vector<int> reps;
    FtsUpdDev* d_upd = nullptr;
    DensDesc* d_dens = nullptr;
    int ndens = 0;
    QDesc* d_q = nullptr;
    // propagation: the group's replicas in chunks (consecutive replicas whose
    // per-bead working set fits chunk_bytes; one chunk when it is 0), each
    // running all its beads before the next chunk starts, so a chunk's
    // slices are read back while they are still cached; per chunk, step s
    // (1..maxN-1) and axis a descriptor list
    std::vector<std::vector<PropDesc*>> d_prop;   // [chunk][step*3 + axis]
    std::vector<std::vector<int>> nprop;          // [chunk][step] chains
    size_t chunk_bytes = 0;
    int maxN = 0;
    bool dirty = false;              // membership changed: rebuild descriptors
    std::vector<void*> allocs;       // descriptor buffers
};

// Replicas live in a deque (stable references); replicas of one grid size
// that are initialised form a group whose launches advance all of them.
// Replicas can join (after their own first pass) and leave between
// iterations; a group's descriptors are rebuilt when its membership changes.
struct FtsBatch {
    std::deque<FtsRep> reps;
    std::vector<FtsGroup> groups;
    hipStream_t stream = nullptr;
    int64_t* d_iter = nullptr;       // iteration counter per slot
    int nslots = 0;
    std::vector<int> free_slots;
    int64_t iter = 0;
    bool inited = false;
    bool plane_xy = true;            // x and y passes as one plane kernel where available
    // one CL iteration captured as a HIP graph and replayed while the
    // launches stay the same (gen counts changes of the groups' launches)
    bool use_graph = true;
    uint64_t gen = 0, graph_gen = ~0ull;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<void*> allocs;
};

template <typename T>
static int dev_alloc(std::vector<void*>& owner, T** p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        set_error(MFT_ERR_OOM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
        return MFT_ERR_OOM;
    }
    owner.push_back((void*)*p);
    return MFT_OK;
}

template <typename T>
static int falloc(FtsBatch* b, T** p, size_t n)
{
    return dev_alloc(b->allocs, p, n);
}

static void free_all(std::vector<void*>& v)
{
    for (void* p : v) (void)hipFree(p);
    v.clear();
}

static int validate_desc(const mft_fts_desc& d)
{
    MFT_REQUIRE(d.nx >= 1 && d.ny >= 1 && d.nz >= 1, MFT_ERR_INVALID, "grid sizes must be positive");
    for (int n : {d.nx, d.ny, d.nz})
        MFT_REQUIRE(mft_fft_supported(n), MFT_ERR_UNSUPPORTED, "grid size %d is not a supported FFT length", n);
    MFT_REQUIRE(d.lx > 0 && d.ly > 0 && d.lz > 0, MFT_ERR_INVALID, "box lengths must be positive");
    MFT_REQUIRE(d.Nr > 1.0, MFT_ERR_INVALID, "Nr must be given and > 1 (the reference has no default)");
    MFT_REQUIRE(d.C > 0.0, MFT_ERR_INVALID, "C must be positive");
    MFT_REQUIRE(d.nspecies >= 1 && d.nspecies <= kFtsMaxSpecies, MFT_ERR_INVALID, "nspecies must be 1..%d",
                kFtsMaxSpecies);
    MFT_REQUIRE(d.npotentials >= 0 && d.npotentials <= kFtsMaxPots, MFT_ERR_INVALID, "too many potentials");
    MFT_REQUIRE(d.nmolecules >= 1 && d.nmolecules <= kFtsMaxMol, MFT_ERR_INVALID, "nmolecules must be 1..%d",
                kFtsMaxMol);
    int noisy = 0;
    for (int p = 0; p < d.npotentials; ++p) {
        const mft_fts_potential& P = d.potentials[p];
        MFT_REQUIRE(P.kind >= 0 && P.kind <= MFT_FTS_EDWARDS, MFT_ERR_UNSUPPORTED, "FTS potential kind %d",
                    P.kind);
        MFT_REQUIRE(P.scheme == MFT_FTS_EM || P.scheme == MFT_FTS_EMPC || P.scheme == MFT_FTS_1S,
                    MFT_ERR_UNSUPPORTED, "update scheme %d", P.scheme);
        if (P.kind == MFT_FTS_FLORY) {
            MFT_REQUIRE(P.I >= 0 && P.I < d.nspecies && P.J >= 0 && P.J < d.nspecies && P.I != P.J, MFT_ERR_INVALID,
                        "Flory species out of range");
            MFT_REQUIRE(P.chiN != 0.0, MFT_ERR_INVALID, "Flory chiN must be nonzero");
            noisy += 2;
        } else if (P.kind == MFT_FTS_HELFAND || P.kind == MFT_FTS_EDWARDS) {
            MFT_REQUIRE(P.kappaN != 0.0, MFT_ERR_INVALID, "Helfand kappaN / Edwards B must be nonzero");
            noisy += 1;
        }
    }
    MFT_REQUIRE(noisy <= 8, MFT_ERR_UNSUPPORTED, "at most 8 noisy fields");
    for (int m = 0; m < d.nmolecules; ++m) {
        const mft_fts_molecule& mol = d.molecules[m];
        MFT_REQUIRE(mol.nblocks >= 1 && mol.nblocks <= kFtsMaxBlocks, MFT_ERR_INVALID, "blocks must be 1..%d",
                    kFtsMaxBlocks);
        for (int b = 0; b < mol.nblocks; ++b) {
            MFT_REQUIRE(mol.N[b] >= 1, MFT_ERR_INVALID, "block length must be positive");
            MFT_REQUIRE(mol.species[b] >= 0 && mol.species[b] < d.nspecies, MFT_ERR_INVALID,
                        "species not found!");
        }
        MFT_REQUIRE(mol.phi > 0.0 || mol.activity > 0.0, MFT_ERR_INVALID,
                    "molecule needs a positive phi or activity");
    }
    return MFT_OK;
}

// exp(-k^2/(Nr-1))/n per axis as an unevaluated float sum hi + lo; k from
// the reference's float wavevector (Box.cu:182-203)
static void axis_table(int n, float L, double Nr, std::vector<float2>& g)
{
    g.resize(n);
    for (int i = 0; i < n; ++i) {
        double k = (double)kcomp(i, n, L);
        double v = std::exp(-k * k / (Nr - 1.0)) / (double)n;
        float hi = (float)v;
        g[i] = make_float2(hi, (float)(v - (double)hi));
    }
}

// |k|^2 of grid index i (Box::get_kD double version, Box.cu:181-203: float
// components, squared and summed in double)
static double k2_of(int i, const mft_fts_desc& d)
{
    const int ix = i % d.nx, iy = (i / d.nx) % d.ny, iz = i / (d.nx * d.ny);
    const double kx = kcomp(ix, d.nx, d.lx), ky = kcomp(iy, d.ny, d.ly), kz = kcomp(iz, d.nz, d.lz);
    return kx * kx + ky * ky + kz * kz;
}

// Debye functions of FTS_Box.cu:537-587 at x = k^2 alpha (f): the whole
// chain (1 at k = 0), one block (0 at k = 0), and two blocks separated by a
// middle block with the factor 2 (0 at k = 0)
static double debye_chain(double k2, double a) { return 2.0 * (std::exp(-k2 * a) + k2 * a - 1.0) / (k2 * k2); }
static double debye_block(double k2, double f, double a)
{
    return 2.0 * (std::exp(-k2 * a * f) + k2 * a * f - 1.0) / (k2 * k2);
}
static double debye_cross(double k2, double fA, double fB, double fC, double a)
{
    return 2.0 * (1.0 - std::exp(-k2 * a * fA)) * std::exp(-k2 * a * fB) * (1.0 - std::exp(-k2 * a * fC)) / (k2 * k2);
}

// Linear coefficients of the 1S fields: initLinearCoeffs
// (fts_potential_helfand.cu:277-280, fts_potential_flory.cu:323-326,
// fts_potential_incompress.cu:249-252) plus every molecule's
// computeLinearTerms (fts_molecule_linear.cu:433-560); the Debye terms go to
// A+ of Helfand and Flory potentials, A- of Flory stays 2C/chiN.
static int linear_coeffs(FtsRep& r)
{
    const mft_fts_desc& d = r.d;
    const size_t M = r.M;
    std::vector<double> k2(M);
    for (size_t i = 0; i < M; ++i) k2[i] = k2_of((int)i, d);
    std::vector<std::vector<double>> Apl(d.npotentials), Ami(d.npotentials);
    for (int p = 0; p < d.npotentials; ++p) {
        const mft_fts_potential& P = d.potentials[p];
        // fts_potential_edwards.cu initLinearCoeffs: 1/B
        const double a0 = P.kind == MFT_FTS_HELFAND
                              ? d.C / P.kappaN
                              : (P.kind == MFT_FTS_FLORY ? 2.0 * d.C / P.chiN
                                                         : (P.kind == MFT_FTS_EDWARDS ? 1.0 / P.kappaN : 0.0));
        Apl[p].assign(M, a0);
        if (P.kind == MFT_FTS_FLORY) Ami[p].assign(M, a0);
    }
    for (int m = 0; m < d.nmolecules; ++m) {
        const mft_fts_molecule& mol = d.molecules[m];
        const int Nt = r.Ntot[m];
        const double alpha = (double)Nt / d.Nr;
        const double pref = r.nmolecs[m] * alpha * (double)Nt / r.V;
        for (int p = 0; p < d.npotentials; ++p) {
            const mft_fts_potential& P = d.potentials[p];
            const bool homo_flory = P.kind == MFT_FTS_FLORY && mol.nblocks == 1;
            // fts_molecule_linear.cu:459: Helfand and Edwards take the chain Debye term
            if (P.kind == MFT_FTS_HELFAND || P.kind == MFT_FTS_EDWARDS || homo_flory) {
                Apl[p][0] += pref;
                for (size_t i = 1; i < M; ++i) Apl[p][i] += pref * debye_chain(k2[i], alpha);
                continue;
            }
            if (P.kind != MFT_FTS_FLORY) continue;
            auto acts = [&](int b) { return mol.species[b] == P.I || mol.species[b] == P.J; };
            for (int j = 0; j < mol.nblocks - 1; ++j) {
                if (!acts(j)) continue;
                const double fA = (double)mol.N[j] / Nt;
                std::vector<double> gMT(M, 0.0);
                for (size_t i = 1; i < M; ++i) gMT[i] = debye_block(k2[i], fA, alpha);
                int nbetween = 0;
                for (int k = j + 1; k < mol.nblocks; ++k) {
                    if (!acts(k)) continue;
                    const double fC = (double)mol.N[k] / Nt, fB = (double)nbetween / Nt;
                    for (size_t i = 1; i < M; ++i) {
                        gMT[i] += debye_block(k2[i], fC, alpha) + debye_cross(k2[i], fA, fB, fC, alpha);
                        Apl[p][i] += pref * gMT[i];
                    }
                    nbetween += mol.N[k];
                }
            }
        }
    }
    for (int p = 0; p < d.npotentials; ++p) {
        if (d.potentials[p].scheme != MFT_FTS_1S) continue;
        MFT_HIP(hipMemcpy(r.Akpl[p], Apl[p].data(), sizeof(double) * M, hipMemcpyHostToDevice));
        if (r.Akmi[p]) MFT_HIP(hipMemcpy(r.Akmi[p], Ami[p].data(), sizeof(double) * M, hipMemcpyHostToDevice));
    }
    return MFT_OK;
}

static int rep_setup(FtsBatch* b, FtsRep& r)
{
    const mft_fts_desc& d = r.d;
    int rc;
    r.M = d.nx * d.ny * d.nz;
    // FTS_Box.cu:466-476 (dx in double from the float box lengths)
    r.V = (double)d.lx * (double)d.ly * (double)d.lz;
    r.gvol = ((double)d.lx / d.nx) * ((double)d.ly / d.ny) * ((double)d.lz / d.nz);
    const size_t M = r.M;
    for (int p = 0; p < d.npotentials; ++p) {
        if ((rc = dev_alloc(r.allocs, &r.wpl[p], M))) return rc;
        MFT_HIP(hipMemset(r.wpl[p], 0, sizeof(float2) * M));
        if (d.potentials[p].kind == MFT_FTS_FLORY) {
            if ((rc = dev_alloc(r.allocs, &r.wmi[p], M))) return rc;
            MFT_HIP(hipMemset(r.wmi[p], 0, sizeof(float2) * M));
        }
        if (d.potentials[p].zero_mean && d.potentials[p].kind != MFT_FTS_FLORY) r.zm = true;
        if (d.potentials[p].scheme == MFT_FTS_1S) {
            r.s1 = true;
            if ((rc = dev_alloc(r.allocs, &r.f1pl[p], M))) return rc;
            if ((rc = dev_alloc(r.allocs, &r.Akpl[p], M))) return rc;
            if (d.potentials[p].kind == MFT_FTS_FLORY) {
                if ((rc = dev_alloc(r.allocs, &r.f1mi[p], M))) return rc;
                if ((rc = dev_alloc(r.allocs, &r.Akmi[p], M))) return rc;
            }
        }
        if (d.potentials[p].scheme == MFT_FTS_EMPC) {
            r.pc = true;
            if ((rc = dev_alloc(r.allocs, &r.wpl_o[p], M))) return rc;
            if ((rc = dev_alloc(r.allocs, &r.fpl_o[p], M))) return rc;
            if (d.potentials[p].kind == MFT_FTS_FLORY) {
                if ((rc = dev_alloc(r.allocs, &r.wmi_o[p], M))) return rc;
                if ((rc = dev_alloc(r.allocs, &r.fmi_o[p], M))) return rc;
            }
        }
    }
    if ((rc = dev_alloc(r.allocs, &r.W, M * d.nspecies))) return rc;
    if ((rc = dev_alloc(r.allocs, &r.E, M * d.nspecies))) return rc;
    r.nblk = 0;
    for (int m = 0; m < d.nmolecules; ++m) {
        const mft_fts_molecule& mol = d.molecules[m];
        int Nt = 0;
        for (int k = 0; k < mol.nblocks; ++k) Nt += mol.N[k];
        r.Ntot[m] = Nt;
        // fts_molecule_linear.cu:81-88
        bool sym = true;
        for (int j = 0; j < mol.nblocks / 2; ++j)
            if (mol.species[j] != mol.species[mol.nblocks - 1 - j] || mol.N[j] != mol.N[mol.nblocks - 1 - j])
                sym = false;
        r.sym[m] = sym;
        // fts_molecule_linear.cu:433-435 with Vfree = V
        // in the grand canonical case phi stays -1 (fts_molecule.cu:19), so
        // this count is negative and only the 1S Debye prefactor uses it
        // (fts_molecule_linear.cu:563-566)
        r.nmolecs[m] = d.C * (mol.activity > 0.0 ? -1.0 : mol.phi) * r.V * d.Nr / (double)Nt;
        r.act_dev[m] = mol.activity > 0.0 ? mol.activity : 0.0;
        if ((rc = dev_alloc(r.allocs, &r.p[m], M * Nt))) return rc;
        if (sym) r.pd[m] = r.p[m];
        else if ((rc = dev_alloc(r.allocs, &r.pd[m], M * Nt))) return rc;
        r.nblk += mol.nblocks;
    }
    MFT_REQUIRE(r.nblk <= kFtsMaxMol * kFtsMaxBlocks, MFT_ERR_INVALID, "too many blocks");
    if (r.s1 && (rc = linear_coeffs(r))) return rc;
    if ((rc = dev_alloc(r.allocs, &r.rhot, M * r.nblk))) return rc;
    MFT_HIP(hipMemset(r.rhot, 0, sizeof(float2) * M * r.nblk));
    if ((rc = dev_alloc(r.allocs, &r.tmp, M * 2 * d.nmolecules))) return rc;
    const int ns[3] = {d.nx, d.ny, d.nz};
    const float Ls[3] = {d.lx, d.ly, d.lz};
    for (int a = 0; a < 3; ++a) {
        std::vector<float2> g;
        axis_table(ns[a], Ls[a], d.Nr, g);
        if ((rc = dev_alloc(r.allocs, &r.g[a], g.size()))) return rc;
        MFT_HIP(hipMemcpy(r.g[a], g.data(), sizeof(float2) * g.size(), hipMemcpyHostToDevice));
    }
    r.nHpart = (r.M + kFtsThreads - 1) / kFtsThreads;
    r.nQpart = r.nHpart;
    if ((rc = dev_alloc(r.allocs, &r.Hpart, (size_t)r.nHpart * kFtsMaxPots * 6))) return rc;
    if ((rc = dev_alloc(r.allocs, &r.Qpart, (size_t)r.nQpart * 2 * d.nmolecules))) return rc;
    if ((rc = dev_alloc(r.allocs, &r.Q, (size_t)d.nmolecules))) return rc;
    return MFT_OK;
}

static void fill_upd(const FtsBatch* b, const FtsRep& r, FtsUpdDev& U, int first, int have_rho, int corrector = 0)
{
    const mft_fts_desc& d = r.d;
    std::memset(&U, 0, sizeof(U));
    U.corrector = corrector;
    U.M = r.M;
    U.nspecies = d.nspecies;
    U.npots = d.npotentials;
    int col = 0;
    for (int p = 0; p < d.npotentials; ++p) {
        const mft_fts_potential& P = d.potentials[p];
        FtsPotDev& D = U.pot[p];
        D.kind = P.kind;
        D.I = P.I;
        D.J = P.J;
        D.wpl = r.wpl[p];
        D.wmi = r.wmi[p];
        D.scheme = P.scheme;
        D.wpl_o = r.wpl_o[p];
        D.wmi_o = r.wmi_o[p];
        D.fpl_o = r.fpl_o[p];
        D.fmi_o = r.fmi_o[p];
        D.f1pl = r.f1pl[p];
        D.f1mi = r.f1mi[p];
        D.dt_pl = P.delt_pl;
        D.dt_mi = P.delt_mi;
        D.col_pl = D.col_mi = -1;
        if (P.kind == MFT_FTS_FLORY) {
            D.cpl = D.cmi = 2.0 * d.C / P.chiN;
            if (d.cl) {
                D.col_pl = col++;
                D.col_mi = col++;
                D.nmag_pl = std::sqrt(2.0 * P.delt_pl / r.gvol);
                D.nmag_mi = std::sqrt(2.0 * P.delt_mi / r.gvol);
            }
        } else if (P.kind == MFT_FTS_HELFAND || P.kind == MFT_FTS_EDWARDS) {
            // the device runs Edwards as Helfand with 1/B and no -iC term
            // (d_makeEdwardsForce, fts_potential_edwards.cu:227-244)
            D.kind = MFT_FTS_HELFAND;
            D.cpl = P.kind == MFT_FTS_EDWARDS ? 1.0 / P.kappaN : d.C / P.kappaN;
            D.cterm = P.kind == MFT_FTS_EDWARDS ? 0.0 : d.C;
            if (d.cl) {
                D.col_pl = col++;
                D.nmag_pl = std::sqrt(2.0 * P.delt_pl / r.gvol);
            }
        }
    }
    int k = 0;
    for (int m = 0; m < d.nmolecules; ++m)
        for (int bb = 0; bb < d.molecules[m].nblocks; ++bb, ++k) {
            U.blk_species[k] = d.molecules[m].species[bb];
            U.blk_mol[k] = m;
        }
    U.nblk = k;
    U.rhot = r.rhot;
    U.Q = r.Q;
    for (int m = 0; m < d.nmolecules; ++m) {
        U.nmolecs[m] = r.nmolecs[m];
        U.activity[m] = r.act_dev[m];
    }
    U.V = r.V;
    U.C = d.C;
    U.Nr = d.Nr;
    U.W = r.W;
    U.E = r.E;
    U.nchain = 0;
    for (int m = 0; m < d.nmolecules; ++m) {
        const mft_fts_molecule& mol = d.molecules[m];
        U.chain_p0[U.nchain] = r.p[m];
        U.chain_species0[U.nchain++] = mol.species[0];
        if (!r.sym[m]) {
            U.chain_p0[U.nchain] = r.pd[m];
            U.chain_species0[U.nchain++] = mol.species[mol.nblocks - 1];
        }
    }
    U.Hpart = r.Hpart;
    U.seed = d.seed;
    U.stream = 0; // noise keyed by the replica's own seed, never by its slot
    U.iter = b->d_iter + r.slot;
    U.cl = d.cl;
    U.first = first;
    U.have_rho = have_rho;
}

// species of bead s of molecule m along the forward (c = 0) or the
// complementary (c = 1) chain
static int bead_species(const mft_fts_molecule& mol, int Ntot, int s, int c)
{
    int bead = c == 0 ? s : Ntot - 1 - s;
    int acc = 0;
    for (int b = 0; b < mol.nblocks; ++b) {
        acc += mol.N[b];
        if (bead < acc) return mol.species[b];
    }
    return mol.species[mol.nblocks - 1];
}

static int group_setup(FtsBatch* b, FtsGroup& G)
{
    int rc;
    ++b->gen;
    free_all(G.allocs);
    G.d_prop.clear();
    G.nprop.clear();
    G.dirty = false;
    const int nr = (int)G.reps.size();
    // update descriptors: [0, nr) regular (the predictor with EMPC),
    // [nr, 2nr) initial pass, [2nr, 3nr) corrector
    std::vector<FtsUpdDev> upd(3 * nr);
    for (int k = 0; k < nr; ++k) {
        fill_upd(b, b->reps[G.reps[k]], upd[k], 0, 1);
        fill_upd(b, b->reps[G.reps[k]], upd[nr + k], 1, 0);
        fill_upd(b, b->reps[G.reps[k]], upd[2 * nr + k], 0, 1, 1);
    }
    if ((rc = dev_alloc(G.allocs, &G.d_upd, upd.size()))) return rc;
    MFT_HIP(hipMemcpy(G.d_upd, upd.data(), sizeof(FtsUpdDev) * upd.size(), hipMemcpyHostToDevice));

    // density and Q descriptors: one per (replica, molecule)
    std::vector<DensDesc> dens;
    std::vector<QDesc> qd;
    for (int k = 0; k < nr; ++k) {
        FtsRep& r = b->reps[G.reps[k]];
        for (int m = 0; m < r.d.nmolecules; ++m) {
            const mft_fts_molecule& mol = r.d.molecules[m];
            DensDesc D;
            std::memset(&D, 0, sizeof(D));
            D.p = r.p[m];
            D.pd = r.pd[m];
            D.W = r.W;
            int off = 0;
            for (int mm = 0; mm < m; ++mm) off += r.d.molecules[mm].nblocks;
            D.rhot = r.rhot + (size_t)off * r.M;
            D.Qpart = r.Qpart + (size_t)m * 2 * r.nQpart;
            D.M = r.M;
            D.Ntot = r.Ntot[m];
            D.nblocks = mol.nblocks;
            for (int bb = 0; bb < mol.nblocks; ++bb) {
                D.N[bb] = mol.N[bb];
                D.species[bb] = mol.species[bb];
            }
            dens.push_back(D);
            QDesc q;
            q.Qpart = D.Qpart;
            // the density pass's blocks (its partial sums; nQpart covers one
            // point per thread)
            q.nparts = (int)((r.M / fts_density_points(r.M) + kFtsThreads - 1) / kFtsThreads);
            q.scale = r.gvol / r.V;
            q.Q = r.Q + m;
            qd.push_back(q);
        }
    }
    G.ndens = (int)dens.size();
    if ((rc = dev_alloc(G.allocs, &G.d_dens, dens.size()))) return rc;
    MFT_HIP(hipMemcpy(G.d_dens, dens.data(), sizeof(DensDesc) * dens.size(), hipMemcpyHostToDevice));
    if ((rc = dev_alloc(G.allocs, &G.d_q, qd.size()))) return rc;
    MFT_HIP(hipMemcpy(G.d_q, qd.data(), sizeof(QDesc) * qd.size(), hipMemcpyHostToDevice));

    // zero-mean fields: w+ of every Helfand / Incompress / Edwards potential
    // with modify zeroMean
    // (a correction re-centres only EMPC fields: correctFields returns early
    // for the others, fts_potential_helfand.cu:155-158)
    std::vector<ZmDesc> zm;
    for (int pass = 0; pass < 2; ++pass) {
        for (int k = 0; k < nr; ++k) {
            FtsRep& r = b->reps[G.reps[k]];
            for (int p = 0; p < r.d.npotentials; ++p) {
                const mft_fts_potential& P = r.d.potentials[p];
                if (!P.zero_mean || P.kind == MFT_FTS_FLORY || (P.scheme == MFT_FTS_EMPC) != (pass == 0)) continue;
                zm.push_back(ZmDesc{r.wpl[p], r.Hpart, r.nHpart, p, r.M});
            }
        }
        if (pass == 0) G.nzm_pc = (int)zm.size();
    }
    G.nzm = (int)zm.size();
    G.d_zm = nullptr;
    if (G.nzm) {
        if ((rc = dev_alloc(G.allocs, &G.d_zm, zm.size()))) return rc;
        MFT_HIP(hipMemcpy(G.d_zm, zm.data(), sizeof(ZmDesc) * zm.size(), hipMemcpyHostToDevice));
    }

    // propagation descriptors per chunk, step and axis
    G.maxN = 0;
    for (int k = 0; k < nr; ++k)
        for (int m = 0; m < b->reps[G.reps[k]].d.nmolecules; ++m) G.maxN = std::max(G.maxN, b->reps[G.reps[k]].Ntot[m]);
    // chunks: a replica's per-bead working set is four slices per chain
    // (q_(s-1) read, the intermediate written and read, e^-W read, q_s written)
    std::vector<int> chunk_end;
    size_t acc = 0;
    for (int k = 0; k < nr; ++k) {
        const FtsRep& r = b->reps[G.reps[k]];
        int chains = 0;
        for (int m = 0; m < r.d.nmolecules; ++m) chains += r.sym[m] ? 1 : 2;
        const size_t ws = (size_t)chains * 4 * sizeof(float2) * (size_t)r.M;
        if (G.chunk_bytes > 0 && k > 0 && acc + ws > G.chunk_bytes) {
            chunk_end.push_back(k);
            acc = 0;
        }
        acc += ws;
    }
    chunk_end.push_back(nr);
    const int nchunks = (int)chunk_end.size();
    G.d_prop.assign(nchunks, std::vector<PropDesc*>((size_t)3 * G.maxN, nullptr));
    G.nprop.assign(nchunks, std::vector<int>(G.maxN, 0));
    for (int ch = 0; ch < nchunks; ++ch)
    for (int s = 1; s < G.maxN; ++s) {
        std::vector<PropDesc> ax[3];
        for (int k = ch == 0 ? 0 : chunk_end[ch - 1]; k < chunk_end[ch]; ++k) {
            FtsRep& r = b->reps[G.reps[k]];
            int chain = 0;
            for (int m = 0; m < r.d.nmolecules; ++m) {
                const mft_fts_molecule& mol = r.d.molecules[m];
                for (int c = 0; c < (r.sym[m] ? 1 : 2); ++c, ++chain) {
                    if (s >= r.Ntot[m]) continue;
                    float2* P = c == 0 ? r.p[m] : r.pd[m];
                    float2* tmp = r.tmp + (size_t)chain * r.M;
                    // q_s = e^{-W_b(s)} IFFT[g FFT[q_(s-1)]] (fts_molecule_linear.cu:166-199)
                    const float2* Epost = r.E + (size_t)bead_species(mol, r.Ntot[m], s, c) * r.M;
                    ax[0].push_back(PropDesc{P + (size_t)(s - 1) * r.M, tmp, nullptr, nullptr, r.g[0]});
                    ax[1].push_back(PropDesc{tmp, tmp, nullptr, nullptr, r.g[1]});
                    ax[2].push_back(PropDesc{tmp, P + (size_t)s * r.M, nullptr, Epost, r.g[2]});
                }
            }
        }
        G.nprop[ch][s] = (int)ax[0].size();
        if (ax[0].empty()) continue;
        for (int a = 0; a < 3; ++a) {
            PropDesc* dp = nullptr;
            if ((rc = dev_alloc(G.allocs, &dp, ax[a].size()))) return rc;
            MFT_HIP(hipMemcpy(dp, ax[a].data(), sizeof(PropDesc) * ax[a].size(), hipMemcpyHostToDevice));
            G.d_prop[ch][(size_t)3 * s + a] = dp;
        }
    }
    return MFT_OK;
}

template <int N>
static void launch_prop_n(const PropDesc* d, int n, const FtsGroup& G, int axis, hipStream_t st)
{
    constexpr int B = 256 / FtsAxisPlan<N>::T;
    AxisGeom geom{G.n[0], G.n[1], G.n[2], axis};
    dim3 grid((geom.nlines() + B - 1) / B, n);
    k_fts_prop_axis<N><<<grid, 256, 0, st>>>(d, G.n[0], G.n[1], G.n[2], axis, fft_twiddles_device(N));
}

static int launch_prop(const PropDesc* d, int n, const FtsGroup& G, int axis, hipStream_t st)
{
    switch (G.n[axis]) {
#define MFT_CASE(n_)                                                                                      \
    case n_: launch_prop_n<n_>(d, n, G, axis, st); break;
        MFT_FOR_EACH_FFT_SIZE(MFT_CASE)
#undef MFT_CASE
    default: set_error(MFT_ERR_UNSUPPORTED, "unsupported FFT length %d", G.n[axis]); return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// whether the x and y passes of this group run as one plane kernel
static bool plane_fused(const FtsGroup& G)
{
    if (G.n[0] != G.n[1]) return false;
    switch (G.n[0]) {
    case 16: case 20: case 24: case 32: case 40: case 48: case 64: case 96: return true;
    default: return false;
    }
}

template <int N>
static void launch_prop_xy_n(const PropDesc* dx, const PropDesc* dy, int n, const FtsGroup& G, hipStream_t st)
{
    dim3 grid(G.n[2], n);
    // enough threads for every row (and column) of the plane in one round,
    // in whole wavefronts: 48-point planes take exactly 48 lines (192
    // threads), so no line needs the scratch region and more workgroups fit in
    // LDS; small planes take one or two wavefronts rather than 256 threads of
    // which most would transform scratch lines
    // (96-point planes: 96 lines of 8 threads, 768 threads, an 81 KB plane)
    constexpr int NT = N == 48 ? 192 : N == 96 ? 768 : (N * FftPlan<N>::T > 256 ? 512 : std::max(64, (N * FftPlan<N>::T + 63) / 64 * 64));
    k_fts_prop_xy<N, NT><<<grid, NT, 0, st>>>(dx, dy, fft_twiddles_device(N));
}

static int launch_prop_xy(const PropDesc* dx, const PropDesc* dy, int n, const FtsGroup& G, hipStream_t st)
{
    switch (G.n[0]) {
    case 16: launch_prop_xy_n<16>(dx, dy, n, G, st); break;
    case 20: launch_prop_xy_n<20>(dx, dy, n, G, st); break;
    case 24: launch_prop_xy_n<24>(dx, dy, n, G, st); break;
    case 32: launch_prop_xy_n<32>(dx, dy, n, G, st); break;
    case 40: launch_prop_xy_n<40>(dx, dy, n, G, st); break;
    case 48: launch_prop_xy_n<48>(dx, dy, n, G, st); break;
    case 64: launch_prop_xy_n<64>(dx, dy, n, G, st); break;
    case 96: launch_prop_xy_n<96>(dx, dy, n, G, st); break;
    default: set_error(MFT_ERR_UNSUPPORTED, "no plane kernel for %d", G.n[0]); return MFT_ERR_UNSUPPORTED;
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// propagators, densities and Q of every replica in the group
static int group_density(FtsBatch* b, FtsGroup& G)
{
    int rc;
    hipStream_t st = b->stream;
    const bool fused = plane_fused(G) && b->plane_xy;
    for (size_t ch = 0; ch < G.d_prop.size(); ++ch) {
        const std::vector<PropDesc*>& dp = G.d_prop[ch];
        for (int s = 1; s < G.maxN; ++s) {
            const int n = G.nprop[ch][s];
            if (n == 0) continue;
            if (fused) {
                if ((rc = launch_prop_xy(dp[(size_t)3 * s], dp[(size_t)3 * s + 1], n, G, st))) return rc;
            } else {
                for (int a = 0; a < 2; ++a)
                    if ((rc = launch_prop(dp[(size_t)3 * s + a], n, G, a, st))) return rc;
            }
            if ((rc = launch_prop(dp[(size_t)3 * s + 2], n, G, 2, st))) return rc;
        }
    }
    const int M = G.n[0] * G.n[1] * G.n[2];
    const int pts = fts_density_points(M);
    dim3 grid((M / pts + kFtsThreads - 1) / kFtsThreads, G.ndens);
    if (pts == 2) k_fts_density<2><<<grid, kFtsThreads, 0, st>>>(G.d_dens);
    else k_fts_density<1><<<grid, kFtsThreads, 0, st>>>(G.d_dens);
    k_fts_qreduce<<<G.ndens, 256, 0, st>>>(G.d_q);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// pass 0: the iteration's update (the predictor with EMPC), 1: the initial
// pass (no update), 2: the EMPC corrector
// w -= mean(w), the mean from the update pass's FP64 block sums, reduced in
// the same fixed order by every workgroup (thrust::reduce / M then subtract,
// fts_potential_helfand.cu:140-146).  256 threads.
__global__ void __launch_bounds__(256) k_fts_zero_mean(const ZmDesc* __restrict__ descs)
{
    const ZmDesc& D = descs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double sx = 0.0, sy = 0.0;
    for (int k = threadIdx.x; k < D.nHpart; k += 256) {
        const double* h = D.Hpart + (long)k * kFtsMaxPots * 6 + D.p * 6;
        sx += h[2];
        sy += h[3];
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    __shared__ double red[2][4];
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = sx;
        red[1][threadIdx.x >> 6] = sy;
    }
    __syncthreads();
    const double mx = (red[0][0] + red[0][1] + red[0][2] + red[0][3]) / (double)D.M;
    const double my = (red[1][0] + red[1][1] + red[1][2] + red[1][3]) / (double)D.M;
    if (i >= D.M) return;
    const float2 w = D.w[i];
    D.w[i] = make_float2((float)((double)w.x - mx), (float)((double)w.y - my));
}

static int group_update(FtsBatch* b, FtsGroup& G, int first);

// zero-mean fields (after a correction only the EMPC ones), then the species
// fields, E and the H sums again
static int group_zero_mean(FtsBatch* b, FtsGroup& G, bool corrector)
{
    const int n = corrector ? G.nzm_pc : G.nzm;
    if (!n) return MFT_OK;
    const int M = G.n[0] * G.n[1] * G.n[2];
    dim3 grid((M + 255) / 256, n);
    k_fts_zero_mean<<<grid, 256, 0, b->stream>>>(G.d_zm);
    MFT_LAUNCH_CHECK();
    return group_update(b, G, 1);
}

static int group_update(FtsBatch* b, FtsGroup& G, int first)
{
    const int M = G.n[0] * G.n[1] * G.n[2];
    const int nr = (int)G.reps.size();
    int ns = 1, np = 1;
    for (int id : G.reps) {
        ns = std::max(ns, b->reps[id].d.nspecies);
        np = std::max(np, b->reps[id].d.npotentials);
    }
    const int NS = ns <= 2 ? 2 : (ns <= 4 ? 4 : 8), NP = np <= 2 ? 2 : (np <= 4 ? 4 : 8);
    dim3 grid((M + kFtsThreads - 1) / kFtsThreads, nr);
    const FtsUpdDev* d = G.d_upd + (size_t)first * nr;
#define MFT_UPD(S_, P_)                                                                                   \
    if (NS == S_ && NP == P_) k_fts_update<S_, P_><<<grid, kFtsThreads, 0, b->stream>>>(d);
    MFT_UPD(2, 2) MFT_UPD(2, 4) MFT_UPD(2, 8) MFT_UPD(4, 2) MFT_UPD(4, 4) MFT_UPD(4, 8) MFT_UPD(8, 2)
    MFT_UPD(8, 4) MFT_UPD(8, 8)
#undef MFT_UPD
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

static int batch_destroy(FtsBatch* b);

static FtsGroup& group_for(FtsBatch* b, const FtsRep& r)
{
    const mft_fts_desc& d = r.d;
    for (auto& G : b->groups)
        if (G.n[0] == d.nx && G.n[1] == d.ny && G.n[2] == d.nz && G.pc == r.pc && G.s1 == r.s1 && G.zm == r.zm)
            return G;
    FtsGroup G;
    G.n[0] = d.nx;
    G.n[1] = d.ny;
    G.n[2] = d.nz;
    G.pc = r.pc;
    G.s1 = r.s1;
    G.zm = r.zm;
    b->groups.push_back(G);
    return b->groups.back();
}

// empty batch with nslots iteration counters
static int batch_create_empty(FtsBatch** out, int nslots)
{
    FtsBatch* b = new FtsBatch();
    // MFT_FTS_PLANE_XY=0 keeps the x and y passes as separate line kernels
    // (same arithmetic; used to check that the plane kernel matches them)
    const char* pe = std::getenv("MFT_FTS_PLANE_XY");
    b->plane_xy = !(pe && pe[0] == '0');
    // MFT_FTS_GRAPH=0 launches every iteration directly (checks the replayed
    // graph against direct launches)
    const char* ge = std::getenv("MFT_FTS_GRAPH");
    b->use_graph = !(ge && ge[0] == '0');
    int rc = falloc(b, &b->d_iter, (size_t)nslots);
    if (!rc) {
        hipError_t e = hipMemset(b->d_iter, 0, sizeof(int64_t) * nslots);
        if (e != hipSuccess) rc = MFT_ERR_DEVICE;
    }
    if (!rc) {
        hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error(MFT_ERR_DEVICE, "stream: %s", hipGetErrorString(e));
            rc = MFT_ERR_DEVICE;
        }
    }
    if (rc) {
        free_all(b->allocs);
        delete b;
        return rc;
    }
    b->nslots = nslots;
    for (int k = nslots - 1; k >= 0; --k) b->free_slots.push_back(k);
    *out = b;
    return MFT_OK;
}

// a new replica: buffers allocated, fields zero, counter at 0; not yet in a
// group (batch_activate runs its first pass and adds it)
static int batch_add(FtsBatch* b, const mft_fts_desc& d, int* rep_out)
{
    int rc = validate_desc(d);
    if (rc) return rc;
    MFT_REQUIRE(!b->free_slots.empty(), MFT_ERR_INVALID, "no free replica slot (%d in use)", b->nslots);
    for (int a = 0; a < 3; ++a)
        MFT_REQUIRE(fft_twiddles_device((&d.nx)[a]) != nullptr, MFT_ERR_OOM, "twiddle table");
    b->reps.emplace_back();
    FtsRep& r = b->reps.back();
    r.d = d;
    r.id = (int)b->reps.size() - 1;
    r.slot = b->free_slots.back();
    if ((rc = rep_setup(b, r))) {
        free_all(r.allocs);
        r.alive = false;
        return rc;
    }
    b->free_slots.pop_back();
    MFT_HIP(hipMemsetAsync(b->d_iter + r.slot, 0, sizeof(int64_t), b->stream));
    *rep_out = r.id;
    return MFT_OK;
}

// first pass (FTS_Box::finishInitialization density pass) of the given new
// replicas, grouped by grid size, then each joins its grid's group
static int batch_activate(FtsBatch* b, const std::vector<int>& reps)
{
    int rc = MFT_OK;
    std::map<std::array<int, 4>, FtsGroup> tmp;
    for (int id : reps) {
        const mft_fts_desc& d = b->reps[id].d;
        const FtsRep& r = b->reps[id];
        FtsGroup& T = tmp[std::array<int, 4>{d.nx, d.ny, d.nz, (r.pc ? 1 : 0) | (r.s1 ? 2 : 0)}];
        T.n[0] = d.nx;
        T.n[1] = d.ny;
        T.n[2] = d.nz;
        T.pc = r.pc;
        T.s1 = r.s1;
        T.reps.push_back(id);
    }
    for (auto& kv : tmp) {
        FtsGroup& T = kv.second;
        if (!rc) rc = group_setup(b, T);
        if (!rc) rc = group_update(b, T, 1);
        if (!rc) rc = group_density(b, T);
    }
    // the temporary descriptors are read by the launches just queued
    hipError_t e = hipStreamSynchronize(b->stream);
    for (auto& kv : tmp) free_all(kv.second.allocs);
    if (rc) return rc;
    if (e != hipSuccess) {
        set_error(MFT_ERR_DEVICE, "first pass: %s", hipGetErrorString(e));
        return MFT_ERR_DEVICE;
    }
    for (int id : reps) {
        FtsGroup& G = group_for(b, b->reps[id]);
        G.reps.push_back(id);
        G.dirty = true;
    }
    ++b->gen;
    return MFT_OK;
}

// a replica leaves: out of its group, buffers freed, slot released (the
// caller has synchronised the stream)
static int batch_retire(FtsBatch* b, int id)
{
    MFT_REQUIRE(id >= 0 && id < (int)b->reps.size() && b->reps[id].alive, MFT_ERR_INVALID, "no such replica");
    FtsRep& r = b->reps[id];
    for (auto& G : b->groups) {
        auto it = std::find(G.reps.begin(), G.reps.end(), id);
        if (it != G.reps.end()) {
            G.reps.erase(it);
            G.dirty = true;
        }
    }
    ++b->gen;
    free_all(r.allocs);
    r.alive = false;
    b->free_slots.push_back(r.slot);
    return MFT_OK;
}

static int batch_create(FtsBatch** out, const mft_fts_desc* descs, int nrep)
{
    MFT_REQUIRE(out && descs && nrep >= 1, MFT_ERR_INVALID, "bad argument");
    *out = nullptr;
    for (int r = 0; r < nrep; ++r) {
        int rc = validate_desc(descs[r]);
        if (rc) return rc;
    }
    FtsBatch* b = nullptr;
    int rc = batch_create_empty(&b, nrep);
    if (rc) return rc;
    for (int r = 0; r < nrep && !rc; ++r) {
        int id = -1;
        rc = batch_add(b, descs[r], &id);
        if (!rc) {
            FtsGroup& G = group_for(b, b->reps[id]);
            G.reps.push_back(id);
            G.dirty = true;
        }
    }
    if (rc) {
        batch_destroy(b);
        return rc;
    }
    *out = b;
    return MFT_OK;
}

static void graph_drop(FtsBatch* b);

static int batch_destroy(FtsBatch* b)
{
    if (!b) return MFT_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    graph_drop(b);
    for (auto& r : b->reps) free_all(r.allocs);
    for (auto& G : b->groups) free_all(G.allocs);
    free_all(b->allocs);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return MFT_OK;
}

static int group_ready(FtsBatch* b, FtsGroup& G)
{
    if (G.dirty || G.allocs.empty()) return group_setup(b, G);
    return MFT_OK;
}

static int batch_init(FtsBatch* b)
{
    int rc;
    for (auto& G : b->groups) {
        if (G.reps.empty()) continue;
        if ((rc = group_ready(b, G))) return rc;
        if ((rc = group_update(b, G, 1))) return rc;
        if ((rc = group_density(b, G))) return rc;
    }
    b->inited = true;
    return MFT_OK;
}

int tild_axis(int sign, int real_in, int real_out, const void* in, void* out, AxisGeom g, int nf, long fs,
              float scale, hipStream_t st);

// d_fts_update1S (fts_updateSchemes.cu:34-52) on k-space arrays
__global__ void k_fts_update1s(float2* __restrict__ w, const float2* __restrict__ f, const double* __restrict__ A,
                               double dt, int M)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const double a = A[i], den = 1.0 + dt * a;
    if (den == 0.0) return;
    const float2 wv = w[i], fv = f[i];
    const double x = ((double)wv.x - dt * ((double)fv.x - a * (double)wv.x)) / den;
    const double y = ((double)wv.y - dt * ((double)fv.y - a * (double)wv.y)) / den;
    w[i] = make_float2((float)x, (float)y);
}

// in-place 3D transform of one complex field of replica r: forward scaled by
// 1/M (Box::cufftWrapperDouble, Box.cu:76-104), inverse unscaled
static int fft3(float2* a, const FtsRep& r, int sign, hipStream_t st)
{
    const mft_fts_desc& d = r.d;
    const float s = sign < 0 ? (float)(1.0 / (double)r.M) : 1.0f;
    for (int ax = 0; ax < 3; ++ax) {
        AxisGeom g{d.nx, d.ny, d.nz, ax};
        const int rc = tild_axis(sign, 0, 0, a, a, g, 1, r.M, ax == 0 ? s : 1.0f, st);
        if (rc) return rc;
    }
    return MFT_OK;
}

// the k-space step of every 1S field of the group (forces from the pass just
// run): w <- IFFT[(w_k - dt (F_k - A w_k)) / (1 + dt A)]
static int group_one_step(FtsBatch* b, FtsGroup& G)
{
    int rc;
    hipStream_t st = b->stream;
    for (int id : G.reps) {
        FtsRep& r = b->reps[id];
        for (int p = 0; p < r.d.npotentials; ++p) {
            const mft_fts_potential& P = r.d.potentials[p];
            if (P.scheme != MFT_FTS_1S) continue;
            struct {
                float2* w;
                float2* f;
                double* A;
                double dt;
            } fields[2] = {{r.wpl[p], r.f1pl[p], r.Akpl[p], P.delt_pl}, {r.wmi[p], r.f1mi[p], r.Akmi[p], P.delt_mi}};
            for (int k = 0; k < (P.kind == MFT_FTS_FLORY ? 2 : 1); ++k) {
                if ((rc = fft3(fields[k].f, r, -1, st))) return rc;
                if ((rc = fft3(fields[k].w, r, -1, st))) return rc;
                k_fts_update1s<<<(r.M + 255) / 256, 256, 0, st>>>(fields[k].w, fields[k].f, fields[k].A, fields[k].dt,
                                                                   r.M);
                if ((rc = fft3(fields[k].w, r, +1, st))) return rc;
            }
        }
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

// one CL iteration of every group: field updates from the previous
// densities, propagators, densities and Q; then every slot's counter
// advances (a slot is reset when a replica takes it)
static int batch_iteration(FtsBatch* b)
{
    int rc;
    for (auto& G : b->groups) {
        if (G.reps.empty()) continue;
        if ((rc = group_update(b, G, 0))) return rc;
        if (G.s1) {
            // k-space step of the 1S fields, then the species fields again
            if ((rc = group_one_step(b, G))) return rc;
            if ((rc = group_update(b, G, 1))) return rc;
        }
        if ((rc = group_zero_mean(b, G, false))) return rc;
        if ((rc = group_density(b, G))) return rc;
        if (G.pc) {
            // FTS_Box.cu:53-81: correct, rebuild species fields, densities again
            if ((rc = group_update(b, G, 2))) return rc;
            if ((rc = group_zero_mean(b, G, true))) return rc;
            if ((rc = group_density(b, G))) return rc;
        }
    }
    for (int o = 0; o < b->nslots; o += 256) {
        int cnt = std::min(256, b->nslots - o);
        k_fts_iter_incr<<<1, 256, 0, b->stream>>>(b->d_iter + o, cnt);
    }
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

static void graph_drop(FtsBatch* b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

// Records one iteration (kernel launches only: no memset, copy or host
// synchronisation happens inside it) into a graph.
static int graph_capture(FtsBatch* b)
{
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error(MFT_ERR_DEVICE, "begin capture: %s", hipGetErrorString(e));
        return MFT_ERR_DEVICE;
    }
    const int rc = batch_iteration(b);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->stream, &g);
    if (rc || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        set_error(MFT_ERR_DEVICE, "capture of the iteration failed");
        return MFT_ERR_DEVICE;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error(MFT_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(e));
        return MFT_ERR_DEVICE;
    }
    b->graph = g;
    b->gexec = x;
    return MFT_OK;
}

// Activity changes wait for one density pass (mft_fts_set_activity): once an
// iteration has run, the descriptors take the new values.
static int apply_pending(FtsBatch* b)
{
    bool any = false;
    for (auto& r : b->reps) any = any || (r.alive && r.act_pending);
    if (!any) return MFT_OK;
    MFT_HIP(hipStreamSynchronize(b->stream));
    for (auto& r : b->reps) {
        if (!r.alive || !r.act_pending) continue;
        for (int m = 0; m < r.d.nmolecules; ++m)
            r.act_dev[m] = r.d.molecules[m].activity > 0.0 ? r.d.molecules[m].activity : 0.0;
        r.act_pending = false;
        group_for(b, r).dirty = true;
    }
    int rc;
    for (auto& G : b->groups)
        if (!G.reps.empty() && (rc = group_ready(b, G))) return rc;
    return MFT_OK;
}

static int batch_step(FtsBatch* b, int niter)
{
    MFT_REQUIRE(b->inited, MFT_ERR_INVALID, "call init before stepping");
    int rc;
    for (auto& G : b->groups)
        if (!G.reps.empty() && (rc = group_ready(b, G))) return rc;
    for (int it = 0; it < niter; ++it) {
        if (b->use_graph && b->graph_gen != b->gen) {
            // the launches changed: run this iteration directly (it also sets
            // up anything created on first use) and capture the next one
            graph_drop(b);
            b->graph_gen = b->gen;
            if ((rc = batch_iteration(b))) return rc;
        } else if (b->use_graph) {
            if (!b->gexec && graph_capture(b)) {
                // no graph on this device or runtime: launch directly from now on
                b->use_graph = false;
                if ((rc = batch_iteration(b))) return rc;
                b->iter++;
                if ((rc = apply_pending(b))) return rc;
                continue;
            }
            MFT_HIP(hipGraphLaunch(b->gexec, b->stream));
        } else if ((rc = batch_iteration(b))) {
            return rc;
        }
        b->iter++;
        if ((rc = apply_pending(b))) return rc;
    }
    return MFT_OK;
}

static int batch_set_fields(FtsBatch* b, int rep, int pot, const float* wpl, const float* wmi)
{
    MFT_REQUIRE(rep >= 0 && rep < (int)b->reps.size() && b->reps[rep].alive, MFT_ERR_INVALID,
                "replica out of range");
    FtsRep& r = b->reps[rep];
    MFT_REQUIRE(pot >= 0 && pot < r.d.npotentials, MFT_ERR_INVALID, "potential out of range");
    MFT_HIP(hipStreamSynchronize(b->stream));
    const size_t M = r.M;
    if (wpl) {
        for (size_t i = 0; i < 2 * M; ++i)
            MFT_REQUIRE(std::isfinite(wpl[i]), MFT_ERR_NONFINITE, "field value %zu is not finite", i);
        MFT_HIP(hipMemcpy(r.wpl[pot], wpl, sizeof(float2) * M, hipMemcpyHostToDevice));
    }
    if (wmi) {
        MFT_REQUIRE(r.wmi[pot], MFT_ERR_INVALID, "potential %d has no w- field", pot);
        for (size_t i = 0; i < 2 * M; ++i)
            MFT_REQUIRE(std::isfinite(wmi[i]), MFT_ERR_NONFINITE, "field value %zu is not finite", i);
        MFT_HIP(hipMemcpy(r.wmi[pot], wmi, sizeof(float2) * M, hipMemcpyHostToDevice));
    }
    b->inited = false;
    return MFT_OK;
}

static int batch_observables(FtsBatch* b, int rep, double* Q, double* H, double* parts)
{
    MFT_REQUIRE(rep >= 0 && rep < (int)b->reps.size() && b->reps[rep].alive, MFT_ERR_INVALID,
                "replica out of range");
    FtsRep& r = b->reps[rep];
    const mft_fts_desc& d = r.d;
    MFT_HIP(hipStreamSynchronize(b->stream));
    std::vector<double> hp((size_t)r.nHpart * kFtsMaxPots * 6);
    std::vector<double2> q(d.nmolecules);
    MFT_HIP(hipMemcpy(hp.data(), r.Hpart, sizeof(double) * hp.size(), hipMemcpyDeviceToHost));
    MFT_HIP(hipMemcpy(q.data(), r.Q, sizeof(double2) * d.nmolecules, hipMemcpyDeviceToHost));
    // partial sums in a fixed order
    double sums[kFtsMaxPots * 6] = {0};
    for (int blk = 0; blk < r.nHpart; ++blk)
        for (int k = 0; k < d.npotentials * 6; ++k) sums[k] += hp[(size_t)blk * kFtsMaxPots * 6 + k];
    std::complex<double> Ht(0.0, 0.0);
    const std::complex<double> I(0.0, 1.0);
    int np = 0;
    for (int p = 0; p < d.npotentials; ++p) {
        const mft_fts_potential& P = d.potentials[p];
        std::complex<double> w2(sums[p * 6 + 0], sums[p * 6 + 1]), w1(sums[p * 6 + 2], sums[p * 6 + 3]),
            m2(sums[p * 6 + 4], sums[p * 6 + 5]);
        std::complex<double> h;
        if (P.kind == MFT_FTS_FLORY)  // fts_potential_flory.cu:252-274
            h = w2 * r.gvol * d.C / P.chiN + m2 * r.gvol * d.C / P.chiN;
        else if (P.kind == MFT_FTS_HELFAND)  // fts_potential_helfand.cu:254-275
            h = w2 * r.gvol * d.C / 2.0 / P.kappaN - I * d.C * (w1 * r.gvol);
        else if (P.kind == MFT_FTS_EDWARDS)  // PotentialEdwards::calcHamiltonian
            h = w2 * r.gvol / 2.0 / P.kappaN;
        else  // fts_potential_incompress.cu:232-247
            h = -I * d.C * (w1 * r.gvol);
        if (parts) {
            parts[2 * np] = h.real();
            parts[2 * np + 1] = h.imag();
        }
        ++np;
        Ht += h;
    }
    for (int m = 0; m < d.nmolecules; ++m) {
        std::complex<double> Qm(q[m].x, q[m].y);
        // LinearMolec::calcHTerm (fts_molecule_linear.cu:368-382)
        std::complex<double> h = d.molecules[m].activity > 0.0 ? -d.molecules[m].activity * r.V * Qm
                                                               : -r.nmolecs[m] * std::log(Qm);
        if (Q) {
            Q[2 * m] = Qm.real();
            Q[2 * m + 1] = Qm.imag();
        }
        if (parts) {
            parts[2 * np] = h.real();
            parts[2 * np + 1] = h.imag();
        }
        ++np;
        Ht += h;
    }
    if (H) {
        H[0] = Ht.real();
        H[1] = Ht.imag();
    }
    // FTS_Box.cu:150-152 dies on NaN; here it is a status
    MFT_REQUIRE(std::isfinite(Ht.real()) && std::isfinite(Ht.imag()), MFT_ERR_NONFINITE,
                "Found Heff = NaN (replica %d)", rep);
    return MFT_OK;
}

static int group_of(FtsBatch* b, int rep)
{
    for (size_t g = 0; g < b->groups.size(); ++g)
        for (size_t k = 0; k < b->groups[g].reps.size(); ++k)
            if (b->groups[g].reps[k] == rep) return (int)(g * 4096 + k);
    return -1;
}

static int batch_densities(FtsBatch* b, int rep, float* rho)
{
    MFT_REQUIRE(rep >= 0 && rep < (int)b->reps.size() && rho && b->reps[rep].alive, MFT_ERR_INVALID, "bad argument");
    FtsRep& r = b->reps[rep];
    int gk = group_of(b, rep);
    MFT_REQUIRE(gk >= 0, MFT_ERR_INVALID, "replica %d is not active", rep);
    FtsGroup& G = b->groups[gk / 4096];
    int k = gk % 4096;
    int rc = group_ready(b, G);
    if (rc) return rc;
    float2* out = nullptr;
    MFT_HIP(hipMalloc((void**)&out, sizeof(float2) * r.M * r.d.nspecies));
    dim3 grid((r.M + kFtsThreads - 1) / kFtsThreads, 1);
    k_fts_species_rho<<<grid, kFtsThreads, 0, b->stream>>>(G.d_upd + k, out);
    MFT_HIP(hipStreamSynchronize(b->stream));
    MFT_HIP(hipMemcpy(rho, out, sizeof(float2) * r.M * r.d.nspecies, hipMemcpyDeviceToHost));
    MFT_HIP(hipFree(out));
    return MFT_OK;
}

static int batch_get_fields(FtsBatch* b, int rep, int pot, float* wpl, float* wmi)
{
    MFT_REQUIRE(rep >= 0 && rep < (int)b->reps.size() && b->reps[rep].alive, MFT_ERR_INVALID,
                "replica out of range");
    FtsRep& r = b->reps[rep];
    MFT_REQUIRE(pot >= 0 && pot < r.d.npotentials, MFT_ERR_INVALID, "potential out of range");
    MFT_HIP(hipStreamSynchronize(b->stream));
    if (wpl) MFT_HIP(hipMemcpy(wpl, r.wpl[pot], sizeof(float2) * r.M, hipMemcpyDeviceToHost));
    if (wmi) {
        MFT_REQUIRE(r.wmi[pot], MFT_ERR_INVALID, "potential %d has no w- field", pot);
        MFT_HIP(hipMemcpy(wmi, r.wmi[pot], sizeof(float2) * r.M, hipMemcpyDeviceToHost));
    }
    return MFT_OK;
}

static int batch_species_fields(FtsBatch* b, int rep, float* W)
{
    MFT_REQUIRE(rep >= 0 && rep < (int)b->reps.size() && b->reps[rep].alive && W, MFT_ERR_INVALID,
                "bad argument");
    FtsRep& r = b->reps[rep];
    MFT_HIP(hipStreamSynchronize(b->stream));
    MFT_HIP(hipMemcpy(W, r.W, sizeof(float2) * r.M * r.d.nspecies, hipMemcpyDeviceToHost));
    return MFT_OK;
}

// Chemical potential of every molecule, complex: ln(n / (V Q)) for a
// canonical molecule (the discrete-chain ideal-gas form in which
// rho = n/(Q V) e^{W} sum q q^dag is written, fts_molecule_linear.cu:286-291),
// ln(activity) for a grand canonical one (its density factor is the activity
// itself, fts_molecule_linear.cu:289-291).  The reference prints no chemical
// potential; this is the quantity its density factor implies.
// Save (restore = 0) or restore (restore = 1) the potential fields and the
// activities of replica rep in snapshot slot `slot`, device to device.  A
// restore is followed by the density pass of the restored fields, so the next
// iteration starts from the saved state (the iteration count, and with it the
// noise stream, carries on).
static int batch_snapshot(FtsBatch* b, int rep, int slot, int restore)
{
    MFT_REQUIRE(rep >= 0 && rep < (int)b->reps.size() && b->reps[rep].alive, MFT_ERR_INVALID,
                "replica out of range");
    MFT_REQUIRE(slot >= 0 && slot < kFtsMaxSnapshots, MFT_ERR_INVALID, "snapshot slot %d out of range (0..%d)", slot,
                kFtsMaxSnapshots - 1);
    FtsRep& r = b->reps[rep];
    std::vector<float2*> fields;
    for (int p = 0; p < r.d.npotentials; ++p) fields.push_back(r.wpl[p]);
    for (int p = 0; p < r.d.npotentials; ++p)
        if (r.wmi[p]) fields.push_back(r.wmi[p]);
    const size_t M = r.M;
    if (!restore) {
        if (!r.snap[slot]) {
            int rc = dev_alloc(r.allocs, &r.snap[slot], fields.size() * M);
            if (rc) return rc;
        }
        for (size_t f = 0; f < fields.size(); ++f)
            MFT_HIP(hipMemcpyAsync(r.snap[slot] + f * M, fields[f], sizeof(float2) * M, hipMemcpyDeviceToDevice,
                                   b->stream));
        for (int m = 0; m < r.d.nmolecules; ++m) r.snap_act[slot][m] = r.d.molecules[m].activity;
        return MFT_OK;
    }
    MFT_REQUIRE(r.snap[slot], MFT_ERR_INVALID, "snapshot slot %d is empty", slot);
    for (size_t f = 0; f < fields.size(); ++f)
        MFT_HIP(hipMemcpyAsync(fields[f], r.snap[slot] + f * M, sizeof(float2) * M, hipMemcpyDeviceToDevice,
                               b->stream));
    for (int m = 0; m < r.d.nmolecules; ++m) {
        r.d.molecules[m].activity = r.snap_act[slot][m];
        r.act_dev[m] = r.snap_act[slot][m];
    }
    r.act_pending = false;
    group_for(b, r).dirty = true;
    return batch_init(b);
}

static int batch_chemical_potentials(FtsBatch* b, int rep, double* mu)
{
    MFT_REQUIRE(rep >= 0 && rep < (int)b->reps.size() && b->reps[rep].alive && mu, MFT_ERR_INVALID,
                "bad argument");
    FtsRep& r = b->reps[rep];
    double2 q[kFtsMaxMol];
    MFT_HIP(hipStreamSynchronize(b->stream));
    MFT_HIP(hipMemcpy(q, r.Q, sizeof(double2) * r.d.nmolecules, hipMemcpyDeviceToHost));
    for (int m = 0; m < r.d.nmolecules; ++m) {
        std::complex<double> v;
        const double z = r.d.molecules[m].activity;
        if (z > 0.0)
            v = std::log(std::complex<double>(z, 0.0));
        else
            v = std::log(std::complex<double>(r.nmolecs[m] / r.V, 0.0)) - std::log(std::complex<double>(q[m].x, q[m].y));
        mu[2 * m] = v.real();
        mu[2 * m + 1] = v.imag();
        MFT_REQUIRE(std::isfinite(v.real()) && std::isfinite(v.imag()), MFT_ERR_NONFINITE,
                    "chemical potential of molecule %d is not finite", m + 1);
    }
    return MFT_OK;
}

} // namespace mft

using namespace mft;

struct mft_fts : FtsBatch {};
struct mft_fts_batch : FtsBatch {};

__global__ void k_fts_bond_factor(float2* g2, int nx, int ny, int nz, float Lx, float Ly, float Lz, double Nr)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long M = (long)nx * ny * nz;
    if (i >= M) return;
    int ix = i % nx, iy = (i / nx) % ny, iz = i / ((long)nx * ny);
    double kx = kcomp(ix, nx, Lx), ky = kcomp(iy, ny, Ly), kz = kcomp(iz, nz, Lz);
    g2[i] = make_float2((float)exp(-(kx * kx + ky * ky + kz * kz) / (Nr - 1.0)), 0.f);
}

extern "C" {

int mft_fts_bond_factor(float* g2, const int Nx[3], const float L[3], double Nr, void* stream)
{
    MFT_REQUIRE(g2 && Nx && L && Nr > 1.0, MFT_ERR_INVALID, "bad argument");
    long M = (long)Nx[0] * Nx[1] * Nx[2];
    k_fts_bond_factor<<<(unsigned)((M + 255) / 256), 256, 0, (hipStream_t)stream>>>((float2*)g2, Nx[0], Nx[1], Nx[2],
                                                                                    L[0], L[1], L[2], Nr);
    MFT_LAUNCH_CHECK();
    return MFT_OK;
}

int mft_fts_create(mft_fts** sys, const mft_fts_desc* desc)
{
    FtsBatch* b = nullptr;
    int rc = batch_create(&b, desc, 1);
    if (sys) *sys = (mft_fts*)b;
    return rc;
}
int mft_fts_destroy(mft_fts* s) { return batch_destroy(s); }
void* mft_fts_stream(mft_fts* s) { return s ? (void*)s->stream : nullptr; }
int mft_fts_set_fields(mft_fts* s, int pot, const float* wpl, const float* wmi)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null system");
    return batch_set_fields(s, 0, pot, wpl, wmi);
}
int mft_fts_get_fields(mft_fts* s, int pot, float* wpl, float* wmi)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null system");
    return batch_get_fields(s, 0, pot, wpl, wmi);
}
int mft_fts_init(mft_fts* s)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null system");
    return batch_init(s);
}
int mft_fts_step(mft_fts* s, int niter)
{
    MFT_REQUIRE(s && niter >= 0, MFT_ERR_INVALID, "bad argument");
    return batch_step(s, niter);
}
int mft_fts_synchronize(mft_fts* s)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null system");
    MFT_HIP(hipStreamSynchronize(s->stream));
    return MFT_OK;
}
int mft_fts_get_densities(mft_fts* s, float* rho)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null system");
    return batch_densities(s, 0, rho);
}
int mft_fts_get_species_fields(mft_fts* s, float* W)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null system");
    return batch_species_fields(s, 0, W);
}
int mft_fts_snapshot(mft_fts* s, int slot, int restore)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null system");
    return batch_snapshot(s, 0, slot, restore);
}
int mft_fts_chemical_potentials(mft_fts* s, double* mu)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null system");
    return batch_chemical_potentials(s, 0, mu);
}
int mft_fts_get_propagators(mft_fts* s, int m, float* q, float* qdag)
{
    MFT_REQUIRE(s && m >= 0 && m < s->reps[0].d.nmolecules, MFT_ERR_INVALID, "bad argument");
    FtsRep& r = s->reps[0];
    MFT_HIP(hipStreamSynchronize(s->stream));
    const mft_fts_molecule& mol = r.d.molecules[m];
    const int Nt = r.Ntot[m];
    const size_t M = r.M;
    (void)mol;
    for (int c = 0; c < 2; ++c) {
        float* out = c == 0 ? q : qdag;
        if (out) MFT_HIP(hipMemcpy(out, c == 0 ? r.p[m] : r.pd[m], sizeof(float2) * M * Nt, hipMemcpyDeviceToHost));
    }
    return MFT_OK;
}
int mft_fts_set_activity(mft_fts* s, int m, double activity)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null system");
    FtsRep& r = s->reps[0];
    MFT_REQUIRE(m >= 0 && m < r.d.nmolecules, MFT_ERR_INVALID, "Invalid molecule index - out of range!");
    MFT_REQUIRE(r.d.molecules[m].activity > 0.0, MFT_ERR_INVALID, "molecule %d is not grand canonical", m + 1);
    MFT_REQUIRE(activity > 0.0 && std::isfinite(activity), MFT_ERR_INVALID, "activity must be positive");
    MFT_HIP(hipStreamSynchronize(s->stream));
    r.d.molecules[m].activity = activity;
    if (s->inited) {
        r.act_pending = true;
    } else {
        r.act_dev[m] = activity;
        group_for(s, r).dirty = true;
    }
    return MFT_OK;
}

int mft_fts_nsites(mft_fts* s, double* nsites)
{
    MFT_REQUIRE(s && nsites, MFT_ERR_INVALID, "null argument");
    FtsRep& r = s->reps[0];
    std::vector<double2> q(r.d.nmolecules);
    MFT_HIP(hipStreamSynchronize(s->stream));
    MFT_HIP(hipMemcpy(q.data(), r.Q, sizeof(double2) * q.size(), hipMemcpyDeviceToHost));
    for (int m = 0; m < r.d.nmolecules; ++m)
        nsites[m] = r.act_dev[m] > 0.0 ? r.act_dev[m] * r.V * q[m].x * r.Ntot[m] : r.nmolecs[m] * r.Ntot[m];
    return MFT_OK;
}

int mft_fts_observables(mft_fts* s, double* Q, double* H, double* parts)
{
    MFT_REQUIRE(s, MFT_ERR_INVALID, "null system");
    return batch_observables(s, 0, Q, H, parts);
}

int mft_fts_batch_create(mft_fts_batch** b, const mft_fts_desc* descs, int nrep)
{
    FtsBatch* x = nullptr;
    int rc = batch_create(&x, descs, nrep);
    if (b) *b = (mft_fts_batch*)x;
    return rc;
}
int mft_fts_batch_destroy(mft_fts_batch* b) { return batch_destroy(b); }
int mft_fts_batch_set_fields(mft_fts_batch* b, int rep, int pot, const float* wpl, const float* wmi)
{
    MFT_REQUIRE(b, MFT_ERR_INVALID, "null batch");
    return batch_set_fields(b, rep, pot, wpl, wmi);
}
int mft_fts_batch_init(mft_fts_batch* b)
{
    MFT_REQUIRE(b, MFT_ERR_INVALID, "null batch");
    return batch_init(b);
}
// Chunk size of every group's propagation (0: one chunk per group).
int mft_fts_batch_set_chunk_bytes(mft_fts_batch* b, double bytes)
{
    MFT_REQUIRE(b && bytes >= 0.0, MFT_ERR_INVALID, "bad argument");
    MFT_HIP(hipStreamSynchronize(b->stream));
    int rc;
    for (auto& G : b->groups) {
        if (G.reps.empty()) continue;
        G.chunk_bytes = (size_t)bytes;
        if ((rc = group_setup(b, G))) return rc;
    }
    return MFT_OK;
}

// Per group, the propagation chunk size whose density pass (propagators,
// densities and Q from the current fields, which it leaves unchanged) runs
// fastest: one chunk, or chunks of 512, 256, 128 or 64 MB of per-bead
// working set.
int mft_fts_batch_autotune(mft_fts_batch* b, int reps, double* ms_per_pass)
{
    MFT_REQUIRE(b && reps >= 1, MFT_ERR_INVALID, "bad argument");
    MFT_REQUIRE(b->inited, MFT_ERR_INVALID, "call init before the autotuner");
    MFT_HIP(hipStreamSynchronize(b->stream));
    hipEvent_t e0, e1;
    MFT_HIP(hipEventCreate(&e0));
    MFT_HIP(hipEventCreate(&e1));
    const size_t MB = (size_t)1 << 20;
    const size_t cand[] = {0, 512 * MB, 256 * MB, 128 * MB, 64 * MB};
    double total = 0.0;
    int rc = MFT_OK;
    for (auto& G : b->groups) {
        if (G.reps.empty()) continue;
        size_t ws = 0;
        for (int k : G.reps) {
            const FtsRep& r = b->reps[k];
            int chains = 0;
            for (int m = 0; m < r.d.nmolecules; ++m) chains += r.sym[m] ? 1 : 2;
            ws += (size_t)chains * 4 * sizeof(float2) * (size_t)r.M;
        }
        float best = 1e30f;
        size_t best_c = 0;
        for (size_t c : cand) {
            if (c != 0 && c >= ws) continue;   // one chunk anyway
            G.chunk_bytes = c;
            if ((rc = group_setup(b, G)) || (rc = group_density(b, G))) break;   // warm
            MFT_HIP(hipEventRecord(e0, b->stream));
            for (int r = 0; r < reps && !rc; ++r) rc = group_density(b, G);
            if (rc) break;
            MFT_HIP(hipEventRecord(e1, b->stream));
            MFT_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            MFT_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                best_c = c;
            }
        }
        if (rc) break;
        G.chunk_bytes = best_c;
        if ((rc = group_setup(b, G))) break;
        total += best / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc && ms_per_pass) *ms_per_pass = total;
    return rc;
}

int mft_fts_batch_step(mft_fts_batch* b, int niter)
{
    MFT_REQUIRE(b && niter >= 0, MFT_ERR_INVALID, "bad argument");
    return batch_step(b, niter);
}
int mft_fts_batch_synchronize(mft_fts_batch* b)
{
    MFT_REQUIRE(b, MFT_ERR_INVALID, "null batch");
    MFT_HIP(hipStreamSynchronize(b->stream));
    return MFT_OK;
}
int mft_fts_batch_observables(mft_fts_batch* b, int rep, double* Q, double* H, double* parts)
{
    MFT_REQUIRE(b, MFT_ERR_INVALID, "null batch");
    return batch_observables(b, rep, Q, H, parts);
}
int mft_fts_batch_get_densities(mft_fts_batch* b, int rep, float* rho)
{
    MFT_REQUIRE(b, MFT_ERR_INVALID, "null batch");
    return batch_densities(b, rep, rho);
}
int mft_fts_batch_get_fields(mft_fts_batch* b, int rep, int pot, float* wpl, float* wmi)
{
    MFT_REQUIRE(b, MFT_ERR_INVALID, "null batch");
    return batch_get_fields(b, rep, pot, wpl, wmi);
}
int mft_fts_batch_get_species_fields(mft_fts_batch* b, int rep, float* W)
{
    MFT_REQUIRE(b, MFT_ERR_INVALID, "null batch");
    return batch_species_fields(b, rep, W);
}
int mft_fts_batch_chemical_potentials(mft_fts_batch* b, int rep, double* mu)
{
    MFT_REQUIRE(b, MFT_ERR_INVALID, "null batch");
    return batch_chemical_potentials(b, rep, mu);
}

} // extern "C"

// ------------------------------------------------- pool for the server ---
// The serving loop keeps one batch whose membership changes between
// iterations (serve.hip); these wrap the batch operations for it.
namespace mft {

int fts_pool_create(FtsBatch** b, int nslots)
{
    int rc = batch_create_empty(b, nslots);
    if (!rc) (*b)->inited = true;
    return rc;
}
int fts_pool_destroy(FtsBatch* b) { return batch_destroy(b); }
int fts_pool_add(FtsBatch* b, const mft_fts_desc& d, int* rep) { return batch_add(b, d, rep); }
int fts_pool_set_fields(FtsBatch* b, int rep, int pot, const float* wpl, const float* wmi)
{
    int rc = batch_set_fields(b, rep, pot, wpl, wmi);
    b->inited = true; // the pool's replicas are initialised one by one
    return rc;
}
int fts_pool_activate(FtsBatch* b, const std::vector<int>& reps) { return batch_activate(b, reps); }
int fts_pool_step(FtsBatch* b, int n) { return batch_step(b, n); }
hipStream_t fts_pool_stream(FtsBatch* b) { return b->stream; }
int fts_pool_observables(FtsBatch* b, int rep, double* Q, double* H, double* parts)
{
    return batch_observables(b, rep, Q, H, parts);
}
int fts_pool_densities(FtsBatch* b, int rep, float* rho) { return batch_densities(b, rep, rho); }
int fts_pool_get_fields(FtsBatch* b, int rep, int pot, float* wpl, float* wmi)
{
    return batch_get_fields(b, rep, pot, wpl, wmi);
}
int fts_pool_chemical_potentials(FtsBatch* b, int rep, double* mu) { return batch_chemical_potentials(b, rep, mu); }
// whether every molecule's Q of the last density pass is finite (a runaway
// CL trajectory overflows the propagators first); the stream must be idle
int fts_pool_q_finite(FtsBatch* b, int rep, bool* ok)
{
    MFT_REQUIRE(rep >= 0 && rep < (int)b->reps.size() && b->reps[rep].alive, MFT_ERR_INVALID, "replica out of range");
    const FtsRep& r = b->reps[rep];
    double2 q[kFtsMaxMol];
    MFT_HIP(hipMemcpy(q, r.Q, sizeof(double2) * r.d.nmolecules, hipMemcpyDeviceToHost));
    *ok = true;
    // Q = 0 (every propagator underflowed) makes -n log Q infinite as well
    for (int m = 0; m < r.d.nmolecules; ++m)
        *ok = *ok && std::isfinite(q[m].x) && std::isfinite(q[m].y) && (q[m].x != 0.0 || q[m].y != 0.0);
    return MFT_OK;
}
int fts_pool_retire(FtsBatch* b, int rep) { return batch_retire(b, rep); }

// device bytes one replica holds (rep_setup)
size_t fts_bytes_estimate(const mft_fts_desc& d)
{
    const size_t M = (size_t)d.nx * d.ny * d.nz;
    size_t fields = 0;
    for (int p = 0; p < d.npotentials; ++p) fields += d.potentials[p].kind == MFT_FTS_FLORY ? 2 : 1;
    size_t prop = 0, nblk = 0;
    for (int m = 0; m < d.nmolecules; ++m) {
        const mft_fts_molecule& mol = d.molecules[m];
        int Nt = 0;
        for (int k = 0; k < mol.nblocks; ++k) Nt += mol.N[k];
        bool sym = true;
        for (int j = 0; j < mol.nblocks / 2; ++j)
            if (mol.species[j] != mol.species[mol.nblocks - 1 - j] || mol.N[j] != mol.N[mol.nblocks - 1 - j])
                sym = false;
        prop += (size_t)Nt * (sym ? 1 : 2);
        nblk += mol.nblocks;
    }
    const size_t cplx = fields + 2 * (size_t)d.nspecies + prop + nblk + 2 * (size_t)d.nmolecules;
    return cplx * M * sizeof(float2) + (M / kFtsThreads + 1) * (kFtsMaxPots * 6 + 8) * sizeof(double) + 4096;
}

} // namespace mft
