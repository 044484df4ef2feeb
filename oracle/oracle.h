/*
 * CPU oracle for the MATILDA.FT hot paths: TILD particle-mesh step and
 * complex-Langevin FTS step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke check and the benchmark's CPU-baseline leg load it.  The
 * product (libmft.so) never links, loads or calls anything here.
 *
 * It restates the reference's algorithms in plain C from the equations in the
 * reference sources (cited per function as file:line under src/ of the
 * reference).  REAL is double for the FP64 oracle (liboracle64.so) and float
 * for the FP32 host build (liboracle32.so) compiled from the same source.
 * Grid indices are always computed in FP32 exactly as the reference does, so
 * they can be compared bit-for-bit.
 */
#ifndef MFT_ORACLE_H
#define MFT_ORACLE_H

#include <stdint.h>
#include <complex.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
typedef ORC_REAL _Complex cplx;

/* ---------------------------------------------------------------- RNG --- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), written from the paper. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* n normals for counters (i, c1, c2, c3), i = 0..ceil(n/4)-1, Box-Muller in
 * double from 24-bit uniforms (the mapping shared with the device). */
void orc_philox_normals(double* out, long n, uint64_t seed, uint32_t c1,
                        uint32_t c2, uint32_t c3);
/* 4 normals for one counter. */
void orc_philox_normal4(const uint32_t ctr[4], const uint32_t key[2], double z[4]);

/* ---------------------------------------------------------------- FFT --- */
/* In-place 1D complex DFT of any length (mixed radix, generic prime stages).
 * sign = -1 forward (cuFFT FORWARD convention), +1 inverse; unnormalised. */
void orc_fft1d(cplx* data, int n, int stride, int sign);
/* 3D (or 2D with nz = 1) transform of an x-fastest grid, unnormalised. */
void orc_fft3d(cplx* data, int nx, int ny, int nz, int sign);
/* Box::cufftWrapperSingle / cufftWrapperDouble (Box.cu:76-127):
 * dir = 1: forward, scaled by 1/M; dir = -1: inverse, unscaled. */
void orc_box_fft(cplx* data, const int Nx[3], int Dim, int dir);

/* -------------------------------------------------------- k-space grid --- */
/* Box::get_kD(int, float*) (Box.cu:157-178): float arithmetic throughout. */
float orc_kvec_f(int id, const int Nx[3], const float L[3], int Dim, float k[3]);
/* Box::get_kD(int, double*) (Box.cu:182-203): k_i computed in float and
 * widened, |k|^2 accumulated in double. */
double orc_kvec_d(int id, const int Nx[3], const float L[3], int Dim, double k[3]);

/* --------------------------------------------------------------- TILD --- */
/* spline_get_weights (ps_grid_utils.cu:17-67), including the p = 4 s^3 -> s^4
 * term at :57. */
void orc_spline_weights(real sx, int p, real* W);
/* d_calcGridWeights (ps_grid_utils.cu:74-204). */
void orc_grid_weights(const float* x, int ns, const int Nx[3], const float dx[3],
                      int pmeorder, int Dim, int* inds, real* W);
/* d_fillDensityGrid / d_fillChargeDensityGrid (device_utils.cu:282-352). */
void orc_fill_density(real* rho, real* rhoq, const float* q, const int* sites,
                      int nsites, const int* inds, const real* W, int G);
/* Potential k-space table u(k) (real part; imag is 0):
 * kind 0 gaussian (ps_potentialGaussian.cu:38-46), 1 erf2
 * (ps_potentialErf2.cu:111-124), 2 erfG (ps_potentialErfG.cu:44-57),
 * 3 charges (ps_potentialCharges.cu:38-53, p1 = l_B, p2 = sigma). */
void orc_uk_table(int kind, double A, double sigma, double Rp, const int Nx[3],
                  const float L[3], int Dim, real* uk);
/* PS_Potential::CalcForces for one source density (ps_potential.cu:64-94):
 * gridF[i*Dim+j] += factor * Re IFFT[-i k_j u(k) FFT(rhoJ)/M]. */
void orc_potential_grid_force(const real* rhoJ, const real* uk, const int Nx[3],
                              const float L[3], int Dim, double factor, real* gridF);
/* PS_Potential::CalcEnergy (ps_potential.cu:133-178): gvol*sum rhoI*(u*rhoJ). */
double orc_potential_energy(const real* rhoI, const real* rhoJ, const real* uk,
                            const int Nx[3], int Dim, double gvol);
/* d_mapGridForcesToPartics / d_mapGridChargeForcesToPartics
 * (device_utils.cu:358-461).  q may be NULL. */
void orc_map_forces(real* f, const real* gridF, const float* q, const int* sites,
                    int nsites, const int* inds, const real* W, int G, double gvol,
                    int Dim);
/* d_bonds (ps_deviceBonds.cu:7-79). */
void orc_bonds(real* f, const int* nBonds, const int* bondedTo, const int* bondType,
               const float* bondReq, const float* bondK, const int* bondStyle,
               const float* x, const float* L, const float* Lh, int ns,
               int maxbonds, int Dim);
/* d_bondStressEnergy energy part (ps_deviceBonds.cu:85-171), per particle,
 * double counted, including the 0.5*Lh minimum image at :132-133. */
void orc_bond_energy(real* e, const int* nBonds, const int* bondedTo,
                     const int* bondType, const float* bondReq, const float* bondK,
                     const int* bondStyle, const float* x, const float* L,
                     const float* Lh, int ns, int maxbonds, int Dim);
/* ps_deviceBonds.cu:85-171, virial terms (xx, yy, zz, xy, xz, yz) per particle */
void orc_bond_virial(real* vir, const int* nBonds, const int* bondedTo,
                     const int* bondType, const float* bondReq, const float* bondK,
                     const int* bondStyle, const float* x, const float* L,
                     const float* Lh, int ns, int maxbonds, int Dim);
/* GJF::finishInitialization coefficients (ps_integratorGJF.cu:56-82). */
void orc_gjf_coeffs(double delt, double mass, double mobility, double out[5]);
/* d_GJF_integrator (ps_integratorGJF.cu:102-163).  coef[type*5 + {a,b,noiseMag,
 * b dt^2/m, b dt/2m}]; xi = unit normals, [ns*Dim] in site order. */
void orc_gjf_step(real* x, real* xo, real* noiseOld, const real* f,
                  const double* coef, const int* type, const float* L,
                  const int* sites, int nsites, int Dim, const double* xi);
/* d_EM_integrator (ps_integratorEM.cu:39-82). */
void orc_em_step(real* x, const real* f, double delt, const float* L,
                 const int* sites, int nsites, int Dim, const double* xi);
/* Threads of the OpenMP build used for the CPU baseline (no effect otherwise). */
void orc_set_threads(int n);

/* ---------------------------------------------------------------- FTS --- */
/* LinearMolec bond factor (fts_molecule_linear.cu:109-122): exp(-k^2/(Nr-1)). */
void orc_fts_bond_fft(const int Nx[3], const float L[3], int Dim, double Nr, cplx* g);
/* LinearMolec::calcPropagators (fts_molecule_linear.cu:144-269).  W holds the
 * species fields [nspecies*M]; blocks given by N[b] and spec[b].  Writes q and
 * qdag [Ntot*M] and returns Q. */
cplx orc_fts_propagators(const cplx* W, int nblocks, const int* N, const int* spec,
                         const cplx* g, const int Nx[3], int Dim, double gvol,
                         double V, cplx* q, cplx* qdag);
/* LinearMolec::calcDensity density part (fts_molecule_linear.cu:278-361):
 * rho_b = factor * e^{W_b} * sum_{s in b} q_s qdag_{Ntot-1-s}, added into the
 * species densities rhoSpec [nspecies*M]. */
void orc_fts_density(const cplx* W, int nblocks, const int* N, const int* spec,
                     const cplx* q, const cplx* qdag, cplx factor, int M,
                     cplx* rhoBlock, cplx* rhoSpec);
/* d_makeFloryForce (fts_potential_flory.cu:279-304). */
void orc_fts_flory_force(cplx* dHdwpl, cplx* dHdwmi, const cplx* wpl, const cplx* wmi,
                         const cplx* rhoI, const cplx* rhoJ, double C, double chiN,
                         double Nr, int M);
/* d_makeHelfandForce (fts_potential_helfand.cu:223-240). */
void orc_fts_helfand_force(cplx* dHdw, const cplx* w, const cplx* rhoTot, double C,
                           double kN, double Nr, int M);
/* d_makeIncompressForce (fts_potential_incompress.cu:202-218). */
void orc_fts_incompress_force(cplx* dHdw, const cplx* w, const cplx* rhoTot,
                              double C, double Nr, int M);
void orc_fts_edwards_force(cplx* dHdw, const cplx* w, const cplx* rhoTot, double B,
                           double Nr, int M);
/* d_fts_updateEM (fts_updateSchemes.cu:8-29); noise may be NULL. */
void orc_fts_update_em(cplx* w, const cplx* dHdw, const cplx* noise, double delt, int M);
/* d_fts_updateEMPC corrector (fts_updateSchemes.cu:56-81):
 * w = wo - delt/2 (dHdw + dHdwP) + noise, with the predictor's noise. */
void orc_fts_update_empc(cplx* w, const cplx* wo, const cplx* dHdw, const cplx* dHdwP, const cplx* noise,
                         double delt, int M);
/* sum_r a(r)*b(r) (complex product, FP64) and sum_r a(r). */
/* d_fts_update1S (fts_updateSchemes.cu:34-52) on k-space arrays (forward
 * transforms scaled by 1/M): w = (w - delt (dHdw - A w)) / (1 + delt A)
 * where the denominator is nonzero; A real; no noise term. */
void orc_fts_update_1s(cplx* w, const cplx* dHdw, const double* Ak, double delt, int M);
cplx orc_csum_prod(const cplx* a, const cplx* b, int M);
cplx orc_csum(const cplx* a, int M);

#endif
