/*
 * Oracle: TILD particle-mesh primitives, restated from the reference sources.
 * Test infrastructure only (see oracle.h).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants.h:9 */
#define REF_PI2 6.2831853071f
/* constants.h:10 (a float literal in the reference; widened here) */
#define REF_PI4 12.56637061435917295384

/* Box::unstack2 (Box.cu:133-154). */
static void unstack(int id, const int Nx[3], int Dim, int nn[3])
{
    if (Dim == 2) {
        nn[1] = id / Nx[0];
        nn[0] = id - nn[1] * Nx[0];
        nn[2] = 0;
    } else {
        nn[2] = id / Nx[1] / Nx[0];
        nn[1] = id / Nx[0] - nn[2] * Nx[1];
        nn[0] = id - (nn[1] + nn[2] * Nx[1]) * Nx[0];
    }
}

float orc_kvec_f(int id, const int Nx[3], const float L[3], int Dim, float k[3])
{
    int n[3];
    unstack(id, Nx, Dim, n);
    volatile float kmag = 0.0f; /* keep float rounding, no contraction */
    for (int i = 0; i < Dim; ++i) {
        volatile float t;
        if ((double)(float)n[i] < (double)(float)Nx[i] / 2.0)
            t = REF_PI2 * (float)n[i];
        else
            t = REF_PI2 * (float)(n[i] - Nx[i]);
        k[i] = t / L[i];
        volatile float sq = k[i] * k[i];
        kmag = kmag + sq;
    }
    return kmag;
}

double orc_kvec_d(int id, const int Nx[3], const float L[3], int Dim, double k[3])
{
    int n[3];
    unstack(id, Nx, Dim, n);
    double kmag = 0.0;
    for (int i = 0; i < Dim; ++i) {
        volatile float t;
        if ((double)(float)n[i] < (double)(float)Nx[i] / 2.0)
            t = REF_PI2 * (float)n[i];
        else
            t = REF_PI2 * (float)(n[i] - Nx[i]);
        volatile float kf = t / L[i];
        k[i] = (double)kf;
        kmag += k[i] * k[i];
    }
    return kmag;
}

void orc_spline_weights(real sx, int p, real* W)
{
    real sx2, sx3, sx4;
    if (p == 0) {
        W[0] = 1;
    } else if (p == 1) {
        W[0] = (real)0.5 - sx;
        W[1] = (real)0.5 + sx;
    } else if (p == 2) {
        sx2 = sx * sx;
        W[0] = (real)0.125 - (real)0.5 * sx + (real)0.5 * sx2;
        W[1] = (real)0.75 - sx2;
        W[2] = (real)0.125 + (real)0.5 * sx + (real)0.5 * sx2;
    } else if (p == 3) {
        sx2 = sx * sx;
        sx3 = sx2 * sx;
        W[0] = (1 - 6 * sx + 12 * sx2 - 8 * sx3) / 48;
        W[1] = (23 - 30 * sx - 12 * sx2 + 24 * sx3) / 48;
        W[2] = (23 + 30 * sx - 12 * sx2 - 24 * sx3) / 48;
        W[3] = (1 + 6 * sx + 12 * sx2 + 8 * sx3) / 48;
    } else if (p == 4) {
        sx2 = sx * sx;
        sx3 = sx2 * sx2; /* ps_grid_utils.cu:57 uses s^4 where s^3 belongs */
        sx4 = sx2 * sx2;
        W[0] = (1 - 8 * sx + 24 * sx2 - 32 * sx3 + 16 * sx4) / 384;
        W[1] = (19 - 44 * sx + 24 * sx2 + 16 * sx3 - 16 * sx4) / 96;
        W[2] = (115 - 120 * sx2 + 48 * sx4) / 192;
        W[3] = (19 + 44 * sx + 24 * sx2 - 16 * sx3 - 16 * sx4) / 96;
        W[4] = (1 + 8 * sx + 24 * sx2 + 32 * sx3 + 16 * sx4) / 384;
    }
}

/* Base grid index of a coordinate, in FP32 exactly as ps_grid_utils.cu:105-113,
 * and the offset gdx computed exactly (double holds the float products). */
static int base_index(float x, float dx, int p, double* gdx)
{
    int g;
    if (p % 2 == 0) {
        volatile float h = 0.5f * dx;
        volatile float s = x + h;
        volatile float q = s / dx;
        g = (int)q;
        *gdx = (double)x - (double)(float)g * (double)dx;
    } else {
        volatile float q = x / dx;
        g = (int)q;
        *gdx = (double)x - ((double)(float)g + 0.5) * (double)dx;
    }
    return g;
}

void orc_grid_weights(const float* x, int ns, const int Nx[3], const float dx[3],
                      int p, int Dim, int* inds, real* W)
{
    int G = 1;
    double gvol = 1.0;
    for (int j = 0; j < Dim; ++j) {
        G *= p + 1;
        gvol *= (double)dx[j];
    }
    int shift = p / 2;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < ns; ++id) {
        real w[3][6];
        int g[3];
        for (int j = 0; j < Dim; ++j) {
            double gdx;
            g[j] = base_index(x[id * Dim + j], dx[j], p, &gdx);
            orc_spline_weights((real)(gdx / (double)dx[j]), p, w[j]);
        }
        int ct = 0;
        if (Dim == 2) {
            for (int ix = 0; ix <= p; ++ix) {
                int n0 = (g[0] + ix - shift + Nx[0]) % Nx[0];
                for (int iy = 0; iy <= p; ++iy) {
                    int n1 = (g[1] + iy - shift + Nx[1]) % Nx[1];
                    inds[(size_t)id * G + ct] = n1 * Nx[0] + n0;
                    W[(size_t)id * G + ct] = (real)((double)w[0][ix] * w[1][iy] / gvol);
                    ++ct;
                }
            }
        } else {
            for (int ix = 0; ix <= p; ++ix) {
                int n0 = (g[0] + ix - shift + Nx[0]) % Nx[0];
                for (int iy = 0; iy <= p; ++iy) {
                    int n1 = (g[1] + iy - shift + Nx[1]) % Nx[1];
                    for (int iz = 0; iz <= p; ++iz) {
                        int n2 = (g[2] + iz - shift + Nx[2]) % Nx[2];
                        inds[(size_t)id * G + ct] = n0 + (n1 + n2 * Nx[1]) * Nx[0];
                        W[(size_t)id * G + ct] =
                            (real)((double)w[0][ix] * w[1][iy] * w[2][iz] / gvol);
                        ++ct;
                    }
                }
            }
        }
    }
}

void orc_fill_density(real* rho, real* rhoq, const float* q, const int* sites,
                      int nsites, const int* inds, const real* W, int G)
{
#ifdef _OPENMP
    /* the CPU-baseline build: private grids per thread, summed at the end */
    int M = 0;
    for (int s = 0; s < nsites; ++s)
        for (int i = 0; i < G; ++i) {
            int g = inds[(size_t)sites[s] * G + i];
            if (g >= M) M = g + 1;
        }
#pragma omp parallel
    {
        real* lr = (real*)calloc((size_t)M, sizeof(real));
        real* lq = rhoq ? (real*)calloc((size_t)M, sizeof(real)) : NULL;
#pragma omp for schedule(static)
        for (int s = 0; s < nsites; ++s) {
            int p = sites[s];
            for (int i = 0; i < G; ++i) {
                size_t idx = (size_t)p * G + i;
                lr[inds[idx]] += W[idx];
                if (lq) lq[inds[idx]] += W[idx] * (real)q[p];
            }
        }
#pragma omp critical
        {
            for (int g = 0; g < M; ++g) rho[g] += lr[g];
            if (lq)
                for (int g = 0; g < M; ++g) rhoq[g] += lq[g];
        }
        free(lr);
        free(lq);
    }
#else
    for (int s = 0; s < nsites; ++s) {
        int p = sites[s];
        for (int i = 0; i < G; ++i) {
            size_t idx = (size_t)p * G + i;
            rho[inds[idx]] += W[idx];
            if (rhoq) rhoq[inds[idx]] += W[idx] * (real)q[p];
        }
    }
#endif
}

void orc_uk_table(int kind, double A, double sigma, double Rp, const int Nx[3],
                  const float L[3], int Dim, real* uk)
{
    int M = Nx[0] * Nx[1] * (Dim == 3 ? Nx[2] : 1);
    double sig2 = sigma * sigma;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) {
        float kf[3];
        double k2 = (double)orc_kvec_f(i, Nx, L, Dim, kf);
        double k = sqrt(k2);
        double u = 0.0;
        if (kind == 0) {
            u = A * exp(-k2 * sig2 / 2.0);
        } else if (kind == 1 || kind == 2) {
            double vol = REF_PI4 * Rp * Rp * Rp / 3.0;
            double shape;
            if (k2 == 0.0) {
                shape = vol;
            } else {
                double x = Rp * k;
                /* 4 pi (sin x - x cos x)/k^3 with a series for small x, where
                 * the direct form cancels catastrophically */
                if (x < 1e-2) {
                    double x2 = x * x;
                    shape = vol * (1.0 - x2 / 10.0 + x2 * x2 / 280.0);
                } else {
                    shape = REF_PI4 * (sin(x) - x * cos(x)) / (k2 * k);
                }
            }
            if (kind == 1)
                u = (k2 == 0.0) ? A * vol * vol : A * exp(-k2 * sig2 * 0.5) * shape * shape;
            else
                u = (k2 == 0.0) ? A * vol : A * exp(-k2 * sig2 * 0.5) * shape;
        } else if (kind == 3) {
            /* A carries l_B */
            u = (k2 > 0.0) ? REF_PI4 * A / k2 * exp(-k2 * sig2) : 0.0;
        }
        uk[i] = (real)u;
    }
}

void orc_potential_grid_force(const real* rhoJ, const real* uk, const int Nx[3],
                              const float L[3], int Dim, double factor, real* gridF)
{
    int M = Nx[0] * Nx[1] * (Dim == 3 ? Nx[2] : 1);
    cplx* a = (cplx*)malloc(sizeof(cplx) * (size_t)M);
    cplx* b = (cplx*)malloc(sizeof(cplx) * (size_t)M);
    for (int i = 0; i < M; ++i) a[i] = rhoJ[i];
    orc_box_fft(a, Nx, Dim, 1);
    for (int j = 0; j < Dim; ++j) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) {
            float kf[3];
            orc_kvec_f(i, Nx, L, Dim, kf);
            /* f_j(k) = -i k_j u(k)  (ps_potentialGaussian.cu:42-45) */
            cplx fk = -I * (real)kf[j] * uk[i];
            b[i] = fk * a[i];
        }
        orc_box_fft(b, Nx, Dim, -1);
        for (int i = 0; i < M; ++i) gridF[(size_t)i * Dim + j] += (real)(factor * creal(b[i]));
    }
    free(a);
    free(b);
}

double orc_potential_energy(const real* rhoI, const real* rhoJ, const real* uk,
                            const int Nx[3], int Dim, double gvol)
{
    int M = Nx[0] * Nx[1] * (Dim == 3 ? Nx[2] : 1);
    cplx* a = (cplx*)malloc(sizeof(cplx) * (size_t)M);
    for (int i = 0; i < M; ++i) a[i] = rhoJ[i];
    orc_box_fft(a, Nx, Dim, 1);
    for (int i = 0; i < M; ++i) a[i] *= uk[i];
    orc_box_fft(a, Nx, Dim, -1);
    double e = 0.0;
    for (int i = 0; i < M; ++i) e += (double)rhoI[i] * creal(a[i]);
    free(a);
    return gvol * e;
}

void orc_map_forces(real* f, const real* gridF, const float* q, const int* sites,
                    int nsites, const int* inds, const real* W, int G, double gvol,
                    int Dim)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < nsites; ++s) {
        int p = sites[s];
        double qi = q ? (double)q[p] : 1.0;
        for (int j = 0; j < Dim; ++j) {
            double acc = 0.0;
            for (int i = 0; i < G; ++i) {
                size_t idx = (size_t)p * G + i;
                acc += (double)gridF[(size_t)inds[idx] * Dim + j] * W[idx] * gvol;
            }
            f[(size_t)p * Dim + j] += (real)(acc * qi);
        }
    }
}

static double min_image(double dr, double L, double half)
{
    if (dr > half) dr -= L;
    else if (dr < -half) dr += L;
    return dr;
}

void orc_bonds(real* f, const int* nBonds, const int* bondedTo, const int* bondType,
               const float* bondReq, const float* bondK, const int* bondStyle,
               const float* x, const float* L, const float* Lh, int ns,
               int maxbonds, int Dim)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ns; ++i) {
        double lf[3] = {0, 0, 0};
        for (int b = 0; b < nBonds[i]; ++b) {
            int id2 = bondedTo[i * maxbonds + b];
            int bt = bondType[i * maxbonds + b];
            double dr[3], mdr2 = 0.0, mf = 0.0;
            for (int j = 0; j < Dim; ++j) {
                dr[j] = min_image((double)x[i * Dim + j] - (double)x[id2 * Dim + j], L[j], Lh[j]);
                mdr2 += dr[j] * dr[j];
            }
            if (mdr2 > 1.0e-4) {
                double mdr = sqrt(mdr2);
                double k = bondK[bt], r0 = bondReq[bt];
                if (bondStyle[bt] == 0)
                    mf = 2.0 * k * (mdr - r0) / mdr;
                else if (bondStyle[bt] == 1)
                    mf = 2.0 * k * mdr / (1.0 - (mdr / r0) * (mdr / r0)) / mdr;
            }
            for (int j = 0; j < Dim; ++j) lf[j] -= mf * dr[j];
        }
        for (int j = 0; j < Dim; ++j) f[i * Dim + j] += (real)lf[j];
    }
}

void orc_bond_energy(real* e, const int* nBonds, const int* bondedTo,
                     const int* bondType, const float* bondReq, const float* bondK,
                     const int* bondStyle, const float* x, const float* L,
                     const float* Lh, int ns, int maxbonds, int Dim)
{
    for (int i = 0; i < ns; ++i) {
        double acc = 0.0;
        for (int b = 0; b < nBonds[i]; ++b) {
            int id2 = bondedTo[i * maxbonds + b];
            int bt = bondType[i * maxbonds + b];
            double dr[3], mdr2 = 0.0;
            for (int j = 0; j < Dim; ++j) {
                /* ps_deviceBonds.cu:132-133 compares against 0.5*Lh */
                dr[j] = min_image((double)x[i * Dim + j] - (double)x[id2 * Dim + j], L[j],
                                  0.5 * (double)Lh[j]);
                mdr2 += dr[j] * dr[j];
            }
            if (mdr2 > 1.0e-4) {
                double mdr = sqrt(mdr2);
                double k = bondK[bt], r0 = bondReq[bt];
                if (bondStyle[bt] == 0) {
                    acc += k * (mdr - r0) * (mdr - r0);
                } else if (bondStyle[bt] == 1 && mdr / r0 < 1.0) {
                    double arg = 1.0 / (1.0 - (mdr / r0) * (mdr / r0));
                    acc += 2.0 * k * arg * mdr;
                }
            }
        }
        e[i] = (real)acc;
    }
}

/* d_bondStressEnergy virial part (ps_deviceBonds.cu:85-171): per particle
 * the six pressure tensor terms -mf dr_a dr_b (xx, yy, zz, xy, xz, yz) summed
 * over its bonds, dr with the 0.5*Lh minimum image (:132-133); mf = 2k(r-r0)/r
 * (harmonic), 2k/(1-(r/r0)^2) (FENE below r0), 0 for r^2 <= 1e-4 and for
 * FENE at or beyond r0 (where the reference leaves mf unset). */
void orc_bond_virial(real* vir, const int* nBonds, const int* bondedTo,
                     const int* bondType, const float* bondReq, const float* bondK,
                     const int* bondStyle, const float* x, const float* L,
                     const float* Lh, int ns, int maxbonds, int Dim)
{
    for (int i = 0; i < ns; ++i) {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int b = 0; b < nBonds[i]; ++b) {
            int id2 = bondedTo[i * maxbonds + b];
            int bt = bondType[i * maxbonds + b];
            double dr[3] = {0, 0, 0}, mdr2 = 0.0;
            for (int j = 0; j < Dim; ++j) {
                dr[j] = min_image((double)x[i * Dim + j] - (double)x[id2 * Dim + j], L[j],
                                  0.5 * (double)Lh[j]);
                mdr2 += dr[j] * dr[j];
            }
            double mf = 0.0;
            if (mdr2 > 1.0e-4) {
                double mdr = sqrt(mdr2);
                double k = bondK[bt], r0 = bondReq[bt];
                if (bondStyle[bt] == 0) mf = 2.0 * k * (mdr - r0) / mdr;
                else if (bondStyle[bt] == 1 && mdr / r0 < 1.0) mf = 2.0 * k / (1.0 - (mdr / r0) * (mdr / r0));
            }
            acc[0] += -mf * dr[0] * dr[0];
            acc[1] += -mf * dr[1] * dr[1];
            acc[2] += -mf * dr[2] * dr[2];
            acc[3] += -mf * dr[0] * dr[1];
            acc[4] += -mf * dr[0] * dr[2];
            acc[5] += -mf * dr[1] * dr[2];
        }
        for (int c = 0; c < 6; ++c) vir[i * 6 + c] = (real)acc[c];
    }
}

void orc_gjf_coeffs(double delt, double m, double D, double out[5])
{
    double b = m / (m + delt / (2.0 * D));
    double a = (1.0 - delt / (2.0 * D * m)) / (1.0 + delt / (2.0 * D * m));
    out[0] = a;
    out[1] = b;
    out[2] = sqrt(2.0 * delt) / sqrt(D);
    out[3] = b * delt * delt / m;
    out[4] = b * delt / (2.0 * m);
}

void orc_gjf_step(real* x, real* xo, real* noiseOld, const real* f,
                  const double* coef, const int* type, const float* L,
                  const int* sites, int nsites, int Dim, const double* xi)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < nsites; ++s) {
        int ind = sites[s];
        const double* c = coef + 5 * type[ind];
        double a = c[0], b = c[1], nmag = c[2], cf = c[3], cn = c[4];
        for (int j = 0; j < Dim; ++j) {
            size_t ai = (size_t)ind * Dim + j;
            double nn = nmag * xi[(size_t)s * Dim + j];
            double xt = x[ai];
            double xold = xo[ai];
            if (xold - xt > L[j] / 2.0) xold -= L[j];
            else if (xold - xt < -L[j] / 2.0) xold += L[j];
            double xn = 2.0 * b * xt - a * xold + cf * (double)f[ai] + cn * (nn + (double)noiseOld[ai]);
            if (xn > L[j]) xn -= L[j];
            else if (xn < 0.0) xn += L[j];
            x[ai] = (real)xn;
            xo[ai] = (real)xt;
            noiseOld[ai] = (real)nn;
        }
    }
}

void orc_em_step(real* x, const real* f, double delt, const float* L,
                 const int* sites, int nsites, int Dim, const double* xi)
{
    double nmag = sqrt(2.0 * delt);
#pragma omp parallel for schedule(static)
    for (int s = 0; s < nsites; ++s) {
        int ind = sites[s];
        for (int j = 0; j < Dim; ++j) {
            size_t ai = (size_t)ind * Dim + j;
            double xt = (double)x[ai] + delt * (double)f[ai] + nmag * xi[(size_t)s * Dim + j];
            if (xt > L[j]) xt -= L[j];
            else if (xt < 0.0) xt += L[j];
            x[ai] = (real)xt;
        }
    }
}

/* Threads of the CPU-baseline build (no effect in the single-thread build). */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}
