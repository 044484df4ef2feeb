/*
 * Oracle: complex-Langevin FTS primitives, restated from the reference
 * sources (complex128 there, REAL here).  Test infrastructure only.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_fts_bond_fft(const int Nx[3], const float L[3], int Dim, double Nr, cplx* g)
{
    int M = Nx[0] * Nx[1] * (Dim == 3 ? Nx[2] : 1);
    for (int i = 0; i < M; ++i) {
        double k[3];
        double k2 = orc_kvec_d(i, Nx, L, Dim, k);
        g[i] = (real)exp(-k2 / (Nr - 1.0));
    }
}

/* One propagator recursion over the blocks in the given order
 * (fts_molecule_linear.cu:150-199 forward, :211-261 complementary). */
static void propagate(const cplx* W, int nblocks, const int* N, const int* spec,
                      int reverse, const cplx* g, const int Nx[3], int Dim, cplx* q)
{
    int M = Nx[0] * Nx[1] * (Dim == 3 ? Nx[2] : 1);
    cplx* expW = (cplx*)malloc(sizeof(cplx) * (size_t)M);
    cplx* a = (cplx*)malloc(sizeof(cplx) * (size_t)M);
    int ind = 0;
    for (int bb = 0; bb < nblocks; ++bb) {
        int b = reverse ? nblocks - 1 - bb : bb;
        const cplx* Wb = W + (size_t)spec[b] * M;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) expW[i] = cexp(-Wb[i]);
        for (int s = 0; s < N[b]; ++s) {
            cplx* qs = q + (size_t)ind * M;
            if (ind == 0) {
                memcpy(qs, expW, sizeof(cplx) * (size_t)M);
            } else {
                memcpy(a, q + (size_t)(ind - 1) * M, sizeof(cplx) * (size_t)M);
                orc_box_fft(a, Nx, Dim, 1);
#pragma omp parallel for schedule(static)
                for (int i = 0; i < M; ++i) a[i] *= g[i];
                orc_box_fft(a, Nx, Dim, -1);
#pragma omp parallel for schedule(static)
                for (int i = 0; i < M; ++i) qs[i] = a[i] * expW[i];
            }
            ++ind;
        }
    }
    free(expW);
    free(a);
}

static int is_symmetric(int nblocks, const int* N, const int* spec)
{
    /* fts_molecule_linear.cu:81-88 */
    for (int j = 0; j < nblocks / 2; ++j)
        if (spec[j] != spec[nblocks - 1 - j] || N[j] != N[nblocks - 1 - j]) return 0;
    return 1;
}

cplx orc_fts_propagators(const cplx* W, int nblocks, const int* N, const int* spec,
                         const cplx* g, const int Nx[3], int Dim, double gvol,
                         double V, cplx* q, cplx* qdag)
{
    int M = Nx[0] * Nx[1] * (Dim == 3 ? Nx[2] : 1);
    int Ntot = 0;
    for (int b = 0; b < nblocks; ++b) Ntot += N[b];
    propagate(W, nblocks, N, spec, 0, g, Nx, Dim, q);
    if (is_symmetric(nblocks, N, spec))
        memcpy(qdag, q, sizeof(cplx) * (size_t)Ntot * M);
    else
        propagate(W, nblocks, N, spec, 1, g, Nx, Dim, qdag);
    /* fts_molecule_linear.cu:265-266 */
    cplx s = orc_csum(q + (size_t)(Ntot - 1) * M, M);
    return s * (real)(gvol / V);
}

void orc_fts_density(const cplx* W, int nblocks, const int* N, const int* spec,
                     const cplx* q, const cplx* qdag, cplx factor, int M,
                     cplx* rhoBlock, cplx* rhoSpec)
{
    int Ntot = 0;
    for (int b = 0; b < nblocks; ++b) Ntot += N[b];
    int ind = 0;
    for (int b = 0; b < nblocks; ++b) {
        cplx* rb = rhoBlock + (size_t)b * M;
        for (int i = 0; i < M; ++i) rb[i] = 0;
        for (int s = 0; s < N[b]; ++s) {
            const cplx* qs = q + (size_t)ind * M;
            const cplx* qd = qdag + (size_t)(Ntot - 1 - ind) * M;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < M; ++i) rb[i] += qs[i] * qd[i];
            ++ind;
        }
        /* e^{+W} of the unsmeared species field (fts_molecule_linear.cu:315-327) */
        const cplx* Wb = W + (size_t)spec[b] * M;
        for (int i = 0; i < M; ++i) rb[i] *= factor * cexp(Wb[i]);
    }
    for (int b = 0; b < nblocks; ++b) {
        cplx* rs = rhoSpec + (size_t)spec[b] * M;
        const cplx* rb = rhoBlock + (size_t)b * M;
        for (int i = 0; i < M; ++i) rs[i] += rb[i];
    }
}

void orc_fts_flory_force(cplx* dHdwpl, cplx* dHdwmi, const cplx* wpl, const cplx* wmi,
                         const cplx* rhoI, const cplx* rhoJ, double C, double chiN,
                         double Nr, int M)
{
    for (int i = 0; i < M; ++i) {
        dHdwpl[i] = (real)(2.0 * C / chiN) * wpl[i] + I * (rhoI[i] + rhoJ[i]) / (real)Nr;
        dHdwmi[i] = (real)(2.0 * C / chiN) * wmi[i] + (rhoJ[i] - rhoI[i]) / (real)Nr;
    }
}

void orc_fts_helfand_force(cplx* dHdw, const cplx* w, const cplx* rhoTot, double C,
                           double kN, double Nr, int M)
{
    for (int i = 0; i < M; ++i)
        dHdw[i] = (real)(C / kN) * w[i] + I * rhoTot[i] / (real)Nr - I * (real)C;
}

void orc_fts_incompress_force(cplx* dHdw, const cplx* w, const cplx* rhoTot,
                              double C, double Nr, int M)
{
    (void)w;
    for (int i = 0; i < M; ++i) dHdw[i] = I * rhoTot[i] / (real)Nr - I * (real)C;
}

/* d_makeEdwardsForce (fts_potential_edwards.cu:227-244): w/B + i rho/Nr */
void orc_fts_edwards_force(cplx* dHdw, const cplx* w, const cplx* rhoTot, double B,
                           double Nr, int M)
{
    for (int i = 0; i < M; ++i) dHdw[i] = w[i] / (real)B + I * rhoTot[i] / (real)Nr;
}

void orc_fts_update_em(cplx* w, const cplx* dHdw, const cplx* noise, double delt, int M)
{
    for (int i = 0; i < M; ++i) {
        w[i] -= (real)delt * dHdw[i];
        if (noise) w[i] += noise[i];
    }
}

void orc_fts_update_empc(cplx* w, const cplx* wo, const cplx* dHdw, const cplx* dHdwP, const cplx* noise,
                         double delt, int M)
{
    const real h = (real)(0.5 * delt);
    for (int i = 0; i < M; ++i) {
        w[i] = wo[i] - h * (dHdw[i] + dHdwP[i]);
        if (noise) w[i] += noise[i];
    }
}

void orc_fts_update_1s(cplx* w, const cplx* dHdw, const double* Ak, double delt, int M)
{
    for (int i = 0; i < M; ++i) {
        const real denom = (real)(1.0 + delt * Ak[i]);
        if (denom != 0)
            w[i] = (w[i] - (real)delt * (dHdw[i] - (real)Ak[i] * w[i])) / denom;
    }
}

cplx orc_csum_prod(const cplx* a, const cplx* b, int M)
{
    double _Complex s = 0;
    for (int i = 0; i < M; ++i) s += (double _Complex)a[i] * (double _Complex)b[i];
    return (cplx)s;
}

cplx orc_csum(const cplx* a, int M)
{
    double _Complex s = 0;
    for (int i = 0; i < M; ++i) s += (double _Complex)a[i];
    return (cplx)s;
}
