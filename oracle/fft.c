/*
 * Reference-free FFT for the oracle: recursive mixed-radix decimation in time
 * with a generic O(p^2) stage for any prime factor p.  Twiddles are computed
 * in double precision from exact integer angles (j mod n)/n and rounded to
 * REAL once.  Conventions follow cuFFT as the reference calls it
 * (Box.cu:76-127): FORWARD has exponent -2 pi i jk/n, INVERSE +2 pi i jk/n,
 * both unnormalised; the reference's Box wrappers scale the forward transform
 * by 1/M (orc_box_fft).
 *
 * Pinned against numpy.fft golden vectors (tests/golden) and Parseval.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int n;
    int sign;
    cplx* w; /* w[j] = exp(sign * 2 pi i j / n) */
} twtab;

#define NTAB 64
static twtab g_tabs[NTAB];
static int g_ntab = 0;

static const cplx* twiddles(int n, int sign)
{
    for (int i = 0; i < g_ntab; ++i)
        if (g_tabs[i].n == n && g_tabs[i].sign == sign) return g_tabs[i].w;
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < n; ++j) {
        double a = 2.0 * M_PI * (double)j / (double)n;
        w[j] = (real)cos(a) + (real)(sign * sin(a)) * I;
    }
    if (g_ntab < NTAB) {
        g_tabs[g_ntab].n = n;
        g_tabs[g_ntab].sign = sign;
        g_tabs[g_ntab].w = w;
        ++g_ntab;
    }
    return w;
}

static int smallest_factor(int n)
{
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* out[0..n) = DFT of in[0], in[istride], ..., with root table w of length N,
 * where this sub-transform of length n uses roots w[(j * (N/n)) % N]. */
static void fft_rec(const cplx* in, int istride, cplx* out, int n, const cplx* w,
                    int N, cplx* tmp)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = smallest_factor(n);
    int m = n / p;
    for (int q = 0; q < p; ++q)
        fft_rec(in + (size_t)q * istride, istride * p, out + (size_t)q * m, m, w, N, tmp);
    int step = N / n; /* root of unity of order n is w[step] */
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            long e = ((long)q * k) % n;
            tmp[q] = out[(size_t)q * m + k] * w[(e * step) % N];
        }
        for (int j = 0; j < p; ++j) {
            cplx acc = 0;
            for (int q = 0; q < p; ++q) {
                long e = ((long)q * j * m) % n;
                acc += tmp[q] * w[(e * step) % N];
            }
            out[(size_t)k + (size_t)m * j] = acc;
        }
    }
}

/* Power-of-two lengths: iterative radix-2 decimation in time on a
 * bit-reversed copy (same roots table). */
static void fft_pow2(cplx* a, int n, const cplx* w)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            cplx t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; ++k) {
                cplx u = a[i + k];
                cplx v = a[i + k + half] * w[k * step];
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
        }
    }
}

static void fft_line(cplx* data, int n, size_t stride, const cplx* w, cplx* buf)
{
    cplx* in = buf;
    cplx* out = buf + n;
    cplx* tmp = buf + 2 * (size_t)n;
    for (int j = 0; j < n; ++j) in[j] = data[(size_t)j * stride];
    if ((n & (n - 1)) == 0) {
        fft_pow2(in, n, w);
        for (int j = 0; j < n; ++j) data[(size_t)j * stride] = in[j];
        return;
    }
    fft_rec(in, 1, out, n, w, n, tmp);
    for (int j = 0; j < n; ++j) data[(size_t)j * stride] = out[j];
}

void orc_fft1d(cplx* data, int n, int stride, int sign)
{
    if (n <= 1) return;
    cplx* buf = (cplx*)malloc(sizeof(cplx) * 3 * (size_t)n);
    fft_line(data, n, (size_t)stride, twiddles(n, sign), buf);
    free(buf);
}

void orc_fft3d(cplx* data, int nx, int ny, int nz, int sign)
{
    /* lines are independent: with OpenMP (the CPU-baseline build) each thread
     * transforms whole lines, so the result does not depend on the thread count */
    size_t plane = (size_t)nx * ny;
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    const cplx* wx = nx > 1 ? twiddles(nx, sign) : NULL;
    const cplx* wy = ny > 1 ? twiddles(ny, sign) : NULL;
    const cplx* wz = nz > 1 ? twiddles(nz, sign) : NULL;
#pragma omp parallel
    {
        cplx* buf = (cplx*)malloc(sizeof(cplx) * 3 * (size_t)nmax);
        if (wx) {
#pragma omp for schedule(static)
            for (int z = 0; z < nz; ++z)
                for (int y = 0; y < ny; ++y) fft_line(data + z * plane + (size_t)y * nx, nx, 1, wx, buf);
        }
        if (wy) {
#pragma omp for schedule(static)
            for (int z = 0; z < nz; ++z)
                for (int x = 0; x < nx; ++x) fft_line(data + z * plane + x, ny, (size_t)nx, wy, buf);
        }
        if (wz) {
#pragma omp for schedule(static)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) fft_line(data + (size_t)y * nx + x, nz, plane, wz, buf);
        }
        free(buf);
    }
}

void orc_box_fft(cplx* data, const int Nx[3], int Dim, int dir)
{
    int nz = Dim == 3 ? Nx[2] : 1;
    size_t M = (size_t)Nx[0] * Nx[1] * nz;
    if (dir == 1) {
        orc_fft3d(data, Nx[0], Nx[1], nz, -1);
        real s = (real)(1.0 / (double)M);
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < M; ++i) data[i] *= s;
    } else {
        orc_fft3d(data, Nx[0], Nx[1], nz, +1);
    }
}
