/*
 * Philox4x32-10 counter-based generator, written from Salmon et al.,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11 (round function,
 * multipliers and Weyl key increments as published there).
 *
 * The reference draws its noise from cuRAND XORWOW states
 * (device_utils.cu:465-472, Box.cu:465-488, fts_make_CLnoise.cu:7-24,
 * ps_integratorGJF.cu:140); those streams cannot be reproduced.  This build
 * replaces them by Philox keyed by (seed) with counters derived from
 * (site, step, stream tag), so that the oracle and the device draw the same
 * uniforms bit-for-bit.  Normal deviates come from Box-Muller on 24-bit
 * uniforms: u1 = ((r0 >> 8) + 1) * 2^-24 in (0,1], u2 = (r1 >> 8) * 2^-24 in
 * [0,1); z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = sqrt(-2 ln u1) sin(2 pi u2);
 * likewise (r2, r3) -> (z2, z3).
 */
#include "oracle.h"
#include <math.h>

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

static void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(PHILOX_M0, c0, &hi0, &lo0);
        mulhilo(PHILOX_M1, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox_normal4(const uint32_t ctr[4], const uint32_t key[2], double z[4])
{
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    const double two_m24 = 1.0 / 16777216.0;
    for (int h = 0; h < 2; ++h) {
        double u1 = (double)((r[2 * h] >> 8) + 1u) * two_m24;
        double u2 = (double)(r[2 * h + 1] >> 8) * two_m24;
        double rad = sqrt(-2.0 * log(u1));
        double ang = 2.0 * M_PI * u2;
        z[2 * h] = rad * cos(ang);
        z[2 * h + 1] = rad * sin(ang);
    }
}

void orc_philox_normals(double* out, long n, uint64_t seed, uint32_t c1,
                        uint32_t c2, uint32_t c3)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (long i = 0; 4 * i < n; ++i) {
        uint32_t ctr[4] = {(uint32_t)i, c1, c2, c3};
        double z[4];
        orc_philox_normal4(ctr, key, z);
        for (int j = 0; j < 4 && 4 * i + j < n; ++j) out[4 * i + j] = z[j];
    }
}
