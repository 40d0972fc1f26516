/*
 * oracle/philox.h -- counter-based random numbers and the synthetic plasma
 * loader used to build identical inputs for the oracle, the compiled
 * reference harness (oracle/_ref) and the test suite.
 *
 * TEST INFRASTRUCTURE: only tests/, the smoke check and the CPU-baseline leg
 * of the benchmark use this file.  The product (quill_amd/) carries its own
 * generator.
 *
 * Philox4x32-10 is the published generator of Salmon, Moraes, Dror and Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3" (SC'11): ten rounds of two
 * 32x32->64 multiplications with the round keys bumped by the Weyl constants.
 *
 * The reference has no such loader (it fills cells on a regular sub-lattice,
 * src/init.cpp:402-478, and draws thermal momenta with its lagged-Fibonacci
 * generator).  The loader below is this project's own, applied identically on
 * every side so that parity tests compare the same initial state.
 */
#ifndef QUILL_ORACLE_PHILOX_H
#define QUILL_ORACLE_PHILOX_H

#include <stdint.h>
#include <math.h>

typedef struct { uint32_t v[4]; } oq_u32x4;

static inline void oq_mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

static inline oq_u32x4 oq_philox4x32_10(oq_u32x4 ctr, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        oq_mulhilo32(0xD2511F53u, ctr.v[0], &hi0, &lo0);
        oq_mulhilo32(0xCD9E8D57u, ctr.v[2], &hi1, &lo1);
        oq_u32x4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return ctr;
}

/* 24-bit uniform in [0, 1 - 2^-24]: exactly representable in float and double. */
static inline double oq_u24(uint32_t r) { return (double)(r >> 8) * (1.0 / 16777216.0); }
/* 24-bit uniform in (0, 1]. */
static inline double oq_u24_open0(uint32_t r) { return (double)((r >> 8) + 1u) * (1.0 / 16777216.0); }

/*
 * One synthetic macro-particle of a uniform thermal plasma.
 *   gid   : global particle id = linear_cell * ppc + slot
 *   off   : in-cell offsets in [0,1) (24-bit uniforms)
 *   u     : momenta u = p/mc, each component N(0, sigma^2) by Box-Muller
 */
static inline void oq_synth_particle(uint64_t seed, uint64_t gid, double sigma,
                                     double off[3], double u[3])
{
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    oq_u32x4 c0 = {{(uint32_t)gid, (uint32_t)(gid >> 32), 0u, 0x51ADu}};
    oq_u32x4 c1 = {{(uint32_t)gid, (uint32_t)(gid >> 32), 1u, 0x51ADu}};
    oq_u32x4 a = oq_philox4x32_10(c0, k0, k1);
    oq_u32x4 b = oq_philox4x32_10(c1, k0, k1);
    off[0] = oq_u24(a.v[0]);
    off[1] = oq_u24(a.v[1]);
    off[2] = oq_u24(a.v[2]);
    const double two_pi = 6.283185307179586;
    double r1 = sqrt(-2.0 * log(oq_u24_open0(a.v[3])));
    double t1 = two_pi * oq_u24(b.v[0]);
    double r2 = sqrt(-2.0 * log(oq_u24_open0(b.v[1])));
    double t2 = two_pi * oq_u24(b.v[2]);
    u[0] = sigma * r1 * cos(t1);
    u[1] = sigma * r1 * sin(t1);
    u[2] = sigma * r2 * cos(t2);
}

#endif
