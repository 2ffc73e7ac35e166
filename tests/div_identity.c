/* The force pass's shared-reciprocal division, checked against IEEE division on the CPU.
 *
 * The kernel forms a / b as quot(a, b, y) = RN(q0 + r y) with y = RN(1 / b), q0 = RN(a y) and
 * r = a - b q0 computed exactly by one fma (Markstein's theorem).  It does so only on steps where
 * every position, velocity, density and pressure is 0 or within [2^-40, 2^40] in magnitude
 * (densities nonzero), so its operands are:
 *   numerators:  sums / differences of two such values: 0 or 2^-63 <= |a| <= 2^41;
 *   divisors:    a density rho in [2^-40, 2^40], 2 rho, and d = dist + 1e-16 in [1e-16, 2^42].
 *
 *   div_identity exhaustive B0 B1    every significand of a against divisor significands [B0, B1)
 *   div_identity sampled N SEED      every significand of a against N random divisor significands
 *   div_identity random N SEED       N random operand pairs drawn from the ranges above
 *
 * Prints the number of pairs and of mismatches (bitwise, except that a = -0 may give +0, which the
 * force sums cannot show); exits 1 on any mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float from_bits(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static uint32_t bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

static float quot(float a, float b, float y) {
    const float q0 = a * y;
    const float r = fmaf(-b, q0, a);
    return fmaf(r, y, q0);
}

static int same(float got, float want, float a) {
    if (bits(got) == bits(want)) return 1;
    return a == 0.0f && got == 0.0f && want == 0.0f; /* sign of a zero quotient */
}

static uint64_t state;
static uint64_t next(void) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
/* random float with |x| = significand * 2^e, e uniform in [lo, hi], random sign */
static float rnd(int lo, int hi) {
    const uint64_t r = next();
    const int e = lo + (int)(r % (uint64_t)(hi - lo + 1));
    const uint32_t m = (uint32_t)(r >> 20) & 0x7fffffu;
    return from_bits(((r >> 63) ? 0x80000000u : 0u) | ((uint32_t)(e + 127) << 23) | m);
}

/* mismatches of a / b over every significand of a in [1, 2) */
static long long all_numerators(uint32_t mb) {
    const float b = from_bits(0x3f800000u | mb);
    const float y = 1.0f / b;
    long long bad = 0;
    for (uint32_t ma = 0; ma < (1u << 23); ++ma) {
        const float a = from_bits(0x3f800000u | ma);
        bad += !same(quot(a, b, y), a / b, a);
    }
    return bad;
}

int main(int argc, char** argv) {
    if (argc == 4 && !strcmp(argv[1], "exhaustive")) {
        const uint32_t lo = (uint32_t)strtoul(argv[2], 0, 0), hi = (uint32_t)strtoul(argv[3], 0, 0);
        long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(dynamic, 4)
        for (uint32_t mb = lo; mb < hi; ++mb) bad += all_numerators(mb);
        printf("%lld %lld\n", (long long)(hi - lo) << 23, bad);
        return bad != 0;
    }
    if (argc == 4 && !strcmp(argv[1], "sampled")) {
        const int n = atoi(argv[2]);
        state = strtoull(argv[3], 0, 0);
        uint32_t* mbs = malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; ++i) mbs[i] = (uint32_t)(next() >> 41); /* 23 random bits */
        long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(dynamic, 4)
        for (int i = 0; i < n; ++i) bad += all_numerators(mbs[i]);
        free(mbs);
        printf("%lld %lld\n", (long long)n << 23, bad);
        return bad != 0;
    }
    if (argc == 4 && !strcmp(argv[1], "random")) {
        const long long n = atoll(argv[2]);
        state = strtoull(argv[3], 0, 0);
        long long bad = 0;
        for (long long i = 0; i < n; ++i) {
            /* operands as the kernel sees them */
            const float u = (next() & 7) == 0 ? 0.0f : rnd(-40, 40), v = (next() & 7) == 0 ? 0.0f : rnd(-40, 40);
            const float a = (next() & 1) ? u - v : u + v; /* 0 or 2^-63 <= |a| <= 2^41 */
            const float rho = fabsf(rnd(-40, 40));
            const float y = 1.0f / rho;
            bad += !same(quot(a, rho, y), a / rho, a);                       /* (v_j - v_i) / rho_j */
            bad += !same(quot(a, 2.0f * rho, y * 0.5f), a / (2.0f * rho), a); /* (p_i + p_j) / (2 rho_j) */
            /* r / d with d = |r| + 1e-16 and |r_x| <= |r| */
            const float rx = a, ry = (next() & 3) == 0 ? 0.0f : rnd(-63, 41), rz = (next() & 3) == 0 ? 0.0f : rnd(-63, 41);
            const float dist = sqrtf((rx * rx + ry * ry) + rz * rz);
            const float d = dist + 1e-16f;
            const float yd = 1.0f / d;
            bad += !same(quot(rx, d, yd), rx / d, rx);
            bad += !same(quot(ry, d, yd), ry / d, ry);
        }
        printf("%lld %lld\n", 4 * n, bad);
        return bad != 0;
    }
    fprintf(stderr, "usage: div_identity exhaustive B0 B1 | sampled N SEED | random N SEED\n");
    return 2;
}
