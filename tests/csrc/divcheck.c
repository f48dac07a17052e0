/* Check of the division used by the Solver update (the Solver kind of the
 * register-blocked sweep, taufactor_amd/csrc/tauhip_sweep.h): for d in
 * {inf, 1..8}, with the table pair (r, d) = (RN(1/d), d) and (0, 0) for inf,
 *   q = s*r;  e = fma(-q, d, s);  q = fma(e, r, q);  if e is subnormal: q = s/d
 * must equal the IEEE quotient s/d bit for bit (a zero of either sign counts
 * as equal: the update c + w (q - c) is the same float for both).
 * argv[1] = "full": all 2^32 inputs; otherwise every mantissa of both signs
 * at a set of exponents (guard boundary, working range, extremes) plus all
 * subnormals.  Prints the number of mismatches; exit status 1 if any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static long long check_range(float df, float r, int finite, uint64_t lo, uint64_t hi) {
    long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long long i = (long long)lo; i < (long long)hi; i++) {
        float s = u2f((uint32_t)i);
        if (isnan(s) || isinf(s)) continue;
        float ref = s / df;
        const float dt = finite ? df : 0.0f;  /* the table's (0, 0) pair for inf */
        float q = s * r;
        float e = fmaf(-q, dt, s);
        q = fmaf(e, r, q);
        if (e != 0.0f && fabsf(e) < 0x1p-126f) q = s / df; /* subnormal residual */
        if (f2u(q) != f2u(ref) && !(q == 0.0f && ref == 0.0f)) bad++;
    }
    return bad;
}

int main(int argc, char** argv) {
    int full = argc > 1 && strcmp(argv[1], "full") == 0;
    static const int exps[] = {-126, -104, -103, -102, -101, -100, -99, -98, -60, -3, -2, -1, 0, 1, 2, 127};
    long long total = 0;
    for (int d = 0; d <= 8; d++) {
        volatile float one = 1.0f;
        float df = d ? (float)d : INFINITY;
        float r = one / df;
        long long bad = 0;
        if (full) {
            bad = check_range(df, r, d != 0, 0, 1ull << 32);
        } else {
            for (int sign = 0; sign < 2; sign++) {
                uint64_t sb = (uint64_t)sign << 31;
                bad += check_range(df, r, d != 0, sb, sb + (1u << 23)); /* subnormals and zero */
                for (unsigned k = 0; k < sizeof(exps) / sizeof(exps[0]); k++) {
                    uint64_t base = sb | ((uint64_t)(exps[k] + 127) << 23);
                    bad += check_range(df, r, d != 0, base, base + (1u << 23));
                }
            }
        }
        printf("d=%d mismatches=%lld\n", d, bad);
        total += bad;
    }
    return total ? 1 : 0;
}
