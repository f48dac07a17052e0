/* Check of the division used by the face-code MultiPhase update
 * (TF_KIND_MULTIPHASE_FACES in taufactor_amd/csrc/tauhip_sweep.h): for each
 * prefactor pair (r, f) read from stdin as two hexadecimal float bit
 * patterns per line,
 *   q = s*r;  e = fma(-q, f, s);  q = fma(e, r, q);  if |s| < 2^-100: q = s/f
 * must equal the IEEE quotient s/f bit for bit for |s| < 2^101 (r is the
 * correctly rounded reciprocal of f, f in [2^-20, 2^20]; pair (0, 0) stands
 * for f = inf, whose quotient 0 gives the same update as s/inf and is not
 * checked here).  Larger |s| are outside the domain: the sums of a converging
 * iteration are bounded by 12 max|c| max(D) (the fast path gives NaN where
 * s/f overflows to inf).
 * argv[1] = "full": every s with |s| < 2^101 per pair; "quick": every
 * mantissa of both signs at the exponents around the 2^-100 guard, near 1
 * and at the top of the domain, plus all subnormals; otherwise the same at
 * more exponents across the working range.
 * Prints the mismatches per pair; exit status 1 if any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static long long check_range(float f, float r, uint64_t lo, uint64_t hi) {
    long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long long i = (long long)lo; i < (long long)hi; i++) {
        float s = u2f((uint32_t)i);
        if (isnan(s) || isinf(s)) continue;
        float ref = s / f;
        float q = s * r;
        float e = fmaf(-q, f, s);
        q = fmaf(e, r, q);
        if (fabsf(s) < 0x1p-100f) q = s / f;
        /* a zero quotient of either sign gives the same update */
        if (f2u(q) != f2u(ref) && !(q == 0.0f && ref == 0.0f)) bad++;
    }
    return bad;
}

int main(int argc, char** argv) {
    int full = argc > 1 && strcmp(argv[1], "full") == 0;
    int quick = argc > 1 && strcmp(argv[1], "quick") == 0;
    static const int exps_all[] = {-126, -101, -100, -99, -98, -60, -20, -3, -2, -1, 0, 1, 2, 20, 60, 100};
    static const int exps_quick[] = {-101, -100, -99, -98, -1, 0, 1, 100};
    const int* exps = quick ? exps_quick : exps_all;
    const unsigned n_exps = quick ? sizeof(exps_quick) / sizeof(int) : sizeof(exps_all) / sizeof(int);
    unsigned ru, fu;
    long long total = 0;
    int pairs = 0;
    while (scanf("%x %x", &ru, &fu) == 2) {
        const float r = u2f(ru), f = u2f(fu);
        long long bad = 0;
        if (full) {
            const uint64_t top = (uint64_t)(101 + 127) << 23; /* |s| < 2^101 */
            bad = check_range(f, r, 0, top) + check_range(f, r, 1ull << 31, (1ull << 31) + top);
        } else {
            for (int sign = 0; sign < 2; sign++) {
                uint64_t sb = (uint64_t)sign << 31;
                bad += check_range(f, r, sb, sb + (1u << 23)); /* subnormals and zero */
                for (unsigned k = 0; k < n_exps; k++) {
                    uint64_t base = sb | ((uint64_t)(exps[k] + 127) << 23);
                    bad += check_range(f, r, base, base + (1u << 23));
                }
            }
        }
        printf("f=%a mismatches=%lld\n", (double)f, bad);
        total += bad;
        pairs++;
    }
    printf("pairs=%d\n", pairs);
    return total ? 1 : 0;
}
