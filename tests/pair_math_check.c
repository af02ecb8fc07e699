/* CPU model check of the short correctly rounded pair arithmetic of alloy_amd/csrc/dfsph_device.h (sqrt_rn, div_h,
 * div_rn): the same float sequences with C's fmaf, the hardware estimates (v_sqrt_f32, v_rcp_f32, each within 1 ulp)
 * stood in for by the correctly rounded value and both of its neighbours, every result compared with IEEE sqrtf and
 * division. Exhaustive over q = r / h in [1/16, 2) for the smoothing lengths of the benchmark scenes and the range
 * ends, sampled down to the 1e-9 h cut-off and over random h in [1e-5, 100]; sqrt exhaustive over [1, 4) (every
 * significand of both exponent parities) and sampled over [2^-96, 2^60); division sampled over the quotient range
 * and exhaustive over the numerator significands for four divisors. Exit status 0 when nothing differs. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static inline float fa(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t ua(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
/* sqrt_rn from an estimate s0 */
static inline float sqrt_fix(float x, float s0) {
    float dn = fa(ua(s0) - 1u), up = fa(ua(s0) + 1u);
    float vp = fmaf(-dn, s0, x), vs = fmaf(-up, s0, x);
    float s = vp <= 0.0f ? dn : s0;
    return vs > 0.0f ? up : s;
}
static inline float div_h(float r, float h, float y) {
    float q0 = r * y;
    float e0 = fmaf(-h, q0, r);
    float q1 = fmaf(e0, y, q0);
    float e1 = fmaf(-h, q1, r);
    return fmaf(e1, y, q1);
}
static inline float div_g(float a, float d, float y0) {
    float e = fmaf(-d, y0, 1.0f);
    float y = fmaf(e, y0, y0);
    float q0 = a * y;
    float r0 = fmaf(-d, q0, a);
    float q1 = fmaf(r0, y, q0);
    float r1 = fmaf(-d, q1, a);
    return fmaf(r1, y, q1);
}
static uint64_t rs = 88172645463325252ull;
static inline uint64_t rnd(void) { rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17; return rs; }
int main(int argc, char** argv) {
    long bad = 0, n = 0;
    /* 1. division by h, exhaustive over q in [2^-4, 2) for a few h, sampled below */
    float hs[] = {0.004f, 0.008f, 0.0026666667f, 0.02f, 1.0e-5f, 100.0f, 0.0123457f};
    for (int k = 0; k < (int)(sizeof(hs) / sizeof(hs[0])); ++k) {
        float h = hs[k], y = 1.0f / h;
        float lo = h * 0.0625f, hi = h * 2.0f;
        for (uint32_t u = ua(lo); u <= ua(hi); ++u) {
            float r = fa(u);
            ++n;
            if (div_h(r, h, y) != r / h) { if (bad++ < 5) printf("div_h h=%a r=%a\n", h, r); }
        }
        for (int t = 0; t < 4000000; ++t) {
            float r = fa(ua(1.0e-9f * h) + (uint32_t)(rnd() % (ua(lo) - ua(1.0e-9f * h))));
            ++n;
            if (div_h(r, h, y) != r / h) { if (bad++ < 5) printf("div_h h=%a r=%a\n", h, r); }
        }
    }
    printf("div_h: %ld checked, %ld wrong\n", n, bad);
    /* random h in [1e-5, 100] */
    long b2 = 0, n2 = 0;
    for (int t = 0; t < 20000; ++t) {
        float h = fa(ua(1.0e-5f) + (uint32_t)(rnd() % (ua(100.0f) - ua(1.0e-5f)))), y = 1.0f / h;
        for (int s = 0; s < 2000; ++s) {
            float r = fa(ua(1.0e-9f * h) + (uint32_t)(rnd() % (ua(2.0f * h) - ua(1.0e-9f * h))));
            ++n2;
            if (div_h(r, h, y) != r / h) { if (b2++ < 5) printf("div_h h=%a r=%a\n", h, r); }
        }
    }
    printf("div_h random h: %ld checked, %ld wrong\n", n2, b2);
    /* 2. sqrt fix-up from the three estimates RN-1, RN, RN+1: exhaustive over [1, 4), sampled over [2^-96, 2^60) */
    long b3 = 0, n3 = 0;
    for (uint32_t u = ua(1.0f); u < ua(4.0f); ++u) {
        float x = fa(u), s = sqrtf(x);
        for (int d = -1; d <= 1; ++d) {
            ++n3;
            if (sqrt_fix(x, fa(ua(s) + d)) != s) { if (b3++ < 5) printf("sqrt x=%a d=%d\n", x, d); }
        }
    }
    for (int t = 0; t < 30000000; ++t) {
        float x = fa(ua(0x1p-96f) + (uint32_t)(rnd() % (ua(0x1p60f) - ua(0x1p-96f)))), s = sqrtf(x);
        for (int d = -1; d <= 1; ++d) {
            ++n3;
            if (sqrt_fix(x, fa(ua(s) + d)) != s) { if (b3++ < 5) printf("sqrt x=%a d=%d\n", x, d); }
        }
    }
    { float x = 0.0f; printf("sqrt(0) from 0: %a\n", sqrt_fix(x, 0.0f)); }
    printf("sqrt: %ld checked, %ld wrong\n", n3, b3);
    /* 3. general division from the rcp estimates RN(1/d) -1, 0, +1 ulp */
    long b4 = 0, n4 = 0;
    for (int t = 0; t < 30000000; ++t) {
        float a = fa(ua(0x1p-100f) + (uint32_t)(rnd() % (ua(0x1p100f) - ua(0x1p-100f))));
        float d = fa(ua(0x1p-70f) + (uint32_t)(rnd() % (ua(0x1p40f) - ua(0x1p-70f))));
        if (t & 1) a = -a;
        float qd = a / d;
        if (!(fabsf(qd) >= 0x1p-100f && fabsf(qd) <= 0x1p100f)) continue;
        float y = 1.0f / d;
        for (int e = -1; e <= 1; ++e) {
            ++n4;
            if (div_g(a, d, fa(ua(y) + e)) != qd) { if (b4++ < 5) printf("div a=%a d=%a e=%d\n", a, d, e); }
        }
    }
    /* exhaustive numerators for a few denominators */
    float ds[] = {0.004f * 0.001f, 3.3e-6f, 1.0f, 0.7f};
    for (int k = 0; k < 4; ++k) {
        float d = ds[k], y = 1.0f / d;
        for (uint32_t u = ua(1.0f); u < ua(2.0f); ++u) {
            float a = fa(u);
            for (int e = -1; e <= 1; ++e) {
                ++n4;
                if (div_g(a, d, fa(ua(y) + e)) != a / d) { if (b4++ < 5) printf("div a=%a d=%a e=%d\n", a, d, e); }
            }
        }
    }
    printf("div: %ld checked, %ld wrong\n", n4, b4);
    return (bad || b2 || b3 || b4) ? 1 : 0;
}
