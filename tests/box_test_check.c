// CPU restatement of the traversal's box test (wavefront-renderer_amd/csrc/wf_traverse.h: boxes4
// of the four-wide walk; slab of the eight-wide walk applies the same rule to one box) for one
// box, checked against exact arithmetic on the true ray o + t d: the test must accept every box
// whose exact slab interval meets [0, tmax] (a dropped box could drop the closest hit). Random
// rays and boxes, most of them built to graze the ray within a few ulps.
//
//   box_test_check <cases> <mode> <seed>
//   mode 0: the device rule (entry planes at fma(x, inv, -next_up(o inv)), exit planes at
//           fma(x, inv, -next_down(o inv)), exit distance times 1 + 2^-21)
//   mode 1: the device rule without the exit-side factor
//   mode 2: the device rule with o inv rounded to nearest on both sides (no outward step)
//   mode 3: planes at (x - o) inv, exit distance times 1 + 2^-21 (drops the boxes whose face
//           holds the origin of a ray with a zero direction component)
// Prints "<cases> <exact accepts> <test accepts> <misses>".
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

// wf_traverse.h next_up: the float next to f towards +inf
static float next_up(float f) {
    const uint32_t b = to_bits(f);
    if ((b & 0x7FFFFFFFu) == 0u) return from_bits(1u);
    return from_bits((int32_t)b >= 0 ? b + 1u : b - 1u);
}

static uint64_t state;
static uint64_t next_u64(void) {
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return state;
}
static float uniform(float a, float b) {
    return a + (b - a) * (float)((next_u64() >> 11) * (1.0 / 9007199254740992.0));
}

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    const long cases = atol(argv[1]);
    const int mode = atoi(argv[2]);
    state = 0x9E3779B97F4A7C15ull ^ (uint64_t)atoll(argv[3]);
    const float slack = mode == 1 ? 1.0f : 1.00000048f;
    const int slab = mode == 3;
    long exact_accepts = 0, test_accepts = 0, misses = 0;
    const float scales[4] = {1000.0f, 1.0f, 0.01f, 100.0f};
    for (long it = 0; it < cases; ++it) {
        // the origin's and the box's distances on independent scales, so that o inv can be
        // far larger than the distances compared (the cancellation case)
        const float so = scales[it % 4], st = scales[(it / 4) % 4];
        float o[3], d[3], inv[3], lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            o[a] = uniform(-so, so);
            d[a] = uniform(-1.0f, 1.0f);
            if (next_u64() % 16 == 0) d[a] = uniform(-1e-6f, 1e-6f);
            if (next_u64() % 64 == 0) d[a] = 0.0f;
            // safe_rcp: |d| clamped to 1e-20, correctly rounded reciprocal
            const float x = fabsf(d[a]) < 1e-20f ? copysignf(1e-20f, d[a]) : d[a];
            inv[a] = 1.0f / x;
        }
        const float t0 = uniform(0.0f, 3.0f * st);
        for (int a = 0; a < 3; ++a) {
            const float X = o[a] + t0 * d[a];
            const float ulp = fabsf(X) * 1.2e-7f + 1e-30f;
            const float w = next_u64() % 3 == 0 ? 0.0f : uniform(0.0f, st * 0.01f);
            const float j = uniform(-4.0f, 4.0f) * ulp;
            switch (next_u64() % 3) {
                case 0: lo[a] = X + j; hi[a] = lo[a] + w; break;
                case 1: hi[a] = X + j; lo[a] = hi[a] - w; break;
                default: lo[a] = X - w; hi[a] = X + w; break;
            }
        }
        const float tmax = next_u64() % 2 ? 1e9f : uniform(0.0f, 3.0f * st);
        __float128 en = 0, ef = tmax;
        float a_c = 0.0f, b_c = tmax;
        for (int a = 0; a < 3; ++a) {
            const int neg = inv[a] < 0.0f;
            const float xn = neg ? hi[a] : lo[a], xf = neg ? lo[a] : hi[a];
            if (d[a] == 0.0f) {
                if (!(o[a] >= lo[a] && o[a] <= hi[a])) { en = 1; ef = 0; }
            } else {
                const __float128 tn = ((__float128)xn - (__float128)o[a]) / (__float128)d[a];
                const __float128 tf = ((__float128)xf - (__float128)o[a]) / (__float128)d[a];
                if (tn > en) en = tn;
                if (tf < ef) ef = tf;
            }
            if (slab) {
                a_c = fmaxf(a_c, (xn - o[a]) * inv[a]);
                b_c = fminf(b_c, (xf - o[a]) * inv[a]);
                continue;
            }
            const float p = o[a] * inv[a];
            const float nup = mode == 2 ? -p : -next_up(p);
            const float ndn = mode == 2 ? -p : next_up(-p);
            a_c = fmaxf(a_c, fmaf(xn, inv[a], nup));
            b_c = fminf(b_c, fmaf(xf, inv[a], ndn));
        }
        const int ex = en <= ef;
        const int ok = a_c <= b_c * slack;
        exact_accepts += ex;
        test_accepts += ok;
        misses += ex && !ok;
    }
    printf("%ld %ld %ld %ld\n", cases, exact_accepts, test_accepts, misses);
    return 0;
}
