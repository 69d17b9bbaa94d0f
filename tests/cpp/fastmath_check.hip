// Bitwise check of the guarded FP64 division and square root
// (titan_amd/csrc/fastmath.h) against the IEEE `/` and `sqrt` on gfx950
// (quotients: equal bits, or both zero -- a zero numerator gives +0).
//
//   fastmath_check [n_per_case]      exit 0 iff every result matches bit for bit
//
// Inputs come from a counter-based generator (splitmix64): random exponents
// and mantissas inside each routine's stated range and its boundary values;
// for unit_axis, spring axes whose components span tiny, ordinary and huge
// magnitudes, exact zeros and -0, so both the fast path and the fallback run.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "fastmath.h"

using namespace titan_mi355x;

#define CHK(x)                                                                    \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
            return 2;                                                             \
        }                                                                         \
    } while (0)

__device__ __forceinline__ uint64_t mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// A double with a uniformly random mantissa and an exponent uniform in [e0, e1].
__device__ __forceinline__ double rnd(uint64_t h, int e0, int e1) {
    const int e = e0 + (int)((h >> 52) % (uint64_t)(e1 - e0 + 1));
    const uint64_t bits = ((uint64_t)(e + 1023) << 52) | (h & 0xFFFFFFFFFFFFFull);
    return __longlong_as_double((long long)bits);
}

__device__ __forceinline__ bool same(double a, double b) {
    return __double_as_longlong(a) == __double_as_longlong(b);
}

// equal bits, or both zero: a zero numerator's quotient is +0 on the fast path
__device__ __forceinline__ bool same_q(double a, double b) { return same(a, b) || (a == 0.0 && b == 0.0); }

struct Counts {
    unsigned long long sqrt_bad, div_bad, axis_bad, axis_fast;
};

__global__ void k_check(uint64_t n, uint64_t seed, Counts *c, double *first_bad) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t h0 = mix(seed ^ (i * 4 + 0)), h1 = mix(seed ^ (i * 4 + 1));
    const uint64_t h2 = mix(seed ^ (i * 4 + 2)), h3 = mix(seed ^ (i * 4 + 3));

    // sqrt on [2^-767, 2^1000]; the first lanes take the end points
    double x = rnd(h0, -767, 999);
    if (i == 0) x = 0x1p-767;
    if (i == 1) x = 0x1p1000;
    if (i == 2) x = 1.0;
    if (!same(fast_sqrt(x), sqrt(x))) {
        atomicAdd(&c->sqrt_bad, 1ull);
        first_bad[0] = x;
    }

    // a / b, b in [2^-500, 2^500], a == +-0 or |a| in [2^-500, 2b]
    const double b = rnd(h1, -500, 499);
    double a;
    const int sel = (int)(h2 & 7);
    if (sel == 0) a = 0.0;
    else if (sel == 1) a = -0.0;
    else if (sel == 2) a = b;  // exact quotient 1
    else {
        // |a| with an exponent at most one above b's
        int eb = (int)((__double_as_longlong(b) >> 52) & 0x7FF) - 1023;
        int lo = eb - 60 < -500 ? -500 : eb - 60;
        a = rnd(h3, lo, eb);
        if (sel == 3) a = fmin(2.0 * b, a * 2.0);
    }
    if (h2 & 8) a = -a;
    const double y = fast_recip(b);
    if (!same_q(fast_div(a, b, y), a / b)) {
        atomicAdd(&c->div_bad, 1ull);
        first_bad[1] = a;
        first_bad[2] = b;
    }

    // unit_axis over all magnitudes
    double t[3];
    for (int q = 0; q < 3; ++q) {
        const uint64_t hq = mix(h3 ^ (uint64_t)(q + 17));
        const int k = (int)(hq & 15);
        if (k == 0) t[q] = 0.0;
        else if (k == 1) t[q] = -0.0;
        else if (k == 2) t[q] = rnd(hq, -700, -480);    // tiny: fallback
        else if (k == 3) t[q] = rnd(hq, 480, 520);      // huge: fallback
        else t[q] = rnd(hq, -30, 10);
        if (hq & 16) t[q] = -t[q];
    }
    double nrm, ux, uy, uz;
    unit_axis(t[0], t[1], t[2], nrm, ux, uy, uz);
    const double rn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    if (!same(nrm, rn) || !same_q(ux, t[0] / rn) || !same_q(uy, t[1] / rn) || !same_q(uz, t[2] / rn)) {
        atomicAdd(&c->axis_bad, 1ull);
        first_bad[3] = t[0];
        first_bad[4] = t[1];
        first_bad[5] = t[2];
    }
    if (fast_range(t[0] * t[0] + t[1] * t[1] + t[2] * t[2], t[0], t[1], t[2])) atomicAdd(&c->axis_fast, 1ull);
}

int main(int argc, char **argv) {
    const uint64_t n = argc > 1 ? strtoull(argv[1], nullptr, 10) : (1ull << 26);
    Counts *c;
    double *fb;
    CHK(hipMalloc(&c, sizeof(Counts)));
    CHK(hipMalloc(&fb, 8 * sizeof(double)));
    CHK(hipMemset(c, 0, sizeof(Counts)));
    CHK(hipMemset(fb, 0, 8 * sizeof(double)));
    const unsigned blocks = (unsigned)((n + 255) / 256);
    printf("fastmath_check: %llu inputs per case\n", (unsigned long long)n);
    fflush(stdout);
    hipLaunchKernelGGL(k_check, dim3(blocks), dim3(256), 0, 0, n, 0x5EEDull, c, fb);
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    Counts h;
    double hb[8];
    CHK(hipMemcpy(&h, c, sizeof h, hipMemcpyDeviceToHost));
    CHK(hipMemcpy(hb, fb, sizeof hb, hipMemcpyDeviceToHost));
    printf("sqrt mismatches %llu\ndiv mismatches %llu\nunit_axis mismatches %llu (fast path %llu of %llu)\n",
           h.sqrt_bad, h.div_bad, h.axis_bad, h.axis_fast, (unsigned long long)n);
    if (h.sqrt_bad) printf("  sqrt x=%a\n", hb[0]);
    if (h.div_bad) printf("  div a=%a b=%a\n", hb[1], hb[2]);
    if (h.axis_bad) printf("  axis t=(%a, %a, %a)\n", hb[3], hb[4], hb[5]);
    CHK(hipFree(c));
    CHK(hipFree(fb));
    const bool ok = !h.sqrt_bad && !h.div_bad && !h.axis_bad && h.axis_fast > 0 && h.axis_fast < n;
    printf(ok ? "PASS\n" : "FAIL\n");
    return ok ? 0 : 1;
}
