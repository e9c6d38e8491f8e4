// quot_check.hip -- test support: compares, bit for bit, the unchecked
// arithmetic of the pair loops' fast path (mcmtop_amd/csrc/mcm_physics.h:
// Divisor::quot, sqrt_noscale) with the device's own division and square
// root over the operand ranges that path relies on.  Built by
// tests/gpu_checks/Makefile into libquot_check.so; called by
// tests/test_gpu_quotients.py.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../mcmtop_amd/csrc/mcm_physics.h"

namespace {
__device__ __forceinline__ uint64_t mix(uint64_t x) {   // splitmix64 finaliser
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// a double with exponent uniform in [elo, ehi] and a random significand
__device__ __forceinline__ double draw(uint64_t r, uint64_t s, int elo, int ehi) {
    const int e = elo + (int)(s % (uint64_t)(ehi - elo + 1));
    const uint64_t bits = ((uint64_t)(e + 1023) << 52) | (r & 0xfffffffffffffull);
    return __longlong_as_double((long long)bits);
}

// which: 0 numerators of the corrected quotients, {0} u +-[2^-502, 2^335],
// over sw^2 in [2^-200, 2^200]; 1 numerators and divisors in Divisor's
// range [2^-300, 2^300] (and +-0 numerators); 2 reciprocals of [2^-300,
// 2^300]; 3 square roots of 0 and [2^-767, 2^1023]
__global__ void k_check(int which, long n, uint64_t seed, unsigned long long* bad, double* first) {
#if defined(__HIP_DEVICE_COMPILE__)
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const uint64_t r0 = mix(seed ^ (uint64_t)t * 4), r1 = mix(seed ^ ((uint64_t)t * 4 + 1));
        const uint64_t r2 = mix(seed ^ ((uint64_t)t * 4 + 2)), r3 = mix(seed ^ ((uint64_t)t * 4 + 3));
        double a = 0.0, b = 1.0, got, want;
        if (which == 3) {
            const double x = (r3 & 63) == 0 ? 0.0 : draw(r0, r1, -767, 1023);
            got = sqrt_noscale(x);
            want = sqrt(x);
            a = x;
        } else {
            if (which == 0) {
                a = draw(r0, r1, -502, 335);
                b = draw(r2, r3 >> 8, -200, 200);
            } else if (which == 1) {
                a = draw(r0, r1, -300, 300);
                b = draw(r2, r3 >> 8, -300, 300);
            } else {
                a = 1.0;
                b = draw(r2, r3 >> 8, -300, 300);
            }
            if (which != 2 && (r3 & 15) == 0) a = 0.0;
            if (which != 2 && (r3 & 16)) a = -a;
            got = Divisor(b).quot(a);
            want = a / b;
        }
        if (__double_as_longlong(got) != __double_as_longlong(want)) {
            if (atomicAdd(bad, 1ull) == 0) {
                first[0] = a;
                first[1] = b;
                first[2] = got;
                first[3] = want;
            }
        }
    }
#endif
}
}  // namespace

// Runs n samples of case `which`; out: [0] mismatches, [1..4] the first
// mismatch (a, b, got, want) as doubles.  Returns 0, or -1 on a HIP error.
extern "C" int quot_check(int which, long n, unsigned long long seed, double* out) {
    unsigned long long* bad = nullptr;
    double* first = nullptr;
    if (hipMalloc(&bad, sizeof(unsigned long long)) != hipSuccess) return -1;
    if (hipMalloc(&first, 4 * sizeof(double)) != hipSuccess) return -1;
    if (hipMemset(bad, 0, sizeof(unsigned long long)) != hipSuccess) return -1;
    if (hipMemset(first, 0, 4 * sizeof(double)) != hipSuccess) return -1;
    k_check<<<4096, 256>>>(which, n, seed, bad, first);
    unsigned long long nb = 0;
    double f[4];
    int rc = hipDeviceSynchronize() == hipSuccess ? 0 : -1;
    if (!rc && hipMemcpy(&nb, bad, sizeof(nb), hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    if (!rc && hipMemcpy(f, first, sizeof(f), hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    (void)hipFree(bad);
    (void)hipFree(first);
    if (rc) return rc;
    out[0] = (double)nb;
    for (int k = 0; k < 4; k++) out[1 + k] = f[k];
    return 0;
}
