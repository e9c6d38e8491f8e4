// expf_check.hip -- test support: compares the cap model's single-precision
// exp (cap_expf, mcmtop_amd/csrc/mcm_cap.h) with the C library's expf, the
// routine the reference's flang build calls, bit for bit over every float
// bit pattern: its host form on the CPU (expf_check_host) and its device
// form on the GPU (expf_check_device).  Built by tests/gpu_checks/Makefile
// into libexpf_check.so; called by tests/test_oracle_closed_form.py (host)
// and tests/test_gpu_quotients.py (device).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../../mcmtop_amd/csrc/mcm_physics.h"

namespace {
bool same(float a, float b) {   // equal bits, or both NaN
    uint32_t ua, ub;
    memcpy(&ua, &a, 4);
    memcpy(&ub, &b, 4);
    return ua == ub || (a != a && b != b);
}

float from_bits(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// checks patterns [lo, lo + n) on nthreads host threads against got (or the
// host cap_expf when got is null); returns the mismatch count and the first
// mismatching pattern in *first
unsigned long long compare(uint64_t lo, uint64_t n, const float* got, int nthreads, uint32_t* first) {
    std::atomic<unsigned long long> bad{0};
    std::atomic<uint32_t> firstbad{0xffffffffu};
    std::vector<std::thread> pool;
    const uint64_t per = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++)
        pool.emplace_back([&, t] {
            const uint64_t a = t * per, b = std::min(n, a + per);
            for (uint64_t k = a; k < b; k++) {
                const float x = from_bits((uint32_t)(lo + k));
                const float g = got ? got[k] : cap_expf(x);
                if (!same(g, expf(x))) {
                    if (bad.fetch_add(1) == 0) firstbad = (uint32_t)(lo + k);
                }
            }
        });
    for (auto& th : pool) th.join();
    *first = firstbad;
    return bad;
}

__global__ void k_expf(uint64_t lo, uint64_t n, float* out) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        out[k] = cap_expf(__uint_as_float((uint32_t)(lo + k)));
}
}  // namespace

// host cap_expf against expf over patterns [lo, lo + n): mismatches into
// out[0], the first mismatching pattern into out[1]
extern "C" int expf_check_host(uint64_t lo, uint64_t n, int nthreads, unsigned long long* out) {
    if (lo + n > (1ull << 32) || nthreads < 1) return -1;
    uint32_t first = 0;
    out[0] = compare(lo, n, nullptr, nthreads, &first);
    out[1] = first;
    return 0;
}

// device cap_expf against the host's expf over patterns [lo, lo + n), in
// chunks of 2^26 patterns
extern "C" int expf_check_device(uint64_t lo, uint64_t n, int nthreads, unsigned long long* out) {
    if (lo + n > (1ull << 32) || nthreads < 1) return -1;
    const uint64_t chunk = 1ull << 26;
    float* d = nullptr;
    if (hipMalloc(&d, chunk * sizeof(float)) != hipSuccess) return -2;
    std::vector<float> h(chunk);
    out[0] = 0;
    out[1] = 0xffffffffu;
    int rc = 0;
    for (uint64_t at = 0; at < n && rc == 0; at += chunk) {
        const uint64_t m = std::min(chunk, n - at);
        k_expf<<<2048, 256>>>(lo + at, m, d);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpy(h.data(), d, m * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
            rc = -3;
            break;
        }
        uint32_t first = 0;
        const unsigned long long b = compare(lo + at, m, h.data(), nthreads, &first);
        if (b && out[0] == 0) out[1] = first;
        out[0] += b;
    }
    (void)hipFree(d);
    return rc;
}
