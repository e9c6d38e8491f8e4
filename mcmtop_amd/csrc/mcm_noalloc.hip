// mcm_noalloc.hip -- the energy totals of mcm_calctotale over particle arrays
// a caller holds, with no device allocation: for callers that cannot call
// hipMalloc (code under HIP graph capture, serving loops).  Included by
// mcm_gpu.hip after the step kernels; it shares k_calctotale and k_c_fold
// with mcm_gpu_calctotale, so both give the same sums bit for bit.
//
// Scratch: the block partials and the scalar block of one call live in the
// file-scope device array `scratch`, of fixed size, which every GPU the
// library runs on holds until the process ends.  Its device address is
// looked up once per device (hipGetSymbolAddress) and used for every launch.
// Every call on one GPU shares that array, so calls that overlap on one GPU
// race; each call clears the scalar block it reads before use.

#include <mutex>

// MCM_NOALLOC_MAX_NP (include/mcm_gpu.h): 16384 blocks of block partials
namespace {
constexpr int kNoallocBlocks = MCM_NOALLOC_MAX_NP / MCM_BLOCK;
static_assert(MCM_NOALLOC_MAX_NP % MCM_BLOCK == 0, "whole blocks");
constexpr int kScratchScalars = 256;   // doubles reserved for the Scalars block
static_assert(sizeof(Scalars) <= kScratchScalars * sizeof(double), "scratch scalar block");
constexpr int kMaxDevices = 64;
void* scratch_dev[kMaxDevices] = {};   // device address of `scratch` per device
std::mutex scratch_mu;
}  // namespace

static __device__ __align__(256) double scratch[kScratchScalars + (size_t)kNoallocBlocks * RED_C];

namespace {
// zero of the scratch scalar block (a kernel rather than a memset node, for
// callers that capture the launches in a graph)
__global__ void k_scratch_clear(double* p, int n) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = 0.0;
}

// out[5 m .. 5 m + 4] = ke, ie, mom(3) of material m + 1; then the totals
__global__ void k_totals_out(const Scalars* sc, int nm, double* out) {
    if (threadIdx.x != 0) return;
    for (int m = 0; m < nm; m++) {
        out[5 * m] = sc->mat_ke[m];
        out[5 * m + 1] = sc->mat_ie[m];
        for (int a = 0; a < 3; a++) out[5 * m + 2 + a] = sc->mat_mom[m][a];
    }
    out[5 * nm] = sc->kinetice;
    out[5 * nm + 1] = sc->internale;
    out[5 * nm + 2] = sc->totale;
}

// the device address of `scratch` on the current device, looked up on the
// first call there
int scratch_address(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return MCM_GPU_ERR_NODEVICE;
    if (dev >= kMaxDevices) return MCM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MCM_GPU_ERR_HIP;
        scratch_dev[dev] = p;
    }
    *out = scratch_dev[dev];
    return MCM_GPU_OK;
}
}  // namespace

extern "C" int mcm_gpu_calctotale_noalloc(const double* mass, const double* vabs, const double* e, const double* vx,
                                          const double* vy, const double* vz, const int32_t* mat, int64_t np,
                                          int32_t nummat, double* out, void* stream) {
    if (!mass || !vabs || !e || !vx || !vy || !vz || !mat || !out) return MCM_GPU_ERR_ARG;
    if (np < 0 || np > MCM_NOALLOC_MAX_NP || nummat < 1 || nummat > MCM_MAX_MAT) return MCM_GPU_ERR_ARG;
    void* base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    DevState d{};
    d.np = (int)np;
    d.nummat = nummat;
    d.nblk = std::max(1, blocks_for(np, MCM_BLOCK));
    d.mass = const_cast<double*>(mass);
    d.vabs = const_cast<double*>(vabs);
    d.e = const_cast<double*>(e);
    d.v[0] = const_cast<double*>(vx);
    d.v[1] = const_cast<double*>(vy);
    d.v[2] = const_cast<double*>(vz);
    d.pmat = const_cast<int*>(mat);
    d.sc = (Scalars*)base;                            // abort flag and results
    d.red = (double*)base + kScratchScalars;          // block partials
    k_scratch_clear<<<1, MCM_BLOCK, 0, st>>>((double*)base, kScratchScalars);
    k_calctotale<<<d.nblk, MCM_BLOCK, 0, st>>>(d);
    switch (nummat) {
        case 1: k_c_fold<1><<<1, MCM_BLOCK, 0, st>>>(d, 0); break;
        case 2: k_c_fold<2><<<1, MCM_BLOCK, 0, st>>>(d, 0); break;
        case 3: k_c_fold<3><<<1, MCM_BLOCK, 0, st>>>(d, 0); break;
        case 4: k_c_fold<4><<<1, MCM_BLOCK, 0, st>>>(d, 0); break;
        case 5: k_c_fold<5><<<1, MCM_BLOCK, 0, st>>>(d, 0); break;
        case 6: k_c_fold<6><<<1, MCM_BLOCK, 0, st>>>(d, 0); break;
        case 7: k_c_fold<7><<<1, MCM_BLOCK, 0, st>>>(d, 0); break;
        default: k_c_fold<8><<<1, MCM_BLOCK, 0, st>>>(d, 0); break;
    }
    k_totals_out<<<1, 64, 0, st>>>(d.sc, nummat, out);
    return hipGetLastError() == hipSuccess ? MCM_GPU_OK : MCM_GPU_ERR_HIP;
}
