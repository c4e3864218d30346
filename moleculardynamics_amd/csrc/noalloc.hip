// noalloc.hip -- mdx_kinetics_noalloc: the kinetics reductions of
// mdx_kinetics (common/DebugCalculations.cpp:46-124) over caller-owned device
// arrays, with the partial sums in a file-scope device array instead of
// memory the library allocates, so that callers which cannot allocate (code
// under HIP graph capture, tight serving loops) can use them.
//
// The array holds 8 partial sums per block of the first and third passes
// (at most kinetics_blocks(n) <= 1024 blocks) and the 8 totals of the second
// pass.  Every block writes its own partials before the next pass reads
// them, so nothing has to be cleared between calls.  Each GPU holds one copy
// for the life of the process; all calls on a GPU share it, so calls that
// overlap on one GPU (different streams, or host threads) race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "mdx_internal.h"

namespace {

constexpr int kMaxBlocks = 1024;
constexpr int kScratchDoubles = 8 * kMaxBlocks + 8;
constexpr int64_t kMaxAtoms = int64_t(1) << 28;

static __device__ __align__(256) double scratch[kScratchDoubles];

// the device address of scratch on each GPU, found once per device (with
// that device current); host code uses these, never the array's name
std::mutex g_mu;
std::vector<double*> g_scratch;
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int scratch_of_current_device(double** out) {
    int dev = 0, ndev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceCount(&ndev) != hipSuccess || dev < 0 || dev >= ndev)
        return fail(MDX_ERR_HIP, "no current HIP device");
    std::lock_guard<std::mutex> lock(g_mu);
    if ((int)g_scratch.size() < ndev) g_scratch.resize(ndev, nullptr);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MDX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return MDX_OK;
}

}  // namespace

extern "C" {

const char* mdx_kinetics_noalloc_last_error(void) { return g_err.c_str(); }

int mdx_kinetics_noalloc(const double* pos3, const double* vel3, const int32_t* type, int32_t n, int64_t stride,
                         const double* mass4, double kboltz, double* out8, void* stream) {
    if (!pos3 || !vel3 || !type || !mass4 || !out8) return fail(MDX_ERR_ARG, "null argument");
    if (n < 1) return fail(MDX_ERR_ARG, "n must be at least 1");
    if ((int64_t)n > kMaxAtoms)
        return fail(MDX_ERR_CAPACITY, "n = " + std::to_string(n) + " exceeds the limit of " +
                                          std::to_string(kMaxAtoms) + " atoms");
    if (stride < n) return fail(MDX_ERR_ARG, "stride must be at least n");
    for (int k = 0; k < mdx::kTypes; k++)
        if (!(mass4[k] > 0) || !std::isfinite(mass4[k])) return fail(MDX_ERR_ARG, "masses must be positive");
    const int32_t nb = mdx::kinetics_blocks(n);
    if (8 * nb + 8 > kScratchDoubles) return fail(MDX_ERR_CAPACITY, "partial sums exceed the scratch array");
    double* part = nullptr;
    int rc = scratch_of_current_device(&part);
    if (rc) return rc;
    mdx::DevState st{};
    for (int d = 0; d < 3; d++) {
        st.r[d] = const_cast<double*>(pos3) + d * stride;
        st.v[d] = const_cast<double*>(vel3) + d * stride;
    }
    st.type = const_cast<int32_t*>(type);
    mdx::MassTable m{};
    for (int k = 0; k < mdx::kTypes; k++) m.mass[k] = mass4[k];
    m.scale = 1.0;
    const hipError_t e =
        mdx::launch_kinetics(static_cast<hipStream_t>(stream), st, n, m, part, nb, out8, kboltz);
    if (e != hipSuccess) return fail(MDX_ERR_HIP, std::string("kinetics: ") + hipGetErrorString(e));
    return MDX_OK;
}

}  // extern "C"
