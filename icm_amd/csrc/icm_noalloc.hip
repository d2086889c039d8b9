// icm_noalloc.hip -- the entry point that allocates nothing.
//
// icm_hip_equity_noalloc computes what icm_hip_equity does for instances that
// fit the one-launch small-instance kernel (k_small), with every device
// buffer it needs -- stacks, payout, quadrature nodes, the per-workgroup
// partial equities and the result -- carved from `scratch`, a file-scope
// device array of fixed size (8.5 MB, held by each GPU the library runs on
// until the process ends).  No hipMalloc, hipFree or stream creation happens
// on this path; it runs on the default stream and returns after the result
// has been copied back.  Every call on one GPU shares that GPU's array:
// calls that overlap on one GPU race.  Nothing in the array needs to start
// at zero (k_small zeroes its partial rows; the column sum overwrites the
// result).
//
// Players keep the caller's order (the small kernel has no leaf division,
// so the stack-descending sort of the multi-launch path only changes
// rounding).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>

#include "../../include/icm_hip.h"
#include "icm_common.hpp"
#include "icm_kernels.hpp"

using namespace icmh;

namespace {

constexpr int kNoAllocMaxN = 4096;     // players
constexpr int kNoAllocMaxQ = 2048;     // quadrature nodes
constexpr int kNoAllocMaxG = 256;      // workgroups (rows of the partial sums)
// scratch layout (doubles)
constexpr long long kOffS = 0;
constexpr long long kOffPay = kOffS + kNoAllocMaxN;
constexpr long long kOffNodes = kOffPay + kNoAllocMaxN;            // logv [MaxQ], then w [MaxQ]
constexpr long long kOffOut = kOffNodes + 2LL * kNoAllocMaxQ;
constexpr long long kOffPartial = kOffOut + kNoAllocMaxN;
constexpr long long kScratchDoubles = kOffPartial + (long long)kNoAllocMaxG * kNoAllocMaxN;

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up once per
// device, with that device current).
double *scratch_ptr(int dev) {
    if (dev < 0 || dev >= kMaxDevices) { set_err("device id out of range for the no-allocation entry"); return nullptr; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            hip_ok(e, "hipGetSymbolAddress(scratch)");
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int icm_hip_equity_noalloc(int n, const double *S, int Q, const double *payout, int k, double *equity) {
    if (!S || !payout || !equity || n < 1 || k < 1 || k > n || Q < 1) {
        set_err("invalid arguments: need 1 <= k <= n, Q >= 1 and non-null arrays");
        return -1;
    }
    if (n > kNoAllocMaxN || Q > kNoAllocMaxQ) {
        char buf[160];
        snprintf(buf, sizeof buf, "icm_hip_equity_noalloc takes n <= %d and Q <= %d (got n=%d, Q=%d)", kNoAllocMaxN,
                 kNoAllocMaxQ, n, Q);
        set_err(buf);
        return -1;
    }
    icm::SmallJob J{};
    if (small_geometry(n, k, J) > icm::kSmallLdsB) {
        char buf[160];
        snprintf(buf, sizeof buf, "icm_hip_equity_noalloc: n=%d k=%d needs %d doubles of LDS, more than %d", n, k,
                 J.lds_doubles, icm::kSmallLdsB);
        set_err(buf);
        return -1;
    }
    double smax = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!(S[i] > 0.0) || !std::isfinite(S[i])) { set_err("stacks must be positive and finite"); return -1; }
        smax = std::max(smax, S[i]);
    }
    if (!ensure_device()) return -1;
    double *base = scratch_ptr(g_device);
    if (!base) return -1;
    // nodes on the host stack: logv then w (src/cpu/icm.c:117-132), compacted
    // to the active ones (w != 0, src/cpu/icm.c:2558)
    double nodes[2 * kNoAllocMaxQ];
    make_nodes(Q, smax, nodes, nodes + kNoAllocMaxQ);
    int Qa = 0;
    for (int q = 0; q < Q; ++q) {
        if (nodes[kNoAllocMaxQ + q] == 0.0) continue;
        nodes[Qa] = nodes[q];
        nodes[kNoAllocMaxQ + Qa] = nodes[kNoAllocMaxQ + q];
        ++Qa;
    }
    hipStream_t s = 0;
    if (!HIP_OK(hipMemcpyAsync(base + kOffS, S, sizeof(double) * n, hipMemcpyHostToDevice, s)) ||
        !HIP_OK(hipMemcpyAsync(base + kOffPay, payout, sizeof(double) * k, hipMemcpyHostToDevice, s)) ||
        !HIP_OK(hipMemcpyAsync(base + kOffNodes, nodes, sizeof(double) * 2 * kNoAllocMaxQ, hipMemcpyHostToDevice, s)))
        return -1;
    J.S = base + kOffS;
    J.logv = base + kOffNodes;
    J.w = base + kOffNodes + kNoAllocMaxQ;
    J.payout = base + kOffPay;
    J.partial = base + kOffPartial;
    J.Qa = Qa;
    int cus = 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g_device) != hipSuccess) cus = 256;
    const int G = std::max(1, std::min({Qa, kNoAllocMaxG, std::max(1, cus)}));
    double *out = base + kOffOut;
    if (!hip_ok(icm::launch_small(J, G, s, nullptr, 0.0), "small-instance kernel") ||
        !hip_ok(icm::launch_small_finalize(J.partial, G, n, nullptr, nullptr, out, nullptr, s, nullptr),
                "small-instance finalize"))
        return -1;
    if (!HIP_OK(hipMemcpyAsync(equity, out, sizeof(double) * n, hipMemcpyDeviceToHost, s)) ||
        !HIP_OK(hipStreamSynchronize(s)))
        return -1;
    return 0;
}

}  // extern "C"
