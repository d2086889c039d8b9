// icm_probe.cpp -- calibration and diagnostic probes of the C ABI:
// the reference's device-copy bandwidth probe and fused-level timings
// (src/gpu/icm_gpu.h:112-116, src/gpu/gpu_api.cu:676-860), and streaming
// rates of this library's own read / write / copy kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/icm_hip.h"
#include "icm_common.hpp"
#include "icm_kernels.hpp"

using namespace icmh;

namespace {

// Device buffers freed on every exit path.
struct Bufs {
    std::vector<void *> v;
    ~Bufs() {
        for (void *x : v) (void)hipFree(x);
    }
    template <class T>
    bool get(T **p, size_t bytes) {
        *p = nullptr;
        if (hipMalloc((void **)p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            set_err("probe: hipMalloc failed");
            return false;
        }
        v.push_back((void *)*p);
        return true;
    }
};

struct Timer {
    hipStream_t s = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    bool ok = false;
    Timer() {
        ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess && hipEventCreate(&a) == hipSuccess &&
             hipEventCreate(&b) == hipSuccess;
    }
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        if (s) (void)hipStreamDestroy(s);
    }
    template <class F>
    double ms(F &&f) {
        if (hipEventRecord(a, s) != hipSuccess) return -1.0;
        if (!f()) return -1.0;
        if (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess) return -1.0;
        float t = 0.f;
        return hipEventElapsedTime(&t, a, b) == hipSuccess ? (double)t : -1.0;
    }
};

double median(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    return v.empty() ? 0.0 : v[v.size() / 2];
}

// Times one LDS-level build and one LDS-level propagate launch over `batch`
// parents at transform size F (children keep F/2 + 1 coefficients, their g
// F/2 entries; parents F + 1 coefficients and F g entries).
int fused_pair_ns(int F, int batch, int quick, double *build_ns, double *corr_ns) {
    if (!build_ns || !corr_ns) return 0;
    *build_ns = *corr_ns = 0.0;
    if (F < 16 || F > icm::kMaxLdsF || (F & (F - 1)) || batch < 1 || batch > (1 << 22)) {
        set_err("fused pair probe: fft_n must be a power of two in 16 .. 4096 and batch in 1 .. 2^22");
        return 0;
    }
    if (!ensure_device()) return 0;
    const int cc = F / 2 + 1, gc = F / 2, cp = F + 1, gp = F;
    const long long nc = 2LL * batch, np = batch;
    Bufs m;
    double *child = nullptr, *parent = nullptr, *gpar = nullptr, *gchild = nullptr;
    double2 *tw = nullptr;
    if (!m.get(&child, sizeof(double) * nc * cc) || !m.get(&parent, sizeof(double) * np * cp) ||
        !m.get(&gpar, sizeof(double) * np * gp) || !m.get(&gchild, sizeof(double) * nc * gc) ||
        !m.get(&tw, sizeof(double2) * 4097))
        return 0;
    std::vector<double2> t(4097);
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < 4096; ++i) {
        const long double ang = -2.0L * PI * (long double)i / 4096.0L;
        t[i] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
    t[4096] = make_double2(1.0, 0.0);
    std::vector<double> hc((size_t)nc * cc), hg((size_t)np * gp);
    for (size_t i = 0; i < hc.size(); ++i) hc[i] = 1.0 / (1.0 + (double)(i % cc));
    for (size_t i = 0; i < hg.size(); ++i) hg[i] = 1.0 / (1.0 + (double)(i % gp));
    if (!HIP_OK(hipMemcpy(tw, t.data(), sizeof(double2) * t.size(), hipMemcpyHostToDevice)) ||
        !HIP_OK(hipMemcpy(child, hc.data(), sizeof(double) * hc.size(), hipMemcpyHostToDevice)) ||
        !HIP_OK(hipMemcpy(gpar, hg.data(), sizeof(double) * hg.size(), hipMemcpyHostToDevice)))
        return 0;
    icm::LevelJob J{};
    J.child = icm::CoefView{child, nc * cc, cc, cc};
    J.parent = icm::CoefOut{parent, np * cp, cp, cp};
    J.gpar = icm::CoefView{gpar, np * gp, gp, gp};
    J.gchild = icm::CoefOut{gchild, nc * gc, gc, gc};
    J.nr_child = (int)nc;
    J.nr_parent = (int)np;
    J.np = (int)np;
    J.plist = nullptr;
    J.F = F;
    J.logF = ilog2i(F);
    J.nq = 1;
    Timer T;
    if (!T.ok) { set_err("probe: stream / event creation failed"); return 0; }
    const int warm = quick ? 1 : 3, reps = quick ? 4 : 12;
    std::vector<double> bs, cs;
    for (int r = 0; r < warm + reps; ++r) {
        const double b = T.ms([&] { return icm::launch_build_lds(J, tw, 12, T.s, nullptr) == hipSuccess; });
        const double c = T.ms([&] { return icm::launch_prop_lds(J, tw, 12, T.s, nullptr) == hipSuccess; });
        if (b < 0 || c < 0) { set_err("fused pair probe: launch failed"); return 0; }
        if (r >= warm) { bs.push_back(b); cs.push_back(c); }
    }
    *build_ns = median(bs) * 1e6 / (double)np;
    *corr_ns = median(cs) * 1e6 / (double)np;
    return 1;
}

}  // namespace

extern "C" {

int icm_gpu_measure_hbm_bandwidth_gbps(double *gbps_out) {
    if (!gbps_out) return 0;
    *gbps_out = 0.0;
    if (!ensure_device()) return 0;
    const size_t bytes = (size_t)1 << 30;
    Bufs m;
    void *a = nullptr, *b = nullptr;
    if (!m.get(&a, bytes) || !m.get(&b, bytes)) return 0;
    Timer T;
    if (!T.ok || hipMemsetAsync(a, 1, bytes, T.s) != hipSuccess || hipMemsetAsync(b, 2, bytes, T.s) != hipSuccess) {
        set_err("hbm probe: setup failed");
        return 0;
    }
    (void)hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, T.s);   // warm-up
    const double ms = T.ms([&] {
        for (int i = 0; i < 16; ++i)
            if (hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, T.s) != hipSuccess) return false;
        return true;
    });
    if (ms <= 0) { set_err("hbm probe: copy failed"); return 0; }
    *gbps_out = (double)bytes * 16.0 / (ms * 1e-3) / 1e9;
    return 1;
}

int icm_gpu_measure_fused_pair_ns(int fft_n, int batch, int quick, double *build_ns_out, double *corr_ns_out) {
    return fused_pair_ns(fft_n, batch, quick, build_ns_out, corr_ns_out);
}

int icm_gpu_measure_fused_r2c_pair_ns(int fft_n, int batch, int quick, double *build_ns_out,
                                      double *corr_ns_out) {
    return fused_pair_ns(fft_n, batch, quick, build_ns_out, corr_ns_out);
}

int icm_hip_measure_stream_gbps(size_t bytes, double *read_gbps, double *write_gbps, double *copy_gbps) {
    if (bytes < 4096) { set_err("stream probe: need at least 4096 bytes per array"); return 0; }
    if (!ensure_device()) return 0;
    const long long n4 = (long long)(bytes / 32);
    const size_t nb = (size_t)n4 * 32;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g_device);
    const int nwg = std::max(1, cus) * 16;
    Bufs m;
    double *a = nullptr, *b = nullptr, *part = nullptr;
    if (!m.get(&a, nb) || !m.get(&b, nb) || !m.get(&part, sizeof(double) * nwg)) return 0;
    Timer T;
    if (!T.ok) { set_err("stream probe: stream / event creation failed"); return 0; }
    double out[3] = {0, 0, 0};
    for (int mode : {1, 0, 2}) {   // fill first: the read and copy probes then read written memory
        std::vector<double> t;
        for (int r = 0; r < 6; ++r) {
            const double ms = T.ms([&] { return icm::launch_stream_probe(mode, a, mode == 1 ? a : b, n4, part, nwg,
                                                                        T.s) == hipSuccess; });
            if (ms <= 0) { set_err("stream probe: launch failed"); return 0; }
            if (r > 0) t.push_back(ms);
        }
        const double moved = (mode == 2 ? 2.0 : 1.0) * (double)nb;
        out[mode] = moved / (median(t) * 1e-3) / 1e9;
    }
    if (read_gbps) *read_gbps = out[0];
    if (write_gbps) *write_gbps = out[1];
    if (copy_gbps) *copy_gbps = out[2];
    return 1;
}

int icm_hip_measure_lds_gbps(double *gbps) {
    if (!gbps) return 0;
    *gbps = 0.0;
    if (!ensure_device()) return 0;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g_device);
    const int nwg = std::max(1, cus) * 8, iters = 4096;
    Bufs m;
    double *src = nullptr, *out = nullptr;
    if (!m.get(&src, sizeof(double) * 4096) || !m.get(&out, sizeof(double) * 256 * (size_t)nwg)) return 0;
    Timer T;
    if (!T.ok || hipMemsetAsync(src, 0, sizeof(double) * 4096, T.s) != hipSuccess) {
        set_err("lds probe: setup failed");
        return 0;
    }
    std::vector<double> t;
    for (int r = 0; r < 4; ++r) {
        const double ms = T.ms([&] { return icm::launch_stream_probe(3, src, out, iters, nullptr, nwg, T.s) == hipSuccess; });
        if (ms <= 0) { set_err("lds probe: launch failed"); return 0; }
        if (r > 0) t.push_back(ms);
    }
    *gbps = (double)nwg * 256 * iters * 16 / (median(t) * 1e-3) / 1e9;
    return 1;
}

}  // extern "C"
