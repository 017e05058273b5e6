// context.cpp -- device allocation, magnetisation upload, kernel timing and
// device metadata for the HIP FDM backend.
#include "context.hpp"

#include <cstring>

namespace fm {

bool ctx_alloc(Context &c, void **ptr, size_t bytes, const char *what) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        c.last_error = std::string("hipMalloc(") + what + ", " + std::to_string(bytes) + " bytes): " +
                       hipGetErrorString(e);
        return false;
    }
    c.allocations.push_back(p);
    c.device_bytes += bytes;
    *ptr = p;
    return true;
}

// Host-to-device copy ordered on the backend's stream and complete on
// return.  The stream is non-blocking (it does not synchronise with the null
// stream), so setup copies go through it rather than a null-stream hipMemcpy.
bool ctx_upload(Context &c, void *dst, const void *src, size_t bytes, const char *what, std::string &err) {
    return hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream), what, err) &&
           hip_ok(hipStreamSynchronize(c.stream), what, err);
}

void ctx_free_all(Context &c) {
    if (c.stream) hipStreamSynchronize(c.stream);
    for (void *p : c.allocations) hipFree(p);
    c.allocations.clear();
    if (c.h_result) hipHostFree(c.h_result);
    c.h_result = nullptr;
    if (c.h_partials) hipHostFree(c.h_partials);
    c.h_partials = nullptr;
    for (auto &pe : c.pending) {
        hipEventDestroy(pe.a);
        hipEventDestroy(pe.b);
    }
    c.pending.clear();
    for (hipEvent_t e : c.event_pool) hipEventDestroy(e);
    c.event_pool.clear();
    if (c.ev_stats) hipEventDestroy(c.ev_stats);
    c.ev_stats = nullptr;
    c.spec_pre = false;
    if (c.stream) hipStreamDestroy(c.stream);
    c.stream = nullptr;
    c.device_bytes = 0;
}

// AoS host array -> SoA device state in the storage precision; inactive
// cells forced to zero (native/backends/fdm/src/context.cu:733-782).
bool ctx_upload_m(Context &c, const void *host, bool host_is_f64) {
    const uint64_t n = c.n;
    auto get = [&](uint64_t i) -> double {
        return host_is_f64 ? static_cast<const double *>(host)[i] : (double)static_cast<const float *>(host)[i];
    };
    bool ok;
    if (c.state_bytes() == 8) {
        std::vector<double> s(3 * n);
        for (uint64_t i = 0; i < n; ++i) {
            const bool act = !c.has_mask || c.mask_host[i] != 0;
            for (int k = 0; k < 3; ++k) s[k * n + i] = act ? get(3 * i + k) : 0.0;
        }
        ok = hip_ok(hipMemcpyAsync(c.m, s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice, c.stream),
                    "hipMemcpy(m)", c.last_error);
        ok = ok && hip_ok(hipStreamSynchronize(c.stream), "hipMemcpy(m)", c.last_error);
    } else {
        std::vector<float> s(3 * n);
        for (uint64_t i = 0; i < n; ++i) {
            const bool act = !c.has_mask || c.mask_host[i] != 0;
            for (int k = 0; k < 3; ++k) s[k * n + i] = act ? (float)get(3 * i + k) : 0.0f;
        }
        ok = hip_ok(hipMemcpyAsync(c.m, s.data(), s.size() * sizeof(float), hipMemcpyHostToDevice, c.stream),
                    "hipMemcpy(m)", c.last_error);
        ok = ok && hip_ok(hipStreamSynchronize(c.stream), "hipMemcpy(m)", c.last_error);
    }
    return ok;
}

bool ctx_query_device_info(Context &c) {
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice", c.last_error)) return false;
    hipDeviceProp_t p;
    if (!hip_ok(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties", c.last_error)) return false;
    std::memset(&c.info, 0, sizeof(c.info));
    // some runtimes leave the marketing name empty; fall back to the ISA name
    const char *nm = p.name[0] ? p.name : p.gcnArchName;
    std::strncpy(c.info.name, nm, sizeof(c.info.name) - 1);
    c.info.compute_capability_major = p.major;
    c.info.compute_capability_minor = p.minor;
    int drv = 0, rt = 0;
    hipDriverGetVersion(&drv);
    hipRuntimeGetVersion(&rt);
    c.info.driver_version = drv;
    c.info.runtime_version = rt;
    c.info_valid = true;
    return true;
}

int timing_entry(Context &c, const char *name) {
    for (size_t i = 0; i < c.timings.size(); ++i)
        if (c.timings[i].name == name) return (int)i;
    c.timings.push_back(TimingEntry{name, 0.0, 0});
    return (int)c.timings.size() - 1;
}

static hipEvent_t take_event(Context &c) {
    if (!c.event_pool.empty()) {
        hipEvent_t e = c.event_pool.back();
        c.event_pool.pop_back();
        return e;
    }
    hipEvent_t e{};
    hipEventCreate(&e);
    return e;
}

void timing_begin(Context &c, int entry, hipEvent_t *a) {
    (void)entry;
    *a = take_event(c);
    hipEventRecord(*a, c.stream);
}

void timing_end(Context &c, int entry, hipEvent_t a) {
    hipEvent_t b = take_event(c);
    hipEventRecord(b, c.stream);
    c.pending.push_back(PendingEvent{entry, a, b});
}

void timing_resolve(Context &c) {
    for (auto &pe : c.pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(pe.b) == hipSuccess && hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
            c.timings[pe.entry].total_ms += ms;
            c.timings[pe.entry].launches += 1;
        }
        c.event_pool.push_back(pe.a);
        c.event_pool.push_back(pe.b);
    }
    c.pending.clear();
}

}  // namespace fm
