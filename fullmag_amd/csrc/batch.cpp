// batch.cpp -- many independent backends stepped as one hipGraph, and the
// launch-schedule autotuner (extensions; the reference runs one problem per
// process and has neither).
//
// A batch captures, for every member backend, `k` consecutive fixed-dt Heun
// steps enqueued on that backend's own stream.  The member streams fork from
// and join back into one origin stream during capture, so the graph holds
// one independent branch per member; a single hipGraphLaunch then advances
// every member by k steps, and the runtime is free to run the branches
// concurrently.  Kernel arguments (buffer pointers, dt) are baked in at
// capture; the Heun path swaps no buffers, so replays are exact.  The graph
// holds no memset nodes.
#include "context.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "fullmag_fdm_ext.h"

namespace fm {
Context *backend_context(fullmag_fdm_backend *h);
}

struct fullmag_fdm_batch {
    std::vector<fm::Context *> members;
    std::vector<double> dts;
    int k = 1;
    hipStream_t origin = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;
    std::string last_error;
};

namespace {

bool fail(fullmag_fdm_batch *b, const std::string &what) {
    b->last_error = what;
    return false;
}

bool capture(fullmag_fdm_batch *b) {
    using fm::hip_ok;
    std::string &err = b->last_error;
    if (!hip_ok(hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture", err))
        return false;
    hipEvent_t fork = b->events[0];
    bool ok = hip_ok(hipEventRecord(fork, b->origin), "hipEventRecord(fork)", err);
    for (size_t i = 0; ok && i < b->members.size(); ++i) {
        fm::Context &c = *b->members[i];
        ok = hip_ok(hipStreamWaitEvent(c.stream, fork, 0), "hipStreamWaitEvent(member)", err);
        if (!ok) break;
        fm::integrator_enqueue_heun(c, b->dts[i], b->k);
        ok = hip_ok(hipEventRecord(b->events[1 + i], c.stream), "hipEventRecord(join)", err) &&
             hip_ok(hipStreamWaitEvent(b->origin, b->events[1 + i], 0), "hipStreamWaitEvent(origin)", err);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->origin, &g);
    if (!ok) {
        if (g) hipGraphDestroy(g);
        return false;
    }
    if (!hip_ok(e, "hipStreamEndCapture", err)) return false;
    b->graph = g;
    return hip_ok(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate", err);
}

}  // namespace

extern "C" {

fullmag_fdm_batch *fullmag_fdm_ext_batch_create(fullmag_fdm_backend *const *handles, uint32_t count,
                                                const double *dts, uint32_t steps_per_graph) {
    auto *b = new fullmag_fdm_batch();
    if (!handles || !dts || count == 0 || steps_per_graph == 0) {
        fail(b, "batch_create needs count > 0 handles, their dt and steps_per_graph > 0");
        return b;
    }
    int dev0 = -1;
    for (uint32_t i = 0; i < count; ++i) {
        fm::Context *c = handles[i] ? fm::backend_context(handles[i]) : nullptr;
        if (!c || !c->ready) {
            fail(b, "batch member " + std::to_string(i) + " is not an initialised backend");
            return b;
        }
        if (!(dts[i] > 0.0)) {
            fail(b, "batch member " + std::to_string(i) + " has dt <= 0");
            return b;
        }
        if (std::find(b->members.begin(), b->members.end(), c) != b->members.end()) {
            fail(b, "batch member " + std::to_string(i) + " appears twice");
            return b;
        }
        if (dev0 < 0) dev0 = c->device;
        if (c->device != dev0) {
            fail(b, "batch members must share one device");
            return b;
        }
        if (!fm::integrator_prepare(*c)) {
            fail(b, "batch member " + std::to_string(i) + ": " + c->last_error);
            return b;
        }
        b->members.push_back(c);
        b->dts.push_back(dts[i]);
    }
    b->k = (int)steps_per_graph;
    if (!fm::hip_ok(hipSetDevice(dev0), "hipSetDevice", b->last_error) ||
        !fm::hip_ok(hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking), "hipStreamCreate(batch)",
                    b->last_error))
        return b;
    b->events.resize(count + 1);
    for (auto &e : b->events)
        if (!fm::hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate(batch)", b->last_error))
            return b;
    capture(b);
    return b;
}

int fullmag_fdm_ext_batch_run(fullmag_fdm_batch *b, uint64_t n_steps, fullmag_fdm_step_stats *stats_out) {
    if (!b) return FULLMAG_FDM_ERR_INVALID;
    if (!b->exec) {
        if (b->last_error.empty()) b->last_error = "batch has no graph";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (n_steps == 0) {
        b->last_error = "batch_run needs n_steps > 0";
        return FULLMAG_FDM_ERR_INVALID;
    }
    const uint64_t full = n_steps / (uint64_t)b->k, rem = n_steps % (uint64_t)b->k;
    for (uint64_t s = 0; s < full; ++s)
        if (!fm::hip_ok(hipGraphLaunch(b->exec, b->origin), "hipGraphLaunch", b->last_error))
            return FULLMAG_FDM_ERR_CUDA;
    if (!fm::hip_ok(hipStreamSynchronize(b->origin), "batch graph", b->last_error)) return FULLMAG_FDM_ERR_CUDA;
    int rc = FULLMAG_FDM_OK;
    for (size_t i = 0; i < b->members.size(); ++i) {
        fm::Context &c = *b->members[i];
        if (rem) fm::integrator_enqueue_heun(c, b->dts[i], (int)rem);
        const int r = fm::integrator_finish(c, b->dts[i], n_steps, stats_out ? &stats_out[i] : nullptr);
        if (r != FULLMAG_FDM_OK && rc == FULLMAG_FDM_OK) {
            rc = r;
            b->last_error = "batch member " + std::to_string(i) + ": " + c.last_error;
        }
    }
    return rc;
}

const char *fullmag_fdm_ext_batch_last_error(fullmag_fdm_batch *b) {
    if (!b) return "null handle";
    return b->last_error.empty() ? nullptr : b->last_error.c_str();
}

void fullmag_fdm_ext_batch_destroy(fullmag_fdm_batch *b) {
    if (!b) return;
    if (b->origin) hipStreamSynchronize(b->origin);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (auto e : b->events)
        if (e) hipEventDestroy(e);
    if (b->origin) hipStreamDestroy(b->origin);
    delete b;
}

// Coordinate search over the launch-time knobs of the demag transforms and
// cell kernels, timing (demag + stats) at the current state; keeps the
// fastest schedule and writes it as "key=value,..." into `chosen`.
int fullmag_fdm_ext_autotune(fullmag_fdm_backend *handle, int reps, char *chosen, uint64_t chosen_len) {
    fm::Context *cp = handle ? fm::backend_context(handle) : nullptr;
    if (!cp || !cp->ready) return FULLMAG_FDM_ERR_INVALID;
    fm::Context &c = *cp;
    reps = std::max(1, reps);
    fm::DemagTuning &t = c.dp.tune;
    struct Knob {
        const char *name;
        int *field;
        std::vector<int> values;
    };
    std::vector<Knob> knobs = {{"fixed", &t.fixed, {1, 0}},         {"y_pipe", &t.y_pipe, {4, 0, 1, 2, 8}},
                               {"y_pipe_f", &t.y_pipe_f, {2, 1, 4, 8}},
                               {"z_quad", &t.z_quad, {0, 1}},         {"z_mirror", &t.z_mirror, {0, 1}},
                               {"xcd", &t.xcd, {1, 0}},               {"zc", &t.zc, {0, 1, 2, 4}}};
    double best = fm::integrator_probe_ms(c, reps);
    for (auto &k : knobs) {
        const int keep = *k.field;
        int best_v = keep;
        for (int v : k.values) {
            if (v == keep) continue;
            *k.field = v;
            const double ms = fm::integrator_probe_ms(c, reps);
            if (ms < best * 0.98) {
                best = ms;
                best_v = v;
            }
        }
        *k.field = best_v;
    }
    if (!fm::hip_ok(hipGetLastError(), "autotune", c.last_error)) return FULLMAG_FDM_ERR_CUDA;
    if (chosen && chosen_len) {
        std::string s;
        for (auto &k : knobs) s += std::string(s.empty() ? "" : ",") + k.name + "=" + std::to_string(*k.field);
        std::strncpy(chosen, s.c_str(), chosen_len - 1);
        chosen[chosen_len - 1] = 0;
    }
    return FULLMAG_FDM_OK;
}

// Applies a schedule string as written by fullmag_fdm_ext_autotune (or the
// FULLMAG_FDM_TUNE syntax) to the launch-time knobs of one backend.
int fullmag_fdm_ext_set_schedule(fullmag_fdm_backend *handle, const char *schedule) {
    fm::Context *cp = handle ? fm::backend_context(handle) : nullptr;
    if (!cp || !cp->ready || !schedule) return FULLMAG_FDM_ERR_INVALID;
    fm::DemagTuning t = cp->dp.tune;
    fm::parse_tuning_string(t, schedule);
    // only launch-time knobs; the planned geometry stays
    cp->dp.tune.fixed = t.fixed;
    cp->dp.tune.y_pipe = t.y_pipe;
    cp->dp.tune.y_pipe_f = t.y_pipe_f;
    cp->dp.tune.z_quad = t.z_quad;
    cp->dp.tune.z_mirror = t.z_mirror;
    cp->dp.tune.xcd = t.xcd;
    cp->dp.tune.zc = t.zc;
    cp->dp.tune.z_pipe = t.z_pipe;
    cp->dp.tune.z_wide = t.z_wide;
    cp->dp.tune.z_kstage = t.z_kstage;
    return FULLMAG_FDM_OK;
}

}  // extern "C"
