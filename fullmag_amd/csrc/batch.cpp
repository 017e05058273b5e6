// batch.cpp -- many independent backends stepped as one hipGraph, and the
// launch-schedule autotuner (extensions; the reference runs one problem per
// process and has neither).
//
// A batch captures, for every member backend, `k` consecutive fixed-dt Heun
// steps enqueued on that backend's own stream.  The member streams fork from
// and join back into one origin stream during capture, so the graph holds
// one independent branch per member; a single hipGraphLaunch then advances
// every member by k steps, and the runtime is free to run the branches
// concurrently.  Kernel arguments (buffer pointers, dt, the external field)
// are baked in at capture; the Heun path swaps no buffers, so replays are
// exact.  The graph holds no memset nodes.  Before each run the batch
// re-establishes what the captured steps assume of every member: H_demag
// of the current m in the member's demag buffer (recomputed when an upload
// or state copy replaced m), and the captured external field (the graph is
// recaptured when set_external_field changed it).
#include "context.hpp"

#include <algorithm>
#include <array>
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
    std::vector<std::array<double, 4>> ext;  // has_ext, H_ext at capture
    std::string last_error;
};

namespace {

bool fail(fullmag_fdm_batch *b, const std::string &what) {
    b->last_error = what;
    return false;
}

std::array<double, 4> ext_of(const fm::Context &c) {
    return {c.has_ext ? 1.0 : 0.0, c.hext[0], c.hext[1], c.hext[2]};
}

bool capture(fullmag_fdm_batch *b) {
    using fm::hip_ok;
    std::string &err = b->last_error;
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->ext.clear();
    for (fm::Context *c : b->members) b->ext.push_back(ext_of(*c));
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
    // members changed since the capture: a new external field needs a new
    // graph; a new m needs H_demag(m) before the first captured stage
    bool recapture = false;
    for (size_t i = 0; i < b->members.size(); ++i) {
        fm::Context &c = *b->members[i];
        if (ext_of(c) != b->ext[i]) recapture = true;
        if (!fm::integrator_prepare(c)) {
            b->last_error = "batch member " + std::to_string(i) + ": " + c.last_error;
            return FULLMAG_FDM_ERR_CUDA;
        }
    }
    if (recapture && !capture(b)) return FULLMAG_FDM_ERR_CUDA;
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

// Coordinate search over the schedule knobs of the demag transforms and the
// cell kernels, timing (demag + stats) at the current state; keeps the
// fastest schedule and writes it as "key=value,..." into `chosen`.  The
// plan-shaping knobs (radices, rows per x workgroup, y lines per workgroup,
// long-z tile) are tried by re-planning the transforms in place; candidates
// whose plan does not fit are skipped.
int fullmag_fdm_ext_autotune(fullmag_fdm_backend *handle, int reps, char *chosen, uint64_t chosen_len) {
    fm::Context *cp = handle ? fm::backend_context(handle) : nullptr;
    if (!cp || !cp->ready) return FULLMAG_FDM_ERR_INVALID;
    fm::Context &c = *cp;
    reps = std::max(1, reps);
    struct Knob {
        const char *name;
        int fm::DemagTuning::*field;
        std::vector<int> values;
        bool plan;  // needs a re-plan
    };
    using T = fm::DemagTuning;
    std::vector<Knob> knobs = {{"x_radix", &T::x_radix, {16, 8, 4}, true},
                               {"x_rows", &T::x_rows, {0, 8, 4, 2}, true},
                               {"y_radix", &T::y_radix, {16, 8, 4}, true},
                               {"y_lpw", &T::y_lpw, {0, 1, 2}, true},
                               {"fixed", &T::fixed, {1, 0}, false},
                               {"y_pipe", &T::y_pipe, {1, 0, 2, 4, 8}, false},
                               {"y_pipe_f", &T::y_pipe_f, {1, 2, 4, 8}, false},
                               {"z_quad", &T::z_quad, {0, 1}, false},
                               {"xcd", &T::xcd, {1, 0}, false},
                               {"zc", &T::zc, {0, 1, 2, 4}, false}};
    if (c.demag && c.dp.three_d && c.dp.pz > 16) knobs.push_back({"z_tile", &T::z_tile, {0, 16, 8, 4, 2}, true});
    std::string err;
    double best = fm::integrator_probe_ms(c, reps);
    for (auto &k : knobs) {
        fm::DemagTuning cur = c.dp.tune;
        const int keep = cur.*k.field;
        int best_v = keep;
        for (int v : k.values) {
            if (v == keep) continue;
            fm::DemagTuning t = cur;
            t.*k.field = v;
            if (k.plan) {
                if (!fm::demag_replan(c, t, err)) continue;
            } else {
                c.dp.tune = t;
            }
            const double ms = fm::integrator_probe_ms(c, reps);
            if (ms < best * 0.98) {
                best = ms;
                best_v = v;
            }
        }
        fm::DemagTuning t = cur;
        t.*k.field = best_v;
        if (!fm::demag_replan(c, t, err)) {
            c.last_error = "autotune: " + err;
            return FULLMAG_FDM_ERR_INTERNAL;
        }
    }
    if (!fm::hip_ok(hipGetLastError(), "autotune", c.last_error)) return FULLMAG_FDM_ERR_CUDA;
    if (chosen && chosen_len) {
        std::string s;
        for (auto &k : knobs)
            s += std::string(s.empty() ? "" : ",") + k.name + "=" + std::to_string(c.dp.tune.*k.field);
        std::strncpy(chosen, s.c_str(), chosen_len - 1);
        chosen[chosen_len - 1] = 0;
    }
    return FULLMAG_FDM_OK;
}

// Applies a schedule string as written by fullmag_fdm_ext_autotune (or the
// FULLMAG_FDM_TUNE syntax) to one backend, re-planning its transforms.
int fullmag_fdm_ext_set_schedule(fullmag_fdm_backend *handle, const char *schedule) {
    fm::Context *cp = handle ? fm::backend_context(handle) : nullptr;
    if (!cp || !cp->ready || !schedule) return FULLMAG_FDM_ERR_INVALID;
    fm::DemagTuning t = cp->dp.tune;
    fm::parse_tuning_string(t, schedule);
    t.y_split = cp->dp.tune.y_split;  // the spectrum layout stays
    std::string err;
    if (!fm::demag_replan(*cp, t, err)) {
        cp->last_error = "set_schedule: " + err;
        return FULLMAG_FDM_ERR_INVALID;
    }
    cp->pre_valid = false;  // the stored predictor came from the old schedule's rounding
    return FULLMAG_FDM_OK;
}

}  // extern "C"
