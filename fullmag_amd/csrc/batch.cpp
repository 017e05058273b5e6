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

// Members whose recorded Heun step has the same launches (kernels, launch
// geometry, argument sizes) form a class: each pass of the step is then one
// launch over all of them, instance blockIdx.z / zdiv reading its arguments
// from a device table.
struct BatchClass {
    std::vector<size_t> members;           // indices into fullmag_fdm_batch::members
    std::vector<fm::LaunchRec> proto;      // the launches of one instance
    std::vector<size_t> table_off;         // byte offset of each launch's table
};

struct fullmag_fdm_batch {
    std::vector<fm::Context *> members;
    std::vector<double> dts;
    int k = 1;
    hipStream_t origin = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;
    std::vector<std::array<double, 4>> ext;  // has_ext, H_ext at capture
    std::vector<BatchClass> classes;         // multi-instance launch groups
    std::vector<size_t> singles;             // members stepped by their own launches
    void *d_tables = nullptr;                // argument tables of all classes
    size_t table_bytes = 0;
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

bool same_launches(const std::vector<fm::LaunchRec> &a, const std::vector<fm::LaunchRec> &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i) {
        const fm::LaunchRec &x = a[i], &y = b[i];
        if (x.multi != y.multi || x.lds != y.lds || x.zdiv != y.zdiv || x.pack.size() != y.pack.size() ||
            x.grid.x != y.grid.x || x.grid.y != y.grid.y || x.grid.z != y.grid.z || x.block.x != y.block.x ||
            x.block.y != y.block.y || x.block.z != y.block.z)
            return false;
    }
    return true;
}

// Records one Heun step of every member, groups members with identical
// launches into classes (every launch having a multi-instance twin), and
// uploads the classes' argument tables.  The others step on their own.
bool plan_classes(fullmag_fdm_batch *b) {
    std::string &err = b->last_error;
    b->classes.clear();
    b->singles.clear();
    std::vector<std::vector<fm::LaunchRec>> recs(b->members.size());
    for (size_t i = 0; i < b->members.size(); ++i) {
        fm::Context &c = *b->members[i];
        fm::LaunchRecorder r;
        c.rec = &r;
        fm::integrator_enqueue_heun(c, b->dts[i], 1);
        c.rec = nullptr;
        recs[i] = std::move(r.recs);
        bool batchable = !recs[i].empty();
        for (const fm::LaunchRec &l : recs[i]) batchable = batchable && l.multi;
        if (!batchable) {
            b->singles.push_back(i);
            continue;
        }
        BatchClass *home = nullptr;
        for (BatchClass &k : b->classes)
            if (same_launches(k.proto, recs[i])) home = &k;
        if (!home) {
            b->classes.push_back(BatchClass{});
            home = &b->classes.back();
            home->proto = recs[i];
        }
        home->members.push_back(i);
    }
    // one table per launch position of each class: the members' packs back
    // to back at the pack's own stride, each table 256-byte aligned
    size_t total = 0;
    for (BatchClass &k : b->classes) {
        k.table_off.clear();
        for (const fm::LaunchRec &l : k.proto) {
            k.table_off.push_back(total);
            total += (l.pack.size() * k.members.size() + 255) / 256 * 256;
        }
    }
    if (total == 0) return true;
    if (b->table_bytes < total) {
        if (b->d_tables) hipFree(b->d_tables);
        b->d_tables = nullptr;
        b->table_bytes = 0;
        if (!fm::hip_ok(hipMalloc(&b->d_tables, total), "hipMalloc(batch tables)", err)) return false;
        b->table_bytes = total;
    }
    std::vector<unsigned char> host(total, 0);
    for (BatchClass &k : b->classes)
        for (size_t p = 0; p < k.proto.size(); ++p)
            for (size_t m = 0; m < k.members.size(); ++m) {
                const std::vector<unsigned char> &pk = recs[k.members[m]][p].pack;
                std::memcpy(host.data() + k.table_off[p] + m * pk.size(), pk.data(), pk.size());
            }
    return fm::hip_ok(hipMemcpy(b->d_tables, host.data(), total, hipMemcpyHostToDevice), "hipMemcpy(batch tables)",
                      err);
}

// Enqueues `steps` Heun steps of class k on `stream`: one launch per pass for
// all of its members.
bool enqueue_class(fullmag_fdm_batch *b, const BatchClass &k, hipStream_t stream, int steps) {
    const int m = (int)k.members.size();
    for (int s = 0; s < steps; ++s)
        for (size_t p = 0; p < k.proto.size(); ++p) {
            const fm::LaunchRec &l = k.proto[p];
            const void *tab = static_cast<const unsigned char *>(b->d_tables) + k.table_off[p];
            int zdiv = l.zdiv;
            void *args[] = {&tab, &zdiv};
            const dim3 grid(l.grid.x, l.grid.y, l.grid.z * (unsigned)m);
            if (!fm::hip_ok(hipLaunchKernel(l.multi, grid, l.block, args, l.lds, stream), "hipLaunchKernel(batch)",
                            b->last_error))
                return false;
        }
    return true;
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
    if (!plan_classes(b)) return false;
    if (!hip_ok(hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture", err))
        return false;
    hipEvent_t fork = b->events[0];
    bool ok = hip_ok(hipEventRecord(fork, b->origin), "hipEventRecord(fork)", err);
    // a class runs on its first member's stream and joins through that
    // member's event
    for (size_t ci = 0; ok && ci < b->classes.size(); ++ci) {
        const BatchClass &k = b->classes[ci];
        const size_t lead = k.members[0];
        fm::Context &c = *b->members[lead];
        ok = hip_ok(hipStreamWaitEvent(c.stream, fork, 0), "hipStreamWaitEvent(class)", err) &&
             enqueue_class(b, k, c.stream, b->k) &&
             hip_ok(hipEventRecord(b->events[1 + lead], c.stream), "hipEventRecord(join)", err) &&
             hip_ok(hipStreamWaitEvent(b->origin, b->events[1 + lead], 0), "hipStreamWaitEvent(origin)", err);
    }
    for (size_t si = 0; ok && si < b->singles.size(); ++si) {
        const size_t i = b->singles[si];
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
        if (c->stack) {
            fail(b, "batch member " + std::to_string(i) + " is a layer of a multilayer stack");
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
    if (b->d_tables) hipFree(b->d_tables);
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
    if (c.stack) {
        c.last_error = "autotune: this backend is a layer of a multilayer stack";
        return FULLMAG_FDM_ERR_INVALID;
    }
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
    if (cp->stack) {
        cp->last_error = "set_schedule: this backend is a layer of a multilayer stack";
        return FULLMAG_FDM_ERR_INVALID;
    }
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
