// Batched many-volume Gaussian (itkg_batch_*): no counterpart in the
// reference; each volume's result equals GaussianFilter{source: volume,
// sigma}.generate_region(largest_region) (src/filters/gaussian.rs:59-90), or
// its adjoint.
//
// Plan: volumes whose radii the fused kernel serves are grouped by radius;
// each group gets a device table of per-volume launch arguments and a packed
// work list of (volume, tile) items, and runs as ONE persistent launch of the
// fused kernel over that list.  Other volumes (anisotropic radii, radius
// beyond the fused set) run the per-axis path with scratch preallocated at
// plan time.  The groups and the per-axis volumes are dealt out to a few
// branches (internal streams forked from and joined back into the caller's
// stream), so that one group's tail overlaps the next group's start.
// itkg_batch_run captures the whole forked sequence into a HIP graph on a
// stream's first run and replays it afterwards.
//
// Autotuner: the z run of a tile (TZ) trades halo recomputation against
// parallelism; itkg_batch_autotune times TZ candidates per radius group on
// the plan's own volumes and keeps the fastest, optionally caching the
// choice in the file named by ITKG_TUNE_CACHE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "itkg_impl.hpp"

namespace itkg {

namespace {

struct BatchClass {
    int r = 0;
    const FusedEntry* entry = nullptr;  // its persistent launcher
    int mode = FUSED_FWD_EXACT;         // kernel mode, which sets the tile shape
    int tz = 0;
    std::vector<int> vol;            // volume indices in this group
    Fused3dArgs* d_args = nullptr;   // one per volume of the group
    int2* d_work = nullptr;          // (index into d_args, tile), XCD-interleaved
    int nwork = 0;
    int grid = 0;
    float measured_ms = -1.0f;  // autotuner's time of the chosen z run (< 0: not timed)
};

// Branches (side streams) a plan forks its launches into.  A process gets
// four hardware queues by default and the caller's stream holds one; a
// fourth side stream would share a queue and run after the others.
constexpr int kBatchBranches = 3;

using TuneKey = std::tuple<int, int, int>;  // direction, arith, radius

// Cache lines: "direction arith radius tz tile_rows".
std::map<TuneKey, std::pair<int, int>> read_cache() {
    std::map<TuneKey, std::pair<int, int>> m;
    const char* path = std::getenv("ITKG_TUNE_CACHE");
    if (!path) return m;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        int d, a, r, tz, ty;
        if (ss >> d >> a >> r >> tz >> ty && tz >= 0) m[{d, a, r}] = {tz, ty};
    }
    return m;
}

void append_cache(const TuneKey& k, int tz, int ty) {
    const char* path = std::getenv("ITKG_TUNE_CACHE");
    if (!path) return;
    std::ofstream f(path, std::ios::app);
    f << std::get<0>(k) << ' ' << std::get<1>(k) << ' ' << std::get<2>(k) << ' ' << tz << ' ' << ty
      << '\n';
}

}  // namespace

struct BatchImpl {
    std::vector<itkg_volume> vols;
    std::vector<Plan> plans;
    itkg_batch_direction dir = ITKG_BATCH_FORWARD;
    itkg_arith arith = ITKG_EXACT;
    int mode = FUSED_FWD_EXACT;
    std::vector<BatchClass> classes;
    std::vector<int> rest;           // volumes on the per-axis path
    void* scratch[kBatchBranches] = {};  // per-axis scratch of each branch
    int64_t voxels = 0;
    int num_cu = 256;
    std::mutex mu;
    std::unordered_map<hipStream_t, hipGraphExec_t> graphs;
    hipStream_t side[kBatchBranches] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[kBatchBranches] = {};
    // per branch: a group (index >= 0) or a per-axis volume v (encoded -1 - v)
    std::vector<std::vector<int>> branch_units;
    std::vector<float> rest_ms;  // autotuner's time per per-axis volume (empty: not timed)

    ~BatchImpl() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (int b = 0; b < kBatchBranches; ++b) {
            if (side[b]) (void)hipStreamDestroy(side[b]);
            if (join_ev[b]) (void)hipEventDestroy(join_ev[b]);
        }
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        for (auto& c : classes) {
            if (c.d_args) (void)hipFree(c.d_args);
            if (c.d_work) (void)hipFree(c.d_work);
        }
        for (void* p : scratch)
            if (p) (void)hipFree(p);
    }

    // The 32-row-tile variant of the plan's mode, or -1 (EXACT has none).
    int small_mode() const {
        return mode == FUSED_FWD_FAST ? FUSED_FWD_FAST_S : (mode == FUSED_ADJ ? FUSED_ADJ_S : -1);
    }

    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    // Modelled time of one group at z run tz, in plane steps per CU: the
    // persistent grid deals the tiles out evenly, and the last tile dealt can
    // finish up to one tile's walk after the others.  Workgroups resident
    // together on one CU share it (they add no throughput).
    double model_cost(const BatchClass& c, int tz, int m) const {
        const int64_t slots = num_cu;
        double steps = 0.0, longest = 0.0;
        for (int i : c.vol) {
            const itkg_volume& v = vols[i];
            const int64_t txy = ((v.size[0] + kFusedTileX - 1) / kFusedTileX) *
                                ((v.size[1] + c.entry->ty[m] - 1) / c.entry->ty[m]);
            const int64_t nz = v.size[2];
            const int64_t tzv = std::min<int64_t>(tz, nz);
            const double walk = (double)(tzv + 2 * c.r + 3);
            steps += (double)(txy * ((nz + tz - 1) / tz)) * walk;
            longest = std::max(longest, walk);
        }
        return steps / (double)slots + longest;
    }

    // Work of one plane step grows with the tile's rows plus their x halo,
    // plus a fixed part (loads, barrier).
    static double step_work(const BatchClass& c, int m) { return 3.0 * c.entry->ty[m] + 2.0 * c.r + 30.0; }

    int model_tz(const BatchClass& c, int m) const {
        int best = 32;
        double best_cost = 1e300;
        for (int tz : {16, 24, 32, 48, 64, 96, 128, 192, 256}) {
            const double cost = model_cost(c, tz, m);
            if (cost < best_cost) {
                best_cost = cost;
                best = tz;
            }
        }
        return best;
    }

    // Relative GPU time of one output voxel, fitted to the measured plain
    // launches on one 512^3 volume: the fused kernel costs about the same up
    // to radius 6 (memory and synchronisation bound) and grows with the taps
    // beyond; the per-axis path makes three passes through HBM.
    static double voxel_cost(int r, bool fused) {
        if (!fused) return 3.0;
        return std::max(1.0, r / 6.0);
    }

    // Estimated time of a group: its voxels at the radius's cost, times the
    // z-ring fill of each volume's tiles (tz + 2r + 3 steps for tz planes).
    double class_work(const BatchClass& c) const {
        double w = 0.0;
        for (int i : c.vol) {
            const itkg_volume& v = vols[i];
            const double nz = (double)v.size[2];
            const int tz = c.tz > 0 ? c.tz
                                    : fused_tz((int)(((v.size[0] + kFusedTileX - 1) / kFusedTileX) *
                                                     ((v.size[1] + c.entry->ty[c.mode] - 1) /
                                                      c.entry->ty[c.mode])),
                                               (int)v.size[2], c.r);
            const double tzv = std::min<double>(tz, nz);
            w += (double)v.size[0] * v.size[1] * nz * voxel_cost(c.r, true) *
                 (tzv + 2 * c.r + 3) / tzv;
        }
        return w;
    }

    // Deals the groups and the per-axis volumes (largest estimated time
    // first) out to the branches, each to the least loaded one.
    // After itkg_batch_autotune the measured times are used instead.
    void assign_branches() {
        std::vector<std::pair<double, int>> units;
        bool timed = rest_ms.size() == rest.size();
        for (const BatchClass& c : classes) timed = timed && c.measured_ms >= 0.0f;
        for (size_t i = 0; i < classes.size(); ++i)
            units.push_back({timed ? (double)classes[i].measured_ms : class_work(classes[i]), (int)i});
        for (size_t k = 0; k < rest.size(); ++k) {
            const Plan& p = plans[rest[k]];
            int64_t r = 0;
            for (int d = 0; d < 3; ++d) r = std::max<int64_t>(r, p.radii[d]);
            units.push_back({timed ? (double)rest_ms[k]
                                   : (double)region_len(3, p.req) * voxel_cost((int)r, false),
                             -1 - rest[k]});
        }
        std::sort(units.begin(), units.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        const int nb = std::min<int>(kBatchBranches, (int)units.size());
        branch_units.assign(nb, {});
        std::vector<double> load(nb, 0.0);
        for (const auto& u : units) {
            const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[b] += u.first;
            branch_units[b].push_back(u.second);
        }
    }

    // (Re)builds the device tables of one group for a z run of `tz`.
    int build_class(BatchClass& c, int tz) {
        std::vector<Fused3dArgs> args(c.vol.size());
        std::vector<int2> items;
        for (size_t i = 0; i < c.vol.size(); ++i) {
            const itkg_volume& v = vols[c.vol[i]];
            const Plan& p = plans[c.vol[i]];
            const itkg_view vin = dense_view(3, const_cast<float*>(v.input), p.req);
            const itkg_view vout = dense_view(3, v.output, p.req);
            fill_fused_args(p, vin, vout, c.mode, 0, tz, 0, 1, &args[i]);
            const int nt = args[i].tiles_x * args[i].tiles_y * args[i].tiles_z;
            for (int t = 0; t < nt; ++t) items.push_back(make_int2((int)i, t));
        }
        // Workgroup b runs items b, b + G, b + 2G, ... and sits on XCD b mod 8
        // (G is a multiple of 8): deal the list out in 8 contiguous chunks so
        // that neighbouring tiles, which share halos, share an XCD's L2.
        const int n = (int)items.size();
        const int q = (n + 7) / 8;
        std::vector<int2> work((size_t)q * 8, make_int2(-1, 0));
        for (int i = 0; i < n; ++i) work[(size_t)(i % q) * 8 + i / q] = items[i];
        if (c.d_args) (void)hipFree(c.d_args);
        if (c.d_work) (void)hipFree(c.d_work);
        c.d_args = nullptr;
        c.d_work = nullptr;
        hipError_t e = hipMalloc(&c.d_args, args.size() * sizeof(Fused3dArgs));
        if (e == hipSuccess) e = hipMalloc(&c.d_work, work.size() * sizeof(int2));
        if (e == hipSuccess)
            e = hipMemcpy(c.d_args, args.data(), args.size() * sizeof(Fused3dArgs),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(c.d_work, work.data(), work.size() * sizeof(int2), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "batch tables");
        c.tz = tz;
        c.nwork = (int)work.size();
        const int per_cu = std::max(1, c.entry->occupancy(c.mode));
        c.grid = std::min(c.nwork, ((num_cu * per_cu + 7) / 8) * 8);
        return ITKG_OK;
    }

    int enqueue_class(const BatchClass& c, hipStream_t s) const {
        if (c.nwork == 0) return ITKG_OK;
        hipError_t e = c.entry->batch(c.d_args, c.d_work, c.nwork, c.grid, c.mode, s);
        if (e != hipSuccess) return hip_fail(e, "batched fused launch");
        return ITKG_OK;
    }

    int enqueue_volume(int i, hipStream_t s, void* scr) const {
        const itkg_volume& v = vols[i];
        const Plan& p = plans[i];
        const itkg_view vin = dense_view(3, const_cast<float*>(v.input), p.req);
        const itkg_view vout = dense_view(3, v.output, p.req);
        return dir == ITKG_BATCH_FORWARD ? forward_impl(p, ITKG_F32, 1, vin, vout, arith, s, scr)
                                         : backward_impl(p, ITKG_F32, 1, vin, vout, 0, s, scr);
    }

    // The whole plan on `s`: fork to the branches, run each branch's units in
    // order, join back into `s`.
    int enqueue(hipStream_t s) {
        hipError_t e = hipEventRecord(fork_ev, s);
        if (e != hipSuccess) return hip_fail(e, "batch fork");
        for (size_t b = 0; b < branch_units.size(); ++b) {
            e = hipStreamWaitEvent(side[b], fork_ev, 0);
            if (e != hipSuccess) return hip_fail(e, "batch fork");
            for (int u : branch_units[b]) {
                const int rc = u < 0 ? enqueue_volume(-1 - u, side[b], scratch[b])
                                     : enqueue_class(classes[u], side[b]);
                if (rc) return rc;
            }
            e = hipEventRecord(join_ev[b], side[b]);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, join_ev[b], 0);
            if (e != hipSuccess) return hip_fail(e, "batch join");
        }
        return ITKG_OK;
    }
};

}  // namespace itkg

struct itkg_batch {
    itkg::BatchImpl impl;
};

using namespace itkg;

extern "C" {

int itkg_batch_create(int32_t n, const itkg_volume* volumes, itkg_batch_direction direction,
                      itkg_arith arith, itkg_batch** plan) {
    clear_error();
    if (n < 0 || (n > 0 && !volumes) || !plan) return fail(ITKG_ERR_INVALID_ARG, "bad arguments");
    if (direction != ITKG_BATCH_FORWARD && direction != ITKG_BATCH_BACKWARD)
        return fail(ITKG_ERR_INVALID_ARG, "bad direction");
    if (arith != ITKG_EXACT && arith != ITKG_FAST) return fail(ITKG_ERR_INVALID_ARG, "bad arith");
    auto* b = new (std::nothrow) itkg_batch;
    if (!b) return fail(ITKG_ERR_OOM, "host allocation");
    BatchImpl& B = b->impl;
    B.dir = direction;
    B.arith = arith;
    B.mode = direction == ITKG_BATCH_BACKWARD ? FUSED_ADJ
                                              : (arith == ITKG_FAST ? FUSED_FWD_FAST : FUSED_FWD_EXACT);
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            cus > 0)
            B.num_cu = cus;
    }
    size_t sc = 0;
    std::map<int, size_t> by_r;
    for (int i = 0; i < n; ++i) {
        const itkg_volume& v = volumes[i];
        itkg_region r{};
        for (int d = 0; d < 3; ++d) r.size[d] = v.size[d];
        Plan p;
        int rc = make_plan(3, v.sigma, v.spacing, &r, r, &p);
        if (!rc && region_len(3, r) > 0 && (!v.input || !v.output))
            rc = fail(ITKG_ERR_INVALID_ARG, "volume %d has a null buffer", i);
        if (rc) {
            delete b;
            return rc;
        }
        B.vols.push_back(v);
        B.plans.push_back(p);
        B.voxels += region_len(3, r);
        if (region_len(3, r) == 0) continue;
        const itkg_view vsrc = dense_view(3, const_cast<float*>(v.input), r);
        const FusedEntry* fe =
            use_fused(p, ITKG_F32, 1) && fused_src_ok(p, vsrc, 1, direction == ITKG_BATCH_BACKWARD)
                ? fused_for(p)
                : nullptr;
        if (fe && fe->batch && p.D == 3 && fe->rz == fe->rxy) {
            const int rad = (int)p.radii[0];
            if (!by_r.count(rad)) {
                by_r[rad] = B.classes.size();
                B.classes.emplace_back();
                B.classes.back().r = rad;
                B.classes.back().entry = fe;
            }
            B.classes[by_r[rad]].vol.push_back(i);
        } else {
            B.rest.push_back(i);
            sc = std::max(sc, scratch_bytes(p, ITKG_F32, 1));
        }
    }
    // every branch may receive per-axis volumes (the autotuner re-deals them)
    for (int k = 0; k < kBatchBranches && sc > 0 && !B.rest.empty(); ++k) {
        hipError_t e = hipMalloc(&B.scratch[k], sc);
        if (e != hipSuccess) {
            delete b;
            return hip_fail(e, "hipMalloc(batch scratch)");
        }
    }
    hipError_t e = hipEventCreateWithFlags(&B.fork_ev, hipEventDisableTiming);
    for (int k = 0; k < kBatchBranches && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&B.side[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&B.join_ev[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        delete b;
        return hip_fail(e, "batch streams");
    }
    const auto cache = read_cache();
    for (BatchClass& c : B.classes) {
        auto it = cache.find({(int)direction, (int)arith, c.r});
        int tz = 0;
        c.mode = B.mode;
        const int small = B.small_mode();
        if (it != cache.end()) {
            tz = it->second.first;
            if (small >= 0 && it->second.second == c.entry->ty[small]) c.mode = small;
        } else {
            // the tile shape and z run the model prefers
            double best = 1e300;
            for (int m : {B.mode, small}) {
                if (m < 0) continue;
                const int t = B.model_tz(c, m);
                const double cost = B.model_cost(c, t, m) * BatchImpl::step_work(c, m);
                if (cost < best) {
                    best = cost;
                    c.mode = m;
                    tz = t;
                }
            }
        }
        const int rc = B.build_class(c, tz);
        if (rc) {
            delete b;
            return rc;
        }
    }
    B.assign_branches();
    *plan = b;
    return ITKG_OK;
}

int itkg_batch_autotune(itkg_batch* plan, void* stream) {
    clear_error();
    if (!plan) return fail(ITKG_ERR_INVALID_ARG, "null plan");
    BatchImpl& B = plan->impl;
    std::lock_guard<std::mutex> lock(B.mu);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(ITKG_ERR_HIP, "hipEventCreate");
    int rc = ITKG_OK;
    for (BatchClass& c : B.classes) {
        int maxz = 1;
        for (int i : c.vol) maxz = std::max<int>(maxz, (int)B.vols[i].size[2]);
        float best_ms = 1e30f;
        int best = c.tz, best_mode = c.mode;
        for (int m : {B.mode, B.small_mode()}) {
        if (m < 0) continue;
        c.mode = m;
        // tz 0: each volume takes the z run of its own plain launch
        for (int tz : {0, 16, 24, 32, 48, 64, 96, 128, 192, 256}) {
            if (tz > maxz && tz != 16) continue;
            if ((rc = B.build_class(c, tz))) break;
            if ((rc = B.enqueue_class(c, s))) break;  // warm-up
            float tot = 0.0f;
            for (int rep = 0; rep < 3 && !rc; ++rep) {
                (void)hipEventRecord(e0, s);
                rc = B.enqueue_class(c, s);
                (void)hipEventRecord(e1, s);
                (void)hipEventSynchronize(e1);
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                tot += ms;
            }
            if (rc) break;
            if (std::getenv("ITKG_DEBUG"))
                std::fprintf(stderr, "itkg: autotune r=%d mode=%d tz=%d: %.4f ms\n", c.r, m, tz, tot / 3.0f);
            if (tot < best_ms) {
                best_ms = tot;
                best = tz;
                best_mode = m;
            }
        }
        if (rc) break;
        }
        if (rc) break;
        c.mode = best_mode;
        if ((rc = B.build_class(c, best))) break;
        c.measured_ms = best_ms / 3.0f;
        append_cache({(int)B.dir, (int)B.arith, c.r}, best, c.entry->ty[c.mode]);
    }
    // the per-axis volumes, for the branch balance
    B.rest_ms.clear();
    for (size_t k = 0; k < B.rest.size() && !rc; ++k) {
        rc = B.enqueue_volume(B.rest[k], s, B.scratch[0]);  // warm-up
        (void)hipEventRecord(e0, s);
        if (!rc) rc = B.enqueue_volume(B.rest[k], s, B.scratch[0]);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        B.rest_ms.push_back(ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B.assign_branches();
    B.drop_graphs();
    return rc;
}

int itkg_batch_run(itkg_batch* plan, void* stream) {
    clear_error();
    if (!plan) return fail(ITKG_ERR_INVALID_ARG, "null plan");
    BatchImpl& B = plan->impl;
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(B.mu);
    if (!s) return B.enqueue(s);  // the legacy null stream cannot be captured
    auto it = B.graphs.find(s);
    if (it == B.graphs.end()) {
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
        const int rc = B.enqueue(s);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        hipGraphExec_t ge = nullptr;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
        it = B.graphs.emplace(s, ge).first;
    }
    hipError_t e = hipGraphLaunch(it->second, s);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    return ITKG_OK;
}

int64_t itkg_batch_voxels(const itkg_batch* plan) { return plan ? plan->impl.voxels : 0; }

int itkg_batch_destroy(itkg_batch* plan) {
    clear_error();
    delete plan;
    return ITKG_OK;
}

}  // extern "C"
