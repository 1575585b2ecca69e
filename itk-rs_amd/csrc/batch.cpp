// Batched many-volume Gaussian (itkg_batch_*): no counterpart in the
// reference; each volume's result equals GaussianFilter{source: volume,
// sigma}.generate_region(largest_region) (src/filters/gaussian.rs:59-90), or
// its adjoint.
//
// FAST forward and adjoint plans: every volume the fused kernel serves
// (isotropic radius up to 24 forward, 15 adjoint) runs in ONE persistent
// launch (fused3d_mega_kernel) over a packed work list of (volume, tile,
// class) items: each volume gets a tile shape and a z run from a cost model
// (or the autotuner), its tiles are dealt in contiguous chunks to 8 per-XCD
// queues, each queue sorted by decreasing modelled cost, and workgroups
// take items dynamically (their own XCD's queue first).  EXACT plans group
// the fused volumes by radius, each group one persistent launch of its own
// instantiation (EXACT needs 8-wave workgroups at some radii, so its radii
// cannot share one kernel).  Volumes outside the fused set run the per-axis
// path with scratch preallocated at plan time, on side streams forked from
// and joined back into the caller's stream.  itkg_batch_run captures the
// whole sequence into a HIP graph on a stream's first run and replays it
// afterwards.
//
// Autotuner: itkg_batch_autotune times candidate plans on the plan's own
// volumes and keeps the fastest -- for the single launch the item size limit
// and the tile-shape policy, for EXACT groups the z run and tile shape per
// group -- optionally caching the choice in the file named by
// ITKG_TUNE_CACHE.
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

// The single persistent launch of a FAST / adjoint plan.
struct MegaPlan {
    bool active = false;
    int family = FUSED_FWD_FAST;  // FUSED_FWD_FAST or FUSED_ADJ
    std::vector<int> elig;        // volumes the fused kernel can serve
    int max_r = 15;               // the launch serves those up to this radius ...
    std::vector<int> big;         // ... and the larger ones run the per-axis path first
    std::vector<int> vol;         // volumes it serves
    std::vector<int> cls;         // their classes
    std::vector<int> tz;          // their z runs
    Fused3dArgsB* d_args = nullptr;
    int4* d_work = nullptr;
    int* d_qstart = nullptr;      // kMegaQueues + 1 queue bounds
    int* d_heads = nullptr;       // kMegaQueues heads + the finish count
    int nwork = 0, grid = 0, lds = 0;
    int split_x100 = 200;         // item size limit = total cost / (CUs * split)
    int shape_policy = 0;         // radii <= 7: 0 model, 1 family shape, 2 32-row tiles
    int only_class = 0;           // > 0: the work list holds this class only (timing breakdown)
    float measured_ms = -1.0f;
};
constexpr int kMegaQueuesHost = 8;

using TuneKey = std::tuple<int, int, int>;  // direction, arith, radius (-1: the single launch)

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
    std::vector<float> big_ms;   // the same for the volumes above the single launch's cap
    MegaPlan mega;
    int mega_lds_tab[kMegaClasses] = {};
    int mega_ty_tab[kMegaClasses] = {};

    void free_mega_tables() {
        if (mega.d_args) (void)hipFree(mega.d_args);
        if (mega.d_work) (void)hipFree(mega.d_work);
        if (mega.d_qstart) (void)hipFree(mega.d_qstart);
        mega.d_args = nullptr;
        mega.d_work = nullptr;
        mega.d_qstart = nullptr;
    }

    ~BatchImpl() {
        free_mega_tables();
        if (mega.d_heads) (void)hipFree(mega.d_heads);
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
        bool timed = rest_ms.size() == rest.size() && big_ms.size() == mega.big.size();
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
        for (size_t k = 0; k < mega.big.size(); ++k) {  // above the single launch's radius cap
            const int v = mega.big[k];
            units.push_back({timed ? (double)big_ms[k]
                                   : (double)region_len(3, plans[v].req) *
                                         voxel_cost((int)plans[v].radii[0], false),
                             -1 - v});
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

    // ---- the single persistent launch ----------------------------------

    // Modelled time of one plane step of a tile with `ty` rows at radius r:
    // the larger of its FMAs (x pass over ty + 2r rows, y and z passes over
    // ty rows, 2r + 1 taps each) and its memory walk (window in, outputs
    // out), in units of one FMA instruction per CU; fitted to the 512^3
    // radius-6 kernel, whose step is about as long as its walk.
    static double mega_step_cost(int r, int ty) {
        const double hy = ty + 2.0 * r;
        const double valu = 1.2 * (hy + 2.0 * ty) * (2 * r + 1);
        const double walk = 0.085 * (hy * (64.0 + 2 * r) * 4.0 + ty * 64.0 * 4.0);
        return std::max(valu, walk);
    }

    int mega_class_r(int cls) const { return cls > kMegaClassSmall ? cls - kMegaClassSmall : cls; }

    // Chooses each served volume's class (tile shape) and z run for the
    // current split and shape policy.
    void mega_pick() {
        MegaPlan& M = mega;
        const size_t n = M.vol.size();
        M.cls.assign(n, 0);
        M.tz.assign(n, 0);
        std::vector<double> step(n), txy(n);
        double total = 0.0;
        for (size_t i = 0; i < n; ++i) {
            const itkg_volume& v = vols[M.vol[i]];
            const int r = (int)plans[M.vol[i]].radii[0];
            auto tiles_of = [&](int c) {
                const int ty = mega_ty_tab[c];
                return (double)((v.size[0] + kFusedTileX - 1) / kFusedTileX) *
                       (double)((v.size[1] + ty - 1) / ty);
            };
            int c = r;
            if (r <= 7) {
                const int small = kMegaClassSmall + r;
                if (M.shape_policy == 2) {
                    c = small;
                } else if (M.shape_policy == 0) {
                    const double big = tiles_of(r) * mega_step_cost(r, mega_ty_tab[r]);
                    const double sm = tiles_of(small) * mega_step_cost(r, mega_ty_tab[small]);
                    if (sm < big) c = small;
                }
            }
            M.cls[i] = c;
            step[i] = mega_step_cost(r, mega_ty_tab[c]);
            txy[i] = tiles_of(c);
            total += txy[i] * (double)(v.size[2] + 2 * r + 3) * step[i];
        }
        // longest item allowed: the total over the CUs, divided by `split`
        const double limit = total / ((double)num_cu * std::max(0.05, M.split_x100 / 100.0));
        for (size_t i = 0; i < n; ++i) {
            const int64_t nz = vols[M.vol[i]].size[2];
            const int r = mega_class_r(M.cls[i]);
            const int64_t tz_max = (int64_t)(limit / step[i]) - 2 * r - 3;
            int64_t tz = nz;
            if (tz_max < nz) {
                const int64_t k = (nz + std::max<int64_t>(tz_max, 8) - 1) / std::max<int64_t>(tz_max, 8);
                tz = (nz + k - 1) / k;
            }
            M.tz[i] = (int)std::max<int64_t>(1, tz);
        }
    }

    // (Re)builds the device tables of the single launch.
    int mega_build() {
        MegaPlan& M = mega;
        free_mega_tables();
        const size_t n = M.vol.size();
        std::vector<Fused3dArgsB> args(n);
        struct Item {
            int4 w;
            double cost;
        };
        std::vector<std::vector<Item>> q(kMegaQueuesHost);
        int lds = 0;
        int rot = 0;  // queue of the first chunk of the next volume
        for (size_t i = 0; i < n; ++i) {
            const itkg_volume& v = vols[M.vol[i]];
            const Plan& p = plans[M.vol[i]];
            const int c = M.cls[i], r = mega_class_r(c);
            const int mode = c > kMegaClassSmall
                                 ? (M.family == FUSED_FWD_FAST ? FUSED_FWD_FAST_S : FUSED_ADJ_S)
                                 : M.family;
            const itkg_view vin = dense_view(3, const_cast<float*>(v.input), p.req);
            const itkg_view vout = dense_view(3, v.output, p.req);
            fill_fused_args(p, vin, vout, mode, 0, M.tz[i], 0, 1, &args[i]);
            Fused3dArgsB& A = args[i];
            A.tiles_y = (A.n_dst[1] + mega_ty_tab[c] - 1) / mega_ty_tab[c];
            lds = std::max(lds, mega_lds_tab[c]);
            const int nt = A.tiles_x * A.tiles_y * A.tiles_z;
            const double step = mega_step_cost(r, mega_ty_tab[c]);
            const int per_plane = A.tiles_x * A.tiles_y;
            // contiguous chunks of the tile list to the queues (neighbouring
            // tiles, which share halos, stay on one XCD's L2)
            const int chunks = std::min(nt, kMegaQueuesHost);
            for (int k = 0; k < chunks; ++k) {
                const int lo = (int)((int64_t)nt * k / chunks), hi = (int)((int64_t)nt * (k + 1) / chunks);
                auto& qq = q[(rot + k) % kMegaQueuesHost];
                for (int t = lo; t < hi; ++t) {
                    if (M.only_class > 0 && c != M.only_class) continue;
                    const int z0 = (t / per_plane) * A.tz;
                    const int planes = std::min(A.tz, A.n_dst[2] - z0);
                    qq.push_back({make_int4((int)i, t, c, 0), (planes + 2 * r + 3) * step});
                }
            }
            rot = (rot + chunks) % kMegaQueuesHost;
        }
        // Every queue runs the classes in the same order (the class with the
        // longest mean item first), and within a class its longest items
        // first: the GPU then sweeps one class at a time, so the CUs that
        // share an instruction cache run the same code (each class's loop is
        // tens of KB), and the short items of the last class fill the tail.
        std::vector<double> cls_sum(kMegaClasses, 0.0), cls_n(kMegaClasses, 0.0);
        for (const auto& qq : q)
            for (const Item& it : qq) {
                cls_sum[it.w.z] += it.cost;
                cls_n[it.w.z] += 1.0;
            }
        std::vector<double> cls_mean(kMegaClasses, 0.0);
        for (int c = 0; c < kMegaClasses; ++c) cls_mean[c] = cls_n[c] > 0 ? cls_sum[c] / cls_n[c] : 0.0;
        std::vector<int4> work;
        std::vector<int> qstart(kMegaQueuesHost + 1, 0);
        for (int k = 0; k < kMegaQueuesHost; ++k) {
            std::stable_sort(q[k].begin(), q[k].end(), [&](const Item& a, const Item& b) {
                if (a.w.z != b.w.z) {
                    const double ma = cls_mean[a.w.z], mb = cls_mean[b.w.z];
                    return ma != mb ? ma > mb : a.w.z < b.w.z;
                }
                return a.cost > b.cost;
            });
            qstart[k] = (int)work.size();
            for (const Item& it : q[k]) work.push_back(it.w);
        }
        qstart[kMegaQueuesHost] = (int)work.size();
        M.nwork = (int)work.size();
        M.lds = lds;
        hipError_t e = M.family == FUSED_FWD_FAST ? mega_prepare_fast(lds) : mega_prepare_adj(lds);
        if (e != hipSuccess) return hip_fail(e, "batched launch: LDS size");
        const int per_cu = M.family == FUSED_FWD_FAST ? mega_occupancy_fast(lds) : mega_occupancy_adj(lds);
        M.grid = std::max(1, std::min(M.nwork, num_cu * std::max(1, per_cu)));
        e = hipMalloc(&M.d_args, std::max<size_t>(1, n) * sizeof(Fused3dArgsB));
        if (e == hipSuccess) e = hipMalloc(&M.d_work, std::max<size_t>(1, work.size()) * sizeof(int4));
        if (e == hipSuccess) e = hipMalloc(&M.d_qstart, qstart.size() * sizeof(int));
        if (e == hipSuccess && n)
            e = hipMemcpy(M.d_args, args.data(), n * sizeof(Fused3dArgsB), hipMemcpyHostToDevice);
        if (e == hipSuccess && !work.empty())
            e = hipMemcpy(M.d_work, work.data(), work.size() * sizeof(int4), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(M.d_qstart, qstart.data(), qstart.size() * sizeof(int), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "batched launch tables");
        return ITKG_OK;
    }

    // Splits the eligible volumes at the radius cap.
    void mega_assign() {
        MegaPlan& M = mega;
        M.vol.clear();
        M.big.clear();
        big_ms.clear();
        for (int v : M.elig) (plans[v].radii[0] <= M.max_r ? M.vol : M.big).push_back(v);
    }

    // The single persistent launch (the volumes above its radius cap run on
    // the branches, see enqueue).
    int enqueue_mega(hipStream_t s) const {
        const MegaPlan& M = mega;
        if (!M.active || M.nwork == 0) return ITKG_OK;
        const hipError_t e =
            M.family == FUSED_FWD_FAST
                ? mega_launch_fast(M.d_args, M.d_work, M.d_qstart, M.d_heads, M.grid, M.lds, s)
                : mega_launch_adj(M.d_args, M.d_work, M.d_qstart, M.d_heads, M.grid, M.lds, s);
        if (e != hipSuccess) return hip_fail(e, "batched persistent launch");
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
    // The whole plan on `s`: fork to the branches and enqueue their units
    // (per-axis volumes, EXACT groups), then the single persistent launch on
    // `s` itself, then join.  The branches' launches come first, so their
    // workgroups take CUs first; the persistent launch's workgroups start on
    // the CUs as they free up and simply take fewer of its items.
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
        }
        if (int rc = enqueue_mega(s)) return rc;
        for (size_t b = 0; b < branch_units.size(); ++b) {
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
        const bool adj = direction == ITKG_BATCH_BACKWARD;
        if (B.mode != FUSED_FWD_EXACT) {
            // FAST forward / adjoint: the single persistent launch
            const int64_t rad = p.radii[0];
            const bool iso = p.D == 3 && p.radii[1] == rad && p.radii[2] == rad;
            if (iso && rad >= 1 && rad <= (adj ? 15 : kFusedMaxRBatch) && fused_path_allowed() &&
                fits_i32(p) && fused_src_ok(p, vsrc, 1, adj)) {
                B.mega.elig.push_back(i);
                if (rad > 15) sc = std::max(sc, scratch_bytes(p, ITKG_F32, 1));  // above the cap
            } else {
                B.rest.push_back(i);
                sc = std::max(sc, scratch_bytes(p, ITKG_F32, 1));
            }
            continue;
        }
        const FusedEntry* fe =
            use_fused(p, ITKG_F32, 1) && fused_src_ok(p, vsrc, 1, adj) ? fused_for(p) : nullptr;
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
    // every branch may receive per-axis volumes (the autotuner re-deals
    // them); branch 0's scratch also serves the volumes above the single
    // launch's radius cap
    for (int k = 0; k < kBatchBranches && sc > 0; ++k) {
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
    if (!B.mega.elig.empty()) {
        MegaPlan& M = B.mega;
        M.active = true;
        M.family = B.mode;
        if (M.family == FUSED_FWD_FAST) mega_classes_fast(B.mega_lds_tab, B.mega_ty_tab);
        else mega_classes_adj(B.mega_lds_tab, B.mega_ty_tab);
        e = hipMalloc(&M.d_heads, (kMegaQueuesHost + 1) * sizeof(int));
        if (e == hipSuccess) e = hipMemset(M.d_heads, 0, (kMegaQueuesHost + 1) * sizeof(int));
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            delete b;
            return hip_fail(e, "batch queue heads");
        }
        auto it = cache.find({(int)direction, (int)arith, -1});
        if (it != cache.end()) {
            M.split_x100 = it->second.first;
            M.shape_policy = it->second.second;
        }
        auto itr = cache.find({(int)direction, (int)arith, -2});
        if (itr != cache.end()) M.max_r = itr->second.first;
        B.mega_assign();
        B.mega_pick();
        const int rc = B.mega_build();
        if (rc) {
            delete b;
            return rc;
        }
    }
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
    if (B.mega.active) {
        MegaPlan& M = B.mega;
        float best_ms = 1e30f;
        int best_split = M.split_x100, best_policy = M.shape_policy, best_maxr = M.max_r;
        int top = 1;
        for (int v : M.elig) top = std::max<int>(top, (int)B.plans[v].radii[0]);
        for (int maxr : {15, 17, kFusedMaxRBatch}) {
          // caps that change which volumes the launch serves
          if ((maxr == 17 && top <= 15) || (maxr == kFusedMaxRBatch && top <= 17)) continue;
          M.max_r = maxr;
          B.mega_assign();
          B.assign_branches();
          for (int policy : {0, 1, 2}) {
            for (int split : {75, 100, 150, 200, 300, 400, 600}) {
                M.split_x100 = split;
                M.shape_policy = policy;
                B.mega_pick();
                if ((rc = B.mega_build())) break;
                if ((rc = B.enqueue(s))) break;  // warm-up
                float tot = 0.0f;
                for (int rep = 0; rep < 3 && !rc; ++rep) {
                    (void)hipEventRecord(e0, s);
                    rc = B.enqueue(s);
                    (void)hipEventRecord(e1, s);
                    (void)hipEventSynchronize(e1);
                    float ms = 0.0f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    tot += ms;
                }
                if (rc) break;
                if (std::getenv("ITKG_DEBUG"))
                    std::fprintf(stderr,
                                 "itkg: autotune single launch max_r=%d policy=%d split=%d: %.4f ms\n",
                                 maxr, policy, split, tot / 3.0f);
                if (tot < best_ms) {
                    best_ms = tot;
                    best_split = split;
                    best_policy = policy;
                    best_maxr = maxr;
                }
            }
            if (rc) break;
          }
          if (rc) break;
        }
        if (!rc) {
            M.split_x100 = best_split;
            M.shape_policy = best_policy;
            M.max_r = best_maxr;
            B.mega_assign();
            B.assign_branches();
            B.mega_pick();
            if (std::getenv("ITKG_DEBUG") && !(rc = B.mega_build()) && M.nwork > 0) {
                // where the chosen plan's launch spends its time: per-item
                // start / end ticks, summed per class, and the idle share
                unsigned long long* d_tr = nullptr;
                std::vector<unsigned long long> tr((size_t)M.nwork * 3);
                if (hipMalloc(&d_tr, tr.size() * sizeof(unsigned long long)) == hipSuccess) {
                    for (const int v : M.big) (void)B.enqueue_volume(v, s, B.scratch[0]);
                    (void)(M.family == FUSED_FWD_FAST
                               ? mega_launch_fast(M.d_args, M.d_work, M.d_qstart, M.d_heads, M.grid,
                                                  M.lds, s, d_tr)
                               : mega_launch_adj(M.d_args, M.d_work, M.d_qstart, M.d_heads, M.grid,
                                                 M.lds, s, d_tr));
                    (void)hipStreamSynchronize(s);
                    (void)hipMemcpy(tr.data(), d_tr, tr.size() * sizeof(unsigned long long),
                                    hipMemcpyDeviceToHost);
                    (void)hipFree(d_tr);
                    std::vector<int4> work(M.nwork);
                    (void)hipMemcpy(work.data(), M.d_work, work.size() * sizeof(int4), hipMemcpyDeviceToHost);
                    unsigned long long t_lo = ~0ull, t_hi = 0;
                    for (int w = 0; w < M.nwork; ++w) {
                        t_lo = std::min(t_lo, tr[3 * w]);
                        t_hi = std::max(t_hi, tr[3 * w + 1]);
                    }
                    std::vector<double> busy(M.grid, 0.0), last(M.grid, 0.0);
                    std::map<int, std::tuple<double, double, int>> per;  // class: sum, max, n
                    for (int w = 0; w < M.nwork; ++w) {
                        const double d = (double)(tr[3 * w + 1] - tr[3 * w]) * 1e-5;  // ms at 100 MHz
                        const int g = (int)tr[3 * w + 2];
                        if (g >= 0 && g < M.grid) {
                            busy[g] += d;
                            last[g] = std::max(last[g], (double)(tr[3 * w + 1] - t_lo) * 1e-5);
                        }
                        auto& e = per[work[w].z];
                        std::get<0>(e) += d;
                        std::get<1>(e) = std::max(std::get<1>(e), d);
                        std::get<2>(e) += 1;
                    }
                    const double span = (double)(t_hi - t_lo) * 1e-5;
                    double tot = 0.0, first_idle = span;
                    for (int g = 0; g < M.grid; ++g) {
                        tot += busy[g];
                        if (busy[g] > 0) first_idle = std::min(first_idle, last[g]);
                    }
                    std::fprintf(stderr,
                                 "itkg: single launch trace: span %.4f ms, busy %.1f%% of %d "
                                 "workgroups, first workgroup idle at %.4f ms\n",
                                 span, 100.0 * tot / (span * M.grid), M.grid, first_idle);
                    for (const auto& kv : per)
                        std::fprintf(stderr, "itkg:   class %d: %d items, %.4f ms summed, longest %.4f ms\n",
                                     kv.first, std::get<2>(kv.second), std::get<0>(kv.second),
                                     std::get<1>(kv.second));
                }
            }
            rc = rc ? rc : B.mega_build();
            M.measured_ms = best_ms / 3.0f;
            append_cache({(int)B.dir, (int)B.arith, -1}, best_split, best_policy);
            append_cache({(int)B.dir, (int)B.arith, -2}, best_maxr, 0);
        }
    }
    for (BatchClass& c : B.classes) {
        if (rc) break;
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
    auto time_volume = [&](int v) {
        rc = B.enqueue_volume(v, s, B.scratch[0]);  // warm-up
        (void)hipEventRecord(e0, s);
        if (!rc) rc = B.enqueue_volume(v, s, B.scratch[0]);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return ms;
    };
    B.rest_ms.clear();
    B.big_ms.clear();
    for (size_t k = 0; k < B.rest.size() && !rc; ++k) B.rest_ms.push_back(time_volume(B.rest[k]));
    for (size_t k = 0; k < B.mega.big.size() && !rc; ++k) B.big_ms.push_back(time_volume(B.mega.big[k]));
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
    static const bool no_graph = [] {
        const char* g = std::getenv("ITKG_BATCH_GRAPH");
        return g && std::atoi(g) == 0;
    }();
    if (no_graph) return B.enqueue(s);
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
