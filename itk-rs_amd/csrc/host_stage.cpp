// Host staging of the *_host entry points: caller buffers in pageable host
// memory are moved to and from the device in chunks, overlapped with the
// kernels of a piecewise computation and with each other.
//
// Pinned form (the default): a process-wide ring of pinned slots per
// direction, touched once when it is created (pinned pages are mapped on
// first touch, at a few GB/s), and a small pool of copy threads per
// direction.  Inbound, the copy threads fill a slot from the caller's
// buffer while the previous slot's DMA runs; outbound, the DMA of the next
// chunks runs while the copy threads empty a finished slot into the
// caller's buffer.  DMA from and to pinned memory runs at the link's rate in
// both directions at once; DMA from and to pageable memory goes through the
// driver's own staging at about half of it inbound and a third outbound.
//
// Pageable form (ITKG_HOST_STAGING=pageable, or when another call holds the
// ring): the same chunk schedule with the copies made straight from and to
// the caller's memory, each on its own thread.
//
// Ordering: the kernels of piece k wait (on the device) for the inbound
// chunk that holds the last input byte the piece reads; the outbound copy of
// a chunk waits for the piece that completes its last byte.  The host waits
// only where a slot is reused or a pageable copy needs its data in place.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "itkg_impl.hpp"

namespace itkg {

namespace {

// Measured on 512^3 f32 host to host (1 GiB over the link): 3 slots of
// 12 MiB per direction 13.4 ms, 4 of 8 MiB 14.3 ms, 4 of 16 MiB 15.3 ms, 8
// or 16 slots 19-21 ms (a deeper inbound queue holds the outbound DMA back),
// 2 slots 15.8-16.9 ms.  Copy threads per direction: f32 the same with 4,
// 6 or 8 (the DMA is the limit); the 512^3 u8 chain 6.1 ms with 4, 5.2 with
// 6 or 8 (copying out into fresh pages is the limit there).
constexpr size_t kStageChunk = size_t(12) << 20;  // bytes per chunk and per slot
constexpr int kStageSlots = 3;                     // slots per direction
constexpr int kStageMaxSlots = 16;
constexpr int kStageThreads = 6;                   // copy threads per direction

// ITKG_STAGE_THREADS / ITKG_STAGE_CHUNK_MIB / ITKG_STAGE_SLOTS override the
// copy threads per direction (1-16), the chunk size (1-256 MiB) and the
// slots per direction (2-16); read when the ring is created.
long env_long(const char* name, long def, long lo, long hi) {
    const char* v = std::getenv(name);
    if (!v || !*v) return def;
    char* end = nullptr;
    const long x = std::strtol(v, &end, 10);
    return (end && *end == '\0' && x >= lo && x <= hi) ? x : def;
}

// Runs fn(part, parts) for parts = workers + 1 (the calling thread takes
// part 0) and waits for all of them.  One job at a time per pool.
class CopyPool {
public:
    explicit CopyPool(int workers) {
        try {
            for (int i = 0; i < workers; ++i) th_.emplace_back([this, i] { loop(i + 1); });
        } catch (...) {
            stop_all();  // the threads already started, before the vector drops them
            throw;
        }
    }
    ~CopyPool() { stop_all(); }
    int parts() const { return (int)th_.size() + 1; }
    void run(const std::function<void(int, int)>& fn) {
        const int P = parts();
        if (P == 1) {
            fn(0, 1);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &fn;
            pending_ = P - 1;
            ++gen_;
        }
        cv_job_.notify_all();
        fn(0, P);
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
    }
    // dst[0, n) = src[0, n), split in page-aligned parts
    void copy(void* dst, const void* src, size_t n) {
        if (n < (size_t(1) << 20)) {
            std::memcpy(dst, src, n);
            return;
        }
        run([&](int part, int P) {
            const size_t per = (n / P + 4095) / 4096 * 4096;
            const size_t a = std::min(n, per * part), b = std::min(n, a + per);
            if (b > a) std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
        });
    }

private:
    void stop_all() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        cv_job_.notify_all();
        for (auto& t : th_)
            if (t.joinable()) t.join();
    }
    void loop(int part) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_job_.wait(lk, [&] { return quit_ || gen_ != seen; });
            if (quit_) return;
            seen = gen_;
            const std::function<void(int, int)>* fn = job_;
            const int P = parts();
            lk.unlock();
            (*fn)(part, P);
            lk.lock();
            if (--pending_ == 0) cv_done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_job_, cv_done_;
    const std::function<void(int, int)>* job_ = nullptr;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool quit_ = false;
};

struct Ring {
    void* slot[kStageMaxSlots] = {};
};

// The process-wide pinned rings and copy pools.  `busy` is held for the
// whole of a staged call; a call that finds it held stages pageable.
constexpr int kStageDevices = 64;

struct Stager {
    std::mutex busy;
    bool ready = false;
    // the copy streams of the pinned form, per device, created once and in
    // pairs: streams created per call were sometimes given one hardware
    // queue for both directions, which serialises the two DMA directions
    // (22 ms for 512^3 f32 host to host instead of 13.4)
    hipStream_t s_in[kStageDevices] = {}, s_out[kStageDevices] = {};
    size_t chunk = kStageChunk;  // slot size of the current ring
    int nslots = kStageSlots;    // slots per direction of the current ring
    Ring in, out;
    CopyPool* pool_in = nullptr;
    CopyPool* pool_out = nullptr;

    // with `busy` held
    hipError_t ensure() {
        if (ready) return hipSuccess;
        if (!pool_in) {
            const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
            const long t = env_long("ITKG_STAGE_THREADS", kStageThreads, 1, 16);
            const int w = std::max(0, std::min<int>((int)t, (int)hw / 2) - 1);
            CopyPool *a = nullptr, *b = nullptr;
            try {
                a = new CopyPool(w);
                b = new CopyPool(w);
            } catch (const std::exception&) {
                delete a;
                return hipErrorOutOfMemory;  // no threads: the caller stages pageable
            }
            pool_in = a;
            pool_out = b;
        }
        chunk = (size_t)env_long("ITKG_STAGE_CHUNK_MIB", (long)(kStageChunk >> 20), 1, 256) << 20;
        nslots = (int)env_long("ITKG_STAGE_SLOTS", kStageSlots, 2, kStageMaxSlots);
        const size_t C = chunk;
        for (Ring* r : {&in, &out})
            for (int i = 0; i < nslots; ++i)
                if (void*& s = r->slot[i]; !s) {
                    hipError_t e = hipHostMalloc(&s, C, hipHostMallocDefault);
                    if (e != hipSuccess) {
                        s = nullptr;
                        release();
                        return e;
                    }
                    // map every page now, in parallel: first touch is slow
                    void* p = s;
                    pool_in->run([&](int part, int P) {
                        const size_t per = (C / P + 4095) / 4096 * 4096;
                        const size_t a = std::min(C, per * part);
                        const size_t b = std::min(C, a + per);
                        if (b > a) std::memset(static_cast<char*>(p) + a, 0, b - a);
                    });
                }
        ready = true;
        return hipSuccess;
    }
    // with `busy` held
    void release() {
        for (Ring* r : {&in, &out})
            for (auto& s : r->slot)
                if (s) {
                    (void)hipHostFree(s);
                    s = nullptr;
                }
        ready = false;
    }
};

Stager& stager() {
    static Stager* g = new Stager();  // lives until the process ends
    return *g;
}

bool pageable_forced() {
    const char* v = std::getenv("ITKG_HOST_STAGING");
    return v && std::strcmp(v, "pageable") == 0;
}

}  // namespace

int stage_trim() {
    Stager& g = stager();
    std::unique_lock<std::mutex> lk(g.busy, std::try_to_lock);
    if (!lk.owns_lock()) return ITKG_OK;  // in use: kept
    g.release();
    return ITKG_OK;
}

namespace {

// One piece and at most a chunk each way: copy in, compute, copy out on the
// caller's stream, with no helper threads or streams.
int run_serial(const StagedJob& job, hipStream_t s) {
    hipError_t e = hipSuccess;
    if (job.in_bytes)
        e = hipMemcpyAsync(job.dev_in, job.host_in, job.in_bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "copy in");
    int rc = job.compute(0, s);
    if (!rc && job.out_bytes) {
        e = hipMemcpyAsync(job.host_out, job.dev_out, job.out_bytes, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = hip_fail(e, "copy out");
    }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess && !rc) rc = hip_fail(e, "hipStreamSynchronize");
    if (job.stats) job.stats->pinned = false;
    return rc;
}

}  // namespace

int run_staged(const StagedJob& job, hipStream_t s_comp) {
    const int K = job.pieces;
    if (K == 1 && job.in_bytes <= kStageChunk && job.out_bytes <= kStageChunk)
        return run_serial(job, s_comp);
    Stager& g = stager();
    std::unique_lock<std::mutex> held(g.busy, std::defer_lock);
    bool pinned = false;
    if (!pageable_forced() && held.try_lock()) {
        if (g.ensure() == hipSuccess) pinned = true;
        else held.unlock();
    }
    const size_t C = pinned ? g.chunk : kStageChunk;
    const int S = pinned ? g.nslots : kStageSlots;
    const int nci = (int)((job.in_bytes + C - 1) / C);
    const int nco = (int)((job.out_bytes + C - 1) / C);
    // the inbound chunk each piece waits for, and the piece each outbound
    // chunk waits for
    std::vector<int> wait_in(K), wait_piece(nco);
    for (int k = 0; k < K; ++k)
        wait_in[k] = job.in_need[k] ? (int)((job.in_need[k] - 1) / C) : -1;
    for (int c = 0, k = 0; c < nco; ++c) {
        const size_t end = std::min(job.out_bytes, (size_t)(c + 1) * C);
        while (k < K - 1 && job.out_end[k] < end) ++k;
        wait_piece[c] = k;
    }

    StageStats st;
    st.pinned = pinned;
    const bool debug = std::getenv("ITKG_DEBUG") != nullptr;
    const auto t_start = std::chrono::steady_clock::now();
    auto ms_since = [&]() {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
            .count();
    };
    double t_in_end = 0, t_comp_end = 0, t_out_end = 0;
    double out_dma_wait = 0, out_copy = 0;  // outbound: waiting for the DMA, copying out

    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kStageDevices) return fail(ITKG_ERR_UNSUPPORTED, "device %d", dev);
    hipStream_t s_in = nullptr, s_out = nullptr;
    const bool own_streams = !pinned;  // the pinned form's streams belong to the ring
    std::vector<hipEvent_t> ev_in(nci, nullptr), ev_out(nco, nullptr), ev_comp(K, nullptr);
    auto cleanup = [&]() {
        for (hipStream_t t : {s_in, s_comp, s_out})
            if (t) (void)hipStreamSynchronize(t);
        for (auto* v : {&ev_in, &ev_out, &ev_comp})
            for (auto ev : *v)
                if (ev) (void)hipEventDestroy(ev);
        if (own_streams)
            for (hipStream_t t : {s_in, s_out})
                if (t) (void)hipStreamDestroy(t);
    };
    if (pinned) {
        for (hipStream_t* t : {&g.s_in[dev], &g.s_out[dev]})
            if (e == hipSuccess && !*t) e = hipStreamCreateWithFlags(t, hipStreamNonBlocking);
        s_in = g.s_in[dev];
        s_out = g.s_out[dev];
    } else {
        for (hipStream_t* t : {&s_in, &s_out})
            if (e == hipSuccess) e = hipStreamCreateWithFlags(t, hipStreamNonBlocking);
    }
    for (auto* v : {&ev_in, &ev_out, &ev_comp})
        for (auto& ev : *v)
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        const int rc = hip_fail(e, "host staging set-up");
        cleanup();
        return rc;
    }

    std::mutex mu;
    std::condition_variable cv;
    int in_rec = 0, comp_rec = 0;  // inbound chunks / pieces whose completion event is recorded
    bool abort = false;
    int rc_in = ITKG_OK, rc_out = ITKG_OK;
    std::string err_in, err_out;
    auto publish = [&](int& counter, int value) {
        {
            std::lock_guard<std::mutex> lk(mu);
            counter = value;
        }
        cv.notify_all();
    };
    auto stop = [&]() {
        {
            std::lock_guard<std::mutex> lk(mu);
            abort = true;
        }
        cv.notify_all();
    };
    auto wait_for = [&](const int& counter, int idx) {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return abort || counter > idx; });
        return !abort;
    };
    auto ready_now = [&](const int& counter, int idx) {
        std::lock_guard<std::mutex> lk(mu);
        return counter > idx;
    };
    const char* src = static_cast<const char*>(job.host_in);
    char* dst = static_cast<char*>(job.host_out);
    char* din = static_cast<char*>(job.dev_in);
    const char* dout = static_cast<const char*>(job.dev_out);

    auto in_fn = [&]() {
        hipError_t te = hipSetDevice(dev);
        for (int c = 0; c < nci && te == hipSuccess; ++c) {
            {
                std::lock_guard<std::mutex> lk(mu);
                if (abort) return;
            }
            const size_t off = (size_t)c * C, len = std::min(C, job.in_bytes - off);
            if (pinned) {
                void* slot = g.in.slot[c % S];
                if (c >= S) te = hipEventSynchronize(ev_in[c - S]);
                if (te != hipSuccess) break;
                g.pool_in->copy(slot, src + off, len);
                te = hipMemcpyAsync(din + off, slot, len, hipMemcpyHostToDevice, s_in);
            } else {
                te = hipMemcpyAsync(din + off, src + off, len, hipMemcpyHostToDevice, s_in);
            }
            if (te == hipSuccess) te = hipEventRecord(ev_in[c], s_in);
            if (te == hipSuccess) publish(in_rec, c + 1);
        }
        t_in_end = ms_since();
        if (te != hipSuccess) {
            rc_in = hip_fail(te, "copy in");
            err_in = last_error_text();
            stop();
        }
    };
    auto out_fn = [&]() {
        hipError_t te = hipSetDevice(dev);
        auto issue = [&](int c) {
            const size_t off = (size_t)c * C, len = std::min(C, job.out_bytes - off);
            hipError_t x = hipStreamWaitEvent(s_out, ev_comp[wait_piece[c]], 0);
            if (x == hipSuccess)
                x = hipMemcpyAsync(g.out.slot[c % S], dout + off, len,
                                   hipMemcpyDeviceToHost, s_out);
            if (x == hipSuccess) x = hipEventRecord(ev_out[c], s_out);
            return x;
        };
        if (pinned) {
            int next = 0;  // next chunk to copy to a slot
            for (int c = 0; c < nco && te == hipSuccess; ++c) {
                while (te == hipSuccess && next < nco && next < c + S) {
                    if (!ready_now(comp_rec, wait_piece[next])) {
                        if (next > c) break;  // empty the slots already filled first
                        if (!wait_for(comp_rec, wait_piece[next])) return;
                    }
                    te = issue(next++);
                }
                const double w0 = debug ? ms_since() : 0.0;
                if (te == hipSuccess) te = hipEventSynchronize(ev_out[c]);
                if (te != hipSuccess) break;
                const double w1 = debug ? ms_since() : 0.0;
                const size_t off = (size_t)c * C, len = std::min(C, job.out_bytes - off);
                g.pool_out->copy(dst + off, g.out.slot[c % S], len);
                if (debug) {
                    out_dma_wait += w1 - w0;
                    out_copy += ms_since() - w1;
                }
            }
        } else {
            for (int c = 0; c < nco && te == hipSuccess; ++c) {
                if (!wait_for(comp_rec, wait_piece[c])) return;
                const size_t off = (size_t)c * C, len = std::min(C, job.out_bytes - off);
                // a copy to pageable memory is made once its data is in place
                te = hipEventSynchronize(ev_comp[wait_piece[c]]);
                if (te == hipSuccess)
                    te = hipMemcpyAsync(dst + off, dout + off, len, hipMemcpyDeviceToHost, s_out);
                if (te == hipSuccess) te = hipStreamSynchronize(s_out);
            }
        }
        if (te == hipSuccess) te = hipStreamSynchronize(s_out);
        t_out_end = ms_since();
        if (te != hipSuccess) {
            rc_out = hip_fail(te, "copy out");
            err_out = last_error_text();
            stop();
        }
    };
    std::thread t_in, t_out;
    try {
        t_in = std::thread(in_fn);
        t_out = std::thread(out_fn);
    } catch (const std::system_error&) {
        stop();
        if (t_in.joinable()) t_in.join();
        cleanup();
        return fail(ITKG_ERR_OOM, "could not start the host staging threads");
    }

    int rc = ITKG_OK;
    for (int k = 0; k < K && !rc; ++k) {
        if (wait_in[k] >= 0) {
            if (!wait_for(in_rec, wait_in[k])) break;
            e = hipStreamWaitEvent(s_comp, ev_in[wait_in[k]], 0);
            if (e != hipSuccess) {
                rc = hip_fail(e, "hipStreamWaitEvent");
                break;
            }
        }
        rc = job.compute(k, s_comp);
        if (!rc) {
            e = hipEventRecord(ev_comp[k], s_comp);
            if (e != hipSuccess) rc = hip_fail(e, "hipEventRecord");
        }
        if (!rc) publish(comp_rec, k + 1);
    }
    if (rc) stop();
    t_comp_end = ms_since();
    t_in.join();
    t_out.join();
    if (debug)
        std::fprintf(stderr,
                     "[itkg] staged %s: %d pieces, %d+%d chunks of %zu MiB; inbound issued "
                     "%.2f ms, pieces enqueued %.2f ms, outbound done %.2f ms (waiting for its "
                     "DMA %.2f ms, copying out %.2f ms), total %.2f ms\n",
                     pinned ? "pinned" : "pageable", K, nci, nco, C >> 20, t_in_end, t_comp_end,
                     t_out_end, out_dma_wait, out_copy, ms_since());
    if (!rc && rc_in) rc = fail(rc_in, "%s", err_in.c_str());
    if (!rc && rc_out) rc = fail(rc_out, "%s", err_out.c_str());
    cleanup();
    if (job.stats) *job.stats = st;
    return rc;
}

}  // namespace itkg
