// Scheduling core of the simulation server (no counterpart in the reference,
// which runs one synchronous vis_gpu call at a time, hera_gpu/vis.py:144).
//
// Everything here is host logic without HIP calls: the job queue, admission
// by device footprint, per-shape-bucket workspace instances, completion,
// the result copies into caller memory, pre-faulting of caller output pages,
// cancellation and shutdown.  The device work goes through an Executor:
// hv_serve.hip drives HIP streams, captured graphs and device buffers; the
// CPU tests drive a fake device, so this same code is built with g++ under
// ThreadSanitizer and AddressSanitizer/UndefinedBehaviorSanitizer.
//
// Instances.  A job runs in an *instance*: the device buffers of one shape
// bucket (mode, N, P, T, F, beam type, FP16 scale, 2^ceil(log2 Kpad),
// layout) at fixed addresses, plus whatever the executor keeps with them (a
// captured graph).  Instances outlive their jobs: a later job of the same
// bucket takes an idle instance on any slot and replays its graph.  Idle
// instances are evicted least recently used first when a new instance does
// not fit the budget.  A slot is a stream; a finished job's result sits in a
// host buffer that the slot hands to the copier, so the slot takes its next
// job while the result is still being copied into caller memory.
//
// Threads and locks.  The worker thread admits and launches jobs and polls
// completions; the copier thread copies finished results into caller memory;
// toucher threads pre-fault caller output pages.  Lock order: `mu`, then
// `touch_mu`; touchers take only `touch_mu`.  No thread waits for touchers
// (stop_touch) or for the device while holding `mu`.
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "heravis.h"

namespace hv {

int fail(int code, const std::string& msg);   // sets the calling thread's hv_last_error()

namespace sched {

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct BucketKey {
    int mode = 0, nant = 0, beam_px = 0, ntimes = 0, nfreq = 0, beam_f64 = 0;
    float scale = 1.0f;
    int64_t kbucket = 0;
    int layout = 0;
    int beam_per_freq = 0;
    int ksplit = 0;            // sources per K-split (fixed by the job's exact size)
    bool operator==(const BucketKey& o) const {
        return std::tie(mode, nant, beam_px, ntimes, nfreq, beam_f64, scale, kbucket, layout, beam_per_freq,
                        ksplit) ==
               std::tie(o.mode, o.nant, o.beam_px, o.ntimes, o.nfreq, o.beam_f64, o.scale, o.kbucket, o.layout,
                        o.beam_per_freq, o.ksplit);
    }
};

enum JobState { J_PENDING, J_RUNNING, J_FINISHING, J_DONE };

struct Job {
    int64_t id = 0;
    int flags = 0;
    BucketKey key;
    size_t footprint = 0;      // device bytes of an instance of the bucket
    size_t out_bytes = 0;      // bytes of the caller's result array
    void* out = nullptr;       // caller memory (may be NULL: no result wanted)
    double flops = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int status = HV_OK;
    int slot = -1;
    int graph_hit = 0;         // set by the executor: a captured graph was replayed
    int graph_captured = 0;    // set by the executor: a graph was captured for this job
    int skipped = 0;           // later jobs admitted ahead of it
    bool cancel = false;       // cancellation requested while running
    JobState state = J_PENDING;
    std::string err;
    void* payload = nullptr;   // the executor's inputs for the job
    // output pre-faulting (under touch_mu)
    size_t touch_next = 0;
    int touch_inflight = 0;
    bool touch_stop = false;
};

// The device side of the server.  Called from the worker thread only, except
// result_data() and release_result(), which the copier calls.
class Executor {
public:
    virtual ~Executor() = default;
    virtual void worker_init() {}
    // An instance for j's bucket: j.footprint device bytes.
    virtual int instance_create(const Job& j, void** inst) = 0;
    virtual void instance_destroy(void* inst) = 0;
    // Stage j's inputs into inst and enqueue its work and the copy of its
    // result to the slot's host buffer (asynchronous).
    virtual int launch(int slot, void* inst, Job& j) = 0;
    // HV_OK: the slot's last launch has finished; HV_ERR_BUSY: not yet; else
    // a device error (message set with fail()).
    virtual int query(int slot) = 0;
    // Wait for everything enqueued on the slot (after a failed launch).
    virtual void drain(int slot) = 0;
    // The host buffer holding the slot's last result (once query() said
    // HV_OK), handed to the caller: the slot's next launch uses another one.
    virtual void* take_result(int slot) = 0;
    virtual const void* result_data(void* result) = 0;
    virtual void release_result(void* result) = 0;
    // Release the job's payload.
    virtual void job_free(Job& j) = 0;
};

enum SlotState { S_FREE, S_RUNNING };

struct Instance {
    BucketKey key;
    void* handle = nullptr;
    size_t bytes = 0;
    bool busy = false;
    uint64_t last_use = 0;
};

struct Slot {
    int idx = 0;
    SlotState state = S_FREE;
    Job* job = nullptr;
    Instance* inst = nullptr;
};

// A finished job whose result waits for the copier, with the host buffer
// that holds it (the slot itself is free again).
struct CopyItem {
    Job* job = nullptr;
    void* result = nullptr;
};

constexpr size_t TOUCH_CHUNK = (size_t)32 << 20;   // bytes a toucher takes at a time
constexpr size_t TOUCH_MIN = (size_t)8 << 20;      // smaller outputs are not pre-faulted

struct Server {
    Executor* ex = nullptr;
    size_t budget = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv;        // worker wake-up
    std::condition_variable done_cv;   // job completion
    std::condition_variable copy_cv;   // copier wake-up
    std::deque<Job*> pending;
    std::deque<CopyItem> to_copy;
    std::unordered_map<int64_t, Job*> live;       // submitted, not yet collected
    std::deque<int64_t> finished_order;
    std::list<Instance> instances;
    size_t inst_bytes = 0;             // bytes of all instances (idle, busy, being created)
    uint64_t clock = 0;
    int64_t next_id = 1;
    bool stopping = false;
    hv_server_stats_t stats{};
    int64_t evictions = 0;
    std::thread worker, copier;
    std::mutex touch_mu;
    std::condition_variable touch_cv;
    std::deque<Job*> touch_q;
    std::vector<std::thread> touchers;
    bool touch_quit = false;
};

// --------------------------------------------------------------- pre-faulting
// Write one byte per 4 KiB page of queued jobs' outputs, front job first (all
// touchers share its chunks).  The bytes written are zeros.
inline void toucher_main(Server* s) {
    std::unique_lock<std::mutex> lk(s->touch_mu);
    for (;;) {
        s->touch_cv.wait(lk, [&] { return s->touch_quit || !s->touch_q.empty(); });
        if (s->touch_quit) break;
        Job* j = s->touch_q.front();
        if (j->touch_stop || j->touch_next >= j->out_bytes) {
            s->touch_q.pop_front();
            continue;
        }
        const size_t off = j->touch_next;
        const size_t n = std::min(TOUCH_CHUNK, j->out_bytes - off);
        j->touch_next += n;
        if (j->touch_next >= j->out_bytes) s->touch_q.pop_front();
        ++j->touch_inflight;
        volatile char* p = static_cast<char*>(j->out) + off;
        lk.unlock();
        for (size_t b = 0; b < n; b += 4096) p[b] = 0;
        lk.lock();
        if (--j->touch_inflight == 0) s->touch_cv.notify_all();
    }
}

// End the pre-faulting of j's output: on return no chunk is handed out any
// more and none is being written.  Called without `mu` held.
inline void stop_touch(Server* s, Job* j) {
    std::unique_lock<std::mutex> lk(s->touch_mu);
    j->touch_stop = true;
    auto it = std::find(s->touch_q.begin(), s->touch_q.end(), j);
    if (it != s->touch_q.end()) s->touch_q.erase(it);
    s->touch_cv.wait(lk, [&] { return j->touch_inflight == 0; });
}

// Record j as finished (mu held; its pre-faulting has been stopped).
inline void finish_locked(Server* s, Job* j) {
    j->t_done = now_s();
    j->state = J_DONE;
    s->finished_order.push_back(j->id);
    if (j->status == HV_OK) s->stats.completed++;
    else s->stats.failed++;
    s->done_cv.notify_all();
}

// Finish j from a thread holding `lk` on mu: release mu while the touchers
// drain, then record the result.
inline void finish_relock(Server* s, Job* j, std::unique_lock<std::mutex>& lk) {
    j->state = J_FINISHING;
    lk.unlock();
    stop_touch(s, j);
    lk.lock();
    finish_locked(s, j);
}

// ------------------------------------------------------------------ admission
struct Pick {
    Job* job = nullptr;
    Instance* inst = nullptr;        // idle instance of the job's bucket, or a new one
    bool fresh = false;              // inst is new (its handle must be created)
    std::vector<void*> victims;      // idle instances evicted to make room
};

// Choose a pending job for a free slot (mu held): the first job in arrival
// order whose bucket has an idle instance, or whose new instance fits the
// budget once idle instances are evicted (least recently used first); a job
// passed over 8 times blocks later ones.
inline bool pick_locked(Server* s, Pick& pk) {
    size_t idle_bytes = 0;
    for (auto& in : s->instances)
        if (!in.busy) idle_bytes += in.bytes;
    for (size_t i = 0; i < s->pending.size(); ++i) {
        Job* j = s->pending[i];
        Instance* reuse = nullptr;
        for (auto& in : s->instances)
            if (!in.busy && in.key == j->key) {
                reuse = &in;
                break;
            }
        const bool fits_now = s->inst_bytes + j->footprint <= s->budget;
        const bool fits_evicting = s->inst_bytes - idle_bytes + j->footprint <= s->budget;
        if (!reuse && !fits_now && !fits_evicting) {
            if (j->skipped >= 8) break;
            continue;
        }
        for (size_t k = 0; k < i; ++k) s->pending[k]->skipped++;
        s->pending.erase(s->pending.begin() + (ptrdiff_t)i);
        pk.job = j;
        if (reuse) {
            pk.inst = reuse;
            reuse->busy = true;
            reuse->last_use = ++s->clock;
            return true;
        }
        while (s->inst_bytes + j->footprint > s->budget) {
            auto victim = s->instances.end();
            for (auto it = s->instances.begin(); it != s->instances.end(); ++it)
                if (!it->busy && (victim == s->instances.end() || it->last_use < victim->last_use)) victim = it;
            if (victim == s->instances.end()) break;   // cannot happen: fits_evicting held
            pk.victims.push_back(victim->handle);
            s->inst_bytes -= victim->bytes;
            s->evictions++;
            s->instances.erase(victim);
        }
        s->instances.push_back(Instance{});
        Instance& in = s->instances.back();
        in.key = j->key;
        in.bytes = j->footprint;
        in.busy = true;
        in.last_use = ++s->clock;
        s->inst_bytes += in.bytes;
        pk.inst = &in;
        pk.fresh = true;
        return true;
    }
    return false;
}

inline void drop_instance_locked(Server* s, Instance* inst) {
    for (auto it = s->instances.begin(); it != s->instances.end(); ++it)
        if (&*it == inst) {
            s->inst_bytes -= it->bytes;
            s->instances.erase(it);
            return;
        }
}

inline size_t workspace_locked(Server* s) { return s->inst_bytes; }

inline int64_t running_locked(Server* s) {
    int64_t n = 0;
    for (auto& sl : s->slots) n += sl.state == S_RUNNING;
    return n;
}

inline void worker_main(Server* s) {
    s->ex->worker_init();
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool busy = false;
        // completions
        for (auto& sl : s->slots) {
            if (sl.state != S_RUNNING) continue;
            busy = true;
            const int q = s->ex->query(sl.idx);
            if (q == HV_ERR_BUSY) continue;
            Job* j = sl.job;
            sl.inst->busy = false;
            sl.inst->last_use = ++s->clock;
            sl.inst = nullptr;
            if (q != HV_OK || j->cancel) {
                j->status = q != HV_OK ? q : HV_ERR_CANCELED;
                j->err = q != HV_OK ? std::string(hv_last_error()) : "cancelled while running";
                sl.job = nullptr;
                sl.state = S_FREE;
                finish_relock(s, j, lk);
                continue;
            }
            // the result is in a host buffer now: the copier takes it and
            // the slot is free for the next job at once
            j->state = J_FINISHING;
            s->to_copy.push_back(CopyItem{j, s->ex->take_result(sl.idx)});
            sl.job = nullptr;
            sl.state = S_FREE;
            s->copy_cv.notify_one();
        }
        // admissions
        for (auto& sl : s->slots) {
            if (sl.state != S_FREE || s->pending.empty()) continue;
            Pick pk;
            if (!pick_locked(s, pk)) continue;
            Job* j = pk.job;
            j->t_admit = now_s();
            j->slot = sl.idx;
            j->state = J_RUNNING;
            sl.job = j;
            sl.inst = pk.inst;
            sl.state = S_RUNNING;
            busy = true;
            s->stats.peak_running = std::max<int64_t>(s->stats.peak_running, running_locked(s));
            const bool fresh = pk.fresh;
            Instance* inst = pk.inst;
            lk.unlock();
            for (void* v : pk.victims) s->ex->instance_destroy(v);
            int rc = HV_OK;
            bool touched = false;   // the executor created or launched something
            void* handle = fresh ? nullptr : inst->handle;
            if (j->flags & HV_JOB_FAIL_TEST) {
                rc = fail(HV_ERR_STATE, "failure injected by the caller (HV_JOB_FAIL_TEST)");
            } else {
                touched = true;
                if (fresh) rc = s->ex->instance_create(*j, &handle);
                if (rc == HV_OK) rc = s->ex->launch(sl.idx, handle, *j);
            }
            std::string msg = rc != HV_OK ? std::string(hv_last_error()) : std::string();
            if (rc != HV_OK && touched) {
                // leave the slot clean: drain its stream and drop the instance,
                // whose captured graph may be incomplete
                s->ex->drain(sl.idx);
                if (handle) s->ex->instance_destroy(handle);
            }
            lk.lock();
            if (fresh && rc == HV_OK) inst->handle = handle;
            if (j->graph_captured) s->stats.graph_captures++;
            if (rc == HV_OK && (j->graph_hit || j->graph_captured)) s->stats.graph_launches++;
            else if (rc == HV_OK) s->stats.direct_launches++;
            if (rc != HV_OK) {
                if (touched || fresh) drop_instance_locked(s, inst);
                else inst->busy = false;   // injected failure: the instance is untouched
                j->status = rc;
                j->err = msg;
                sl.job = nullptr;
                sl.inst = nullptr;
                sl.state = S_FREE;
                finish_relock(s, j, lk);
            }
        }
        s->stats.workspace_bytes = workspace_locked(s);
        if (s->stopping && s->pending.empty()) {
            bool any = false;
            for (auto& sl : s->slots) any |= sl.state != S_FREE;
            if (!any) break;
        }
        if (busy) s->cv.wait_for(lk, std::chrono::microseconds(20));
        else s->cv.wait_for(lk, std::chrono::milliseconds(50));
    }
}

// Copy a finished result into caller memory.  Large results are split over
// several threads: the copy into a fresh caller buffer is bound by first-touch
// page faults as much as by memory bandwidth, and both spread over cores.
inline void copy_result(void* dst, const void* src, size_t bytes) {
    constexpr size_t CHUNK_MIN = (size_t)16 << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(8, hw), bytes / CHUNK_MIN);
    if (nt <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t per = (bytes / nt + 4095) / 4096 * 4096;
    std::vector<std::thread> th;
    th.reserve(nt);
    for (size_t i = 0; i < nt; ++i) {
        const size_t off = i * per;
        if (off >= bytes) break;
        const size_t n = std::min(per, bytes - off);
        th.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, n); });
    }
    for (auto& t : th) t.join();
}

inline bool all_idle_locked(Server* s) {
    for (auto& sl : s->slots)
        if (sl.state != S_FREE) return false;
    return true;
}

inline void copier_main(Server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->copy_cv.wait(lk, [&] { return !s->to_copy.empty() || (s->stopping && s->pending.empty() && all_idle_locked(s)); });
        if (s->to_copy.empty()) break;
        const CopyItem it = s->to_copy.front();
        s->to_copy.pop_front();
        Job* j = it.job;
        lk.unlock();
        stop_touch(s, j);
        if (j->out) copy_result(j->out, s->ex->result_data(it.result), j->out_bytes);
        s->ex->release_result(it.result);
        lk.lock();
        finish_locked(s, j);
        s->cv.notify_one();
    }
}

// ------------------------------------------------------------------- the API
inline Server* create(Executor* ex, int nslots, size_t budget, int ntouchers) {
    auto s = std::make_unique<Server>();
    s->ex = ex;
    s->budget = budget;
    s->slots.resize((size_t)nslots);
    for (int i = 0; i < nslots; ++i) s->slots[(size_t)i].idx = i;
    Server* raw = s.release();
    raw->worker = std::thread(worker_main, raw);
    raw->copier = std::thread(copier_main, raw);
    for (int i = 0; i < ntouchers; ++i) raw->touchers.emplace_back(toucher_main, raw);
    return raw;
}

// Queue j (its fields filled by the caller: key, footprint, out, out_bytes,
// flops, flags, payload).  On failure the job is not queued and its payload
// stays with the caller.
inline int submit(Server* s, std::unique_ptr<Job> j, int64_t* id) {
    if (j->footprint > s->budget)
        return fail(HV_ERR_NOMEM, "job footprint of " + std::to_string(j->footprint) +
                                      " bytes exceeds the server budget of " + std::to_string(s->budget));
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return fail(HV_ERR_STATE, "server is shutting down");
    j->id = s->next_id++;
    j->t_submit = now_s();
    j->state = J_PENDING;
    *id = j->id;
    Job* raw = j.release();
    s->live[raw->id] = raw;
    if (raw->out && raw->out_bytes >= TOUCH_MIN) {
        std::lock_guard<std::mutex> tl(s->touch_mu);
        s->touch_q.push_back(raw);
        s->touch_cv.notify_all();
    }
    s->pending.push_back(raw);
    s->stats.submitted++;
    s->cv.notify_one();
    return HV_OK;
}

inline void fill_info(const Job* j, hv_job_info* info) {
    std::memset(info, 0, sizeof(hv_job_info));
    info->id = j->id;
    info->status = j->status;
    info->slot = j->slot;
    info->graph_hit = j->graph_hit;
    info->t_submit = j->t_submit;
    info->t_admit = j->t_admit;
    info->t_done = j->t_done;
    info->footprint = (int64_t)j->footprint;
    info->flops = j->flops;
    std::strncpy(info->error, j->err.c_str(), sizeof(info->error) - 1);
}

// Remove a finished job from the books (mu held) and release it.
inline void collect_locked(Server* s, Job* j) {
    s->live.erase(j->id);
    s->ex->job_free(*j);
    delete j;
}

inline int poll(Server* s, hv_job_info* infos, int max, int* n) {
    std::lock_guard<std::mutex> lk(s->mu);
    int k = 0;
    while (k < max && !s->finished_order.empty()) {
        const int64_t id = s->finished_order.front();
        s->finished_order.pop_front();
        auto it = s->live.find(id);
        if (it == s->live.end()) continue;   // already collected by wait
        fill_info(it->second, &infos[k]);
        collect_locked(s, it->second);
        ++k;
    }
    *n = k;
    return HV_OK;
}

inline int wait(Server* s, int64_t id, double timeout_s, hv_job_info* info) {
    std::unique_lock<std::mutex> lk(s->mu);
    if (id <= 0 || id >= s->next_id) return fail(HV_ERR_ARG, "unknown job id");
    auto pred = [&] {
        auto it = s->live.find(id);
        return it == s->live.end() || it->second->state == J_DONE;
    };
    if (timeout_s < 0) s->done_cv.wait(lk, pred);
    else if (!s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
        return fail(HV_ERR_BUSY, "job not finished within the timeout");
    auto it = s->live.find(id);
    if (it == s->live.end()) return fail(HV_ERR_ARG, "job already collected");
    if (info) fill_info(it->second, info);
    collect_locked(s, it->second);
    return HV_OK;
}

// Cancel a job: a queued job finishes at once with HV_ERR_CANCELED; a running
// job finishes with HV_ERR_CANCELED when its device work ends (the work is not
// interrupted) and its result is not copied.  HV_ERR_ARG if the job is not
// queued or running (its device work has ended, or it was collected).
inline int cancel(Server* s, int64_t id) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(id);
    if (it == s->live.end()) return fail(HV_ERR_ARG, "unknown or collected job id");
    Job* j = it->second;
    if (j->state == J_PENDING) {
        auto p = std::find(s->pending.begin(), s->pending.end(), j);
        if (p != s->pending.end()) s->pending.erase(p);
        j->status = HV_ERR_CANCELED;
        j->err = "cancelled before it started";
        s->stats.cancelled++;
        finish_relock(s, j, lk);
        return HV_OK;
    }
    if (j->state == J_RUNNING && !j->cancel) {
        j->cancel = true;
        s->stats.cancelled++;
        return HV_OK;
    }
    return fail(HV_ERR_ARG, "job is not queued or running");
}

inline int stats(Server* s, hv_server_stats_t* out) {
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->pending = (int64_t)s->pending.size();
    out->running = 0;
    for (auto& sl : s->slots) out->running += sl.state != S_FREE;
    out->workspace_bytes = s->inst_bytes;
    out->budget_bytes = s->budget;
    out->instances = (int64_t)s->instances.size();
    out->evictions = s->evictions;
    return HV_OK;
}

// Fail queued jobs, wait for running ones, stop the threads and release every
// instance.  Jobs not yet collected are released too.
inline void destroy(Server* s) {
    std::vector<Job*> dropped;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->stopping = true;
        for (Job* j : s->pending) {
            j->status = HV_ERR_STATE;
            j->err = "server destroyed before the job started";
            j->state = J_FINISHING;
            dropped.push_back(j);
        }
        s->pending.clear();
        lk.unlock();
        for (Job* j : dropped) stop_touch(s, j);
        lk.lock();
        for (Job* j : dropped) finish_locked(s, j);
        s->cv.notify_all();
        s->copy_cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->copy_cv.notify_all();
    }
    if (s->copier.joinable()) s->copier.join();
    {
        // every job has finished, which ended its pre-faulting
        std::lock_guard<std::mutex> lk(s->touch_mu);
        s->touch_quit = true;
        s->touch_cv.notify_all();
    }
    for (auto& t : s->touchers) t.join();
    for (auto& in : s->instances)
        if (in.handle) s->ex->instance_destroy(in.handle);
    s->instances.clear();
    for (auto& kv : s->live) {
        s->ex->job_free(*kv.second);
        delete kv.second;
    }
    s->live.clear();
    delete s;
}

}  // namespace sched
}  // namespace hv
