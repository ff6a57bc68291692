// Continuous-batching simulation server (no counterpart in the reference,
// which runs one synchronous vis_gpu call at a time, hera_gpu/vis.py:144).
//
// Jobs name a resident sky catalog (first nsrc sources) and a resident beam,
// and bring their own antenna positions, times and frequencies.  A worker
// thread admits queued jobs into free slots as slots and device memory free
// up; each slot has its own stream, workspace and pinned result buffer, so
// admitted jobs run concurrently on the device.  Per slot, the kernel
// sequence of a job is captured once per shape bucket
//   (mode, N, P, T, F, beam type, FP16 scale, 2^ceil(log2 Kpad))
// into a hipGraph and replayed for later jobs of the bucket: the job's own
// inputs are staged into the slot's fixed buffers (outside the graph) and
// its true source count is read by the kernels from device memory, so a
// replay does exactly the arithmetic of a stand-alone run and results are
// bitwise identical alone and in a mix.  A copier thread moves results from
// pinned memory to the caller's array, so large copies do not hold up
// admission.  Toucher threads pre-fault the caller's output pages from
// submission on (one write per 4 KiB page), so the final copy, which would
// otherwise pay for first-touch faults and page zeroing, runs at memory speed
// while the device works on this or earlier jobs.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <vector>
#include <memory>
#include <thread>
#include <tuple>
#include <unordered_map>

#include "hv_host.h"

using namespace hv;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Operand bytes one serving job may hold at once (one batch of units): four
// HERA-350 units of 2^20 sources per launch.
constexpr size_t SERVE_OPERAND_BUDGET = (size_t)12 << 30;

struct BucketKey {
    int mode, nant, beam_px, ntimes, nfreq, beam_f64;
    float scale;
    int64_t kbucket;
    int layout;
    bool beam_per_freq;
    bool operator<(const BucketKey& o) const {
        return std::tie(mode, nant, beam_px, ntimes, nfreq, beam_f64, scale, kbucket, layout, beam_per_freq) <
               std::tie(o.mode, o.nant, o.beam_px, o.ntimes, o.nfreq, o.beam_f64, o.scale, o.kbucket, o.layout,
                        o.beam_per_freq);
    }
};

struct SrvJob {
    int64_t id = 0;
    int mode = 0, flags = 0;
    const hv_sky* sky = nullptr;
    const hv_beam* beam = nullptr;
    int64_t nsrc = 0, kpad = 0;
    int nant = 0, ntimes = 0, nfreq = 0;
    int layout = 0;            // output layout from the job flags
    std::vector<double> antpos, invlam, eq2tops;
    void* out = nullptr;
    BucketKey key{};
    Plan plan;                 // planned at the bucket's padded size
    size_t footprint = 0;      // device bytes of a slot sized for the bucket
    size_t out_bytes = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int status = HV_OK;
    int slot = -1;
    int graph_hit = 0;
    int skipped = 0;           // times later jobs were admitted ahead of it
    std::string err;
    // output pre-faulting (all under hv_server::touch_mu)
    size_t touch_next = 0;     // next output byte handed to a toucher
    int touch_inflight = 0;    // chunks being touched now
    bool touch_stop = false;   // no further chunks: the result copy or the job's end is due
};

enum SlotState { FREE, RUNNING, COPYING };

struct Slot {
    int idx = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Workspace ws;
    void* pinned = nullptr;
    size_t pinned_cap = 0;
    int64_t* dyn_host = nullptr;       // pinned source count staged to the device
    std::map<BucketKey, hipGraphExec_t> graphs;
    SrvJob* job = nullptr;
    SlotState state = FREE;
};

size_t slot_bytes_needed(const SrvJob& j) {
    const Plan& p = j.plan;
    const size_t beam_bytes = (size_t)j.beam->nfreq * j.nant * j.key.beam_px * j.key.beam_px * (j.key.beam_f64 ? 8 : 4);
    return p.footprint() + (size_t)4 * p.kpad * 8 + beam_bytes + (size_t)j.nant * 24 + (size_t)j.ntimes * 72 +
           (size_t)j.nfreq * 8 + 256 * 6;
}

}  // namespace

struct hv_server {
    hv_ctx* ctx = nullptr;
    size_t budget = 0;
    bool use_graphs = true;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv;        // worker wake-up
    std::condition_variable done_cv;   // job completion
    std::condition_variable copy_cv;   // copier wake-up
    std::deque<SrvJob*> pending;
    std::deque<Slot*> to_copy;
    std::unordered_map<int64_t, SrvJob*> finished;   // completed, not yet collected
    std::map<BucketKey, int> seen;     // jobs admitted per shape bucket
    std::deque<int64_t> finished_order;
    int64_t next_id = 1;
    bool stopping = false;
    hv_server_stats_t stats{};
    std::thread worker, copier;
    // output pre-faulting (lock order: mu, then touch_mu; touchers take only touch_mu)
    std::mutex touch_mu;
    std::condition_variable touch_cv;
    std::deque<SrvJob*> touch_q;       // jobs with output pages left to pre-fault
    std::vector<std::thread> touchers;
    bool touch_quit = false;
};

constexpr size_t TOUCH_CHUNK = (size_t)32 << 20;   // bytes a toucher takes at a time
constexpr size_t TOUCH_MIN = (size_t)8 << 20;      // smaller outputs are not pre-faulted

// Write one byte per 4 KiB page of the chunks of queued jobs' outputs, front
// job first (all touchers share its chunks).
static void toucher_main(hv_server* s) {
    std::unique_lock<std::mutex> lk(s->touch_mu);
    for (;;) {
        s->touch_cv.wait(lk, [&] { return s->touch_quit || !s->touch_q.empty(); });
        if (s->touch_quit) break;
        SrvJob* j = s->touch_q.front();
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

// End the pre-faulting of a job's output: no chunk is handed out after this
// returns and none is being written.  Called before the result copy and
// before a job is reported finished (its caller may then free the array).
static void stop_touch(hv_server* s, SrvJob* j) {
    std::unique_lock<std::mutex> lk(s->touch_mu);
    j->touch_stop = true;
    auto it = std::find(s->touch_q.begin(), s->touch_q.end(), j);
    if (it != s->touch_q.end()) s->touch_q.erase(it);
    s->touch_cv.wait(lk, [&] { return j->touch_inflight == 0; });
}

static size_t workspace_total(hv_server* s) {
    size_t t = 0;
    for (auto& sl : s->slots) t += sl->ws.bytes();
    return t;
}

static void clear_graphs(Slot& sl) {
    for (auto& kv : sl.graphs) (void)hipGraphExecDestroy(kv.second);
    sl.graphs.clear();
}

// Stage a job's inputs into the slot and launch its kernels (graph replay
// or direct) followed by the result copy to pinned memory.  Runs on the
// worker thread without the server lock.
static int launch_job(hv_server* s, Slot& sl, SrvJob& j) {
    hv_ctx* c = s->ctx;
    HV_HIP(hipSetDevice(c->device));
    const Plan& p = j.plan;
    bool moved = false;
    HV_CHECK(prepare_workspace(c, sl.ws, p, sl.stream, &moved));
    const size_t es = j.key.beam_f64 ? 8 : 4;
    const size_t beam_bytes = (size_t)j.beam->nfreq * j.nant * j.key.beam_px * j.key.beam_px * es;
    HV_CHECK(ensure(c, sl.ws, B_CRD_EQ, (size_t)3 * p.kpad * 8, sl.stream, &moved));
    HV_CHECK(ensure(c, sl.ws, B_ISQRT, (size_t)p.kpad * 8, sl.stream, &moved));
    HV_CHECK(ensure(c, sl.ws, B_BEAM, beam_bytes, sl.stream, &moved));
    HV_CHECK(ensure(c, sl.ws, B_ANTPOS, (size_t)j.nant * 24, sl.stream, &moved));
    HV_CHECK(ensure(c, sl.ws, B_EQ2TOPS, (size_t)j.ntimes * 72, sl.stream, &moved));
    HV_CHECK(ensure(c, sl.ws, B_INVLAM, (size_t)j.nfreq * 8, sl.stream, &moved));
    HV_CHECK(ensure(c, sl.ws, B_OUT, p.out_bytes, sl.stream, &moved));
    HV_CHECK(ensure(c, sl.ws, B_DYN, 256, sl.stream, &moved));
    if (moved) clear_graphs(sl);   // captured graphs hold the old addresses
    if (sl.pinned_cap < p.out_bytes) {
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        sl.pinned = nullptr;
        sl.pinned_cap = 0;
        HV_HIP(hipHostMalloc(&sl.pinned, p.out_bytes, hipHostMallocDefault));
        sl.pinned_cap = p.out_bytes;
    }

    // ---- stage inputs (outside any graph)
    Workspace& w = sl.ws;
    double* crd = (double*)w.buf[B_CRD_EQ].p;
    double* isq = (double*)w.buf[B_ISQRT].p;
    for (int d = 0; d < 3; ++d)
        HV_HIP(hipMemcpyAsync(crd + d * p.kpad, j.sky->crd_eq + d * j.sky->kpad, j.nsrc * 8,
                              hipMemcpyDeviceToDevice, sl.stream));
    HV_HIP(hipMemcpyAsync(isq, j.sky->isqrt, j.nsrc * 8, hipMemcpyDeviceToDevice, sl.stream));
    if (j.kpad > j.nsrc) {
        for (int d = 0; d < 3; ++d)
            HV_HIP(hipMemsetAsync(crd + d * p.kpad + j.nsrc, 0, (j.kpad - j.nsrc) * 8, sl.stream));
        HV_HIP(hipMemsetAsync(isq + j.nsrc, 0, (j.kpad - j.nsrc) * 8, sl.stream));
    }
    HV_HIP(hipMemcpyAsync(w.buf[B_BEAM].p, j.beam->data, beam_bytes, hipMemcpyDeviceToDevice, sl.stream));
    HV_HIP(hipMemcpyAsync(w.buf[B_ANTPOS].p, j.antpos.data(), j.antpos.size() * 8, hipMemcpyHostToDevice,
                          sl.stream));
    HV_HIP(hipMemcpyAsync(w.buf[B_EQ2TOPS].p, j.eq2tops.data(), j.eq2tops.size() * 8, hipMemcpyHostToDevice,
                          sl.stream));
    HV_HIP(hipMemcpyAsync(w.buf[B_INVLAM].p, j.invlam.data(), j.invlam.size() * 8, hipMemcpyHostToDevice,
                          sl.stream));
    *sl.dyn_host = j.kpad;
    HV_HIP(hipMemcpyAsync(w.buf[B_DYN].p, sl.dyn_host, 8, hipMemcpyHostToDevice, sl.stream));

    JobDev jd{};
    jd.mode = j.mode;
    jd.nant = j.nant;
    jd.ntimes = j.ntimes;
    jd.nfreq = j.nfreq;
    jd.beam_px = j.key.beam_px;
    jd.beam_f64 = j.key.beam_f64;
    jd.kpad = p.kpad;
    jd.antpos = (const double*)w.buf[B_ANTPOS].p;
    jd.inv_lambda = (const double*)w.buf[B_INVLAM].p;
    jd.eq2tops = (const double*)w.buf[B_EQ2TOPS].p;
    jd.crd_eq = crd;
    jd.isqrt = isq;
    jd.beam = w.buf[B_BEAM].p;
    jd.vis = w.buf[B_OUT].p;
    jd.amp_scale = j.key.scale;
    jd.dyn_kpad = (const int64_t*)w.buf[B_DYN].p;
    jd.layout = j.layout;
    jd.beam_fstride = j.beam->nfreq > 1 ? (size_t)j.nant * j.key.beam_px * j.key.beam_px : 0;

    // A bucket is captured the second time it is seen: a shape that comes
    // once runs direct (capture and instantiation of a large job's graph
    // costs more than its launches), a recurring one is replayed.
    bool recurring = false;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        recurring = s->seen[j.key]++ > 0;
    }
    StageTimer tm;
    auto it = sl.graphs.find(j.key);
    if (s->use_graphs && (it != sl.graphs.end() || recurring)) {
        if (it == sl.graphs.end()) {
            hipGraph_t g = nullptr;
            HV_HIP(hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_job(jd, p, w, sl.stream, tm);
            hipError_t e = hipStreamEndCapture(sl.stream, &g);
            if (rc != HV_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(HV_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            hipGraphExec_t ex = nullptr;
            e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(HV_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            it = sl.graphs.emplace(j.key, ex).first;
            std::lock_guard<std::mutex> lk(s->mu);
            s->stats.graph_captures++;
        } else {
            j.graph_hit = 1;
        }
        HV_HIP(hipGraphLaunch(it->second, sl.stream));
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.graph_launches++;
    } else {
        HV_CHECK(enqueue_job(jd, p, w, sl.stream, tm));
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.direct_launches++;
    }
    HV_HIP(hipMemcpyAsync(sl.pinned, w.buf[B_OUT].p, p.out_bytes, hipMemcpyDeviceToHost, sl.stream));
    HV_HIP(hipEventRecord(sl.done, sl.stream));
    return HV_OK;
}

static void finish_locked(hv_server* s, SrvJob* j) {
    stop_touch(s, j);
    j->t_done = now_s();
    s->finished[j->id] = j;
    s->finished_order.push_back(j->id);
    if (j->status == HV_OK) s->stats.completed++;
    else s->stats.failed++;
    s->done_cv.notify_all();
}

// Pick the job to admit into a free slot: first pending job (in arrival
// order) whose slot footprint fits the memory left; a job passed over 8
// times blocks later ones until it fits.
static SrvJob* pick_locked(hv_server* s, Slot& sl, size_t* grow) {
    const size_t total = workspace_total(s);
    for (size_t i = 0; i < s->pending.size(); ++i) {
        SrvJob* j = s->pending[i];
        const size_t need = slot_bytes_needed(*j);
        const size_t have = sl.ws.bytes();
        const size_t extra = need > have ? need - have : 0;
        if (total + extra <= s->budget) {
            for (size_t k = 0; k < i; ++k) s->pending[k]->skipped++;
            s->pending.erase(s->pending.begin() + i);
            *grow = extra;
            return j;
        }
        if (j->skipped >= 8) break;
    }
    return nullptr;
}

static void worker_main(hv_server* s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool busy = false;
        // completions
        for (auto& up : s->slots) {
            Slot& sl = *up;
            if (sl.state != RUNNING) continue;
            busy = true;
            hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                sl.job->status = HV_ERR_HIP;
                sl.job->err = std::string("device error: ") + hipGetErrorString(q);
                finish_locked(s, sl.job);
                sl.job = nullptr;
                sl.state = FREE;
                continue;
            }
            sl.state = COPYING;
            s->to_copy.push_back(&sl);
            s->copy_cv.notify_one();
        }
        // admissions
        for (auto& up : s->slots) {
            Slot& sl = *up;
            if (sl.state != FREE || s->pending.empty()) continue;
            size_t grow = 0;
            SrvJob* j = pick_locked(s, sl, &grow);
            if (!j) {
                // make room: drop the workspaces of idle slots, then retry once
                bool freed = false;
                for (auto& other : s->slots)
                    if (other->state == FREE && other->ws.bytes() > 0) {
                        clear_graphs(*other);
                        release(s->ctx, other->ws);
                        freed = true;
                    }
                if (freed) j = pick_locked(s, sl, &grow);
                if (!j) continue;
            }
            j->t_admit = now_s();
            j->slot = sl.idx;
            sl.job = j;
            sl.state = RUNNING;
            busy = true;
            s->stats.peak_running = std::max<int64_t>(
                s->stats.peak_running,
                (int64_t)std::count_if(s->slots.begin(), s->slots.end(),
                                       [](const std::unique_ptr<Slot>& x) { return x->state == RUNNING; }));
            int rc;
            if (j->flags & HV_JOB_FAIL_TEST) {
                rc = fail(HV_ERR_STATE, "failure injected by the caller (HV_JOB_FAIL_TEST)");
            } else {
                lk.unlock();
                rc = launch_job(s, sl, *j);
                lk.lock();
            }
            if (rc != HV_OK) {
                j->status = rc;
                j->err = hv_last_error();
                // leave the slot clean: drain its stream and drop its graphs
                (void)hipStreamSynchronize(sl.stream);
                clear_graphs(sl);
                finish_locked(s, j);
                sl.job = nullptr;
                sl.state = FREE;
            }
        }
        s->stats.workspace_bytes = workspace_total(s);
        if (s->stopping && s->pending.empty()) {
            bool any = false;
            for (auto& up : s->slots) any |= up->state != FREE;
            if (!any) break;
        }
        if (busy) s->cv.wait_for(lk, std::chrono::microseconds(20));
        else s->cv.wait_for(lk, std::chrono::milliseconds(50));
    }
}

// Copy a finished result from pinned memory into the caller's buffer.  Large
// results are split over several threads: the copy of a fresh caller buffer
// is bound by first-touch page faults as much as by memory bandwidth, and
// both spread over cores.
static void copy_result(void* dst, const void* src, size_t bytes) {
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

static void copier_main(hv_server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->copy_cv.wait(lk, [&] { return !s->to_copy.empty() || (s->stopping && s->pending.empty() && [&] {
                                             for (auto& up : s->slots)
                                                 if (up->state != FREE) return false;
                                             return true;
                                         }()); });
        if (s->to_copy.empty()) break;
        Slot* sl = s->to_copy.front();
        s->to_copy.pop_front();
        SrvJob* j = sl->job;
        lk.unlock();
        stop_touch(s, j);
        if (j->out) copy_result(j->out, sl->pinned, j->out_bytes);
        lk.lock();
        finish_locked(s, j);
        sl->job = nullptr;
        sl->state = FREE;
        s->cv.notify_one();
    }
}

// ------------------------------------------------------------------ C ABI
extern "C" int hv_server_create(hv_ctx* c, int nslots, size_t budget, int use_graphs, hv_server** out) {
    if (!c || !out) return fail(HV_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (nslots < 1 || nslots > 16) return fail(HV_ERR_ARG, "nslots must be in [1, 16]");
    HV_HIP(hipSetDevice(c->device));
    auto s = std::make_unique<hv_server>();
    s->ctx = c;
    {
        std::lock_guard<std::mutex> lk(c->mem_mu);
        const size_t left = c->budget - c->used;
        s->budget = budget ? std::min(budget, left) : left / 2;
    }
    s->use_graphs = use_graphs != 0;
    for (int i = 0; i < nslots; ++i) {
        auto sl = std::make_unique<Slot>();
        sl->idx = i;
        HV_HIP(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
        HV_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
        HV_HIP(hipHostMalloc((void**)&sl->dyn_host, 64, hipHostMallocDefault));
        s->slots.push_back(std::move(sl));
    }
    hv_server* raw = s.release();
    raw->worker = std::thread(worker_main, raw);
    raw->copier = std::thread(copier_main, raw);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    for (unsigned i = 0; i < std::min(4u, std::max(1u, hw / 2)); ++i)
        raw->touchers.emplace_back(toucher_main, raw);
    *out = raw;
    return HV_OK;
}

extern "C" int hv_server_submit(hv_server* s, const hv_job* job, int64_t* id) {
    if (!s || !job || !id) return fail(HV_ERR_ARG, "NULL argument");
    const hv_sky* sky = job->sky;
    const hv_beam* beam = job->beam;
    if (!sky || !beam || !job->antpos || !job->freqs || !job->eq2tops) return fail(HV_ERR_ARG, "NULL job array");
    if (sky->ctx != s->ctx || beam->ctx != s->ctx) return fail(HV_ERR_ARG, "sky/beam belong to another context");
    if (job->mode < HV_MODE_FP32 || job->mode > HV_MODE_FP64) return fail(HV_ERR_ARG, "unknown mode");
    if (job->nant <= 0 || job->nfreq <= 0 || job->ntimes <= 0) return fail(HV_ERR_ARG, "bad job shape");
    if (job->nsrc < 0 || job->nsrc > sky->nsrc) return fail(HV_ERR_ARG, "nsrc outside the sky catalog");
    if (beam->nant != job->nant) return fail(HV_ERR_ARG, "beam antenna count does not match antpos");
    if (beam->nfreq != 1 && beam->nfreq != job->nfreq)
        return fail(HV_ERR_ARG, "a per-frequency beam needs one plane per frequency of the job");
    if (job->mode == HV_MODE_FP64 && !beam->f64) return fail(HV_ERR_ARG, "FP64 mode needs a float64 beam");
    auto j = std::make_unique<SrvJob>();
    j->mode = job->mode;
    j->flags = job->flags;
    j->sky = sky;
    j->beam = beam;
    j->nsrc = job->nsrc;
    j->kpad = round_up(std::max<int64_t>(job->nsrc, 1), KALIGN);
    j->nant = job->nant;
    j->ntimes = job->ntimes;
    j->nfreq = job->nfreq;
    j->antpos.assign(job->antpos, job->antpos + (size_t)job->nant * 3);
    j->eq2tops.assign(job->eq2tops, job->eq2tops + (size_t)job->ntimes * 9);
    j->invlam.resize(job->nfreq);
    for (int i = 0; i < job->nfreq; ++i) j->invlam[i] = job->freqs[i] / C_LIGHT;
    j->out = job->vis_out;
    int64_t kb = 1024;
    while (kb < j->kpad) kb <<= 1;
    const float scale = (j->mode == HV_MODE_SPLIT_FP16) ? fp16_scale(beam->max_abs * sky->max_isqrt) : 1.0f;
    j->layout = (job->flags & HV_JOB_TBF) ? LAYOUT_TBF : (job->flags & HV_JOB_FTB) ? LAYOUT_FTB : LAYOUT_FULL;
    j->key = BucketKey{j->mode, j->nant, beam->beam_px, j->ntimes, j->nfreq, beam->f64, scale, kb, j->layout,
                       beam->nfreq > 1};
    j->plan = plan_job(j->mode, j->nant, kb, j->ntimes, j->nfreq, SERVE_OPERAND_BUDGET, j->layout, beam->nfreq > 1);
    j->out_bytes = j->plan.out_bytes;
    j->footprint = slot_bytes_needed(*j);
    if (j->footprint > s->budget)
        return fail(HV_ERR_NOMEM, "job footprint of " + std::to_string(j->footprint) +
                                      " bytes exceeds the server budget of " + std::to_string(s->budget));
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return fail(HV_ERR_STATE, "server is shutting down");
    j->id = s->next_id++;
    j->t_submit = now_s();
    *id = j->id;
    SrvJob* raw = j.release();
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

static void fill_info(const SrvJob* j, hv_job_info* info) {
    info->id = j->id;
    info->status = j->status;
    info->slot = j->slot;
    info->graph_hit = j->graph_hit;
    info->t_submit = j->t_submit;
    info->t_admit = j->t_admit;
    info->t_done = j->t_done;
    info->footprint = (int64_t)j->footprint;
    info->flops = 4.0 * j->nant * (j->nant + 1.0) * (double)j->nsrc * j->ntimes * j->nfreq;
    std::strncpy(info->error, j->err.c_str(), sizeof(info->error) - 1);
    info->error[sizeof(info->error) - 1] = 0;
}

extern "C" int hv_server_poll(hv_server* s, hv_job_info* infos, int max, int* n) {
    if (!s || !n || (max > 0 && !infos)) return fail(HV_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    int k = 0;
    while (k < max && !s->finished_order.empty()) {
        const int64_t id = s->finished_order.front();
        s->finished_order.pop_front();
        auto it = s->finished.find(id);
        if (it == s->finished.end()) continue;   // already collected by hv_server_wait
        std::memset(&infos[k], 0, sizeof(hv_job_info));
        fill_info(it->second, &infos[k]);
        delete it->second;
        s->finished.erase(it);
        ++k;
    }
    *n = k;
    return HV_OK;
}

extern "C" int hv_server_wait(hv_server* s, int64_t id, double timeout_s, hv_job_info* info) {
    if (!s) return fail(HV_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (id <= 0 || id >= s->next_id) return fail(HV_ERR_ARG, "unknown job id");
    auto pred = [&] { return s->finished.count(id) > 0; };
    if (timeout_s < 0) s->done_cv.wait(lk, pred);
    else if (!s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
        return fail(HV_ERR_BUSY, "job not finished within the timeout (or already collected)");
    SrvJob* j = s->finished[id];
    if (info) {
        std::memset(info, 0, sizeof(hv_job_info));
        fill_info(j, info);
    }
    s->finished.erase(id);
    delete j;
    return HV_OK;
}

extern "C" int hv_server_stats(hv_server* s, hv_server_stats_t* out) {
    if (!s || !out) return fail(HV_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->pending = (int64_t)s->pending.size();
    out->running = 0;
    for (auto& sl : s->slots) out->running += sl->state != FREE;
    out->budget_bytes = s->budget;
    return HV_OK;
}

extern "C" int hv_server_destroy(hv_server* s) {
    if (!s) return HV_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        // queued jobs that never started fail; running ones finish
        for (SrvJob* j : s->pending) {
            j->status = HV_ERR_STATE;
            j->err = "server destroyed before the job started";
            finish_locked(s, j);
        }
        s->pending.clear();
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
    (void)hipSetDevice(s->ctx->device);
    for (auto& up : s->slots) {
        Slot& sl = *up;
        (void)hipStreamSynchronize(sl.stream);
        clear_graphs(sl);
        release(s->ctx, sl.ws);
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        if (sl.dyn_host) (void)hipHostFree(sl.dyn_host);
        (void)hipEventDestroy(sl.done);
        (void)hipStreamDestroy(sl.stream);
    }
    for (auto& kv : s->finished) delete kv.second;
    delete s;
    return HV_OK;
}
