// Continuous-batching simulation server: the HIP executor and the C ABI
// (no counterpart in the reference, which runs one synchronous vis_gpu call
// at a time, hera_gpu/vis.py:144).
//
// The scheduling (queue, admission by footprint, instances per shape bucket,
// result copies, pre-faulting, cancellation) is hv_sched.h.  This file gives
// it a device:
//   - a slot is a HIP stream and a completion event; each launch copies its
//     result into a pinned host buffer from a pool, handed to the copier when
//     the job completes, so the slot takes its next job at once;
//   - an instance is a workspace (every device buffer of one job of the
//     bucket, at fixed addresses) and the hipGraph captured from the kernel
//     sequence of the first job that ran in it;
//   - a job's own inputs are staged into the instance's buffers outside the
//     graph (catalog sources, beam, antenna positions, times, frequencies and
//     the true source count), and the kernels read the source count from
//     device memory, so a replay does exactly the arithmetic of a
//     stand-alone run: results are bitwise identical alone and in a mix.
// Graph replays replace the reference's per-call JIT and per-call uploads
// (vis.py:183-206).
#include <map>

#include "hv_host.h"
#include "hv_sched.h"

using namespace hv;

namespace {

// Batch bytes (partial tiles, or operands) one serving job holds at once:
// 18 HERA-350 units of 2^20 sources per contraction launch.  Instances stay
// resident between jobs, so this bounds the memory each bucket keeps: the
// largest instance of the mixed stream is ~6 GB (12 GB of batch gave
// ~15 GB, and the 277 buckets of a 2000-job stream ~600 GB in all).
constexpr size_t SERVE_OPERAND_BUDGET = (size_t)2 << 30;

struct HipJob {
    const hv_sky* sky = nullptr;
    const hv_beam* beam = nullptr;
    int64_t nsrc = 0, kpad = 0;
    int mode = 0, nant = 0, ntimes = 0, nfreq = 0, layout = 0;
    std::vector<double> antpos, invlam, eq2tops;
    Plan plan;                 // planned at the bucket's padded size
};

struct HipInstance {
    Workspace ws;
    hipGraphExec_t graph = nullptr;
};

struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct HipSlot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    PinnedBuf* result = nullptr;   // pinned buffer of the slot's current job
    int64_t* dyn_host = nullptr;   // pinned source count staged to the device
};

size_t beam_bytes(const HipJob& j) {
    return (size_t)j.beam->nfreq * j.nant * j.beam->beam_px * j.beam->beam_px * (j.beam->f64 ? 8 : 4);
}

// Size class of a device block: the next of 2^k x {1, 1.25, 1.5, 1.75, 2}
// (at least 256 B), so blocks freed by evicted instances serve later ones.
size_t size_class(size_t b) {
    if (b <= 256) return 256;
    const int k = 63 - __builtin_clzll((unsigned long long)(b - 1));   // 2^k < b <= 2^(k+1)
    const size_t base = (size_t)1 << k;
    for (size_t q = 1; q <= 4; ++q)
        if (base + q * (base / 4) >= b) return base + q * (base / 4);
    return base << 1;
}

// The device buffers of an instance of the job's bucket (the plan's and the
// staged inputs'), each rounded to its size class.
std::vector<std::pair<Buf, size_t>> instance_buffers(const HipJob& j) {
    const Plan& p = j.plan;
    const std::pair<Buf, size_t> raw[] = {
        {B_CRD_TOP, p.crd_top_bytes}, {B_W, p.w_bytes}, {B_PART, p.part_bytes},
        {B_CRD_EQ, (size_t)3 * p.kpad * 8}, {B_ISQRT, (size_t)p.kpad * 8}, {B_BEAM, beam_bytes(j)},
        {B_ANTPOS, (size_t)j.nant * 24}, {B_EQ2TOPS, (size_t)j.ntimes * 72}, {B_INVLAM, (size_t)j.nfreq * 8},
        {B_OUT, p.out_bytes}, {B_DYN, 256}};
    std::vector<std::pair<Buf, size_t>> out;
    for (const auto& r : raw) out.emplace_back(r.first, size_class(r.second));
    return out;
}

size_t instance_bytes(const HipJob& j) {
    size_t t = 0;
    for (const auto& b : instance_buffers(j)) t += b.second;
    return t;
}

class HipExecutor : public sched::Executor {
public:
    HipExecutor(hv_ctx* c, bool graphs) : ctx(c), use_graphs(graphs) {}

    int init_slots(int n) {
        slots.resize((size_t)n);
        for (auto& sl : slots) {
            HV_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
            HV_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
            HV_HIP(hipHostMalloc((void**)&sl.dyn_host, 64, hipHostMallocDefault));
        }
        return HV_OK;
    }

    ~HipExecutor() override {
        (void)hipSetDevice(ctx->device);
        for (auto& sl : slots) {
            if (sl.stream) (void)hipStreamSynchronize(sl.stream);
            if (sl.result) pool.push_back(sl.result);
            if (sl.dyn_host) (void)hipHostFree(sl.dyn_host);
            if (sl.done) (void)hipEventDestroy(sl.done);
            if (sl.stream) (void)hipStreamDestroy(sl.stream);
        }
        for (PinnedBuf* b : pool) {
            (void)hipHostFree(b->p);
            delete b;
        }
        cache_trim();
    }

    void worker_init() override { (void)hipSetDevice(ctx->device); }

    int instance_create(const sched::Job& j, void** out) override {
        *out = nullptr;
        HV_HIP(hipSetDevice(ctx->device));
        const HipJob& hj = *static_cast<const HipJob*>(j.payload);
        auto in = std::make_unique<HipInstance>();
        Workspace& w = in->ws;
        for (const auto& b : instance_buffers(hj)) {
            void* p = nullptr;
            const int rc = block_get(b.second, &p);
            if (rc != HV_OK) {
                put_blocks(w);
                return rc;
            }
            w.buf[b.first].p = p;
            w.buf[b.first].cap = b.second;
        }
        *out = in.release();
        return HV_OK;
    }

    // The instance's blocks go back to the cache, not to hipFree (which
    // waits for the whole device): a later instance of any bucket reuses them.
    void instance_destroy(void* h) override {
        auto* in = static_cast<HipInstance*>(h);
        (void)hipSetDevice(ctx->device);
        if (in->graph) (void)hipGraphExecDestroy(in->graph);
        put_blocks(in->ws);
        delete in;
    }

    int launch(int slot, void* h, sched::Job& j) override;

    int query(int slot) override {
        const hipError_t q = hipEventQuery(slots[(size_t)slot].done);
        if (q == hipErrorNotReady) return HV_ERR_BUSY;
        if (q != hipSuccess) return fail(HV_ERR_HIP, std::string("device error: ") + hipGetErrorString(q));
        return HV_OK;
    }

    void drain(int slot) override { (void)hipStreamSynchronize(slots[(size_t)slot].stream); }

    void* take_result(int slot) override {
        PinnedBuf* b = slots[(size_t)slot].result;
        slots[(size_t)slot].result = nullptr;
        return b;
    }
    const void* result_data(void* r) override { return static_cast<PinnedBuf*>(r)->p; }
    void release_result(void* r) override {
        PinnedBuf* drop = nullptr;
        {
            std::lock_guard<std::mutex> lk(pool_mu);
            pool.push_back(static_cast<PinnedBuf*>(r));
            // keep at most three buffers per slot; drop the smallest beyond that
            if (pool.size() > 3 * slots.size()) {
                auto it = std::min_element(pool.begin(), pool.end(),
                                           [](const PinnedBuf* a, const PinnedBuf* b) { return a->cap < b->cap; });
                drop = *it;
                pool.erase(it);
            }
        }
        if (drop) {
            (void)hipHostFree(drop->p);
            delete drop;
        }
    }

    void job_free(sched::Job& j) override {
        delete static_cast<HipJob*>(j.payload);
        j.payload = nullptr;
    }

private:
    // Device blocks of one size class: a cached free block of exactly that
    // class, else a new allocation; when the context budget is exhausted the
    // cache is released (hipFree) and the allocation tried again.  Worker
    // thread only.
    int block_get(size_t bytes, void** p) {
        auto it = cache.find(bytes);
        if (it != cache.end()) {
            *p = it->second;
            cache.erase(it);
            return HV_OK;
        }
        int rc = dev_alloc(ctx, p, bytes);
        if (rc != HV_OK && !cache.empty()) {
            cache_trim();
            rc = dev_alloc(ctx, p, bytes);
        }
        return rc;
    }
    void put_blocks(Workspace& w) {
        for (auto& b : w.buf)
            if (b.p) {
                cache.emplace(b.cap, b.p);
                b.p = nullptr;
                b.cap = 0;
            }
    }
    void cache_trim() {
        for (auto& kv : cache) dev_free(ctx, kv.second, kv.first);
        cache.clear();
    }

    // A free pinned buffer of at least `bytes` (the smallest that fits), or a
    // new one rounded up to a power of two (at least 16 MB) so the pool holds
    // few sizes.
    int pinned_get(size_t bytes, PinnedBuf** out) {
        {
            std::lock_guard<std::mutex> lk(pool_mu);
            auto best = pool.end();
            for (auto it = pool.begin(); it != pool.end(); ++it)
                if ((*it)->cap >= bytes && (best == pool.end() || (*it)->cap < (*best)->cap)) best = it;
            if (best != pool.end()) {
                *out = *best;
                pool.erase(best);
                return HV_OK;
            }
        }
        size_t cap = (size_t)16 << 20;
        while (cap < bytes) cap <<= 1;
        auto b = std::make_unique<PinnedBuf>();
        HV_HIP(hipHostMalloc(&b->p, cap, hipHostMallocDefault));
        b->cap = cap;
        *out = b.release();
        return HV_OK;
    }

    hv_ctx* ctx;
    bool use_graphs;
    std::vector<HipSlot> slots;
    std::mutex pool_mu;               // the copier returns buffers from its thread
    std::vector<PinnedBuf*> pool;     // free pinned result buffers
    std::multimap<size_t, void*> cache;   // free device blocks by size class
};

// Stage the job's inputs into the instance, run its kernels (graph replay,
// capture on the instance's first job, or direct launches) and copy the
// result to the slot's pinned buffer.
int HipExecutor::launch(int slot, void* h, sched::Job& job) {
    HV_HIP(hipSetDevice(ctx->device));
    HipSlot& sl = slots[(size_t)slot];
    auto* in = static_cast<HipInstance*>(h);
    const HipJob& j = *static_cast<const HipJob*>(job.payload);
    const Plan& p = j.plan;
    Workspace& w = in->ws;
    if (sl.result && sl.result->cap < p.out_bytes) {
        release_result(sl.result);
        sl.result = nullptr;
    }
    if (!sl.result) HV_CHECK(pinned_get(p.out_bytes, &sl.result));

    // ---- stage inputs (outside any graph)
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
    HV_HIP(hipMemcpyAsync(w.buf[B_BEAM].p, j.beam->data, beam_bytes(j), hipMemcpyDeviceToDevice, sl.stream));
    HV_HIP(hipMemcpyAsync(w.buf[B_ANTPOS].p, j.antpos.data(), j.antpos.size() * 8, hipMemcpyHostToDevice,
                          sl.stream));
    HV_HIP(hipMemcpyAsync(w.buf[B_EQ2TOPS].p, j.eq2tops.data(), j.eq2tops.size() * 8, hipMemcpyHostToDevice,
                          sl.stream));
    HV_HIP(hipMemcpyAsync(w.buf[B_INVLAM].p, j.invlam.data(), j.invlam.size() * 8, hipMemcpyHostToDevice,
                          sl.stream));
    // the slot's previous job has finished (the slot is free), so its
    // staging word may be rewritten
    *sl.dyn_host = j.kpad;
    HV_HIP(hipMemcpyAsync(w.buf[B_DYN].p, sl.dyn_host, 8, hipMemcpyHostToDevice, sl.stream));

    JobDev jd{};
    jd.mode = j.mode;
    jd.nant = j.nant;
    jd.ntimes = j.ntimes;
    jd.nfreq = j.nfreq;
    jd.beam_px = j.beam->beam_px;
    jd.beam_f64 = j.beam->f64;
    jd.kpad = p.kpad;
    jd.antpos = (const double*)w.buf[B_ANTPOS].p;
    jd.inv_lambda = (const double*)w.buf[B_INVLAM].p;
    jd.eq2tops = (const double*)w.buf[B_EQ2TOPS].p;
    jd.crd_eq = crd;
    jd.isqrt = isq;
    jd.beam = w.buf[B_BEAM].p;
    jd.vis = w.buf[B_OUT].p;
    jd.amp_scale = job.key.scale;
    jd.dyn_kpad = (const int64_t*)w.buf[B_DYN].p;
    jd.layout = j.layout;
    jd.beam_fstride = j.beam->nfreq > 1 ? (size_t)j.nant * j.beam->beam_px * j.beam->beam_px : 0;

    StageTimer tm;
    if (use_graphs) {
        if (!in->graph) {
            hipGraph_t g = nullptr;
            HV_HIP(hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_job(jd, p, w, sl.stream, tm);
            const hipError_t e = hipStreamEndCapture(sl.stream, &g);
            if (rc != HV_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(HV_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess)
                return fail(HV_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
            in->graph = ex;
            job.graph_captured = 1;
        } else {
            job.graph_hit = 1;
        }
        HV_HIP(hipGraphLaunch(in->graph, sl.stream));
    } else {
        HV_CHECK(enqueue_job(jd, p, w, sl.stream, tm));
    }
    HV_HIP(hipMemcpyAsync(sl.result->p, w.buf[B_OUT].p, p.out_bytes, hipMemcpyDeviceToHost, sl.stream));
    HV_HIP(hipEventRecord(sl.done, sl.stream));
    return HV_OK;
}

}  // namespace

struct hv_server {
    hv_ctx* ctx = nullptr;
    std::unique_ptr<HipExecutor> ex;
    sched::Server* core = nullptr;
};

// ------------------------------------------------------------------ C ABI
extern "C" int hv_server_create(hv_ctx* c, int nslots, size_t budget, int use_graphs, hv_server** out) {
    if (!c || !out) return fail(HV_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (nslots < 1 || nslots > 16) return fail(HV_ERR_ARG, "nslots must be in [1, 16]");
    HV_HIP(hipSetDevice(c->device));
    auto s = std::make_unique<hv_server>();
    s->ctx = c;
    size_t b;
    {
        std::lock_guard<std::mutex> lk(c->mem_mu);
        const size_t left = c->budget - c->used;
        b = budget ? std::min(budget, left) : left / 2;
    }
    s->ex = std::make_unique<HipExecutor>(c, use_graphs != 0);
    HV_CHECK(s->ex->init_slots(nslots));
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    s->core = sched::create(s->ex.get(), nslots, b, (int)std::min(4u, std::max(1u, hw / 2)));
    *out = s.release();
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
    auto hj = std::make_unique<HipJob>();
    hj->sky = sky;
    hj->beam = beam;
    hj->mode = job->mode;
    hj->nsrc = job->nsrc;
    hj->kpad = round_up(std::max<int64_t>(job->nsrc, 1), KALIGN);
    hj->nant = job->nant;
    hj->ntimes = job->ntimes;
    hj->nfreq = job->nfreq;
    hj->antpos.assign(job->antpos, job->antpos + (size_t)job->nant * 3);
    hj->eq2tops.assign(job->eq2tops, job->eq2tops + (size_t)job->ntimes * 9);
    hj->invlam.resize((size_t)job->nfreq);
    for (int i = 0; i < job->nfreq; ++i) hj->invlam[(size_t)i] = job->freqs[i] / C_LIGHT;
    int64_t kb = 1024;
    while (kb < hj->kpad) kb <<= 1;
    const float scale = (hj->mode == HV_MODE_SPLIT_FP16) ? fp16_scale(beam->max_abs * sky->max_isqrt) : 1.0f;
    hj->layout = (job->flags & HV_JOB_TBF) ? LAYOUT_TBF : (job->flags & HV_JOB_FTB) ? LAYOUT_FTB : LAYOUT_FULL;
    // the K-split length follows the job's exact size (as a direct call's),
    // not the bucket's padded one; it is part of the bucket key
    const int ksplit = job_ksplit(hj->mode, hj->nant, (int64_t)hj->ntimes * hj->nfreq, hj->kpad);
    hj->plan = plan_job(hj->mode, hj->nant, kb, hj->ntimes, hj->nfreq, SERVE_OPERAND_BUDGET, hj->layout,
                        beam->nfreq > 1, false, ksplit);
    auto j = std::make_unique<sched::Job>();
    j->flags = job->flags;
    j->key = sched::BucketKey{hj->mode, hj->nant, beam->beam_px, hj->ntimes, hj->nfreq, beam->f64, scale, kb,
                              hj->layout, beam->nfreq > 1 ? 1 : 0, ksplit};
    j->footprint = instance_bytes(*hj);
    j->out_bytes = hj->plan.out_bytes;
    j->out = job->vis_out;
    j->flops = 4.0 * hj->nant * (hj->nant + 1.0) * (double)hj->nsrc * hj->ntimes * hj->nfreq;
    j->payload = hj.get();
    const int rc = sched::submit(s->core, std::move(j), id);
    if (rc == HV_OK) (void)hj.release();   // the server owns the payload now
    return rc;
}

extern "C" int hv_server_poll(hv_server* s, hv_job_info* infos, int max, int* n) {
    if (!s || !n || (max > 0 && !infos)) return fail(HV_ERR_ARG, "NULL argument");
    return sched::poll(s->core, infos, max, n);
}

extern "C" int hv_server_wait(hv_server* s, int64_t id, double timeout_s, hv_job_info* info) {
    if (!s) return fail(HV_ERR_ARG, "NULL server");
    return sched::wait(s->core, id, timeout_s, info);
}

extern "C" int hv_server_cancel(hv_server* s, int64_t id) {
    if (!s) return fail(HV_ERR_ARG, "NULL server");
    return sched::cancel(s->core, id);
}

extern "C" int hv_server_stats(hv_server* s, hv_server_stats_t* out) {
    if (!s || !out) return fail(HV_ERR_ARG, "NULL argument");
    return sched::stats(s->core, out);
}

extern "C" int hv_server_destroy(hv_server* s) {
    if (!s) return HV_OK;
    sched::destroy(s->core);   // joins the threads and destroys every instance
    s->ex.reset();             // streams, events, pinned buffers
    delete s;
    return HV_OK;
}
