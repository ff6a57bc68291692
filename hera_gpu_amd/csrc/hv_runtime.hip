// Host runtime and C ABI of libheravis (see include/heravis.h).
//
// One hv_ctx owns a device, a stream and a grow-only workspace.  A
// simulation is planned as: one rotation launch for all times, then for each
// batch of (time, frequency) units a phasor build, the Hermitian MFMA
// contraction and the K-split reduction, all on the context stream.  The
// plan replaces the reference's per-call JIT, texture upload, per-chunk
// streams and host-side chunk sum (hera_gpu/vis.py:176-258).
#include <cstdlib>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <mutex>
#include <string>
#include <vector>

#include "hv_host.h"

using namespace hv;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

int hv::fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

extern "C" int hv_version(void) { return 100; }
extern "C" const char* hv_last_error(void) { return g_err.c_str(); }

extern "C" int hv_device_count(int* count) {
    if (!count) return fail(HV_ERR_ARG, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return HV_OK;
}

// --------------------------------------------------------------- workspace
static int set_device(hv_ctx* c) {
    HV_HIP(hipSetDevice(c->device));
    return HV_OK;
}

int hv::dev_alloc(hv_ctx* c, void** p, size_t bytes) {
    if (bytes == 0) bytes = 256;
    {
        std::lock_guard<std::mutex> lk(c->mem_mu);
        if (c->used + bytes > c->budget)
            return fail(HV_ERR_NOMEM, "device memory request of " + std::to_string(bytes) +
                                          " bytes exceeds the HBM budget (" + std::to_string(c->budget - c->used) +
                                          " of " + std::to_string(c->budget) + " bytes left)");
        c->used += bytes;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(c->mem_mu);
        c->used -= bytes;
        *p = nullptr;
        return fail(HV_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return HV_OK;
}

void hv::dev_free(hv_ctx* c, void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    std::lock_guard<std::mutex> lk(c->mem_mu);
    c->used -= bytes;
}

// ------------------------------------------------------ staged host copies
namespace {
constexpr size_t STAGE_CHUNK = (size_t)64 << 20;
constexpr size_t STAGE_MIN = (size_t)8 << 20;   // below: one plain hipMemcpyAsync

// memcpy spread over up to 8 threads (pages of a fresh caller buffer fault in
// on first touch, which also spreads over cores)
void par_memcpy(void* dst, const void* src, size_t bytes) {
    constexpr size_t PER_MIN = (size_t)4 << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(8, hw), bytes / PER_MIN);
    if (nt <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t per = (bytes / nt + 4095) / 4096 * 4096;
    std::vector<std::thread> th;
    th.reserve(nt);
    size_t done = 0;   // bytes handed to threads
    try {
        for (size_t i = 0; i < nt; ++i) {
            const size_t off = i * per;
            if (off >= bytes) break;
            const size_t n = std::min(per, bytes - off);
            th.emplace_back(
                [=] { std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, n); });
            done = off + n;
        }
    } catch (...) {
        // no more threads: the rest on this one
    }
    if (done < bytes) std::memcpy(static_cast<char*>(dst) + done, static_cast<const char*>(src) + done, bytes - done);
    for (auto& t : th) t.join();
}

int ensure_stage(hv_ctx* c) {
    for (int k = 0; k < 2; ++k) {
        if (!c->stage[k]) HV_HIP(hipHostMalloc(&c->stage[k], STAGE_CHUNK, hipHostMallocDefault));
        if (!c->stage_ev[k]) HV_HIP(hipEventCreateWithFlags(&c->stage_ev[k], hipEventDisableTiming));
    }
    return HV_OK;
}
}  // namespace

int hv::copy_h2d(hv_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes < STAGE_MIN) {
        HV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        return HV_OK;
    }
    HV_CHECK(ensure_stage(c));
    int k = 0;
    for (size_t off = 0; off < bytes; off += STAGE_CHUNK, k ^= 1) {
        const size_t n = std::min(STAGE_CHUNK, bytes - off);
        HV_HIP(hipEventSynchronize(c->stage_ev[k]));   // the buffer's last DMA (this call or an earlier one)
        par_memcpy(c->stage[k], static_cast<const char*>(src) + off, n);
        HV_HIP(hipMemcpyAsync(static_cast<char*>(dst) + off, c->stage[k], n, hipMemcpyHostToDevice, s));
        HV_HIP(hipEventRecord(c->stage_ev[k], s));
    }
    return HV_OK;
}

int hv::copy_d2h(hv_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes < STAGE_MIN) {
        HV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        HV_HIP(hipStreamSynchronize(s));
        return HV_OK;
    }
    HV_CHECK(ensure_stage(c));
    for (int k = 0; k < 2; ++k) HV_HIP(hipEventSynchronize(c->stage_ev[k]));
    const size_t nch = (bytes + STAGE_CHUNK - 1) / STAGE_CHUNK;
    auto issue = [&](size_t i) -> int {
        const size_t off = i * STAGE_CHUNK, n = std::min(STAGE_CHUNK, bytes - off);
        HV_HIP(hipMemcpyAsync(c->stage[i & 1], static_cast<const char*>(src) + off, n, hipMemcpyDeviceToHost, s));
        HV_HIP(hipEventRecord(c->stage_ev[i & 1], s));
        return HV_OK;
    };
    HV_CHECK(issue(0));
    if (nch > 1) HV_CHECK(issue(1));
    for (size_t i = 0; i < nch; ++i) {
        const size_t off = i * STAGE_CHUNK, n = std::min(STAGE_CHUNK, bytes - off);
        HV_HIP(hipEventSynchronize(c->stage_ev[i & 1]));
        par_memcpy(static_cast<char*>(dst) + off, c->stage[i & 1], n);
        if (i + 2 < nch) HV_CHECK(issue(i + 2));
    }
    return HV_OK;
}

int hv::ensure(hv_ctx* c, Workspace& ws, Buf b, size_t bytes, hipStream_t stream, bool* moved) {
    DevBuf& d = ws.buf[b];
    if (d.cap >= bytes && d.p) return HV_OK;
    if (moved) *moved = true;
    if (d.p) {
        HV_HIP(hipStreamSynchronize(stream));
        dev_free(c, d.p, d.cap);
        d.p = nullptr;
        d.cap = 0;
    }
    // the context's own workspace (direct calls) grows geometrically (x1.5)
    // so a stream of growing requests reallocates O(log) times, falling back
    // to the exact size when the budget is tight; serving slots take exact
    // sizes, which their admission accounts for
    size_t want = std::max(bytes, (size_t)256);
    const size_t grown = (&ws == &c->ws) ? std::max(want, d.cap_seen + d.cap_seen / 2) : want;
    if (grown > want && dev_alloc(c, &d.p, grown) == HV_OK) {
        d.cap = d.cap_seen = grown;
        return HV_OK;
    }
    HV_CHECK(dev_alloc(c, &d.p, want));
    d.cap = d.cap_seen = want;
    return HV_OK;
}

void hv::release(hv_ctx* c, Workspace& ws) {
    for (auto& d : ws.buf) {
        dev_free(c, d.p, d.cap);
        d.p = nullptr;
        d.cap = 0;
    }
}

static int ensure(hv_ctx* c, Buf b, size_t bytes) { return hv::ensure(c, c->ws, b, bytes, c->stream); }

extern "C" int hv_ctx_create(int device, size_t hbm_budget, hv_ctx** out) {
    if (!out) return fail(HV_ERR_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(HV_ERR_HIP, "no HIP device visible");
    if (device < 0 || device >= n) return fail(HV_ERR_ARG, "device index out of range");
    hv_ctx* c = new hv_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return fail(HV_ERR_HIP, "hipSetDevice failed");
    }
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        delete c;
        return fail(HV_ERR_HIP, "hipMemGetInfo failed");
    }
    c->budget = hbm_budget ? std::min(hbm_budget, fr) : (size_t)(0.9 * (double)fr);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(HV_ERR_HIP, "hipStreamCreate failed");
    }
    for (auto& e : c->ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            delete c;
            return fail(HV_ERR_HIP, "hipEventCreate failed");
        }
    herk_init();   // occupancy queries now, never inside a stream capture
    *out = c;
    return HV_OK;
}

extern "C" int hv_ctx_destroy(hv_ctx* c) {
    if (!c) return HV_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->stream2);
    release(c, c->ws);
    for (int k = 0; k < 2; ++k) {
        if (c->stage_ev[k]) (void)hipEventSynchronize(c->stage_ev[k]);
        if (c->stage_ev[k]) (void)hipEventDestroy(c->stage_ev[k]);
        if (c->stage[k]) (void)hipHostFree(c->stage[k]);
    }
    for (auto e : c->ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream2);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return HV_OK;
}

// ------------------------------------------------------- host conversions
static void to_f64(const void* src, int dtype, size_t n, std::vector<double>& dst) {
    dst.resize(n);
    if (dtype == HV_F64) std::memcpy(dst.data(), src, n * sizeof(double));
    else {
        const float* f = static_cast<const float*>(src);
        for (size_t i = 0; i < n; ++i) dst[i] = (double)f[i];
    }
}

static int check_dtype(int dtype) {
    if (dtype != HV_F32 && dtype != HV_F64) return fail(HV_ERR_ARG, "dtype must be HV_F32 or HV_F64");
    return HV_OK;
}
static int check_mode(int mode) {
    if (mode < HV_MODE_FP32 || mode > HV_MODE_FP64) return fail(HV_ERR_ARG, "unknown mode");
    return HV_OK;
}

// Power-of-two scale that puts the largest |A sqrt(I)| near 2^14 (fp16 range).
float hv::fp16_scale(double max_amp) {
    if (!(max_amp > 0.0) || !std::isfinite(max_amp)) return 1.0f;
    int e = (int)std::floor(std::log2(16384.0 / max_amp));
    e = std::max(-60, std::min(60, e));
    return std::ldexp(1.0f, e);
}

// -------------------------------------------------------------- the job
int StageTimer::begin(int stage, hipStream_t s) {
    if (!on) return HV_OK;
    hipEvent_t e;
    HV_HIP(hipEventCreate(&e));
    HV_HIP(hipEventRecord(e, s));
    ev[stage][0].push_back(e);
    return HV_OK;
}
int StageTimer::end(int stage, hipStream_t s) {
    if (!on) return HV_OK;
    hipEvent_t e;
    HV_HIP(hipEventCreate(&e));
    HV_HIP(hipEventRecord(e, s));
    ev[stage][1].push_back(e);
    return HV_OK;
}
int StageTimer::collect(hv_stage_times* t) {
    double ms[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        for (size_t i = 0; i < ev[k][0].size(); ++i) {
            float m = 0;
            HV_HIP(hipEventElapsedTime(&m, ev[k][0][i], ev[k][1][i]));
            ms[k] += m;
        }
    }
    if (t) {
        t->coords_ms = ms[0]; t->phasor_ms = ms[1]; t->herk_ms = ms[2]; t->reduce_ms = ms[3];
        t->coords_n = (int64_t)ev[0][0].size(); t->phasor_n = (int64_t)ev[1][0].size();
        t->herk_n = (int64_t)ev[2][0].size(); t->reduce_n = (int64_t)ev[3][0].size();
    }
    return HV_OK;
}
StageTimer::~StageTimer() {
    for (auto& st : ev)
        for (auto& v : st)
            for (auto e : v) (void)hipEventDestroy(e);
}

// Upper bound on operand bytes held at once by a direct (non-serving) call.
static size_t operand_budget(hv_ctx* c) {
    return std::min((size_t)24 << 30, c->budget / 3);
}

Plan hv::plan_job(int mode, int nant, int64_t kpad, int ntimes, int nfreq, size_t wbudget, int layout,
                  bool beam_per_freq, bool operands, int ksplit) {
    Plan p;
    p.mode = mode;
    p.nant = nant;
    p.ntimes = ntimes;
    p.nfreq = nfreq;
    p.kpad = kpad;
    p.fused = herk_fused(mode, nant) && !operands;
    p.tile = p.fused ? herk_tile(mode, nant) : herk_tile_operands(mode, nant);
    p.nb = (nant + p.tile - 1) / p.tile;
    p.ntiles = p.nb * (p.nb + 1) / 2;
    // polarised jobs contract operands built in HBM (few units per launch):
    // short splits
    p.ksplit = ksplit ? ksplit : operands ? KSPLIT : job_ksplit(mode, nant, (int64_t)ntimes * nfreq, kpad);
    p.nsplit = (int)((kpad + p.ksplit - 1) / p.ksplit);
    p.row_bytes = operand_row_bytes(mode, kpad);
    p.unit_bytes = (size_t)nant * p.row_bytes;
    const size_t acc_bytes = mode_is_f64(mode) ? 8 : 4;
    p.part_unit = (size_t)p.nsplit * p.ntiles * 2 * p.tile * p.tile * acc_bytes;
    // batch: all frequencies of a time if they fit, then as many times as fit.
    // Operands per unit bound a batch, or, when the contraction builds them,
    // the partial tiles per unit and A sqrt(I) per time (per unit with a beam
    // per frequency: then the frequencies of a time do not share it)
    p.amp_unit = p.fused && beam_per_freq;
    p.amp_time = p.fused ? (size_t)nant * kpad * sizeof(float) : 0;
    const size_t per_unit = p.fused ? p.part_unit + (p.amp_unit ? p.amp_time : 0) : p.unit_bytes;
    const size_t per_time = p.amp_unit ? 0 : p.amp_time;
    p.fb = (int)std::max<size_t>(1, std::min<size_t>((size_t)nfreq, wbudget / std::max<size_t>(per_unit, 1)));
    p.tb = 1;
    if (p.fb == nfreq)
        p.tb = (int)std::max<size_t>(
            1, std::min<size_t>((size_t)ntimes, wbudget / std::max<size_t>(per_unit * nfreq + per_time, 1)));
    // keep one contraction launch within ~64k work items and 2^44 (unit,
    // source, antenna pair) products, 128 units of 2^20 sources at 350
    // antennas (HERA-350 with 32768-source splits: 2 times per launch
    // 2580-2586 ms per step, 7 times 2597-2604)
    while (p.tb > 1 && ((int64_t)p.tb * p.fb * p.nsplit * p.ntiles > 65536 ||
                        (double)p.tb * p.fb * (double)kpad * nant * nant > 17592186044416.0))
        --p.tb;
    const int units = p.fb * p.tb;
    p.crd_top_bytes = (size_t)ntimes * 3 * kpad * sizeof(double);
    p.w_bytes = p.fused ? p.amp_time * (p.amp_unit ? units : p.tb) : p.unit_bytes * units;
    p.part_bytes = p.part_unit * units;
    p.out_bytes = out_elems(layout, nfreq, ntimes, nant) * (mode_is_f64(mode) ? 16 : 8);
    return p;
}

int hv::prepare_workspace(hv_ctx* c, Workspace& ws, const Plan& p, hipStream_t s, bool* moved) {
    HV_CHECK(ensure(c, ws, B_CRD_TOP, p.crd_top_bytes, s, moved));
    HV_CHECK(ensure(c, ws, B_W, p.w_bytes, s, moved));
    HV_CHECK(ensure(c, ws, B_PART, p.part_bytes, s, moved));
    return HV_OK;
}

static int enqueue_coords(const JobDev& j, const Plan& p, Workspace& ws, hipStream_t stream, StageTimer& tm) {
    CoordsArgs ca{j.eq2tops, j.crd_eq, static_cast<double*>(ws.buf[B_CRD_TOP].p), p.kpad, j.ntimes, j.dyn_kpad};
    HV_CHECK(tm.begin(0, stream));
    HV_HIP(launch_coords(ca, stream));
    HV_CHECK(tm.end(0, stream));
    return HV_OK;
}

static int enqueue_phasor(const JobDev& j, const Plan& p, Workspace& ws, void* W, int t0, int nt, int f0, int nf,
                          hipStream_t stream, StageTimer& tm) {
    PhasorArgs pa{};
    pa.crd_top = static_cast<double*>(ws.buf[B_CRD_TOP].p) + (size_t)t0 * 3 * p.kpad;
    pa.isqrt = j.isqrt;
    pa.beam = static_cast<const char*>(j.beam) + (size_t)f0 * j.beam_fstride * (j.beam_f64 ? 8 : 4);
    pa.beam_fstride = j.beam_fstride;
    pa.antpos = j.antpos;
    pa.inv_lambda = j.inv_lambda + f0;
    pa.W = W;
    pa.unit_stride = p.unit_bytes;
    pa.kpad = p.kpad;
    pa.nant = j.nant;
    pa.beam_px = j.beam_px;
    pa.nfreq = nf;
    pa.ntime = nt;
    pa.amp_scale = j.amp_scale;
    pa.in_f64 = j.beam_f64;
    pa.dyn_kpad = j.dyn_kpad;
    HV_CHECK(tm.begin(1, stream));
    HV_HIP(launch_phasor(j.mode, pa, stream));
    HV_CHECK(tm.end(1, stream));
    return HV_OK;
}

static int enqueue_reduce(const JobDev& j, const Plan& p, void* part, int t0, int nt, int f0, int nf,
                          hipStream_t stream, StageTimer& tm);

static int enqueue_contract(const JobDev& j, const Plan& p, const void* W, void* part, int t0, int nt, int f0,
                            int nf, hipStream_t stream, StageTimer& tm) {
    HerkArgs ha{};
    ha.W = W;
    ha.unit_stride = p.unit_bytes;
    ha.partials = part;
    ha.nant = j.nant;
    ha.nb = p.nb;
    ha.ntiles = p.ntiles;
    ha.kpad = p.kpad;
    ha.nsplit = p.nsplit;
    ha.ksplit = p.ksplit;
    ha.nunits = nt * nf;
    ha.dyn_kpad = j.dyn_kpad;
    HV_CHECK(tm.begin(2, stream));
    HV_HIP(launch_herk(j.mode, p.tile, ha, stream));
    HV_CHECK(tm.end(2, stream));
    return enqueue_reduce(j, p, part, t0, nt, f0, nf, stream, tm);
}

static int enqueue_reduce(const JobDev& j, const Plan& p, void* part, int t0, int nt, int f0, int nf,
                          hipStream_t stream, StageTimer& tm) {
    ReduceArgs ra{};
    ra.partials = part;
    ra.vis = j.vis;
    ra.t0 = j.t_base + t0;
    ra.f0 = f0;
    ra.nfreq_launch = nf;
    ra.ntimes_total = j.ntimes_out ? j.ntimes_out : j.ntimes;
    ra.nant = j.nant;
    ra.nb = p.nb;
    ra.ntiles = p.ntiles;
    ra.nsplit = p.nsplit;
    ra.ksplit = p.ksplit;
    ra.nunits = nt * nf;
    ra.tile = p.tile;
    ra.scale = 1.0 / ((double)j.amp_scale * (double)j.amp_scale);
    ra.dyn_kpad = j.dyn_kpad;
    ra.layout = j.layout;
    ra.nfreq_total = j.nfreq;
    HV_CHECK(tm.begin(3, stream));
    HV_HIP(launch_reduce(j.mode, ra, stream));
    HV_CHECK(tm.end(3, stream));
    return HV_OK;
}

// Fused plans: A sqrt(I) of the batch's times (stage 1), then the
// contraction that builds its operands from it (stage 2).
static int enqueue_amp(const JobDev& j, const Plan& p, Workspace& ws, int t0, int nt, int f0, int nf,
                       hipStream_t stream, StageTimer& tm) {
    AmpArgs aa{};
    aa.crd_top = static_cast<double*>(ws.buf[B_CRD_TOP].p) + (size_t)t0 * 3 * p.kpad;
    aa.isqrt = j.isqrt;
    aa.beam = static_cast<const char*>(j.beam) + (size_t)f0 * j.beam_fstride * (j.beam_f64 ? 8 : 4);
    aa.beam_fstride = j.beam_fstride;
    aa.nfreq = nf;
    aa.amp = static_cast<float*>(ws.buf[B_W].p);
    aa.kpad = p.kpad;
    aa.nant = j.nant;
    aa.beam_px = j.beam_px;
    aa.ntime = nt;
    aa.in_f64 = j.beam_f64;
    aa.amp_scale = j.amp_scale;
    aa.dyn_kpad = j.dyn_kpad;
    HV_CHECK(tm.begin(1, stream));
    HV_HIP(launch_amp(aa, stream));
    HV_CHECK(tm.end(1, stream));
    return HV_OK;
}

static int enqueue_fused(const JobDev& j, const Plan& p, Workspace& ws, int t0, int nt, int f0, int nf,
                         hipStream_t stream, StageTimer& tm) {
    FusedArgs fa{};
    fa.amp = static_cast<const float*>(ws.buf[B_W].p);
    fa.crd_top = static_cast<double*>(ws.buf[B_CRD_TOP].p) + (size_t)t0 * 3 * p.kpad;
    fa.antpos = j.antpos;
    fa.inv_lambda = j.inv_lambda + f0;
    fa.partials = ws.buf[B_PART].p;
    fa.kpad = p.kpad;
    fa.nant = j.nant;
    fa.nb = p.nb;
    fa.ntiles = p.ntiles;
    fa.nsplit = p.nsplit;
    fa.ksplit = p.ksplit;
    fa.nunits = nt * nf;
    fa.nfreq = nf;
    fa.dyn_kpad = j.dyn_kpad;
    fa.amp_per_unit = p.amp_unit ? 1 : 0;
    fa.job_nfreq = j.nfreq;
    HV_CHECK(tm.begin(2, stream));
    HV_HIP(launch_fused(j.mode, p.tile, fa, stream));
    HV_CHECK(tm.end(2, stream));
    return HV_OK;
}

int hv::enqueue_job(const JobDev& j, const Plan& p, Workspace& ws, hipStream_t stream, StageTimer& tm) {
    HV_CHECK(enqueue_coords(j, p, ws, stream, tm));
    if (p.fused) {
        int amp_t0 = -1;
        for (int t0 = 0; t0 < j.ntimes; t0 += p.tb) {
            const int nt = std::min(p.tb, j.ntimes - t0);
            for (int f0 = 0; f0 < j.nfreq; f0 += p.fb) {
                const int nf = std::min(p.fb, j.nfreq - f0);
                if (amp_t0 != t0 || p.amp_unit) {          // per time, or per batch of units
                    HV_CHECK(enqueue_amp(j, p, ws, t0, nt, f0, nf, stream, tm));
                    amp_t0 = t0;
                }
                HV_CHECK(enqueue_fused(j, p, ws, t0, nt, f0, nf, stream, tm));
                HV_CHECK(enqueue_reduce(j, p, ws.buf[B_PART].p, t0, nt, f0, nf, stream, tm));
            }
        }
        return HV_OK;
    }
    for (int t0 = 0; t0 < j.ntimes; t0 += p.tb) {
        const int nt = std::min(p.tb, j.ntimes - t0);
        for (int f0 = 0; f0 < j.nfreq; f0 += p.fb) {
            const int nf = std::min(p.fb, j.nfreq - f0);
            HV_CHECK(enqueue_phasor(j, p, ws, ws.buf[B_W].p, t0, nt, f0, nf, stream, tm));
            HV_CHECK(enqueue_contract(j, p, ws.buf[B_W].p, ws.buf[B_PART].p, t0, nt, f0, nf, stream, tm));
        }
    }
    return HV_OK;
}

int hv::enqueue_job_pipelined(hv_ctx* c, const JobDev& j, const Plan& p, Workspace& ws, StageTimer& tm) {
    hipStream_t sm = c->stream, sp = c->stream2;
    hipEvent_t start = c->ev[0], ph_done[2] = {c->ev[1], c->ev[2]}, hk_done[2] = {c->ev[3], c->ev[4]};
    char* W[2] = {static_cast<char*>(ws.buf[B_W].p), static_cast<char*>(ws.buf[B_W].p) + p.w_bytes};
    char* P[2] = {static_cast<char*>(ws.buf[B_PART].p), static_cast<char*>(ws.buf[B_PART].p) + p.part_bytes};
    // the phasor stream starts after everything already queued on the main stream
    HV_HIP(hipEventRecord(start, sm));
    HV_HIP(hipStreamWaitEvent(sp, start, 0));
    HV_CHECK(enqueue_coords(j, p, ws, sp, tm));
    int b = 0;
    for (int t0 = 0; t0 < j.ntimes; t0 += p.tb) {
        const int nt = std::min(p.tb, j.ntimes - t0);
        for (int f0 = 0; f0 < j.nfreq; f0 += p.fb, ++b) {
            const int nf = std::min(p.fb, j.nfreq - f0);
            const int k = b & 1;
            if (b >= 2) HV_HIP(hipStreamWaitEvent(sp, hk_done[k], 0));   // contraction b-2 done with W[k]
            HV_CHECK(enqueue_phasor(j, p, ws, W[k], t0, nt, f0, nf, sp, tm));
            HV_HIP(hipEventRecord(ph_done[k], sp));
            HV_HIP(hipStreamWaitEvent(sm, ph_done[k], 0));
            HV_CHECK(enqueue_contract(j, p, W[k], P[k], t0, nt, f0, nf, sm, tm));
            HV_HIP(hipEventRecord(hk_done[k], sm));
        }
    }
    return HV_OK;
}

// Batches of a direct call run on one stream (phasor build, then
// contraction) unless HERAVIS_OVERLAP=1 asks for the two-stream schedule.
// Measured on MI355X (HERA-350, 2^20 sources): the phasor build writes its
// operands at close to the HBM write ceiling (~3.6 TB/s), so running it
// beside the contraction slows both, and one stream is ~6% faster.
static bool overlap_batches() {
    const char* e = std::getenv("HERAVIS_OVERLAP");
    return e && e[0] == '1';
}

static int run_job(hv_ctx* c, const JobDev& j, StageTimer& tm) {
    Plan p = plan_job(j.mode, j.nant, j.kpad, j.ntimes, j.nfreq, operand_budget(c), j.layout, j.beam_fstride != 0);
    const int64_t batches = (int64_t)((j.ntimes + p.tb - 1) / p.tb) * ((j.nfreq + p.fb - 1) / p.fb);
    if (batches < 2 || !overlap_batches() || p.fused) {
        HV_CHECK(prepare_workspace(c, c->ws, p, c->stream));
        return enqueue_job(j, p, c->ws, c->stream, tm);
    }
    // several batches: double-buffer operands and partials at half the budget each
    p = plan_job(j.mode, j.nant, j.kpad, j.ntimes, j.nfreq, operand_budget(c) / 2, j.layout, j.beam_fstride != 0);
    HV_CHECK(ensure(c, c->ws, B_CRD_TOP, p.crd_top_bytes, c->stream));
    HV_CHECK(ensure(c, c->ws, B_W, 2 * p.w_bytes, c->stream));
    HV_CHECK(ensure(c, c->ws, B_PART, 2 * p.part_bytes, c->stream));
    return enqueue_job_pipelined(c, j, p, c->ws, tm);
}

// ---------------------------------------------------------- argument checks
// Beams up to 2048 pixels a side: planes that fit in LDS are staged there
// (P <= 202 for float32, 143 for float64; polarised beams stage four planes,
// P <= 101 / 71), larger ones are read from global memory by the same
// bilinear arithmetic.
constexpr int MAX_BEAM_PX = 2048;
static int check_job_args(int mode, int dtype, int nant, int nfreq, int ntimes, int64_t nsrc, int beam_px) {
    HV_CHECK(check_mode(mode));
    HV_CHECK(check_dtype(dtype));
    if (nant <= 0) return fail(HV_ERR_ARG, "nant must be positive");
    if (nfreq <= 0) return fail(HV_ERR_ARG, "nfreq must be positive");
    if (ntimes <= 0) return fail(HV_ERR_ARG, "ntimes must be positive");
    if (nsrc < 0) return fail(HV_ERR_ARG, "nsrc must be non-negative");
    if (beam_px < 1 || beam_px > MAX_BEAM_PX) return fail(HV_ERR_ARG, "beam_px must be in [1, 2048]");
    if (nsrc > ((int64_t)1 << 36)) return fail(HV_ERR_ARG, "nsrc too large");
    return HV_OK;
}

static void inv_lambdas(const double* freqs, int nfreq, std::vector<double>& out) {
    out.resize(nfreq);
    for (int i = 0; i < nfreq; ++i) out[i] = freqs[i] / C_LIGHT;
}

static int upload(hv_ctx* c, Buf b, const void* host, size_t bytes) {
    HV_CHECK(ensure(c, b, bytes));
    HV_HIP(hipMemcpyAsync(c->ws.buf[b].p, host, bytes, hipMemcpyHostToDevice, c->stream));
    return HV_OK;
}

// Pad (3,K) directions and K sqrt-intensities to kpad with zeros (a zero
// direction has z = 0, i.e. not above the horizon, so it contributes 0).
static void pad_sky(const void* crd_eq, const void* isqrt, int dtype, int64_t nsrc, int64_t kpad,
                    std::vector<double>& crd, std::vector<double>& isq) {
    std::vector<double> tmp;
    to_f64(crd_eq, dtype, (size_t)3 * nsrc, tmp);
    crd.assign((size_t)3 * kpad, 0.0);
    for (int d = 0; d < 3; ++d)
        std::memcpy(crd.data() + d * kpad, tmp.data() + d * nsrc, nsrc * sizeof(double));
    to_f64(isqrt, dtype, (size_t)nsrc, tmp);
    isq.assign((size_t)kpad, 0.0);
    std::memcpy(isq.data(), tmp.data(), nsrc * sizeof(double));
}

static double max_abs(const void* p, int dtype, size_t n) {
    double m = 0;
    if (dtype == HV_F64) {
        const double* d = static_cast<const double*>(p);
        for (size_t i = 0; i < n; ++i) m = std::max(m, std::fabs(d[i]));
    } else {
        const float* f = static_cast<const float*>(p);
        for (size_t i = 0; i < n; ++i) m = std::max(m, (double)std::fabs(f[i]));
    }
    return m;
}

// ------------------------------------------------------------ hv_vis
extern "C" int hv_vis(hv_ctx* c, int mode, int dtype, const void* antpos, int nant, const double* freqs,
                      int nfreq, const void* eq2tops, int ntimes, const void* crd_eq, const void* isqrt,
                      int64_t nsrc, const void* bm_cube, int beam_px, void* vis_out) {
    return hv_vis_ex(c, mode, dtype, antpos, nant, freqs, nfreq, eq2tops, ntimes, crd_eq, isqrt, nsrc, bm_cube,
                     beam_px, 1, HV_LAYOUT_FULL, vis_out);
}

extern "C" int hv_vis_ex(hv_ctx* c, int mode, int dtype, const void* antpos, int nant, const double* freqs,
                         int nfreq, const void* eq2tops, int ntimes, const void* crd_eq, const void* isqrt,
                         int64_t nsrc, const void* bm_cube, int beam_px, int beam_nfreq, int layout,
                         void* vis_out) {
    if (!c) return fail(HV_ERR_ARG, "ctx is NULL");
    if (layout < HV_LAYOUT_FULL || layout > HV_LAYOUT_TBF) return fail(HV_ERR_ARG, "unknown output layout");
    if (beam_nfreq != 1 && beam_nfreq != nfreq)
        return fail(HV_ERR_ARG, "beam_nfreq must be 1 (one beam) or nfreq (a beam per frequency)");
    HV_CHECK(check_job_args(mode, dtype, nant, nfreq, ntimes, nsrc, beam_px));
    if (!antpos || !freqs || !eq2tops || !bm_cube || !vis_out || (nsrc > 0 && (!crd_eq || !isqrt)))
        return fail(HV_ERR_ARG, "NULL array argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_CHECK(set_device(c));
    const int64_t kpad = round_up(std::max<int64_t>(nsrc, 1), KALIGN);

    std::vector<double> ap, eq, il, crd, isq;
    to_f64(antpos, dtype, (size_t)nant * 3, ap);
    to_f64(eq2tops, dtype, (size_t)ntimes * 9, eq);
    inv_lambdas(freqs, nfreq, il);
    pad_sky(crd_eq, isqrt, dtype, nsrc, kpad, crd, isq);
    const size_t beam_plane = (size_t)nant * beam_px * beam_px;
    const size_t beam_bytes = beam_plane * beam_nfreq * (dtype == HV_F64 ? 8 : 4);

    HV_CHECK(upload(c, B_ANTPOS, ap.data(), ap.size() * 8));
    HV_CHECK(upload(c, B_EQ2TOPS, eq.data(), eq.size() * 8));
    HV_CHECK(upload(c, B_INVLAM, il.data(), il.size() * 8));
    HV_CHECK(upload(c, B_CRD_EQ, crd.data(), crd.size() * 8));
    HV_CHECK(upload(c, B_ISQRT, isq.data(), isq.size() * 8));
    HV_CHECK(upload(c, B_BEAM, bm_cube, beam_bytes));
    const size_t cplx = mode_is_f64(mode) ? 16 : 8;
    const size_t out_bytes = out_elems(layout, nfreq, ntimes, nant) * cplx;
    HV_CHECK(ensure(c, B_OUT, out_bytes));

    JobDev j{};
    j.layout = layout;
    j.beam_fstride = beam_nfreq > 1 ? beam_plane : 0;
    j.mode = mode;
    j.nant = nant; j.ntimes = ntimes; j.nfreq = nfreq; j.beam_px = beam_px; j.beam_f64 = dtype == HV_F64;
    j.kpad = kpad;
    j.antpos = (const double*)c->ws.buf[B_ANTPOS].p;
    j.inv_lambda = (const double*)c->ws.buf[B_INVLAM].p;
    j.eq2tops = (const double*)c->ws.buf[B_EQ2TOPS].p;
    j.crd_eq = (const double*)c->ws.buf[B_CRD_EQ].p;
    j.isqrt = (const double*)c->ws.buf[B_ISQRT].p;
    j.beam = c->ws.buf[B_BEAM].p;
    j.vis = c->ws.buf[B_OUT].p;
    j.amp_scale = (mode == HV_MODE_SPLIT_FP16)
                      ? fp16_scale(max_abs(bm_cube, dtype, beam_bytes / (dtype == HV_F64 ? 8 : 4)) *
                                   max_abs(isq.data(), HV_F64, isq.size()))
                      : 1.0f;
    StageTimer tm;
    HV_CHECK(run_job(c, j, tm));
    HV_CHECK(copy_d2h(c, vis_out, j.vis, out_bytes, c->stream));
    return HV_OK;
}

// --------------------------------------------------------- polarised beams
// A polarised job is the Hermitian contraction of 2N rows (feed p of antenna
// a: row p N + a) over 2 Kpad columns (sky axis k of source s: column
// 2 s + k); operands from the polarised phasor build, LDS-ring contraction,
// reduce into [F][2][2][T][N][N].
static int run_pol_job(hv_ctx* c, const JobDev& j, const void* jones, StageTimer& tm) {
    const int nv = 2 * j.nant;
    const int64_t kv = 2 * j.kpad;
    Plan p = plan_job(j.mode, nv, kv, j.ntimes, j.nfreq, operand_budget(c), LAYOUT_POL, false, true);
    HV_CHECK(prepare_workspace(c, c->ws, p, c->stream));
    hipStream_t s = c->stream;
    double* crd_top = static_cast<double*>(c->ws.buf[B_CRD_TOP].p);   // [T][3][kpad] of the real sources
    CoordsArgs ca{j.eq2tops, j.crd_eq, crd_top, j.kpad, j.ntimes, nullptr};
    HV_CHECK(tm.begin(0, s));
    HV_HIP(launch_coords(ca, s));
    HV_CHECK(tm.end(0, s));
    JobDev jv = j;                  // the contraction and reduce see the 2N x 2Kpad problem
    jv.nant = nv;
    jv.kpad = kv;
    jv.layout = LAYOUT_POL;
    for (int t0 = 0; t0 < j.ntimes; t0 += p.tb) {
        const int nt = std::min(p.tb, j.ntimes - t0);
        for (int f0 = 0; f0 < j.nfreq; f0 += p.fb) {
            const int nf = std::min(p.fb, j.nfreq - f0);
            PhasorPolArgs pa{};
            pa.crd_top = crd_top + (size_t)t0 * 3 * j.kpad;
            pa.isqrt = j.isqrt;
            pa.jones = jones;
            pa.antpos = j.antpos;
            pa.inv_lambda = j.inv_lambda + f0;
            pa.W = c->ws.buf[B_W].p;
            pa.unit_stride = p.unit_bytes;
            pa.kpad = j.kpad;
            pa.nant = j.nant;
            pa.beam_px = j.beam_px;
            pa.nfreq = nf;
            pa.ntime = nt;
            pa.in_f64 = j.beam_f64;
            pa.amp_scale = j.amp_scale;
            HV_CHECK(tm.begin(1, s));
            HV_HIP(launch_phasor_pol(j.mode, pa, s));
            HV_CHECK(tm.end(1, s));
            HV_CHECK(enqueue_contract(jv, p, c->ws.buf[B_W].p, c->ws.buf[B_PART].p, t0, nt, f0, nf, s, tm));
        }
    }
    return HV_OK;
}

extern "C" int hv_vis_pol(hv_ctx* c, int mode, int dtype, const void* antpos, int nant, const double* freqs,
                          int nfreq, const void* eq2tops, int ntimes, const void* crd_eq, const void* isqrt,
                          int64_t nsrc, const void* jones, int beam_px, void* vis_out) {
    if (!c) return fail(HV_ERR_ARG, "ctx is NULL");
    HV_CHECK(check_job_args(mode, dtype, nant, nfreq, ntimes, nsrc, beam_px));
    if (!antpos || !freqs || !eq2tops || !jones || !vis_out || (nsrc > 0 && (!crd_eq || !isqrt)))
        return fail(HV_ERR_ARG, "NULL array argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_CHECK(set_device(c));
    const int64_t kpad = round_up(std::max<int64_t>(nsrc, 1), KALIGN);
    std::vector<double> ap, eq, il, crd, isq;
    to_f64(antpos, dtype, (size_t)nant * 3, ap);
    to_f64(eq2tops, dtype, (size_t)ntimes * 9, eq);
    inv_lambdas(freqs, nfreq, il);
    pad_sky(crd_eq, isqrt, dtype, nsrc, kpad, crd, isq);
    const size_t nreal = (size_t)nant * 4 * beam_px * beam_px * 2;   // re, im of [N][2][2][P][P]
    const size_t jones_bytes = nreal * (dtype == HV_F64 ? 8 : 4);
    HV_CHECK(upload(c, B_ANTPOS, ap.data(), ap.size() * 8));
    HV_CHECK(upload(c, B_EQ2TOPS, eq.data(), eq.size() * 8));
    HV_CHECK(upload(c, B_INVLAM, il.data(), il.size() * 8));
    HV_CHECK(upload(c, B_CRD_EQ, crd.data(), crd.size() * 8));
    HV_CHECK(upload(c, B_ISQRT, isq.data(), isq.size() * 8));
    HV_CHECK(upload(c, B_BEAM, jones, jones_bytes));
    const size_t cplx = mode_is_f64(mode) ? 16 : 8;
    const size_t out_bytes = out_elems(LAYOUT_FULL, nfreq, ntimes, 2 * nant) * cplx;
    HV_CHECK(ensure(c, B_OUT, out_bytes));

    JobDev j{};
    j.layout = LAYOUT_POL;
    j.mode = mode;
    j.nant = nant; j.ntimes = ntimes; j.nfreq = nfreq; j.beam_px = beam_px; j.beam_f64 = dtype == HV_F64;
    j.kpad = kpad;
    j.antpos = (const double*)c->ws.buf[B_ANTPOS].p;
    j.inv_lambda = (const double*)c->ws.buf[B_INVLAM].p;
    j.eq2tops = (const double*)c->ws.buf[B_EQ2TOPS].p;
    j.crd_eq = (const double*)c->ws.buf[B_CRD_EQ].p;
    j.isqrt = (const double*)c->ws.buf[B_ISQRT].p;
    j.beam = c->ws.buf[B_BEAM].p;
    j.vis = c->ws.buf[B_OUT].p;
    // |J| <= sqrt(2) max |Re, Im| bounds the operand magnitude for the FP16 scale
    j.amp_scale = (mode == HV_MODE_SPLIT_FP16)
                      ? fp16_scale(1.4142135623730951 * max_abs(jones, dtype, nreal) *
                                   max_abs(isq.data(), HV_F64, isq.size()))
                      : 1.0f;
    StageTimer tm;
    HV_CHECK(run_pol_job(c, j, j.beam, tm));
    HV_CHECK(copy_d2h(c, vis_out, j.vis, out_bytes, c->stream));
    return HV_OK;
}

// --------------------------------------------------------- resident inputs
extern "C" int hv_sky_upload(hv_ctx* c, int dtype, const void* crd_eq, const void* isqrt, int64_t nsrc,
                             hv_sky** out) {
    if (!c || !out) return fail(HV_ERR_ARG, "NULL argument");
    *out = nullptr;
    HV_CHECK(check_dtype(dtype));
    if (nsrc < 0 || (nsrc > 0 && (!crd_eq || !isqrt))) return fail(HV_ERR_ARG, "bad sky arrays");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_CHECK(set_device(c));
    const int64_t kpad = round_up(std::max<int64_t>(nsrc, 1), KALIGN);
    std::vector<double> crd, isq;
    pad_sky(crd_eq, isqrt, dtype, nsrc, kpad, crd, isq);
    hv_sky* s = new hv_sky();
    s->ctx = c;
    s->nsrc = nsrc;
    s->kpad = kpad;
    s->max_isqrt = max_abs(isq.data(), HV_F64, isq.size());
    int r = dev_alloc(c, (void**)&s->crd_eq, crd.size() * 8);
    if (r == HV_OK) r = dev_alloc(c, (void**)&s->isqrt, isq.size() * 8);
    if (r != HV_OK) {
        if (s->crd_eq) dev_free(c, s->crd_eq, crd.size() * 8);
        delete s;
        return r;
    }
    HV_HIP(hipMemcpy(s->crd_eq, crd.data(), crd.size() * 8, hipMemcpyHostToDevice));
    HV_HIP(hipMemcpy(s->isqrt, isq.data(), isq.size() * 8, hipMemcpyHostToDevice));
    *out = s;
    return HV_OK;
}

extern "C" int hv_sky_free(hv_sky* s) {
    if (!s) return HV_OK;
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    dev_free(s->ctx, s->crd_eq, (size_t)3 * s->kpad * 8);
    dev_free(s->ctx, s->isqrt, (size_t)s->kpad * 8);
    delete s;
    return HV_OK;
}

extern "C" int hv_beam_upload(hv_ctx* c, int dtype, const void* bm_cube, int nant, int beam_px, hv_beam** out) {
    return hv_beam_upload_freq(c, dtype, bm_cube, 1, nant, beam_px, out);
}

extern "C" int hv_beam_upload_freq(hv_ctx* c, int dtype, const void* bm_cube, int nfreq, int nant, int beam_px,
                                   hv_beam** out) {
    if (!c || !out || !bm_cube) return fail(HV_ERR_ARG, "NULL argument");
    *out = nullptr;
    HV_CHECK(check_dtype(dtype));
    if (nant <= 0 || nfreq <= 0 || beam_px < 1 || beam_px > MAX_BEAM_PX) return fail(HV_ERR_ARG, "bad beam shape");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_CHECK(set_device(c));
    const size_t n = (size_t)nfreq * nant * beam_px * beam_px;
    const size_t bytes = n * (dtype == HV_F64 ? 8 : 4);
    hv_beam* b = new hv_beam();
    b->ctx = c;
    b->nant = nant;
    b->nfreq = nfreq;
    b->beam_px = beam_px;
    b->f64 = dtype == HV_F64;
    b->max_abs = max_abs(bm_cube, dtype, n);
    int r = dev_alloc(c, &b->data, bytes);
    if (r != HV_OK) {
        delete b;
        return r;
    }
    HV_HIP(hipMemcpy(b->data, bm_cube, bytes, hipMemcpyHostToDevice));
    *out = b;
    return HV_OK;
}

extern "C" int hv_beam_free(hv_beam* b) {
    if (!b) return HV_OK;
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    dev_free(b->ctx, b->data, (size_t)b->nfreq * b->nant * b->beam_px * b->beam_px * (b->f64 ? 8 : 4));
    delete b;
    return HV_OK;
}

extern "C" int hv_simulate(hv_ctx* c, int mode, const hv_sky* sky, int64_t nsrc_used, const hv_beam* beam,
                           int dtype, const void* antpos, int nant, const double* freqs, int nfreq,
                           const void* eq2tops, int ntimes, void* vis_out, int reps, hv_stage_times* times) {
    if (!c || !sky || !beam || !antpos || !freqs || !eq2tops) return fail(HV_ERR_ARG, "NULL argument");
    if (sky->ctx != c || beam->ctx != c) return fail(HV_ERR_ARG, "sky/beam belong to another context");
    HV_CHECK(check_job_args(mode, dtype, nant, nfreq, ntimes, nsrc_used, beam->beam_px));
    if (beam->nant != nant) return fail(HV_ERR_ARG, "beam antenna count does not match antpos");
    if (beam->nfreq != 1 && beam->nfreq != nfreq)
        return fail(HV_ERR_ARG, "a per-frequency beam needs one plane per frequency of the call");
    if (nsrc_used > sky->nsrc) return fail(HV_ERR_ARG, "nsrc_used exceeds the sky catalog");
    if (mode == HV_MODE_FP64 && !beam->f64) return fail(HV_ERR_ARG, "FP64 mode needs a float64 beam");
    if (reps < 1) reps = 1;
    std::lock_guard<std::mutex> lk(c->mu);
    HV_CHECK(set_device(c));
    // Sources beyond nsrc_used must read as zero: use a zero-padded view when
    // the catalog is truncated at a non-multiple of KALIGN.
    const int64_t kpad = round_up(std::max<int64_t>(nsrc_used, 1), KALIGN);
    const double* crd_eq = sky->crd_eq;
    const double* isq = sky->isqrt;
    if (nsrc_used != sky->nsrc) {
        HV_CHECK(ensure(c, B_CRD_EQ, (size_t)3 * kpad * 8));
        HV_CHECK(ensure(c, B_ISQRT, (size_t)kpad * 8));
        double* dc = (double*)c->ws.buf[B_CRD_EQ].p;
        double* di = (double*)c->ws.buf[B_ISQRT].p;
        HV_HIP(hipMemsetAsync(dc, 0, (size_t)3 * kpad * 8, c->stream));
        HV_HIP(hipMemsetAsync(di, 0, (size_t)kpad * 8, c->stream));
        for (int d = 0; d < 3; ++d)
            HV_HIP(hipMemcpyAsync(dc + d * kpad, sky->crd_eq + d * sky->kpad, nsrc_used * 8,
                                  hipMemcpyDeviceToDevice, c->stream));
        HV_HIP(hipMemcpyAsync(di, sky->isqrt, nsrc_used * 8, hipMemcpyDeviceToDevice, c->stream));
        crd_eq = dc;
        isq = di;
    }
    std::vector<double> ap, eq, il;
    to_f64(antpos, dtype, (size_t)nant * 3, ap);
    to_f64(eq2tops, dtype, (size_t)ntimes * 9, eq);
    inv_lambdas(freqs, nfreq, il);
    HV_CHECK(upload(c, B_ANTPOS, ap.data(), ap.size() * 8));
    HV_CHECK(upload(c, B_EQ2TOPS, eq.data(), eq.size() * 8));
    HV_CHECK(upload(c, B_INVLAM, il.data(), il.size() * 8));
    const size_t cplx = mode_is_f64(mode) ? 16 : 8;
    const size_t out_bytes = (size_t)nfreq * ntimes * nant * nant * cplx;
    HV_CHECK(ensure(c, B_OUT, out_bytes));

    JobDev j{};
    j.mode = mode;
    j.nant = nant; j.ntimes = ntimes; j.nfreq = nfreq; j.beam_px = beam->beam_px; j.beam_f64 = beam->f64;
    j.kpad = kpad;
    j.antpos = (const double*)c->ws.buf[B_ANTPOS].p;
    j.inv_lambda = (const double*)c->ws.buf[B_INVLAM].p;
    j.eq2tops = (const double*)c->ws.buf[B_EQ2TOPS].p;
    j.crd_eq = crd_eq;
    j.isqrt = isq;
    j.beam = beam->data;
    j.beam_fstride = beam->nfreq > 1 ? (size_t)nant * beam->beam_px * beam->beam_px : 0;
    j.vis = c->ws.buf[B_OUT].p;
    j.amp_scale = (mode == HV_MODE_SPLIT_FP16) ? fp16_scale(beam->max_abs * sky->max_isqrt) : 1.0f;

    // Warm allocation outside the timed loop.
    HV_HIP(hipStreamSynchronize(c->stream));
    hipEvent_t e0, e1;
    HV_HIP(hipEventCreate(&e0));
    HV_HIP(hipEventCreate(&e1));
    StageTimer tm;
    tm.on = times != nullptr;
    HV_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) HV_CHECK(run_job(c, j, tm));
    HV_HIP(hipEventRecord(e1, c->stream));
    HV_HIP(hipEventSynchronize(e1));
    float total = 0;
    HV_HIP(hipEventElapsedTime(&total, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (times) {
        HV_CHECK(tm.collect(times));
        times->total_ms = total;
    }
    if (vis_out) {
        HV_CHECK(copy_d2h(c, vis_out, j.vis, out_bytes, c->stream));
    }
    return HV_OK;
}

// ------------------------------------------------------ kernel-level entries
extern "C" int hv_k_rotate(hv_ctx* c, int dtype, const void* eq2top, const void* crd_eq, int64_t nsrc,
                           void* crd_top_out) {
    if (!c || !eq2top || !crd_top_out || (nsrc > 0 && !crd_eq)) return fail(HV_ERR_ARG, "NULL argument");
    HV_CHECK(check_dtype(dtype));
    if (nsrc <= 0) return fail(HV_ERR_ARG, "nsrc must be positive");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_CHECK(set_device(c));
    std::vector<double> eq, crd;
    to_f64(eq2top, dtype, 9, eq);
    to_f64(crd_eq, dtype, (size_t)3 * nsrc, crd);
    HV_CHECK(upload(c, B_EQ2TOPS, eq.data(), 72));
    HV_CHECK(upload(c, B_CRD_EQ, crd.data(), crd.size() * 8));
    HV_CHECK(ensure(c, B_CRD_TOP, crd.size() * 8));
    CoordsArgs a{(const double*)c->ws.buf[B_EQ2TOPS].p, (const double*)c->ws.buf[B_CRD_EQ].p,
                 (double*)c->ws.buf[B_CRD_TOP].p, nsrc, 1};
    HV_HIP(launch_coords(a, c->stream));
    std::vector<double> out(crd.size());
    HV_HIP(hipMemcpyAsync(out.data(), c->ws.buf[B_CRD_TOP].p, out.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HV_HIP(hipStreamSynchronize(c->stream));
    if (dtype == HV_F64) std::memcpy(crd_top_out, out.data(), out.size() * 8);
    else {
        float* f = static_cast<float*>(crd_top_out);
        for (size_t i = 0; i < out.size(); ++i) f[i] = (float)out[i];
    }
    return HV_OK;
}

extern "C" int hv_k_beam(hv_ctx* c, int dtype, const void* bm_cube, int nant, int beam_px, const void* crd_top,
                         int64_t nsrc, void* A_out) {
    if (!c || !bm_cube || !crd_top || !A_out) return fail(HV_ERR_ARG, "NULL argument");
    HV_CHECK(check_dtype(dtype));
    if (nant <= 0 || nsrc <= 0 || beam_px < 1 || beam_px > MAX_BEAM_PX) return fail(HV_ERR_ARG, "bad shape");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_CHECK(set_device(c));
    const size_t es = dtype == HV_F64 ? 8 : 4;
    HV_CHECK(upload(c, B_BEAM, bm_cube, (size_t)nant * beam_px * beam_px * es));
    HV_CHECK(upload(c, B_CRD_TOP, crd_top, (size_t)3 * nsrc * es));
    HV_CHECK(ensure(c, B_AUX, (size_t)nant * nsrc * es));
    BeamArgs a{c->ws.buf[B_BEAM].p, c->ws.buf[B_CRD_TOP].p, c->ws.buf[B_AUX].p, nsrc, nant, beam_px, dtype == HV_F64};
    HV_HIP(launch_beam(a, c->stream));
    HV_HIP(hipMemcpyAsync(A_out, c->ws.buf[B_AUX].p, (size_t)nant * nsrc * es, hipMemcpyDeviceToHost, c->stream));
    HV_HIP(hipStreamSynchronize(c->stream));
    return HV_OK;
}

// Decode stage-major operands of `mode` into complex values (float or double).
template <typename R>
static void decode_operands(int mode, const std::vector<uint8_t>& W, int nant, int64_t nsrc, double inv_scale,
                            R* out) {
    const int sst = stage_src(mode == HV_MODE_FP64);
    for (int a = 0; a < nant; ++a) {
        for (int64_t s = 0; s < nsrc; ++s) {
            const uint8_t* st = W.data() + operand_offset(s / sst, nant, a);
            const int i = (int)(s % sst);
            double re, im;
            if (mode == HV_MODE_FP64) {
                const double* d = reinterpret_cast<const double*>(st) + 2 * i;
                re = d[0];
                im = d[1];
            } else if (mode == HV_MODE_FP32) {
                const float* f = reinterpret_cast<const float*>(st) + 2 * i;
                re = f[0];
                im = f[1];
            } else {
                const uint16_t* g = reinterpret_cast<const uint16_t*>(st + (i / 4) * 32);
                const int e = i % 4;
                auto dec = [&](uint16_t h) -> double {
                    if (mode == HV_MODE_SPLIT_BF16) {
                        uint32_t u = (uint32_t)h << 16;
                        float f;
                        std::memcpy(&f, &u, 4);
                        return f;
                    }
                    _Float16 x;
                    std::memcpy(&x, &h, 2);
                    return (double)(float)x;
                };
                re = dec(g[2 * e]) + dec(g[8 + 2 * e]);
                im = dec(g[2 * e + 1]) + dec(g[8 + 2 * e + 1]);
            }
            out[2 * ((size_t)a * nsrc + s)] = (R)(re * inv_scale);
            out[2 * ((size_t)a * nsrc + s) + 1] = (R)(im * inv_scale);
        }
    }
}

extern "C" int hv_k_phasor(hv_ctx* c, int mode, int dtype, const void* bm_cube, int nant, int beam_px,
                           const void* crd_top, const void* isqrt, int64_t nsrc, const void* antpos, double freq,
                           void* v_out) {
    if (!c || !bm_cube || !crd_top || !isqrt || !antpos || !v_out) return fail(HV_ERR_ARG, "NULL argument");
    HV_CHECK(check_job_args(mode, dtype, nant, 1, 1, nsrc, beam_px));
    if (nsrc <= 0) return fail(HV_ERR_ARG, "nsrc must be positive");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_CHECK(set_device(c));
    const int64_t kpad = round_up(nsrc, KALIGN);
    std::vector<double> ct, isq, ap, il{freq / C_LIGHT};
    pad_sky(crd_top, isqrt, dtype, nsrc, kpad, ct, isq);
    to_f64(antpos, dtype, (size_t)nant * 3, ap);
    const size_t es = dtype == HV_F64 ? 8 : 4;
    HV_CHECK(upload(c, B_BEAM, bm_cube, (size_t)nant * beam_px * beam_px * es));
    HV_CHECK(upload(c, B_CRD_TOP, ct.data(), ct.size() * 8));
    HV_CHECK(upload(c, B_ISQRT, isq.data(), isq.size() * 8));
    HV_CHECK(upload(c, B_ANTPOS, ap.data(), ap.size() * 8));
    HV_CHECK(upload(c, B_INVLAM, il.data(), 8));
    const size_t row_bytes = operand_row_bytes(mode, kpad);
    HV_CHECK(ensure(c, B_W, row_bytes * nant));
    const float scale = (mode == HV_MODE_SPLIT_FP16)
                            ? fp16_scale(max_abs(bm_cube, dtype, (size_t)nant * beam_px * beam_px) *
                                         max_abs(isq.data(), HV_F64, isq.size()))
                            : 1.0f;
    PhasorArgs pa{};
    pa.crd_top = (const double*)c->ws.buf[B_CRD_TOP].p;
    pa.isqrt = (const double*)c->ws.buf[B_ISQRT].p;
    pa.beam = c->ws.buf[B_BEAM].p;
    pa.antpos = (const double*)c->ws.buf[B_ANTPOS].p;
    pa.inv_lambda = (const double*)c->ws.buf[B_INVLAM].p;
    pa.W = c->ws.buf[B_W].p;
    pa.unit_stride = row_bytes * nant;
    pa.kpad = kpad;
    pa.nant = nant;
    pa.beam_px = beam_px;
    pa.nfreq = 1;
    pa.ntime = 1;
    pa.amp_scale = scale;
    pa.in_f64 = dtype == HV_F64;
    HV_HIP(launch_phasor(mode, pa, c->stream));
    std::vector<uint8_t> W(row_bytes * nant);
    HV_HIP(hipMemcpyAsync(W.data(), c->ws.buf[B_W].p, W.size(), hipMemcpyDeviceToHost, c->stream));
    HV_HIP(hipStreamSynchronize(c->stream));
    if (mode == HV_MODE_FP64) decode_operands<double>(mode, W, nant, nsrc, 1.0 / scale, (double*)v_out);
    else decode_operands<float>(mode, W, nant, nsrc, 1.0 / scale, (float*)v_out);
    return HV_OK;
}

extern "C" int hv_k_herk(hv_ctx* c, int mode, const void* v, int nant, int64_t nsrc, void* vis_out) {
    if (!c || !v || !vis_out) return fail(HV_ERR_ARG, "NULL argument");
    HV_CHECK(check_mode(mode));
    if (nant <= 0 || nsrc <= 0) return fail(HV_ERR_ARG, "bad shape");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_CHECK(set_device(c));
    const bool f64 = mode == HV_MODE_FP64;
    const int64_t kpad = round_up(nsrc, KALIGN);
    const size_t cplx = f64 ? 16 : 8;
    HV_CHECK(upload(c, B_AUX, v, (size_t)nant * nsrc * cplx));
    const size_t row_bytes = operand_row_bytes(mode, kpad);
    HV_CHECK(ensure(c, B_W, row_bytes * nant));
    float scale = 1.0f;
    if (mode == HV_MODE_SPLIT_FP16) {
        const float* f = static_cast<const float*>(v);
        double m = 0;
        for (size_t i = 0; i < (size_t)nant * nsrc * 2; ++i) m = std::max(m, (double)std::fabs(f[i]));
        scale = fp16_scale(m);
    }
    SplitArgs sa{c->ws.buf[B_AUX].p, c->ws.buf[B_W].p, nsrc, kpad, nant, scale};
    HV_HIP(launch_split(mode, sa, c->stream));
    const int tile = herk_tile_operands(mode, nant);
    const int nb = (nant + tile - 1) / tile;
    const int ntiles = nb * (nb + 1) / 2;
    const int ksplit = job_ksplit(mode, nant, 1, kpad);
    const int nsplit = (int)((kpad + ksplit - 1) / ksplit);
    HV_CHECK(ensure(c, B_PART, (size_t)nsplit * ntiles * 2 * tile * tile * (f64 ? 8 : 4)));
    HV_CHECK(ensure(c, B_OUT, (size_t)nant * nant * cplx));
    HerkArgs ha{};
    ha.W = c->ws.buf[B_W].p;
    ha.unit_stride = row_bytes * nant;
    ha.partials = c->ws.buf[B_PART].p;
    ha.nant = nant;
    ha.nb = nb;
    ha.ntiles = ntiles;
    ha.kpad = kpad;
    ha.nsplit = nsplit;
    ha.ksplit = ksplit;
    ha.nunits = 1;
    HV_HIP(launch_herk(mode, tile, ha, c->stream));
    ReduceArgs ra{};
    ra.partials = c->ws.buf[B_PART].p;
    ra.vis = c->ws.buf[B_OUT].p;
    ra.t0 = 0;
    ra.f0 = 0;
    ra.nfreq_launch = 1;
    ra.ntimes_total = 1;
    ra.nant = nant;
    ra.nb = nb;
    ra.ntiles = ntiles;
    ra.nsplit = nsplit;
    ra.ksplit = ksplit;
    ra.nunits = 1;
    ra.tile = tile;
    ra.scale = 1.0 / ((double)scale * scale);
    HV_HIP(launch_reduce(mode, ra, c->stream));
    HV_HIP(hipMemcpyAsync(vis_out, c->ws.buf[B_OUT].p, (size_t)nant * nant * cplx, hipMemcpyDeviceToHost,
                          c->stream));
    HV_HIP(hipStreamSynchronize(c->stream));
    return HV_OK;
}
