// Host-side internals shared by the runtime (hv_runtime.hip) and the
// serving executor (hv_serve.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "heravis.h"
#include "hv_internal.h"

namespace hv {

int fail(int code, const std::string& msg);

#define HV_HIP(expr)                                                                                    \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess)                                                                           \
            return ::hv::fail(HV_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(_e));    \
    } while (0)

#define HV_CHECK(expr)              \
    do {                            \
        int _r = (expr);            \
        if (_r != HV_OK) return _r; \
    } while (0)

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

enum Buf {
    B_ANTPOS, B_INVLAM, B_EQ2TOPS, B_CRD_EQ, B_ISQRT, B_BEAM, B_CRD_TOP, B_W, B_PART,
    B_OUT, B_AUX, B_DYN, B_EXT, B_COUNT   // B_EXT: Omnical and the median filter
};

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    size_t cap_seen = 0;   // largest capacity held so far (growth base)
};

// A set of grow-only device buffers.  The context owns one; every serving
// slot owns its own, so concurrent jobs never share scratch memory.
struct Workspace {
    DevBuf buf[B_COUNT];
    size_t bytes() const {
        size_t s = 0;
        for (auto& b : buf) s += b.cap;
        return s;
    }
};

}  // namespace hv

struct hv_ctx {
    int device = 0;
    size_t budget = 0;
    size_t used = 0;
    hipStream_t stream = nullptr;    // main stream (contraction, reduce, copies)
    hipStream_t stream2 = nullptr;   // phasor stream of the two-stream pipeline
    hipEvent_t ev[5] = {};           // pipeline events (no timing)
    std::mutex mu;       // guards `ws` and calls on `stream`
    std::mutex mem_mu;   // guards `used`
    hv::Workspace ws;
    void* stage[2] = {};           // pinned staging buffers for copies of caller memory
    hipEvent_t stage_ev[2] = {};   // last DMA that used each staging buffer
};

struct hv_sky {
    hv_ctx* ctx;
    double* crd_eq;   // [3][Kpad]
    double* isqrt;    // [Kpad]
    int64_t nsrc, kpad;
    double max_isqrt;
};

struct hv_beam {
    hv_ctx* ctx;
    void* data;       // [nfreq][N][P][P] float or double
    int nant, beam_px, f64;
    int nfreq;        // 1: one beam for every frequency
    double max_abs;
};

namespace hv {

int dev_alloc(hv_ctx* c, void** p, size_t bytes);
void dev_free(hv_ctx* c, void* p, size_t bytes);
// Grow buffer b of ws to at least `bytes` (synchronises `stream` before
// releasing an old buffer).  Returns whether the buffer moved in *moved.
int ensure(hv_ctx* c, Workspace& ws, Buf b, size_t bytes, hipStream_t stream, bool* moved = nullptr);
void release(hv_ctx* c, Workspace& ws);
// Copies between caller (pageable) host memory and the device on stream s.
// Large ones go through the context's two pinned staging buffers: the host
// side of each chunk is spread over threads while the DMA of the other runs.
// copy_h2d returns once the source may be reused; copy_d2h once dst holds the
// data (it waits for the stream's earlier work).  Calls on one context are
// serialised by its lock.
int copy_h2d(hv_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t s);
int copy_d2h(hv_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t s);
// Device address of the no-allocation entry point's scratch array on the
// current device (looked up once per device).
int noalloc_scratch(int device, void** ptr);

float fp16_scale(double max_amp);

// Launch plan of one job.
struct Plan {
    int mode = 0, nant = 0, ntimes = 0, nfreq = 0;
    int64_t kpad = 0;              // padded sources (buffers and grids)
    int tile = 0, nb = 0, ntiles = 0, nsplit = 0;
    int ksplit = 0;                // sources per K-split (job_ksplit of the whole job)
    size_t row_bytes = 0, unit_bytes = 0, part_unit = 0;
    int fb = 1, tb = 1;            // frequencies / times per batch
    bool fused = false;            // operands built inside the contraction (B_W holds A sqrt(I))
    size_t amp_time = 0;           // bytes of A sqrt(I) per time or per unit (fused)
    bool amp_unit = false;         // A sqrt(I) per unit (frequency-dependent beams)
    // device bytes per workspace buffer
    size_t crd_top_bytes = 0, w_bytes = 0, part_bytes = 0, out_bytes = 0;
    size_t footprint() const { return crd_top_bytes + w_bytes + part_bytes + out_bytes; }
};

// operands: build the operands in HBM (phasor build + LDS-ring contraction)
// even where the built-in choice builds them inside the contraction.
// ksplit: sources per K-split, job_ksplit of the whole job (0: of this
// call's shape); passed where the plan covers a part of a job's times or a
// serving bucket's padded size.
Plan plan_job(int mode, int nant, int64_t kpad, int ntimes, int nfreq, size_t operand_budget, int layout = 0,
              bool beam_per_freq = false, bool operands = false, int ksplit = 0);

struct JobDev {
    int mode;
    int nant, ntimes, nfreq, beam_px, beam_f64;
    int64_t kpad;
    const double* antpos;
    const double* inv_lambda;   // [F]
    const double* eq2tops;
    const double* crd_eq;       // [3][kpad]
    const double* isqrt;        // [kpad]
    const void* beam;
    void* vis;                  // [F][T][N][N] complex, device
    float amp_scale;
    const int64_t* dyn_kpad;    // optional device-side source count
    int layout = 0;             // output layout (Layout)
    size_t beam_fstride = 0;    // elements between per-frequency beam planes (0: one beam)
    int t_base = 0;             // time index of this job's first time in `vis`
    int ntimes_out = 0;         // times of the `vis` array (0: ntimes)
};

struct StageTimer {
    bool on = false;
    std::vector<hipEvent_t> ev[4][2];
    int begin(int stage, hipStream_t s);
    int end(int stage, hipStream_t s);
    int collect(hv_stage_times* t);
    ~StageTimer();
};

// Allocate the plan's buffers in ws (outside any graph capture).
int prepare_workspace(hv_ctx* c, Workspace& ws, const Plan& p, hipStream_t s, bool* moved = nullptr);
// Enqueue the kernels of one job on stream s (capturable: no allocation,
// no memset, no synchronisation).
int enqueue_job(const JobDev& j, const Plan& p, Workspace& ws, hipStream_t s, StageTimer& tm);
// Two-stream variant for direct calls: the phasor build of batch b+1 (on
// ctx->stream2) overlaps the contraction and reduce of batch b (on
// ctx->stream) through double-buffered operands and partials.  `p` must be
// planned with half the operand budget; ws holds 2x its W and partials.
int enqueue_job_pipelined(hv_ctx* c, const JobDev& j, const Plan& p, Workspace& ws, StageTimer& tm);

}  // namespace hv
