// hv_vis_device_noalloc: the whole simulation on caller device arrays,
// without any device allocation.  No counterpart in the reference; for callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid allocation).
//
// The kernels and their order are those of hv_vis (coordinates, A sqrt(I) or
// the phasor build, the contraction, the FP64 split reduce); only their
// working memory differs: rotated coordinates, A sqrt(I) (or operands),
// partial tiles and 1/lambda live in `scratch`, a file-scope device array of
// HV_NOALLOC_SCRATCH_BYTES per GPU, carved per call.  Times are processed in
// batches small enough for the scratch; results do not depend on batching
// (K-splits fixed by the whole job, summed in split order), so they equal
// hv_vis's bit for bit.  Every byte a kernel reads from the scratch is written earlier
// in the same call, so nothing in it has to start at zero.  All calls on one
// GPU share its array: calls that overlap on one GPU race.
#include <mutex>
#include <vector>

#include "hv_host.h"

using namespace hv;

namespace {

constexpr size_t SCRATCH_BYTES = (size_t)HV_NOALLOC_SCRATCH_BYTES;

// Working memory of hv_vis_device_noalloc.  In host code this name is a
// host-side stand-in, not a device address: the address of each device's copy
// comes from hipGetSymbolAddress (noalloc_scratch) and only that is used.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

__global__ void inv_lambda_kernel(const double* __restrict__ freqs, double* __restrict__ out, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) out[i] = freqs[i] / C_LIGHT;
}

std::mutex g_scratch_mu;
std::vector<void*> g_scratch;   // per device index

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

// Looked up once per device, with that device current (the caller has set
// it); later calls reuse the address.
int hv::noalloc_scratch(int device, void** ptr) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch.size() <= device) g_scratch.resize((size_t)device + 1, nullptr);
    if (!g_scratch[(size_t)device]) {
        void* p = nullptr;
        HV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(HV_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
        g_scratch[(size_t)device] = p;
    }
    *ptr = g_scratch[(size_t)device];
    return HV_OK;
}

extern "C" int hv_vis_device_noalloc(hv_ctx* c, int mode, const double* antpos, int nant, const double* freqs,
                                     int nfreq, const double* eq2tops, int ntimes, const double* crd_eq,
                                     const double* isqrt, int64_t nsrc, const float* bm_cube, int beam_px,
                                     void* vis_out, void* stream) {
    if (!c) return fail(HV_ERR_ARG, "NULL context");
    if (mode != HV_MODE_FP32 && mode != HV_MODE_SPLIT_BF16)
        return fail(HV_ERR_ARG, "hv_vis_device_noalloc takes HV_MODE_FP32 or HV_MODE_SPLIT_BF16");
    if (!antpos || !freqs || !eq2tops || !crd_eq || !isqrt || !bm_cube || !vis_out)
        return fail(HV_ERR_ARG, "NULL array");
    if (nant <= 0 || nfreq <= 0 || ntimes <= 0) return fail(HV_ERR_ARG, "nant, nfreq and ntimes must be positive");
    if (nsrc <= 0 || nsrc % KALIGN != 0)
        return fail(HV_ERR_ARG, "nsrc must be a positive multiple of 64 (pad with zero directions, which "
                                "are below the horizon and contribute nothing)");
    if (beam_px < 1 || beam_px > 2048) return fail(HV_ERR_ARG, "beam_px must be in [1, 2048]");
    HV_HIP(hipSetDevice(c->device));
    void* base = nullptr;
    HV_CHECK(noalloc_scratch(c->device, &base));
    char* sbase = static_cast<char*>(base);

    // carve: 1/lambda, then coordinates, A sqrt(I) or operands, partial tiles
    const int64_t kpad = nsrc;
    const size_t il_bytes = align256((size_t)nfreq * 8);
    if (il_bytes >= SCRATCH_BYTES) return fail(HV_ERR_NOMEM, "too many frequencies for the scratch");
    const size_t avail = SCRATCH_BYTES - il_bytes;
    int nt = ntimes;
    const int ksplit = job_ksplit(mode, nant, (int64_t)ntimes * nfreq, kpad);
    Plan p;
    size_t need = 0;
    for (;;) {
        const size_t crd = (size_t)nt * 3 * kpad * 8;
        const size_t wb = avail > crd ? avail - crd : 0;
        p = plan_job(mode, nant, kpad, nt, nfreq, wb, LAYOUT_FULL, false, false, ksplit);
        need = align256(p.crd_top_bytes) + align256(p.w_bytes) + align256(p.part_bytes);
        if (need <= avail) break;
        if (nt == 1)
            return fail(HV_ERR_NOMEM, "the simulation needs " + std::to_string(need + il_bytes) +
                                          " bytes of working memory for one time and one frequency batch, "
                                          "beyond the no-allocation scratch of " +
                                          std::to_string(SCRATCH_BYTES) + " bytes; use hv_vis");
        nt = (nt + 1) / 2;
    }
    Workspace ws{};
    size_t off = il_bytes;
    const Buf bufs[3] = {B_CRD_TOP, B_W, B_PART};
    const size_t sizes[3] = {p.crd_top_bytes, p.w_bytes, p.part_bytes};
    for (int i = 0; i < 3; ++i) {
        ws.buf[bufs[i]].p = sbase + off;
        ws.buf[bufs[i]].cap = sizes[i];
        off += align256(sizes[i]);
    }

    hipStream_t s = static_cast<hipStream_t>(stream);
    double* il = reinterpret_cast<double*>(sbase);
    hipLaunchKernelGGL(inv_lambda_kernel, dim3((unsigned)((nfreq + 255) / 256)), dim3(256), 0, s, freqs, il, nfreq);
    HV_HIP(hipGetLastError());

    JobDev jd{};
    jd.mode = mode;
    jd.nant = nant;
    jd.nfreq = nfreq;
    jd.beam_px = beam_px;
    jd.beam_f64 = 0;
    jd.kpad = kpad;
    jd.antpos = antpos;
    jd.inv_lambda = il;
    jd.crd_eq = crd_eq;
    jd.isqrt = isqrt;
    jd.beam = bm_cube;
    jd.vis = vis_out;
    jd.amp_scale = 1.0f;
    jd.dyn_kpad = nullptr;
    jd.layout = LAYOUT_FULL;
    jd.ntimes_out = ntimes;
    StageTimer tm;
    for (int t0 = 0; t0 < ntimes; t0 += nt) {
        jd.ntimes = std::min(nt, ntimes - t0);
        jd.eq2tops = eq2tops + (size_t)t0 * 9;
        jd.t_base = t0;
        HV_CHECK(enqueue_job(jd, p, ws, s, tm));
    }
    return HV_OK;
}
