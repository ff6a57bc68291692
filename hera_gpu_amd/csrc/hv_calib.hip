// Calibration microbenchmarks (no reference counterpart): a plain device copy
// of known size (HBM bandwidth, and the yardstick that scales rocprofv3
// FETCH_SIZE / WRITE_SIZE counters to bytes), and register-resident MFMA
// loops that measure the matrix-core peak of each arithmetic used by the
// contraction.
#include "hv_host.h"

using namespace hv;

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Each wave runs `iters` rounds of 8 independent MFMA chains and writes one
// value per lane so the work is not removed.
template <int MODE>
__global__ __launch_bounds__(256) void mfma_kernel(float* out, int iters, float seed) {
    const int lane = threadIdx.x & 63;
    float r = 0.f;
    if constexpr (MODE == MODE_FP32) {
        f32x16 acc[4] = {};
        float a = seed + lane, b = seed - lane;
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[k], 0, 0, 0);
        for (int k = 0; k < 4; ++k) r += acc[k][0] + acc[k][15];
    } else if constexpr (MODE == MODE_SPLIT_BF16) {
        f32x16 acc[4] = {};
        bf16x8 a, b;
        for (int e = 0; e < 8; ++e) { a[e] = (__bf16)(seed + e); b[e] = (__bf16)(seed - lane); }
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[k], 0, 0, 0);
        for (int k = 0; k < 4; ++k) r += acc[k][0] + acc[k][15];
    } else if constexpr (MODE == MODE_SPLIT_FP16) {
        f32x16 acc[4] = {};
        f16x8 a, b;
        for (int e = 0; e < 8; ++e) { a[e] = (_Float16)(seed + e); b[e] = (_Float16)(seed - lane); }
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[k], 0, 0, 0);
        for (int k = 0; k < 4; ++k) r += acc[k][0] + acc[k][15];
    } else {
        f64x4 acc[8] = {};
        double a = seed + lane, b = seed - lane;
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[k], 0, 0, 0);
        for (int k = 0; k < 8; ++k) r += (float)(acc[k][0] + acc[k][3]);
    }
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = r;
}

// flops of one kernel-loop iteration per wave
double mfma_flops_per_iter(int mode) {
    switch (mode) {
        case MODE_FP32: return 4.0 * 32 * 32 * 2 * 2;
        case MODE_SPLIT_BF16:
        case MODE_SPLIT_FP16: return 4.0 * 32 * 32 * 16 * 2;
        default: return 8.0 * 16 * 16 * 4 * 2;
    }
}

// Random finite operands in the storage format of `mode` (values in [-1, 1]).
template <int MODE>
__global__ __launch_bounds__(256) void fill_kernel(uint32_t* w, size_t nwords, uint32_t seed) {
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nwords; i += stride) {
        // FP64 operands hash per 8-byte element, the others per 4-byte word
        const uint32_t e = (uint32_t)(MODE == MODE_FP64 ? i >> 1 : i);
        uint32_t h = e * 2654435761u ^ seed;
        h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
        const float x = (float)(int)(h & 0xffff) * (1.0f / 32768.0f) - 1.0f;
        const float y = (float)(int)(h >> 16) * (1.0f / 32768.0f) - 1.0f;
        uint32_t v;
        if constexpr (MODE == MODE_SPLIT_BF16) {
            v = (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)x) | ((uint32_t)__builtin_bit_cast(uint16_t, (__bf16)y) << 16);
        } else if constexpr (MODE == MODE_SPLIT_FP16) {
            v = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x) | ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)y) << 16);
        } else if constexpr (MODE == MODE_FP32) {
            v = __float_as_uint(x);
        } else {
            const uint64_t b = __builtin_bit_cast(uint64_t, (double)x);
            v = (i & 1) ? (uint32_t)(b >> 32) : (uint32_t)b;
        }
        w[i] = v;
    }
}

}  // namespace

extern "C" int hv_bench_copy(hv_ctx* c, size_t bytes, int reps, double* ms_per_copy) {
    if (!c || !ms_per_copy || reps < 1 || bytes < 16) return fail(HV_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_HIP(hipSetDevice(c->device));
    bytes = bytes / 16 * 16;
    void *a = nullptr, *b = nullptr;
    HV_CHECK(dev_alloc(c, &a, bytes));
    int rc = dev_alloc(c, &b, bytes);
    if (rc != HV_OK) {
        dev_free(c, a, bytes);
        return rc;
    }
    const size_t n = bytes / 16;
    const int grid = 256 * 32;
    hipEvent_t e0, e1;
    rc = HV_OK;
    if (hipMemsetAsync(a, 1, bytes, c->stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        rc = fail(HV_ERR_HIP, "setup failed");
    if (rc == HV_OK) {
        hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, c->stream, (const uint4*)a, (uint4*)b, n);
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, c->stream, (const uint4*)a, (uint4*)b, n);
        (void)hipEventRecord(e1, c->stream);
        float ms = 0;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(HV_ERR_HIP, "copy timing failed");
        *ms_per_copy = ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    dev_free(c, a, bytes);
    dev_free(c, b, bytes);
    return rc;
}

extern "C" int hv_bench_mfma(hv_ctx* c, int mode, int iters, double* tflops) {
    if (!c || !tflops || iters < 1 || mode < HV_MODE_FP32 || mode > HV_MODE_FP64)
        return fail(HV_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_HIP(hipSetDevice(c->device));
    int dev = 0, cus = 0;
    HV_HIP(hipGetDevice(&dev));
    HV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = cus * 8;   // 8 workgroups of 4 waves per CU
    void* out = nullptr;
    HV_CHECK(dev_alloc(c, &out, (size_t)grid * 256 * 4));
    hipEvent_t e0, e1;
    HV_HIP(hipEventCreate(&e0));
    HV_HIP(hipEventCreate(&e1));
    auto launch = [&](int it) {
        switch (mode) {
            case MODE_FP32: hipLaunchKernelGGL(mfma_kernel<MODE_FP32>, dim3(grid), dim3(256), 0, c->stream, (float*)out, it, 1.0f); break;
            case MODE_SPLIT_BF16: hipLaunchKernelGGL(mfma_kernel<MODE_SPLIT_BF16>, dim3(grid), dim3(256), 0, c->stream, (float*)out, it, 1.0f); break;
            case MODE_SPLIT_FP16: hipLaunchKernelGGL(mfma_kernel<MODE_SPLIT_FP16>, dim3(grid), dim3(256), 0, c->stream, (float*)out, it, 1.0f); break;
            default: hipLaunchKernelGGL(mfma_kernel<MODE_FP64>, dim3(grid), dim3(256), 0, c->stream, (float*)out, it, 1.0f); break;
        }
    };
    launch(16);
    (void)hipEventRecord(e0, c->stream);
    launch(iters);
    (void)hipEventRecord(e1, c->stream);
    float ms = 0;
    int rc = HV_OK;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        rc = fail(HV_ERR_HIP, "mfma timing failed");
    const double waves = (double)grid * 4;
    *tflops = waves * iters * mfma_flops_per_iter(mode) / (ms * 1e-3) / 1e12;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dev_free(c, out, (size_t)grid * 256 * 4);
    return rc;
}

// The Hermitian contraction alone on synthetic operands (random finite values
// in [-1, 1] in the mode's storage format), for tile and prefetch variants;
// tile = pf = 0 selects the production choice.
extern "C" int hv_bench_herk(hv_ctx* c, int mode, int nant, int64_t nsrc, int units, int tile, int pf, int reps,
                             double* ms_per_launch) {
    if (!c || !ms_per_launch || nant < 1 || nsrc < 1 || units < 1 || reps < 1 || mode < HV_MODE_FP32 ||
        mode > HV_MODE_FP64)
        return fail(HV_ERR_ARG, "bad argument");
    if (tile == 192) return fail(HV_ERR_ARG, "192-tiles build their operands in place: time them through hv_simulate");
    if (tile == 0) tile = herk_tile_operands(mode, nant);   // the HBM-operand kernel
    if (pf == 5 || pf == 6) return fail(HV_ERR_ARG, "pf 5 and 6 build their operands in place: time them through hv_simulate");
    if (pf == 0) pf = herk_prefetch(mode, tile) == 5 ? 4 : herk_prefetch(mode, tile);   // same MFMA loop
    std::lock_guard<std::mutex> lk(c->mu);
    HV_HIP(hipSetDevice(c->device));
    const int64_t kpad = round_up(nsrc, KALIGN);
    const size_t unit_bytes = (size_t)nant * operand_row_bytes(mode, kpad);
    const int nb = (nant + tile - 1) / tile;
    const int ntiles = nb * (nb + 1) / 2;
    const int ksplit = job_ksplit(mode, nant, units, kpad);
    const int nsplit = (int)((kpad + ksplit - 1) / ksplit);
    const size_t part = (size_t)units * nsplit * ntiles * 2 * tile * tile * (mode == HV_MODE_FP64 ? 8 : 4);
    void *W = nullptr, *P = nullptr;
    HV_CHECK(dev_alloc(c, &W, unit_bytes * units));
    int rc = dev_alloc(c, &P, part);
    if (rc != HV_OK) {
        dev_free(c, W, unit_bytes * units);
        return rc;
    }
    HerkArgs ha{};
    ha.W = W;
    ha.unit_stride = unit_bytes;
    ha.partials = P;
    ha.nant = nant;
    ha.nb = nb;
    ha.ntiles = ntiles;
    ha.kpad = kpad;
    ha.nsplit = nsplit;
    ha.ksplit = ksplit;
    ha.nunits = units;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t nwords = unit_bytes * units / 4;
    switch (mode) {
        case HV_MODE_FP32: hipLaunchKernelGGL(fill_kernel<MODE_FP32>, dim3(4096), dim3(256), 0, c->stream, (uint32_t*)W, nwords, 12345u); break;
        case HV_MODE_SPLIT_BF16: hipLaunchKernelGGL(fill_kernel<MODE_SPLIT_BF16>, dim3(4096), dim3(256), 0, c->stream, (uint32_t*)W, nwords, 12345u); break;
        case HV_MODE_SPLIT_FP16: hipLaunchKernelGGL(fill_kernel<MODE_SPLIT_FP16>, dim3(4096), dim3(256), 0, c->stream, (uint32_t*)W, nwords, 12345u); break;
        default: hipLaunchKernelGGL(fill_kernel<MODE_FP64>, dim3(4096), dim3(256), 0, c->stream, (uint32_t*)W, nwords, 12345u); break;
    }
    if (hipGetLastError() != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(HV_ERR_HIP, "setup failed");
    if (rc == HV_OK && launch_herk_variant(mode, tile, pf, ha, c->stream) != hipSuccess)
        rc = fail(HV_ERR_ARG, "no such contraction variant (mode, tile, prefetch)");
    if (rc == HV_OK) {
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps; ++r) (void)launch_herk_variant(mode, tile, pf, ha, c->stream);
        (void)hipEventRecord(e1, c->stream);
        float ms = 0;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(HV_ERR_HIP, "contraction timing failed");
        *ms_per_launch = ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(c, W, unit_bytes * units);
    dev_free(c, P, part);
    return rc;
}

extern "C" int hv_set_herk_variant(int tile, int pf) {
    if (!(tile == 0 || tile == 64 || tile == 128 || tile == 192) || pf < 0 || pf > 6)
        return fail(HV_ERR_ARG, "tile must be 0, 64, 128 or 192 and pf 0..6");
    if (tile == 192 && pf != 0 && pf != 5 && pf != 6)
        return fail(HV_ERR_ARG, "192-tiles exist only as in-place operand builds (pf 5 or 6)");
    if (pf == 6 && tile != 192 && tile != 0)
        return fail(HV_ERR_ARG, "pf 6 (tile-pair lists) is a 192-tile variant");
    set_herk_variant(tile, pf);
    return HV_OK;
}

extern "C" int hv_job_ksplit(int mode, int nant, int64_t units, int64_t nsrc, int* ksplit) {
    if (!ksplit || nant < 1 || units < 1 || nsrc < 0) return fail(HV_ERR_ARG, "bad argument");
    if (mode < MODE_FP32 || mode > MODE_FP64) return fail(HV_ERR_ARG, "unknown contraction mode");
    *ksplit = job_ksplit(mode, nant, units, round_up(std::max<int64_t>(nsrc, 1), KALIGN));
    return HV_OK;
}

extern "C" int hv_set_ksplit_items(int64_t items) {
    if (items < 0) return fail(HV_ERR_ARG, "items must be >= 0 (0 restores the built-in threshold)");
    set_ksplit_items(items);
    return HV_OK;
}

extern "C" int hv_herk_choice(int mode, int nant, int* tile, int* pf) {
    if (!tile || !pf || nant < 1) return fail(HV_ERR_ARG, "bad argument");
    if (mode < MODE_FP32 || mode > MODE_FP64) return fail(HV_ERR_ARG, "unknown contraction mode");
    *tile = herk_tile(mode, nant);
    *pf = herk_prefetch(mode, *tile);
    return HV_OK;
}

