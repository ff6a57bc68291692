// 2-D median filter (hera_gpu/median_filter.py:8-365): for every pixel of a
// batch of float32 images, the element of rank (rows*cols)/2 of the
// rows x cols window whose top-left corner is rows/2 rows and cols/2 columns
// before it, with zeros outside the image (the reference pads on the host,
// :318, and trims afterwards, :360-363; here the padding is never stored).
//
// The reference gives each thread a private copy of its window and runs a
// Floyd-Wirth selection with data-dependent branches (:96-170, :192-213).
// Here a workgroup stages a 64 x 32 output tile plus its halo in LDS as
// integer keys whose order is IEEE totalOrder (so selection is plain integer
// min/max, and -0/+0 and NaNs have a defined place), and for the windows of
// odd sides up to 11 each thread evaluates a selection network fixed at
// compile time: Batcher's odd-even merge sort (K. E. Batcher, "Sorting
// networks and their applications", AFIPS 1968) pruned, at compile time, to
// the comparators that reach the wanted ranks.  Two vertically adjacent
// outputs share all but one row of their windows: the shared rows are
// reduced once to the only ranks that can still be the median of either
// window (rank k - cols ... k of the shared part), and each output then takes
// the median of those cols + 1 keys and its own cols keys.  No branches
// depend on the data.  Other window shapes (even sides, sides 13 to 127) take
// a radix selection over the keys in LDS: 32 counting passes, one per bit.
#include <algorithm>
#include <utility>

#include "hv_host.h"

using namespace hv;

namespace {

constexpr int MF_TX = 64;        // output columns per workgroup, one per lane
constexpr int MF_TY = 32;        // output rows per workgroup: 4 row pairs per wave
constexpr int MF_TY_SMALL = 16;  // for batches of fewer than MF_SMALL_TILES 32-row tiles
constexpr int MF_SMALL_TILES = 1024;
constexpr int MF_THREADS = 256;
constexpr int MF_MAX_SIDE = 127;

struct MfArgs {
    const float* in;      // [nimg][nrows][ncols]
    float* out;           // [nimg][nrows][ncols]
    int nrows, ncols, wr, wc;
    int ty;               // output rows per tile (MF_TY or MF_TY_SMALL)
    int tiles_x;
    int64_t tiles_per_img, ntiles, chunk;   // chunk = tiles handled by one XCD
};

// ---- selection networks, built at compile time -----------------------------

struct Cmp {
    short i, j;    // i < j: the smaller key goes to i, the larger to j
    bool lo, hi;   // whether the min (to i) / max (to j) is still needed
};

constexpr int pow2_ceil(int n) {
    int p = 1;
    while (p < n) p *= 2;
    return p;
}

// Batcher's odd-even merge sort of n = 2^m keys, comparators in network order.
template <typename F>
constexpr void batcher(int n, F&& emit) {
    for (int p = 1; p < n; p *= 2)
        for (int k = p; k >= 1; k /= 2)
            for (int j = k % p; j + k < n; j += 2 * k)
                for (int i = 0; i < k && i + j + k < n; ++i)
                    if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) emit(i + j, i + j + k);
}

constexpr int batcher_size(int n) {
    int c = 0;
    batcher(n, [&](int, int) { ++c; });
    return c;
}

// Comparators of a sort of W keys that decide ranks LO..HI.  The W keys sit
// in a network of P = 2^m inputs whose others hold +infinity; a comparator
// touching one of those moves nothing and is dropped.  Walking backwards from
// the wanted ranks keeps only comparators whose min or max is read later.
template <int W, int LO, int HI>
struct SelectNet {
    static constexpr int P = pow2_ceil(W);
    static constexpr int CAP = batcher_size(P) > 0 ? batcher_size(P) : 1;
    struct Ops {
        Cmp op[CAP];
        int n;
    };
    static constexpr Ops make() {
        Cmp all[CAP] = {};
        int na = 0;
        batcher(P, [&](int a, int b) {
            if (b < W) all[na++] = Cmp{(short)a, (short)b, true, true};
        });
        bool need[P] = {};
        for (int q = LO; q <= HI; ++q) need[q] = true;
        Cmp kept[CAP] = {};
        int nk = 0;
        for (int q = na - 1; q >= 0; --q) {
            const bool lo = need[all[q].i], hi = need[all[q].j];
            if (lo || hi) {
                kept[nk++] = Cmp{all[q].i, all[q].j, lo, hi};
                need[all[q].i] = need[all[q].j] = true;
            }
        }
        Ops o{};
        o.n = nk;
        for (int q = 0; q < nk; ++q) o.op[q] = kept[nk - 1 - q];
        return o;
    }
    static constexpr Ops ops = make();
};

template <int I, int J, bool LO, bool HI>
__device__ __forceinline__ void cmp_step(int* v) {
    const int a = v[I], b = v[J];
    if constexpr (LO) v[I] = min(a, b);
    if constexpr (HI) v[J] = max(a, b);
}

template <typename Net, int... Q>
__device__ __forceinline__ void run_net(int* v, std::integer_sequence<int, Q...>) {
    (cmp_step<Net::ops.op[Q].i, Net::ops.op[Q].j, Net::ops.op[Q].lo, Net::ops.op[Q].hi>(v), ...);
}

// Leaves ranks LO..HI of v[0..W) in v[LO..HI].
template <int W, int LO, int HI>
__device__ __forceinline__ void select_ranks(int* v) {
    using Net = SelectNet<W, LO, HI>;
    run_net<Net>(v, std::make_integer_sequence<int, Net::ops.n>{});
}

// ---- tiles -----------------------------------------------------------------

// float32 bits -> int32 key whose signed order is IEEE totalOrder (an
// involution: the same map turns a key back into the float's bits).
__device__ __forceinline__ int total_order_key(int u) { return u ^ ((u >> 31) & 0x7fffffff); }

struct TileAt {
    const float* img;
    float* out;
    int row0, col0;
};

// Blocks are dealt to the 8 XCDs in turn; block b handles tile
// (b % 8) * chunk + b / 8, so each XCD walks one contiguous run of row-major
// tiles and the halo rows shared by vertical neighbours stay in its L2.
__device__ __forceinline__ bool tile_at(const MfArgs& a, TileAt& t) {
    const int64_t b = blockIdx.x;
    const int64_t tile = (b % 8) * a.chunk + b / 8;
    if (tile >= a.ntiles) return false;
    const int64_t img = tile / a.tiles_per_img;
    const int64_t rem = tile - img * a.tiles_per_img;
    const int ty = (int)(rem / a.tiles_x), tx = (int)(rem - (int64_t)ty * a.tiles_x);
    const int64_t plane = (int64_t)a.nrows * a.ncols;
    t.img = a.in + img * plane;
    t.out = a.out + img * plane;
    t.row0 = ty * a.ty;
    t.col0 = tx * MF_TX;
    return true;
}

// Keys of the tile and its halo, zero (the key of +0.0) outside the image.
__device__ __forceinline__ void stage(const MfArgs& a, const TileAt& t, int* lds, int lh, int lw, int hr, int hc) {
    for (int e = threadIdx.x; e < lh * lw; e += MF_THREADS) {
        const int r = e / lw, c = e - r * lw;
        const int gr = t.row0 - hr + r, gc = t.col0 - hc + c;
        int key = 0;
        if (gr >= 0 && gr < a.nrows && gc >= 0 && gc < a.ncols)
            key = total_order_key(__float_as_int(t.img[(int64_t)gr * a.ncols + gc]));
        lds[e] = key;
    }
}

__device__ __forceinline__ void put(const MfArgs& a, const TileAt& t, int r, int c, int key) {
    if (t.row0 + r < a.nrows && t.col0 + c < a.ncols)
        t.out[(int64_t)(t.row0 + r) * a.ncols + t.col0 + c] = __int_as_float(total_order_key(key));
}

// Windows of WX rows x WY columns fixed at compile time, tiles of TY rows
// (TY / 8 row pairs per thread).
template <int WX, int WY, int TY>
__global__ __launch_bounds__(MF_THREADS) void medfilt_kernel(MfArgs a) {
    constexpr int LW = MF_TX + WY - 1, LH = TY + WX - 1;
    constexpr int W = WX * WY, K = W / 2;
    constexpr int PAIRS = TY / 2 / (MF_THREADS / MF_TX);
    __shared__ int lds[LH * LW];
    TileAt t;
    if (!tile_at(a, t)) return;    // whole workgroup, before any barrier
    stage(a, t, lds, LH, LW, WX / 2, WY / 2);
    __syncthreads();
    const int c = threadIdx.x & (MF_TX - 1);
    const int* col = lds + c;
#pragma unroll 1
    for (int p = 0; p < PAIRS; ++p) {
        const int r0 = 2 * (PAIRS * (threadIdx.x / MF_TX) + p);   // window rows r0 .. r0 + WX: two outputs
        int res0, res1;
        if constexpr (WX >= 3) {
            // rows r0 + 1 .. r0 + WX - 1 are in both windows: keep ranks K - WY .. K
            constexpr int C = (WX - 1) * WY;
            int v[C];
#pragma unroll
            for (int i = 0; i < WX - 1; ++i)
#pragma unroll
                for (int j = 0; j < WY; ++j) v[i * WY + j] = col[(r0 + 1 + i) * LW + j];
            select_ranks<C, K - WY, K>(v);
            int u[2 * WY + 1], w[2 * WY + 1];
#pragma unroll
            for (int q = 0; q <= WY; ++q) u[q] = w[q] = v[K - WY + q];
#pragma unroll
            for (int j = 0; j < WY; ++j) {
                u[WY + 1 + j] = col[r0 * LW + j];
                w[WY + 1 + j] = col[(r0 + WX) * LW + j];
            }
            // the K - WY keys below the kept ranks are below both medians: the
            // median of each window is rank WY of its 2 WY + 1 remaining keys
            select_ranks<2 * WY + 1, WY, WY>(u);
            select_ranks<2 * WY + 1, WY, WY>(w);
            res0 = u[WY];
            res1 = w[WY];
        } else {
            int u[WY], w[WY];
#pragma unroll
            for (int j = 0; j < WY; ++j) {
                u[j] = col[r0 * LW + j];
                w[j] = col[(r0 + 1) * LW + j];
            }
            select_ranks<WY, K, K>(u);
            select_ranks<WY, K, K>(w);
            res0 = u[K];
            res1 = w[K];
        }
        put(a, t, r0, c, res0);
        put(a, t, r0 + 1, c, res1);
    }
}

// Any window up to MF_MAX_SIDE per side: the rank-K key by a binary search on
// its bits, counting the window keys below each candidate.
__global__ __launch_bounds__(MF_THREADS) void medfilt_generic_kernel(MfArgs a) {
    extern __shared__ int lds_dyn[];
    const int lw = MF_TX + a.wc - 1, lh = a.ty + a.wr - 1;
    const int K = (a.wr * a.wc) / 2;
    const int pairs = a.ty / 2 / (MF_THREADS / MF_TX);
    TileAt t;
    if (!tile_at(a, t)) return;
    stage(a, t, lds_dyn, lh, lw, a.wr / 2, a.wc / 2);
    __syncthreads();
    const int c = threadIdx.x & (MF_TX - 1);
    const int r0 = 2 * pairs * (threadIdx.x / MF_TX);
    for (int o = 0; o < 2 * pairs; ++o) {
        const int* win = lds_dyn + (r0 + o) * lw + c;
        unsigned p = 0;   // unsigned key (sign bit flipped) of the answer, built from the top bit
        for (int b = 31; b >= 0; --b) {
            const unsigned cand = p | (1u << b);
            int below = 0;
            for (int i = 0; i < a.wr; ++i)
                for (int j = 0; j < a.wc; ++j) below += ((unsigned)win[i * lw + j] ^ 0x80000000u) < cand;
            if (below <= K) p = cand;
        }
        put(a, t, r0 + o, c, (int)(p ^ 0x80000000u));
    }
}

using KernelFn = void (*)(MfArgs);

template <int WX, int TY>
KernelFn pick_cols(int wc) {
    switch (wc) {
        case 1: return medfilt_kernel<WX, 1, TY>;
        case 3: return medfilt_kernel<WX, 3, TY>;
        case 5: return medfilt_kernel<WX, 5, TY>;
        case 7: return medfilt_kernel<WX, 7, TY>;
        case 9: return medfilt_kernel<WX, 9, TY>;
        case 11: return medfilt_kernel<WX, 11, TY>;
        default: return nullptr;
    }
}

template <int TY>
KernelFn pick_ty(int wr, int wc) {
    switch (wr) {
        case 1: return pick_cols<1, TY>(wc);
        case 3: return pick_cols<3, TY>(wc);
        case 5: return pick_cols<5, TY>(wc);
        case 7: return pick_cols<7, TY>(wc);
        case 9: return pick_cols<9, TY>(wc);
        case 11: return pick_cols<11, TY>(wc);
        default: return nullptr;
    }
}

KernelFn pick(int wr, int wc, int ty) {
    return ty == MF_TY_SMALL ? pick_ty<MF_TY_SMALL>(wr, wc) : pick_ty<MF_TY>(wr, wc);
}

int check_args(hv_ctx* c, int nimg, int nrows, int ncols, int wr, int wc) {
    if (!c) return fail(HV_ERR_ARG, "NULL context");
    if (nimg < 1 || nrows < 1 || ncols < 1) return fail(HV_ERR_ARG, "medfilt2d: empty image batch");
    if (wr < 1 || wc < 1 || wr > MF_MAX_SIDE || wc > MF_MAX_SIDE)
        return fail(HV_ERR_ARG, "medfilt2d: window sides must be in [1, 127]");
    const int64_t tiles = (int64_t)nimg * ((nrows + MF_TY - 1) / MF_TY) * ((ncols + MF_TX - 1) / MF_TX);
    if (tiles + 8 > INT32_MAX) return fail(HV_ERR_ARG, "medfilt2d: batch too large for one launch");
    return HV_OK;
}

// Enqueue the filter of nimg device images on stream s.
int launch_medfilt(const float* in, float* out, int nimg, int nrows, int ncols, int wr, int wc, hipStream_t s) {
    MfArgs a{};
    a.in = in;
    a.out = out;
    a.nrows = nrows;
    a.ncols = ncols;
    a.wr = wr;
    a.wc = wc;
    a.tiles_x = (ncols + MF_TX - 1) / MF_TX;
    // batches of fewer than MF_SMALL_TILES 32-row tiles (4 workgroups per CU)
    // take 16-row tiles: twice the workgroups for the same halo per column
    const int64_t tiles32 = (int64_t)a.tiles_x * ((nrows + MF_TY - 1) / MF_TY) * nimg;
    a.ty = tiles32 < MF_SMALL_TILES ? MF_TY_SMALL : MF_TY;
    a.tiles_per_img = (int64_t)a.tiles_x * ((nrows + a.ty - 1) / a.ty);
    a.ntiles = a.tiles_per_img * nimg;
    a.chunk = (a.ntiles + 7) / 8;
    const dim3 grid((unsigned)(8 * a.chunk)), block(MF_THREADS);
    if (KernelFn fn = pick(wr, wc, a.ty)) {
        hipLaunchKernelGGL(fn, grid, block, 0, s, a);
    } else {
        const size_t lds = (size_t)(a.ty + wr - 1) * (MF_TX + wc - 1) * sizeof(int);
        if (lds > 64 * 1024)
            HV_HIP(hipFuncSetAttribute((const void*)medfilt_generic_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(medfilt_generic_kernel, grid, block, lds, s, a);
    }
    HV_HIP(hipGetLastError());
    return HV_OK;
}

}  // namespace

extern "C" int hv_medfilt2d(hv_ctx* c, const float* images, float* out, int nimg, int nrows, int ncols, int wrows,
                            int wcols) {
    HV_CHECK(check_args(c, nimg, nrows, ncols, wrows, wcols));
    if (!images || !out) return fail(HV_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_HIP(hipSetDevice(c->device));
    const size_t bytes = (size_t)nimg * nrows * ncols * sizeof(float);
    const size_t slice = (bytes + 255) / 256 * 256;
    HV_CHECK(ensure(c, c->ws, B_EXT, 2 * slice, c->stream));
    char* base = static_cast<char*>(c->ws.buf[B_EXT].p);
    float* d_in = (float*)base;
    float* d_out = (float*)(base + slice);
    HV_CHECK(copy_h2d(c, d_in, images, bytes, c->stream));
    HV_CHECK(launch_medfilt(d_in, d_out, nimg, nrows, ncols, wrows, wcols, c->stream));
    return copy_d2h(c, out, d_out, bytes, c->stream);
}

extern "C" int hv_bench_medfilt2d(hv_ctx* c, const float* images, int nimg, int nrows, int ncols, int wrows,
                                  int wcols, int reps, double* ms_per_launch) {
    HV_CHECK(check_args(c, nimg, nrows, ncols, wrows, wcols));
    if (!images || !ms_per_launch || reps < 1) return fail(HV_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_HIP(hipSetDevice(c->device));
    const size_t bytes = (size_t)nimg * nrows * ncols * sizeof(float);
    const size_t slice = (bytes + 255) / 256 * 256;
    HV_CHECK(ensure(c, c->ws, B_EXT, 2 * slice, c->stream));
    char* base = static_cast<char*>(c->ws.buf[B_EXT].p);
    float* d_in = (float*)base;
    float* d_out = (float*)(base + slice);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto run = [&]() -> int {
        HV_HIP(hipEventCreate(&e0));
        HV_HIP(hipEventCreate(&e1));
        HV_CHECK(copy_h2d(c, d_in, images, bytes, c->stream));
        HV_CHECK(launch_medfilt(d_in, d_out, nimg, nrows, ncols, wrows, wcols, c->stream));   // warm-up
        HV_HIP(hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps; ++r)
            HV_CHECK(launch_medfilt(d_in, d_out, nimg, nrows, ncols, wrows, wcols, c->stream));
        HV_HIP(hipEventRecord(e1, c->stream));
        HV_HIP(hipEventSynchronize(e1));
        float ms = 0;
        HV_HIP(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_launch = ms / reps;
        return HV_OK;
    };
    const int rc = run();
    (void)hipStreamSynchronize(c->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}
