// Hermitian visibility contraction V = v v^H on gfx950 MFMA.
//
// Replaces the full Cgemm/Zgemm of hera_gpu/vis.py:246-250 plus the host
// chunk sum of vis.py:258.  Only tiles (bi <= bj) of the upper triangle are
// computed.
//
// Formulation.  A phasor row is stored interleaved, w[a] = (x_0, y_0, x_1,
// y_1, ...) with v = x + i y.  Then
//     Re V[a][b] = w[a] . w[b]                      (sum x_a x_b + y_a y_b)
//     Im V[a][b] = sum y_a x_b - x_a y_b
// are real dot products that share the loaded fragments.  For FP32 and FP64
// the two halves of Im run as separate MFMA chains P1 = sum y_a x_b and
// P2 = sum x_a y_b (odd k-slot of A against even k-slot of B and vice versa)
// whose difference is exactly zero for equal rows, as the reference's
// co-located test requires (test_vis_gpu.py:218-221).  For the split modes a
// lane holds 8 halves per MFMA, so Im is one chain of w'[a] . w[b] with the
// pair rotation w' = (y_0, -x_0, y_1, -x_1, ...) formed in registers.
//
// Arithmetic modes:
//   FP32        v_mfma_f32_32x32x2_f32 on complex64 rows;
//   SPLIT_BF16  rows split as hi + lo bf16; three bf16 products per term
//               (hi.hi + hi.lo + lo.hi) on v_mfma_f32_32x32x16_bf16;
//   SPLIT_FP16  as above with fp16 halves (rows pre-scaled by a power of 2);
//   FP64        v_mfma_f64_16x16x4_f64 on complex128 rows.
// 64-tiles fold their FP32 accumulators into a second FP32 register total
// every 512 sources (short MFMA accumulation chains); K is cut into fixed
// splits of job_ksplit(...) sources (8192, or 32768 for split-mode jobs
// that still hold 32768 work items with them) whose partial tiles are summed in FP64, in
// split order, by the reduce stage, so results do not depend on how jobs are
// padded or batched.
//
// Workgroup: 4 waves (2 x 2) on a TILE x TILE complex tile; a K stage is
// 128 bytes of every row (16 complex64 sources, 8 complex128), staged
// through LDS with a padded 144-byte row stride (conflict-free 16-byte
// reads) and double buffering.  Workgroups are ordered so the tiles of one
// K split run back to back on one XCD (shared L2 lines).
#include "hv_internal.h"
#include "hv_phasor_math.h"

namespace hv {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int LDS_STRIDE = 144;

// Built-in variants (measured on MI355X): split modes build their operands
// inside the contraction, in items of several frequency units at 33..128
// antennas (herk_mu_kernel), in the one or two items of the triangle
// decomposition at 129..352 (herk_t192_kernel with run or triangle/strip
// lists) and in 192-tiles above; 64-tiles up to 32 antennas and with a beam
// per frequency below 129.  Contractions of operands already in HBM take
// the LDS ring with 128-tiles above 64 antennas (the tiles of a split share
// their stages in L2: reads ~1.1x the operand bytes against ~3.5x for
// 64-tiles).  FP32 and FP64 keep the register-staged 64-tile kernel, their
// fastest.
static int g_tile_override = 0, g_pf_override = 0;   // 0 = built-in choice

void set_herk_variant(int tile, int pf) {
    g_tile_override = tile;
    g_pf_override = pf;
}

static bool split_mode(int mode) { return mode == MODE_SPLIT_BF16 || mode == MODE_SPLIT_FP16; }

// pf of the built-in choice or of the override where it exists for the mode
// (pf 5, operands built in place, is a split-mode variant; FP64 has no pf 2).
static int pf_choice(int mode) {
    // pf 6: in place (pf 5) with the 192-tile kernel's tile-pair lists only
    const int pfo = g_pf_override == 6 ? 5 : g_pf_override;
    if (pfo && (pfo != 5 || split_mode(mode)) && !(mode == MODE_FP64 && pfo == 2)) return pfo;
    // split modes build their operands in place (whole simulations, 2^18
    // sources: faster than the phasor build + LDS ring at every array size
    // measured, 1.2x at N = 37 and 350, 1.6x at N = 128, and no operand buffer)
    return split_mode(mode) ? 5 : 1;
}

// Tile edge: 64 (FP32, FP64, small arrays), 128 (split modes at 64 < N <=
// 128), 192 (split modes above 128 antennas; in-place builds only: the
// items of the triangle decomposition up to 352 antennas, 192-tiles above).
// Whole simulations on MI355X, split-BF16, 2^20 sources, before the triangle
// decomposition: N = 160 258 TFLOP/s with 192-tiles against 120 with
// 128-tiles (two 128-blocks, the second a quarter full); N = 96, 128, 200,
// 256: 128-tiles 148 / 240 / 186 / 271 against 108 / 190 / 118 / 192 with
// 192-tiles; N = 300: 261 against 184; N = 37 and 64: 64-tiles 45 / 130
// against 22 / 67.  Overrides apply where the variant exists for the mode.
int herk_tile(int mode, int nant) {
    // 129..352 antennas: the items of the triangle decomposition, which store
    // into 192-tiles; pf 6 keeps 128-tiles at 193..256
    const bool t192 = nant > 256 || (nant > 128 && (nant <= 192 || g_pf_override != 6));
    int t = split_mode(mode) ? (t192 ? 192 : nant > 64 ? 128 : 64) : 64;
    if (g_tile_override && (g_tile_override == 64 || split_mode(mode))) t = g_tile_override;
    if (t == 192 && pf_choice(mode) != 5) t = 128;
    return t;
}

// (x, y) pairs of 16-bit values -> (-y, x): rotate each 32-bit word by 16 and
// flip the sign bit of the lower half.  Applied to the B operand b = (x, y)
// of a split-mode contraction, a . rot(b) = a_x (-b_y) + a_y b_x, the
// imaginary part of a conj(b).  The 192-tile kernel rotates A instead where
// that side has fewer fragments (rot_a16: the same two products in swapped
// halves of each 32-bit pair); the matrix cores return the same bits either
// way on gfx950, which the variant tests check (the 128- and 192-tile
// kernels' results are bitwise equal; 64-tiles fold their accumulators every
// FOLD_SRC sources, so their bits differ at the 1e-6 level).
__device__ inline u32x4 rot_neg16(u32x4 v) {
    u32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = __builtin_amdgcn_alignbit(v[i], v[i], 16) ^ 0x00008000u;
    return r;
}

template <int MODE>
__device__ inline f32x16 mfma16(u32x4 a, u32x4 b, f32x16 c) {
    if constexpr (MODE == MODE_SPLIT_BF16)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                      __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

__device__ inline void decode_tile(int t, int nb, int& bi, int& bj) {
    int i = 0;
    while (t >= nb - i) { t -= nb - i; ++i; }
    bi = i;
    bj = i + t;
}

// Work decode.  Per unit there are nsplit * ntiles workgroups; with nsplit a
// multiple of 8 the hardware's round-robin XCD assignment (block id mod 8)
// gives XCD x the splits x, x+8, ... with all their tiles consecutive.
__device__ inline void decode_work(int nsplit, int ntiles, int64_t L, int& unit, int& split, int& tile) {
    const int64_t per = (int64_t)nsplit * ntiles;
    unit = (int)(L / per);
    const int64_t rem = L - (int64_t)unit * per;
    if ((nsplit & 7) == 0) {
        const int x = (int)(rem & 7);
        const int64_t p = rem >> 3;
        split = x + 8 * (int)(p / ntiles);
        tile = (int)(p % ntiles);
    } else {
        split = (int)(rem / ntiles);
        tile = (int)(rem % ntiles);
    }
}
__device__ inline void decode_work(const HerkArgs& a, int64_t L, int& unit, int& split, int& tile) {
    decode_work(a.nsplit, a.ntiles, L, unit, split, tile);
}

// Compile-time shape of one kernel variant.
template <int MODE, int TILE>
struct Cfg {
    static constexpr bool F64 = (MODE == MODE_FP64);
    static constexpr int BLK = F64 ? 16 : 32;             // MFMA block edge
    static constexpr int WT = TILE / 2;                   // wave tile edge (complex)
    static constexpr int NS = WT / BLK;                   // subtiles per wave per dim
    static constexpr int NCH = TILE * (STAGE_BYTES / 16) / 256;   // 16-B chunks / thread / operand
    static constexpr int SRC_STAGE = stage_src(F64);
    // 64-tiles fold the MFMA accumulators into FP32 register totals every
    // FOLD_SRC sources; 128-tiles (split modes only) have no registers to
    // spare for totals and accumulate each K-split in one chain.
    static constexpr bool FOLD_ON = !F64 && TILE == 64;
    static constexpr int FOLD = FOLD_ON ? FOLD_SRC / SRC_STAGE : (1 << 30);
    // Im as P1 - P2 (FP32, FP64) or as one chain of w'.w (split modes)
    static constexpr bool PAIR_IM = (MODE == MODE_FP32 || MODE == MODE_FP64);
    using ACC = typename std::conditional<F64, f64x4, f32x16>::type;
    using AR = typename std::conditional<F64, double, float>::type;
    static constexpr int NACC = F64 ? 4 : 16;
};

// MFMAs of one LDS stage for one wave: Re += w_a . w_b and the Im chain(s)
// for every active subtile (bit i*NS+j of smask).
// Byte offset of 16-byte chunk q of tile row r in an LDS operand buffer:
// padded rows (register-staged kernel) or 128-byte rows with the chunk index
// XOR-swizzled by the row (ring kernel; conflict-free for 8 consecutive rows).
template <bool SWZ>
__device__ __forceinline__ int lds_off(int r, int q) {
    if constexpr (SWZ) return r * STAGE_BYTES + ((q ^ (r & 7)) << 4);
    else return r * LDS_STRIDE + (q << 4);
}

template <int MODE, int TILE, bool ALLON, bool SWZ = false>
__device__ __forceinline__ void stage_mma(const char* __restrict__ As, const char* __restrict__ Bs, int wr, int wc,
                                          int lrow, int lh, unsigned smask,
                                          typename Cfg<MODE, TILE>::ACC (&re)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS],
                                          typename Cfg<MODE, TILE>::ACC (&im)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS],
                                          typename Cfg<MODE, TILE>::ACC (&p2)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS]) {
    using C = Cfg<MODE, TILE>;
    constexpr int NS = C::NS, BLK = C::BLK, WT = C::WT;
#define ON(i, j) (ALLON || (smask & (1u << ((i) * NS + (j)))))
    if constexpr (MODE == MODE_FP32) {
        // lane half lh holds (x_s, y_s, x_s+1, y_s+1), s = 4q + 2 lh
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            f32x4 av[NS], bv[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                av[i] = *reinterpret_cast<const f32x4*>(As + lds_off<SWZ>(wr * WT + i * BLK + lrow, 2 * q + lh));
                bv[i] = *reinterpret_cast<const f32x4*>(Bs + lds_off<SWZ>(wc * WT + i * BLK + lrow, 2 * q + lh));
            }
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j < NS; ++j)
                    if (ON(i, j)) {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            re[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i][e], bv[j][e], re[i][j], 0, 0, 0);
#pragma unroll
                        for (int e = 0; e < 4; e += 2) {
                            im[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i][e + 1], bv[j][e], im[i][j], 0, 0, 0);
                            p2[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i][e], bv[j][e + 1], p2[i][j], 0, 0, 0);
                        }
                    }
        }
    } else if constexpr (MODE == MODE_SPLIT_BF16 || MODE == MODE_SPLIT_FP16) {
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
            const int g = 2 * jj + lh;
            u32x4 ahi[NS], alo[NS], bhi[NS], blo[NS], bhr[NS], blr[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                const int ra = wr * WT + i * BLK + lrow, rb = wc * WT + i * BLK + lrow;
                ahi[i] = *reinterpret_cast<const u32x4*>(As + lds_off<SWZ>(ra, 2 * g));
                alo[i] = *reinterpret_cast<const u32x4*>(As + lds_off<SWZ>(ra, 2 * g + 1));
                bhi[i] = *reinterpret_cast<const u32x4*>(Bs + lds_off<SWZ>(rb, 2 * g));
                blo[i] = *reinterpret_cast<const u32x4*>(Bs + lds_off<SWZ>(rb, 2 * g + 1));
                bhr[i] = rot_neg16(bhi[i]);
                blr[i] = rot_neg16(blo[i]);
            }
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j < NS; ++j)
                    if (ON(i, j)) {
                        re[i][j] = mfma16<MODE>(ahi[i], bhi[j], re[i][j]);
                        im[i][j] = mfma16<MODE>(ahi[i], bhr[j], im[i][j]);
                        re[i][j] = mfma16<MODE>(ahi[i], blo[j], re[i][j]);
                        im[i][j] = mfma16<MODE>(ahi[i], blr[j], im[i][j]);
                        re[i][j] = mfma16<MODE>(alo[i], bhi[j], re[i][j]);
                        im[i][j] = mfma16<MODE>(alo[i], bhr[j], im[i][j]);
                    }
        }
    } else {  // FP64: lane quarter lh holds (x_s, y_s, x_s+1, y_s+1), s = 2 lh
        f64x4 av[NS], bv[NS];
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const int ra = wr * WT + i * BLK + lrow, rb = wc * WT + i * BLK + lrow;
            const f64x2 a0 = *reinterpret_cast<const f64x2*>(As + lds_off<SWZ>(ra, 2 * lh));
            const f64x2 a1 = *reinterpret_cast<const f64x2*>(As + lds_off<SWZ>(ra, 2 * lh + 1));
            const f64x2 b0 = *reinterpret_cast<const f64x2*>(Bs + lds_off<SWZ>(rb, 2 * lh));
            const f64x2 b1 = *reinterpret_cast<const f64x2*>(Bs + lds_off<SWZ>(rb, 2 * lh + 1));
            av[i] = f64x4{a0[0], a0[1], a1[0], a1[1]};
            bv[i] = f64x4{b0[0], b0[1], b1[0], b1[1]};
        }
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int j = 0; j < NS; ++j)
                if (ON(i, j)) {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i][e], bv[j][e], re[i][j], 0, 0, 0);
#pragma unroll
                    for (int e = 0; e < 4; e += 2) {
                        im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i][e + 1], bv[j][e], im[i][j], 0, 0, 0);
                        p2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i][e], bv[j][e + 1], p2[i][j], 0, 0, 0);
                    }
                }
    }
#undef ON
}

// Loader state of one thread (see herk_kernel): two register sets so that
// the loads of stage st+2 are in flight while stage st is computed and stage
// st+1 is written to LDS.
template <int NCH>
struct LoopState {
    const char* gA;
    const char* gB;
    size_t stage_bytes;
    u32x4 rA[2][NCH], rB[2][NCH];
    uint32_t mA[NCH], mB[NCH];
    int roffA[NCH], roffB[NCH], loff[NCH];
};

template <int NCH, int SET>
__device__ __forceinline__ void load_stage(LoopState<NCH>& ls, int st) {
    const size_t o = (size_t)st * ls.stage_bytes;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        ls.rA[SET][c] = *reinterpret_cast<const u32x4*>(ls.gA + o + ls.roffA[c]);
        ls.rB[SET][c] = *reinterpret_cast<const u32x4*>(ls.gB + o + ls.roffB[c]);
    }
}

template <int NCH, int SET>
__device__ __forceinline__ void store_stage(LoopState<NCH>& ls, char* dst, int op_bytes) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        *reinterpret_cast<u32x4*>(dst + ls.loff[c]) = ls.rA[SET][c] & ls.mA[c];
        *reinterpret_cast<u32x4*>(dst + op_bytes + ls.loff[c]) = ls.rB[SET][c] & ls.mB[c];
    }
}

// Two stages (st in LDS buffer 0, st+1 in buffer 1) of one accumulation
// group.  PF = 2: stage st's operands are in LDS and stage st+1's in register
// set 1 on entry, and each half issues the loads two stages ahead.  PF = 1:
// only stage st is staged on entry and each half loads one stage ahead.
template <int MODE, int TILE, int KIND, int PF>
__device__ __forceinline__ void two_stages(LoopState<Cfg<MODE, TILE>::NCH>& ls, char* lds, int st, int s_end, int wr,
                                           int wc, int lrow, int lh, unsigned smask,
                                           typename Cfg<MODE, TILE>::ACC (&re)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS],
                                           typename Cfg<MODE, TILE>::ACC (&im)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS],
                                           typename Cfg<MODE, TILE>::ACC (&p2)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS]) {
    constexpr int NCH = Cfg<MODE, TILE>::NCH;
    constexpr int OP = TILE * LDS_STRIDE;        // bytes of one operand buffer
    char* b0 = lds;
    char* b1 = lds + 2 * OP;
    __syncthreads();
    if constexpr (PF == 2) {
        if (st + 2 < s_end) load_stage<NCH, 0>(ls, st + 2);
    } else {
        load_stage<NCH, 1>(ls, st + 1);                // s_end - st is even: st + 1 < s_end
    }
    if constexpr (KIND != 0) stage_mma<MODE, TILE, KIND == 1>(b0, b0 + OP, wr, wc, lrow, lh, smask, re, im, p2);
    store_stage<NCH, 1>(ls, b1, OP);                   // stage st+1
    __syncthreads();
    if constexpr (PF == 2) {
        if (st + 3 < s_end) load_stage<NCH, 1>(ls, st + 3);
    } else {
        if (st + 2 < s_end) load_stage<NCH, 0>(ls, st + 2);
    }
    if constexpr (KIND != 0) stage_mma<MODE, TILE, KIND == 1>(b1, b1 + OP, wr, wc, lrow, lh, smask, re, im, p2);
    if (st + 2 < s_end) store_stage<NCH, 0>(ls, b0, OP);
}

// Stages [s_begin, s_end) of one split (an even count).  KIND 0: the wave
// only stages operands; 1: every subtile of the wave is computed; 2: the
// subtiles of smask.  Keeping the MFMAs free of branches in the common case
// (KIND 1) lets the accumulators live in AGPRs across stages.  With the fold
// on, each FOLD-stage group accumulates from zero and is added to the FP32
// totals (tre, tim); without it the whole split is one group written to
// (tre, tim) at the end.
template <int MODE, int TILE, int KIND, int PF>
__device__ __forceinline__ void split_loop(LoopState<Cfg<MODE, TILE>::NCH>& ls, char* lds, int s_begin, int s_end,
                                           int wr, int wc, int lrow, int lh, unsigned smask,
                                           typename Cfg<MODE, TILE>::ACC (&tre)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS],
                                           typename Cfg<MODE, TILE>::ACC (&tim)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS]) {
    using C = Cfg<MODE, TILE>;
    constexpr int NS = C::NS, NACC = C::NACC;
    using ACC = typename C::ACC;
    auto zero = [](ACC (&x)[NS][NS]) {
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int j = 0; j < NS; ++j)
#pragma unroll
                for (int e = 0; e < NACC; ++e) x[i][j][e] = 0;
    };
    if constexpr (C::FOLD_ON) {
        zero(tre);
        zero(tim);
        for (int st = s_begin; st < s_end;) {
            const int g_end = min(st + C::FOLD, s_end);
            ACC re[NS][NS], im[NS][NS], p2[NS][NS];
            zero(re);
            zero(im);
            zero(p2);
            for (; st < g_end; st += 2)
                two_stages<MODE, TILE, KIND, PF>(ls, lds, st, s_end, wr, wc, lrow, lh, smask, re, im, p2);
            if constexpr (KIND != 0) {
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int j = 0; j < NS; ++j) {
                        tre[i][j] += re[i][j];
                        if constexpr (C::PAIR_IM) tim[i][j] += im[i][j] - p2[i][j];   // P1 - P2
                        else tim[i][j] += im[i][j];
                    }
            }
        }
    } else {
        ACC p2[NS][NS];
        zero(tre);
        zero(tim);
        zero(p2);
        for (int st = s_begin; st < s_end; st += 2)
            two_stages<MODE, TILE, KIND, PF>(ls, lds, st, s_end, wr, wc, lrow, lh, smask, tre, tim, p2);
        if constexpr (C::PAIR_IM && KIND != 0) {
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j < NS; ++j) tim[i][j] -= p2[i][j];                 // P1 - P2
        }
    }
}

template <int MODE, int TILE, int PF>
__global__ __launch_bounds__(256) void herk_kernel(HerkArgs a) {
    using C = Cfg<MODE, TILE>;
    constexpr int NS = C::NS, BLK = C::BLK, WT = C::WT, NCH = C::NCH, NACC = C::NACC;
    using ACC = typename C::ACC;
    using AR = typename C::AR;

    __shared__ __attribute__((aligned(16))) char lds[2][2][TILE * LDS_STRIDE];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;

    int unit, split, tile, bi, bj;
    decode_work(a, blockIdx.x, unit, split, tile);
    decode_tile(tile, a.nb, bi, bj);
    const bool diag = (bi == bj);

    const int kst = (int)(eff_kpad(a.kpad, a.dyn_kpad) / C::SRC_STAGE);
    const int s_begin = split * (a.ksplit / C::SRC_STAGE);
    const int s_end = min(s_begin + a.ksplit / C::SRC_STAGE, kst);
    if (s_begin >= s_end) return;   // split beyond the job's sources (uniform per workgroup)

    // A subtile that lies strictly below the diagonal of a diagonal tile, or
    // wholly in the antenna padding, is skipped (no MFMAs, no partial write;
    // the reduce stage reads the mirrored entry instead).  A wave with no
    // subtile left still helps stage operands.
    unsigned smask = 0;
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int r0 = wr * WT + i * BLK, c0 = wc * WT + j * BLK;   // tile coordinates
            const bool on = (bi * TILE + r0 < a.nant) && (bj * TILE + c0 < a.nant) && !(diag && r0 > c0);
            smask |= (unsigned)on << (i * NS + j);
        }
    smask = __builtin_amdgcn_readfirstlane(smask);

    // ---- loader: a stage of TILE antennas is one contiguous TILE*128-byte
    // block (stage-major operands); thread chunk c covers bytes
    // [16 idx, 16 idx + 16) of it, i.e. row idx/8, 16-byte column idx%8.
    // Loads are unconditional (rows past the last antenna re-read the last
    // antenna's row) and padding rows are zeroed when the registers go to
    // LDS, after the stage's MFMAs, so the loads stay in flight meanwhile.
    const char* Wu = reinterpret_cast<const char*>(a.W) + (size_t)unit * a.unit_stride;
    const size_t stage_bytes = (size_t)a.nant * STAGE_BYTES;
    const char* gA = Wu + (size_t)bi * TILE * STAGE_BYTES + (size_t)s_begin * stage_bytes;
    const char* gB = Wu + (size_t)bj * TILE * STAGE_BYTES + (size_t)s_begin * stage_bytes;
    // (staging uses a native vector type and a bitwise mask: a ternary on a
    // struct type would be lowered through private memory)
    int roffA[NCH], roffB[NCH], loff[NCH];
    uint32_t mA[NCH], mB[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int idx = tid + 256 * c;
        const int r = idx >> 3, c16 = idx & 7;
        const int ra = bi * TILE + r, rb = bj * TILE + r;
        mA[c] = ra < a.nant ? 0xffffffffu : 0u;
        mB[c] = rb < a.nant ? 0xffffffffu : 0u;
        roffA[c] = (min(ra, a.nant - 1) - bi * TILE) * STAGE_BYTES + c16 * 16;
        roffB[c] = (min(rb, a.nant - 1) - bj * TILE) * STAGE_BYTES + c16 * 16;
        loff[c] = r * LDS_STRIDE + c16 * 16;
    }
    LoopState<NCH> ls;
    ls.gA = gA;
    ls.gB = gB;
    ls.stage_bytes = stage_bytes;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        ls.mA[c] = mA[c]; ls.mB[c] = mB[c];
        ls.roffA[c] = roffA[c]; ls.roffB[c] = roffB[c]; ls.loff[c] = loff[c];
    }
    // prologue: stage s_begin to LDS buffer 0, stage s_begin+1 in flight
    char* lbase = &lds[0][0][0];
    load_stage<NCH, 0>(ls, 0);
    store_stage<NCH, 0>(ls, lbase, TILE * LDS_STRIDE);
    if constexpr (PF == 2) load_stage<NCH, 1>(ls, 1);   // splits hold a multiple of 4 stages

    // lane -> (row within block, k slot group)
    const int lrow = C::F64 ? (lane & 15) : (lane & 31);
    const int lh = C::F64 ? (lane >> 4) : (lane >> 5);

    ACC tre[NS][NS], tim[NS][NS];
    const unsigned full = (1u << (NS * NS)) - 1;
    // stage indices below are relative to s_begin (ls.gA/gB already point there)
    const int n_st = s_end - s_begin;
    if (smask == full)
        split_loop<MODE, TILE, 1, PF>(ls, lbase, 0, n_st, wr, wc, lrow, lh, smask, tre, tim);
    else if (smask)
        split_loop<MODE, TILE, 2, PF>(ls, lbase, 0, n_st, wr, wc, lrow, lh, smask, tre, tim);
    else
        split_loop<MODE, TILE, 0, PF>(ls, lbase, 0, n_st, wr, wc, lrow, lh, smask, tre, tim);
    if (!smask) return;

    // ---- write the (Re, Im) partial tile
    constexpr size_t TSZ = (size_t)TILE * TILE;
    AR* out = reinterpret_cast<AR*>(a.partials) +
              (((size_t)unit * a.nsplit + split) * a.ntiles + tile) * 2 * TSZ;
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (!(smask & (1u << (i * NS + j)))) continue;
#pragma unroll
            for (int e = 0; e < NACC; ++e) {
                int r, c;
                if constexpr (C::F64) {        // 16x16 f64 block: col = lane%16, row = lane/16 + 4e
                    r = (lane >> 4) + 4 * e;
                    c = lane & 15;
                } else {                       // 32x32 f32 block
                    r = 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                    c = lane & 31;
                }
                const int row = wr * WT + i * BLK + r;
                const int col = wc * WT + j * BLK + c;
                out[(size_t)row * TILE + col] = tre[i][j][e];
                out[TSZ + (size_t)row * TILE + col] = tim[i][j][e];
            }
        }
}

// ---------------------------------------------------------------------------
// Ring variant: operands go from global memory straight into LDS
// (global_load_lds_dwordx4, no staging registers) through a ring of RING
// slots, issued RING-1 stages ahead.  Every stage issues the same number of
// loads (past the last stage the last one is re-read into a slot nobody reads
// again), so one explicit `s_waitcnt vmcnt` count retires exactly the stage
// about to be used; a raw s_barrier (no fence, so the in-flight loads stay in
// flight) then publishes it to the other waves.  Each slot is its own LDS
// array so the compiler can tell the slots apart and adds no waits of its
// own.  Rows are 128 bytes with XOR-swizzled 16-byte chunks.  Rows past the
// last antenna load the last antenna's row: their MFMA outputs land only in
// rows/columns that are never read.
constexpr int RING = 4;

template <int TILE>
struct RingLds {
    static constexpr int OP = TILE * STAGE_BYTES;      // bytes of one operand tile
};

// vmcnt(n) with no wait on the other counters (gfx9 encoding: vmcnt[3:0] in
// bits 3:0, vmcnt[5:4] in bits 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt out of range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

#define HV_LDS(p) ((__attribute__((address_space(3))) void*)(p))

// s_barrier without a memory fence (the waits before it are explicit, so
// loads issued for later stages stay in flight), fenced for the compiler
// only: no memory access moves across it, as LDS written by other waves
// before the barrier is read after it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Issue the direct-to-LDS loads of one stage (A and B tiles) into `slot`.
template <int TILE>
__device__ __forceinline__ void ring_issue(char* slot, const char* gA, const char* gB, const int* offA, const int* offB,
                                           int wave) {
    constexpr int NCH = TILE * (STAGE_BYTES / 16) / 256;
    constexpr int OP = RingLds<TILE>::OP;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        char* dst = slot + (c * 4 + wave) * 1024;     // 64 lanes x 16 B, lane-contiguous
        __builtin_amdgcn_global_load_lds(gA + offA[c], HV_LDS(dst), 16, 0, 0);
        __builtin_amdgcn_global_load_lds(gB + offB[c], HV_LDS(dst + OP), 16, 0, 0);
    }
}

// s_waitcnt with vmcnt(VM) and lgkmcnt(LGKM), expcnt left alone.
template <int VM, int LGKM>
__device__ __forceinline__ void wait_vm_lgkm() {
    static_assert(VM >= 0 && VM < 64 && LGKM >= 0 && LGKM < 16, "count out of range");
    __builtin_amdgcn_s_waitcnt((VM & 0xF) | ((VM >> 4) << 14) | (0x7 << 4) | (LGKM << 8));
}

// Split-mode operand fragments of one 8-source MFMA step for one wave: hi
// and lo words of the wave's A rows and B rows (swizzled ring layout).
template <int NS>
struct SplitFrag {
    u32x4 ahi[NS], alo[NS], bhi[NS], blo[NS];
};

// Read the fragments of source group g (chunks 2g = hi, 2g+1 = lo).
template <int NS>
__device__ __forceinline__ void read_frag(SplitFrag<NS>& f, const char* As, const char* Bs, int ra0, int rb0, int g) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        f.ahi[i] = *reinterpret_cast<const u32x4*>(As + lds_off<true>(ra0 + 32 * i, 2 * g));
        f.alo[i] = *reinterpret_cast<const u32x4*>(As + lds_off<true>(ra0 + 32 * i, 2 * g + 1));
        f.bhi[i] = *reinterpret_cast<const u32x4*>(Bs + lds_off<true>(rb0 + 32 * i, 2 * g));
        f.blo[i] = *reinterpret_cast<const u32x4*>(Bs + lds_off<true>(rb0 + 32 * i, 2 * g + 1));
    }
}

// Re += a.b and Im += a.b' with b' = (-y, x), as hi.hi + hi.lo + lo.hi.
template <int MODE, int NS, bool ALLON>
__device__ __forceinline__ void frag_mma(const SplitFrag<NS>& f, unsigned smask, f32x16 (&re)[NS][NS],
                                         f32x16 (&im)[NS][NS]) {
    u32x4 bhr[NS], blr[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        bhr[j] = rot_neg16(f.bhi[j]);
        blr[j] = rot_neg16(f.blo[j]);
    }
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j)
            if (ALLON || (smask & (1u << (i * NS + j)))) {
                re[i][j] = mfma16<MODE>(f.ahi[i], f.bhi[j], re[i][j]);
                im[i][j] = mfma16<MODE>(f.ahi[i], bhr[j], im[i][j]);
                re[i][j] = mfma16<MODE>(f.ahi[i], f.blo[j], re[i][j]);
                im[i][j] = mfma16<MODE>(f.ahi[i], blr[j], im[i][j]);
                re[i][j] = mfma16<MODE>(f.alo[i], f.bhi[j], re[i][j]);
                im[i][j] = mfma16<MODE>(f.alo[i], bhr[j], im[i][j]);
            }
}

// Scheduling pattern for one MFMA group of NMF MFMAs issued together with
// NV vector-memory loads and ND LDS reads: one MFMA, then one load and one
// read, and so on, so the wave never leaves the matrix pipe idle while it
// issues memory instructions (sched_group_barrier masks: 0x008 MFMA, 0x020
// VMEM read, 0x100 DS read).
template <int NMF, int NV, int ND>
__device__ __forceinline__ void interleave_mfma() {
#pragma unroll
    for (int i = 0; i < NMF; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if (i < NV) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
        if (i < ND) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
}

// One split of the ring loop (stages [0, n_st) of the split at gA/gB).
// KIND 0: the wave only issues its share of the loads; 1: it computes every
// block of its wave tile (blocks outside smask compute values nobody
// writes); 2: only the blocks of smask.  The kind is fixed per wave before
// the loop, so the MFMA chains have no branches and the accumulators stay in
// AGPRs.  On exit every load of the split has landed and every wave has
// passed its last LDS read, so the slots can take the next split.
template <int MODE, int TILE, int KIND>
__device__ __forceinline__ void ring_split(char* s0, char* s1, char* s2, char* s3, const char* gA, const char* gB,
                                           size_t stage_bytes, int n_st, const int* offA, const int* offB, int wave,
                                           int wr, int wc, int lrow, int lh, unsigned smask,
                                           typename Cfg<MODE, TILE>::ACC (&tre)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS],
                                           typename Cfg<MODE, TILE>::ACC (&tim)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS]) {
    using C = Cfg<MODE, TILE>;
    constexpr int NS = C::NS, NACC = C::NACC, NCH = C::NCH;
    constexpr int OP = RingLds<TILE>::OP;
    constexpr int PER_STAGE = 2 * NCH;                 // loads per thread per stage
    using ACC = typename C::ACC;
    char* slots[RING] = {s0, s1, s2, s3};
    // stages past the last one re-read the last stage (keeps the count per
    // stage constant for the vmcnt arithmetic)
#define HV_ISSUE(slot, st)                                                                    \
    ring_issue<TILE>(slot, gA + (size_t)min(st, n_st - 1) * stage_bytes,                      \
                     gB + (size_t)min(st, n_st - 1) * stage_bytes, offA, offB, wave)
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j)
#pragma unroll
            for (int e = 0; e < NACC; ++e) { tre[i][j][e] = 0; tim[i][j][e] = 0; }

    if constexpr (MODE == MODE_SPLIT_BF16 || MODE == MODE_SPLIT_FP16) {
        // Software-pipelined split modes: a stage is two 8-source steps; the
        // LDS reads of the next step are issued before the MFMAs of the
        // current one.  The stage barrier sits between the two steps: by then
        // every wave has its reads of stage k back (lgkmcnt 0), so slot k
        // takes stage k+RING, and stage k+1 has landed for every wave.  The
        // ring starts RING stages deep; in steady state stages k+2..k+RING
        // are in flight at the wait.
        const int ra0 = wr * C::WT + lrow, rb0 = wc * C::WT + lrow;
        HV_ISSUE(s0, 0);
        HV_ISSUE(s1, 1);
        HV_ISSUE(s2, 2);
        HV_ISSUE(s3, 3);
        wait_vmcnt<(RING - 1) * PER_STAGE>();
        lds_barrier();
        // 64-tiles: groups of FOLD_SRC sources accumulate from zero and are
        // added to the FP32 totals; 128-tiles (one wave per SIMD, no
        // registers for a second set of accumulators) run one chain per split
        // straight into the totals.
        constexpr int RFOLD = (TILE == 64) ? FOLD_SRC / C::SRC_STAGE : (1 << 30);
        constexpr int NMF = 6 * NS * NS;               // MFMAs per 8-source step
        constexpr int NRD = 4 * NS;                    // LDS reads per 8-source step
        SplitFrag<NS> fa, fb;
        read_frag<NS>(fa, s0, s0 + OP, ra0, rb0, lh);
        for (int st = 0; st < n_st;) {
            const int g_end = min(st + RFOLD, n_st);
            ACC gre[NS][NS], gim[NS][NS];
            if constexpr (TILE == 64) {
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int j = 0; j < NS; ++j)
#pragma unroll
                        for (int e = 0; e < NACC; ++e) { gre[i][j][e] = 0; gim[i][j][e] = 0; }
            }
            ACC (&re)[NS][NS] = (TILE == 64) ? gre : tre;
            ACC (&im)[NS][NS] = (TILE == 64) ? gim : tim;
            for (; st < g_end; st += RING) {
#pragma unroll
                for (int k = 0; k < RING; ++k) {
                    char* cur = slots[k];
                    char* nxt = slots[(k + 1) % RING];
                    // fa's reads went out a whole MFMA group ago: retire them
                    // before the next reads, so no wait below has to let
                    // younger reads stay in flight
                    wait_vm_lgkm<63, 0>();
                    __builtin_amdgcn_sched_barrier(0);
                    read_frag<NS>(fb, cur, cur + OP, ra0, rb0, 2 + lh);
                    if constexpr (KIND != 0) {
                        frag_mma<MODE, NS, KIND == 1>(fa, smask, re, im);
                        interleave_mfma<NMF, 0, NRD>();
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    wait_vm_lgkm<(RING - 2) * PER_STAGE, 0>();
                    lds_barrier();
                    __builtin_amdgcn_sched_barrier(0);
                    HV_ISSUE(cur, st + k + RING);
                    read_frag<NS>(fa, nxt, nxt + OP, ra0, rb0, lh);   // past the end: a re-read, unused
                    if constexpr (KIND != 0) {
                        frag_mma<MODE, NS, KIND == 1>(fb, smask, re, im);
                        interleave_mfma<NMF, PER_STAGE, NRD>();
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            if constexpr (KIND != 0 && TILE == 64) {
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int j = 0; j < NS; ++j) {
                        tre[i][j] += gre[i][j];
                        tim[i][j] += gim[i][j];
                    }
            }
        }
        wait_vm_lgkm<0, 0>();
        lds_barrier();
        return;
    }

    HV_ISSUE(s0, 0);
    HV_ISSUE(s1, 1);
    HV_ISSUE(s2, 2);

    if constexpr (C::FOLD_ON) {
        // groups of FOLD stages accumulate from zero and are added to the
        // FP32 totals
        for (int st = 0; st < n_st;) {
            const int g_end = min(st + C::FOLD, n_st);
            ACC re[NS][NS], im[NS][NS], p2[NS][NS];
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j < NS; ++j)
#pragma unroll
                    for (int e = 0; e < NACC; ++e) { re[i][j][e] = 0; im[i][j][e] = 0; p2[i][j][e] = 0; }
            for (; st < g_end; st += RING) {
#pragma unroll
                for (int k = 0; k < RING; ++k) {
                    wait_vmcnt<(RING - 2) * PER_STAGE>();  // stage st+k has landed (this wave)
                    lds_barrier();          // ... for every wave; slot k-1 is free
                    HV_ISSUE(slots[(k + RING - 1) % RING], st + k + RING - 1);
                    if constexpr (KIND != 0)
                        stage_mma<MODE, TILE, KIND == 1, true>(slots[k], slots[k] + OP, wr, wc, lrow, lh, smask, re,
                                                               im, p2);
                }
            }
            if constexpr (KIND != 0) {
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int j = 0; j < NS; ++j) {
                        tre[i][j] += re[i][j];
                        if constexpr (C::PAIR_IM) tim[i][j] += im[i][j] - p2[i][j];   // P1 - P2
                        else tim[i][j] += im[i][j];
                    }
            }
        }
    } else {
        // one MFMA chain per split, straight into the totals
        ACC p2[NS][NS];
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int j = 0; j < NS; ++j)
#pragma unroll
                for (int e = 0; e < NACC; ++e) p2[i][j][e] = 0;
        for (int st = 0; st < n_st; st += RING) {
#pragma unroll
            for (int k = 0; k < RING; ++k) {
                wait_vmcnt<(RING - 2) * PER_STAGE>();
                lds_barrier();
                HV_ISSUE(slots[(k + RING - 1) % RING], st + k + RING - 1);
                if constexpr (KIND != 0)
                    stage_mma<MODE, TILE, KIND == 1, true>(slots[k], slots[k] + OP, wr, wc, lrow, lh, smask, tre, tim,
                                                           p2);
            }
        }
        if constexpr (C::PAIR_IM && KIND != 0) {
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j < NS; ++j) tim[i][j] -= p2[i][j];
        }
    }
#undef HV_ISSUE
    wait_vmcnt<0>();                 // the re-reads past the end have landed too
    lds_barrier();    // every wave is done with the slots
}

// Work item L (a (unit, split, tile) triple, numbered as decode_work reads
// it) per loop trip; the grid is either one workgroup per item or a
// persistent grid of a multiple of 8 workgroups, which keeps every
// workgroup's items on its own XCD and the tiles of one split running
// side by side (their shared operand stages hit in L2).
template <int MODE, int TILE>
__global__ __launch_bounds__(256) void herk_ring_kernel(HerkArgs a) {
    using C = Cfg<MODE, TILE>;
    constexpr int NS = C::NS, BLK = C::BLK, WT = C::WT, NCH = C::NCH, NACC = C::NACC;
    constexpr int OP = RingLds<TILE>::OP;
    using ACC = typename C::ACC;
    using AR = typename C::AR;

    __shared__ __attribute__((aligned(1024))) char slot0[2 * OP];
    __shared__ __attribute__((aligned(1024))) char slot1[2 * OP];
    __shared__ __attribute__((aligned(1024))) char slot2[2 * OP];
    __shared__ __attribute__((aligned(1024))) char slot3[2 * OP];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int lrow = C::F64 ? (lane & 15) : (lane & 31);
    const int lh = C::F64 ? (lane >> 4) : (lane >> 5);
    const int kst = (int)(eff_kpad(a.kpad, a.dyn_kpad) / C::SRC_STAGE);
    const int64_t nitems = (int64_t)a.nunits * a.nsplit * a.ntiles;
    const size_t stage_bytes = (size_t)a.nant * STAGE_BYTES;
    const unsigned full = (1u << (NS * NS)) - 1;

    for (int64_t L = blockIdx.x; L < nitems; L += gridDim.x) {
        int unit, split, tile, bi, bj;
        decode_work(a, L, unit, split, tile);
        decode_tile(tile, a.nb, bi, bj);
        const bool diag = (bi == bj);
        const int s_begin = split * (a.ksplit / C::SRC_STAGE);
        const int s_end = min(s_begin + a.ksplit / C::SRC_STAGE, kst);
        if (s_begin >= s_end) continue;   // split beyond the job's sources (uniform per workgroup)
        const int n_st = s_end - s_begin; // a multiple of RING

        unsigned smask = 0;
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                const int r0 = wr * WT + i * BLK, c0 = wc * WT + j * BLK;
                const bool on = (bi * TILE + r0 < a.nant) && (bj * TILE + c0 < a.nant) && !(diag && r0 > c0);
                smask |= (unsigned)on << (i * NS + j);
            }
        smask = __builtin_amdgcn_readfirstlane(smask);

        // global offsets of this lane's chunks: chunk id (c*4 + wave)*64 + lane
        // lands at LDS byte 16*id = row*128 + 16*(lane & 7), so it must hold
        // the row's logical chunk (lane & 7) ^ (row & 7)
        const char* Wu = reinterpret_cast<const char*>(a.W) + (size_t)unit * a.unit_stride;
        const char* gA = Wu + (size_t)bi * TILE * STAGE_BYTES + (size_t)s_begin * stage_bytes;
        const char* gB = Wu + (size_t)bj * TILE * STAGE_BYTES + (size_t)s_begin * stage_bytes;
        int offA[NCH], offB[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int r = (c * 4 + wave) * 8 + (lane >> 3);
            const int q = (lane & 7) ^ (r & 7);
            offA[c] = (min(bi * TILE + r, a.nant - 1) - bi * TILE) * STAGE_BYTES + q * 16;
            offB[c] = (min(bj * TILE + r, a.nant - 1) - bj * TILE) * STAGE_BYTES + q * 16;
        }

        ACC tre[NS][NS], tim[NS][NS];
        // 128-tiles run one wave per SIMD: a wave with a partial mask
        // computes its whole wave tile, which costs the workgroup no time
        // (another wave of the tile has a full mask or as much work)
        if (smask == full || (TILE == 128 && smask))
            ring_split<MODE, TILE, 1>(slot0, slot1, slot2, slot3, gA, gB, stage_bytes, n_st, offA, offB, wave, wr, wc,
                                      lrow, lh, smask, tre, tim);
        else if (smask)
            ring_split<MODE, TILE, 2>(slot0, slot1, slot2, slot3, gA, gB, stage_bytes, n_st, offA, offB, wave, wr, wc,
                                      lrow, lh, smask, tre, tim);
        else
            ring_split<MODE, TILE, 0>(slot0, slot1, slot2, slot3, gA, gB, stage_bytes, n_st, offA, offB, wave, wr, wc,
                                      lrow, lh, smask, tre, tim);
        if (!smask) continue;

        constexpr size_t TSZ = (size_t)TILE * TILE;
        AR* out = reinterpret_cast<AR*>(a.partials) +
                  (((size_t)unit * a.nsplit + split) * a.ntiles + tile) * 2 * TSZ;
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                if (!(smask & (1u << (i * NS + j)))) continue;
#pragma unroll
                for (int e = 0; e < NACC; ++e) {
                    int r, c;
                    if constexpr (C::F64) {
                        r = (lane >> 4) + 4 * e;
                        c = lane & 15;
                    } else {
                        r = 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                        c = lane & 31;
                    }
                    const int row = wr * WT + i * BLK + r;
                    const int col = wc * WT + j * BLK + c;
                    out[(size_t)row * TILE + col] = tre[i][j][e];
                    out[TSZ + (size_t)row * TILE + col] = tim[i][j][e];
                }
            }
    }
}
#undef HV_LDS

// Resident workgroups of a ring kernel on the whole device (a multiple of 8),
// or 0 if the runtime cannot say.
// Computed once per process (thread-safe static initialisation), for the
// device current at the first call; the library drives one device type.
template <int MODE, int TILE>
static int64_t ring_resident() {
    static const int64_t cached = [] {
        int dev = 0, ncu = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, herk_ring_kernel<MODE, TILE>, 256, 0) != hipSuccess)
            return (int64_t)0;
        return (int64_t)ncu * per_cu / 8 * 8;
    }();
    return cached;
}

// pf 3: one workgroup per work item; pf 4: persistent grid.
template <int MODE, int TILE>
static hipError_t launch_herk_ring_t(const HerkArgs& a, bool persistent, hipStream_t s) {
    const int64_t nblk = (int64_t)a.nunits * a.nsplit * a.ntiles;
    if (nblk <= 0) return hipSuccess;
    int64_t grid = nblk;
    if (persistent) {
        const int64_t res = ring_resident<MODE, TILE>();
        if (res > 0 && res < nblk) grid = res;
    }
    hipLaunchKernelGGL((herk_ring_kernel<MODE, TILE>), dim3((unsigned)grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Split-mode contraction that builds its operands in place.  Instead of
// reading phasor stages written by phasor_kernel, each workgroup generates
// the hi/lo operand rows of its two antenna blocks for every stage from the
// beam amplitude A sqrt(I) (amp_kernel, one per time) and the FP64 source
// directions, with the same arithmetic as phasor_kernel (bit for bit the same
// operands), and contracts them as herk_ring_kernel does (the same MFMA
// sequence, so the same partial sums).  The 8 B per (antenna, source, unit)
// operand never goes to HBM; the VALU work of the build runs beside the
// matrix pipe of the same waves.
//
// LDS: an input ring of FRING slots (amplitude rows of both antenna blocks,
// 64 B per row, and the stage's x, y, z) filled by global_load_lds, and two
// operand buffers (A then B rows, 128-byte swizzled rows as in the ring
// kernel).  Per stage k, after one barrier: issue the inputs of stage
// k + FRING, build operand stage k+1 (wave w builds source group w of every
// row, lane l rows l, l+64, ...) and run the MFMAs of stage k.
#define HV_LDS(p) ((__attribute__((address_space(3))) void*)(p))
constexpr int FRING = 4;

// Amplitude rows in LDS: the 16-byte part p of row r (4 sources) sits at
// chunk 4 r + (p ^ amp_swz(r)), so the 8 lanes of a read phase, which read
// part g of 8 consecutive rows, hit 8 distinct bank groups.
__device__ __forceinline__ int amp_swz(int r) { return (r >> 1) & 3; }

template <int TILE>
struct FusedLds {
    static constexpr int ROWS = 2 * TILE;                // A rows, then B rows
    static constexpr int AMP = ROWS * 64;                // amplitude bytes per stage
    static constexpr int CRD = 4 * 1024;                 // one KB per wave; x, y, z in waves 0-2
    static constexpr int IN = AMP + CRD;                 // one input slot
    static constexpr int OP = TILE * STAGE_BYTES;        // one operand tile
    static constexpr int NAMP = ROWS * 4 / 256;          // amplitude loads per thread per stage
    static constexpr int PER_STAGE = NAMP + 1;           // loads per thread per stage
};

// Issue the input loads of one stage into an input slot.  amp_st / crd_st
// point at the stage in global memory; ampoff[i] (floats) and crdoff
// (doubles) are this lane's offsets from there.
template <int TILE>
__device__ __forceinline__ void fused_issue(char* slot, const float* amp_st, const double* crd_st, const int* ampoff,
                                            int crdoff, int wave) {
    using L = FusedLds<TILE>;
#pragma unroll
    for (int i = 0; i < L::NAMP; ++i)
        __builtin_amdgcn_global_load_lds(amp_st + ampoff[i], HV_LDS(slot + (i * 4 + wave) * 1024), 16, 0, 0);
    __builtin_amdgcn_global_load_lds(crd_st + crdoff, HV_LDS(slot + L::AMP + wave * 1024), 16, 0, 0);
}

// Build operand stage rows for source group g (= wave) from an input slot:
// lane rows rr = lane + 64 i, i < RPT (rows < TILE go to the A tile, the
// rest to the B tile).
template <int MODE, int TILE, int RPT>
__device__ __forceinline__ void fused_build(const char* in, char* op, int lane, int g, const double (&ax)[RPT],
                                            const double (&ay)[RPT], const double (&az)[RPT]) {
    using L = FusedLds<TILE>;
    const f64x2* cx = reinterpret_cast<const f64x2*>(in + L::AMP + g * 32);
    const f64x2* cy = reinterpret_cast<const f64x2*>(in + L::AMP + 1024 + g * 32);
    const f64x2* cz = reinterpret_cast<const f64x2*>(in + L::AMP + 2048 + g * 32);
    const f64x2 x01 = cx[0], x23 = cx[1], y01 = cy[0], y23 = cy[1], z01 = cz[0], z23 = cz[1];
    const double xs[4] = {x01[0], x01[1], x23[0], x23[1]};
    const double ys[4] = {y01[0], y01[1], y23[0], y23[1]};
    const double zs[4] = {z01[0], z01[1], z23[0], z23[1]};
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int rr = lane + 64 * i;
        const f32x4 am = *reinterpret_cast<const f32x4*>(in + rr * 64 + ((g ^ amp_swz(rr)) << 4));
        u32x4 hi, lo;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float sn, cs;
            sincos_turns_hw(phase_turns(ax[i], ay[i], az[i], xs[j], ys[j], zs[j]), &sn, &cs);
            uint32_t h, l;
            split_phasor<MODE>(am[j], cs, sn, h, l);
            hi[j] = h;
            lo[j] = l;
        }
        char* base = (rr < TILE) ? op : op + L::OP;
        const int r = (rr < TILE) ? rr : rr - TILE;
        *reinterpret_cast<u32x4*>(base + lds_off<true>(r, 2 * g)) = hi;
        *reinterpret_cast<u32x4*>(base + lds_off<true>(r, 2 * g + 1)) = lo;
    }
}

// One split: stages [0, n_st).  RPT rows per lane (TILE/64 on diagonal
// tiles, whose B tile is the A tile; TILE/32 otherwise).  KIND 0: the wave
// builds but runs no MFMAs; 1: it runs the MFMAs of its whole wave tile.
template <int MODE, int TILE, int RPT, int KIND>
__device__ __forceinline__ void fused_split(char* in0, char* in1, char* in2, char* in3, char* op0, char* op1,
                                            const float* amp_t, const double* crd_t, size_t amp_stage, int s_begin,
                                            int n_st, const int* ampoff, int crdoff, int lane, int wave, int wr,
                                            int wc, const double (&ax)[RPT], const double (&ay)[RPT],
                                            const double (&az)[RPT], bool diag,
                                            f32x16 (&tre)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS],
                                            f32x16 (&tim)[Cfg<MODE, TILE>::NS][Cfg<MODE, TILE>::NS]) {
    using C = Cfg<MODE, TILE>;
    using L = FusedLds<TILE>;
    constexpr int NS = C::NS, NACC = C::NACC;
    constexpr int NMF = 6 * NS * NS;
    constexpr int PER = L::PER_STAGE;
    constexpr int VPM = (RPT * 4 * 36) / (2 * NMF);     // build VALU ops per MFMA (~36 per element)
    constexpr int RFOLD = (TILE == 64) ? FOLD_SRC / C::SRC_STAGE : (1 << 30);
    char* ins[FRING] = {in0, in1, in2, in3};
    char* ops[2] = {op0, op1};
    const int lrow = lane & 31, lh = lane >> 5;
    const int ra0 = wr * C::WT + lrow, rb0 = wc * C::WT + lrow;
    const int boff = diag ? 0 : L::OP;                  // B tile of a diagonal tile is its A tile
#define HV_FISSUE(slot, st)                                                                         \
    fused_issue<TILE>(slot, amp_t + (size_t)(s_begin + min(st, n_st - 1)) * amp_stage,               \
                      crd_t + (size_t)(s_begin + min(st, n_st - 1)) * 16, ampoff, crdoff, wave)

#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j)
#pragma unroll
            for (int e = 0; e < NACC; ++e) { tre[i][j][e] = 0; tim[i][j][e] = 0; }

    HV_FISSUE(in0, 0);
    HV_FISSUE(in1, 1);
    HV_FISSUE(in2, 2);
    HV_FISSUE(in3, 3);
    wait_vmcnt<(FRING - 1) * PER>();
    lds_barrier();
    fused_build<MODE, TILE, RPT>(in0, op0, lane, wave, ax, ay, az);

    for (int st = 0; st < n_st;) {
        const int g_end = min(st + RFOLD, n_st);
        f32x16 gre[NS][NS], gim[NS][NS];
        if constexpr (TILE == 64) {
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j < NS; ++j)
#pragma unroll
                    for (int e = 0; e < NACC; ++e) { gre[i][j][e] = 0; gim[i][j][e] = 0; }
        }
        f32x16 (&re)[NS][NS] = (TILE == 64) ? gre : tre;
        f32x16 (&im)[NS][NS] = (TILE == 64) ? gim : tim;
        for (; st < g_end; st += FRING) {
#pragma unroll
            for (int k = 0; k < FRING; ++k) {
                char* cur = ops[k & 1];
                char* nxt = ops[(k + 1) & 1];
                // inputs of stage st+k+1 have landed (this wave); every
                // wave's LDS writes and reads of the previous stage are done
                wait_vm_lgkm<(FRING - 2) * PER, 0>();
                lds_barrier();
                HV_FISSUE(ins[k], st + k + FRING);           // slot of stage st+k, consumed
                SplitFrag<NS> fa, fb;
                read_frag<NS>(fa, cur, cur + boff, ra0, rb0, lh);
                read_frag<NS>(fb, cur, cur + boff, ra0, rb0, 2 + lh);
                // operand stage st+k+1 (past the end: rebuilt from a re-read, unused)
                fused_build<MODE, TILE, RPT>(ins[(k + 1) % FRING], nxt, lane, wave, ax, ay, az);
                if constexpr (KIND != 0) {
                    frag_mma<MODE, NS, true>(fa, 0u, re, im);
                    frag_mma<MODE, NS, true>(fb, 0u, re, im);
                    // spread the build's VALU work between the MFMAs so
                    // the matrix pipe and the VALU run side by side
#pragma unroll
                    for (int m = 0; m < 2 * NMF; ++m) {
                        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                        __builtin_amdgcn_sched_group_barrier(0x002, VPM, 0);
                    }
                }
            }
        }
        if constexpr (KIND != 0 && TILE == 64) {
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j < NS; ++j) {
                    tre[i][j] += gre[i][j];
                    tim[i][j] += gim[i][j];
                }
        }
    }
#undef HV_FISSUE
    wait_vm_lgkm<0, 0>();
    lds_barrier();
}

template <int MODE, int TILE>
__global__ __launch_bounds__(256) void herk_fused_kernel(FusedArgs a) {
    using C = Cfg<MODE, TILE>;
    using L = FusedLds<TILE>;
    constexpr int NS = C::NS, BLK = C::BLK, WT = C::WT, NACC = C::NACC;
    constexpr int RD = TILE / 64, RO = TILE / 32;       // rows per lane: diagonal / other tiles

    __shared__ __attribute__((aligned(1024))) char in0[L::IN];
    __shared__ __attribute__((aligned(1024))) char in1[L::IN];
    __shared__ __attribute__((aligned(1024))) char in2[L::IN];
    __shared__ __attribute__((aligned(1024))) char in3[L::IN];
    __shared__ __attribute__((aligned(1024))) char op0[2 * L::OP];
    __shared__ __attribute__((aligned(1024))) char op1[2 * L::OP];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int kst = (int)(eff_kpad(a.kpad, a.dyn_kpad) / C::SRC_STAGE);
    const int64_t nitems = (int64_t)a.nunits * a.nsplit * a.ntiles;
    const size_t amp_stage = (size_t)a.nant * 16;                 // floats per amplitude stage
    const size_t amp_time = (size_t)a.kpad * a.nant;              // floats per time
    const unsigned full = (1u << (NS * NS)) - 1;
    (void)full;

    for (int64_t Lw = blockIdx.x; Lw < nitems; Lw += gridDim.x) {
        int unit, split, tile, bi, bj;
        decode_work(a.nsplit, a.ntiles, Lw, unit, split, tile);
        decode_tile(tile, a.nb, bi, bj);
        const bool diag = (bi == bj);
        const int s_begin = split * (a.ksplit / C::SRC_STAGE);
        const int s_end = min(s_begin + a.ksplit / C::SRC_STAGE, kst);
        if (s_begin >= s_end) continue;   // split beyond the job's sources (uniform per workgroup)
        const int n_st = s_end - s_begin;
        const int tb = unit / a.nfreq, fl = unit - tb * a.nfreq;
        const double il = a.inv_lambda[fl];   // turns per metre
        const float* amp_t = a.amp + (size_t)(a.amp_per_unit ? unit : tb) * amp_time;
        const double* crd_t = a.crd_top + (size_t)tb * 3 * a.kpad;

        unsigned smask = 0;
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                const int r0 = wr * WT + i * BLK, c0 = wc * WT + j * BLK;
                const bool on = (bi * TILE + r0 < a.nant) && (bj * TILE + c0 < a.nant) && !(diag && r0 > c0);
                smask |= (unsigned)on << (i * NS + j);
            }
        smask = __builtin_amdgcn_readfirstlane(smask);

        // input loads: amplitude chunk c = (i*4 + wave)*64 + lane is row
        // c/4 (A rows, then B rows; past the last antenna: the last one),
        // 16-byte part c%4; coordinate chunk: wave w < 3 loads coordinate w,
        // lanes past 8 repeat lane & 7
        int ampoff[L::NAMP];
#pragma unroll
        for (int i = 0; i < L::NAMP; ++i) {
            const int c = (i * 4 + wave) * 64 + lane;
            const int rr = c >> 2;
            const int ant = min((rr < TILE ? bi * TILE + rr : bj * TILE + rr - TILE), a.nant - 1);
            ampoff[i] = ant * 16 + ((c & 3) ^ amp_swz(rr)) * 4;
        }
        const int crdoff = min(wave, 2) * (int)a.kpad + 2 * (lane & 7);

        f32x16 tre[NS][NS], tim[NS][NS];
        if (diag) {
            double ax[RD], ay[RD], az[RD];
#pragma unroll
            for (int i = 0; i < RD; ++i) {
                const int ant = min(bi * TILE + lane + 64 * i, a.nant - 1);
                ax[i] = a.antpos[3 * ant] * il; ay[i] = a.antpos[3 * ant + 1] * il; az[i] = a.antpos[3 * ant + 2] * il;
            }
            if (smask)
                fused_split<MODE, TILE, RD, 1>(in0, in1, in2, in3, op0, op1, amp_t, crd_t, amp_stage, s_begin, n_st,
                                               ampoff, crdoff, lane, wave, wr, wc, ax, ay, az, true, tre, tim);
            else
                fused_split<MODE, TILE, RD, 0>(in0, in1, in2, in3, op0, op1, amp_t, crd_t, amp_stage, s_begin, n_st,
                                               ampoff, crdoff, lane, wave, wr, wc, ax, ay, az, true, tre, tim);
        } else {
            double ax[RO], ay[RO], az[RO];
#pragma unroll
            for (int i = 0; i < RO; ++i) {
                const int rr = lane + 64 * i;
                const int ant = min((rr < TILE ? bi * TILE + rr : bj * TILE + rr - TILE), a.nant - 1);
                ax[i] = a.antpos[3 * ant] * il; ay[i] = a.antpos[3 * ant + 1] * il; az[i] = a.antpos[3 * ant + 2] * il;
            }
            if (smask)
                fused_split<MODE, TILE, RO, 1>(in0, in1, in2, in3, op0, op1, amp_t, crd_t, amp_stage, s_begin, n_st,
                                               ampoff, crdoff, lane, wave, wr, wc, ax, ay, az, false, tre, tim);
            else
                fused_split<MODE, TILE, RO, 0>(in0, in1, in2, in3, op0, op1, amp_t, crd_t, amp_stage, s_begin, n_st,
                                               ampoff, crdoff, lane, wave, wr, wc, ax, ay, az, false, tre, tim);
        }
        if (!smask) continue;

        constexpr size_t TSZ = (size_t)TILE * TILE;
        float* out = reinterpret_cast<float*>(a.partials) +
                     (((size_t)unit * a.nsplit + split) * a.ntiles + tile) * 2 * TSZ;
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                if (!(smask & (1u << (i * NS + j)))) continue;
#pragma unroll
                for (int e = 0; e < NACC; ++e) {
                    const int r = 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                    const int c = lane & 31;
                    const int row = wr * WT + i * BLK + r;
                    const int col = wc * WT + j * BLK + c;
                    out[(size_t)row * TILE + col] = tre[i][j][e];
                    out[TSZ + (size_t)row * TILE + col] = tim[i][j][e];
                }
            }
    }
}

// ---------------------------------------------------------------------------
// 128-tile build-in-place contraction with 8 waves (two per SIMD): each wave
// owns a 64x32 wave tile (two 32x32 MFMA blocks, 64 accumulator registers),
// so two waves share a SIMD and one's MFMAs run while the other builds
// operands or waits on LDS.  Operand build, input ring and per-block MFMA
// chains are those of herk_fused_kernel (same bits).
//
// Waves: MFMA tile rows wr = wave / 4 (64 rows), columns wc = wave % 4 (32);
// build group g = wave % 4 (4 sources), row half h = wave / 4.
struct Frag21 {
    u32x4 ahi[2], alo[2], bhi, blo;
};

// NA: A row blocks read (2, or 1 for a wave that owns one block)
template <int NA = 2>
__device__ __forceinline__ void read_frag21(Frag21& f, const char* As, const char* Bs, int ra0, int rb0, int g) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        f.ahi[i] = *reinterpret_cast<const u32x4*>(As + lds_off<true>(ra0 + 32 * i, 2 * g));
        f.alo[i] = *reinterpret_cast<const u32x4*>(As + lds_off<true>(ra0 + 32 * i, 2 * g + 1));
    }
    f.bhi = *reinterpret_cast<const u32x4*>(Bs + lds_off<true>(rb0, 2 * g));
    f.blo = *reinterpret_cast<const u32x4*>(Bs + lds_off<true>(rb0, 2 * g + 1));
}

// Same chain per 32x32 block as frag_mma: Re += hi.hi + hi.lo + lo.hi and
// Im += a.b' with the same operand order.
template <int MODE, int NA = 2>
__device__ __forceinline__ void frag21_mma(const Frag21& f, f32x16 (&re)[2], f32x16 (&im)[2]) {
    // one B fragment per step: rotating B costs half of rotating the two A fragments
    const u32x4 bhr = rot_neg16(f.bhi), blr = rot_neg16(f.blo);
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        re[i] = mfma16<MODE>(f.ahi[i], f.bhi, re[i]);
        im[i] = mfma16<MODE>(f.ahi[i], bhr, im[i]);
        re[i] = mfma16<MODE>(f.ahi[i], f.blo, re[i]);
        im[i] = mfma16<MODE>(f.ahi[i], blr, im[i]);
        re[i] = mfma16<MODE>(f.alo[i], f.bhi, re[i]);
        im[i] = mfma16<MODE>(f.alo[i], bhr, im[i]);
    }
}

struct Fused8Lds {
    static constexpr int TILE = 128;
    static constexpr int ROWS = 2 * TILE;
    static constexpr int AMP = ROWS * 64;                // 16 KB
    static constexpr int CRD = 4 * 1024;                 // waves w and w+4 fill region w % 4
    static constexpr int IN = AMP + CRD;
    static constexpr int OP = TILE * STAGE_BYTES;
    static constexpr int NAMP = ROWS * 4 / 512;          // amplitude loads per thread per stage
    static constexpr int PER_STAGE = NAMP + 1;
};

__device__ __forceinline__ void fused8_issue(char* slot, const float* amp_st, const double* crd_st,
                                             const int* ampoff, int crdoff, int wave) {
    using L = Fused8Lds;
#pragma unroll
    for (int i = 0; i < L::NAMP; ++i)
        __builtin_amdgcn_global_load_lds(amp_st + ampoff[i], HV_LDS(slot + (i * 8 + wave) * 1024), 16, 0, 0);
    // waves w and w + 4 write the same bytes of coordinate w % 4 (w % 4 = 3:
    // a copy of z nobody reads)
    __builtin_amdgcn_global_load_lds(crd_st + crdoff, HV_LDS(slot + L::AMP + (wave & 3) * 1024), 16, 0, 0);
}

// Build operand rows rbase + 64 i (i < RPT) of source group g from an
// input slot (as fused_build).
template <int MODE, int RPT>
__device__ __forceinline__ void fused8_build(const char* in, char* op, int rbase, int g, const double (&ax)[RPT],
                                             const double (&ay)[RPT], const double (&az)[RPT]) {
    using L = Fused8Lds;
    constexpr int TILE = 128;
    const double* cx = reinterpret_cast<const double*>(in + L::AMP + g * 32);
    const double* cy = reinterpret_cast<const double*>(in + L::AMP + 1024 + g * 32);
    const double* cz = reinterpret_cast<const double*>(in + L::AMP + 2048 + g * 32);
    f32x4 am[RPT];
#pragma unroll
    for (int i = 0; i < RPT; ++i)
        am[i] = *reinterpret_cast<const f32x4*>(in + (rbase + 64 * i) * 64 + ((g ^ amp_swz(rbase + 64 * i)) << 4));
    u32x4 hi[RPT], lo[RPT];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double x = cx[j], y = cy[j], z = cz[j];      // one source, every row of the lane
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            float sn, cs;
            sincos_turns_hw(phase_turns(ax[i], ay[i], az[i], x, y, z), &sn, &cs);
            uint32_t hw, lw;
            split_phasor<MODE>(am[i][j], cs, sn, hw, lw);
            hi[i][j] = hw;
            lo[i][j] = lw;
        }
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int rr = rbase + 64 * i;
        char* base = (rr < TILE) ? op : op + L::OP;
        const int r = (rr < TILE) ? rr : rr - TILE;
        *reinterpret_cast<u32x4*>(base + lds_off<true>(r, 2 * g)) = hi[i];
        *reinterpret_cast<u32x4*>(base + lds_off<true>(r, 2 * g + 1)) = lo[i];
    }
}

// The MFMAs of one stage for one wave: both 8-source steps, the second
// step's fragments read while the first step's MFMAs run.  NA: the wave's
// row blocks (1: only block row ra0).
template <int MODE, int NA = 2>
__device__ __forceinline__ void fused8_mma(const char* As, const char* Bs, int ra0, int rb0, int lh,
                                           f32x16 (&re)[2], f32x16 (&im)[2]) {
    Frag21 f0, f1;
    read_frag21<NA>(f0, As, Bs, ra0, rb0, lh);
    read_frag21<NA>(f1, As, Bs, ra0, rb0, 2 + lh);   // in flight during the first step's MFMAs
    frag21_mma<MODE, NA>(f0, re, im);
    frag21_mma<MODE, NA>(f1, re, im);
}

// Diagonal 128-tiles: the 10 blocks (r, c), r <= c, of the 4x4 grid of
// 32x32 blocks as 3, 3, 2, 2 per SIMD (SIMD = wave % 4) instead of the
// quadrant layout's 2, 2, 4, 4 (a wave runs its blocks' MFMAs even when one
// of its two is below the diagonal).  Wave -> block rows rp .. rp + nb - 1
// of block column cb; nb = 0: no blocks.
__device__ inline int fused8_diag_blocks(int wave, int& rp, int& cb) {
    if (wave < 4) {                 // pairs (0-1, 1), (0-1, 2), (0-1, 3), (2-3, 3)
        rp = wave == 3 ? 2 : 0;
        cb = wave == 3 ? 3 : wave + 1;
        return 2;
    }
    rp = wave == 5 ? 2 : 0;         // singles (0, 0) and (2, 2)
    cb = rp;
    return wave < 6 ? 1 : 0;
}

// KIND 0: build only; 1: MFMAs of the wave's two row blocks; 2: of its
// first row block only.  ra0 / rb0: the lane's first A row / its B row.
template <int MODE, int RPT, int KIND>
__device__ __forceinline__ void fused8_split(char* in0, char* in1, char* in2, char* in3, char* op0, char* op1,
                                             const float* amp_t, const double* crd_t, size_t amp_stage,
                                             int s_begin, int n_st, const int* ampoff, int crdoff, int lane,
                                             int wave, const double (&ax)[RPT], const double (&ay)[RPT],
                                             const double (&az)[RPT], bool diag, int ra0, int rb0,
                                             f32x16 (&re)[2], f32x16 (&im)[2]) {
    using L = Fused8Lds;
    constexpr int TILE = 128;
    constexpr int PER = L::PER_STAGE;
    char* ins[FRING] = {in0, in1, in2, in3};
    char* ops[2] = {op0, op1};
    const int g = wave & 3, h = wave >> 2;
    const int lh = lane >> 5;
    constexpr int NA = KIND == 2 ? 1 : 2;
    const int boff = diag ? 0 : L::OP;
    // rows this lane builds: RPT = 2 (other tiles): h*128 + lane + 64 i;
    // RPT = 1 (diagonal tiles, A rows only): h*64 + lane
    const int rbase = (RPT == 2) ? h * 128 + lane : h * 64 + lane;
#define HV_F8ISSUE(slot, st)                                                                         \
    fused8_issue(slot, amp_t + (size_t)(s_begin + min(st, n_st - 1)) * amp_stage,                     \
                 crd_t + (size_t)(s_begin + min(st, n_st - 1)) * 16, ampoff, crdoff, wave)


#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 16; ++e) { re[i][e] = 0; im[i][e] = 0; }

    HV_F8ISSUE(in0, 0);
    HV_F8ISSUE(in1, 1);
    HV_F8ISSUE(in2, 2);
    HV_F8ISSUE(in3, 3);
    wait_vmcnt<(FRING - 1) * PER>();
    lds_barrier();
    fused8_build<MODE, RPT>(in0, op0, rbase, g, ax, ay, az);

    for (int st = 0; st < n_st; st += FRING) {
#pragma unroll
        for (int k = 0; k < FRING; ++k) {
            char* cur = ops[k & 1];
            char* nxt = ops[(k + 1) & 1];
            wait_vm_lgkm<(FRING - 2) * PER, 0>();
            lds_barrier();
            HV_F8ISSUE(ins[k], st + k + FRING);
            // the two waves of a SIMD (h = 0 and 1) take the build and the
            // MFMAs in opposite order, so one's matrix work runs while the
            // other's VALU work does
            if (h == 0) {
                fused8_build<MODE, RPT>(ins[(k + 1) % FRING], nxt, rbase, g, ax, ay, az);
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (KIND != 0) fused8_mma<MODE, NA>(cur, cur + boff, ra0, rb0, lh, re, im);
            } else {
                if constexpr (KIND != 0) fused8_mma<MODE, NA>(cur, cur + boff, ra0, rb0, lh, re, im);
                __builtin_amdgcn_sched_barrier(0);
                fused8_build<MODE, RPT>(ins[(k + 1) % FRING], nxt, rbase, g, ax, ay, az);
            }
        }
    }
#undef HV_F8ISSUE
    wait_vm_lgkm<0, 0>();
    lds_barrier();
}

template <int MODE>
__global__ __launch_bounds__(512) void herk_fused8_kernel(FusedArgs a) {
    using L = Fused8Lds;
    constexpr int TILE = 128;
    __shared__ __attribute__((aligned(1024))) char in0[L::IN];
    __shared__ __attribute__((aligned(1024))) char in1[L::IN];
    __shared__ __attribute__((aligned(1024))) char in2[L::IN];
    __shared__ __attribute__((aligned(1024))) char in3[L::IN];
    __shared__ __attribute__((aligned(1024))) char op0[2 * L::OP];
    __shared__ __attribute__((aligned(1024))) char op1[2 * L::OP];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = wave >> 2, wc = wave & 3;
    const int kst = (int)(eff_kpad(a.kpad, a.dyn_kpad) / 16);
    const int64_t nitems = (int64_t)a.nunits * a.nsplit * a.ntiles;
    const size_t amp_stage = (size_t)a.nant * 16;
    const size_t amp_time = (size_t)a.kpad * a.nant;
    const int st_split = a.ksplit / 16;   // stages per K-split

    for (int64_t Lw = blockIdx.x; Lw < nitems; Lw += gridDim.x) {
        int unit, split, tile, bi, bj;
        decode_work(a.nsplit, a.ntiles, Lw, unit, split, tile);
        decode_tile(tile, a.nb, bi, bj);
        const bool diag = (bi == bj);
        const int s_begin = split * st_split;
        const int s_end = min(s_begin + st_split, kst);
        if (s_begin >= s_end) continue;   // split beyond the job's sources (uniform per workgroup)
        const int n_st = s_end - s_begin;
        const int tb = unit / a.nfreq, fl = unit - tb * a.nfreq;
        const double il = a.inv_lambda[fl];   // turns per metre
        const float* amp_t = a.amp + (size_t)(a.amp_per_unit ? unit : tb) * amp_time;
        const double* crd_t = a.crd_top + (size_t)tb * 3 * a.kpad;

        // the wave's blocks: block rows rp, rp + 1 (nb of them) of column cb
        int rp = 2 * wr, cb = wc, nb = 2;
        if (diag) nb = fused8_diag_blocks(wave, rp, cb);
        unsigned smask = 0;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bool on = i < nb && (bi * TILE + (rp + i) * 32 < a.nant) && (bj * TILE + cb * 32 < a.nant);
            smask |= (unsigned)on << i;
        }
        smask = __builtin_amdgcn_readfirstlane(smask);
        const int ra0 = rp * 32 + (lane & 31), rb0 = cb * 32 + (lane & 31);

        // amplitude chunk c = (i*8 + wave)*64 + lane: row c/4, part c%4
        int ampoff[L::NAMP];
#pragma unroll
        for (int i = 0; i < L::NAMP; ++i) {
            const int c = (i * 8 + wave) * 64 + lane;
            const int rr = c >> 2;
            const int ant = min((rr < TILE ? bi * TILE + rr : bj * TILE + rr - TILE), a.nant - 1);
            ampoff[i] = ant * 16 + ((c & 3) ^ amp_swz(rr)) * 4;
        }
        const int crdoff = min(wave & 3, 2) * (int)a.kpad + 2 * (lane & 7);

        f32x16 re[2], im[2];
        const int h = wave >> 2;
        if (diag) {
            double ax[1], ay[1], az[1];
            const int ant = min(bi * TILE + h * 64 + lane, a.nant - 1);
            ax[0] = a.antpos[3 * ant] * il; ay[0] = a.antpos[3 * ant + 1] * il; az[0] = a.antpos[3 * ant + 2] * il;
            if (smask == 3)
                fused8_split<MODE, 1, 1>(in0, in1, in2, in3, op0, op1, amp_t, crd_t, amp_stage, s_begin, n_st,
                                         ampoff, crdoff, lane, wave, ax, ay, az, true, ra0, rb0, re, im);
            else if (smask)   // only block row rp (the second is off the list or past the antennas)
                fused8_split<MODE, 1, 2>(in0, in1, in2, in3, op0, op1, amp_t, crd_t, amp_stage, s_begin, n_st,
                                         ampoff, crdoff, lane, wave, ax, ay, az, true, ra0, rb0, re, im);
            else
                fused8_split<MODE, 1, 0>(in0, in1, in2, in3, op0, op1, amp_t, crd_t, amp_stage, s_begin, n_st,
                                         ampoff, crdoff, lane, wave, ax, ay, az, true, ra0, rb0, re, im);
        } else {
            double ax[2], ay[2], az[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int rr = h * 128 + lane + 64 * i;
                const int ant = min((rr < TILE ? bi * TILE + rr : bj * TILE + rr - TILE), a.nant - 1);
                ax[i] = a.antpos[3 * ant] * il; ay[i] = a.antpos[3 * ant + 1] * il; az[i] = a.antpos[3 * ant + 2] * il;
            }
            if (smask)
                fused8_split<MODE, 2, 1>(in0, in1, in2, in3, op0, op1, amp_t, crd_t, amp_stage, s_begin, n_st,
                                         ampoff, crdoff, lane, wave, ax, ay, az, false, ra0, rb0, re, im);
            else
                fused8_split<MODE, 2, 0>(in0, in1, in2, in3, op0, op1, amp_t, crd_t, amp_stage, s_begin, n_st,
                                         ampoff, crdoff, lane, wave, ax, ay, az, false, ra0, rb0, re, im);
        }
        if (!smask) continue;

        constexpr size_t TSZ = (size_t)TILE * TILE;
        float* out = reinterpret_cast<float*>(a.partials) +
                     (((size_t)unit * a.nsplit + split) * a.ntiles + tile) * 2 * TSZ;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (!(smask & (1u << i))) continue;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int r = 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                const int c = lane & 31;
                const int row = (rp + i) * 32 + r;
                const int col = cb * 32 + c;
                out[(size_t)row * TILE + col] = re[i][e];
                out[TSZ + (size_t)row * TILE + col] = im[i][e];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 192-tile build-in-place contraction: 8 waves, two per SIMD.  The SIMD's
// pair of waves owns a set of 32x32 MFMA blocks; one wave of the pair
// accumulates Re, the other Im, of the same blocks (144 accumulator
// registers each at most), reading the same operand fragments.  Against the
// 128-tile kernel a tile covers 2.25x the output for 1.5x the operand rows
// it builds: at N = 350 (two 192-blocks, three tiles) each antenna's
// operands are built twice per unit instead of three times.  Blocks:
//   off-diagonal tiles: SIMD q takes the 96x96 quadrant (q / 2, q % 2),
//     3x3 blocks;
//   diagonal tiles: the 21 blocks (r, c), r <= c, of the 6x6 block grid, 6
//     or 5 per SIMD (q 0: row 0; q 1: row 1; q 2: row 2 and (5, 5); q 3:
//     rows 3 and 4), so a diagonal tile costs 2/3 of an off-diagonal one.
// Each block runs the MFMA chain of the 128-tile kernels (per stage: source
// groups 0/1, then 2/3; hi.hi, hi.lo, lo.hi; one chain per split), so the
// partial sums and the visibilities are those kernels' bit for bit.
//
// LDS: two input slots (amplitudes of the 384 rows, the stage's x, y, z)
// filled by global_load_lds one stage ahead and two operand buffers (A rows,
// then B rows); each lane keeps the antenna positions of its rows.  Per stage k, after one
// barrier: issue the inputs of stage k+2, build operand rows of stage k+1
// (rows per lane: T192Rows), then run the MFMAs of stage k.
//
// Work order: the workgroups resident on one XCD take the same (split,
// tile) of up to 32 units (frequencies of one time) at once, so the
// amplitude and coordinate rows they read come from that XCD's L2.
struct T192 {
    static constexpr int TILE = 192;
    static constexpr int ROWS = 2 * TILE;                // A rows, then B rows
    static constexpr int AMP = ROWS * 64;                // 24 KB of amplitudes per stage
    static constexpr int CRD = 512;                      // x, y, z of 16 sources (128 B each)
    static constexpr int IN = AMP + CRD;
    static constexpr int OP = TILE * STAGE_BYTES;        // 24 KB per operand block
    static constexpr int NAMP = ROWS * 4 / 512;          // amplitude loads per thread per stage
};

// LDS pointers kept in the LDS address space (32-bit addresses).
typedef __attribute__((address_space(3))) char lds_char;
template <typename T>
__device__ __forceinline__ T lds_ld(const lds_char* p) {
    return *reinterpret_cast<const __attribute__((address_space(3))) T*>(p);
}
template <typename T>
__device__ __forceinline__ void lds_st(lds_char* p, const T& v) {
    *reinterpret_cast<__attribute__((address_space(3))) T*>(p) = v;
}

// Inputs of one stage: amplitude chunks c = (i*8 + wave)*64 + lane of the
// NROW rows (row c/4), and coordinate w of the stage's 16 sources by lanes
// 0-7 of wave w < 3.
template <int NROW>
__device__ __forceinline__ void t192_issue(lds_char* slot, const float* amp_st, const double* crd_st,
                                           const int* ampoff, int crdoff, int wave, int lane) {
#pragma unroll
    for (int i = 0; i < (NROW * 4 + 511) / 512; ++i)
        if ((i * 8 + wave) * 64 < NROW * 4)
            __builtin_amdgcn_global_load_lds(amp_st + ampoff[i], slot + (i * 8 + wave) * 1024, 16, 0, 0);
    if (wave < 3 && lane < 8)
        __builtin_amdgcn_global_load_lds(crd_st + crdoff, slot + T192::AMP + wave * 128, 16, 0, 0);
}

// Rows a lane builds: lane l takes source group l / 16 (4 sources) of rows
// row0 + (l % 16) + 16 k, k < NR.  Off-diagonal tiles: wave w builds rows
// 48 w .. 48 w + 47 (NR 3; rows 0-191 are the A block, 192-383 the B block);
// diagonal tiles (192 rows): waves 0-3 rows 32 w .. 32 w + 31 (NR 2), waves
// 4-7 rows 128 + 16 (w - 4) .. + 15 (NR 1).
template <bool DIAG, bool IM>
struct T192Rows {
    static constexpr int NR = DIAG ? (IM ? 1 : 2) : 3;
    __device__ static int row0(int wave) { return DIAG ? (IM ? 128 + 16 * (wave - 4) : 32 * wave) : 48 * wave; }
};

// Operand buffers of the 192-tile kernel (its own LDS layout; no other kernel
// reads it).  Row r (A block rows 0-191, then B block rows, 128 bytes each)
// holds the stage's 16 sources as four 16-byte chunks of 4 sources, hi and lo
// halves: chunk 2 g + h for source group g.  A chunk holds its 4 sources
// planar, (x0 x1 x2 x3 y0 y1 y2 y3) as 16-bit values: both operands of every
// MFMA use the same order, so each output sums the products of the
// interleaved order, and a rotation (y, -x) is a swap of word pairs and two
// sign flips.  The chunk index is XOR-ed with (r >> 1) & 7, which spreads the
// 16 rows of a 16-lane phase of ds_read_b128 / ds_write_b128 over all 64
// banks; every term but the row's 32-row block is then fixed per lane, so
// block offsets ride in the instructions' immediate offsets.
__device__ __forceinline__ int t192_off(int r, int q) { return r * STAGE_BYTES + ((q ^ ((r >> 1) & 7)) << 4); }

// hi and lo words of two values a = am_a c_a, b = am_b c_b, a in the low half:
// hi = RNE(round(am c)) and lo = RNE(fma(am, c, -hi)) per value, the
// arithmetic of split_phasor (no contraction), so the halves are those of
// every other operand builder.
template <int MODE>
__device__ __forceinline__ void split_two(float am_a, float c_a, float am_b, float c_b, uint32_t& hi, uint32_t& lo) {
#pragma clang fp contract(off)
    const float va = am_a * c_a, vb = am_b * c_b;
    if constexpr (MODE == MODE_SPLIT_BF16) {
        hi = __builtin_bit_cast(uint32_t, __builtin_convertvector((hv_f32x2){va, vb}, hv_bf16x2));
        const float ha = __uint_as_float(hi << 16), hb = __uint_as_float(hi & 0xffff0000u);
        lo = __builtin_bit_cast(uint32_t,
                                __builtin_convertvector((hv_f32x2){fmaf(am_a, c_a, -ha), fmaf(am_b, c_b, -hb)}, hv_bf16x2));
    } else {
        const hv_f16x2 h = __builtin_convertvector((hv_f32x2){va, vb}, hv_f16x2);
        hi = __builtin_bit_cast(uint32_t, h);
        const hv_f32x2 b = __builtin_convertvector(h, hv_f32x2);
        lo = __builtin_bit_cast(uint32_t,
                                __builtin_convertvector((hv_f32x2){fmaf(am_a, c_a, -b[0]), fmaf(am_b, c_b, -b[1])}, hv_f16x2));
    }
}

// Operand rows of one lane (T192Rows) for its source group of the stage;
// ax/ay/az: the rows' antenna positions in wavelengths.  Phases and halves
// are those of fused8_build; only their places in LDS differ.
template <int MODE, int NR>
__device__ __forceinline__ void t192_build(const lds_char* in, lds_char* op, int row0, int lane,
                                           const double (&ax)[3], const double (&ay)[3], const double (&az)[3]) {
    using L = T192;
    const int g = lane >> 4, l16 = lane & 15;
    const lds_char* cin = in + L::AMP + 32 * g;                                     // x; y at +128, z at +256
    const lds_char* ain = in + (row0 + l16) * 64 + ((g ^ ((l16 >> 1) & 3)) << 4);   // +1024 per 16 rows
    lds_char* wop = op + t192_off(row0 + l16, 2 * g);                               // +2048 per 16 rows
    const int wlo = (((2 * g + 1) ^ ((l16 >> 1) & 7)) - ((2 * g) ^ ((l16 >> 1) & 7))) << 4;
    double x[4], y[4], z[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        x[j] = lds_ld<double>(cin + 8 * j);
        y[j] = lds_ld<double>(cin + 128 + 8 * j);
        z[j] = lds_ld<double>(cin + 256 + 8 * j);
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const f32x4 am = lds_ld<f32x4>(ain + 1024 * k);
        float sn[4], cs[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) sincos_turns_hw(phase_turns(ax[k], ay[k], az[k], x[j], y[j], z[j]), &sn[j], &cs[j]);
        uint32_t h0, h1, h2, h3, l0, l1, l2, l3;
        split_two<MODE>(am[0], cs[0], am[1], cs[1], h0, l0);
        split_two<MODE>(am[2], cs[2], am[3], cs[3], h1, l1);
        split_two<MODE>(am[0], sn[0], am[1], sn[1], h2, l2);
        split_two<MODE>(am[2], sn[2], am[3], sn[3], h3, l3);
        lds_st(wop + 2048 * k, u32x4{h0, h1, h2, h3});
        lds_st(wop + 2048 * k + wlo, u32x4{l0, l1, l2, l3});
    }
}

// The same for rows whose amplitude rows are not their operand rows (items
// of several units: operand row v takes amplitude row v mod 32 NBK):
// arow[k], the amplitude row of the lane's group k (multiple of 16).
template <int MODE, int NR>
__device__ __forceinline__ void t192_build_mu(const lds_char* in, lds_char* op, int row0, int lane,
                                              const int (&arow)[3], const double (&ax)[3], const double (&ay)[3],
                                              const double (&az)[3]) {
    using L = T192;
    const int g = lane >> 4, l16 = lane & 15;
    const lds_char* cin = in + L::AMP + 32 * g;
    const int aswz = (g ^ ((l16 >> 1) & 3)) << 4;
    lds_char* wop = op + t192_off(row0 + l16, 2 * g);
    const int wlo = (((2 * g + 1) ^ ((l16 >> 1) & 7)) - ((2 * g) ^ ((l16 >> 1) & 7))) << 4;
    double x[4], y[4], z[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        x[j] = lds_ld<double>(cin + 8 * j);
        y[j] = lds_ld<double>(cin + 128 + 8 * j);
        z[j] = lds_ld<double>(cin + 256 + 8 * j);
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const f32x4 am = lds_ld<f32x4>(in + (arow[k] + l16) * 64 + aswz);
        float sn[4], cs[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) sincos_turns_hw(phase_turns(ax[k], ay[k], az[k], x[j], y[j], z[j]), &sn[j], &cs[j]);
        uint32_t h0, h1, h2, h3, l0, l1, l2, l3;
        split_two<MODE>(am[0], cs[0], am[1], cs[1], h0, l0);
        split_two<MODE>(am[2], cs[2], am[3], cs[3], h1, l1);
        split_two<MODE>(am[0], sn[0], am[1], sn[1], h2, l2);
        split_two<MODE>(am[2], sn[2], am[3], sn[3], h3, l3);
        lds_st(wop + 2048 * k, u32x4{h0, h1, h2, h3});
        lds_st(wop + 2048 * k + wlo, u32x4{l0, l1, l2, l3});
    }
}

// acc += a . b (Re) or a . rot(b) (Im) as hi.hi + hi.lo + lo.hi.
template <int MODE>
__device__ __forceinline__ void t192_mac(const u32x4& ahi, const u32x4& alo, const u32x4& bhi, const u32x4& blo,
                                         f32x16& acc) {
    acc = mfma16<MODE>(ahi, bhi, acc);
    acc = mfma16<MODE>(ahi, blo, acc);
    acc = mfma16<MODE>(alo, bhi, acc);
}

// Planar chunks (x0 x1 x2 x3 y0 y1 y2 y3): a' = (a_y, -a_x), with
// a' . b = a . rot(b) = a_y b_x - a_x b_y (rotating the A side), and
// b' = (-b_y, b_x) (rotating the B side).  The same two products per source
// land in the same MFMA lanes either way.
__device__ __forceinline__ u32x4 t192_rot_a(u32x4 v) {
    return u32x4{v[2], v[3], v[0] ^ 0x80008000u, v[1] ^ 0x80008000u};
}
__device__ __forceinline__ u32x4 t192_rot_b(u32x4 v) {
    return u32x4{v[2] ^ 0x80008000u, v[3] ^ 0x80008000u, v[0], v[1]};
}

// Fragment of 32-row block R (rows 32 R + lane % 32) for one step; fo: the
// lane's byte offsets of its hi and lo chunks inside a block (T192Frag).
// ROT 1: rotated as B, 2: as A.
struct T192Frag {
    int hi[2], lo[2];   // per step s
    __device__ explicit T192Frag(int lane) {
        const int lrow = lane & 31, lh = lane >> 5;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            hi[s] = t192_off(lrow, 2 * (2 * s + lh));
            lo[s] = t192_off(lrow, 2 * (2 * s + lh) + 1);
        }
    }
};
template <int ROT>
__device__ __forceinline__ void t192_frag(const lds_char* S, int R, int fh, int fl, u32x4& hi, u32x4& lo) {
    hi = lds_ld<u32x4>(S + fh + 32 * STAGE_BYTES * R);
    lo = lds_ld<u32x4>(S + fl + 32 * STAGE_BYTES * R);
    if constexpr (ROT == 1) {
        hi = t192_rot_b(hi);
        lo = t192_rot_b(lo);
    } else if constexpr (ROT == 2) {
        hi = t192_rot_a(hi);
        lo = t192_rot_a(lo);
    }
}

// Off-diagonal tiles: the 3x3 blocks of one quadrant (As, Bs: its first A and
// B block rows), one component, both steps of a stage.
template <int MODE, bool IM>
__device__ __forceinline__ void t192_quad_mma(const lds_char* As, const lds_char* Bs, const T192Frag& fo,
                                              f32x16 (&acc)[9]) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        // fragments in order of first use (LDS reads return in order)
        u32x4 bhi[3], blo[3];
        u32x4 ahi, alo;
        t192_frag<0>(As, 0, fo.hi[s], fo.lo[s], ahi, alo);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            t192_frag<IM ? 1 : 0>(Bs, j, fo.hi[s], fo.lo[s], bhi[j], blo[j]);
            t192_mac<MODE>(ahi, alo, bhi[j], blo[j], acc[j]);
        }
#pragma unroll
        for (int i = 1; i < 3; ++i) {
            t192_frag<0>(As, i, fo.hi[s], fo.lo[s], ahi, alo);
#pragma unroll
            for (int j = 0; j < 3; ++j) t192_mac<MODE>(ahi, alo, bhi[j], blo[j], acc[3 * i + j]);
        }
    }
}

// The first two fragments of a stage's first step read ahead of the build
// (tri_pre, runs_pre): for a rectangle or a row run its A row (rotated for
// Im) and first B column, for a triangle its first two rows.
struct TriPre {
    u32x4 h1, l1, h2, l2;
};

// Diagonal tiles: block row R, columns C0..C1 into acc[K0 ...], one step
// (PRE: its A fragment and first B fragment already read, in pre).
template <int MODE, bool IM, int R, int C0, int C1, int K0, bool PRE = false>
__device__ __forceinline__ void t192_diag_row(const lds_char* S, int fh, int fl, f32x16 (&acc)[9],
                                              const TriPre* pre = nullptr) {
    // Im: the one A fragment of the row is rotated instead of its B fragments
    constexpr int RA = IM ? 2 : 0;
    u32x4 ahi, alo;
    if constexpr (PRE) {
        ahi = pre->h1;
        alo = pre->l1;
    } else {
        t192_frag<RA>(S, R, fh, fl, ahi, alo);
    }
#pragma unroll
    for (int c = C0; c <= C1; ++c) {
        u32x4 bhi, blo;
        if (PRE && c == C0) {
            bhi = pre->h2;
            blo = pre->l2;
        } else {
            t192_frag<0>(S, c, fh, fl, bhi, blo);
        }
        t192_mac<MODE>(ahi, alo, bhi, blo, acc[K0 + c - C0]);
    }
}

// The A row R and B column C of a run list's first run, step 0.
template <bool IM>
__device__ __forceinline__ TriPre runs_pre(const lds_char* S, const T192Frag& fo, int R, int C) {
    TriPre p;
    t192_frag<IM ? 2 : 0>(S, R, fo.hi[0], fo.lo[0], p.h1, p.l1);
    t192_frag<0>(S, C, fo.hi[0], fo.lo[0], p.h2, p.l2);
    return p;
}

template <int MODE, bool IM, int Q>
__device__ __forceinline__ void t192_diag_mma(const lds_char* S, const T192Frag& fo, f32x16 (&acc)[9]) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int fh = fo.hi[s], fl = fo.lo[s];
        if constexpr (Q == 0) t192_diag_row<MODE, IM, 0, 0, 5, 0>(S, fh, fl, acc);
        if constexpr (Q == 1) t192_diag_row<MODE, IM, 1, 1, 5, 0>(S, fh, fl, acc);
        if constexpr (Q == 2) {
            t192_diag_row<MODE, IM, 2, 2, 5, 0>(S, fh, fl, acc);
            t192_diag_row<MODE, IM, 5, 5, 5, 4>(S, fh, fl, acc);
        }
        if constexpr (Q == 3) {
            t192_diag_row<MODE, IM, 3, 3, 5, 0>(S, fh, fl, acc);
            t192_diag_row<MODE, IM, 4, 4, 5, 3>(S, fh, fl, acc);
        }
        // the 15 blocks of a diagonal tile with five block rows inside the
        // array (the last tile at N = 321..352), 4, 4, 4, 3 per SIMD
        if constexpr (Q == 4) t192_diag_row<MODE, IM, 0, 0, 3, 0>(S, fh, fl, acc);
        if constexpr (Q == 5) t192_diag_row<MODE, IM, 1, 1, 4, 0>(S, fh, fl, acc);
        if constexpr (Q == 6) {
            t192_diag_row<MODE, IM, 2, 2, 4, 0>(S, fh, fl, acc);
            t192_diag_row<MODE, IM, 0, 4, 4, 3>(S, fh, fl, acc);
        }
        if constexpr (Q == 7) {
            t192_diag_row<MODE, IM, 3, 3, 4, 0>(S, fh, fl, acc);
            t192_diag_row<MODE, IM, 4, 4, 4, 2>(S, fh, fl, acc);
        }
        // the 10 blocks of a diagonal tile with at most four block rows
        // inside the array (N <= 128 at the last block), 3, 3, 2, 2 per SIMD
        if constexpr (Q == 8) t192_diag_row<MODE, IM, 0, 0, 2, 0>(S, fh, fl, acc);
        if constexpr (Q == 9) {
            t192_diag_row<MODE, IM, 0, 3, 3, 0>(S, fh, fl, acc);
            t192_diag_row<MODE, IM, 1, 1, 2, 1>(S, fh, fl, acc);
        }
        if constexpr (Q == 10) {
            t192_diag_row<MODE, IM, 1, 3, 3, 0>(S, fh, fl, acc);
            t192_diag_row<MODE, IM, 2, 2, 2, 1>(S, fh, fl, acc);
        }
        if constexpr (Q == 11) {
            t192_diag_row<MODE, IM, 2, 3, 3, 0>(S, fh, fl, acc);
            t192_diag_row<MODE, IM, 3, 3, 3, 1>(S, fh, fl, acc);
        }
    }
}

// Off-diagonal tiles whose B block has at most five 32-column blocks inside
// the array (N = 321..352 at the last block): block rows R0..R1 x columns
// C0..C1 per SIMD (8, 8, 8, 6 blocks instead of 9, 9, 9, 9 with 6 wasted),
// B fragments held, A streamed.
template <int MODE, bool IM, int R0, int R1, int C0, int C1>
__device__ __forceinline__ void t192_rect_mma(const lds_char* As, const lds_char* Bs, const T192Frag& fo,
                                              f32x16 (&acc)[9]) {
    constexpr int NC = C1 - C0 + 1;
    // Im: rotate the side with fewer fragments per step
    constexpr bool ROTA = NC > R1 - R0 + 1;
    constexpr int RA = (IM && ROTA) ? 2 : 0, RB = (IM && !ROTA) ? 1 : 0;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int fh = fo.hi[s], fl = fo.lo[s];
        u32x4 bhi[NC], blo[NC];
        u32x4 ahi, alo;
        t192_frag<RA>(As, R0, fh, fl, ahi, alo);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            t192_frag<RB>(Bs, C0 + c, fh, fl, bhi[c], blo[c]);
            t192_mac<MODE>(ahi, alo, bhi[c], blo[c], acc[c]);
        }
#pragma unroll
        for (int r = R0 + 1; r <= R1; ++r) {
            t192_frag<RA>(As, r, fh, fl, ahi, alo);
#pragma unroll
            for (int c = 0; c < NC; ++c) t192_mac<MODE>(ahi, alo, bhi[c], blo[c], acc[(r - R0) * NC + c]);
        }
    }
}

template <int MODE, bool IM, int Q>
__device__ __forceinline__ void t192_rect5_mma(const lds_char* As, const lds_char* Bs, const T192Frag& fo,
                                               f32x16 (&acc)[9]) {
    if constexpr (Q == 0) t192_rect_mma<MODE, IM, 0, 1, 0, 3>(As, Bs, fo, acc);
    if constexpr (Q == 1) t192_rect_mma<MODE, IM, 2, 3, 0, 3>(As, Bs, fo, acc);
    if constexpr (Q == 2) t192_rect_mma<MODE, IM, 4, 5, 0, 3>(As, Bs, fo, acc);
    if constexpr (Q == 3) t192_rect_mma<MODE, IM, 0, 5, 4, 4>(As, Bs, fo, acc);
    // at most four B column blocks inside the array: 6 blocks per SIMD
    if constexpr (Q == 4) t192_rect_mma<MODE, IM, 0, 1, 0, 2>(As, Bs, fo, acc);
    if constexpr (Q == 5) t192_rect_mma<MODE, IM, 2, 3, 0, 2>(As, Bs, fo, acc);
    if constexpr (Q == 6) t192_rect_mma<MODE, IM, 4, 5, 0, 2>(As, Bs, fo, acc);
    if constexpr (Q == 7) t192_rect_mma<MODE, IM, 0, 5, 3, 3>(As, Bs, fo, acc);
}

// The first and the last diagonal tile as one item (As: the first block's
// rows, Bs: the last block's, whose last block column inside the array is
// BC = 4 or 3): the 21 blocks r <= c of the first and the 15 or 10 of the
// last, 9, 9, 9, 9 (8, 8, 8, 7) per SIMD, where the two tiles apart take 6
// and 4 (3) per SIMD in two items.
template <int MODE, bool IM, int Q, int BC>
__device__ __forceinline__ void t192_pair_mma(const lds_char* As, const lds_char* Bs, const T192Frag& fo,
                                              f32x16 (&acc)[9]) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int fh = fo.hi[s], fl = fo.lo[s];
        if constexpr (Q == 0) {
            t192_diag_row<MODE, IM, 0, 0, 5, 0>(As, fh, fl, acc);
            t192_diag_row<MODE, IM, 3, 3, BC, 6>(Bs, fh, fl, acc);
            if constexpr (BC == 4) t192_diag_row<MODE, IM, 4, 4, 4, 8>(Bs, fh, fl, acc);
        }
        if constexpr (Q == 1) {
            t192_diag_row<MODE, IM, 1, 1, 5, 0>(As, fh, fl, acc);
            t192_diag_row<MODE, IM, 1, 1, BC, 5>(Bs, fh, fl, acc);
        }
        if constexpr (Q == 2) {
            t192_diag_row<MODE, IM, 2, 2, 5, 0>(As, fh, fl, acc);
            t192_diag_row<MODE, IM, 0, 0, BC, 4>(Bs, fh, fl, acc);
        }
        if constexpr (Q == 3) {
            t192_diag_row<MODE, IM, 3, 3, 5, 0>(As, fh, fl, acc);
            t192_diag_row<MODE, IM, 4, 4, 5, 3>(As, fh, fl, acc);
            t192_diag_row<MODE, IM, 5, 5, 5, 5>(As, fh, fl, acc);
            t192_diag_row<MODE, IM, 2, 2, BC, 6>(Bs, fh, fl, acc);
        }
    }
}

// Arrays of 321..352 antennas (eleven 32-row blocks, two 192-blocks) as two
// items per (unit, split) that split the upper triangle of blocks by column:
// the triangle of blocks r <= c < 8 (36 blocks, antenna rows 0-255 built) and
// the strip of columns 8-10 (30 blocks, rows 0-351 built).  Together they
// build 608 operand rows per (unit, split), against 768 for the paired
// diagonal item and the off-diagonal item (384 rows each).  Both items keep
// the operand buffer's rows in antenna order (bi = 0, bj = 1), so block b
// sits at row 32 b of the buffer, and their blocks are stored to the three
// 192-tiles they belong to.  Each block runs the same MFMA chain as in the
// other lists, so the partial sums are the same bits.

// Blocks (r, c), r in R0..R1 and c in C0..C1 (all with r <= c), one step,
// into acc[K0 + (r - R0) * NC + c - C0]: B fragments held, A streamed (Im:
// the A fragments rotated); bhi/blo keep the B fragments for the caller.
template <int MODE, bool IM, int R0, int R1, int C0, int C1, int K0, bool PRE = false>
__device__ __forceinline__ void tri_rect(const lds_char* S, int fh, int fl, u32x4 (&bhi)[C1 - C0 + 1],
                                         u32x4 (&blo)[C1 - C0 + 1], f32x16 (&acc)[9],
                                         const TriPre* pre = nullptr) {
    constexpr int NC = C1 - C0 + 1;
    constexpr int RA = IM ? 2 : 0;
    u32x4 ahi, alo;
    if constexpr (PRE) {
        ahi = pre->h1;
        alo = pre->l1;
    } else {
        t192_frag<RA>(S, R0, fh, fl, ahi, alo);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if (PRE && c == 0) {
            bhi[0] = pre->h2;
            blo[0] = pre->l2;
        } else {
            t192_frag<0>(S, C0 + c, fh, fl, bhi[c], blo[c]);
        }
        t192_mac<MODE>(ahi, alo, bhi[c], blo[c], acc[K0 + c]);
    }
#pragma unroll
    for (int r = R0 + 1; r <= R1; ++r) {
        t192_frag<RA>(S, r, fh, fl, ahi, alo);
#pragma unroll
        for (int c = 0; c < NC; ++c) t192_mac<MODE>(ahi, alo, bhi[c], blo[c], acc[K0 + (r - R0) * NC + c]);
    }
}

// Blocks (r, c), R0 <= r <= c <= R1 (a triangle on the diagonal), in row
// order into acc[K0 ...], one step: each fragment read once, and rotated as
// the A side of its own row for Im.
template <int MODE, bool IM, int R0, int R1, int K0, bool PRE = false>
__device__ __forceinline__ void tri_tri(const lds_char* S, int fh, int fl, f32x16 (&acc)[9],
                                        const TriPre* pre = nullptr) {
    constexpr int n = R1 - R0 + 1;
    u32x4 hi[n], lo[n];
#pragma unroll
    for (int i = 0; i < n; ++i) {
        if (PRE && i == 0) {
            hi[0] = pre->h1;
            lo[0] = pre->l1;
        } else if (PRE && i == 1) {
            hi[1] = pre->h2;
            lo[1] = pre->l2;
        } else {
            t192_frag<0>(S, R0 + i, fh, fl, hi[i], lo[i]);
        }
    }
#pragma unroll
    for (int r = 0; r < n; ++r) {
        const u32x4 ahi = IM ? t192_rot_a(hi[r]) : hi[r];
        const u32x4 alo = IM ? t192_rot_a(lo[r]) : lo[r];
#pragma unroll
        for (int c = r; c < n; ++c) t192_mac<MODE>(ahi, alo, hi[c], lo[c], acc[K0 + r * n - r * (r - 1) / 2 + c - r]);
    }
}

// One diagonal block (B, B) or off-diagonal (A, B) from fragments already
// held (the strip's B fragments), A rotated for Im.
template <int MODE, bool IM>
__device__ __forceinline__ void tri_held(const u32x4& ahi0, const u32x4& alo0, const u32x4& bhi, const u32x4& blo,
                                         f32x16& acc) {
    const u32x4 ahi = IM ? t192_rot_a(ahi0) : ahi0;
    const u32x4 alo = IM ? t192_rot_a(alo0) : alo0;
    t192_mac<MODE>(ahi, alo, bhi, blo, acc);
}

// The two items, per SIMD q (9, 9, 9, 9 and 8, 8, 7, 7 blocks):
//   triangle q0: rows 0-2 x cols 5-7; q1: rows 0-2 x cols 2-4;
//            q2: (0,0) (0,1) (1,1) + rows 3-4 x cols 5-7;
//            q3: (3,3) (3,4) (4,4) + the triangle of rows 5-7;
//   strip    q: rows 2q, 2q+1 x cols 8-10, then q0: (8,8) (8,9);
//            q1: (8,10) (9,9); q2: (9,10); q3: (10,10).
// The step's first two fragments of list Q, read ahead of the build (the
// stage's operands are complete once the stage's barrier has passed).
template <int MODE, bool IM, int Q>
__device__ __forceinline__ TriPre tri_pre(const lds_char* S, const T192Frag& fo) {
    TriPre p;
    const int fh = fo.hi[0], fl = fo.lo[0];
    if constexpr (Q == 2 || Q == 3) {
        constexpr int R0 = Q == 2 ? 0 : 3;
        t192_frag<0>(S, R0, fh, fl, p.h1, p.l1);
        t192_frag<0>(S, R0 + 1, fh, fl, p.h2, p.l2);
    } else {
        constexpr int R0 = Q < 2 ? 0 : 2 * (Q - 4);
        constexpr int C0 = Q == 0 ? 5 : Q == 1 ? 2 : 8;
        t192_frag<IM ? 2 : 0>(S, R0, fh, fl, p.h1, p.l1);
        t192_frag<0>(S, C0, fh, fl, p.h2, p.l2);
    }
    return p;
}

template <int MODE, bool IM, int Q, bool P>
__device__ __forceinline__ void tri_mma_step(const lds_char* S, int fh, int fl, f32x16 (&acc)[9],
                                             const TriPre* pre) {
    {
        u32x4 bhi[3], blo[3];
        if constexpr (Q == 0) tri_rect<MODE, IM, 0, 2, 5, 7, 0, P>(S, fh, fl, bhi, blo, acc, pre);
        if constexpr (Q == 1) tri_rect<MODE, IM, 0, 2, 2, 4, 0, P>(S, fh, fl, bhi, blo, acc, pre);
        if constexpr (Q == 2) {
            tri_tri<MODE, IM, 0, 1, 0, P>(S, fh, fl, acc, pre);
            tri_rect<MODE, IM, 3, 4, 5, 7, 3>(S, fh, fl, bhi, blo, acc);
        }
        if constexpr (Q == 3) {
            tri_tri<MODE, IM, 3, 4, 0, P>(S, fh, fl, acc, pre);
            tri_tri<MODE, IM, 5, 7, 3>(S, fh, fl, acc);
        }
        if constexpr (Q >= 4) {
            constexpr int q = Q - 4;
            tri_rect<MODE, IM, 2 * q, 2 * q + 1, 8, 10, 0, P>(S, fh, fl, bhi, blo, acc, pre);
            if constexpr (q == 0) {
                tri_held<MODE, IM>(bhi[0], blo[0], bhi[0], blo[0], acc[6]);
                tri_held<MODE, IM>(bhi[0], blo[0], bhi[1], blo[1], acc[7]);
            }
            if constexpr (q == 1) {
                tri_held<MODE, IM>(bhi[0], blo[0], bhi[2], blo[2], acc[6]);
                tri_held<MODE, IM>(bhi[1], blo[1], bhi[1], blo[1], acc[7]);
            }
            if constexpr (q == 2) tri_held<MODE, IM>(bhi[1], blo[1], bhi[2], blo[2], acc[6]);
            if constexpr (q == 3) tri_held<MODE, IM>(bhi[2], blo[2], bhi[2], blo[2], acc[6]);
        }
    }
}

template <int MODE, bool IM, int Q, bool PRE = false>
__device__ __forceinline__ void tri_mma(const lds_char* S, const T192Frag& fo, f32x16 (&acc)[9],
                                        const TriPre* pre = nullptr) {
    tri_mma_step<MODE, IM, Q, PRE>(S, fo.hi[0], fo.lo[0], acc, pre);
    tri_mma_step<MODE, IM, Q, false>(S, fo.hi[1], fo.lo[1], acc, pre);
}

// Block k of list Q of tri_mma -> absolute 32-row block (br, bc), or false.
__host__ __device__ constexpr bool tri_block(int Q, int k, int& br, int& bc) {
    switch (Q) {
        case 0: br = k / 3; bc = 5 + k % 3; return k < 9;
        case 1: br = k / 3; bc = 2 + k % 3; return k < 9;
        case 2:
            if (k < 3) { br = k < 2 ? 0 : 1; bc = k < 1 ? 0 : 1; return true; }
            br = 3 + (k - 3) / 3; bc = 5 + (k - 3) % 3; return k < 9;
        case 3:
            if (k < 3) { br = k < 2 ? 3 : 4; bc = k < 1 ? 3 : 4; return true; }
            br = k < 6 ? 5 : k < 8 ? 6 : 7;
            bc = k < 6 ? 5 + (k - 3) : k < 8 ? 6 + (k - 6) : 7;
            return k < 9;
        default: {
            const int q = Q - 4;
            if (k < 6) { br = 2 * q + k / 3; bc = 8 + k % 3; return true; }
            if (q == 0) { br = 8; bc = 8 + (k - 6); return k < 8; }
            if (q == 1) { br = k == 6 ? 8 : 9; bc = k == 6 ? 10 : 9; return k < 8; }
            br = q == 2 ? 9 : 10; bc = 10;
            return k < 7;
        }
    }
}

// Build rows of the triangle/strip items: triangle (256 rows): wave w rows
// 32 w .. 32 w + 31; strip (352 rows): waves 0-5 rows 48 w .. 48 w + 47,
// waves 6-7 rows 288 + 32 (w - 6) .. + 31.
template <bool STRIP, bool IM, int Q>
struct TriRows {
    static constexpr int NR = !STRIP ? 2 : (IM && Q >= 2) ? 2 : 3;
    __device__ static int row0(int wave) { return !STRIP ? 32 * wave : wave < 6 ? 48 * wave : 288 + 32 * (wave - 6); }
};

// Arrays of 129..320 antennas (NBK = 5..10 blocks of 32 rows) take the same
// decomposition with lists of row runs (t192_diag_row: one block row against
// a range of block columns): up to eight blocks, one triangle item holding
// every block (each row built once per (unit, split)); at nine and ten
// blocks, the triangle of the first six or seven block columns and the strip
// of the last three.  A SIMD's list is at most three runs of at most nine
// blocks in all; rows are built in antenna order and spread over the eight
// waves in 16-row groups.
struct TRun {
    int r, c0, c1;   // block row r, block columns c0..c1; r < 0: no run
};
__host__ __device__ constexpr int tri_n1(int nbk) { return nbk <= 8 ? nbk : nbk - 3; }
__host__ __device__ constexpr int tri_items(int nbk) { return nbk <= 8 ? 1 : 2; }
// Run i of SIMD q's list in item `part` (0: triangle, 1: strip) for NBK = 5..10.
__host__ __device__ constexpr TRun tri_run(int nbk, int part, int q, int i) {
    constexpr TRun X{-1, 0, -1};
    if (part == 0) {
        const int n = tri_n1(nbk);
        if (n == 8) return i == 0 ? TRun{q, q, 7} : i == 1 ? TRun{7 - q, 7 - q, 7} : X;   // 9, 9, 9, 9
        if (n == 7) {                                                                        // 7, 7, 7, 7
            if (q == 0) return i == 0 ? TRun{0, 0, 6} : X;
            return i == 0 ? TRun{q, q, 6} : i == 1 ? TRun{7 - q, 7 - q, 6} : X;
        }
        if (n == 6) {                                                                        // 6, 5, 5, 5
            if (q < 2) return i == 0 ? TRun{q, q, 5} : X;
            return i == 0 ? TRun{q, q, 5} : i == 1 ? TRun{7 - q, 7 - q, 5} : X;
        }
        // n == 5: 4, 4, 4, 3
        if (q == 0) return i == 0 ? TRun{0, 0, 3} : X;
        if (q == 1) return i == 0 ? TRun{0, 4, 4} : i == 1 ? TRun{1, 1, 3} : X;
        if (q == 2) return i == 0 ? TRun{1, 4, 4} : i == 1 ? TRun{2, 2, 4} : X;
        return i == 0 ? TRun{3, 3, 4} : i == 1 ? TRun{4, 4, 4} : X;
    }
    if (nbk == 10) {   // columns 7-9: 7, 7, 7, 6
        if (q < 3) return i < 2 ? TRun{2 * q + i, 7, 9} : i == 2 ? TRun{7, 7 + q, 7 + q} : X;
        return i == 0 ? TRun{6, 7, 9} : i == 1 ? TRun{8, 8, 9} : i == 2 ? TRun{9, 9, 9} : X;
    }
    // nbk == 9: columns 6-8: 6, 6, 6, 6
    if (q < 3) return i < 2 ? TRun{2 * q + i, 6, 8} : X;
    return i == 0 ? TRun{6, 6, 8} : i == 1 ? TRun{7, 7, 8} : i == 2 ? TRun{8, 8, 8} : X;
}
__host__ __device__ constexpr int tri_list_blocks(int nbk, int part, int q) {
    int n = 0;
    for (int i = 0; i < 3; ++i) {
        const TRun r = tri_run(nbk, part, q, i);
        if (r.r >= 0) n += r.c1 - r.c0 + 1;
    }
    return n;
}

// Block k of a run list -> (block row, block column), or false past it.
__host__ __device__ constexpr bool tri_run_block(int nbk, int part, int q, int k, int& br, int& bc) {
    for (int i = 0; i < 3; ++i) {
        const TRun r = tri_run(nbk, part, q, i);
        if (r.r < 0) return false;
        const int n = r.c1 - r.c0 + 1;
        if (k < n) {
            br = r.r;
            bc = r.c0 + k;
            return true;
        }
        k -= n;
    }
    return false;
}

template <int MODE, bool IM, int NBK, int PART, int Q, int I, int K0, bool PRE = false>
__device__ __forceinline__ void tri_runs_step(const lds_char* S, int fh, int fl, f32x16 (&acc)[9],
                                              const TriPre* pre = nullptr) {
    if constexpr (I < 3) {
        constexpr TRun run = tri_run(NBK, PART, Q, I);
        if constexpr (run.r >= 0) {
            t192_diag_row<MODE, IM, run.r, run.c0, run.c1, K0, PRE && I == 0>(S, fh, fl, acc, pre);
            tri_runs_step<MODE, IM, NBK, PART, Q, I + 1, K0 + run.c1 - run.c0 + 1>(S, fh, fl, acc);
        }
    }
}
template <int MODE, bool IM, int NBK, int PART, int Q, bool PRE = false>
__device__ __forceinline__ void tri_runs_mma(const lds_char* S, const T192Frag& fo, f32x16 (&acc)[9],
                                             const TriPre* pre = nullptr) {
    static_assert(tri_list_blocks(NBK, PART, Q) <= 9, "a SIMD list holds at most nine blocks");
    tri_runs_step<MODE, IM, NBK, PART, Q, 0, 0, PRE>(S, fo.hi[0], fo.lo[0], acc, pre);
    tri_runs_step<MODE, IM, NBK, PART, Q, 0, 0>(S, fo.hi[1], fo.lo[1], acc);
}

// Items of several units (arrays of 33..128 antennas, NBK = 2..4): the
// U x NBK row runs (unit u, block row r: that unit's columns r .. NBK-1, as
// operand rows u NBK + r against u NBK + r .. u NBK + NBK-1) dealt to the
// four SIMDs longest first, each to the SIMD with the fewest blocks (the
// lowest index on ties).  Run i of SIMD q, or r < 0 past the list.
__host__ __device__ constexpr TRun mu_run(int nbk, int U, int q, int i) {
    int nblk[4] = {0, 0, 0, 0}, nrun[4] = {0, 0, 0, 0};
    for (int len = nbk; len >= 1; --len)
        for (int u = 0; u < U; ++u) {
            int b = 0;
            for (int k = 1; k < 4; ++k)
                if (nblk[k] < nblk[b]) b = k;
            if (b == q && nrun[b] == i) {
                const int r = u * nbk + nbk - len;
                return TRun{r, r, u * nbk + nbk - 1};
            }
            nblk[b] += len;
            ++nrun[b];
        }
    return TRun{-1, 0, -1};
}
__host__ __device__ constexpr int mu_list_blocks(int nbk, int U, int q) {
    int n = 0;
    for (int i = 0; i < 8; ++i) {
        const TRun r = mu_run(nbk, U, q, i);
        if (r.r >= 0) n += r.c1 - r.c0 + 1;
    }
    return n;
}
__host__ __device__ constexpr bool mu_run_block(int nbk, int U, int q, int k, int& br, int& bc) {
    for (int i = 0; i < 3; ++i) {
        const TRun r = mu_run(nbk, U, q, i);
        if (r.r < 0) return false;
        const int n = r.c1 - r.c0 + 1;
        if (k < n) {
            br = r.r;
            bc = r.c0 + k;
            return true;
        }
        k -= n;
    }
    return false;
}
template <int MODE, bool IM, int NBK, int U, int Q, int I, int K0, bool PRE = false>
__device__ __forceinline__ void mu_runs_step(const lds_char* S, int fh, int fl, f32x16 (&acc)[9],
                                             const TriPre* pre = nullptr) {
    if constexpr (I < 3) {
        constexpr TRun run = mu_run(NBK, U, Q, I);
        if constexpr (run.r >= 0) {
            t192_diag_row<MODE, IM, run.r, run.c0, run.c1, K0, PRE && I == 0>(S, fh, fl, acc, pre);
            mu_runs_step<MODE, IM, NBK, U, Q, I + 1, K0 + run.c1 - run.c0 + 1>(S, fh, fl, acc);
        }
    }
}
template <int MODE, bool IM, int NBK, int U, int Q, bool PRE = false>
__device__ __forceinline__ void mu_runs_mma(const lds_char* S, const T192Frag& fo, f32x16 (&acc)[9],
                                            const TriPre* pre = nullptr) {
    static_assert(mu_list_blocks(NBK, U, Q) <= 9, "a SIMD list holds at most nine blocks");
    static_assert(mu_run(NBK, U, Q, 3).r < 0, "a SIMD list holds at most three runs");
    mu_runs_step<MODE, IM, NBK, U, Q, 0, 0, PRE>(S, fo.hi[0], fo.lo[0], acc, pre);
    mu_runs_step<MODE, IM, NBK, U, Q, 0, 0>(S, fo.hi[1], fo.lo[1], acc);
}

// Build rows of the run-list items: the item's 16-row groups (rows 0 ..
// 32 n - 1 of the triangle's n block columns, or all NBK blocks' rows for
// the strip) split evenly over the waves, wave 4 IM + Q taking groups
// G0 .. G1 - 1 (one to three).
template <int NBK, int PART, bool IM, int Q>
struct TriRowsG {
    static constexpr int G = 2 * (PART == 0 ? tri_n1(NBK) : NBK);
    static constexpr int W = (IM ? 4 : 0) + Q;
    static constexpr int G0 = W * G / 8, G1 = (W + 1) * G / 8;
    static constexpr int NR = G1 - G0;
    static constexpr int ROWS = 16 * G;
    static_assert(NR >= 1 && NR <= 3, "one to three 16-row groups per wave");
    __device__ static int row0(int) { return 16 * G0; }
};

// Block k of a wave's accumulator set -> (block row, block column) in the
// tile (list DQ of t192_split; DQ < 0: quadrant q), or false past the set;
// `second`: the block lies in the second tile of a paired item (DQ >= 20).
__device__ inline bool t192_block(int dq, int q, int k, int& br, int& bc, int& second) {
    second = 0;
    if (dq >= 20) {
        const int BC = dq < 24 ? 4 : 3;
        const int Q = dq & 3;
        if (Q == 0) {
            if (k < 6) { br = 0; bc = k; return true; }
            second = 1;
            br = k < 8 ? 3 : 4;
            bc = k < 8 ? 3 + (k - 6) : 4;
            return BC == 4 ? k < 9 : k < 7;
        }
        if (Q == 1) {
            if (k < 5) { br = 1; bc = 1 + k; return true; }
            second = 1; br = 1; bc = 1 + (k - 5);
            return k < 5 + BC;
        }
        if (Q == 2) {
            if (k < 4) { br = 2; bc = 2 + k; return true; }
            second = 1; br = 0; bc = k - 4;
            return k < 5 + BC;
        }
        if (k < 6) {
            br = k < 3 ? 3 : k < 5 ? 4 : 5;
            bc = k < 3 ? 3 + k : k < 5 ? 4 + (k - 3) : 5;
            return true;
        }
        second = 1; br = 2; bc = 2 + (k - 6);
        return k < 5 + BC;
    }
    switch (dq) {
        case 0: br = 0; bc = k; return k < 6;
        case 1: br = 1; bc = 1 + k; return k < 5;
        case 2: br = k < 4 ? 2 : 5; bc = k < 4 ? 2 + k : 5; return k < 5;
        case 3: br = k < 3 ? 3 : 4; bc = k < 3 ? 3 + k : 4 + (k - 3); return k < 5;
        case 4: br = 0; bc = k; return k < 4;
        case 5: br = 1; bc = 1 + k; return k < 4;
        case 6: br = k < 3 ? 2 : 0; bc = k < 3 ? 2 + k : 4; return k < 4;
        case 7: br = k < 2 ? 3 : 4; bc = k < 2 ? 3 + k : 4; return k < 3;
        case 8: case 9: case 10: br = 2 * (dq - 8) + k / 4; bc = k % 4; return k < 8;
        case 11: br = k; bc = 4; return k < 6;
        case 12: br = 0; bc = k; return k < 3;
        case 13: br = k < 1 ? 0 : 1; bc = k < 1 ? 3 : k; return k < 3;
        case 14: br = k < 1 ? 1 : 2; bc = k < 1 ? 3 : 2; return k < 2;
        case 15: br = k < 1 ? 2 : 3; bc = 3; return k < 2;
        case 16: case 17: case 18: br = 2 * (dq - 16) + k / 3; bc = k % 3; return k < 6;
        case 19: br = k; bc = 3; return k < 6;
        default:
            br = (q >> 1) * 3 + k / 3;
            bc = (q & 1) * 3 + k % 3;
            return true;
    }
}

// One split of one tile for one wave.  DQ < 0: off-diagonal tile (quadrant
// q); DQ = 0..3: diagonal tile, block list DQ (21 blocks); DQ = 4..7:
// diagonal tile with at most five block rows inside the array (15 blocks);
// DQ = 8..11: off-diagonal tile with at most five B column blocks inside
// the array (30 blocks); DQ = 12..15: diagonal tile with at most four block
// rows inside (10 blocks); DQ = 16..19: off-diagonal tile with at most four
// B column blocks inside (24 blocks); DQ = 20..23 / 24..27: the first and
// the last diagonal tile as one item, at most five / four block rows of the
// last inside (t192_pair_mma).  Paired items build both blocks' rows
// as an off-diagonal tile does.  DQ = 28..31 / 32..35: the triangle / strip
// items of the triangle decomposition (NBK = 11: tri_mma lists 0..3 /
// 4..7; NBK = 5..10: the run lists of item 0 / 1, SIMD DQ & 3).
template <int MODE, bool IM, int DQ, int NBK>
__device__ __forceinline__ void t192_split(lds_char* in0, lds_char* in1, lds_char* op0, lds_char* op1,
                                           const double* antpos, int nant, int bi, int bj, const float* amp_t,
                                           const double* crd_t, size_t amp_stage, int s_begin, int n_st,
                                           const int* ampoff, int crdoff, int lane, int wave, double il,
                                           f32x16 (&acc)[9]) {
    using L = T192;
    constexpr bool DIAG = (DQ >= 0 && DQ < 8) || (DQ >= 12 && DQ < 16);
    constexpr bool RECT5 = (DQ >= 8 && DQ < 12) || (DQ >= 16 && DQ < 20);
    constexpr bool PAIR = DQ >= 20 && DQ < 28;
    constexpr bool TRI = DQ >= 28, STRIP = DQ >= 32;
    constexpr int QP = DQ & 3, BCP = DQ < 24 ? 4 : 3;   // t192_pair_mma list
    constexpr int QD = DQ < 8 ? DQ : DQ - 4;     // t192_diag_mma list
    constexpr int QR = DQ < 12 ? DQ - 8 : DQ - 12;   // t192_rect5_mma list
    constexpr int QT = DQ - 28;                  // tri_mma list
    constexpr bool RUNS = TRI && NBK < 11;       // run lists (tri_runs_mma)
    using RW = typename std::conditional<
        RUNS, TriRowsG<(RUNS ? NBK : 5), (STRIP ? 1 : 0), IM, (DQ & 3)>,
        typename std::conditional<TRI, TriRows<STRIP, IM, (DQ & 3)>, T192Rows<DIAG, IM>>::type>::type;
    constexpr int NROW = DIAG ? L::TILE : RUNS ? TriRowsG<(RUNS ? NBK : 5), (STRIP ? 1 : 0), IM, (DQ & 3)>::ROWS
                                      : TRI ? (STRIP ? 352 : 256) : L::ROWS;
    constexpr int NR = RW::NR;
    const int row0 = RW::row0(wave);
    // antenna positions of the rows this lane builds, in wavelengths
    double ax[3], ay[3], az[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int r = row0 + (lane & 15) + 16 * (k < NR ? k : 0);
        const int ant = min((r < L::TILE ? bi * L::TILE + r : bj * L::TILE + r - L::TILE), nant - 1);
        ax[k] = antpos[3 * ant] * il;
        ay[k] = antpos[3 * ant + 1] * il;
        az[k] = antpos[3 * ant + 2] * il;
    }
    const int q = wave & 3;
    const T192Frag fo(lane);
    // off-diagonal quadrant (q / 2, q % 2): its first A and B block rows
    const int qa = (q >> 1) * 96 * STAGE_BYTES, qb = L::OP + (q & 1) * 96 * STAGE_BYTES;
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[k][e] = 0;
#define HV_T192ISSUE(slot, st)                                                                               \
    t192_issue<NROW>(slot, amp_t + (size_t)(s_begin + (st)) * amp_stage, crd_t + (size_t)(s_begin + (st)) * 16, \
                     ampoff, crdoff, wave, lane)
#define HV_T192MMA(buf)                                                                   \
    do {                                                                                   \
        if constexpr (DIAG) t192_diag_mma<MODE, IM, (DIAG ? QD : 0)>(buf, fo, acc);          \
        else if constexpr (RECT5) t192_rect5_mma<MODE, IM, (RECT5 ? QR : 0)>(buf, buf + L::OP, fo, acc); \
        else if constexpr (PAIR) t192_pair_mma<MODE, IM, QP, BCP>(buf, buf + L::OP, fo, acc); \
        else if constexpr (RUNS) tri_runs_mma<MODE, IM, (RUNS ? NBK : 5), (STRIP ? 1 : 0), (DQ & 3), (NBK <= 8)>(buf, fo, acc, &pre); \
        else if constexpr (TRI) tri_mma<MODE, IM, (TRI ? QT : 0), true>(buf, fo, acc, &pre);   \
        else t192_quad_mma<MODE, IM>(buf + qa, buf + qb, fo, acc);                             \
    } while (0)

    HV_T192ISSUE(in0, 0);
    wait_vmcnt<0>();
    __syncthreads();                       // stage 0's inputs are in LDS
    HV_T192ISSUE(in1, min(1, n_st - 1));   // a one-stage split re-reads stage 0 (never built)
    t192_build<MODE, NR>(in0, op0, row0, lane, ax, ay, az);
    for (int st = 0; st < n_st; ++st) {
        const bool odd = st & 1;
        const lds_char* cur = odd ? op1 : op0;
        // stage st+1's inputs have landed and every wave's build of stage st
        // is in LDS; slot st & 1 (read by that build) and the operand buffer
        // of stage st-1 are free
        wait_vm_lgkm<0, 0>();
        lds_barrier();
        if (st + 2 < n_st) HV_T192ISSUE(odd ? in1 : in0, st + 2);
        // build stage st+1 (at the last stage: an earlier stage again, from a
        // slot nobody refills, into the buffer no MFMA reads)
        const lds_char* bin = odd ? in0 : in1;
        lds_char* bop = odd ? op0 : op1;
        // the first two fragments of this stage's MFMAs go out before the
        // build, so the MFMAs start without waiting on LDS (measured on
        // MI355X, 128 units, in one run each: N = 350 157.0 against 159.2 ms,
        // the whole first group read ahead spilled: 196 ms; N = 160 40.8
        // against 41.8, N = 256 81.1-82.2 against 81.1-83.1)
        TriPre pre;
        if constexpr (TRI && !RUNS) {
            pre = tri_pre<MODE, IM, (TRI ? QT : 0)>(cur, fo);
            __builtin_amdgcn_sched_barrier(0);
        } else if constexpr (RUNS && NBK <= 8) {
            // one item of every block (NBK <= 8); with a strip item too
            // (NBK 9, 10) the read-ahead spilled (N = 300: 153.0 against
            // 145.8 ms)
            constexpr TRun run0 = tri_run((RUNS ? NBK : 5), (STRIP ? 1 : 0), (DQ & 3), 0);
            if constexpr (run0.r >= 0) pre = runs_pre<IM>(cur, fo, run0.r, run0.c0);
            __builtin_amdgcn_sched_barrier(0);
        }
        // both waves of a SIMD build, then both run their MFMAs (measured on
        // MI355X: 7% faster than staggering the build and the MFMAs of the
        // pair, whose LDS traffic then competes)
        t192_build<MODE, NR>(bin, bop, row0, lane, ax, ay, az);
        __builtin_amdgcn_sched_barrier(0);
        HV_T192MMA(cur);
    }
#undef HV_T192MMA
#undef HV_T192ISSUE
    wait_vm_lgkm<0, 0>();
    lds_barrier();
}

// Work item L -> (unit, split, tile), or false for a slot with no work.
// Combos (unit group, split, tile) go round-robin to the 8 XCDs (block id
// mod 8); within an XCD, consecutive items are the ug units of one combo, so
// the workgroups resident on an XCD at once share their input rows.
__device__ inline bool t192_decode(int nunits, int nsplit, int ntiles, int ug, int64_t L, int& unit, int& split,
                                   int& tile) {
    const int x = (int)(L & 7);
    const int64_t q = L >> 3;
    const int64_t combo = (q / ug) * 8 + x;
    const int u_in = (int)(q % ug);
    const int nug = (nunits + ug - 1) / ug;
    if (combo >= (int64_t)nug * nsplit * ntiles) return false;
    tile = (int)(combo % ntiles);
    const int64_t rest = combo / ntiles;
    split = (int)(rest % nsplit);
    unit = (int)(rest / nsplit) * ug + u_in;
    return unit < nunits;
}

// Units per combo for a grid of g workgroups: at most g / 8 (the workgroups
// of one XCD), spread evenly over the launch's units.
__host__ __device__ inline int t192_unit_group(int nunits, int64_t g) {
    const int per_xcd = (int)(g / 8 > 0 ? g / 8 : 1);
    const int nug = (nunits + per_xcd - 1) / per_xcd;
    return (nunits + nug - 1) / nug;
}

// Arrays whose last block holds at most five 32-row blocks of antennas run
// the first and the last diagonal tile as one item (work tile 0: tiles 0 and
// ntiles - 1 of the decode order); work tile w = 1 .. ntiles - 2 is tile w.
// Two blocks: 2 items per (unit, split) instead of 3.
__host__ __device__ inline bool t192_pairs(int nb, int nant) { return nb >= 2 && nant - (nb - 1) * 192 <= 160; }
// Arrays of 129..352 antennas (5..11 blocks of 32 rows) take the items of
// the triangle decomposition (tri_mma, tri_runs_mma) in place of 192-tiles
// unless `lists` is 1; the block count, or 0 for the tile lists.
__host__ __device__ inline int t192_tri_nbk(int nant, int lists) {
    const int nbk = (nant + 31) / 32;
    return (lists != 1 && nbk >= 5 && nbk <= 11) ? nbk : 0;
}
__host__ __device__ inline int t192_work_tiles(int nb, int nant, int ntiles) {
    return t192_pairs(nb, nant) ? ntiles - 1 : ntiles;
}

__host__ __device__ inline int64_t t192_items(int nunits, int nsplit, int ntiles, int ug) {
    const int64_t nug = (nunits + ug - 1) / ug;
    const int64_t ncombo = nug * nsplit * ntiles;
    return (ncombo + 7) / 8 * 8 * ug;
}

// NBKT: 0 for the 192-tile lists, else the array's block count (5..11) for
// the items of the triangle decomposition; each instantiation holds only
// the lists it runs.
template <int MODE, int NBKT>
__global__ __launch_bounds__(512) void herk_t192_kernel(FusedArgs a) {
    using L = T192;
    constexpr int TILE = L::TILE;
    __shared__ __attribute__((aligned(1024))) char s_in0[L::IN];
    __shared__ __attribute__((aligned(1024))) char s_in1[L::IN];
    __shared__ __attribute__((aligned(1024))) char s_op0[2 * L::OP];
    __shared__ __attribute__((aligned(1024))) char s_op1[2 * L::OP];
    lds_char* in0 = (lds_char*)s_in0;
    lds_char* in1 = (lds_char*)s_in1;
    lds_char* op0 = (lds_char*)s_op0;
    lds_char* op1 = (lds_char*)s_op1;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int q = wave & 3;                // SIMD slot: quadrant or diagonal block list
    const bool im = wave >= 4;             // component this wave accumulates
    const int kst = (int)(eff_kpad(a.kpad, a.dyn_kpad) / 16);
    const size_t amp_stage = (size_t)a.nant * 16;
    const size_t amp_time = (size_t)a.kpad * a.nant;
    const int st_split = a.ksplit / 16;   // stages per K-split
    const int ug = t192_unit_group(a.nunits, gridDim.x);
    const bool pairs = t192_pairs(a.nb, a.nant);
    constexpr bool tri = NBKT > 0;   // work tile 0: triangle, 1: strip
    const int ntw = tri ? tri_items(NBKT) : t192_work_tiles(a.nb, a.nant, a.ntiles);
    const int64_t nitems = t192_items(a.nunits, a.nsplit, ntw, ug);

    for (int64_t Lw = blockIdx.x; Lw < nitems; Lw += gridDim.x) {
        int unit, split, tile, bi, bj;
        if (!t192_decode(a.nunits, a.nsplit, ntw, ug, Lw, unit, split, tile)) continue;
        // a paired item (work tile 0) builds the rows of the first and the
        // last block; its blocks are stored to the diagonal tiles 0 and
        // ntiles - 1
        const bool pair = pairs && tile == 0;
        if (tri) {
            bi = 0;   // rows in antenna order
            bj = 1;
        } else if (pair) {
            bi = 0;
            bj = a.nb - 1;
        } else {
            decode_tile(tile, a.nb, bi, bj);
        }
        const bool diag = (bi == bj);
        const int s_begin = split * st_split;
        const int s_end = min(s_begin + st_split, kst);
        if (s_begin >= s_end) continue;   // split beyond the job's sources (uniform per workgroup)
        const int n_st = s_end - s_begin;
        const int tb = unit / a.nfreq, fl = unit - tb * a.nfreq;
        const double il = a.inv_lambda[fl];   // turns per metre
        const float* amp_t = a.amp + (size_t)(a.amp_per_unit ? unit : tb) * amp_time;
        const double* crd_t = a.crd_top + (size_t)tb * 3 * a.kpad;

        // amplitude chunk c = (i*8 + wave)*64 + lane: row c/4, stored part
        // c%4 holds the row's part (c%4) ^ amp_swz(row)
        int ampoff[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int c = (i * 8 + wave) * 64 + lane;
            const int rr = c >> 2;
            const int ant = min((rr < TILE ? bi * TILE + rr : bj * TILE + rr - TILE), a.nant - 1);
            ampoff[i] = ant * 16 + ((c & 3) ^ amp_swz(rr)) * 4;
        }
        const int crdoff = min(wave, 2) * (int)a.kpad + 2 * lane;   // used by lanes 0-7 of waves 0-2

        f32x16 acc[9];
#define HV_T192SPLIT(IM, DQ)                                                                                 \
    t192_split<MODE, IM, DQ, NBKT>(in0, in1, op0, op1, a.antpos, a.nant, bi, bj, amp_t, crd_t, amp_stage, s_begin, n_st, \
                             ampoff, crdoff, lane, wave, il, acc)
#define HV_T192BOTH(DQ)              \
    do {                             \
        if (im) HV_T192SPLIT(true, DQ); \
        else HV_T192SPLIT(false, DQ);   \
    } while (0)
        // diagonal tiles: the 15-block lists when at most five 32-row blocks
        // of the tile hold antennas
        // off-diagonal tiles: the 30-block lists when at most five 32-column
        // blocks of the B block hold antennas
        // the lists that skip blocks past the last antenna: at most five or
        // four 32-row blocks of the last block (diagonal tiles) or of the B
        // block (off-diagonal tiles) hold antennas
        const int rem = a.nant - bj * TILE;
        const bool five = rem <= 160, four = rem <= 128;
        const int dq = tri ? (tile == 0 ? 28 : 32) + q
                     : pair ? (four ? 24 : 20) + q
                            : diag ? (four ? 12 : five ? 4 : 0) + q : (four ? 16 + q : five ? 8 + q : -1);
        if constexpr (tri) {
            if (dq == 28) HV_T192BOTH(28);
            else if (dq == 29) HV_T192BOTH(29);
            else if (dq == 30) HV_T192BOTH(30);
            else if (dq == 31) HV_T192BOTH(31);
            else if constexpr (tri_items(NBKT) == 2) {
                if (dq == 32) HV_T192BOTH(32);
                else if (dq == 33) HV_T192BOTH(33);
                else if (dq == 34) HV_T192BOTH(34);
                else HV_T192BOTH(35);
            }
        } else if (dq < 0) HV_T192BOTH(-1);
        else if (dq == 0) HV_T192BOTH(0);
        else if (dq == 1) HV_T192BOTH(1);
        else if (dq == 2) HV_T192BOTH(2);
        else if (dq == 3) HV_T192BOTH(3);
        else if (dq == 4) HV_T192BOTH(4);
        else if (dq == 5) HV_T192BOTH(5);
        else if (dq == 6) HV_T192BOTH(6);
        else if (dq == 7) HV_T192BOTH(7);
        else if (dq == 8) HV_T192BOTH(8);
        else if (dq == 9) HV_T192BOTH(9);
        else if (dq == 10) HV_T192BOTH(10);
        else if (dq == 11) HV_T192BOTH(11);
        else if (dq == 12) HV_T192BOTH(12);
        else if (dq == 13) HV_T192BOTH(13);
        else if (dq == 14) HV_T192BOTH(14);
        else if (dq == 15) HV_T192BOTH(15);
        else if (dq == 16) HV_T192BOTH(16);
        else if (dq == 17) HV_T192BOTH(17);
        else if (dq == 18) HV_T192BOTH(18);
        else if (dq == 19) HV_T192BOTH(19);
        else if (dq == 20) HV_T192BOTH(20);
        else if (dq == 21) HV_T192BOTH(21);
        else if (dq == 22) HV_T192BOTH(22);
        else if (dq == 23) HV_T192BOTH(23);
        else if (dq == 24) HV_T192BOTH(24);
        else if (dq == 25) HV_T192BOTH(25);
        else if (dq == 26) HV_T192BOTH(26);
        else HV_T192BOTH(27);
#undef HV_T192BOTH
#undef HV_T192SPLIT

        constexpr size_t TSZ = (size_t)TILE * TILE;
        float* const out0 = reinterpret_cast<float*>(a.partials) +
                            (((size_t)unit * a.nsplit + split) * a.ntiles) * 2 * TSZ + (im ? TSZ : 0);
        // the list's SIMD slot, opaque to the compiler so that the block
        // coordinates below are derived per item instead of being hoisted
        // out of the item loop (144 store offsets held across it spill)
        int qs = q;
        asm volatile("" : "+v"(qs));
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            int br, bc, second;
            if constexpr (tri) {
                // absolute 32-row blocks -> their 192-tile (0, 0), (0, 1) or (1, 1)
                const bool in = NBKT == 11 ? tri_block(dq - 28, k, br, bc)
                                           : tri_run_block(NBKT, (dq - 28) >> 2, qs, k, br, bc);
                if (!in || br * 32 >= a.nant || bc * 32 >= a.nant) continue;
                float* out = out0 + (size_t)(br < 6 ? (bc < 6 ? 0 : 1) : 2) * 2 * TSZ;
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int row = (br % 6) * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                    const int col = (bc % 6) * 32 + (lane & 31);
                    out[(size_t)row * TILE + col] = acc[k][e];
                }
                continue;
            }
            if (!t192_block(dq, q, k, br, bc, second)) continue;
            // a paired item's blocks: tile 0 = (0, 0) or tile ntiles - 1 = (nb - 1, nb - 1)
            const int ot = pair ? second * (a.ntiles - 1) : tile;
            const int ob0 = pair ? second * (a.nb - 1) : bi, ob1 = pair ? second * (a.nb - 1) : bj;
            if (ob0 * TILE + br * 32 >= a.nant || ob1 * TILE + bc * 32 >= a.nant) continue;
            float* out = out0 + (size_t)ot * 2 * TSZ;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int row = br * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                const int col = bc * 32 + (lane & 31);
                out[(size_t)row * TILE + col] = acc[k][e];
            }
        }
    }
}

#undef HV_LDS

// ---------------------------------------------------------------------------
// Items of U frequency units of one time for arrays of 33..128 antennas
// (NBK = 2..4 blocks of 32 rows, U = 6, 4, 3: 384 operand rows per stage).
// The units share the time's amplitude rows and source directions, so one
// input DMA and one barrier per stage serve all of them, and a SIMD's list
// holds up to nine blocks as at HERA-350 (a 128-antenna tile alone has 3,
// 3, 2, 2).  Operand row v is antenna v mod 32 NBK of unit v / (32 NBK) of
// the item; the MFMA chain of every block is that of the other kernels, so
// the partial sums are theirs bit for bit.  Partials: the unit's TILE x
// TILE tile (TILE = 64 for NBK = 2, 128 for 3 and 4).
template <int MODE, bool IM, int Q, int NBK, int U>
__device__ __forceinline__ void mu_split(lds_char* in0, lds_char* in1, lds_char* op0, lds_char* op1,
                                         const double* antpos, int nant, const double* inv_lambda, int f0, int nu,
                                         const float* amp_t, const double* crd_t, size_t amp_stage, int s_begin,
                                         int n_st, const int* ampoff, int crdoff, int lane, int wave,
                                         f32x16 (&acc)[9]) {
    constexpr int R32 = NBK * 32;
    constexpr int G = U * R32 / 16;                      // 16-row groups of the item
    constexpr int W = (IM ? 4 : 0) + Q;
    constexpr int G0 = W * G / 8, G1 = (W + 1) * G / 8;
    constexpr int NR = G1 - G0;
    static_assert(NR >= 1 && NR <= 3, "one to three 16-row groups per wave");
    const int row0 = 16 * G0;
    // antenna positions (wavelengths of the row's unit) and amplitude rows
    double ax[3], ay[3], az[3];
    int arow[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int kk = k < NR ? k : 0;
        const int v = row0 + (lane & 15) + 16 * kk;
        const int u = v / R32, r = v - u * R32;
        const int ant = min(r, nant - 1);
        const double il = inv_lambda[f0 + min(u, nu - 1)];
        ax[k] = antpos[3 * ant] * il;
        ay[k] = antpos[3 * ant + 1] * il;
        az[k] = antpos[3 * ant + 2] * il;
        arow[k] = (row0 + 16 * kk) % R32;
    }
    const T192Frag fo(lane);
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[k][e] = 0;
#define HV_MUISSUE(slot, st)                                                                                  \
    t192_issue<R32>(slot, amp_t + (size_t)(s_begin + (st)) * amp_stage, crd_t + (size_t)(s_begin + (st)) * 16, \
                    ampoff, crdoff, wave, lane)
    HV_MUISSUE(in0, 0);
    wait_vmcnt<0>();
    __syncthreads();                       // stage 0's inputs are in LDS
    HV_MUISSUE(in1, min(1, n_st - 1));
    t192_build_mu<MODE, NR>(in0, op0, row0, lane, arow, ax, ay, az);
    for (int st = 0; st < n_st; ++st) {
        const bool odd = st & 1;
        const lds_char* cur = odd ? op1 : op0;
        wait_vm_lgkm<0, 0>();
        lds_barrier();
        if (st + 2 < n_st) HV_MUISSUE(odd ? in1 : in0, st + 2);
        // the first two fragments of this stage's MFMAs go out before the
        // build (N = 128, 128 units: 27.0-27.5 against 27.2-28.0 ms)
        constexpr TRun run0 = mu_run(NBK, U, Q, 0);
        TriPre pre;
        if constexpr (run0.r >= 0) pre = runs_pre<IM>(cur, fo, run0.r, run0.c0);
        __builtin_amdgcn_sched_barrier(0);
        t192_build_mu<MODE, NR>(odd ? in0 : in1, odd ? op0 : op1, row0, lane, arow, ax, ay, az);
        __builtin_amdgcn_sched_barrier(0);
        mu_runs_mma<MODE, IM, NBK, U, Q, true>(cur, fo, acc, &pre);
    }
#undef HV_MUISSUE
    wait_vm_lgkm<0, 0>();
    lds_barrier();
}

template <int MODE, int NBK, int U>
__global__ __launch_bounds__(512) void herk_mu_kernel(FusedArgs a) {
    using L = T192;
    constexpr int R32 = NBK * 32;
    constexpr int PT = NBK <= 2 ? 64 : 128;
    static_assert(U * R32 <= L::ROWS, "an item's operand rows fit the operand buffer");
    __shared__ __attribute__((aligned(1024))) char s_in0[L::IN];
    __shared__ __attribute__((aligned(1024))) char s_in1[L::IN];
    __shared__ __attribute__((aligned(1024))) char s_op0[2 * L::OP];
    __shared__ __attribute__((aligned(1024))) char s_op1[2 * L::OP];
    lds_char* in0 = (lds_char*)s_in0;
    lds_char* in1 = (lds_char*)s_in1;
    lds_char* op0 = (lds_char*)s_op0;
    lds_char* op1 = (lds_char*)s_op1;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int q = wave & 3;
    const bool im = wave >= 4;
    const int kst = (int)(eff_kpad(a.kpad, a.dyn_kpad) / 16);
    const size_t amp_stage = (size_t)a.nant * 16;
    const size_t amp_time = (size_t)a.kpad * a.nant;
    const int st_split = a.ksplit / 16;   // stages per K-split
    const int nt = a.nunits / a.nfreq;                   // times of the launch
    const int gpt = (a.nfreq + U - 1) / U;               // unit groups per time
    const int ngroups = nt * gpt;
    const int ug = t192_unit_group(ngroups, gridDim.x);
    const int64_t nitems = t192_items(ngroups, a.nsplit, 1, ug);
    // amplitude chunk c = (i*8 + wave)*64 + lane: antenna row c/4 (< 32 NBK)
    int ampoff[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int c = (i * 8 + wave) * 64 + lane;
        const int rr = c >> 2;
        ampoff[i] = min(rr, a.nant - 1) * 16 + ((c & 3) ^ amp_swz(rr)) * 4;
    }
    const int crdoff = min(wave, 2) * (int)a.kpad + 2 * lane;   // used by lanes 0-7 of waves 0-2

    for (int64_t Lw = blockIdx.x; Lw < nitems; Lw += gridDim.x) {
        int grp, split, tile;
        if (!t192_decode(ngroups, a.nsplit, 1, ug, Lw, grp, split, tile)) continue;
        const int s_begin = split * st_split;
        const int s_end = min(s_begin + st_split, kst);
        if (s_begin >= s_end) continue;   // split beyond the job's sources (uniform per workgroup)
        const int n_st = s_end - s_begin;
        const int tb = grp / gpt, f0 = (grp - tb * gpt) * U;
        const int nu = min(U, a.nfreq - f0);
        const float* amp_t = a.amp + (size_t)tb * amp_time;
        const double* crd_t = a.crd_top + (size_t)tb * 3 * a.kpad;

        f32x16 acc[9];
#define HV_MUSPLIT(IM, Q)                                                                                      \
    mu_split<MODE, IM, Q, NBK, U>(in0, in1, op0, op1, a.antpos, a.nant, a.inv_lambda, f0, nu, amp_t, crd_t,      \
                                  amp_stage, s_begin, n_st, ampoff, crdoff, lane, wave, acc)
        if (im) {
            if (q == 0) HV_MUSPLIT(true, 0);
            else if (q == 1) HV_MUSPLIT(true, 1);
            else if (q == 2) HV_MUSPLIT(true, 2);
            else HV_MUSPLIT(true, 3);
        } else {
            if (q == 0) HV_MUSPLIT(false, 0);
            else if (q == 1) HV_MUSPLIT(false, 1);
            else if (q == 2) HV_MUSPLIT(false, 2);
            else HV_MUSPLIT(false, 3);
        }
#undef HV_MUSPLIT

        constexpr size_t TSZ = (size_t)PT * PT;
        // the SIMD slot, opaque to the compiler so that the block coordinates
        // are derived per item instead of being hoisted out of the item loop
        int qs = q;
        asm volatile("" : "+v"(qs));
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            int br = 0, bc = 0;
            if (!mu_run_block(NBK, U, qs, k, br, bc)) continue;
            const int u = br / NBK, r = br - u * NBK, c = bc - u * NBK;
            if (u >= nu || r * 32 >= a.nant || c * 32 >= a.nant) continue;
            const int unit = tb * a.nfreq + f0 + u;
            float* out = reinterpret_cast<float*>(a.partials) + ((size_t)unit * a.nsplit + split) * 2 * TSZ +
                         (im ? TSZ : 0);
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int row = r * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                const int col = c * 32 + (lane & 31);
                out[(size_t)row * PT + col] = acc[k][e];
            }
        }
    }
}

// Resident workgroups of a fused kernel on the whole device (a multiple of 8).
template <int MODE, int TILE>
static int64_t fused_resident() {
    static const int64_t cached = [] {
        int dev = 0, ncu = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, herk_fused_kernel<MODE, TILE>, 256, 0) !=
                hipSuccess)
            return (int64_t)0;
        return (int64_t)ncu * per_cu / 8 * 8;
    }();
    return cached;
}

template <int MODE>
static int64_t fused8_resident() {
    static const int64_t cached = [] {
        int dev = 0, ncu = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, herk_fused8_kernel<MODE>, 512, 0) != hipSuccess)
            return (int64_t)0;
        return (int64_t)ncu * per_cu / 8 * 8;
    }();
    return cached;
}

template <int MODE>
static hipError_t launch_fused8_t(const FusedArgs& a, hipStream_t s) {
    const int64_t nblk = (int64_t)a.nunits * a.nsplit * a.ntiles;
    if (nblk <= 0) return hipSuccess;
    if (a.kpad % 16) return hipErrorInvalidValue;
    const int64_t res = fused8_resident<MODE>();
    const int64_t grid = (res > 0 && res < nblk) ? res : nblk;
    hipLaunchKernelGGL((herk_fused8_kernel<MODE>), dim3((unsigned)grid), dim3(512), 0, s, a);
    return hipGetLastError();
}

template <int MODE, int NBKT>
static int64_t t192_resident() {
    static const int64_t cached = [] {
        int dev = 0, ncu = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, herk_t192_kernel<MODE, NBKT>, 512, 0) != hipSuccess)
            return (int64_t)0;
        return (int64_t)ncu * per_cu / 8 * 8;
    }();
    return cached;
}

template <int MODE, int NBKT>
static hipError_t launch_t192_nbk(const FusedArgs& a, hipStream_t s) {
    const int64_t res = t192_resident<MODE, NBKT>();
    const int64_t g0 = res > 0 ? res : 256;
    const int ntw = NBKT > 0 ? tri_items(NBKT) : t192_work_tiles(a.nb, a.nant, a.ntiles);
    const int64_t items = t192_items(a.nunits, a.nsplit, ntw, t192_unit_group(a.nunits, g0));
    const int64_t grid = g0 < items ? g0 : items;
    hipLaunchKernelGGL((herk_t192_kernel<MODE, NBKT>), dim3((unsigned)grid), dim3(512), 0, s, a);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_t192_t(const FusedArgs& a0, hipStream_t s) {
    if ((int64_t)a0.nunits * a0.nsplit * a0.ntiles <= 0) return hipSuccess;
    if (a0.kpad % 16) return hipErrorInvalidValue;
    FusedArgs a = a0;
    if (g_pf_override == 6) a.lists = 1;
    // the triangle decomposition stores into 192-tiles of one or two blocks
    const int nbk = (a.nb <= 2) ? t192_tri_nbk(a.nant, a.lists) : 0;
    switch (nbk) {
        case 5: return launch_t192_nbk<MODE, 5>(a, s);
        case 6: return launch_t192_nbk<MODE, 6>(a, s);
        case 7: return launch_t192_nbk<MODE, 7>(a, s);
        case 8: return launch_t192_nbk<MODE, 8>(a, s);
        case 9: return launch_t192_nbk<MODE, 9>(a, s);
        case 10: return launch_t192_nbk<MODE, 10>(a, s);
        case 11: return launch_t192_nbk<MODE, 11>(a, s);
        default: return launch_t192_nbk<MODE, 0>(a, s);
    }
}

template <int MODE, int NBK, int U>
static int64_t mu_resident() {
    static const int64_t cached = [] {
        int dev = 0, ncu = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, herk_mu_kernel<MODE, NBK, U>, 512, 0) !=
                hipSuccess)
            return (int64_t)0;
        return (int64_t)ncu * per_cu / 8 * 8;
    }();
    return cached;
}

template <int MODE, int NBK, int U>
static hipError_t launch_mu_t(const FusedArgs& a, hipStream_t s) {
    if (a.nfreq < 1 || a.nunits % a.nfreq) return hipErrorInvalidValue;
    const int64_t res = mu_resident<MODE, NBK, U>();
    const int64_t g0 = res > 0 ? res : 256;
    const int ngroups = a.nunits / a.nfreq * ((a.nfreq + U - 1) / U);
    const int64_t items = t192_items(ngroups, a.nsplit, 1, t192_unit_group(ngroups, g0));
    const int64_t grid = g0 < items ? g0 : items;
    hipLaunchKernelGGL((herk_mu_kernel<MODE, NBK, U>), dim3((unsigned)grid), dim3(512), 0, s, a);
    return hipGetLastError();
}

// Items of several units for split modes at 33..128 antennas, when the plan's
// tile is the one the items store into and every unit of a time shares one
// A sqrt(I) (one beam); pf 6 keeps the 64- and 128-tile kernels.
template <int MODE>
static bool launch_mu(int tile, const FusedArgs& a, hipStream_t s, hipError_t& err) {
    const int nbk = (a.nant + 31) / 32;
    if (a.amp_per_unit || a.lists == 1 || g_pf_override == 6 || nbk < 2 || nbk > 4) return false;
    // an item computes all U unit slots: a job with fewer frequencies per
    // time than U takes the tile kernels (one unit per item) instead
    if ((a.job_nfreq > 0 ? a.job_nfreq : a.nfreq) < herk_mu_units(a.nant)) return false;
    if (tile != (nbk <= 2 ? 64 : 128) || a.kpad % 16) return false;
    err = nbk == 2 ? launch_mu_t<MODE, 2, 6>(a, s) : nbk == 3 ? launch_mu_t<MODE, 3, 4>(a, s)
                                                   : launch_mu_t<MODE, 4, 3>(a, s);
    return true;
}

template <int MODE, int TILE>
static hipError_t launch_fused_t(const FusedArgs& a, hipStream_t s) {
    const int64_t nblk = (int64_t)a.nunits * a.nsplit * a.ntiles;
    if (nblk <= 0) return hipSuccess;
    if (a.kpad % 16) return hipErrorInvalidValue;
    const int64_t res = fused_resident<MODE, TILE>();
    const int64_t grid = (res > 0 && res < nblk) ? res : nblk;
    hipLaunchKernelGGL((herk_fused_kernel<MODE, TILE>), dim3((unsigned)grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fused(int mode, int tile, const FusedArgs& a, hipStream_t s) {
    if ((int64_t)a.nunits * a.nsplit * a.ntiles <= 0) return hipSuccess;
    if (!ksplit_valid(a.ksplit)) return hipErrorInvalidValue;
    hipError_t err = hipSuccess;
    if (mode == MODE_SPLIT_BF16 && launch_mu<MODE_SPLIT_BF16>(tile, a, s, err)) return err;
    if (mode == MODE_SPLIT_FP16 && launch_mu<MODE_SPLIT_FP16>(tile, a, s, err)) return err;
    if (mode == MODE_SPLIT_BF16 && tile == 64) return launch_fused_t<MODE_SPLIT_BF16, 64>(a, s);
    if (mode == MODE_SPLIT_BF16 && tile == 128) return launch_fused8_t<MODE_SPLIT_BF16>(a, s);
    if (mode == MODE_SPLIT_FP16 && tile == 64) return launch_fused_t<MODE_SPLIT_FP16, 64>(a, s);
    if (mode == MODE_SPLIT_FP16 && tile == 128) return launch_fused8_t<MODE_SPLIT_FP16>(a, s);
    if (mode == MODE_SPLIT_BF16 && tile == 192) return launch_t192_t<MODE_SPLIT_BF16>(a, s);
    if (mode == MODE_SPLIT_FP16 && tile == 192) return launch_t192_t<MODE_SPLIT_FP16>(a, s);
    return hipErrorInvalidValue;
}

template <int MODE, int TILE, int PF>
static hipError_t launch_herk_t(const HerkArgs& a, hipStream_t s) {
    const int64_t nblk = (int64_t)a.nunits * a.nsplit * a.ntiles;
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL((herk_kernel<MODE, TILE, PF>), dim3((unsigned)nblk), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Kernel variants: tile edge (64 / 128) x prefetch depth (1 / 2 stages).
// variant 0 selects the default for the mode and antenna count.
hipError_t launch_herk_variant(int mode, int tile, int pf, const HerkArgs& a, hipStream_t s) {
    if (!ksplit_valid(a.ksplit)) return hipErrorInvalidValue;
#define HV_V(M, T, P) if (mode == M && tile == T && pf == P) return launch_herk_t<M, T, P>(a, s)
    HV_V(MODE_FP32, 64, 1); HV_V(MODE_FP32, 64, 2);
    HV_V(MODE_SPLIT_BF16, 64, 1); HV_V(MODE_SPLIT_BF16, 64, 2);
    HV_V(MODE_SPLIT_BF16, 128, 1); HV_V(MODE_SPLIT_BF16, 128, 2);
    HV_V(MODE_SPLIT_FP16, 64, 1); HV_V(MODE_SPLIT_FP16, 64, 2);
    HV_V(MODE_SPLIT_FP16, 128, 1); HV_V(MODE_SPLIT_FP16, 128, 2);
    HV_V(MODE_FP64, 64, 1);
#undef HV_V
    // pf = 3: ring variant (direct-to-LDS loads, 3 stages ahead), one
    // workgroup per work item; pf = 4: the same on a persistent grid
    if (pf == 3 || pf == 4) {
        const bool p = (pf == 4);
        if (mode == MODE_FP32 && tile == 64) return launch_herk_ring_t<MODE_FP32, 64>(a, p, s);
        if (mode == MODE_SPLIT_BF16 && tile == 64) return launch_herk_ring_t<MODE_SPLIT_BF16, 64>(a, p, s);
        if (mode == MODE_SPLIT_BF16 && tile == 128) return launch_herk_ring_t<MODE_SPLIT_BF16, 128>(a, p, s);
        if (mode == MODE_SPLIT_FP16 && tile == 64) return launch_herk_ring_t<MODE_SPLIT_FP16, 64>(a, p, s);
        if (mode == MODE_SPLIT_FP16 && tile == 128) return launch_herk_ring_t<MODE_SPLIT_FP16, 128>(a, p, s);
        if (mode == MODE_FP64 && tile == 64) return launch_herk_ring_t<MODE_FP64, 64>(a, p, s);
    }
    return hipErrorInvalidValue;
}

void herk_init() {
    (void)ring_resident<MODE_FP32, 64>();
    (void)ring_resident<MODE_SPLIT_BF16, 64>();
    (void)ring_resident<MODE_SPLIT_BF16, 128>();
    (void)ring_resident<MODE_SPLIT_FP16, 64>();
    (void)ring_resident<MODE_SPLIT_FP16, 128>();
    (void)ring_resident<MODE_FP64, 64>();
    (void)fused_resident<MODE_SPLIT_BF16, 64>();
    (void)fused8_resident<MODE_SPLIT_BF16>();
    (void)fused_resident<MODE_SPLIT_FP16, 64>();
    (void)fused8_resident<MODE_SPLIT_FP16>();
    (void)t192_resident<MODE_SPLIT_BF16, 0>();
    (void)t192_resident<MODE_SPLIT_FP16, 0>();
    (void)t192_resident<MODE_SPLIT_BF16, 11>();
    (void)t192_resident<MODE_SPLIT_FP16, 11>();
}

int herk_prefetch(int mode, int tile) {
    (void)tile;
    return pf_choice(mode);
}

// Contraction of operands already in HBM (the built-in variant, or with
// pf 5, which builds operands in place, the LDS ring on a persistent grid).
// Tile edge for contracting operands already in device memory (LDS-ring and
// register-staged kernels): the in-place choice, with 192 (in place only)
// taken as 128.
int herk_tile_operands(int mode, int nant) {
    const int t = herk_tile(mode, nant);
    return t == 192 ? 128 : t;
}

hipError_t launch_herk(int mode, int tile, const HerkArgs& a, hipStream_t s) {
    const int pf = herk_prefetch(mode, tile);
    return launch_herk_variant(mode, tile, pf == 5 ? 4 : pf, a, s);
}

bool herk_fused(int mode, int nant) {
    return split_mode(mode) && herk_prefetch(mode, herk_tile(mode, nant)) == 5;
}

int herk_mu_units(int nant) {
    const int nbk = (nant + 31) / 32;
    return nbk == 2 ? 6 : nbk == 3 ? 4 : nbk == 4 ? 3 : 0;
}

static int64_t g_ksplit_items = KSPLIT_LONG_ITEMS;
void set_ksplit_items(int64_t items) { g_ksplit_items = items > 0 ? items : KSPLIT_LONG_ITEMS; }

int job_ksplit(int mode, int nant, int64_t units, int64_t kpad) {
    if (!split_mode(mode)) return KSPLIT;
    // work items per (unit, split) of the built-in kernels, times 12: a
    // multi-unit item holds 6, 4 or 3 units at 2, 3 or 4 blocks of 32 rows;
    // 5..8 blocks run one triangle item, 9..11 a triangle and a strip item;
    // otherwise one item per tile (192-tiles beyond 352 antennas, 64-tiles
    // up to 32).  Variant overrides do not enter.
    const int nbk = (nant + 31) / 32;
    int64_t ipu12;
    if (nbk >= 2 && nbk <= 4) {
        ipu12 = nbk == 2 ? 2 : nbk == 3 ? 3 : 4;
    } else if (nbk >= 5 && nbk <= 11) {
        ipu12 = nbk <= 8 ? 12 : 24;
    } else {
        const int tile = nbk > 11 ? 192 : 64;
        const int64_t nb = (nant + tile - 1) / tile;
        ipu12 = 12 * (nb * (nb + 1) / 2);
    }
    const int64_t nlong = (kpad + KSPLIT_LONG - 1) / KSPLIT_LONG;
    return units * ipu12 * nlong >= 12 * g_ksplit_items ? KSPLIT_LONG : KSPLIT;
}

}  // namespace hv
