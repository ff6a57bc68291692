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
// FP32 accumulators are folded into a second FP32 register total every 512
// sources (short MFMA accumulation chains); K is cut into fixed 8192-source
// splits whose partial tiles are summed in FP64, in split order, by the
// reduce stage, so results do not depend on how jobs are padded or batched.
//
// Workgroup: 4 waves (2 x 2) on a TILE x TILE complex tile; a K stage is
// 128 bytes of every row (16 complex64 sources, 8 complex128), staged
// through LDS with a padded 144-byte row stride (conflict-free 16-byte
// reads) and double buffering.  Workgroups are ordered so the tiles of one
// K split run back to back on one XCD (shared L2 lines).
#include "hv_internal.h"

namespace hv {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int LDS_STRIDE = 144;

int herk_tile(int /*mode*/) { return 64; }

// (x, y) pairs of 16-bit values -> (y, -x): rotate each 32-bit word by 16 and
// flip the sign bit of the upper half.
__device__ inline u32x4 rot_neg16(u32x4 v) {
    u32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = __builtin_amdgcn_alignbit(v[i], v[i], 16) ^ 0x80000000u;
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
__device__ inline void decode_work(const HerkArgs& a, int64_t L, int& unit, int& split, int& tile) {
    const int64_t per = (int64_t)a.nsplit * a.ntiles;
    unit = (int)(L / per);
    const int64_t rem = L - (int64_t)unit * per;
    if ((a.nsplit & 7) == 0) {
        const int x = (int)(rem & 7);
        const int64_t p = rem >> 3;
        split = x + 8 * (int)(p / a.ntiles);
        tile = (int)(p % a.ntiles);
    } else {
        split = (int)(rem / a.ntiles);
        tile = (int)(rem % a.ntiles);
    }
}

template <int MODE, int TILE>
__global__ __launch_bounds__(256) void herk_kernel(HerkArgs a) {
    constexpr bool F64 = (MODE == MODE_FP64);
    constexpr int BLK = F64 ? 16 : 32;            // MFMA block edge
    constexpr int WT = TILE / 2;                  // wave tile edge (complex)
    constexpr int NS = WT / BLK;                  // subtiles per wave per dim
    constexpr int NCH = TILE * (STAGE_BYTES / 16) / 256;   // 16-B chunks / thread / operand
    constexpr int SRC_STAGE = stage_src(F64);
    constexpr int STAGES_SPLIT = KSPLIT / SRC_STAGE;
    constexpr int FOLD = FOLD_SRC / SRC_STAGE;
    using ACC = typename std::conditional<F64, f64x4, f32x16>::type;
    using AR = typename std::conditional<F64, double, float>::type;
    constexpr int NACC = F64 ? 4 : 16;

    __shared__ __attribute__((aligned(16))) char lds[2][2][TILE * LDS_STRIDE];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;

    int unit, split, tile, bi, bj;
    decode_work(a, blockIdx.x, unit, split, tile);
    decode_tile(tile, a.nb, bi, bj);
    const bool diag = (bi == bj);

    const int kst = (int)(eff_kpad(a.kpad, a.dyn_kpad) / SRC_STAGE);
    const int s_begin = split * STAGES_SPLIT;
    const int s_end = min(s_begin + STAGES_SPLIT, kst);
    if (s_begin >= s_end) return;   // split beyond the job's sources (uniform per workgroup)

    const char* Wu = reinterpret_cast<const char*>(a.W) + (size_t)unit * a.unit_stride;

    // ---- loader addressing: a stage of TILE antennas is one contiguous
    // TILE*128-byte block (stage-major operands); thread chunk c covers bytes
    // [16 idx, 16 idx + 16) of it, i.e. row idx/8, 16-byte column idx%8
    int roff[NCH];
    int loff[NCH];
    bool vA[NCH], vB[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int idx = tid + 256 * c;
        const int r = idx >> 3, c16 = idx & 7;
        vA[c] = bi * TILE + r < a.nant;
        vB[c] = bj * TILE + r < a.nant;
        roff[c] = idx * 16;
        loff[c] = r * LDS_STRIDE + c16 * 16;
    }
    const size_t stage_bytes = (size_t)a.nant * STAGE_BYTES;
    const char* gA0 = Wu + (size_t)bi * TILE * STAGE_BYTES;
    const char* gB0 = Wu + (size_t)bj * TILE * STAGE_BYTES;
    uint4 rA[NCH], rB[NCH];
    const uint4 zero4 = make_uint4(0, 0, 0, 0);
    auto gload = [&](int st) {
        const size_t o = (size_t)st * stage_bytes;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            rA[c] = vA[c] ? *reinterpret_cast<const uint4*>(gA0 + o + roff[c]) : zero4;
            if (!diag) rB[c] = vB[c] ? *reinterpret_cast<const uint4*>(gB0 + o + roff[c]) : zero4;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            *reinterpret_cast<uint4*>(&lds[buf][0][loff[c]]) = rA[c];
            if (!diag) *reinterpret_cast<uint4*>(&lds[buf][1][loff[c]]) = rB[c];
        }
    };

    // Re chain; Im either as one chain of w'.w (split modes) or as P1 - P2
    // with P1 = sum y_a x_b and P2 = sum x_a y_b (FP32, FP64), which cancel
    // exactly when rows a and b are equal.
    constexpr bool PAIR_IM = (MODE == MODE_FP32 || MODE == MODE_FP64);
    ACC re[NS][NS], im[NS][NS], p1[NS][NS], p2[NS][NS];
    ACC tre[NS][NS], tim[NS][NS];
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j)
#pragma unroll
            for (int e = 0; e < NACC; ++e) {
                re[i][j][e] = 0; im[i][j][e] = 0; p1[i][j][e] = 0; p2[i][j][e] = 0;
                tre[i][j][e] = 0; tim[i][j][e] = 0;
            }

    // A wave whose 32x32 (complex) region lies strictly below the diagonal of
    // a diagonal tile, or wholly in the antenna padding, skips its MFMAs and
    // its partial write (the reduce stage reads the mirrored entry instead);
    // it still helps stage operands.  Co-resident workgroups use the idle
    // SIMD time.
    const bool active = (bi * TILE + wr * WT < a.nant) && (bj * TILE + wc * WT < a.nant) && !(diag && wr > wc);

    // lane -> (row within block, k slot group)
    const int lrow = F64 ? (lane & 15) : (lane & 31);
    const int lh = F64 ? (lane >> 4) : (lane >> 5);

    gload(s_begin);
    sstore(0);
    int since_fold = 0;
    for (int st = s_begin; st < s_end; ++st) {
        const int buf = (st - s_begin) & 1;
        __syncthreads();
        if (st + 1 < s_end) gload(st + 1);
        const char* As = lds[buf][0];
        const char* Bs = diag ? lds[buf][0] : lds[buf][1];

        if (!active) {
            // nothing to compute for this wave
        } else if constexpr (MODE == MODE_FP32) {
            // lane half lh holds (x_s, y_s, x_s+1, y_s+1), s = 4q + 2 lh
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                f32x4 av[NS], bv[NS];
#pragma unroll
                for (int i = 0; i < NS; ++i) {
                    av[i] = *reinterpret_cast<const f32x4*>(As + (wr * WT + i * BLK + lrow) * LDS_STRIDE + q * 32 + lh * 16);
                    bv[i] = *reinterpret_cast<const f32x4*>(Bs + (wc * WT + i * BLK + lrow) * LDS_STRIDE + q * 32 + lh * 16);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int i = 0; i < NS; ++i)
#pragma unroll
                        for (int j = 0; j < NS; ++j)
                            re[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i][e], bv[j][e], re[i][j], 0, 0, 0);
#pragma unroll
                for (int e = 0; e < 4; e += 2)
#pragma unroll
                    for (int i = 0; i < NS; ++i)
#pragma unroll
                        for (int j = 0; j < NS; ++j) {
                            p1[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i][e + 1], bv[j][e], p1[i][j], 0, 0, 0);
                            p2[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i][e], bv[j][e + 1], p2[i][j], 0, 0, 0);
                        }
            }
        } else if constexpr (MODE == MODE_SPLIT_BF16 || MODE == MODE_SPLIT_FP16) {
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                const int g = 2 * jj + lh;
                u32x4 ahi[NS], alo[NS], phi[NS], plo[NS], bhi[NS], blo[NS];
#pragma unroll
                for (int i = 0; i < NS; ++i) {
                    const char* pa = As + (wr * WT + i * BLK + lrow) * LDS_STRIDE + g * 32;
                    ahi[i] = *reinterpret_cast<const u32x4*>(pa);
                    alo[i] = *reinterpret_cast<const u32x4*>(pa + 16);
                    phi[i] = rot_neg16(ahi[i]);
                    plo[i] = rot_neg16(alo[i]);
                    const char* pb = Bs + (wc * WT + i * BLK + lrow) * LDS_STRIDE + g * 32;
                    bhi[i] = *reinterpret_cast<const u32x4*>(pb);
                    blo[i] = *reinterpret_cast<const u32x4*>(pb + 16);
                }
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int j = 0; j < NS; ++j) {
                        re[i][j] = mfma16<MODE>(ahi[i], bhi[j], re[i][j]);
                        im[i][j] = mfma16<MODE>(phi[i], bhi[j], im[i][j]);
                    }
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int j = 0; j < NS; ++j) {
                        re[i][j] = mfma16<MODE>(ahi[i], blo[j], re[i][j]);
                        im[i][j] = mfma16<MODE>(phi[i], blo[j], im[i][j]);
                    }
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int j = 0; j < NS; ++j) {
                        re[i][j] = mfma16<MODE>(alo[i], bhi[j], re[i][j]);
                        im[i][j] = mfma16<MODE>(plo[i], bhi[j], im[i][j]);
                    }
            }
        } else {  // FP64: lane quarter lh holds (x_s, y_s, x_s+1, y_s+1), s = 2 lh
            f64x4 av[NS], bv[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                av[i] = *reinterpret_cast<const f64x4*>(As + (wr * WT + i * BLK + lrow) * LDS_STRIDE + lh * 32);
                bv[i] = *reinterpret_cast<const f64x4*>(Bs + (wc * WT + i * BLK + lrow) * LDS_STRIDE + lh * 32);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int j = 0; j < NS; ++j)
                        re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i][e], bv[j][e], re[i][j], 0, 0, 0);
#pragma unroll
            for (int e = 0; e < 4; e += 2)
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int j = 0; j < NS; ++j) {
                        p1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i][e + 1], bv[j][e], p1[i][j], 0, 0, 0);
                        p2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i][e], bv[j][e + 1], p2[i][j], 0, 0, 0);
                    }
        }

        if (st + 1 < s_end) sstore(buf ^ 1);
        if constexpr (!F64) {
            if (++since_fold == FOLD) {
                since_fold = 0;
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int j = 0; j < NS; ++j) {
                        tre[i][j] += re[i][j];
                        if constexpr (PAIR_IM) tim[i][j] += p1[i][j] - p2[i][j];
                        else tim[i][j] += im[i][j];
#pragma unroll
                        for (int e = 0; e < NACC; ++e) { re[i][j][e] = 0; im[i][j][e] = 0; p1[i][j][e] = 0; p2[i][j][e] = 0; }
                    }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            tre[i][j] += re[i][j];
            if constexpr (PAIR_IM) tim[i][j] += p1[i][j] - p2[i][j];
            else tim[i][j] += im[i][j];
        }

    if (!active) return;
    // ---- write the (Re, Im) partial tile
    constexpr size_t TSZ = (size_t)TILE * TILE;
    AR* out = reinterpret_cast<AR*>(a.partials) +
              (((size_t)unit * a.nsplit + split) * a.ntiles + tile) * 2 * TSZ;
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j)
#pragma unroll
            for (int e = 0; e < NACC; ++e) {
                int r, c;
                if constexpr (F64) {           // 16x16 f64 block: col = lane%16, row = lane/16 + 4e
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

template <int MODE, int TILE>
static hipError_t launch_herk_t(const HerkArgs& a, hipStream_t s) {
    const int64_t nblk = (int64_t)a.nunits * a.nsplit * a.ntiles;
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL((herk_kernel<MODE, TILE>), dim3((unsigned)nblk), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_herk(int mode, const HerkArgs& a, hipStream_t s) {
    switch (mode) {
        case MODE_FP32: return launch_herk_t<MODE_FP32, 64>(a, s);
        case MODE_SPLIT_BF16: return launch_herk_t<MODE_SPLIT_BF16, 64>(a, s);
        case MODE_SPLIT_FP16: return launch_herk_t<MODE_SPLIT_FP16, 64>(a, s);
        case MODE_FP64: return launch_herk_t<MODE_FP64, 64>(a, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace hv
