// Memory-bound stages of the visibility pipeline: rotation, beam, phasor
// build, operand split for caller phasors, and the K-split reduction.
//
// Reference semantics (hera_gpu/vis.py):
//   rotation   crd_top = eq2top . crd_eq                       vis.py:228-232
//   beam       bilinear at texel (P-1)(0.5 x + 0.5), clamp,   vis.py:46-89
//              A = 0 unless z > 0
//   delay      tau = antpos . crd_top / c                      vis.py:233-236
//   phasor     v = A sqrt(I) exp(+i 2 pi nu tau)               vis.py:92-110,159
// Here the delay is never materialised: each phasor thread forms the path
// difference d = antpos . crd_top in FP64, converts it to quarter turns with
// 4 nu / c, keeps only the fractional part (exact in FP64 for |turns| < 2^49)
// and the quadrant, and evaluates sin/cos of that small angle.
#include <algorithm>

#include "hv_internal.h"
#include "hv_phasor_math.h"

namespace hv {

// Write 8 consecutive phasors (one thread's share) of one antenna row.
template <int MODE, typename R>
__device__ inline void store_phasors8(char* dst, const R* re, const R* im) {
    if constexpr (MODE == MODE_FP64) {
        double2* d = reinterpret_cast<double2*>(dst);
#pragma unroll
        for (int j = 0; j < 8; ++j) d[j] = make_double2(re[j], im[j]);
    } else if constexpr (MODE == MODE_FP32) {
        float4* d = reinterpret_cast<float4*>(dst);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            d[j] = make_float4(re[2 * j], im[2 * j], re[2 * j + 1], im[2 * j + 1]);
    } else {
        uint4* d = reinterpret_cast<uint4*>(dst);
        uint4 g[2];
        split_group<MODE>(re, im, g);
        d[0] = g[0];
        d[1] = g[1];
        split_group<MODE>(re + 4, im + 4, g);
        d[2] = g[0];
        d[3] = g[1];
    }
}

// ------------------------------------------------------------- rotation
// One thread per (time, source): crd_top[t][:, s] = eq2top[t] . crd_eq[:, s].
__global__ __launch_bounds__(256) void coords_kernel(CoordsArgs a) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int t = blockIdx.y;
    if (s >= eff_kpad(a.kpad, a.dyn_kpad)) return;
    const double* R = a.eq2tops + 9 * t;
    const double x = a.crd_eq[s], y = a.crd_eq[a.kpad + s], z = a.crd_eq[2 * a.kpad + s];
    double* o = a.crd_top + (size_t)t * 3 * a.kpad;
    o[s] = fma(R[0], x, fma(R[1], y, R[2] * z));
    o[a.kpad + s] = fma(R[3], x, fma(R[4], y, R[5] * z));
    o[2 * a.kpad + s] = fma(R[6], x, fma(R[7], y, R[8] * z));
}

hipError_t launch_coords(const CoordsArgs& a, hipStream_t s) {
    dim3 grid((unsigned)((a.kpad + 255) / 256), (unsigned)a.ntimes);
    hipLaunchKernelGGL(coords_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------ fused phasor build
// Grid (source chunks, antennas).  The antenna's beam is staged in LDS once
// per workgroup; each thread then handles 8 consecutive sources: beam
// amplitude A sqrt(I) (zero below the horizon) and FP64 path difference are
// formed once and reused for every frequency of the launch.
// LDS of one workgroup on gfx950; beam planes that do not fit are read from
// global memory instead (the same taps through L2).
constexpr size_t LDS_BYTES = (size_t)160 * 1024;
constexpr int PH_THREADS = 256;
constexpr int PH_SRC_PER_THREAD = 8;
constexpr int PH_SRC_PER_WG = 8192;

constexpr bool split_mode(int mode) { return mode == MODE_SPLIT_BF16 || mode == MODE_SPLIT_FP16; }

// Phasors of 8 consecutive sources of one antenna at one frequency written to
// their operand slots: ph holds the delay phases in turns (split modes) or
// the path differences in metres (FP32 / FP64: scaled by il4 here).
template <int MODE, typename R>
__device__ __forceinline__ void phasor_store(char* dst, const R* amp, const double* ph, double il4) {
    if constexpr (split_mode(MODE)) {
        // two 32-byte groups [4 hi words | 4 lo words]
        uint32_t hw[8], lw[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float sn, cs;
            sincos_turns_hw(ph[j], &sn, &cs);
            split_phasor<MODE>(amp[j], cs, sn, hw[j], lw[j]);
        }
        uint4* o = reinterpret_cast<uint4*>(dst);
        o[0] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
        o[1] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
        o[2] = make_uint4(hw[4], hw[5], hw[6], hw[7]);
        o[3] = make_uint4(lw[4], lw[5], lw[6], lw[7]);
    } else {
        R re[8], im[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            R sn, cs;
            sincos_quarter_turns(ph[j] * il4, &sn, &cs);
            re[j] = amp[j] * cs;
            im[j] = amp[j] * sn;
        }
        store_phasors8<MODE, R>(dst, re, im);
    }
}

// GB: the beam plane does not fit in LDS; the taps come from global memory.
template <int MODE, typename BT, bool GB>
__global__ __launch_bounds__(PH_THREADS) void phasor_kernel(PhasorArgs a) {
    using R = typename std::conditional<MODE == MODE_FP64, double, float>::type;
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const BT* sb = reinterpret_cast<const BT*>(smem);
    const int64_t kact = eff_kpad(a.kpad, a.dyn_kpad);
    // antenna is the fastest grid index: the workgroups of one source chunk
    // run side by side and share its coordinates in L2
    const int64_t wg0 = (int64_t)blockIdx.y * PH_SRC_PER_WG;
    if (wg0 >= kact) return;                       // uniform per workgroup
    const int ant = blockIdx.x;
    const int tb = blockIdx.z;
    const int P = a.beam_px;
    const BT* gb = reinterpret_cast<const BT*>(a.beam) + (size_t)ant * P * P;
    if constexpr (GB) {
        sb = gb;
    } else {
        BT* st = reinterpret_cast<BT*>(smem);
        for (int i = threadIdx.x; i < P * P; i += PH_THREADS) st[i] = gb[i];
        __syncthreads();
    }

    const double ax = a.antpos[3 * ant], ay = a.antpos[3 * ant + 1], az = a.antpos[3 * ant + 2];
    const double* cx = a.crd_top + (size_t)tb * 3 * a.kpad;
    const double* cy = cx + a.kpad;
    const double* cz = cx + 2 * a.kpad;
    const double* isq = a.isqrt;
    const int64_t wg1 = min(wg0 + (int64_t)PH_SRC_PER_WG, kact);
    char* wunit = reinterpret_cast<char*>(a.W) + (size_t)tb * a.nfreq * a.unit_stride;
    constexpr int SST = stage_src(MODE == MODE_FP64);
    constexpr size_t elem = (MODE == MODE_FP64) ? 16 : 8;

    if (a.beam_fstride) {
        // a beam per frequency: frequency-outer, each frequency's plane in
        // LDS in turn, the same arithmetic per element as below
        for (int u = 0; u < a.nfreq; ++u) {
            if (u > 0) {
                const BT* gu = reinterpret_cast<const BT*>(a.beam) + u * a.beam_fstride + (size_t)ant * P * P;
                if constexpr (GB) {
                    sb = gu;
                } else {
                    __syncthreads();                   // every thread is done with plane u-1
                    BT* st = reinterpret_cast<BT*>(smem);
                    for (int i = threadIdx.x; i < P * P; i += PH_THREADS) st[i] = gu[i];
                    __syncthreads();
                }
            }
            const double il = a.inv_lambda[u];
            const double il4 = 4.0 * il;
            const double axl = ax * il, ayl = ay * il, azl = az * il;
            for (int64_t s0 = wg0 + (int64_t)threadIdx.x * PH_SRC_PER_THREAD; s0 < wg1;
                 s0 += PH_THREADS * PH_SRC_PER_THREAD) {
                R amp[8];
                double ph[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int64_t s = s0 + j;
                    const double x = cx[s], y = cy[s], z = cz[s];
                    R A = (R)0;
                    if (z > 0.0) A = (R)beam_bilinear<BT>(sb, P, x, y);
                    amp[j] = A * (R)isq[s] * (R)a.amp_scale;
                    ph[j] = split_mode(MODE) ? phase_turns(axl, ayl, azl, x, y, z) : fma(ax, x, fma(ay, y, az * z));
                }
                phasor_store<MODE, R>(wunit + (size_t)u * a.unit_stride + operand_offset(s0 / SST, a.nant, ant) +
                                          (size_t)(s0 % SST) * elem,
                                      amp, ph, il4);
            }
        }
        return;
    }

    for (int64_t s0 = wg0 + (int64_t)threadIdx.x * PH_SRC_PER_THREAD; s0 < wg1;
         s0 += PH_THREADS * PH_SRC_PER_THREAD) {
        R amp[8];
        double d[8], sx[8], sy[8], sz[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t s = s0 + j;
            const double x = cx[s], y = cy[s], z = cz[s];
            R A = (R)0;
            if (z > 0.0) A = (R)beam_bilinear<BT>(sb, P, x, y);
            amp[j] = A * (R)isq[s] * (R)a.amp_scale;
            sx[j] = x;
            sy[j] = y;
            sz[j] = z;
            d[j] = fma(ax, x, fma(ay, y, az * z));   // FP32 / FP64: metres, shared by the frequencies
        }
        for (int u = 0; u < a.nfreq; ++u) {
            const double il = a.inv_lambda[u];          // turns per metre
            const double il4 = 4.0 * il;                // quarter turns per metre (exact scaling)
            double ph[8];
            if constexpr (split_mode(MODE)) {
                const double axl = ax * il, ayl = ay * il, azl = az * il;
#pragma unroll
                for (int j = 0; j < 8; ++j) ph[j] = phase_turns(axl, ayl, azl, sx[j], sy[j], sz[j]);
            }
            phasor_store<MODE, R>(wunit + (size_t)u * a.unit_stride + operand_offset(s0 / SST, a.nant, ant) +
                                      (size_t)(s0 % SST) * elem,
                                  amp, split_mode(MODE) ? ph : d, il4);
        }
    }
}

template <int MODE, typename BT>
static hipError_t launch_phasor_t(const PhasorArgs& a, hipStream_t s) {
    const size_t lds = (size_t)a.beam_px * a.beam_px * sizeof(BT);
    const int64_t nchunk = (a.kpad + PH_SRC_PER_WG - 1) / PH_SRC_PER_WG;
    if (nchunk > 65535 || a.ntime > 65535) return hipErrorInvalidConfiguration;
    dim3 grid((unsigned)a.nant, (unsigned)nchunk, (unsigned)a.ntime);
    if (lds > LDS_BYTES) {   // the antenna's beam plane does not fit: taps from global memory
        hipLaunchKernelGGL((phasor_kernel<MODE, BT, true>), grid, dim3(PH_THREADS), 0, s, a);
        return hipGetLastError();
    }
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)phasor_kernel<MODE, BT, false>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((phasor_kernel<MODE, BT, false>), grid, dim3(PH_THREADS), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_phasor(int mode, const PhasorArgs& a, hipStream_t s) {
    const bool f64 = a.in_f64 != 0;
    switch (mode) {
        case MODE_FP32:
            return f64 ? launch_phasor_t<MODE_FP32, double>(a, s) : launch_phasor_t<MODE_FP32, float>(a, s);
        case MODE_SPLIT_BF16:
            return f64 ? launch_phasor_t<MODE_SPLIT_BF16, double>(a, s)
                       : launch_phasor_t<MODE_SPLIT_BF16, float>(a, s);
        case MODE_SPLIT_FP16:
            return f64 ? launch_phasor_t<MODE_SPLIT_FP16, double>(a, s)
                       : launch_phasor_t<MODE_SPLIT_FP16, float>(a, s);
        case MODE_FP64:
            return f64 ? launch_phasor_t<MODE_FP64, double>(a, s) : launch_phasor_t<MODE_FP64, float>(a, s);
    }
    return hipErrorInvalidValue;
}

// ------------------------------------------------- polarised phasor build
// Grid (virtual row p N + a, chunk of 4096 real sources, time).  The four
// real planes (Re, Im of J^{p0} and J^{p1}) of antenna a, feed p go to LDS;
// each thread takes 4 consecutive sources (8 consecutive columns 2 s + k):
// one bilinear weight set per source for the four planes, the FP64 path
// difference once, then per frequency the phase in the arithmetic of the
// mode (whole turns with hardware sine/cosine for the split modes, quarter
// turns otherwise) and the complex product J sqrt(I) exp(i phase).
constexpr int PP_SRC_PER_WG = 4096;

// Four planes at once with one set of texel corners and weights (the
// arithmetic of beam_bilinear).  CPLX false: planes b + c P^2 (real, as
// staged in LDS); true: the two complex planes of a feed in global memory,
// (re, im) interleaved, plane c = real (c even) or imaginary part of complex
// plane c / 2.  The values, and so the results, are the same either way.
template <typename BT, bool CPLX = false, typename OUT>
__device__ inline void beam_bilinear4(const BT* __restrict__ b, int P, double x, double y, OUT (&out)[4]) {
    const double bx = (double)(P - 1) * (0.5 * x + 0.5);
    const double by = (double)(P - 1) * (0.5 * y + 0.5);
    const double px = floor(bx), py = floor(by);
    const BT fx = (BT)(bx - px), fy = (BT)(by - py);
    const double pm = (double)(P - 1);
    const int ix0 = (int)clampr(px, 0.0, pm), ix1 = (int)clampr(px + 1.0, 0.0, pm);
    const int iy0 = (int)clampr(py, 0.0, pm), iy1 = (int)clampr(py + 1.0, 0.0, pm);
    const size_t pl = (size_t)P * P;
    constexpr size_t ST = CPLX ? 2 : 1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const BT* q = CPLX ? b + (size_t)(c >> 1) * 2 * pl + (c & 1) : b + c * pl;
        const BT v00 = q[(iy0 * P + ix0) * ST], v01 = q[(iy0 * P + ix1) * ST];
        const BT v10 = q[(iy1 * P + ix0) * ST], v11 = q[(iy1 * P + ix1) * ST];
        const BT top = fma(fx, v01 - v00, v00);
        const BT bot = fma(fx, v11 - v10, v10);
        out[c] = (OUT)fma(fy, bot - top, top);
    }
}

// sin and cos of the delay phase of antenna (ax, ay, az) m and source (x, y,
// z) at 1/lambda = il, as the scalar-beam builders form it.
template <int MODE, typename R>
__device__ inline void phase_sincos(double ax, double ay, double az, double x, double y, double z, double il, R* s,
                                    R* c) {
    if constexpr (split_mode(MODE)) {
        sincos_turns_hw(phase_turns(ax * il, ay * il, az * il, x, y, z), s, c);
    } else {
        sincos_quarter_turns(fma(ax, x, fma(ay, y, az * z)) * (4.0 * il), s, c);
    }
}

// GB: the planes do not fit in LDS and the bilinear taps come from the
// Jones array in global memory (same values, same arithmetic).
template <int MODE, typename BT, bool GB>
__global__ __launch_bounds__(PH_THREADS) void phasor_pol_kernel(PhasorPolArgs a) {
    using R = typename std::conditional<MODE == MODE_FP64, double, float>::type;
    extern __shared__ __attribute__((aligned(16))) char smem[];
    BT* sb = reinterpret_cast<BT*>(smem);                 // [4][P][P]: Re J0, Im J0, Re J1, Im J1
    const int vrow = blockIdx.x;                          // p N + a
    const int p = vrow / a.nant, ant = vrow - p * a.nant;
    const int tb = blockIdx.z;
    const int P = a.beam_px;
    const size_t pl = (size_t)P * P;
    const BT* gj = reinterpret_cast<const BT*>(a.jones) + ((size_t)ant * 4 + (size_t)p * 2) * pl * 2;
    if constexpr (!GB) {
        for (size_t i = threadIdx.x; i < 2 * pl; i += PH_THREADS) {
            const size_t k = i / pl, e = i - k * pl;      // sky axis k, element e
            sb[(2 * k) * pl + e] = gj[(k * pl + e) * 2];
            sb[(2 * k + 1) * pl + e] = gj[(k * pl + e) * 2 + 1];
        }
        __syncthreads();
    }

    const double ax = a.antpos[3 * ant], ay = a.antpos[3 * ant + 1], az = a.antpos[3 * ant + 2];
    const double* cx = a.crd_top + (size_t)tb * 3 * a.kpad;
    const double* cy = cx + a.kpad;
    const double* cz = cx + 2 * a.kpad;
    const int nrow = 2 * a.nant;
    const int64_t kv = 2 * a.kpad;                        // columns
    constexpr int SST = stage_src(MODE == MODE_FP64);
    constexpr size_t elem = (MODE == MODE_FP64) ? 16 : 8;
    char* wunit = reinterpret_cast<char*>(a.W) + (size_t)tb * a.nfreq * a.unit_stride;
    const int64_t s_lo = (int64_t)blockIdx.y * PP_SRC_PER_WG;
    const int64_t s_hi = min(s_lo + (int64_t)PP_SRC_PER_WG, a.kpad);
    for (int64_t s0 = s_lo + (int64_t)threadIdx.x * 4; s0 < s_hi; s0 += PH_THREADS * 4) {
        R jr[4][2], ji[4][2], amp[4];
        double sx[4], sy[4], sz[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t s = s0 + j;
            const double x = cx[s], y = cy[s], z = cz[s];
            R bj[4] = {0, 0, 0, 0};
            if (z > 0.0) {
                if constexpr (GB) beam_bilinear4<BT, true>(gj, P, x, y, bj);
                else beam_bilinear4<BT>(sb, P, x, y, bj);
            }
            jr[j][0] = bj[0]; ji[j][0] = bj[1]; jr[j][1] = bj[2]; ji[j][1] = bj[3];
            amp[j] = (R)a.isqrt[s] * (R)a.amp_scale;
            sx[j] = x;
            sy[j] = y;
            sz[j] = z;
        }
        const int64_t col0 = 2 * s0;                      // 8 columns, all in one stage (SST >= 8)
        for (int u = 0; u < a.nfreq; ++u) {
            const double il = a.inv_lambda[u];
            R re[8], im[8];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                R sn, cs;
                phase_sincos<MODE, R>(ax, ay, az, sx[j], sy[j], sz[j], il, &sn, &cs);
                const R er = amp[j] * cs, ei = amp[j] * sn;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    re[2 * j + k] = jr[j][k] * er - ji[j][k] * ei;
                    im[2 * j + k] = jr[j][k] * ei + ji[j][k] * er;
                }
            }
            store_phasors8<MODE, R>(wunit + (size_t)u * a.unit_stride + operand_offset(col0 / SST, nrow, vrow) +
                                        (size_t)(col0 % SST) * elem,
                                    re, im);
        }
    }
    (void)kv;
}

template <int MODE, typename BT>
static hipError_t launch_phasor_pol_t(const PhasorPolArgs& a, hipStream_t s) {
    const size_t lds = (size_t)4 * a.beam_px * a.beam_px * sizeof(BT);
    const int64_t nchunk = (a.kpad + PP_SRC_PER_WG - 1) / PP_SRC_PER_WG;
    if (nchunk > 65535 || a.ntime > 65535 || a.kpad % 4) return hipErrorInvalidConfiguration;
    dim3 grid((unsigned)(2 * a.nant), (unsigned)nchunk, (unsigned)a.ntime);
    if (lds > LDS_BYTES) {   // the four planes of a feed do not fit: taps from global memory
        hipLaunchKernelGGL((phasor_pol_kernel<MODE, BT, true>), grid, dim3(PH_THREADS), 0, s, a);
        return hipGetLastError();
    }
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)phasor_pol_kernel<MODE, BT, false>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((phasor_pol_kernel<MODE, BT, false>), grid, dim3(PH_THREADS), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_phasor_pol(int mode, const PhasorPolArgs& a, hipStream_t s) {
    const bool f64 = a.in_f64 != 0;
    switch (mode) {
        case MODE_FP32:
            return f64 ? launch_phasor_pol_t<MODE_FP32, double>(a, s) : launch_phasor_pol_t<MODE_FP32, float>(a, s);
        case MODE_SPLIT_BF16:
            return f64 ? launch_phasor_pol_t<MODE_SPLIT_BF16, double>(a, s)
                       : launch_phasor_pol_t<MODE_SPLIT_BF16, float>(a, s);
        case MODE_SPLIT_FP16:
            return f64 ? launch_phasor_pol_t<MODE_SPLIT_FP16, double>(a, s)
                       : launch_phasor_pol_t<MODE_SPLIT_FP16, float>(a, s);
        case MODE_FP64:
            return f64 ? launch_phasor_pol_t<MODE_FP64, double>(a, s) : launch_phasor_pol_t<MODE_FP64, float>(a, s);
    }
    return hipErrorInvalidValue;
}

// ------------------------------------------- beam amplitude per time
// A sqrt(I) (times the FP16 operand scale) per (antenna, source, time) in
// FP32, for the contraction that builds its operands in place: the same
// arithmetic, in the same order, as the first half of phasor_kernel, so the
// operands built from it are bit for bit those phasor_kernel writes.  A
// workgroup serves AMP_ANT antennas (their beams in LDS): per source, the
// coordinates, the horizon test and the texel position and weights are
// formed once and applied to each antenna's beam.  Grid (antenna group,
// source chunk, time or unit).
// antennas per workgroup, at most (their beams must fit the 160 KB of LDS),
// and threads per workgroup: 8 antennas share each source's coordinates,
// horizon test and texel weights, with 8 waves to cover the latency that
// one 127 KB workgroup per CU leaves (measured on MI355X against 4 antennas
// and 256 threads: 1.14 -> 1.02 ms for 128 antennas x 65536 sources x 64
// times, 2.58 -> 2.24 ms for 350 x 2^20 x 2)
constexpr int AMP_ANT = 8;
constexpr int AMP_THREADS = 512;
// sources per workgroup: the beam planes staged in LDS serve 4x the sources
// of a phasor_kernel chunk (1.30 -> 1.10 ms and 2.84 -> 2.53 ms on the same
// two launches), unless that leaves fewer than 1024 workgroups (small jobs
// keep 8192)
constexpr int AMP_SRC_PER_WG = 32768;

// GB: not even one antenna's beam plane fits in LDS; the taps come from
// global memory (same values, same arithmetic).
template <typename BT, bool GB>
__global__ __launch_bounds__(AMP_THREADS) void amp_kernel(AmpArgs a) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    BT* sb = reinterpret_cast<BT*>(smem);                 // [AMP_ANT][P][P]
    const int64_t kact = eff_kpad(a.kpad, a.dyn_kpad);
    const int64_t wg0 = (int64_t)blockIdx.y * a.src_per_wg;
    if (wg0 >= kact) return;                       // uniform per workgroup
    const int ant0 = blockIdx.x * a.ants_per_wg;
    const int nq = min(a.ants_per_wg, a.nant - ant0);
    // z: the time of the launch, or (frequency-dependent beams) its unit
    const int z = blockIdx.z;
    const int tb = a.beam_fstride ? z / a.nfreq : z;
    const int fu = a.beam_fstride ? z - tb * a.nfreq : 0;
    const int P = a.beam_px;
    const int pp = P * P;
    const BT* gb = reinterpret_cast<const BT*>(a.beam) + fu * a.beam_fstride + (size_t)ant0 * pp;
    if constexpr (!GB) {
        for (int i = threadIdx.x; i < nq * pp; i += AMP_THREADS) sb[i] = gb[i];
        __syncthreads();
    }
    const double* cx = a.crd_top + (size_t)tb * 3 * a.kpad;
    const double* cy = cx + a.kpad;
    const double* cz = cx + 2 * a.kpad;
    const int64_t wg1 = min(wg0 + (int64_t)a.src_per_wg, kact);
    float* out = a.amp + (size_t)z * a.kpad * a.nant;
    const double pm = (double)(P - 1);
    for (int64_t s0 = wg0 + (int64_t)threadIdx.x * PH_SRC_PER_THREAD; s0 < wg1;
         s0 += AMP_THREADS * PH_SRC_PER_THREAD) {
        // texel corners and weights of each source (beam_bilinear's arithmetic)
        int o00[8], o01[8], o10[8], o11[8];
        BT fx[8], fy[8];
        float is[8];
        bool up[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t s = s0 + j;
            const double x = cx[s], y = cy[s], zz = cz[s];
            up[j] = zz > 0.0;
            const double bx = (double)(P - 1) * (0.5 * x + 0.5);
            const double by = (double)(P - 1) * (0.5 * y + 0.5);
            const double px = floor(bx), py = floor(by);
            fx[j] = (BT)(bx - px);
            fy[j] = (BT)(by - py);
            const int ix0 = (int)clampr(px, 0.0, pm), ix1 = (int)clampr(px + 1.0, 0.0, pm);
            const int iy0 = (int)clampr(py, 0.0, pm), iy1 = (int)clampr(py + 1.0, 0.0, pm);
            o00[j] = iy0 * P + ix0; o01[j] = iy0 * P + ix1;
            o10[j] = iy1 * P + ix0; o11[j] = iy1 * P + ix1;
            is[j] = (float)a.isqrt[s];
        }
        for (int q = 0; q < nq; ++q) {
            const BT* b = (GB ? gb : sb) + q * pp;
            float amp[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                float A = 0.f;
                if (up[j]) {
                    const BT v00 = b[o00[j]], v01 = b[o01[j]], v10 = b[o10[j]], v11 = b[o11[j]];
                    const BT top = fma(fx[j], v01 - v00, v00);
                    const BT bot = fma(fx[j], v11 - v10, v10);
                    A = (float)fma(fy[j], bot - top, top);
                }
                amp[j] = A * is[j] * a.amp_scale;
            }
            float4* d = reinterpret_cast<float4*>(out + ((size_t)(s0 / 16) * a.nant + ant0 + q) * 16 + (s0 % 16));
            d[0] = make_float4(amp[0], amp[1], amp[2], amp[3]);
            d[1] = make_float4(amp[4], amp[5], amp[6], amp[7]);
        }
    }
}

hipError_t launch_amp(const AmpArgs& args, hipStream_t s) {
    AmpArgs a = args;
    const size_t plane = (size_t)a.beam_px * a.beam_px * (a.in_f64 ? 8 : 4);
    const bool gb = plane > LDS_BYTES;   // not one plane fits: taps from global memory
    a.ants_per_wg = gb ? AMP_ANT : (int)std::max<size_t>(1, std::min<size_t>(AMP_ANT, LDS_BYTES / plane));
    const int64_t nz = a.beam_fstride ? (int64_t)a.ntime * a.nfreq : a.ntime;
    const int64_t ngroups = (a.nant + a.ants_per_wg - 1) / a.ants_per_wg;
    a.src_per_wg = ngroups * ((a.kpad + AMP_SRC_PER_WG - 1) / AMP_SRC_PER_WG) * nz >= 1024 ? AMP_SRC_PER_WG
                                                                                        : PH_SRC_PER_WG;
    const int64_t nchunk = (a.kpad + a.src_per_wg - 1) / a.src_per_wg;
    if (nchunk > 65535 || nz > 65535 || a.kpad % 16) return hipErrorInvalidConfiguration;
    dim3 grid((unsigned)((a.nant + a.ants_per_wg - 1) / a.ants_per_wg), (unsigned)nchunk, (unsigned)nz);
    if (gb) {
        if (a.in_f64) hipLaunchKernelGGL((amp_kernel<double, true>), grid, dim3(AMP_THREADS), 0, s, a);
        else hipLaunchKernelGGL((amp_kernel<float, true>), grid, dim3(AMP_THREADS), 0, s, a);
        return hipGetLastError();
    }
    const size_t lds = (size_t)a.ants_per_wg * plane;
    const void* k = a.in_f64 ? (const void*)amp_kernel<double, false> : (const void*)amp_kernel<float, false>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (a.in_f64)
        hipLaunchKernelGGL((amp_kernel<double, false>), grid, dim3(AMP_THREADS), lds, s, a);
    else
        hipLaunchKernelGGL((amp_kernel<float, false>), grid, dim3(AMP_THREADS), lds, s, a);
    return hipGetLastError();
}

// ----------------------------------------------------- standalone beam stage
// Kernel-level beam entry (one workgroup per (source chunk, antenna)), same
// device function as the fused phasor build.
template <typename R, bool GB>
__global__ __launch_bounds__(256) void beam_kernel(BeamArgs a) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int ant = blockIdx.y;
    const int P = a.beam_px;
    const R* gb = reinterpret_cast<const R*>(a.beam) + (size_t)ant * P * P;
    const R* sb = gb;   // GB: taps from global memory
    if constexpr (!GB) {
        R* st = reinterpret_cast<R*>(smem);
        for (int i = threadIdx.x; i < P * P; i += 256) st[i] = gb[i];
        __syncthreads();
        sb = st;
    }
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= a.nsrc) return;
    const R* c = reinterpret_cast<const R*>(a.crd_top);
    const double x = c[s], y = c[a.nsrc + s], z = c[2 * a.nsrc + s];
    R A = (R)0;
    if (z > 0.0) A = beam_bilinear<R>(sb, P, x, y);
    reinterpret_cast<R*>(a.A)[(size_t)ant * a.nsrc + s] = A;
}

hipError_t launch_beam(const BeamArgs& a, hipStream_t s) {
    dim3 grid((unsigned)((a.nsrc + 255) / 256), (unsigned)a.nant);
    const size_t lds = (size_t)a.beam_px * a.beam_px * (a.f64 ? 8 : 4);
    if (lds > LDS_BYTES) {
        if (a.f64) hipLaunchKernelGGL((beam_kernel<double, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((beam_kernel<float, true>), grid, dim3(256), 0, s, a);
        return hipGetLastError();
    }
    const void* k = a.f64 ? (const void*)beam_kernel<double, false> : (const void*)beam_kernel<float, false>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (a.f64) hipLaunchKernelGGL((beam_kernel<double, false>), grid, dim3(256), lds, s, a);
    else hipLaunchKernelGGL((beam_kernel<float, false>), grid, dim3(256), lds, s, a);
    return hipGetLastError();
}

// ------------------------------------------- caller phasors -> operand rows
// One thread per (antenna, 8 sources); zero beyond nsrc up to kpad.
template <int MODE>
__global__ __launch_bounds__(256) void split_kernel(SplitArgs a) {
    using R = typename std::conditional<MODE == MODE_FP64, double, float>::type;
    const int64_t s0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 8;
    const int ant = blockIdx.y;
    if (s0 >= a.kpad) return;
    R re[8], im[8];
    const R* v = reinterpret_cast<const R*>(a.v) + (size_t)ant * a.nsrc * 2;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int64_t s = s0 + j;
        re[j] = s < a.nsrc ? v[2 * s] * (R)a.scale : (R)0;
        im[j] = s < a.nsrc ? v[2 * s + 1] * (R)a.scale : (R)0;
    }
    constexpr int SST = stage_src(MODE == MODE_FP64);
    constexpr size_t elem = (MODE == MODE_FP64) ? 16 : 8;
    store_phasors8<MODE, R>(reinterpret_cast<char*>(a.W) + operand_offset(s0 / SST, a.nant, ant) +
                                (size_t)(s0 % SST) * elem, re, im);
}

hipError_t launch_split(int mode, const SplitArgs& a, hipStream_t s) {
    dim3 grid((unsigned)((a.kpad / 8 + 255) / 256), (unsigned)a.nant);
    switch (mode) {
        case MODE_FP32: hipLaunchKernelGGL(split_kernel<MODE_FP32>, grid, dim3(256), 0, s, a); break;
        case MODE_SPLIT_BF16: hipLaunchKernelGGL(split_kernel<MODE_SPLIT_BF16>, grid, dim3(256), 0, s, a); break;
        case MODE_SPLIT_FP16: hipLaunchKernelGGL(split_kernel<MODE_SPLIT_FP16>, grid, dim3(256), 0, s, a); break;
        case MODE_FP64: hipLaunchKernelGGL(split_kernel<MODE_FP64>, grid, dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------- K-split reduce
// V[a][b] = sum over splits, in split order and in FP64, of the (Re, Im)
// partial tiles; entries of the lower triangle are read from the mirrored
// upper-triangle entry and conjugated (the Hermitian fill of
// test_vis_gpu.py:354-355).  The diagonal is real.
__device__ inline int tile_id(int bi, int bj, int nb) { return bi * nb - (bi * (bi - 1)) / 2 + (bj - bi); }

template <typename ACC, typename OUT>
__global__ __launch_bounds__(256) void reduce_kernel(ReduceArgs r) {
    const int b = blockIdx.x * 32 + (threadIdx.x & 31);
    const int a = blockIdx.y * 8 + (threadIdx.x >> 5);
    const int u = blockIdx.z;
    if (a >= r.nant || b >= r.nant) return;
    if ((r.layout == LAYOUT_FTB || r.layout == LAYOUT_TBF) && a > b) return;   // baseline layouts: a <= b only
    const int T = r.tile;
    const int ta = a / T, tb = b / T;
    const bool mirror = (ta > tb) || (ta == tb && a > b);
    const int row = mirror ? b : a, col = mirror ? a : b;
    const int tid = mirror ? tile_id(tb, ta, r.nb) : tile_id(ta, tb, r.nb);
    const size_t tsz = (size_t)T * T;
    const size_t off = (size_t)(row % T) * T + (col % T);
    const ACC* p = reinterpret_cast<const ACC*>(r.partials) + ((size_t)u * r.nsplit * r.ntiles + tid) * 2 * tsz + off;
    const size_t split_stride = (size_t)r.ntiles * 2 * tsz;
    const int nsplit = (int)((eff_kpad((int64_t)r.nsplit * r.ksplit, r.dyn_kpad) + r.ksplit - 1) / r.ksplit);
    double re = 0.0, im = 0.0;
    for (int s = 0; s < nsplit; ++s) {
        re += (double)p[(size_t)s * split_stride];
        im += (double)p[(size_t)s * split_stride + tsz];
    }
    if (mirror) im = -im;
    if (a == b) im = 0.0;
    const int64_t t = r.t0 + u / r.nfreq_launch, f = r.f0 + u % r.nfreq_launch;
    int64_t idx;
    if (r.layout == LAYOUT_FULL) {
        idx = (f * r.ntimes_total + t) * (int64_t)r.nant * r.nant + (int64_t)a * r.nant + b;
    } else if (r.layout == LAYOUT_POL) {
        const int n = r.nant / 2;                         // antennas per feed
        const int p = a / n, q = b / n;
        idx = (((f * 2 + p) * 2 + q) * r.ntimes_total + t) * (int64_t)n * n + (int64_t)(a - p * n) * n + (b - q * n);
    } else {
        const int64_t nbl = (int64_t)r.nant * (r.nant + 1) / 2;
        const int64_t bl = (int64_t)a * r.nant - (int64_t)a * (a - 1) / 2 + (b - a);
        idx = (r.layout == LAYOUT_FTB) ? (f * r.ntimes_total + t) * nbl + bl
                                       : (t * nbl + bl) * r.nfreq_total + f;
    }
    OUT* o = reinterpret_cast<OUT*>(r.vis) + 2 * idx;
    o[0] = (OUT)(re * r.scale);
    o[1] = (OUT)(im * r.scale);
}

// Full-matrix layouts (LAYOUT_FULL, LAYOUT_POL): one workgroup per 32x32
// output block (i, j), i <= j, of one unit.  Each partial is read once; the
// block is written as summed and, through LDS, transposed and conjugated
// into block (j, i) - the values reduce_kernel writes, bit for bit, with
// half its reads.
__device__ inline int64_t full_index(const ReduceArgs& r, int64_t f, int64_t t, int a, int b) {
    if (r.layout == LAYOUT_POL) {
        const int n = r.nant / 2;                         // antennas per feed
        const int p = a / n, q = b / n;
        return (((f * 2 + p) * 2 + q) * r.ntimes_total + t) * (int64_t)n * n + (int64_t)(a - p * n) * n + (b - q * n);
    }
    return (f * r.ntimes_total + t) * (int64_t)r.nant * r.nant + (int64_t)a * r.nant + b;
}

template <typename ACC, typename OUT>
__global__ __launch_bounds__(256) void reduce_sym_kernel(ReduceArgs r) {
    __shared__ OUT sre[32][33], sim[32][33];
    const int n32 = (r.nant + 31) / 32;
    int i = 0, rem = blockIdx.x;
    while (rem >= n32 - i) { rem -= n32 - i; ++i; }
    const int j = i + rem;                                // 32-blocks i <= j
    const int u = blockIdx.y;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int T = r.tile;
    const int tid = tile_id((i * 32) / T, (j * 32) / T, r.nb);
    const size_t tsz = (size_t)T * T;
    const ACC* base = reinterpret_cast<const ACC*>(r.partials) + ((size_t)u * r.nsplit * r.ntiles + tid) * 2 * tsz;
    const size_t split_stride = (size_t)r.ntiles * 2 * tsz;
    const int nsplit = (int)((eff_kpad((int64_t)r.nsplit * r.ksplit, r.dyn_kpad) + r.ksplit - 1) / r.ksplit);
    const int64_t t = r.t0 + u / r.nfreq_launch, f = r.f0 + u % r.nfreq_launch;
    OUT* vis = reinterpret_cast<OUT*>(r.vis);
    const int b = j * 32 + tx;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int ly = ty + 8 * k;
        const int a = i * 32 + ly;
        double re = 0.0, im = 0.0;
        const bool in = a < r.nant && b < r.nant;
        if (in) {
            const ACC* p = base + (size_t)(a % T) * T + (b % T);
            for (int sp = 0; sp < nsplit; ++sp) {
                re += (double)p[(size_t)sp * split_stride];
                im += (double)p[(size_t)sp * split_stride + tsz];
            }
        }
        const OUT vr = (OUT)(re * r.scale), vi = (a == b) ? (OUT)0 : (OUT)(im * r.scale);
        if (in && (i < j || ly <= tx)) {
            OUT* o = vis + 2 * full_index(r, f, t, a, b);
            o[0] = vr;
            o[1] = vi;
        }
        sre[ly][tx] = vr;
        sim[ly][tx] = vi;
    }
    __syncthreads();
    // block (j, i): V[A][B] = conj(V[B][A]) for A > B
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int ry = ty + 8 * k;
        const int A = j * 32 + ry, B = i * 32 + tx;
        if (A < r.nant && B < r.nant && (i < j || tx < ry)) {
            OUT* o = vis + 2 * full_index(r, f, t, A, B);
            o[0] = sre[tx][ry];
            o[1] = -sim[tx][ry];
        }
    }
}

hipError_t launch_reduce(int mode, const ReduceArgs& r, hipStream_t s) {
    if (!ksplit_valid(r.ksplit)) return hipErrorInvalidValue;
    if (r.layout == LAYOUT_FULL || r.layout == LAYOUT_POL) {
        const int n32 = (r.nant + 31) / 32;
        dim3 grid((unsigned)(n32 * (n32 + 1) / 2), (unsigned)r.nunits);
        if (mode == MODE_FP64)
            hipLaunchKernelGGL((reduce_sym_kernel<double, double>), grid, dim3(256), 0, s, r);
        else
            hipLaunchKernelGGL((reduce_sym_kernel<float, float>), grid, dim3(256), 0, s, r);
        return hipGetLastError();
    }
    dim3 grid((unsigned)((r.nant + 31) / 32), (unsigned)((r.nant + 7) / 8), (unsigned)r.nunits);
    if (mode == MODE_FP64)
        hipLaunchKernelGGL((reduce_kernel<double, double>), grid, dim3(256), 0, s, r);
    else
        hipLaunchKernelGGL((reduce_kernel<float, float>), grid, dim3(256), 0, s, r);
    return hipGetLastError();
}

}  // namespace hv
