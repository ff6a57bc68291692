// Internal declarations shared by the HIP translation units of libheravis.
//
// Device data layout (one simulation "job"):
//   antpos     double[N][3]            metres
//   inv_lambda double[F]               nu / c  (1/m)
//   eq2tops    double[T][9]            row-major 3x3
//   crd_eq     double[3][Kpad]         zero beyond K
//   isqrt      double[Kpad]            sqrt(I_sky), zero beyond K
//   beam       real[N][P][P]           [ant][y][x]
//   crd_top    double[T][3][Kpad]      rotated source directions
//   W          operand[units][stage][N][128 B]  phasors, stage-major (below)
//   partials   acc[units][nsplit][ntiles][2][TILE][TILE]   (Re, Im) per split
//              of job_ksplit(...) sources
//   vis        complex[F][T][N][N]
//
// Phasor operands are stage-major: the sources are cut into stages of
// STAGE_SRC(mode) sources (128 bytes per antenna), and stage st of antenna a
// sits at byte ((st * N + a) * 128) of its unit, so one contraction stage of
// a 64- or 128-antenna tile is one contiguous 8 or 16 KB block.  Within a
// stage, (re, im)
// interleave per source:
//   FP32 / FP64 : complex64 (16 sources) / complex128 (8 sources);
//   split modes : 4 groups of 4 sources, each 32 bytes [8 hi | 8 lo] halves
//                 ordered (x0 y0 x1 y1 x2 y2 x3 y3).
// Bytes per unit: N * Kpad * 8 (FP32 and split), N * Kpad * 16 (FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hv {

constexpr int KALIGN = 64;          // Kpad multiple (sources)
constexpr int KSPLIT = 8192;        // sources per K-split (short splits)
constexpr int KSPLIT_LONG = 32768;  // sources per K-split of large jobs
constexpr int64_t KSPLIT_LONG_ITEMS = 32768;
// Sources per K-split of a job of `units` (time, frequency) units over kpad
// sources.  A split is one FP32 accumulator chain: the contraction writes
// one partial tile set per split and the reduce sums the splits in FP64, in
// split order.  The split length is a function of the job alone (arithmetic,
// antennas, units, padded sources), never of how its units are batched into
// launches, of variant overrides or of the serving buckets, so a job's
// results are the same bits however it runs.  A split-BF16/FP16 job that
// still has KSPLIT_LONG_ITEMS contraction work items (128 per CU of a
// 256-CU device) with 32768-source splits takes them: a quarter of the
// partial-tile writes and reduce reads.  Smaller jobs, and FP32 and FP64
// jobs (whose launches hold few units: their operands go through HBM), keep
// 8192-source splits for parallelism.
int job_ksplit(int mode, int nant, int64_t units, int64_t kpad);
// Work-item threshold of job_ksplit (0: KSPLIT_LONG_ITEMS); process-wide,
// for tests and tuning.
void set_ksplit_items(int64_t items);
inline bool ksplit_valid(int ksplit) { return ksplit == KSPLIT || ksplit == KSPLIT_LONG; }
constexpr int FOLD_SRC = 512;       // sources per FP32 accumulator chain
constexpr double C_LIGHT = 299792458.0;

enum Mode { MODE_FP32 = 0, MODE_SPLIT_BF16 = 1, MODE_SPLIT_FP16 = 2, MODE_FP64 = 3 };

constexpr int STAGE_BYTES = 128;
inline bool mode_is_f64(int mode) { return mode == MODE_FP64; }
__host__ __device__ constexpr int stage_src(bool f64) { return f64 ? 8 : 16; }
// bytes of one antenna over all sources (the unit holds N of these)
inline size_t operand_row_bytes(int mode, int64_t kpad) {
    return (size_t)kpad * (mode_is_f64(mode) ? 16 : 8);
}
// byte offset of (stage, antenna) inside a unit
__host__ __device__ inline size_t operand_offset(int64_t stage, int nant, int ant) {
    return ((size_t)stage * nant + ant) * STAGE_BYTES;
}

// Every stage takes the padded source count `kpad` that sizes its buffers
// and grid, and optionally `dyn_kpad`, a device pointer to the number of
// sources actually present (a multiple of KALIGN, <= kpad).  Work beyond
// *dyn_kpad is skipped, so a graph captured for a bucket of source counts
// serves every job of the bucket with the same arithmetic as an exact-size
// run.
__device__ __host__ inline int64_t eff_kpad(int64_t kpad, const int64_t* dyn) {
#ifdef __HIP_DEVICE_COMPILE__
    return dyn ? *dyn : kpad;
#else
    (void)dyn;
    return kpad;
#endif
}

struct CoordsArgs {
    const double* eq2tops;   // [T][9]
    const double* crd_eq;    // [3][Kpad]
    double* crd_top;         // [T][3][Kpad]
    int64_t kpad;
    int ntimes;
    const int64_t* dyn_kpad;
};

struct PhasorArgs {
    const double* crd_top;   // [3][Kpad] for the first time of the launch;
                             // time tb of the launch at crd_top + tb*3*Kpad
    const double* isqrt;     // [Kpad]
    const void* beam;        // [N][P][P] float or double (in_f64)
    const double* antpos;    // [N][3]
    const double* inv_lambda;// [nfreq] nu/c of the frequencies of this launch
    void* W;                 // operand base of unit 0
    size_t unit_stride;      // bytes between units
    int64_t kpad;
    int nant;
    int beam_px;
    int nfreq;               // frequencies per time in this launch
    int ntime;               // times in this launch (grid z); unit = tb*nfreq + f
    float amp_scale;         // multiplies A*sqrt(I) (power of two; FP16 range)
    int in_f64;              // beam element type is double
    const int64_t* dyn_kpad;
    size_t beam_fstride;     // elements between the beam planes of consecutive
                             // frequencies (0: one beam for all frequencies);
                             // beam points at the launch's first frequency
};

struct HerkArgs {
    const void* W;           // operand base of unit 0
    size_t unit_stride;
    void* partials;          // [units][nsplit][ntiles][2][TILE][TILE]
    int nant;                // valid rows
    int nb;                  // tile blocks per dim
    int ntiles;              // nb (nb+1) / 2
    int64_t kpad;
    int nsplit;
    int ksplit;              // sources per split (job_ksplit)
    int nunits;
    const int64_t* dyn_kpad;
};

// Output layouts of the visibilities (HV_LAYOUT_* in heravis.h).  LAYOUT_POL:
// a polarised job's (2N) x (2N) correlation matrix (row p N + a: feed p of
// antenna a) written as [F][2][2][T][N][N] (feed p, feed q, time, a, b).
enum Layout { LAYOUT_FULL = 0, LAYOUT_FTB = 1, LAYOUT_TBF = 2, LAYOUT_POL = 3 };
// Complex elements of an output: full N x N per unit, or the N (N+1) / 2
// baselines a <= b (index a N - a (a-1) / 2 + (b - a)).
inline size_t out_elems(int layout, int nfreq, int ntimes, int nant) {
    const size_t per = (layout == LAYOUT_FULL) ? (size_t)nant * nant : (size_t)nant * (nant + 1) / 2;
    return (size_t)nfreq * ntimes * per;
}

// Polarised phasor build: rows (feed p, antenna a) = p N + a of 2N, columns
// (source s, sky axis k) = 2 s + k of 2 Kpad; operand
//   v[p N + a][2 s + k] = J_a^{pk}(s) sqrt(I_s) exp(+i 2 pi nu d_a(s) / c)
// with the complex Jones beam J_a^{pk} interpolated bilinearly (real and
// imaginary parts with the weights of the scalar beam).
struct PhasorPolArgs {
    const double* crd_top;   // [3][kpad] (real sources) of the launch's first time; tb at + tb*3*kpad
    const double* isqrt;     // [kpad]
    const void* jones;       // [N][2 feeds][2 axes][P][P] complex (re, im), float or double (in_f64)
    const double* antpos;    // [N][3]
    const double* inv_lambda;// [nfreq] nu/c of the launch's frequencies
    void* W;                 // operands of unit 0 (2N rows, 2 kpad columns, stage-major)
    size_t unit_stride;
    int64_t kpad;            // real sources (padded)
    int nant;                // real antennas
    int beam_px, nfreq, ntime, in_f64;
    float amp_scale;
};

struct ReduceArgs {
    const void* partials;
    void* vis;               // complex[F][T][N][N], [F][T][Nbl] or [T][Nbl][F] (layout)
    int t0, f0;              // first time / frequency of the launch
    int nfreq_launch;        // unit u of the launch = (t0 + u / nfreq_launch,
                             //                        f0 + u % nfreq_launch)
    int ntimes_total;        // T of the output array
    int nant;
    int nb;
    int ntiles;
    int nsplit;
    int ksplit;              // sources per split (job_ksplit)
    int nunits;
    int tile;
    double scale;            // 1 / amp_scale^2
    const int64_t* dyn_kpad; // splits summed = ceil(*dyn_kpad / ksplit)
    int layout;              // Layout
    int nfreq_total;         // F of the output array
};
hipError_t launch_phasor_pol(int mode, const PhasorPolArgs& a, hipStream_t s);

struct BeamArgs {
    const void* beam;
    const void* crd_top;     // [3][K] in dtype
    void* A;                 // [N][K] in dtype
    int64_t nsrc;
    int nant;
    int beam_px;
    int f64;
};

struct SplitArgs {           // complex phasors -> operands (hv_k_herk)
    const void* v;           // [N][K] complex64 or complex128
    void* W;
    int64_t nsrc;
    int64_t kpad;
    int nant;
    float scale;
};

// Beam amplitude A sqrt(I) (times the FP16 operand scale) per (antenna,
// source, time): the input of the contraction that builds its operands in
// place.  amp is [ntime][stage][N][16] FP32 (stage = 16 sources).
struct AmpArgs {
    const double* crd_top;   // [3][Kpad] of the first time; time tb at + tb*3*Kpad
    const double* isqrt;     // [Kpad]
    const void* beam;        // [N][P][P] float or double (in_f64), first frequency's plane
    float* amp;              // [ntime][Kpad/16][N][16], or [ntime*nfreq][...] per unit
    int64_t kpad;
    int nant, beam_px, ntime, in_f64;
    float amp_scale;
    const int64_t* dyn_kpad;
    size_t beam_fstride;     // 0: one beam for all frequencies (amp per time);
                             // else elements between frequency planes (amp per unit)
    int nfreq;               // frequencies per time of the launch (per-unit amp)
    int ants_per_wg;         // set by launch_amp (beams that fit in LDS)
    int src_per_wg;          // set by launch_amp (sources per workgroup)
};

// Contraction of split-mode operands generated in place (see hv_herk.hip).
// Unit u of the launch is time u / nfreq, frequency u % nfreq.
struct FusedArgs {
    const float* amp;         // AmpArgs::amp of the launch's times
    const double* crd_top;    // [3][Kpad] of the first time; time tb at + tb*3*Kpad
    const double* antpos;     // [N][3]
    const double* inv_lambda; // [nfreq] nu/c of the launch's frequencies
    void* partials;           // as HerkArgs::partials
    int64_t kpad;
    int nant, nb, ntiles, nsplit, nunits, nfreq;
    int ksplit;               // sources per split (job_ksplit)
    const int64_t* dyn_kpad;
    int amp_per_unit;         // amp holds one block per unit (frequency-dependent beams)
    int lists;                // 192-tile block lists: 0 built-in, 1 tile pairs only (no triangle/strip)
    int job_nfreq;            // frequencies per time of the whole job (items of several units need
                              // herk_mu_units of them; 0: the launch's nfreq)
};

hipError_t launch_amp(const AmpArgs& a, hipStream_t s);
hipError_t launch_fused(int mode, int tile, const FusedArgs& a, hipStream_t s);
bool herk_fused(int mode, int nant);   // does the built-in choice build operands in place?
// Units per multi-unit work item at this antenna count (6, 4, 3 at 2, 3, 4
// blocks of 32 rows), 0 where there are none.
int herk_mu_units(int nant);

int herk_tile(int mode, int nant);
int herk_tile_operands(int mode, int nant);

hipError_t launch_coords(const CoordsArgs& a, hipStream_t s);
hipError_t launch_phasor(int mode, const PhasorArgs& a, hipStream_t s);
hipError_t launch_herk(int mode, int tile, const HerkArgs& a, hipStream_t s);
hipError_t launch_herk_variant(int mode, int tile, int pf, const HerkArgs& a, hipStream_t s);
int herk_prefetch(int mode, int tile);
void set_herk_variant(int tile, int pf);
void herk_init();   // caches per-kernel launch facts (resident workgroups) for the current device
hipError_t launch_reduce(int mode, const ReduceArgs& a, hipStream_t s);
hipError_t launch_beam(const BeamArgs& a, hipStream_t s);
hipError_t launch_split(int mode, const SplitArgs& a, hipStream_t s);

}  // namespace hv
