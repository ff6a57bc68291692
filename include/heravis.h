/*
 * heravis.h -- C ABI of the MI355X (gfx950) visibility simulator.
 *
 * Drop-in target: hera_gpu.vis.vis_gpu (hera_gpu/vis.py:144-262), the only
 * public entry of the reference's visibility path (re-exported at
 * hera_gpu/__init__.py:3).  The reference is Python calling PyCUDA and
 * cuBLAS; it has no C ABI of its own, so the entry points below are what its
 * ctypes-style FFI for that path binds: one call for the whole simulation
 * plus one call per device stage that its kernel-level tests
 * (hera_gpu/tests/test_vis_gpu.py:17-169) launch on their own.
 *
 * Conventions (all entry points):
 *   - return 0 on success, a negative HV_ERR_* code on failure; the message
 *     is available from hv_last_error() on the calling thread;
 *   - host pointers are caller-owned, read (or written) only during the call;
 *   - arrays are C-contiguous; complex outputs are interleaved (re, im);
 *   - "dtype" selects float32 (HV_F32) or float64 (HV_F64) element type for
 *     real inputs, and complex64 / complex128 for complex outputs;
 *   - shapes follow the reference: antpos (N,3), eq2tops (T,3,3),
 *     crd_eq (3,K), I_sky / isqrt (K,), bm_cube (N,P,P) indexed [ant,y,x],
 *     1 <= P <= 2048 (beam planes that fit in the 160 KB of LDS are staged
 *     there, larger ones are read from device memory by the same bilinear
 *     arithmetic, as the reference's texture path takes any P,
 *     vis.py:113-138),
 *     visibilities (F,T,N,N) with V[a,b] = sum_s v_a(s) conj(v_b(s)).
 */
#ifndef HERAVIS_H
#define HERAVIS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HV_OK 0
#define HV_ERR_ARG (-1)     /* bad shape / argument: Python raises AssertionError */
#define HV_ERR_HIP (-2)     /* HIP runtime error */
#define HV_ERR_NOMEM (-3)   /* request exceeds the HBM budget */
#define HV_ERR_STATE (-4)   /* wrong object state (e.g. closed server) */
#define HV_ERR_BUSY (-5)    /* non-blocking call would block */
#define HV_ERR_CANCELED (-6) /* the job was cancelled (hv_server_cancel) */

#define HV_F32 0
#define HV_F64 1

/* Arithmetic of the Hermitian contraction.  The delay phase is FP64 in all. */
#define HV_MODE_FP32 0        /* complex64 phasors, FP32 MFMA               */
#define HV_MODE_SPLIT_BF16 1  /* bf16 hi+lo phasors, 3 bf16 MFMA products   */
#define HV_MODE_SPLIT_FP16 2  /* fp16 hi+lo phasors (power-of-2 scaled)     */
#define HV_MODE_FP64 3        /* complex128 phasors, FP64 MFMA (precision=2) */

typedef struct hv_ctx hv_ctx;

/* Library version as an integer (major*10000 + minor*100 + patch). */
int hv_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* hv_last_error(void);

/* Number of visible HIP devices (0 on a machine without a GPU). */
int hv_device_count(int* count);

/* Context: one device, a stream, a grow-only device workspace bounded by
 * hbm_budget bytes (0 = 90% of free device memory).  Replaces the import-time
 * pycuda.autoinit context and per-call cublasCreate (vis.py:1,196). */
int hv_ctx_create(int device, size_t hbm_budget, hv_ctx** out);
int hv_ctx_destroy(hv_ctx* ctx);

/* ---------------------------------------------------------------------------
 * Whole simulation, host arrays in and out (vis.py:144-262).
 *   dtype  : element type of antpos, eq2tops, crd_eq, isqrt and bm_cube
 *            (HV_F32 for precision=1, HV_F64 for precision=2);
 *   freqs  : F frequencies in Hz (the reference takes one scalar, vis.py:159);
 *   isqrt  : sqrt(I_sky), taken by the caller as the reference does
 *            (vis.py:174);
 *   vis_out: (F,T,N,N) complex64 when mode != HV_MODE_FP64, complex128 else.
 * Unlike the reference, every source is summed (vis.py:176-177 drops the
 * K mod chunk tail).
 * ------------------------------------------------------------------------- */
int hv_vis(hv_ctx* ctx, int mode, int dtype,
           const void* antpos, int nant,
           const double* freqs, int nfreq,
           const void* eq2tops, int ntimes,
           const void* crd_eq, const void* isqrt, int64_t nsrc,
           const void* bm_cube, int beam_px,
           void* vis_out);

/* Output layouts.  The baseline layouts hold the N (N+1) / 2 baselines
 * (a, b), a <= b (autos included), at index a N - a (a-1) / 2 + (b - a), i.e.
 * V[a][b] = sum v_a conj(v_b) of the full layout; half the bytes of the full
 * matrix.  No counterpart in the reference (which returns (T,N,N),
 * vis.py:212,262): the output writer that turns (T,N,N) into the baseline-
 * major, time-major (Nblts, Nfreqs) order of a UVData data array. */
#define HV_LAYOUT_FULL 0   /* (F, T, N, N)                 */
#define HV_LAYOUT_FTB 1    /* (F, T, Nbl)                  */
#define HV_LAYOUT_TBF 2    /* (T, Nbl, F): UVData blt order */

/* hv_vis with per-frequency beams and an output layout (hv_vis =
 * beam_nfreq 1, HV_LAYOUT_FULL).  beam_nfreq = nfreq: bm_cube is
 * (F, N, P, P), one beam per frequency; no counterpart in the reference,
 * whose bm_cube is (N, P, P) for every frequency (vis.py:167,175). */
int hv_vis_ex(hv_ctx* ctx, int mode, int dtype,
              const void* antpos, int nant,
              const double* freqs, int nfreq,
              const void* eq2tops, int ntimes,
              const void* crd_eq, const void* isqrt, int64_t nsrc,
              const void* bm_cube, int beam_px, int beam_nfreq,
              int layout, void* vis_out);

/* ---------------------------------------------------------------------------
 * Kernel-level entry points, one device stage each, host arrays in and out.
 * ------------------------------------------------------------------------- */

/* Rotation: crd_top (3,K) = eq2top (3,3) . crd_eq (3,K), computed in FP64,
 * written in dtype.  Replaces the rotation Sgemm/Dgemm (vis.py:228-232);
 * pinned like test_cublas (test_vis_gpu.py:140-169). */
int hv_k_rotate(hv_ctx* ctx, int dtype, const void* eq2top, const void* crd_eq,
                int64_t nsrc, void* crd_top_out);

/* Beam: A (N,K) = clamp-to-edge bilinear of bm_cube[ant] at crd_top (x,y),
 * zero unless z > 0, from an LDS-staged copy of each antenna's beam.
 * Replaces InterpolateBeam (vis.py:46-89); pinned like test_bm_interp_flat /
 * test_bm_interp_middle (test_vis_gpu.py:31-102). */
int hv_k_beam(hv_ctx* ctx, int dtype, const void* bm_cube, int nant, int beam_px,
              const void* crd_top, int64_t nsrc, void* A_out);

/* Phasor: v (N,K) = A * isqrt * exp(+i 2 pi nu (antpos . crd_top) / c) with
 * the beam A as in hv_k_beam and the delay phase reduced in FP64 turns;
 * written in the operand format of `mode` and returned as complex64
 * (complex128 for HV_MODE_FP64; split modes return hi+lo).  Replaces the
 * delay Sgemm (vis.py:233-236) and MeasEq (vis.py:92-110); pinned like
 * test_meas_eq (test_vis_gpu.py:103-139). */
int hv_k_phasor(hv_ctx* ctx, int mode, int dtype,
                const void* bm_cube, int nant, int beam_px,
                const void* crd_top, const void* isqrt, int64_t nsrc,
                const void* antpos, double freq, void* v_out);

/* Hermitian product: V (N,N) = v . v^H of caller phasors v (N,K) (complex64,
 * or complex128 for HV_MODE_FP64), split to the operand format of `mode`.
 * Replaces the visibility Cgemm/Zgemm (vis.py:246-250) and the host chunk
 * sum (vis.py:258). */
int hv_k_herk(hv_ctx* ctx, int mode, const void* v, int nant, int64_t nsrc,
              void* vis_out);

/* ---------------------------------------------------------------------------
 * Resident inputs and timed runs (no reference counterpart: the reference
 * re-uploads everything per call, vis.py:198-206).  Used by the serving loop
 * and by measurement, so timed regions start with inputs already in HBM.
 * ------------------------------------------------------------------------- */
typedef struct hv_sky hv_sky;     /* crd_eq (3,K) + isqrt (K), FP64 in HBM */
typedef struct hv_beam hv_beam;   /* bm_cube (N,P,P) in HBM                */

int hv_sky_upload(hv_ctx* ctx, int dtype, const void* crd_eq, const void* isqrt,
                  int64_t nsrc, hv_sky** out);
int hv_sky_free(hv_sky* sky);
int hv_beam_upload(hv_ctx* ctx, int dtype, const void* bm_cube, int nant,
                   int beam_px, hv_beam** out);
/* A beam per frequency, (nfreq, nant, P, P): calls and jobs using it must
 * have exactly nfreq frequencies (no reference counterpart, as hv_vis_ex). */
int hv_beam_upload_freq(hv_ctx* ctx, int dtype, const void* bm_cube, int nfreq,
                        int nant, int beam_px, hv_beam** out);
int hv_beam_free(hv_beam* beam);

/* Per-stage device times of the last hv_simulate call with timing on, in
 * milliseconds, summed over all launches of that stage, plus launch counts. */
typedef struct hv_stage_times {
    double coords_ms, phasor_ms, herk_ms, reduce_ms, total_ms;
    int64_t coords_n, phasor_n, herk_n, reduce_n;
} hv_stage_times;

/* Run a simulation on resident sky and beam, using the first nsrc_used sky
 * sources.  antpos/eq2tops are host arrays in `dtype`.  If vis_out is not
 * NULL the (F,T,N,N) result is copied to it; reps >= 1 repeats the run (for
 * timing) and, if times != NULL, records per-stage HIP events. */
int hv_simulate(hv_ctx* ctx, int mode, const hv_sky* sky, int64_t nsrc_used,
                const hv_beam* beam, int dtype, const void* antpos, int nant,
                const double* freqs, int nfreq, const void* eq2tops, int ntimes,
                void* vis_out, int reps, hv_stage_times* times);

/* ---------------------------------------------------------------------------
 * The whole simulation on device arrays, allocating nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels as hv_vis and results bitwise equal to hv_vis on
 * the same values.
 *   mode   : HV_MODE_FP32 or HV_MODE_SPLIT_BF16;
 *   antpos (N,3) m, freqs (F) Hz, eq2tops (T,3,3), crd_eq (3,nsrc) and isqrt
 *            (nsrc): device float64; bm_cube (N,P,P): device float32;
 *   vis_out: device (F,T,N,N) complex64;
 *   nsrc   : a multiple of 64 (pad with zero directions: below the horizon,
 *            they contribute nothing);
 *   stream : a hipStream_t (NULL: the default stream).  The call only
 *            enqueues work on it; the arrays must stay valid until it ends.
 * Working memory (rotated coordinates, A sqrt(I) or operands, partial tiles)
 * is a static device array of HV_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; times are processed in batches that fit it, and a call
 * whose single time does not fit fails with HV_ERR_NOMEM (e.g. N = 350 fits up
 * to ~1.5e5 sources, N = 37 up to ~1.5e6).  All calls on one GPU share that
 * array: calls that overlap on one GPU race.
 * ------------------------------------------------------------------------- */
#define HV_NOALLOC_SCRATCH_BYTES (256u << 20)
int hv_vis_device_noalloc(hv_ctx* ctx, int mode, const double* antpos, int nant, const double* freqs, int nfreq,
                          const double* eq2tops, int ntimes, const double* crd_eq, const double* isqrt,
                          int64_t nsrc, const float* bm_cube, int beam_px, void* vis_out, void* stream);

/* ---------------------------------------------------------------------------
 * Polarised (Jones) beams.  No counterpart in the reference, whose beam is a
 * real scalar per antenna (vis.py:162-175, "XXX complex?" at vis.py:175).
 * jones: complex [nant][2 feeds p][2 sky axes k][beam_px][beam_px], (re, im)
 * interleaved, float32 (dtype HV_F32) or float64; the other arrays as for
 * hv_vis.  For an unpolarised sky of intensity I_s:
 *   V_ab^{pq} = sum_s sum_k J_a^{pk}(s) conj(J_b^{qk}(s)) I_s exp(i 2 pi nu (tau_a - tau_b))
 * with J interpolated bilinearly per real and imaginary part (the scalar
 * beam's rule; zero below the horizon).  vis_out: complex [nfreq][2][2]
 * [ntimes][nant][nant] (feed p, feed q, time, a, b).  J_a^{pk} = A_a
 * delta_pk gives the scalar result on p = q and zeros on p != q.  Runs the
 * phasor build + LDS-ring contraction on 2 nant rows and 2 nsrc columns.
 * ------------------------------------------------------------------------- */
int hv_vis_pol(hv_ctx* ctx, int mode, int dtype, const void* antpos, int nant, const double* freqs, int nfreq,
               const void* eq2tops, int ntimes, const void* crd_eq, const void* isqrt, int64_t nsrc,
               const void* jones, int beam_px, void* vis_out);

/* ---------------------------------------------------------------------------
 * Omnical redundant calibration, replacing hera_gpu.redcal.omnical
 * (redcal.py:358-743; kernels redcal.py:70-355).  ndata independent pixels,
 * each with nbls baselines (ggu_indices[b] = (i, j, k): data ~ g_i conj(g_j)
 * u_k), nants gains and nubls unique baselines.  precision 1: float32 /
 * complex64 arrays, 2: float64 / complex128.  gains [ndata][nants] and
 * ubls [ndata][nubls] complex are read as the starting point and overwritten
 * with the solution; data [ndata][nbls] complex, wgts [ndata][nbls] real;
 * chisq, conv [ndata] real and iters [ndata] out.  conv_crit, maxiter,
 * check_every (> 1), check_after and gain have the meaning of the
 * reference's arguments.  Deterministic (fixed summation order, no atomics);
 * a pixel that converges or whose chi^2 rises keeps its previous solution.
 * Needs 5 (nants + nubls) + 256 values of LDS per workgroup (<= 160 KB); the
 * per-baseline arrays (weighted products, data weights) go to LDS when they
 * fit and at least two workgroups per CU stay resident, else to device memory.
 * The device buffer stays with the context for later calls (as hv_medfilt2d).
 * ------------------------------------------------------------------------- */
int hv_omnical(hv_ctx* ctx, int precision, const uint32_t* ggu_indices, int nbls, void* gains, int nants,
               void* ubls, int nubls, const void* data, const void* wgts, int64_t ndata, double conv_crit,
               int maxiter, int check_every, int check_after, double gain, void* chisq, uint32_t* iters,
               void* conv);
/* hv_omnical with the inputs uploaded once and the solve run 1 + reps times
 * from the same starting point: ms per solve (kernel only, mean of the last
 * reps); the outputs are those of the last solve.  No reference counterpart
 * (measurement). */
int hv_bench_omnical(hv_ctx* ctx, int precision, const uint32_t* ggu_indices, int nbls, void* gains, int nants,
                     void* ubls, int nubls, const void* data, const void* wgts, int64_t ndata, double conv_crit,
                     int maxiter, int check_every, int check_after, double gain, void* chisq, uint32_t* iters,
                     void* conv, int reps, double* ms_per_solve);
/* Process-wide cap on where hv_omnical keeps its per-baseline arrays (no
 * reference counterpart; tests and tuning): -1 automatic (as much in LDS as
 * fits), 0 all in device memory, 1 the weighted products in LDS, 2 products
 * and data weights in LDS.  A cap above what fits is lowered to what fits. */
int hv_set_omnical_place(int place);

/* ---------------------------------------------------------------------------
 * 2-D median filter, replacing hera_gpu.median_filter.MedianFilter
 * (median_filter.py:8-365; kernel mf, :176-216).  images and out are host
 * arrays [nimg][nrows][ncols] float32; out[i][r][c] is the element of rank
 * (wrows*wcols)/2 of the wrows x wcols window whose top-left corner is at
 * (r - wrows/2, c - wcols/2), zeros outside the image (the reference's zero
 * padding, :318).  Ranks follow IEEE totalOrder (-0 < +0, NaNs at the ends by
 * sign).  Window sides 1..127; odd sides up to 11 use compile-time selection
 * networks, other shapes a radix selection.  All images in one launch; the
 * device buffer (input and output of the batch) stays with the context for
 * later calls, growing as needed, until hv_ctx_destroy.
 * ------------------------------------------------------------------------- */
int hv_medfilt2d(hv_ctx* ctx, const float* images, float* out, int nimg, int nrows, int ncols, int wrows,
                 int wcols);
/* The filter alone over `images` uploaded once: ms per launch over `reps`
 * launches (after one warm-up), inputs and outputs resident in HBM. */
int hv_bench_medfilt2d(hv_ctx* ctx, const float* images, int nimg, int nrows, int ncols, int wrows, int wcols,
                       int reps, double* ms_per_launch);

/* ---------------------------------------------------------------------------
 * Calibration (no reference counterpart): a plain device copy of `bytes`
 * (ms per copy; HBM bandwidth = 2*bytes/time, and the yardstick for
 * rocprofv3 byte counters) and register-resident MFMA loops of each
 * arithmetic (measured matrix-core TFLOP/s).
 * ------------------------------------------------------------------------- */
int hv_bench_copy(hv_ctx* ctx, size_t bytes, int reps, double* ms_per_copy);
int hv_bench_mfma(hv_ctx* ctx, int mode, int iters, double* tflops);
/* The Hermitian contraction alone on synthetic operands of `units` units of
 * nant x nsrc; tile (64/128) and pf select a kernel variant (as for
 * hv_set_herk_variant), 0 = the production choice (its HBM-operand form: 128
 * where the production choice is a 192-tile in-place build).  Returns ms per
 * launch. */
int hv_bench_herk(hv_ctx* ctx, int mode, int nant, int64_t nsrc, int units, int tile, int pf, int reps,
                  double* ms_per_launch);
/* Process-wide choice of contraction variant for later calls (tuning and
 * tests): tile 64/128/192, pf 1/2 (register-staged prefetch depth), 3 (LDS
 * ring, one workgroup per work item), 4 (LDS ring on a persistent grid) or 5
 * (split modes: operands built inside the contraction from A sqrt(I) per
 * time, no operand buffer; hv_k_herk and hv_bench_herk, which take operands,
 * use 4) or 6 (as 5, with the 192-tile kernel's tile-pair work items also for
 * arrays of 321..352 antennas, which otherwise take triangle and strip items;
 * the same bits).  192-tiles exist only with pf 5, 6 (or 0).  0 restores the
 * built-in choice.  Not safe while other calls run. */
int hv_set_herk_variant(int tile, int pf);
/* The contraction variant (tile edge, pf as above) that calls of this mode and
 * antenna count use now. */
int hv_herk_choice(int mode, int nant, int* tile, int* pf);
/* Sources per K-split (8192 or 32768) of a job of `units` (time, frequency)
 * units over nsrc sources.  Each split is one FP32 accumulator chain whose
 * partial tiles the reduce sums in FP64; the length is fixed by the job
 * alone (never by batching, variants or serving buckets): 32768 for a
 * split-BF16/FP16 job that still has 32768 contraction work items with it,
 * 8192 else (FP32, FP64 and polarised jobs always). */
int hv_job_ksplit(int mode, int nant, int64_t units, int64_t nsrc, int* ksplit);
/* Process-wide work-item threshold of that rule for later calls (tests and
 * tuning; 0 restores 32768).  Not safe while other calls run. */
int hv_set_ksplit_items(int64_t items);

/* ---------------------------------------------------------------------------
 * Serving (no reference counterpart: the reference has no request loop; it
 * recompiles and re-uploads on every call, vis.py:183-206).
 * A server admits queued jobs into slots as slots and device memory free up;
 * admitted jobs run concurrently on per-slot streams.  A job runs in an
 * instance of its shape bucket (mode, N, P, T, F, beam type, FP16 scale,
 * 2^ceil(log2 Kpad), layout): the device buffers of such a job at fixed
 * addresses and the hipGraph captured from the first job that ran in it.
 * Instances outlive their jobs, so a later job of the bucket replays the
 * graph on whichever slot admits it; idle instances are evicted least
 * recently used first when a new one does not fit the budget.  Results are
 * bitwise identical to hv_simulate on the same inputs, whatever else runs.
 * ------------------------------------------------------------------------- */
typedef struct hv_server hv_server;

#define HV_JOB_FAIL_TEST 1   /* flags: fail this job at admission (host-side test hook) */
#define HV_JOB_FTB 2         /* flags: write vis_out in HV_LAYOUT_FTB                  */
#define HV_JOB_TBF 4         /* flags: write vis_out in HV_LAYOUT_TBF                  */

typedef struct hv_job {
    int mode;                  /* HV_MODE_*                                        */
    int flags;                 /* HV_JOB_* bits                                    */
    const hv_sky* sky;         /* resident catalog; the first nsrc sources are used */
    int64_t nsrc;
    const hv_beam* beam;       /* resident beam, nant x P x P                      */
    const double* antpos;      /* host (nant,3), copied at submit                  */
    int nant;
    const double* freqs;       /* host (nfreq,) Hz, copied at submit               */
    int nfreq;
    const double* eq2tops;     /* host (ntimes,3,3), copied at submit              */
    int ntimes;
    void* vis_out;             /* host (F,T,N,N) complex64 (complex128 for FP64),
                                  or a baseline layout (flags), written before the
                                  job is reported finished; from submission on,
                                  its pages may be pre-faulted (zero bytes written)
                                  until the result is copied in.  Undefined after a
                                  failed or cancelled job (some pages may hold
                                  zeros); untouched by the server once the job is
                                  reported finished                                 */
} hv_job;

typedef struct hv_job_info {
    int64_t id;
    int status;                /* HV_OK or the HV_ERR_* code of the failure        */
    int slot;
    int graph_hit;             /* 1 if a previously captured graph was replayed    */
    int reserved;
    double t_submit, t_admit, t_done;   /* seconds on the CLOCK_MONOTONIC clock   */
    int64_t footprint;         /* device bytes of the slot the job needed          */
    double flops;              /* useful Hermitian flops 4 N (N+1) K T F           */
    char error[160];
} hv_job_info;

typedef struct hv_server_stats_t {
    int64_t submitted, completed, failed;   /* failed includes cancelled jobs      */
    int64_t graph_captures, graph_launches, direct_launches;
    int64_t peak_running, pending, running;
    size_t workspace_bytes, budget_bytes;    /* device bytes of all instances      */
    int64_t cancelled;                       /* hv_server_cancel calls that took    */
    int64_t instances;                       /* bucket instances held now           */
    int64_t evictions;                       /* idle instances dropped for room     */
} hv_server_stats_t;

/* nslots in [1,16]; budget = device bytes for instances (0 = half of what
 * the context has left); use_graphs = 0 launches kernels directly (same
 * instances, no capture). */
int hv_server_create(hv_ctx* ctx, int nslots, size_t budget, int use_graphs, hv_server** out);
/* Queue a job (non-blocking).  Shape errors return HV_ERR_ARG; a job that can
 * never fit the budget returns HV_ERR_NOMEM. */
int hv_server_submit(hv_server* srv, const hv_job* job, int64_t* id);
/* Collect up to max finished jobs (each is reported once). */
int hv_server_poll(hv_server* srv, hv_job_info* infos, int max, int* n);
/* Block until job id finishes (timeout_s < 0: no limit); HV_ERR_BUSY on timeout. */
int hv_server_wait(hv_server* srv, int64_t id, double timeout_s, hv_job_info* info);
/* Cancel job id.  A queued job finishes at once with HV_ERR_CANCELED; a running
 * one finishes with HV_ERR_CANCELED when its device work ends (the kernels are
 * not interrupted) and its result is not copied.  HV_ERR_ARG if the job is not
 * queued or running (finished, collected or unknown). */
int hv_server_cancel(hv_server* srv, int64_t id);
int hv_server_stats(hv_server* srv, hv_server_stats_t* out);
/* Fail queued jobs, wait for running ones, release everything. */
int hv_server_destroy(hv_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* HERAVIS_H */
