// Omnical redundant calibration (hera_gpu/redcal.py:358-743 and the kernels
// of redcal.py:70-355) for many independent pixels (time, frequency).
//
// The reference launches 8-12 small kernels per iteration for chunks of 64
// pixels, accumulates with global atomics and copies the active mask to the
// host at every convergence check.  Here one workgroup solves one pixel from
// start to finish inside a single launch: its gains and unique baselines live
// in LDS, the per-baseline model and weighted products in a per-workgroup
// scratch slice, and every sum runs in a fixed order without atomics:
//   * per antenna, its baselines as first index in baseline order, then as
//     second index (conjugated) in baseline order (CSR lists built on the
//     host); per unique baseline, its baselines in order;
//   * pixel sums (chi^2, convergence) as 256 strided sequential partial sums
//     and a pairwise tree over them.
// Complex products follow cuCmul's component order with contraction off, so
// results are deterministic and follow the CPU restatement of the oracle.
// A pixel stops when the relative change of its solution falls below
// conv_crit or chi^2 rises (it then keeps the previous solution), as
// redcal.py:317-355 intends (without the cross-block race of update_active).
#include <algorithm>
#include <vector>

#include "hv_host.h"

using namespace hv;

namespace {

constexpr int OMNI_THREADS = 256;

template <typename R>
struct OmniArgs {
    const uint32_t* ggu;                  // [nbls][3] (i, j, k)
    const int *iptr, *ib;                 // CSR: baselines with first antenna a
    const int *jptr, *jb;                 // CSR: baselines with second antenna a
    const int *kptr, *kb;                 // CSR: baselines of unique baseline k
    const R* data;                        // [ndata][nbls] complex (re, im)
    const R* wgts;                        // [ndata][nbls]
    R* gains;                             // [ndata][nants] complex, in/out
    R* ubls;                              // [ndata][nubls] complex, in/out
    R* scratch;                           // per workgroup: the per-baseline arrays kept outside LDS
    R* chisq;                             // [ndata]
    uint32_t* iters;                      // [ndata]
    R* conv;                              // [ndata]
    int nbls, nants, nubls;
    int64_t ndata;
    int maxiter, check_every, check_after;
    R one_m, gs, crit;
};

// Where a workgroup keeps its per-baseline arrays: the weighted products
// (2 values per baseline, gathered per antenna and unique baseline) and the
// data weights |model|^2 w of the last check (1 value).  Chosen per call from
// what fits in LDS beside the gains and unique baselines.
enum OmniPlace { OMNI_GLOBAL = 0, OMNI_PROD_LDS = 1, OMNI_ALL_LDS = 2 };

template <typename R>
__device__ __forceinline__ R mag2(R x, R y) {
#pragma clang fp contract(off)
    return x * x + y * y;
}

template <typename R>
__device__ __forceinline__ void cmul(R ar, R ai, R br, R bi, R& cr, R& ci) {
#pragma clang fp contract(off)
    cr = ar * br - ai * bi;
    ci = ar * bi + ai * br;
}

// Sum of one value per thread-strided index: partial p = v(p) + v(p+256) + ...
// in order, then the pairwise tree p[i] += p[i + s], s = 128, 64, ..., 1 (the
// last six levels inside wave 0 by lane shuffles, the same additions).
// Every thread gets the total.
template <typename R, typename F>
__device__ __forceinline__ R fixed_sum(R* red, int n, F v) {
#pragma clang fp contract(off)
    R acc = 0;
#pragma unroll 4
    for (int b = threadIdx.x; b < n; b += OMNI_THREADS) acc = acc + v(b);
    const int t = threadIdx.x;
    red[t] = acc;
    __syncthreads();
    if (t < 128) red[t] = red[t] + red[t + 128];
    __syncthreads();
    if (t < 64) {
        R x = red[t] + red[t + 64];
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) x = x + __shfl_down(x, s, 64);
        if (t == 0) red[0] = x;
    }
    __syncthreads();
    const R total = red[0];
    __syncthreads();
    return total;
}

// g_i conj(g_j) u_k of baseline b (redcal.py:70-93), G and U in LDS
template <typename R>
__device__ __forceinline__ void model_of(const OmniArgs<R>& a, int b, const R* G, const R* U, R& mr, R& mi) {
#pragma clang fp contract(off)
    const uint32_t i = a.ggu[3 * b], j = a.ggu[3 * b + 1], k = a.ggu[3 * b + 2];
    R tr, ti;
    cmul(G[2 * i], G[2 * i + 1], G[2 * j], -G[2 * j + 1], tr, ti);
    cmul(tr, ti, U[2 * k], U[2 * k + 1], mr, mi);
}

// s + v[idx[q0]] + v[idx[q0 + 1]] + ... + v[idx[q1 - 1]] in that order; the
// indices and values of 8 terms are loaded before they are added.
template <typename R>
__device__ __forceinline__ R gather_sum(const int* idx, int q0, int q1, const R* v, R s) {
#pragma clang fp contract(off)
    int q = q0;
    for (; q + 8 <= q1; q += 8) {
        int b[8];
        R x[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) b[t] = idx[q + t];
#pragma unroll
        for (int t = 0; t < 8; ++t) x[t] = v[b[t]];
#pragma unroll
        for (int t = 0; t < 8; ++t) s = s + x[t];
    }
    for (; q < q1; ++q) s = s + v[idx[q]];
    return s;
}

// The same over complex pairs v[2b], v[2b + 1]; NEG: the imaginary parts enter negated.
template <typename R, bool NEG>
__device__ __forceinline__ void gather_sum2(const int* idx, int q0, int q1, const R* v, R& sr, R& si) {
#pragma clang fp contract(off)
    int q = q0;
    for (; q + 8 <= q1; q += 8) {
        int b[8];
        R xr[8], xi[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) b[t] = idx[q + t];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            xr[t] = v[2 * b[t]];
            xi[t] = v[2 * b[t] + 1];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            sr = sr + xr[t];
            si = NEG ? si + -xi[t] : si + xi[t];
        }
    }
    for (; q < q1; ++q) {
        const int b = idx[q];
        sr = sr + v[2 * b];
        si = NEG ? si + -v[2 * b + 1] : si + v[2 * b + 1];
    }
}

template <typename R, int PLACE>
__global__ __launch_bounds__(OMNI_THREADS) void omnical_kernel(OmniArgs<R> a) {
#pragma clang fp contract(off)
    extern __shared__ __attribute__((aligned(16))) char smem[];
    R* g = reinterpret_cast<R*>(smem);   // [nants][2]
    R* gn = g + 2 * a.nants;             // [nants][2] candidate
    R* u = gn + 2 * a.nants;             // [nubls][2]
    R* un = u + 2 * a.nubls;             // [nubls][2] candidate
    R* gw = un + 2 * a.nubls;            // [nants]
    R* uw = gw + a.nants;                // [nubls]
    R* red = uw + a.nubls;               // [256]
    R* lds_b = red + OMNI_THREADS;       // per-baseline arrays held in LDS
    // weighted products [nbls][2] and weights [nbls]
    constexpr int NGLOBAL = PLACE == OMNI_GLOBAL ? 3 : PLACE == OMNI_PROD_LDS ? 1 : 0;
    R* gsc = a.scratch + (size_t)blockIdx.x * a.nbls * NGLOBAL;
    R* prod = PLACE == OMNI_GLOBAL ? gsc + a.nbls : lds_b;
    R* dw = PLACE == OMNI_ALL_LDS ? lds_b + 2 * a.nbls : gsc;
    const int tid = threadIdx.x;
    const int nsol = a.nants + a.nubls;

    for (int64_t px = blockIdx.x; px < a.ndata; px += gridDim.x) {
        const R* data = a.data + (size_t)px * a.nbls * 2;
        const R* wgts = a.wgts + (size_t)px * a.nbls;
        for (int n = tid; n < 2 * a.nants; n += OMNI_THREADS) g[n] = a.gains[(size_t)px * a.nants * 2 + n];
        for (int n = tid; n < 2 * a.nubls; n += OMNI_THREADS) u[n] = a.ubls[(size_t)px * a.nubls * 2 + n];
        __syncthreads();
        // chi^2 of the model of (G, U) (redcal.py:96-114)
        auto chisq_of = [&](const R* G, const R* U) {
            return fixed_sum(red, a.nbls, [&](int b) {
                R mr, mi;
                model_of(a, b, G, U, mr, mi);
                const R er = data[2 * b] - mr, ei = data[2 * b + 1] - mi;
                return mag2(er, ei) * wgts[b];
            });
        };
        R chisq = chisq_of(g, u);
        uint32_t iters = 0;
        R conv = 0;
        for (int it = 1; it <= a.maxiter; ++it) {
            if (it % a.check_every == 1) {
                // data weights |model|^2 w and their sums per antenna and
                // unique baseline (redcal.py:117-163)
                for (int b = tid; b < a.nbls; b += OMNI_THREADS) {
                    R mr, mi;
                    model_of(a, b, g, u, mr, mi);
                    dw[b] = mag2(mr, mi) * wgts[b];
                }
                __syncthreads();
                for (int n = tid; n < nsol; n += OMNI_THREADS) {
                    R s = 0;
                    if (n < a.nants) {
                        s = gather_sum(a.ib, a.iptr[n], a.iptr[n + 1], dw, s);
                        gw[n] = gather_sum(a.jb, a.jptr[n], a.jptr[n + 1], dw, s);
                    } else {
                        const int k = n - a.nants;
                        uw[k] = gather_sum(a.kb, a.kptr[k], a.kptr[k + 1], dw, s);
                    }
                }
            }
            // weighted products data conj(model) dwgts / |model|^2 (redcal.py:166-199)
            for (int b0 = tid; b0 < a.nbls; b0 += 4 * OMNI_THREADS) {
                R pr[4], pi[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int b = b0 + t * OMNI_THREADS;
                    if (b >= a.nbls) break;
                    R mr, mi;
                    model_of(a, b, g, u, mr, mi);
                    const R wt = dw[b] / mag2(mr, mi);
                    cmul(data[2 * b], data[2 * b + 1], mr, -mi, pr[t], pi[t]);
                    pr[t] = pr[t] * wt;
                    pi[t] = pi[t] * wt;
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int b = b0 + t * OMNI_THREADS;
                    if (b >= a.nbls) break;
                    prod[2 * b] = pr[t];
                    prod[2 * b + 1] = pi[t];
                }
            }
            __syncthreads();
            // candidate solution: x * ((1 - g) + g buf / wgt) (redcal.py:202-314)
            for (int n = tid; n < nsol; n += OMNI_THREADS) {
                R br = 0, bi = 0;
                if (n < a.nants) {
                    gather_sum2<R, false>(a.ib, a.iptr[n], a.iptr[n + 1], prod, br, bi);
                    gather_sum2<R, true>(a.jb, a.jptr[n], a.jptr[n + 1], prod, br, bi);
                    const R wr = a.one_m + a.gs * br / gw[n], wi = a.gs * bi / gw[n];
                    cmul(g[2 * n], g[2 * n + 1], wr, wi, gn[2 * n], gn[2 * n + 1]);
                } else {
                    const int k = n - a.nants;
                    gather_sum2<R, false>(a.kb, a.kptr[k], a.kptr[k + 1], prod, br, bi);
                    const R wr = a.one_m + a.gs * br / uw[k], wi = a.gs * bi / uw[k];
                    cmul(u[2 * k], u[2 * k + 1], wr, wi, un[2 * k], un[2 * k + 1]);
                }
            }
            __syncthreads();
            const bool fast = it < a.maxiter && (it < a.check_after || it % a.check_every != 0);
            if (!fast) {
                const R new_chisq = chisq_of(gn, un);
                auto dsol = [&](int n) {
                    return n < a.nants ? mag2(gn[2 * n] - g[2 * n], gn[2 * n + 1] - g[2 * n + 1])
                                       : mag2(un[2 * (n - a.nants)] - u[2 * (n - a.nants)],
                                              un[2 * (n - a.nants) + 1] - u[2 * (n - a.nants) + 1]);
                };
                auto nsq = [&](int n) {
                    return n < a.nants ? mag2(gn[2 * n], gn[2 * n + 1])
                                       : mag2(un[2 * (n - a.nants)], un[2 * (n - a.nants) + 1]);
                };
                const R csum = fixed_sum(red, nsol, dsol);
                const R cwgt = fixed_sum(red, nsol, nsq);
                conv = sqrt(csum / cwgt);
                if (conv < a.crit || new_chisq > chisq) break;   // uniform: keep the previous solution
                chisq = new_chisq;
                iters = (uint32_t)it;
            }
            for (int n = tid; n < 2 * a.nants; n += OMNI_THREADS) g[n] = gn[n];
            for (int n = tid; n < 2 * a.nubls; n += OMNI_THREADS) u[n] = un[n];
            __syncthreads();
        }
        for (int n = tid; n < 2 * a.nants; n += OMNI_THREADS) a.gains[(size_t)px * a.nants * 2 + n] = g[n];
        for (int n = tid; n < 2 * a.nubls; n += OMNI_THREADS) a.ubls[(size_t)px * a.nubls * 2 + n] = u[n];
        if (tid == 0) {
            a.chisq[px] = chisq;
            a.iters[px] = iters;
            a.conv[px] = conv;
        }
        __syncthreads();   // LDS and scratch are reused by the next pixel
    }
}

// CSR lists of baselines per key (baselines in increasing order).
void csr(const uint32_t* ggu, int nbls, int col, int nkeys, std::vector<int>& ptr, std::vector<int>& idx) {
    ptr.assign(nkeys + 1, 0);
    for (int b = 0; b < nbls; ++b) ++ptr[ggu[3 * b + col] + 1];
    for (int k = 0; k < nkeys; ++k) ptr[k + 1] += ptr[k];
    idx.resize(nbls);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int b = 0; b < nbls; ++b) idx[fill[ggu[3 * b + col]]++] = b;
}

int g_omni_place = -1;   // process-wide override (hv_set_omnical_place), -1: automatic

template <typename R, int PLACE>
int omni_occupancy(size_t lds, int* per_cu) {
    if (lds > 64 * 1024)
        HV_HIP(hipFuncSetAttribute((const void*)omnical_kernel<R, PLACE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    HV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, omnical_kernel<R, PLACE>, OMNI_THREADS, lds));
    return HV_OK;
}

// One solve of every pixel.  reps > 0 (timing): the solve runs 1 + reps times
// from the same starting point, *ms = mean kernel time of the last reps.
template <typename R>
int omnical_t(hv_ctx* c, const uint32_t* ggu, int nbls, void* gains, int nants, void* ubls, int nubls,
              const void* data, const void* wgts, int64_t ndata, double conv_crit, int maxiter, int check_every,
              int check_after, double gain, void* chisq, uint32_t* iters, void* conv, int reps = 0,
              double* ms = nullptr) {
    constexpr size_t LDS_MAX = 160 * 1024;
    const size_t lds_base = ((size_t)4 * nants + 4 * nubls + nants + nubls + OMNI_THREADS) * sizeof(R);
    if (lds_base > LDS_MAX)
        return fail(HV_ERR_ARG, "omnical: gains, unique baselines and their candidates must fit in LDS "
                                "(5 (nants + nubls) + 256 values <= 160 KB)");
    int dev = 0, ncu = 0;
    HV_HIP(hipGetDevice(&dev));
    HV_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // Per-baseline arrays in LDS as far as they fit, unless that leaves fewer
    // than two resident workgroups per CU for a batch that fills the device:
    // the solve is latency-bound, and on MI355X more resident pixels win over
    // LDS placement (121 antennas, 2000 pixels: float 7.9 ms with the products
    // in LDS at 2 per CU against 10.9 ms all in LDS at 1; double 15.5 ms in
    // device memory at 3 per CU against 19.2 ms at 1).
    // values per baseline held in LDS: 3 (all), 2 (the products) or none
    auto lds_of = [&](int pl) {
        return lds_base + (size_t)(pl == OMNI_ALL_LDS ? 3 : pl == OMNI_PROD_LDS ? 2 : 0) * nbls * sizeof(R);
    };
    auto occupancy = [&](int pl, int* per_cu) -> int {
        const size_t l = lds_of(pl);
        if (pl == OMNI_ALL_LDS) return omni_occupancy<R, OMNI_ALL_LDS>(l, per_cu);
        if (pl == OMNI_PROD_LDS) return omni_occupancy<R, OMNI_PROD_LDS>(l, per_cu);
        return omni_occupancy<R, OMNI_GLOBAL>(l, per_cu);
    };
    int place = lds_of(OMNI_ALL_LDS) <= LDS_MAX ? OMNI_ALL_LDS : lds_of(OMNI_PROD_LDS) <= LDS_MAX ? OMNI_PROD_LDS
                                                                                               : OMNI_GLOBAL;
    if (g_omni_place >= 0) place = std::min(place, g_omni_place);
    int per_cu = 0;
    HV_CHECK(occupancy(place, &per_cu));
    if (g_omni_place < 0)
        while (place > OMNI_GLOBAL && per_cu < 2 && ndata > ncu) HV_CHECK(occupancy(--place, &per_cu));
    const int nglobal = place == OMNI_GLOBAL ? 3 : place == OMNI_PROD_LDS ? 1 : 0;
    const size_t lds = lds_of(place);
    std::vector<int> iptr, ib, jptr, jb, kptr, kb;
    csr(ggu, nbls, 0, nants, iptr, ib);
    csr(ggu, nbls, 1, nants, jptr, jb);
    csr(ggu, nbls, 2, nubls, kptr, kb);
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(ndata, (int64_t)ncu * std::max(per_cu, 1)));

    const size_t b_ggu = (size_t)nbls * 3 * 4, b_csr = ((size_t)2 * (nants + 1) + nubls + 1 + 3 * (size_t)nbls) * 4;
    const size_t b_data = (size_t)ndata * nbls * 2 * sizeof(R), b_wgts = (size_t)ndata * nbls * sizeof(R);
    const size_t b_g = (size_t)ndata * nants * 2 * sizeof(R), b_u = (size_t)ndata * nubls * 2 * sizeof(R);
    const size_t b_sc = std::max<size_t>(1, (size_t)grid * nbls * nglobal) * sizeof(R);
    const size_t b_start = reps > 0 ? b_g + b_u + 512 : 0;   // starting points kept for repeated solves
    const size_t total = b_ggu + b_csr + b_data + b_wgts + b_g + b_u + b_sc + (size_t)ndata * (2 * sizeof(R) + 4) +
                         16 * 256 + b_start;   // + alignment of 10 slices
    HV_CHECK(ensure(c, c->ws, B_EXT, total, c->stream));
    char* base = static_cast<char*>(c->ws.buf[B_EXT].p);
    auto carve = [&](size_t bytes) {
        char* p = base;
        base += (bytes + 255) / 256 * 256;
        return p;
    };
    uint32_t* d_ggu = (uint32_t*)carve(b_ggu);
    int* d_csr = (int*)carve(b_csr);
    R* d_data = (R*)carve(b_data);
    R* d_wgts = (R*)carve(b_wgts);
    R* d_g = (R*)carve(b_g);
    R* d_u = (R*)carve(b_u);
    R* d_sc = (R*)carve(b_sc);
    R* d_chisq = (R*)carve(ndata * sizeof(R));
    R* d_conv = (R*)carve(ndata * sizeof(R));
    uint32_t* d_iters = (uint32_t*)carve(ndata * 4);
    R* d_g0 = reps > 0 ? (R*)carve(b_g) : nullptr;
    R* d_u0 = reps > 0 ? (R*)carve(b_u) : nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<int> all;
    all.insert(all.end(), iptr.begin(), iptr.end());
    all.insert(all.end(), jptr.begin(), jptr.end());
    all.insert(all.end(), kptr.begin(), kptr.end());
    all.insert(all.end(), ib.begin(), ib.end());
    all.insert(all.end(), jb.begin(), jb.end());
    all.insert(all.end(), kb.begin(), kb.end());
    hipStream_t s = c->stream;
    auto run = [&]() -> int {
        HV_HIP(hipMemcpyAsync(d_ggu, ggu, b_ggu, hipMemcpyHostToDevice, s));
        HV_HIP(hipMemcpyAsync(d_csr, all.data(), all.size() * 4, hipMemcpyHostToDevice, s));
        HV_CHECK(copy_h2d(c, d_data, data, b_data, s));
        HV_CHECK(copy_h2d(c, d_wgts, wgts, b_wgts, s));
        HV_HIP(hipMemcpyAsync(d_g, gains, b_g, hipMemcpyHostToDevice, s));
        HV_HIP(hipMemcpyAsync(d_u, ubls, b_u, hipMemcpyHostToDevice, s));
        OmniArgs<R> a{};
        a.ggu = d_ggu;
        a.iptr = d_csr;
        a.jptr = a.iptr + nants + 1;
        a.kptr = a.jptr + nants + 1;
        a.ib = a.kptr + nubls + 1;
        a.jb = a.ib + nbls;
        a.kb = a.jb + nbls;
        a.data = d_data;
        a.wgts = d_wgts;
        a.gains = d_g;
        a.ubls = d_u;
        a.scratch = d_sc;
        a.chisq = d_chisq;
        a.iters = d_iters;
        a.conv = d_conv;
        a.nbls = nbls;
        a.nants = nants;
        a.nubls = nubls;
        a.ndata = ndata;
        a.maxiter = maxiter;
        a.check_every = check_every;
        a.check_after = check_after;
        const float gs = (float)gain;
        a.one_m = (R)(1.0f - gs);          // (1 - gstep) in float32, redcal.py:274
        a.gs = (R)gs;
        a.crit = (R)conv_crit;
        const dim3 gr((unsigned)grid), bl(OMNI_THREADS);
        auto launch = [&]() -> int {
            if (place == OMNI_ALL_LDS) hipLaunchKernelGGL((omnical_kernel<R, OMNI_ALL_LDS>), gr, bl, lds, s, a);
            else if (place == OMNI_PROD_LDS) hipLaunchKernelGGL((omnical_kernel<R, OMNI_PROD_LDS>), gr, bl, lds, s, a);
            else hipLaunchKernelGGL((omnical_kernel<R, OMNI_GLOBAL>), gr, bl, lds, s, a);
            HV_HIP(hipGetLastError());
            return HV_OK;
        };
        if (reps > 0) {
            HV_HIP(hipEventCreate(&e0));
            HV_HIP(hipEventCreate(&e1));
            HV_HIP(hipMemcpyAsync(d_g0, d_g, b_g, hipMemcpyDeviceToDevice, s));
            HV_HIP(hipMemcpyAsync(d_u0, d_u, b_u, hipMemcpyDeviceToDevice, s));
            HV_CHECK(launch());   // warm-up
            double total_ms = 0;
            for (int r = 0; r < reps; ++r) {
                HV_HIP(hipMemcpyAsync(d_g, d_g0, b_g, hipMemcpyDeviceToDevice, s));
                HV_HIP(hipMemcpyAsync(d_u, d_u0, b_u, hipMemcpyDeviceToDevice, s));
                HV_HIP(hipEventRecord(e0, s));
                HV_CHECK(launch());
                HV_HIP(hipEventRecord(e1, s));
                HV_HIP(hipEventSynchronize(e1));
                float t = 0;
                HV_HIP(hipEventElapsedTime(&t, e0, e1));
                total_ms += t;
            }
            *ms = total_ms / reps;
        } else {
            HV_CHECK(launch());
        }
        HV_HIP(hipMemcpyAsync(gains, d_g, b_g, hipMemcpyDeviceToHost, s));
        HV_HIP(hipMemcpyAsync(ubls, d_u, b_u, hipMemcpyDeviceToHost, s));
        HV_HIP(hipMemcpyAsync(chisq, d_chisq, ndata * sizeof(R), hipMemcpyDeviceToHost, s));
        HV_HIP(hipMemcpyAsync(conv, d_conv, ndata * sizeof(R), hipMemcpyDeviceToHost, s));
        HV_HIP(hipMemcpyAsync(iters, d_iters, ndata * 4, hipMemcpyDeviceToHost, s));
        HV_HIP(hipStreamSynchronize(s));
        return HV_OK;
    };
    const int rc = run();
    if (rc != HV_OK) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace

extern "C" int hv_set_omnical_place(int place) {
    if (place < -1 || place > 2) return fail(HV_ERR_ARG, "omnical placement must be -1, 0, 1 or 2");
    g_omni_place = place;
    return HV_OK;
}

static int omnical_entry(hv_ctx* c, int precision, const uint32_t* ggu_indices, int nbls, void* gains, int nants,
                         void* ubls, int nubls, const void* data, const void* wgts, int64_t ndata, double conv_crit,
                         int maxiter, int check_every, int check_after, double gain, void* chisq, uint32_t* iters,
                         void* conv, int reps, double* ms) {
    if (!c || !ggu_indices || !gains || !ubls || !data || !wgts || !chisq || !iters || !conv)
        return fail(HV_ERR_ARG, "NULL argument");
    if (precision != 1 && precision != 2) return fail(HV_ERR_ARG, "precision must be 1 or 2");
    if (nbls < 1 || nants < 1 || nubls < 1 || ndata < 1 || maxiter < 1 || check_after < 0)
        return fail(HV_ERR_ARG, "bad omnical shape or iteration arguments");
    if (check_every <= 1) return fail(HV_ERR_ARG, "check_every must be > 1");
    for (int b = 0; b < nbls; ++b)
        if (ggu_indices[3 * b] >= (uint32_t)nants || ggu_indices[3 * b + 1] >= (uint32_t)nants ||
            ggu_indices[3 * b + 2] >= (uint32_t)nubls)
            return fail(HV_ERR_ARG, "ggu_indices out of range");
    std::lock_guard<std::mutex> lk(c->mu);
    HV_HIP(hipSetDevice(c->device));
    if (precision == 1)
        return omnical_t<float>(c, ggu_indices, nbls, gains, nants, ubls, nubls, data, wgts, ndata, conv_crit,
                                maxiter, check_every, check_after, gain, chisq, iters, conv, reps, ms);
    return omnical_t<double>(c, ggu_indices, nbls, gains, nants, ubls, nubls, data, wgts, ndata, conv_crit, maxiter,
                             check_every, check_after, gain, chisq, iters, conv, reps, ms);
}

extern "C" int hv_omnical(hv_ctx* c, int precision, const uint32_t* ggu_indices, int nbls, void* gains, int nants,
                          void* ubls, int nubls, const void* data, const void* wgts, int64_t ndata,
                          double conv_crit, int maxiter, int check_every, int check_after, double gain,
                          void* chisq, uint32_t* iters, void* conv) {
    return omnical_entry(c, precision, ggu_indices, nbls, gains, nants, ubls, nubls, data, wgts, ndata, conv_crit,
                         maxiter, check_every, check_after, gain, chisq, iters, conv, 0, nullptr);
}

extern "C" int hv_bench_omnical(hv_ctx* c, int precision, const uint32_t* ggu_indices, int nbls, void* gains,
                                int nants, void* ubls, int nubls, const void* data, const void* wgts, int64_t ndata,
                                double conv_crit, int maxiter, int check_every, int check_after, double gain,
                                void* chisq, uint32_t* iters, void* conv, int reps, double* ms_per_solve) {
    if (reps < 1 || !ms_per_solve) return fail(HV_ERR_ARG, "bad argument");
    return omnical_entry(c, precision, ggu_indices, nbls, gains, nants, ubls, nubls, data, wgts, ndata, conv_crit,
                         maxiter, check_every, check_after, gain, chisq, iters, conv, reps, ms_per_solve);
}
