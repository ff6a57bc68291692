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
    R* scratch;                           // [grid][nbls][5]: model (2), dwgts (1), weighted product (2)
    R* chisq;                             // [ndata]
    uint32_t* iters;                      // [ndata]
    R* conv;                              // [ndata]
    int nbls, nants, nubls;
    int64_t ndata;
    int maxiter, check_every, check_after;
    R one_m, gs, crit;
};

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
// in order, then the pairwise tree.  Every thread gets the total.
template <typename R, typename F>
__device__ __forceinline__ R fixed_sum(R* red, int n, F v) {
#pragma clang fp contract(off)
    R acc = 0;
    for (int b = threadIdx.x; b < n; b += OMNI_THREADS) acc = acc + v(b);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = OMNI_THREADS / 2; s >= 1; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    const R total = red[0];
    __syncthreads();
    return total;
}

// model g_i conj(g_j) u_k of every baseline into the scratch slice
template <typename R>
__device__ __forceinline__ void gen_model(const OmniArgs<R>& a, const R* g, const R* u, R* sc) {
#pragma clang fp contract(off)
    for (int b = threadIdx.x; b < a.nbls; b += OMNI_THREADS) {
        const uint32_t i = a.ggu[3 * b], j = a.ggu[3 * b + 1], k = a.ggu[3 * b + 2];
        R tr, ti, mr, mi;
        cmul(g[2 * i], g[2 * i + 1], g[2 * j], -g[2 * j + 1], tr, ti);
        cmul(tr, ti, u[2 * k], u[2 * k + 1], mr, mi);
        sc[5 * b] = mr;
        sc[5 * b + 1] = mi;
    }
}

template <typename R>
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
    R* sc = a.scratch + (size_t)blockIdx.x * a.nbls * 5;
    const int tid = threadIdx.x;
    const int nsol = a.nants + a.nubls;

    for (int64_t px = blockIdx.x; px < a.ndata; px += gridDim.x) {
        const R* data = a.data + (size_t)px * a.nbls * 2;
        const R* wgts = a.wgts + (size_t)px * a.nbls;
        for (int n = tid; n < 2 * a.nants; n += OMNI_THREADS) g[n] = a.gains[(size_t)px * a.nants * 2 + n];
        for (int n = tid; n < 2 * a.nubls; n += OMNI_THREADS) u[n] = a.ubls[(size_t)px * a.nubls * 2 + n];
        __syncthreads();
        gen_model(a, g, u, sc);
        __syncthreads();
        auto resid = [&](int b) {
            const R er = data[2 * b] - sc[5 * b], ei = data[2 * b + 1] - sc[5 * b + 1];
            return mag2(er, ei) * wgts[b];
        };
        R chisq = fixed_sum(red, a.nbls, resid);
        uint32_t iters = 0;
        R conv = 0;
        for (int it = 1; it <= a.maxiter; ++it) {
            if (it % a.check_every == 1) {
                for (int b = tid; b < a.nbls; b += OMNI_THREADS) sc[5 * b + 2] = mag2(sc[5 * b], sc[5 * b + 1]) * wgts[b];
                __syncthreads();
                for (int n = tid; n < nsol; n += OMNI_THREADS) {
                    R s = 0;
                    if (n < a.nants) {
                        for (int q = a.iptr[n]; q < a.iptr[n + 1]; ++q) s = s + sc[5 * a.ib[q] + 2];
                        for (int q = a.jptr[n]; q < a.jptr[n + 1]; ++q) s = s + sc[5 * a.jb[q] + 2];
                        gw[n] = s;
                    } else {
                        const int k = n - a.nants;
                        for (int q = a.kptr[k]; q < a.kptr[k + 1]; ++q) s = s + sc[5 * a.kb[q] + 2];
                        uw[k] = s;
                    }
                }
            }
            // weighted products data conj(model) dwgts / |model|^2
            for (int b = tid; b < a.nbls; b += OMNI_THREADS) {
                const R mr = sc[5 * b], mi = sc[5 * b + 1];
                const R wt = sc[5 * b + 2] / mag2(mr, mi);
                R pr, pi;
                cmul(data[2 * b], data[2 * b + 1], mr, -mi, pr, pi);
                sc[5 * b + 3] = pr * wt;
                sc[5 * b + 4] = pi * wt;
            }
            __syncthreads();
            // candidate solution: x * ((1 - g) + g buf / wgt)
            for (int n = tid; n < nsol; n += OMNI_THREADS) {
                R br = 0, bi = 0;
                if (n < a.nants) {
                    for (int q = a.iptr[n]; q < a.iptr[n + 1]; ++q) {
                        br = br + sc[5 * a.ib[q] + 3];
                        bi = bi + sc[5 * a.ib[q] + 4];
                    }
                    for (int q = a.jptr[n]; q < a.jptr[n + 1]; ++q) {
                        br = br + sc[5 * a.jb[q] + 3];
                        bi = bi + -sc[5 * a.jb[q] + 4];
                    }
                    const R wr = a.one_m + a.gs * br / gw[n], wi = a.gs * bi / gw[n];
                    cmul(g[2 * n], g[2 * n + 1], wr, wi, gn[2 * n], gn[2 * n + 1]);
                } else {
                    const int k = n - a.nants;
                    for (int q = a.kptr[k]; q < a.kptr[k + 1]; ++q) {
                        br = br + sc[5 * a.kb[q] + 3];
                        bi = bi + sc[5 * a.kb[q] + 4];
                    }
                    const R wr = a.one_m + a.gs * br / uw[k], wi = a.gs * bi / uw[k];
                    cmul(u[2 * k], u[2 * k + 1], wr, wi, un[2 * k], un[2 * k + 1]);
                }
            }
            __syncthreads();
            const bool fast = it < a.maxiter && (it < a.check_after || it % a.check_every != 0);
            if (fast) {
                for (int n = tid; n < 2 * a.nants; n += OMNI_THREADS) g[n] = gn[n];
                for (int n = tid; n < 2 * a.nubls; n += OMNI_THREADS) u[n] = un[n];
                __syncthreads();
                gen_model(a, g, u, sc);
                __syncthreads();
                continue;
            }
            gen_model(a, gn, un, sc);
            __syncthreads();
            const R new_chisq = fixed_sum(red, a.nbls, resid);
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

template <typename R>
int omnical_t(hv_ctx* c, const uint32_t* ggu, int nbls, void* gains, int nants, void* ubls, int nubls,
              const void* data, const void* wgts, int64_t ndata, double conv_crit, int maxiter, int check_every,
              int check_after, double gain, void* chisq, uint32_t* iters, void* conv) {
    const size_t lds = ((size_t)4 * nants + 4 * nubls + nants + nubls + OMNI_THREADS) * sizeof(R);
    if (lds > 160 * 1024)
        return fail(HV_ERR_ARG, "omnical: gains, unique baselines and their candidates must fit in LDS "
                                "(5 (nants + nubls) + 256 values <= 160 KB)");
    std::vector<int> iptr, ib, jptr, jb, kptr, kb;
    csr(ggu, nbls, 0, nants, iptr, ib);
    csr(ggu, nbls, 1, nants, jptr, jb);
    csr(ggu, nbls, 2, nubls, kptr, kb);
    int dev = 0, ncu = 0, per_cu = 0;
    HV_HIP(hipGetDevice(&dev));
    HV_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (lds > 64 * 1024)
        HV_HIP(hipFuncSetAttribute((const void*)omnical_kernel<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    HV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, omnical_kernel<R>, OMNI_THREADS, lds));
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(ndata, (int64_t)ncu * std::max(per_cu, 1)));

    const size_t b_ggu = (size_t)nbls * 3 * 4, b_csr = ((size_t)2 * (nants + 1) + nubls + 1 + 3 * (size_t)nbls) * 4;
    const size_t b_data = (size_t)ndata * nbls * 2 * sizeof(R), b_wgts = (size_t)ndata * nbls * sizeof(R);
    const size_t b_g = (size_t)ndata * nants * 2 * sizeof(R), b_u = (size_t)ndata * nubls * 2 * sizeof(R);
    const size_t b_sc = (size_t)grid * nbls * 5 * sizeof(R), b_out = (size_t)ndata * (2 * sizeof(R) + 4);
    char* base = nullptr;
    const size_t total = b_ggu + b_csr + b_data + b_wgts + b_g + b_u + b_sc + b_out + 16 * 256;   // + alignment of 10 slices
    HV_CHECK(dev_alloc(c, (void**)&base, total));
    auto carve = [&](size_t bytes) {
        char* p = base;
        base += (bytes + 255) / 256 * 256;
        return p;
    };
    char* const mem = base;
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
    std::vector<int> all;
    all.insert(all.end(), iptr.begin(), iptr.end());
    all.insert(all.end(), jptr.begin(), jptr.end());
    all.insert(all.end(), kptr.begin(), kptr.end());
    all.insert(all.end(), ib.begin(), ib.end());
    all.insert(all.end(), jb.begin(), jb.end());
    all.insert(all.end(), kb.begin(), kb.end());
    hipStream_t s = c->stream;
    int rc = HV_OK;
    auto run = [&]() -> int {
        HV_HIP(hipMemcpyAsync(d_ggu, ggu, b_ggu, hipMemcpyHostToDevice, s));
        HV_HIP(hipMemcpyAsync(d_csr, all.data(), all.size() * 4, hipMemcpyHostToDevice, s));
        HV_HIP(hipMemcpyAsync(d_data, data, b_data, hipMemcpyHostToDevice, s));
        HV_HIP(hipMemcpyAsync(d_wgts, wgts, b_wgts, hipMemcpyHostToDevice, s));
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
        hipLaunchKernelGGL((omnical_kernel<R>), dim3((unsigned)grid), dim3(OMNI_THREADS), lds, s, a);
        HV_HIP(hipGetLastError());
        HV_HIP(hipMemcpyAsync(gains, d_g, b_g, hipMemcpyDeviceToHost, s));
        HV_HIP(hipMemcpyAsync(ubls, d_u, b_u, hipMemcpyDeviceToHost, s));
        HV_HIP(hipMemcpyAsync(chisq, d_chisq, ndata * sizeof(R), hipMemcpyDeviceToHost, s));
        HV_HIP(hipMemcpyAsync(conv, d_conv, ndata * sizeof(R), hipMemcpyDeviceToHost, s));
        HV_HIP(hipMemcpyAsync(iters, d_iters, ndata * 4, hipMemcpyDeviceToHost, s));
        HV_HIP(hipStreamSynchronize(s));
        return HV_OK;
    };
    rc = run();
    dev_free(c, mem, total);
    return rc;
}

}  // namespace

extern "C" int hv_omnical(hv_ctx* c, int precision, const uint32_t* ggu_indices, int nbls, void* gains, int nants,
                          void* ubls, int nubls, const void* data, const void* wgts, int64_t ndata,
                          double conv_crit, int maxiter, int check_every, int check_after, double gain,
                          void* chisq, uint32_t* iters, void* conv) {
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
                                maxiter, check_every, check_after, gain, chisq, iters, conv);
    return omnical_t<double>(c, ggu_indices, nbls, gains, nants, ubls, nubls, data, wgts, ndata, conv_crit, maxiter,
                             check_every, check_after, gain, chisq, iters, conv);
}
