"""CPU restatement of the reference's GPU Omnical solver (test infrastructure:
only tests, smoke() and bench.py's CPU leg use it).

Follows hera_gpu/redcal.py:358-743 (host loop) and the kernels of
redcal.py:70-355 (gen_dmdl, calc_chisq, calc_dwgts, calc_gu_wgt, calc_gu_buf,
update_gains, update_ubls, calc_conv, update_active), one independent
problem per pixel, with the arithmetic made deterministic:

  * complex products are formed component by component in the working
    precision, in the order of cuCmul ((ax bx - ay by), (ax by + ay bx)),
    each operation rounded on its own (no fused multiply-add);
  * the reference's atomic accumulations become sums in a fixed order:
    per antenna, its baselines as first index in baseline order, then as
    second index in baseline order; per unique baseline, its baselines in
    order; pixel sums (chi^2, convergence) as 256 strided sequential partial
    sums followed by a pairwise tree (``fixed_sum``);
  * update_active keeps its intent without the cross-block race of the
    reference (redcal.py:345): a pixel that stops (converged or diverged)
    keeps its previous gains, unique baselines and chi^2.

The reference computes (1 - gstep) with gstep a float32 (redcal.py:274,
passed as np.float32(gain) at :618) before promoting to the working
precision; that is kept.  iters and conv start at 0 (the reference leaves
them uninitialised).
"""
from __future__ import annotations

import numpy as np

NPART = 256   # partial sums of fixed_sum (the GPU workgroup size)


def fixed_sum(x: np.ndarray) -> np.ndarray:
    """Sum over the last axis: partial p = x[p] + x[p + 256] + ... in order,
    then partials pairwise (p += p + s for s = 128, 64, ..., 1).  x real."""
    n = x.shape[-1]
    m = -(-n // NPART) * NPART
    pad = np.zeros(x.shape[:-1] + (m,), dtype=x.dtype)
    pad[..., :n] = x
    rows = pad.reshape(x.shape[:-1] + (m // NPART, NPART))
    part = np.zeros(x.shape[:-1] + (NPART,), dtype=x.dtype)
    for r in range(rows.shape[-2]):
        part = part + rows[..., r, :]
    s = NPART // 2
    while s >= 1:
        part = part.copy()
        part[..., :s] = part[..., :s] + part[..., s:2 * s]
        s //= 2
    return part[..., 0]


def cmul(ar, ai, br, bi):
    """Complex product in the order of cuCmul, every operation rounded."""
    return ar * br - ai * bi, ar * bi + ai * br


def gen_dmdl(ggu, gr, gi, ur, ui):
    """g_i conj(g_j) u_k per baseline (redcal.py:70-93)."""
    i, j, k = ggu[:, 0], ggu[:, 1], ggu[:, 2]
    tr, ti = cmul(gr[:, i], gi[:, i], gr[:, j], -gi[:, j])
    return cmul(tr, ti, ur[:, k], ui[:, k])


def mag2(xr, xi):
    return xr * xr + xi * xi


def chisq_of(dr, di, mr, mi, w):
    """sum_b |data - dmdl|^2 wgts (redcal.py:96-114)."""
    er, ei = dr - mr, di - mi
    return fixed_sum(mag2(er, ei) * w)


def _gather(nout, idx, val):
    """out[idx[n]] += val[n] in order of n (idx sorted so that each output's
    terms come in the required order)."""
    out = np.zeros((val.shape[0], nout), dtype=val.dtype)
    for n in range(idx.size):
        out[:, idx[n]] = out[:, idx[n]] + val[:, n]
    return out


def omnical(ggu_indices, gains, ubls, data, wgts, conv_crit=1e-10, maxiter=50, check_every=4, check_after=1,
            gain=0.3, precision=1):
    """Same arguments and returns as hera_gpu.redcal.omnical (redcal.py:358)."""
    ggu = np.asarray(ggu_indices, dtype=np.int64)
    nbls = ggu.shape[0]
    ndata = data.shape[0]
    nants, nubls = gains.shape[1], ubls.shape[1]
    assert ggu.shape == (nbls, 3) and data.shape == (ndata, nbls) and wgts.shape == (ndata, nbls)
    assert gains.shape == (ndata, nants) and ubls.shape == (ndata, nubls) and precision in (1, 2)
    assert check_every > 1
    R = np.float32 if precision == 1 else np.float64
    gr, gi = np.real(gains).astype(R), np.imag(gains).astype(R)
    ur, ui = np.real(ubls).astype(R), np.imag(ubls).astype(R)
    dr, di = np.real(data).astype(R), np.imag(data).astype(R)
    w = np.asarray(wgts).astype(R)
    gstep = np.float32(gain)
    one_m = R(np.float32(1) - gstep)          # (1 - gstep) in float32 (redcal.py:274)
    gs = R(gstep)
    crit = R(conv_crit)
    # accumulation orders: antenna a takes its baselines as i (in order), then as j
    order_i = np.argsort(ggu[:, 0], kind="stable")
    order_j = np.argsort(ggu[:, 1], kind="stable")
    order_k = np.argsort(ggu[:, 2], kind="stable")
    ant_idx = np.concatenate([ggu[order_i, 0], ggu[order_j, 1]])
    # per antenna the i-terms must precede the j-terms: sort the concatenation by antenna, stable
    perm = np.argsort(ant_idx, kind="stable")
    src = np.concatenate([order_i, order_j])[perm]              # baseline of each term
    conj_term = np.concatenate([np.zeros(nbls, bool), np.ones(nbls, bool)])[perm]
    ant_of = ant_idx[perm]
    ubl_of = ggu[order_k, 2]

    def acc_ant(vr, vi):
        tr = vr[:, src]
        ti = np.where(conj_term[None, :], -vi[:, src], vi[:, src])
        return _gather(nants, ant_of, tr), _gather(nants, ant_of, ti)

    def acc_ant_real(v):
        return _gather(nants, ant_of, v[:, src])

    def acc_ubl(v):
        return _gather(nubls, ubl_of, v[:, order_k])

    mr, mi = gen_dmdl(ggu, gr, gi, ur, ui)
    chisq = chisq_of(dr, di, mr, mi, w)
    active = np.ones(ndata, bool)
    iters = np.zeros(ndata, np.uint32)
    conv = np.zeros(ndata, R)
    dw = np.zeros_like(w)
    gw = np.zeros((ndata, nants), R)
    uw = np.zeros((ndata, nubls), R)
    for it in range(1, maxiter + 1):
        a = active
        if not a.any():
            break
        if it % check_every == 1:
            dw[a] = mag2(mr[a], mi[a]) * w[a]
            gw[a] = acc_ant_real(dw[a])
            uw[a] = acc_ubl(dw[a])
        wt = dw[a] / mag2(mr[a], mi[a])
        pr, pi = cmul(dr[a], di[a], mr[a], -mi[a])
        pr, pi = pr * wt, pi * wt
        gbr, gbi = acc_ant(pr, pi)
        ubr, ubi = acc_ubl(pr), acc_ubl(pi)
        # wc = (1 - g) + g buf / wgt  (redcal.py:274-275, 296-297)
        gwr, gwi = one_m + gs * gbr / gw[a], gs * gbi / gw[a]
        uwr, uwi = one_m + gs * ubr / uw[a], gs * ubi / uw[a]
        ngr, ngi = cmul(gr[a], gi[a], gwr, gwi)
        nur, nui = cmul(ur[a], ui[a], uwr, uwi)
        if it < maxiter and (it < check_after or it % check_every != 0):
            gr[a], gi[a], ur[a], ui[a] = ngr, ngi, nur, nui
            mr[a], mi[a] = gen_dmdl(ggu, gr[a], gi[a], ur[a], ui[a])
            continue
        nmr, nmi = gen_dmdl(ggu, ngr, ngi, nur, nui)
        mr[a], mi[a] = nmr, nmi
        new_chisq = chisq_of(dr[a], di[a], nmr, nmi, w[a])
        csum = fixed_sum(np.concatenate([mag2(ngr - gr[a], ngi - gi[a]), mag2(nur - ur[a], nui - ui[a])], axis=1))
        cwgt = fixed_sum(np.concatenate([mag2(ngr, ngi), mag2(nur, nui)], axis=1))
        c = np.sqrt(csum / cwgt)
        idx = np.flatnonzero(a)
        conv[idx] = c
        stop_conv = c < crit
        stop_div = (~stop_conv) & (new_chisq > chisq[idx])
        keep = ~(stop_conv | stop_div)
        k = idx[keep]
        chisq[k] = new_chisq[keep]
        iters[k] = it
        gr[k], gi[k], ur[k], ui[k] = ngr[keep], ngi[keep], nur[keep], nui[keep]
        active[idx[~keep]] = False
    cd = np.complex64 if precision == 1 else np.complex128
    return {"gains": (gr + 1j * gi).astype(cd), "ubls": (ur + 1j * ui).astype(cd), "chisq": chisq,
            "iters": iters, "conv": conv}


def redundant_array(n=4, seed=0, nd=3, noise=0.0):
    """Synthetic test problem: an n x n grid of antennas, every pair i < j,
    unique baseline = separation class (dx, dy); random gains near 1 and
    unique-baseline values; data = the model plus optional complex noise."""
    rng = np.random.default_rng(seed)
    pos = [(x, y) for y in range(n) for x in range(n)]
    classes, ggu = {}, []
    for i in range(len(pos)):
        for j in range(i + 1, len(pos)):
            d = (pos[j][0] - pos[i][0], pos[j][1] - pos[i][1])
            k = classes.setdefault(d, len(classes))
            ggu.append((i, j, k))
    ggu = np.array(ggu, dtype=np.uint32)
    nants, nubls = len(pos), len(classes)
    g = (1 + 0.1 * rng.standard_normal((nd, nants))) * np.exp(1j * 0.3 * rng.standard_normal((nd, nants)))
    u = rng.standard_normal((nd, nubls)) + 1j * rng.standard_normal((nd, nubls))
    i, j, k = ggu[:, 0], ggu[:, 1], ggu[:, 2]
    data = g[:, i] * np.conj(g[:, j]) * u[:, k]
    data = data + noise * (rng.standard_normal(data.shape) + 1j * rng.standard_normal(data.shape))
    wgts = np.ones(data.shape)
    return ggu, g, u, data, wgts
