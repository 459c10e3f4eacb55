"""The oracle's auto_qs calibration loop (oracle/qs_oracle_calib.cpp).

Pins, each chosen so that a plausible mistake fails one of them:
- the offsets order: every offset of the kernel box exactly once, the origin
  first, distances non-decreasing (the order of equal distances is the
  reference's std::sort on the same sequence and comparator, by construction);
- the whole loop against an independent Python restatement of
  include/calibration.hpp:30-253 on tiny inputs: numpy's MT19937 stream
  (RandomState's raw 32-bit words are std::mt19937's for the same seed, and
  std::uniform_real_distribution<float> is float(word) / 2^32), brute-force
  numpy maps (E = -sum_j w_j (x_j - y(c + d_j))^2 for a continuous variable,
  sum_j w_j [x_j == y(c + d_j)] for a categorical one), the verbatim disk by
  its definition, the skip rule and the accumulation by their definitions.
  A dropped or reordered draw, a wrong cell formula, a lost exclusion, a
  wrong setting index or an off-by-one in the skip test changes the logged
  sequence or the outputs.
- the exclusion disk against its definition on random maps.
"""
import ctypes
import ctypes.util

import numpy as np
import pytest

import oracle

# C's powf and pow (numpy turns a power of 0.5 into a square root, which can
# differ from powf in the last place)
_LIBM = ctypes.CDLL(ctypes.util.find_library("m"))
_LIBM.powf.restype = ctypes.c_float
_LIBM.powf.argtypes = [ctypes.c_float, ctypes.c_float]
_LIBM.pow.restype = ctypes.c_double
_LIBM.pow.argtypes = [ctypes.c_double, ctypes.c_double]


def _powf(x, y):
    return np.float32(_LIBM.powf(float(x), float(y)))


class _MT:
    """std::mt19937 + std::uniform_real_distribution<float>(0, 1)."""

    def __init__(self, seed):
        self.rs = np.random.RandomState(seed)

    def u(self):
        x = int(np.frombuffer(self.rs.bytes(4), "<u4")[0])
        v = np.float32(np.float32(x) / np.float32(4294967296.0))
        if v >= np.float32(1):
            v = np.nextafter(np.float32(1), np.float32(0))
        return v


def _box_offsets(kdims):
    rng = [range(-((k + 1) // 2 - 1), (k + 1) // 2) for k in kdims]
    return sorted(tuple(int(v) for v in o) for o in np.array(np.meshgrid(*rng, indexing="ij")).reshape(len(kdims), -1).T)


@pytest.mark.parametrize("kdims", [(5,), (7, 5), (4, 6), (3, 5, 3), (9, 9)])
def test_offsets_order(kdims):
    o = oracle.calib_offsets(kdims)
    assert sorted(map(tuple, o.tolist())) == _box_offsets(kdims)
    assert tuple(o[0]) == (0,) * len(kdims)
    d2 = (o.astype(np.int64) ** 2).sum(1)
    assert np.all(np.diff(d2) >= 0)


def test_exclude_disk():
    rng = np.random.default_rng(3)
    for dims in ((13,), (11, 9), (7, 6, 5)):
        S = int(np.prod(dims))
        for r in (0.0, 0.5, 1.0, 2.5, 4.0):
            m = rng.standard_normal(S).astype(np.float32)
            o = int(rng.integers(0, S))
            ref = m.copy()
            oc = np.unravel_index(o, dims[::-1])[::-1]
            cc = np.array(np.unravel_index(np.arange(S), dims[::-1])[::-1]).T
            d2 = ((cc - np.array(oc)) ** 2).sum(1)
            ref[d2 <= (r * r if r >= 1 else 0)] = -np.inf
            oracle.lib().orc_exclude_disk(len(dims), oracle.u32(dims), m, o, r)
            assert np.array_equal(m, ref)


def _py_calibration(tis, dims, vtype, kernels, kdims, nbc, dens, K, power, max_iter, min_iter, noise, seed,
                    quantum):
    """include/calibration.hpp:30-253 for one variable, from the reference's
    text, with numpy brute-force maps."""
    D = len(dims)
    S = int(np.prod(dims))
    nTI, nK, nC, nD = len(tis), len(kernels), len(nbc), len(dens)
    offs = oracle.calib_offsets(kdims[0])
    coords = np.array(np.unravel_index(np.arange(S), dims[::-1])[::-1]).T  # (S, D), dim 0 first
    radius = np.float32(np.power(np.float64(S) * 0.005, np.float64(np.float32(1) / np.float32(D))))
    r2 = radius * radius if radius >= 1 else np.float32(0)
    settings = nC * nK * nD
    cum = np.zeros(settings * K)
    cum2 = np.zeros(settings * K)
    cnt = np.zeros(settings * K, np.int64)
    best = np.full(nD, np.float32(np.inf), np.float32)
    devb = np.zeros(nD, np.float32)
    mt = _MT(seed)
    log = []
    f32 = np.float32
    with np.errstate(divide="ignore", invalid="ignore"):
        for _ in range(max_iter):
            snap_b, snap_d = best.copy(), devb.copy()
            for s in range(settings):
                cfg, ker, den = s % nC, (s // nC) % nK, s // (nC * nK)
                eb = f32(cum[s * K] / cnt[s * K])
                db = f32(np.sqrt(cum2[s * K] / cnt[s * K] - np.float64(f32(eb * eb))))
                nbs = f32(1) / f32(power)
                if cnt[s * K] > min_iter and (eb - nbs * db) > (snap_b[den] + nbs * snap_d[den]):
                    continue
                n = int(nbc[cfg])
                t = int(np.floor(mt.u() * f32(nTI)))
                cell = int(np.floor(mt.u() * f32(S) / f32(1)))
                c0 = coords[cell]
                sel = []
                ps = 0
                while len(sel) < n and ps < len(offs):
                    if mt.u() < f32(dens[den]) and np.all((c0 + offs[ps] >= 0) & (c0 + offs[ps] < np.array(dims))):
                        sel.append(ps)
                    ps += 1
                if ps >= len(offs):
                    continue
                d = offs[sel].copy()
                x = np.array([tis[t][np.ravel_multi_index(tuple((c0 + o)[::-1]), dims[::-1])] for o in d], np.float32)
                if len(d) > 1 and noise > 0:
                    frac = f32(noise) / f32(len(d)) / f32(2) if noise >= 1 else f32(noise) / f32(2)
                    for _j in range(len(d)):
                        if mt.u() < frac:
                            a = int(np.floor(mt.u() * f32(len(d))))
                            b = int(np.floor(mt.u() * f32(len(d))))
                            d[[a, b]] = d[[b, a]]
                if len(d) <= 1:
                    continue
                sseed = mt.u()
                # kernel weights at centre dims/2 + d (0 outside the kernel)
                kd = np.array(kdims[ker])
                kc = kd // 2 + d
                inside = np.all((kc >= 0) & (kc < kd), axis=1)
                w = np.array([kernels[ker][np.ravel_multi_index(tuple(k[::-1]), tuple(kd[::-1]))] if ok else 0.0
                              for k, ok in zip(kc, inside)], np.float64)
                lo = np.maximum(0, -d.min(0))
                hi = np.array(dims) - 1 - np.maximum(0, d.max(0))
                cand = []
                for ti in range(nTI):
                    img = tis[ti].astype(np.float64)
                    m = np.full(S, -np.inf)
                    ok = np.all((coords >= lo) & (coords <= hi), axis=1)
                    E = np.zeros(ok.sum())
                    for j in range(len(d)):
                        y = img[np.ravel_multi_index(tuple((coords[ok] + d[j]).T[::-1]), dims[::-1])]
                        if vtype == 0:
                            E -= w[j] * (np.float64(x[j]) - y) ** 2
                        else:
                            E += w[j] * (x[j] == y)
                    m[ok] = E
                    m = m.astype(np.float32)
                    if ti == t:
                        dd = ((coords - c0).astype(np.float32) ** 2).sum(1)
                        m[dd <= r2] = -np.inf
                    v, p, _ = oracle.select_device(dims, m, [0] * D, [s_ - 1 for s_ in dims], K, sseed, quantum)
                    cand += [((np.rint(vv / quantum) if quantum > 0 else vv), ti, int(pp))
                             for vv, pp in zip(v, p) if pp != 0xFFFFFFFF]
                order = sorted(range(len(cand)), key=lambda i: -cand[i][0])  # stable
                picks = [(cand[i][1], cand[i][2]) for i in order[:K]]
                ref = tis[t][cell]
                for i, (ti, pp) in enumerate(picks):
                    e = np.float64(_powf(np.abs(f32(tis[ti][pp]) - f32(ref)), f32(power)))
                    cum[s * K + i] += e
                    cum2[s * K + i] += e * e
                    cnt[s * K + i] += 1
                log.append((s, t, cell, picks + [(0xFFFFFFFF, 0xFFFFFFFF)] * (K - len(picks))))
                ea = f32(cum[s * K] / cnt[s * K])
                if eb > ea and best[den] > ea:
                    best[den] = ea
                    devb[den] = f32(np.sqrt(cum2[s * K] / cnt[s * K] - np.float64(f32(ea * ea))))
        mf = (cum / cnt).astype(np.float32)
        mean = np.array([_powf(v, f32(1) / f32(power)) for v in mf], np.float32)
        dev = np.array([_LIBM.pow(float(a), 0.5 / float(power))
                        for a in cum2 / cnt - (mf * mf).astype(np.float64)]).astype(np.float32)
    shape = (nD, nK, nC, K)
    return mean.reshape(shape), dev.reshape(shape), cnt.reshape(shape), log


CASES = {
    # name: (n TIs, variable type, kernels, configs, densities, K, power, iterations, min_iter, noise)
    "continuous": (1, 0, 1, [2, 4, 6], [1.0, 0.4], 2, 2.0, 4, 10**6, 0.0),
    "skip_rule": (1, 0, 1, [1, 3, 5], [0.6, 0.3], 1, 2.0, 12, 2, 0.0),
    "two_tis_noise": (2, 0, 2, [3, 6], [0.5], 3, 1.0, 3, 10**6, 0.7),
    "categorical": (1, 1, 1, [2, 5], [0.8], 2, 1.0, 3, 10**6, 2.0),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_loop_matches_python_restatement(name):
    nti, vtype, nker, nbc, dens, K, power, iters, min_iter, noise = CASES[name]
    rng = np.random.default_rng(len(name))
    dims = (9, 7)
    S = int(np.prod(dims))
    if vtype == 0:
        tis = [rng.standard_normal(S).astype(np.float32) for _ in range(nti)]
    else:
        tis = [rng.integers(0, 3, S).astype(np.float32) * 2 + 1 for _ in range(nti)]
    kdims = [(5, 5)] + [(3, 5)] * (nker - 1)
    kernels = [rng.uniform(0.5, 1.5, int(np.prod(k))).astype(np.float32) for k in kdims]
    quantum = 0.0 if vtype == 0 else 2.0 ** -10
    seed = 7 + len(name)
    got = oracle.calibration(tis, dims, [vtype], kernels, kdims, np.array(nbc)[:, None], dens, K, power, iters,
                             min_iter, noise, seed, quantum=quantum, log_cap=100000)
    mean, dev, cnt, log = _py_calibration(tis, dims, vtype, kernels, kdims, nbc, dens, K, power, iters, min_iter,
                                          noise, seed, quantum)
    assert got["distributions"] == len(log)
    assert len(log) > 5
    for i, (s, t, c, picks) in enumerate(log):
        assert got["log_setting"][i] == s and tuple(got["log_cell"][i]) == (t, c), i
        assert [tuple(p) for p in got["log_pos"][i].tolist()] == picks, i
    assert np.array_equal(got["samples"], cnt)
    assert np.array_equal(got["mean"], mean, equal_nan=True)
    assert np.array_equal(got["dev"], dev, equal_nan=True)
    if name == "skip_rule":
        # the rule fired: some setting has fewer samples than the most sampled
        assert cnt[..., 0].min() < cnt[..., 0].max()
