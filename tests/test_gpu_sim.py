"""Whole-path QuickSampling on the GPU (qsmi_qs_run_in_memory) against the
oracle's serial node loop, node by node ("lockstep").

The GPU run packs path nodes into waves of independent nodes; its result must
be the serial one. For every path node i the oracle recomputes, from the GPU
realisation restricted to the cells the serial loop would see, the node's
neighbours, its FP64 map, the device-rule K best and the draw.

Runs with the exact selection (no tie quantum, no NaN in the TI) must equal
the oracle at every node, with no tolerance. Runs that keep the FP32
selection may differ only where the two positions' FP64 values are within
quantum + 2 eps (tie quantum; eps the stated FP32 map bound) or, in the
cross-measurement form, within MAP_RTOL (s(a) + s(g)); every such node is
reported.
"""
import sys

import numpy as np
import pytest

import g2s_amd
import oracle
from g2s_amd import synth
from qs_helpers import ERR_C, MAP_RTOL

pytestmark = pytest.mark.gpu


def _log(m):
    print(m, flush=True)
    sys.stdout.flush()


def default_kernel(ti_dims, nexp):
    kd = [d // 2 + 1 for d in ti_dims]
    return kd, np.ones(int(np.prod(kd)) * nexp, np.float32)


def lockstep(ti, ti_dims, di0, sim_dims, types, n, k, seed, quantum, nodes=None,
             weights=None, wave_size=64, threads=8, cross=False, no_verbatim=False, n_image=None, k_image=None,
             circular_sim=False, circular_ti=False):
    """nodes: path-relative node indices to check (None: every node);
    weights: one kernel weight per variable (None: the default uniform 1)."""
    types = list(types)
    nvar = len(types)
    gkw = {}
    if weights is not None:
        kd0 = [d // 2 + 1 for d in ti_dims]
        gkw = dict(kernel=np.tile(np.asarray(weights, np.float32), int(np.prod(kd0))), kernel_dims=kd0)
    sim, idx, st = g2s_amd.qs(ti, ti_dims, di0, sim_dims, types, n, k, seed, wave_size=wave_size,
                              tie_quantum=quantum, no_verbatim=no_verbatim, n_image=n_image, k_image=k_image,
                              circular_sim=circular_sim, circular_ti=circular_ti, **gkw)
    _log(f"  gpu run: {st}")
    # oracle set-up identical to the reference's qs.cpp
    ncat, catvals = [], []
    for v in range(nvar):
        if types[v] != 0:
            c = oracle.categories(ti.reshape(-1, nvar), nvar, v)
            ncat.append(len(c))
            catvals.extend(c.tolist())
    nexp = sum(1 if t == 0 else ncat[list(np.nonzero(types)[0]).index(i)] for i, t in enumerate(types))
    kd, kern = default_kernel(ti_dims, nexp)
    if weights is not None:
        # expanded channels: a categorical variable's classes share its weight
        ch = []
        for v, t in enumerate(types):
            nc = 1 if t == 0 else ncat[list(np.nonzero(types)[0]).index(v)]
            ch.extend([float(weights[v])] * nc)
        kern = np.tile(np.asarray(ch, np.float32), int(np.prod(kd)))
    offs = oracle.offsets(kd)
    di0f = di0.reshape(-1, nvar)
    informed = (~np.isnan(di0f).any(axis=1)).astype(np.uint8)
    p, s, b = oracle.path(informed, seed)
    gp, gs, gb = g2s_amd.path(informed, seed)
    assert np.array_equal(p, gp) and np.array_equal(s, gs) and b == gb
    orc = oracle.Sim(ti, ti_dims, types, ncat, catvals, sim_dims, kd, kern, offs, n, k, p, s, b, mode=0,
                     quantum=quantum, cross=cross, no_verbatim=no_verbatim, n_image=n_image, k_image=k_image,
                     circular_sim=circular_sim, circular_ti=circular_ti, di0=di0)
    orc.set_index(idx)  # -nV reads the positions the neighbours were copied from
    K = int(np.ceil(k))
    simf = sim.reshape(-1, nvar)
    # hard data kept
    hard = informed.astype(bool)
    assert np.array_equal(simf[hard], di0f[hard])
    assert not np.isnan(simf).any()
    # partially informed cells keep their values (include/simulation.hpp:658-666)
    given = ~np.isnan(di0f)
    assert np.array_equal(simf[given], di0f[given])
    tif = ti.reshape(-1, nvar).copy()
    if cross:  # cells with any NaN variable lose all of them (src/qs.cpp:1447-1466)
        tif[np.isnan(tif).any(axis=1)] = np.nan
    # every simulated cell holds the TI values at its index where it held none
    sc = p[b:]
    assert np.array_equal(simf[sc], np.where(np.isnan(di0f[sc]), tif[idx[sc]], di0f[sc]))
    exact = quantum == 0.0 and not cross
    mismatches = 0
    worst_gap = 0.0
    check = list(range(len(p) - b)) if nodes is None else [int(i) for i in nodes]
    if threads > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=threads) as ex:
            results = list(ex.map(lambda i: orc.node(simf, i), check))
    else:
        results = [orc.node(simf, i) for i in check]
    for i, (chosen, v, pos, nb) in zip(check, results):
        cell = p[b + i]
        if chosen == idx[cell]:
            continue
        mismatches += 1
        assert nb > 1 or not np.isnan(di0f[cell]).all(), f"marginal draw differs at node {i}"
        m, sc, _ = orc.node_map(simf, i)
        a, g = int(chosen), int(idx[cell])
        assert np.isfinite(m[g]), f"node {i}: GPU chose an invalid position {g}"
        assert not exact, f"node {i}: gpu {g} ({m[g]}) oracle {a} ({m[a]}): the exact selection allows no difference"
        # FP32 selection: explained only by FP64 values within the bound
        gap = abs(m[a] - m[g])
        if cross:
            bound = MAP_RTOL * (sc[a] + sc[g])
        else:
            bound = quantum + 2.0 * ERR_C * orc.node_norm(simf, i)
        assert gap <= bound, f"node {i}: gpu {g} ({m[g]}) oracle {a} ({m[a]}) gap {gap} > {bound}"
        worst_gap = max(worst_gap, gap / bound)
        # the value written is the TI value at the GPU's choice (the cell's
        # own values kept)
        assert np.array_equal(simf[cell], np.where(np.isnan(di0f[cell]), tif[idx[cell]], di0f[cell]))
    _log(f"  lockstep mismatches {mismatches} of {len(check)} checked nodes (of {len(p) - b}; largest gap / "
         f"bound {worst_gap:.3f})")
    return sim, idx, st


def _verbatim_hits(orc, ti_dims, sim_dims, offs, p, b, idx):
    """Nodes whose chosen position is the one -nV removes (the TI position
    neighbour 1 was copied from, moved back by its offset with wrap-around,
    through the pinned index maps with the node's bounding box)."""
    L = oracle.lib()
    f = oracle.fft_size(ti_dims)
    s = oracle.u32(ti_dims)
    hits = total = 0
    for i in range(len(p) - b):
        g = orc.gather(i)
        if len(g) < 2:
            continue
        o = offs[g]
        c = np.array(np.unravel_index(int(p[b + i]), tuple(reversed(sim_dims)))[::-1])
        t1 = int(np.ravel_multi_index(tuple((c + o[1])[::-1]), tuple(reversed(sim_dims))))
        src = np.array(np.unravel_index(int(idx[t1]), tuple(reversed(ti_dims)))[::-1])
        v = (src - o[1]) % np.array(ti_dims)
        vpos = int(np.ravel_multi_index(tuple(v[::-1]), tuple(reversed(ti_dims))))
        mn = np.minimum(o.min(0), 0).astype(np.int32)
        excl = L.orc_index_to_position(len(ti_dims), f, s, mn, L.orc_position_to_index(len(ti_dims), f, s, mn, vpos))
        hits += int(idx[p[b + i]] == excl)
        total += 1
    return hits, total


@pytest.mark.parametrize("kind", ["continuous", "categorical_hard_data"])
def test_sim_no_verbatim(kind):
    """-nV: lockstep with the oracle's serial loop (which removes the
    position in its own index maps), and no node takes its verbatim
    position; the plain run on the same TI takes it at some nodes."""
    if kind == "continuous":
        g = synth.grf((40, 44), 4.0, seed=31)
        ti, dims = synth.to_g2s(np.exp(g))
        types, n, k, q = [0], 12, 1.5, 0.0
        sdims = (36, 30)
        di = np.full(int(np.prod(sdims)), np.nan, np.float32)
    else:
        g = synth.grf((44, 40), 5.0, seed=32)
        ti, dims = synth.to_g2s(np.digitize(g, np.quantile(g, [0.4, 0.75])).astype(np.float32))
        types, n, k, q = [1], 10, 1.2, float(2.0 ** (np.ceil(np.log2(10)) - 16))
        sdims = (32, 34)
        rng = np.random.default_rng(9)
        di = np.full(int(np.prod(sdims)), np.nan, np.float32)
        pick = rng.choice(di.size, di.size // 10, replace=False)
        di[pick] = ti.reshape(-1)[rng.integers(0, ti.size, pick.size)]
    _log(f"-nV {kind}")
    sim, idx, st = lockstep(ti, dims, di, sdims, types, n, k, 5, q, no_verbatim=True)
    kd, kern = default_kernel(dims, 1 if types[0] == 0 else len(np.unique(ti)))
    offs = oracle.offsets(kd)
    informed = (~np.isnan(di)).astype(np.uint8)
    p, s_, b = oracle.path(informed, 5)
    orc = oracle.Sim(ti, dims, types, [] if types[0] == 0 else [len(np.unique(ti))],
                     [] if types[0] == 0 else np.unique(ti).tolist(), sdims, kd, kern, offs, n, k, p, s_, b)
    hits, total = _verbatim_hits(orc, dims, sdims, offs, p, b, idx)
    _log(f"  verbatim positions taken with -nV: {hits} of {total}")
    assert hits == 0 and total > 100
    _, idxp, _ = g2s_amd.qs(ti, dims, di, sdims, types, n, k, 5, wave_size=64, tie_quantum=q)
    hp, tp = _verbatim_hits(orc, dims, sdims, offs, p, b, idxp)
    _log(f"  without -nV: {hp} of {tp}")
    assert hp > 0


def test_sim_continuous_2d():
    g = synth.grf((48, 56), 4.0, seed=21)
    ti, dims = synth.to_g2s(np.exp(g))
    sdims = (40, 36)
    di = np.full(int(np.prod(sdims)), np.nan, np.float32)
    _log("continuous 2-D")
    lockstep(ti, dims, di, sdims, [0], 20, 1.5, 7, 0.0)


def test_sim_categorical_2d():
    g = synth.grf((50, 50), 5.0, seed=22, anis=(3.0,), angle_deg=30.0)
    ti, dims = synth.to_g2s((g > np.quantile(g, 0.72)).astype(np.float32))
    sdims = (40, 40)
    di = np.full(int(np.prod(sdims)), np.nan, np.float32)
    q = float(2.0 ** (np.ceil(np.log2(15)) - 16))
    _log("categorical 2-D")
    lockstep(ti, dims, di, sdims, [1], 15, 1.2, 3, q)


def test_sim_multivariate_hard_data():
    a = synth.grf((40, 44), 4.0, seed=23)
    b_ = synth.grf((40, 44), 4.0, seed=24)
    c = 0.7 * a + 0.71 * b_
    cat = np.digitize(c, np.quantile(c, [0.25, 0.5, 0.75])).astype(np.float32)
    ti, dims = synth.to_g2s(a, cat)
    sdims = (30, 32)
    rng = np.random.default_rng(5)
    di = np.full((int(np.prod(sdims)), 2), np.nan, np.float32)
    pick = rng.choice(di.shape[0], size=di.shape[0] // 10, replace=False)
    di[pick] = ti[rng.integers(0, ti.shape[0], size=pick.size)]
    _log("multivariate 2-D with hard data")
    lockstep(ti, dims, di.reshape(-1), sdims, [0, 1], 12, 1.5, 4, 0.0)


def test_sim_categorical_3d():
    g1 = synth.grf((12, 16, 18), (3.0, 5.0, 5.0), seed=25)
    fac = (g1 > np.quantile(g1, 0.5)).astype(np.float32) + (g1 > np.quantile(g1, 0.8)).astype(np.float32)
    ti, dims = synth.to_g2s(fac)
    sdims = (14, 12, 8)
    di = np.full(int(np.prod(sdims)), np.nan, np.float32)
    q = float(2.0 ** (np.ceil(np.log2(20)) - 16))
    _log("categorical 3-D")
    lockstep(ti, dims, di, sdims, [1], 20, 1.5, 9, q)


def test_sim_1d():
    g = synth.grf((300,), 6.0, seed=26)
    ti, dims = synth.to_g2s(np.exp(g))
    sdims = (200,)
    di = np.full(200, np.nan, np.float32)
    _log("continuous 1-D")
    lockstep(ti, dims, di, sdims, [0], 10, 1.5, 11, 0.0)


def test_sim_nan_ti_continuous_2d():
    """A TI with NaN holes: the cross-measurement form (-|E/C^4|)."""
    g = synth.grf((48, 56), 4.0, seed=27)
    ti = np.exp(g).astype(np.float32).reshape(-1)
    rng = np.random.default_rng(27)
    ti[rng.random(ti.size) < 0.1] = np.nan
    ti[(np.arange(ti.size) % 56 > 40) & (np.arange(ti.size) // 56 < 10)] = np.nan  # a NaN block
    sdims = (40, 36)
    di = np.full(int(np.prod(sdims)), np.nan, np.float32)
    _log("continuous 2-D, TI with NaN")
    sim, idx, st = lockstep(ti, (56, 48), di, sdims, [0], 20, 1.5, 7, 0.0, cross=True)
    assert not np.isnan(ti[idx]).any()


def test_sim_nan_ti_two_continuous_hard_data():
    a = synth.grf((40, 44), 4.0, seed=28)
    b_ = synth.grf((40, 44), 6.0, seed=29)
    ti, dims = synth.to_g2s(a, np.exp(b_))
    rng = np.random.default_rng(28)
    ti[rng.random(ti.shape[0]) < 0.08, 0] = np.nan
    ti[rng.random(ti.shape[0]) < 0.05, 1] = np.nan
    sdims = (30, 32)
    di = np.full((int(np.prod(sdims)), 2), np.nan, np.float32)
    tiok = ti[~np.isnan(ti).any(axis=1)]
    pick = rng.choice(di.shape[0], size=di.shape[0] // 10, replace=False)
    di[pick] = tiok[rng.integers(0, tiok.shape[0], size=pick.size)]
    _log("two continuous variables, TI with NaN, hard data")
    lockstep(ti.reshape(-1), dims, di.reshape(-1), sdims, [0, 0], 12, 1.5, 4, 0.0, cross=True)


def test_sim_several_tis():
    """Several TIs: lockstep with the oracle's merge across TIs (each TI's K
    best, ordered together, rank floor(seed k ceil(nTI/k) / nTI)); a TI that
    never matches changes nothing but the index encoding; three copies of a
    TI at k = 1.5 give the one-TI realisation at k = 1."""
    ga = synth.grf((40, 44), 4.0, seed=41)
    gb = synth.grf((40, 44), 4.0, seed=42)
    a, dims = synth.to_g2s(np.exp(ga))
    bb, _ = synth.to_g2s(np.exp(gb))
    a, bb = a.reshape(-1), bb.reshape(-1)
    sdims = (36, 30)
    cells = int(np.prod(sdims))
    rng = np.random.default_rng(6)
    di = np.full(cells, np.nan, np.float32)
    pick = rng.choice(cells, cells // 20, replace=False)
    di[pick] = a[rng.integers(0, a.size, pick.size)]
    _log("several TIs: 2 x 40x44 continuous, sim 36x30, 5 % hard data, n=12, k=1.5")
    sim, idx, st = g2s_amd.qs([a, bb], dims, di, sdims, [0], 12, 1.5, 4, wave_size=64)
    kd = [d // 2 + 1 for d in dims]
    offs = oracle.offsets(kd)
    informed = (~np.isnan(di)).astype(np.uint8)
    p, s_, b = oracle.path(informed, 4)
    orc = oracle.Sim(np.concatenate([a, bb]), dims, [0], [], [], sdims, kd, np.ones(int(np.prod(kd)), np.float32),
                     offs, 12, 1.5, p, s_, b, n_ti=2)
    free = p[b:]
    both = np.concatenate([a, bb])
    assert np.array_equal(sim[free], both[(idx[free] % 2) * a.size + idx[free] // 2])
    mism = sum(int(orc.node(sim, i)[0] != idx[p[b + i]]) for i in range(len(free)))
    used = np.bincount(idx[free] % 2, minlength=2)
    _log(f"  lockstep mismatches {mism} of {len(free)}; nodes per TI {used.tolist()}")
    assert mism == 0  # exact selection on every TI
    assert used.min() > 0
    far = np.full_like(a, 1e4)
    s1, i1, _ = g2s_amd.qs(a, dims, di, sdims, [0], 12, 1.5, 4, wave_size=64)
    for order, t in (([a, far], 0), ([far, a], 1)):
        s2, i2, _ = g2s_amd.qs(order, dims, di, sdims, [0], 12, 1.5, 4, wave_size=64)
        assert np.all(i2[free] % 2 == t)
        same2 = float(np.mean(i2[free] == i1[free] * 2 + t))
        _log(f"  far TI at {t}: {same2:.4f} of the nodes identical to the one-TI run")
        assert same2 == 1.0
    sk1, ik1, _ = g2s_amd.qs(a, dims, di, sdims, [0], 12, 1.0, 4, wave_size=64)
    s3, i3, _ = g2s_amd.qs([a, a, a], dims, di, sdims, [0], 12, 1.5, 4, wave_size=64)
    same = float(np.mean(i3[free] // 3 == ik1[free]))
    _log(f"  three copies at k=1.5 vs one at k=1: {same:.4f} of the nodes identical")
    assert same == 1.0


@pytest.mark.parametrize("kind", ["continuous", "categorical"])
def test_sim_n_and_k_images(kind):
    """qs -ni / -kvi: per-cell n and k, lockstep with the oracle's serial loop."""
    if kind == "continuous":
        g = synth.grf((48, 56), 4.0, seed=51)
        ti, dims = synth.to_g2s(np.exp(g))
        types, q = [0], 0.0
    else:
        g = synth.grf((50, 50), 5.0, seed=52)
        ti, dims = synth.to_g2s((g > np.quantile(g, 0.6)).astype(np.float32))
        types, q = [1], float(2.0 ** (np.ceil(np.log2(20)) - 16))
    sdims = (40, 36)
    cells = int(np.prod(sdims))
    rng = np.random.default_rng(7)
    nimg = rng.choice([4, 9, 20], cells).astype(np.float32)
    kimg = rng.choice([1.0, 2.0, 3.5, 0.0], cells).astype(np.float32)
    di = np.full(cells, np.nan, np.float32)
    _log(f"n / k images {kind}: n in (4, 9, 20), k in (1, 2, 3.5, run's 1.5)")
    lockstep(ti, dims, di, sdims, types, 20, 1.5, 8, q, n_image=nimg, k_image=kimg)


def test_sim_kernel_index_image():
    """qs -ki a b -kii image: each cell's kernel (weights and its own offset
    order); lockstep with the oracle node by node (one oracle per kernel,
    the node's kernel deciding), and an index image of zeros gives the
    one-kernel run."""
    g = synth.grf((48, 56), 4.0, seed=61)
    ti, dims = synth.to_g2s(np.exp(g))
    ti = ti.reshape(-1)
    sdims = (40, 36)
    cells = int(np.prod(sdims))
    kd = [(21, 21), (15, 11)]
    k0 = np.ones(21 * 21, np.float32)
    yy, xx = np.mgrid[-5:6, -7:8]
    k1 = np.exp(-np.sqrt(xx ** 2 + yy ** 2) / 4.0).astype(np.float32).reshape(-1)
    rng = np.random.default_rng(11)
    kii = (rng.random(cells) < 0.5).astype(np.float32)
    di = np.full(cells, np.nan, np.float32)
    pick = rng.choice(cells, cells // 25, replace=False)
    di[pick] = ti[rng.integers(0, ti.size, pick.size)]
    n, k, seed = 16, 1.5, 12
    _log("kernel-index image: 21x21 uniform / 15x11 exponential, random per cell")
    sim, idx, st = g2s_amd.qs(ti, dims, di, sdims, [0], n, k, seed, wave_size=64, kernels=[k0, k1], kernels_dims=kd,
                              kernel_index_image=kii)
    informed = (~np.isnan(di)).astype(np.uint8)
    p, s_, b = oracle.path(informed, seed)
    sims = [oracle.Sim(ti, dims, [0], [], [], sdims, kdq, kq, oracle.offsets(kdq, kq), n, k, p, s_, b)
            for kdq, kq in zip(kd, (k0, k1))]
    mism, worst = 0, 0.0
    for i in range(len(p) - b):
        cell = p[b + i]
        o = sims[int(kii[cell])]
        chosen, _, _, nb = o.node(sim, i)
        if chosen == idx[cell]:
            continue
        mism += 1
        _log(f"  node {i}: gpu {int(idx[cell])} oracle {int(chosen)}")
    _log(f"  lockstep mismatches {mism} of {len(p) - b}")
    assert mism == 0  # exact selection
    s0, i0, _ = g2s_amd.qs(ti, dims, di, sdims, [0], n, k, seed, wave_size=64, kernel=k0, kernel_dims=kd[0])
    sz, iz, _ = g2s_amd.qs(ti, dims, di, sdims, [0], n, k, seed, wave_size=64, kernels=[k0, k1], kernels_dims=kd,
                           kernel_index_image=np.zeros(cells, np.float32))
    assert np.array_equal(iz, i0) and np.array_equal(sz, s0, equal_nan=True)


@pytest.mark.parametrize("kind", ["continuous", "categorical_no_verbatim"])
def test_sim_circular_simulation(kind):
    """qs -csim: neighbours wrap around the simulation grid; lockstep with
    the oracle's serial loop (with -nV on top in the categorical case)."""
    if kind == "continuous":
        g = synth.grf((48, 56), 4.0, seed=71)
        ti, dims = synth.to_g2s(np.exp(g))
        types, q, nv = [0], 0.0, False
    else:
        g = synth.grf((50, 50), 5.0, seed=72)
        ti, dims = synth.to_g2s((g > np.quantile(g, 0.6)).astype(np.float32))
        types, q, nv = [1], float(2.0 ** (np.ceil(np.log2(15)) - 16)), True
    sdims = (40, 36)
    di = np.full(int(np.prod(sdims)), np.nan, np.float32)
    _log(f"circular simulation {kind}")
    lockstep(ti, dims, di, sdims, types, 15, 1.5, 9, q, circular_sim=True, no_verbatim=nv)


@pytest.mark.parametrize("kind", ["continuous", "categorical"])
def test_sim_circular_ti(kind):
    """qs -cti: every TI position valid, neighbourhoods wrapped around the
    TI (extents that are FFT sizes, so the FFT grid is the TI grid);
    lockstep with the oracle's serial loop on the wrapped closed-form map."""
    if kind == "continuous":
        g = synth.grf((48, 56), 4.0, seed=81)
        ti, dims = synth.to_g2s(np.exp(g))
        types, q = [0], 0.0
    else:
        g = synth.grf((50, 50), 5.0, seed=82)
        ti, dims = synth.to_g2s((g > np.quantile(g, 0.6)).astype(np.float32))
        types, q = [1], float(2.0 ** (np.ceil(np.log2(15)) - 16))
    assert all(g2s_amd.fft_size(d) == d for d in dims)
    sdims = (40, 36)
    di = np.full(int(np.prod(sdims)), np.nan, np.float32)
    _log(f"circular TI {kind}")
    sim, idx, st = lockstep(ti, dims, di, sdims, types, 15, 1.5, 10, q, circular_ti=True)
    assert idx.max() < int(np.prod(dims))


def test_sim_ti_index_image():
    """qs -ii with several TIs: each cell scans only its TI; lockstep with the
    oracle, and an image of ones gives the one-TI run on the second TI."""
    ga = synth.grf((40, 44), 4.0, seed=91)
    gb = synth.grf((40, 44), 4.0, seed=92)
    a, dims = synth.to_g2s(np.exp(ga))
    bb, _ = synth.to_g2s(np.exp(gb))
    a, bb = a.reshape(-1), bb.reshape(-1)
    sdims = (36, 30)
    cells = int(np.prod(sdims))
    rng = np.random.default_rng(8)
    di = np.full(cells, np.nan, np.float32)
    pick = rng.choice(cells, cells // 20, replace=False)
    di[pick] = a[rng.integers(0, a.size, pick.size)]
    mixed = rng.integers(0, 2, cells).astype(np.float32)
    _log("TI-index image: 2 x 40x44 continuous, random TI per cell")
    sim, idx, _ = g2s_amd.qs([a, bb], dims, di, sdims, [0], 12, 1.5, 4, wave_size=64, ti_index_image=mixed)
    kd = [d // 2 + 1 for d in dims]
    offs = oracle.offsets(kd)
    p, s_, b = oracle.path((~np.isnan(di)).astype(np.uint8), 4)
    orc = oracle.Sim(np.concatenate([a, bb]), dims, [0], [], [], sdims, kd, np.ones(int(np.prod(kd)), np.float32),
                     offs, 12, 1.5, p, s_, b, n_ti=2, ti_image=mixed)
    free = p[b:]
    assert np.array_equal(idx[free] % 2, mixed[free].astype(np.uint32))
    mism = sum(int(orc.node(sim, i)[0] != idx[p[b + i]]) for i in range(len(free)))
    _log(f"  lockstep mismatches {mism} of {len(free)}")
    assert mism == 0
    s1, i1, _ = g2s_amd.qs(bb, dims, di, sdims, [0], 12, 1.5, 4, wave_size=64)
    s2, i2, _ = g2s_amd.qs([a, bb], dims, di, sdims, [0], 12, 1.5, 4, wave_size=64,
                           ti_index_image=np.ones(cells, np.float32))
    same = float(np.mean(i2[free] == i1[free] * 2 + 1))
    _log(f"  image of ones vs the one-TI run on B: {same:.4f} of the nodes identical")
    assert same == 1.0


def test_sim_several_tis_of_their_own_extents():
    """Several TIs of different extents: lockstep with the oracle (each TI's
    boxes on its own extent), and a never-matching second TI of another
    extent leaves the one-TI realisation unchanged."""
    ga = synth.grf((40, 44), 4.0, seed=95)
    gb = synth.grf((30, 36), 4.0, seed=96)
    a, dims = synth.to_g2s(np.exp(ga))
    bb, dims_b = synth.to_g2s(np.exp(gb))
    a, bb = a.reshape(-1), bb.reshape(-1)
    sdims = (36, 30)
    cells = int(np.prod(sdims))
    rng = np.random.default_rng(9)
    di = np.full(cells, np.nan, np.float32)
    pick = rng.choice(cells, cells // 20, replace=False)
    di[pick] = a[rng.integers(0, a.size, pick.size)]
    _log(f"several TIs of their own extents: {dims} and {dims_b}")
    sim, idx, _ = g2s_amd.qs([a, bb], dims, di, sdims, [0], 12, 1.5, 4, wave_size=64, ti_dims_all=[dims, dims_b])
    kd = [d // 2 + 1 for d in dims]
    offs = oracle.offsets(kd)
    p, s_, b = oracle.path((~np.isnan(di)).astype(np.uint8), 4)
    orc = oracle.Sim(np.concatenate([a, bb]), dims, [0], [], [], sdims, kd, np.ones(int(np.prod(kd)), np.float32),
                     offs, 12, 1.5, p, s_, b, n_ti=2, ti_dims_all=[dims, dims_b])
    free = p[b:]
    src = np.where(idx[free] % 2 == 0, a[np.minimum(idx[free] // 2, a.size - 1)], bb[np.minimum(idx[free] // 2, bb.size - 1)])
    assert np.array_equal(sim[free], src)
    mism = sum(int(orc.node(sim, i)[0] != idx[p[b + i]]) for i in range(len(free)))
    used = np.bincount(idx[free] % 2, minlength=2)
    _log(f"  lockstep mismatches {mism} of {len(free)}; nodes per TI {used.tolist()}")
    assert mism == 0 and used.min() > 0
    far = np.full(30 * 36, 1e4, np.float32)
    s1, i1, _ = g2s_amd.qs(a, dims, di, sdims, [0], 12, 1.5, 4, wave_size=64)
    s2, i2, _ = g2s_amd.qs([a, far], dims, di, sdims, [0], 12, 1.5, 4, wave_size=64, ti_dims_all=[dims, (30, 36)])
    same = float(np.mean(i2[free] == i1[free] * 2))
    _log(f"  far TI of another extent: {same:.4f} of the nodes identical to the one-TI run")
    assert same == 1.0


@pytest.mark.parametrize("kind", ["continuous_1000", "categorical_generic"])
def test_sim_bit_reproducible(kind):
    """One seeded realisation run twice gives identical bytes (simulation and
    index image): every FP32 sum of the scan has a fixed order, and the exact
    selection depends on FP64 values only."""
    if kind == "continuous_1000":
        w = synth.config2_continuous_2d()
        ti, dims, types, n, k, q = w["ti"], w["ti_dims"], [0], 80, 1.5, 0.0
        sdims = (120, 100)
    else:
        g = synth.grf((50, 50), 5.0, seed=22, anis=(3.0,), angle_deg=30.0)
        ti, dims = synth.to_g2s((g > np.quantile(g, 0.72)).astype(np.float32))
        types, n, k, q = [1], 15, 1.2, float(2.0 ** (np.ceil(np.log2(15)) - 16))
        sdims = (40, 40)
    di = np.full(int(np.prod(sdims)), np.nan, np.float32)
    _log(f"bit reproducibility {kind}: two runs of seed 3")
    s1, i1, _ = g2s_amd.qs(ti, dims, di, sdims, types, n, k, 3, wave_size=256, tie_quantum=q)
    s2, i2, _ = g2s_amd.qs(ti, dims, di, sdims, types, n, k, 3, wave_size=256, tie_quantum=q)
    assert s1.tobytes() == s2.tobytes() and i1.tobytes() == i2.tobytes()


@pytest.mark.parametrize("kind", ["continuous_1000", "categorical_generic"])
def test_sim_wave_size_invariant(kind):
    """The realisation does not depend on how path nodes are packed into
    waves: 16-node waves (several waves per dependency level, alternating
    between the loop's two streams), 256-node waves, and one wave per level
    (one stream) give identical bytes."""
    if kind == "continuous_1000":
        w = synth.config2_continuous_2d()
        ti, dims, types, n, k, q = w["ti"], w["ti_dims"], [0], 80, 1.5, 0.0
        sdims = (200, 160)
    else:
        g = synth.grf((50, 50), 5.0, seed=22, anis=(3.0,), angle_deg=30.0)
        ti, dims = synth.to_g2s((g > np.quantile(g, 0.72)).astype(np.float32))
        types, n, k, q = [1], 15, 1.2, float(2.0 ** (np.ceil(np.log2(15)) - 16))
        sdims = (60, 50)
    di = np.full(int(np.prod(sdims)), np.nan, np.float32)
    runs = []
    for ws in (16, 256, 1 << 20):
        _log(f"wave-size invariance {kind}: waves of <= {ws} nodes")
        s, i, st = g2s_amd.qs(ti, dims, di, sdims, types, n, k, 5, wave_size=ws, tie_quantum=q)
        runs.append((s.tobytes(), i.tobytes(), st["waves"]))
    assert runs[0][2] > runs[2][2] and runs[1][2] >= runs[2][2]
    for s, i, _ in runs[1:]:
        assert s == runs[0][0] and i == runs[0][1]


@pytest.mark.parametrize("kind", ["continuous_1000", "multivariate_500"])
def test_sim_graph_steps_identical(kind):
    """Replaying each wave's kernels as a captured HIP graph (tuning key
    loop.graph: one graph per lane, node counts and wave starts read on the
    device) gives the same bytes as launching them: 16-node waves (many waves
    of different sizes per level, on both lanes) and 256-node waves."""
    if kind == "continuous_1000":
        w = synth.config2_continuous_2d()
        ti, dims, types, n, k, gkw = w["ti"], w["ti_dims"], [0], 80, 1.5, {}
        sdims = (200, 160)
        di = np.full(int(np.prod(sdims)), np.nan, np.float32)
    else:
        c = synth.config4_multivariate_2d()
        ti, dims, types, n, k = c["ti"], c["ti_dims"], c["types"], c["n"], c["k"]
        sdims, di = c["sim_dims"], c["di"]
        kd0 = [d // 2 + 1 for d in dims]  # weights 0.5 / 0.5
        gkw = dict(kernel=np.tile(np.asarray([0.5, 0.5], np.float32), int(np.prod(kd0))), kernel_dims=kd0)
    try:
        for ws in (16, 256):
            runs = []
            for graph in (0, 1):
                g2s_amd.set_default_tuning("loop.graph", graph)
                _log(f"graph steps {kind}: waves of <= {ws} nodes, loop.graph={graph}")
                s, i, st = g2s_amd.qs(ti, dims, di, sdims, types, n, k, 9, wave_size=ws, **gkw)
                runs.append((s.tobytes(), i.tobytes(), st["waves"]))
            assert runs[0][2] == runs[1][2] and runs[0][2] > 1
            assert runs[1][0] == runs[0][0] and runs[1][1] == runs[0][1]
    finally:
        g2s_amd.set_default_tuning("loop.graph", 0)


def _partial_grid(sdims, nvar, rng, frac_partial=0.3, frac_full=0.05, src=None):
    """A simulation grid with hard data and partially informed cells (each
    holding a random non-empty proper subset of the variables), values copied
    from `src` cells."""
    cells = int(np.prod(sdims))
    di = np.full((cells, nvar), np.nan, np.float32)
    u = rng.random(cells)
    full = u < frac_full
    part = (u >= frac_full) & (u < frac_full + frac_partial)
    pick = rng.integers(0, src.shape[0], size=cells)
    di[full] = src[pick[full]]
    for c in np.flatnonzero(part):
        keep = rng.random(nvar) < 0.5
        if keep.all() or not keep.any():
            keep = np.zeros(nvar, bool)
            keep[rng.integers(0, nvar)] = True
        di[c, keep] = src[pick[c], keep]
    return di


@pytest.mark.parametrize("kind", ["multivariate_generic", "continuous_1024", "centre_only"])
def test_sim_partially_informed_cells(kind):
    """Partially informed simulation cells (include/simulation.hpp:324-335,
    581-584, 658-666): simulated at their place in the path with their own
    values as the centre neighbour, keeping them. Lockstep against the
    oracle; centre_only uses n = 1, so a partially informed cell is scanned
    with its own values alone and an empty cell draws from the marginal."""
    rng = np.random.default_rng(61)
    if kind == "multivariate_generic":
        a = synth.grf((40, 44), 4.0, seed=63)
        b_ = synth.grf((40, 44), 4.0, seed=64)
        c = 0.7 * a + 0.71 * b_
        cat = np.digitize(c, np.quantile(c, [0.25, 0.5, 0.75])).astype(np.float32)
        ti, dims = synth.to_g2s(a, cat)
        types, n, k, sdims = [0, 1], 12, 1.5, (30, 32)
    elif kind == "continuous_1024":
        a = synth.grf((360, 400), 12.0, seed=65)
        b_ = synth.grf((360, 400), 12.0, seed=66)
        ti, dims = synth.to_g2s(np.exp(a), 0.6 * a + 0.8 * b_)
        types, n, k, sdims = [0, 0], 20, 1.5, (60, 50)
    else:
        a = synth.grf((30, 36), 4.0, seed=67)
        b_ = synth.grf((30, 36), 4.0, seed=68)
        ti, dims = synth.to_g2s(a, np.exp(b_))
        types, n, k, sdims = [0, 0], 1, 1.5, (20, 18)
    nvar = len(types)
    di = _partial_grid(sdims, nvar, rng, src=ti.reshape(-1, nvar)).reshape(-1)
    _log(f"partially informed cells {kind}: TI {dims}, sim {sdims}, n={n}")
    free = int(np.isnan(di.reshape(-1, nvar)).any(axis=1).sum())  # the path nodes
    nodes = None if kind != "continuous_1024" else _sample_nodes(free, 400, 5)
    lockstep(ti, dims, di, sdims, types, n, k, 13, 0.0, nodes=nodes, wave_size=64, threads=8)


def _sample_nodes(total, count, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(total, size=min(count, total), replace=False))


@pytest.mark.parametrize("nvar", [1, 2])
def test_sim_nan_ti_marginal_draw(nvar):
    """n = 1 (only the cell itself): every node draws from the marginal. With
    NaN cells in the TI the draw walks the complete cells
    (include/simulation.hpp:586-654); the GPU's index of every cell equals
    the oracle's draw for that node's seed."""
    rng = np.random.default_rng(40 + nvar)
    dims = (30, 20)
    cells = dims[0] * dims[1]
    ti = rng.standard_normal(cells * nvar).astype(np.float32)
    ti[rng.random(ti.size) < 0.45] = np.nan
    ti[0] = np.nan  # cell 0 incomplete
    sdims = (17, 13)
    di = np.full(int(np.prod(sdims)) * nvar, np.nan, np.float32)
    _log(f"marginal draw, TI with NaN, {nvar} variable(s)")
    sim, idx, st = g2s_amd.qs(ti, dims, di, sdims, [0] * nvar, 1, 1.5, 11)
    complete = (~np.isnan(ti.reshape(cells, nvar)).any(axis=1)).astype(np.uint8)
    p, s, b = g2s_amd.path(np.zeros(int(np.prod(sdims)), np.uint8), 11)
    want = np.array([oracle.marginal_cell(complete, nvar, float(s[i])) for i in range(len(p))], np.int64)
    got = idx[p].astype(np.int64)
    walked = int(sum(not complete[int(np.floor(np.float32(s[i]) * np.float32(cells)))] for i in range(len(p))))
    _log(f"  {len(p)} nodes, {walked} walked the complete cells")
    assert walked > 0
    assert np.array_equal(got, want)
    assert np.array_equal(sim.reshape(-1, nvar), ti.reshape(cells, nvar)[idx], equal_nan=True)
