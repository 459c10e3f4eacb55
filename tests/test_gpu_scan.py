"""GPU parity of the MI355X scan against the oracle (called through the C-ABI).

Maps (FP32): |E_gpu - E_oracle| <= eps, the library's stated bound
ERR_C * sum_a rms(Y_a) * sum_j |v_ja| (the FP32 FFT error scales with the
operands' norms); the error relative to s(p) is reported.
Selection of continuous maps: exact. The K best positions and their values
equal the oracle's device rule on the FP64 map, with no tolerance (the
library rescores every position within 2 eps of the K-th best in FP64).
Selection with a tie quantum (categorical TIs) stays in FP32: bit-exact
against the device rule on the same GPU map, and against the FP64 map equal
except where the two FP64 values are within quantum + 2 eps.
"""
import sys

import numpy as np
import pytest

import g2s_amd
import oracle
from g2s_amd import synth
from qs_helpers import ERR_C, MAP_RTOL, Problem, grf2d, neighbourhood, neighbourhood_low, valid_box

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_problem(kind, dims, rng):
    cells = int(np.prod(dims))
    shape = tuple(reversed(dims))
    if kind == "cont":
        ti = np.exp(grf2d(shape, rng, smooth=2.0)).reshape(-1)
        return Problem(ti, dims, [0])
    if kind == "heavy":  # heavy tail: exp(1.5 GRF)
        ti = np.exp(1.5 * grf2d(shape, rng, smooth=2.0)).reshape(-1)
        return Problem(ti, dims, [0])
    if kind == "cat":
        g = grf2d(shape, rng, smooth=2.0).reshape(-1)
        ti = (g > np.quantile(g, 0.6)).astype(np.float32) + (g > np.quantile(g, 0.85)).astype(np.float32)
        return Problem(ti, dims, [1])
    if kind == "multi":
        g1 = grf2d(shape, rng, smooth=2.0).reshape(-1)
        g2 = grf2d(shape, rng, smooth=2.0).reshape(-1)
        c = np.digitize(0.7 * g1 + 0.7 * g2, np.quantile(0.7 * g1 + 0.7 * g2, [0.25, 0.5, 0.75])).astype(np.float32)
        ti = np.stack([g1, c], axis=1)
        return Problem(ti, dims, [0, 1])
    raise ValueError(kind)


def check_map(prob, dev, off, x, circular=False):
    """The GPU map against the FP64 closed form: within the stated bound.
    Returns (gpu map, oracle map, u, delta0, error / bound, error / s(p))."""
    u, d0 = prob.convert(x)
    assert dev.candidate_for_pattern(off, u, d0)
    gm = dev.error_map().astype(np.float64)
    om = prob.oracle_map(off, u, d0, circular)
    valid = np.isfinite(om)
    assert np.array_equal(np.isfinite(gm), valid), "valid-position masks differ"
    eps = prob.eps_bound(u)
    err = np.abs(gm[valid] - om[valid])
    ratio = float(err.max() / eps) if err.size else 0.0
    assert ratio <= 1.0, f"map error {err.max():.3e} above the stated bound {eps:.3e}"
    sc = prob.scale_map(off, x)
    rel = float(np.max(err / np.maximum(sc[valid], 1e-30))) if err.size else 0.0
    return gm, om, u, d0, ratio, rel


def check_select_exact(dev, prob, om, lo, hi, K, seed):
    """search_k_biggest on a continuous map equals the device rule on the FP64 map."""
    err, idx = dev.search_k_biggest(K, seed)
    pos = np.array([dev.cvt_index_to_position(int(i)) for i in idx])
    ov, op, _ = oracle.select_device_f64(prob.dims, om, lo, hi, K, seed)
    assert np.array_equal(pos, op.astype(np.int64)), (pos, op)
    assert np.array_equal(err, ov.astype(np.float32)), (err, ov)


def check_select_quantum(dev, prob, gm, om, u, lo, hi, K, seed, quantum):
    """FP32 selection with a tie quantum: bit-exact on the GPU map; against the
    FP64 map the same set unless the values are within quantum + 2 eps."""
    err, idx = dev.search_k_biggest(K, seed)
    pos = np.array([dev.cvt_index_to_position(int(i)) for i in idx])
    ov, op, _ = oracle.select_device(prob.dims, gm.astype(np.float32), lo, hi, K, seed, quantum)
    assert np.array_equal(pos, op.astype(np.int64)), (pos, op)
    assert np.array_equal(err, ov)
    ov2, op2, _ = oracle.select_device_f64(prob.dims, om, lo, hi, K + 1, seed, quantum)
    if set(pos.tolist()) != set(op2[:K].tolist()):
        gap = float(om[op2[K - 1]] - om[op2[K]])
        bound = quantum + 2.0 * prob.eps_bound(u)
        assert gap <= bound, f"k-best sets differ with gap {gap} > {bound}"


CASES = [
    ("cont", (64, 48), 20, 6),
    ("cont", (100, 60), 40, 9),
    ("cat", (64, 48), 25, 7),
    ("multi", (50, 40), 16, 5),
    ("cont", (250,), 12, 30),
    ("cont", (24, 20, 18), 30, 3),
    ("cat", (24, 20, 18), 30, 3),
]


@pytest.mark.parametrize("kind,dims,nb,radius", CASES)
def test_map_and_select(kind, dims, nb, radius):
    _log(f"case {kind} {dims} nb={nb} r={radius}")
    rng = np.random.default_rng(hash((kind, dims)) & 0xFFFF)
    prob = make_problem(kind, dims, rng)
    ti = g2s_amd.TrainingImage(prob.sub, prob.dims)
    dev = g2s_amd.AcceleratorDevice(ti, prob.alpha)
    quantum = 0.0
    if kind == "cat":
        quantum = float(2.0 ** (np.ceil(np.log2(nb)) - 16))
        dev.set_tie_quantum(quantum)
    worst = worst_rel = 0.0
    for trial in range(6):
        if trial < 4 or kind == "cat" or prob.D == 1:
            off, x = neighbourhood(prob, rng, nb, radius)
        else:  # dense low-valued copied neighbourhoods
            off, x = neighbourhood_low(prob, rng, nb)
        gm, om, u, d0, w, rel = check_map(prob, dev, off, x)
        worst, worst_rel = max(worst, w), max(worst_rel, rel)
        lo, hi = valid_box(prob.dims, off)
        for K in (1, 2, 5):
            seed = float(rng.random())
            if quantum > 0.0:
                check_select_quantum(dev, prob, gm, om, u, lo, hi, K, seed, quantum)
            else:
                check_select_exact(dev, prob, om, lo, hi, K, seed)
    _log(f"  worst map error / bound = {worst:.3f}, / s(p) = {worst_rel:.3e}")
    dev.close()
    ti.close()


@pytest.mark.parametrize("dims", [(160, 120), (40, 32, 24)])
def test_low_valued_neighbourhoods(dims):
    """Heavy-tailed TI (exp(1.5 GRF)), dense neighbourhoods copied from its
    lowest values: the FP32 error exceeds the gaps between the best
    positions, the exact selection still returns the FP64 K best."""
    _log(f"low-valued neighbourhoods {dims}")
    rng = np.random.default_rng(77)
    prob = make_problem("heavy", dims, rng)
    ti = g2s_amd.TrainingImage(prob.sub, prob.dims)
    dev = g2s_amd.AcceleratorDevice(ti, prob.alpha)
    worst = worst_rel = 0.0
    nodes, refs = [], []
    for trial in range(8):
        nb = 40 if prob.D == 2 else 60
        off, x = neighbourhood_low(prob, rng, nb, keep=[1.0, 0.9, 0.7][trial % 3])
        gm, om, u, d0, w, rel = check_map(prob, dev, off, x)
        worst, worst_rel = max(worst, w), max(worst_rel, rel)
        lo, hi = valid_box(prob.dims, off)
        for K in (1, 2, 4):
            check_select_exact(dev, prob, om, lo, hi, K, float(rng.random()))
        seed = float(rng.random())
        nodes.append((off, u, d0, seed))
        refs.append((om, lo, hi, seed))
    err, pos, ch = dev.scan_wave(nodes, 1.5)
    for i, (om, lo, hi, seed) in enumerate(refs):
        ov, op, _ = oracle.select_device_f64(prob.dims, om, lo, hi, 2, seed)
        assert list(pos[i]) == list(op), (i, pos[i], op)
        assert np.array_equal(err[i], ov.astype(np.float32))
    _log(f"  worst map error / bound = {worst:.3f}, / s(p) = {worst_rel:.3e}")
    dev.close()
    ti.close()


def test_index_maps_roundtrip():
    rng = np.random.default_rng(5)
    prob = make_problem("cont", (40, 30), rng)
    ti = g2s_amd.TrainingImage(prob.sub, prob.dims)
    dev = g2s_amd.AcceleratorDevice(ti, prob.alpha)
    off, x = neighbourhood(prob, rng, 10, 4)
    u, d0 = prob.convert(x)
    dev.candidate_for_pattern(off, u, d0)
    lo, hi = valid_box(prob.dims, off)
    f = [oracle.lib().orc_fft_size_1d(d) for d in prob.dims]
    mn = np.minimum(off.min(axis=0), 0).astype(np.int32)
    for c1 in range(lo[1], hi[1] + 1, 3):
        for c0 in range(lo[0], hi[0] + 1, 5):
            p = c0 + prob.dims[0] * c1
            idx = dev.cvt_position_to_index(p)
            assert dev.cvt_index_to_position(idx) == p
            assert idx == oracle.lib().orc_position_to_index(2, oracle.u32(f), oracle.u32(prob.dims), mn, p)
    assert dev.array_size() == int(np.prod(f))
    # setValueInErrorArray moves the selection
    err, idx = dev.search_k_biggest(1, 0.3)
    dev.set_value_in_error_array(int(idx[0]), -np.inf)
    err2, idx2 = dev.search_k_biggest(1, 0.3)
    assert idx2[0] != idx[0] and err2[0] <= err[0]
    dev.close()
    ti.close()


def test_self_retrieval_like_reference():
    """src/test.cpp:191-459: the full neighbourhood of a TI cell must retrieve
    that cell (k=2, draw seed 0) — 2-D 256x128 and 1-D 1500, 21^d kernel."""
    import oracle as O

    for dims in ((256, 128), (1500,)):
        gen = np.random.default_rng(1)
        ti = gen.random(int(np.prod(dims))).astype(np.float32)
        prob = Problem(ti, dims, [0])
        tih = g2s_amd.TrainingImage(prob.sub, prob.dims)
        dev = g2s_amd.AcceleratorDevice(tih, prob.alpha)
        offs = O.offsets([21] * len(dims))[1:]  # centre excluded, as the test
        hits = 0
        for trial in range(10):
            cell = int(gen.integers(0, ti.size))
            coords = np.unravel_index(cell, tuple(reversed(dims)))[::-1]
            inside = np.all((np.asarray(coords)[None, :] + offs >= 0)
                            & (np.asarray(coords)[None, :] + offs < np.asarray(dims)[None, :]), axis=1)
            o = offs[inside]
            t = (np.asarray(coords)[None, :] + o)
            flat = np.ravel_multi_index(tuple(t[:, ::-1].T), tuple(reversed(dims)))
            x = ti[flat].reshape(-1, 1)
            u, d0 = prob.convert(x)
            nodes = [(o, u, d0, 0.0)]
            err, pos, ch = dev.scan_wave(nodes, 2.0)
            hits += int(ch[0] == cell)
            assert ch[0] == cell, (dims, cell, ch[0], err)
        dev.close()
        tih.close()


@pytest.mark.parametrize("kind,dims", [("cont", (120, 90)), ("cat", (80, 70)), ("cont", (30, 24, 20))])
def test_scan_wave_matches_per_node(kind, dims):
    """A wave of uneven neighbour counts gives each node the oracle's K best."""
    _log(f"wave {kind} {dims}")
    rng = np.random.default_rng(11)
    prob = make_problem(kind, dims, rng)
    ti = g2s_amd.TrainingImage(prob.sub, prob.dims)
    dev = g2s_amd.AcceleratorDevice(ti, prob.alpha)
    quantum = 0.0
    if kind == "cat":
        quantum = float(2.0 ** (np.ceil(np.log2(40)) - 16))
        dev.set_tie_quantum(quantum)
    nodes, refs = [], []
    for i in range(37):
        nb = int(rng.integers(2, 40))
        if kind == "cont" and i % 4 == 3:
            off, x = neighbourhood_low(prob, rng, max(nb, 8))
        else:
            off, x = neighbourhood(prob, rng, nb, int(rng.integers(4, 8)) if len(dims) < 3 else 3)
        u, d0 = prob.convert(x)
        seed = float(rng.random())
        nodes.append((off, u, d0, seed))
        refs.append((off, x, u, d0, seed))
    err, pos, ch = dev.scan_wave(nodes, 1.5)
    for i, (off, x, u, d0, seed) in enumerate(refs):
        om = prob.oracle_map(off, u, d0)
        lo, hi = valid_box(prob.dims, off)
        ov, op, _ = oracle.select_device_f64(prob.dims, om, lo, hi, 3, seed, quantum)
        if quantum == 0.0:
            assert list(pos[i]) == list(op[:2]), f"node {i}: {pos[i]} vs oracle {op[:2]}"
            assert np.array_equal(err[i], ov[:2].astype(np.float32))
        elif list(pos[i]) != list(op[:2]):
            # FP32 keys: a swap must be explained by FP64 values within quantum + 2 eps
            bound = quantum + 2.0 * prob.eps_bound(u)
            for a, b in ((0, 1), (1, 2)):
                if float(om[op[a]] - om[op[b]]) <= bound:
                    break
            else:
                raise AssertionError(f"node {i}: {pos[i]} vs oracle {op[:2]}")
        r = int(np.floor(np.float32(seed) * np.float32(1.5)))
        assert ch[i] == pos[i][min(r, 1)]
    dev.close()
    ti.close()


@pytest.mark.parametrize("kind", ["cont", "cat"])
def test_fast_path_1000(kind):
    """The specialised 1000 x 1000 kernels (one FFT-1000 per block) against the
    FP64 oracle: maps, select, and a wave with neighbourhood spans up to 250."""
    _log(f"fast path 1000x1000 {kind}")
    rng = np.random.default_rng(1000)
    prob = make_problem(kind, (1000, 1000), rng)
    ti = g2s_amd.TrainingImage(prob.sub, prob.dims)
    dev = g2s_amd.AcceleratorDevice(ti, prob.alpha)
    quantum = 0.0
    if kind == "cat":
        quantum = float(2.0 ** (np.ceil(np.log2(80)) - 16))
        dev.set_tie_quantum(quantum)
    nodes, refs = [], []
    for i, radius in enumerate((5, 12, 40, 120, 250)):
        off, x = neighbourhood(prob, rng, 80, radius)
        gm, om, u, d0, w, rel = check_map(prob, dev, off, x)
        _log(f"  radius {radius}: map error / bound = {w:.3f}, / s(p) = {rel:.3e}")
        lo, hi = valid_box(prob.dims, off)
        if quantum > 0.0:
            check_select_quantum(dev, prob, gm, om, u, lo, hi, 2, 0.37, quantum)
        else:
            check_select_exact(dev, prob, om, lo, hi, 2, 0.37)
        seed = float(rng.random())
        nodes.append((off, u, d0, seed))
        refs.append((off, x, u, om, seed))
    err, pos, ch = dev.scan_wave(nodes, 1.5)
    for i, (off, x, u, om, seed) in enumerate(refs):
        lo, hi = valid_box(prob.dims, off)
        ov, op, _ = oracle.select_device_f64(prob.dims, om, lo, hi, 3, seed, quantum)
        if quantum == 0.0:
            assert list(pos[i]) == list(op[:2]), f"node {i}: {pos[i]} vs oracle {op[:2]}"
        elif list(pos[i]) != list(op[:2]):
            bound = quantum + 2.0 * prob.eps_bound(u)
            for a, b in ((0, 1), (1, 2)):
                if float(om[op[a]] - om[op[b]]) <= bound:
                    break
            else:
                raise AssertionError(f"node {i}: {pos[i]} vs oracle {op[:2]}")
    dev.close()
    ti.close()


def test_fast_path_1000_low_valued():
    """The benchmark TI (1000^2 log-normal, range 20) with dense 80-neighbour
    neighbourhoods copied from its lowest 2-10 % of values: the case whose
    FP32 map error exceeds the gaps between the best positions. Maps within
    the stated bound; K best and the wave's draws equal the FP64 oracle's."""
    _log("fast path 1000x1000, low-valued dense neighbourhoods on the benchmark TI")
    w = synth.config2_continuous_2d()
    prob = Problem(w["ti"], w["ti_dims"], [0])
    ti = g2s_amd.TrainingImage(prob.sub, prob.dims)
    dev = g2s_amd.AcceleratorDevice(ti, prob.alpha)
    rng = np.random.default_rng(924505)
    nodes, refs = [], []
    worst = worst_rel = 0.0
    for trial in range(8):
        off, x = neighbourhood_low(prob, rng, 80, keep=[1.0, 0.92, 0.8, 0.6][trial % 4])
        gm, om, u, d0, wr, rel = check_map(prob, dev, off, x)
        worst, worst_rel = max(worst, wr), max(worst_rel, rel)
        lo, hi = valid_box(prob.dims, off)
        check_select_exact(dev, prob, om, lo, hi, 2, float(rng.random()))
        seed = float(rng.random())
        nodes.append((off, u, d0, seed))
        refs.append((om, lo, hi, seed))
    err, pos, ch = dev.scan_wave(nodes, 1.5)
    for i, (om, lo, hi, seed) in enumerate(refs):
        ov, op, _ = oracle.select_device_f64(prob.dims, om, lo, hi, 2, seed)
        assert list(pos[i]) == list(op), (i, pos[i], op)
        r = min(int(np.floor(np.float32(seed) * np.float32(1.5))), 1)
        assert ch[i] == op[r]
    _log(f"  worst map error / bound = {worst:.3f}, / s(p) = {worst_rel:.3e}")
    dev.close()
    ti.close()


def test_reflected_layout_ti():
    """The TI built from the sub-images exactly as G2S's SharedMemoryManager
    hands them to allocAndInitSharedMemory (the compiled reference's
    DataImage::convertInput4Xcorr output, point-reflected on FFT grids larger
    than the TI: golden vectors) gives the oracle's maps and K best."""
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))
    rng = np.random.default_rng(4)
    ci = 0
    while f"conv_{ci}_dims" in gold:
        g = {k[len(f"conv_{ci}_"):]: gold[k] for k in gold.files if k.startswith(f"conv_{ci}_")}
        dims, fdims, types = tuple(int(d) for d in g["dims"]), g["fdims"], g["types"]
        _log(f"reflected layout case {ci}: TI {dims}, FFT grid {tuple(fdims)}, types {list(types)}")
        prob = Problem(g["data"], dims, types)
        imgs = g["images"].reshape(-1, int(np.prod(fdims)))
        ti = g2s_amd.TrainingImage.from_reflected(list(imgs), dims, fdims)
        dev = g2s_amd.AcceleratorDevice(ti, prob.alpha)
        quantum = 0.0
        if np.all(types != 0):
            quantum = float(2.0 ** (np.ceil(np.log2(4)) - 16))
            dev.set_tie_quantum(quantum)
        assert dev.array_size() == int(np.prod(fdims))
        for trial in range(3):
            off, x = neighbourhood(prob, rng, 4, 1 if len(dims) > 1 else 2)
            gm, om, u, d0, w, rel = check_map(prob, dev, off, x)
            lo, hi = valid_box(prob.dims, off)
            if quantum > 0.0:
                check_select_quantum(dev, prob, gm, om, u, lo, hi, 2, 0.3, quantum)
            else:
                check_select_exact(dev, prob, om, lo, hi, 2, 0.3)
        dev.close()
        ti.close()
        ci += 1
    assert ci >= 4


def _cross_setup(dims, nv, rng, nan_frac=0.2):
    """A TI with NaN in the cross layout: (raw TI (cells, nv), sub-images
    [y^2, y, m] per variable, alpha, alphaX) from the library's set-up, which
    the CPU tests pin to the oracle's restatement of convertInput4Xcorr /
    generateCoefMatrix4Xcorr."""
    import ctypes as C

    from g2s_amd._lib import lib

    shape = tuple(reversed(dims))
    ti = np.stack([np.exp(grf2d(shape, rng, smooth=2.0)).reshape(-1) for _ in range(nv)], axis=1)
    ti = ti.astype(np.float32)
    ti[rng.random(ti.shape) < nan_frac] = np.nan
    n = C.c_int(0)
    pa = np.zeros(256, np.float32)
    pax = np.zeros(256, np.float32)
    pct = np.zeros(16, np.int32)
    pso = np.zeros(16, np.int32)
    assert lib().qsmi_coef_matrix_x(nv, np.zeros(nv, np.int32), pa, pax, pct, pso, C.byref(n)) == 0
    ns = n.value
    alpha = pa[:ns * ns].reshape(ns, ns).T.copy()  # [A, B]
    alpha_x = pax[:ns * ns].reshape(ns, ns).T.copy()
    sub = np.zeros(ns * ti.shape[0], np.float32)
    assert lib().qsmi_ti_subimages_x(len(dims), oracle.u32(dims), nv, np.zeros(nv, np.int32), ti.reshape(-1),
                                     sub) == 0
    return ti, sub.reshape(ns, -1), alpha, alpha_x


@pytest.mark.parametrize("dims,nv,nb", [((60, 48), 1, 18), ((60, 48), 2, 14), ((400, 340), 1, 24),
                                        ((400, 340), 2, 16)])
def test_cross_layer_per_node(dims, nv, nb):
    """The per-node cross-measurement layer (an AcceleratorDevice built
    withCrossMeasurement): both layers against their FP64 closed forms within
    the stated FP32 bound, maskLayerWithVariable + compensateMissingData
    against -|E / (C m)^4| (-inf where the masked count is zero), and the
    K best of the compensated map by the device rule. 60x48 scans on the TI's
    own grid, 400x340 on the 1024 grid."""
    rng = np.random.default_rng(dims[0] + 7 * nv)
    ti, sub, alpha, alpha_x = _cross_setup(dims, nv, rng)
    ns = sub.shape[0]
    tih = g2s_amd.TrainingImage(sub, dims)
    dev = g2s_amd.AcceleratorDevice(tih, alpha)
    dev.set_cross_layer(alpha_x)
    n_fft = int(np.prod([oracle.lib().orc_fft_size_1d(d) for d in dims]))
    rms = np.sqrt((sub.astype(np.float64) ** 2).sum(axis=1) / n_fft)
    offs_all = oracle.offsets([9] * len(dims))
    finite = np.flatnonzero(~np.isnan(ti).any(axis=1))
    for trial in range(3):
        sel = np.sort(rng.choice(np.arange(1, offs_all.shape[0]), size=nb, replace=False))
        off = offs_all[sel]
        x = ti[rng.choice(finite, size=nb)].astype(np.float64) * (1.0 + 0.05 * rng.standard_normal((nb, nv)))
        u = np.zeros((nb, ns), np.float32)
        for v in range(nv):  # conversions P0, P1, P2 of the cross layout, weight 1
            u[:, 3 * v] = 1.0
            u[:, 3 * v + 1] = x[:, v]
            u[:, 3 * v + 2] = x[:, v] ** 2
        assert dev.candidate_for_pattern(off, u, 0.0)
        E = dev.error_map().astype(np.float64)
        Cm = dev.cross_map().astype(np.float64)
        E_ref = oracle.map_direct(dims, sub, alpha, off, u, 0.0)
        C_ref = oracle.map_direct(dims, sub, alpha_x, off, u, 0.0)
        valid = np.isfinite(E_ref)
        assert np.array_equal(np.isfinite(E), valid) and np.array_equal(np.isfinite(Cm), valid)
        uu = u.astype(np.float64)
        eps_e = ERR_C * float(np.sum(rms * np.abs(uu @ alpha.astype(np.float64).T).sum(axis=0)))
        eps_c = ERR_C * float(np.sum(rms * np.abs(uu @ alpha_x.astype(np.float64).T).sum(axis=0)))
        de = np.abs(E[valid] - E_ref[valid]).max()
        dc = np.abs(Cm[valid] - C_ref[valid]).max()
        _log(f"cross layer {dims} nv={nv} trial {trial}: |dE| {de:.3e} (bound {eps_e:.3e}), "
             f"|dC| {dc:.3e} (bound {eps_c:.3e})")
        assert de <= eps_e and dc <= eps_c
        var = trial % nv
        dev.mask_layer_with_variable(1, var)
        dev.compensate_missing_data()
        M = dev.error_map().astype(np.float64)
        cm = C_ref * sub[3 * var + 2].astype(np.float64)  # the count times the centre's mask of `var`
        ok = valid & (cm != 0.0)  # counts are sums of unit weights: exact zeros in FP64
        assert np.array_equal(np.isfinite(M), ok)
        ref = -np.abs(E_ref[ok] / cm[ok] ** 4)
        tol = eps_e / cm[ok] ** 4 + 4.0 * np.abs(E_ref[ok]) * eps_c / np.abs(cm[ok]) ** 5 + 1e-6 * np.abs(ref)
        assert np.all(np.abs(M[ok] - ref) <= tol)
        lo, hi = valid_box(dims, off)
        K = 3
        err, idx = dev.search_k_biggest(K, 0.37)
        pos = np.array([dev.cvt_index_to_position(int(i)) for i in idx])
        ov, op, _ = oracle.select_device(dims, M.astype(np.float32), lo, hi, K, 0.37)
        assert np.array_equal(pos, op.astype(np.int64)) and np.array_equal(err, ov)
    dev.close()
