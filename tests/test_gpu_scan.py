"""GPU parity of the MI355X scan against the oracle (called through the C-ABI).

Maps: |E_gpu - E_oracle| <= 1e-4 * s(p) (FP32 FFT vs FP64 closed form).
Selection: bit-exact against the oracle's device rule on the same map;
k-best sets equal to the FP64 oracle's except where the K-th/(K+1)-th gap
is within the map tolerance.
"""
import sys

import numpy as np
import pytest

import g2s_amd
import oracle
from qs_helpers import MAP_RTOL, Problem, grf2d, neighbourhood, valid_box

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
    u, d0 = prob.convert(x)
    assert dev.candidate_for_pattern(off, u, d0)
    gm = dev.error_map().astype(np.float64)
    om = prob.oracle_map(off, u, d0, circular)
    valid = np.isfinite(om)
    assert np.array_equal(np.isfinite(gm), valid), "valid-position masks differ"
    sc = prob.scale_map(off, x)
    err = np.abs(gm[valid] - om[valid])
    tol = MAP_RTOL * sc[valid]
    worst = float(np.max(err / np.maximum(sc[valid], 1e-30))) if err.size else 0.0
    assert np.all(err <= tol), f"map error {worst:.3e} (relative to s(p)) above {MAP_RTOL}"
    return gm, om, u, d0, worst


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
    worst = 0.0
    for trial in range(4):
        off, x = neighbourhood(prob, rng, nb, radius)
        gm, om, u, d0, w = check_map(prob, dev, off, x)
        worst = max(worst, w)
        lo, hi = valid_box(prob.dims, off)
        for K in (1, 2, 5):
            seed = float(rng.random())
            err, idx = dev.search_k_biggest(K, seed)
            pos = np.array([dev.cvt_index_to_position(int(i)) for i in idx])
            # bit-exact against the device rule on the same (GPU) map
            ov, op, _ = oracle.select_device(prob.dims, gm.astype(np.float32), lo, hi, K, seed, quantum)
            assert np.array_equal(pos, op.astype(np.int64)), (pos, op)
            assert np.array_equal(err, ov)
            # against the FP64 map: same set unless the K-th/(K+1)-th gap is within tolerance
            ov2, op2, _ = oracle.select_device(prob.dims, om.astype(np.float32), lo, hi, K + 1, seed, quantum)
            if set(pos.tolist()) != set(op2[:K].tolist()):
                sc = prob.scale_map(off, x)
                gap = float(om[op2[K - 1]] - om[op2[K]])
                bound = MAP_RTOL * float(sc[op2[K - 1]] + sc[op2[K]])
                assert gap <= bound, f"k-best sets differ with gap {gap} > {bound}"
    _log(f"  worst map error / s(p) = {worst:.3e}")
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
        off, x = neighbourhood(prob, rng, nb, int(rng.integers(4, 8)) if len(dims) < 3 else 3)
        u, d0 = prob.convert(x)
        seed = float(rng.random())
        nodes.append((off, u, d0, seed))
        refs.append((off, x, u, d0, seed))
    err, pos, ch = dev.scan_wave(nodes, 1.5)
    for i, (off, x, u, d0, seed) in enumerate(refs):
        om = prob.oracle_map(off, u, d0)
        lo, hi = valid_box(prob.dims, off)
        ov, op, _ = oracle.select_device(prob.dims, om.astype(np.float32), lo, hi, 3, seed, quantum)
        if list(pos[i]) != list(op[:2]):
            sc = prob.scale_map(off, x)
            # any swap must be explained by a gap within the map tolerance
            for a, b in ((0, 1), (1, 2)):
                gap = float(om[op[a]] - om[op[b]])
                if gap <= MAP_RTOL * float(sc[op[a]] + sc[op[b]]):
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
        gm, om, u, d0, w = check_map(prob, dev, off, x)
        _log(f"  radius {radius}: map error / s(p) = {w:.3e}")
        lo, hi = valid_box(prob.dims, off)
        err, idx = dev.search_k_biggest(2, 0.37)
        pos = np.array([dev.cvt_index_to_position(int(q)) for q in idx])
        ov, op, _ = oracle.select_device(prob.dims, gm.astype(np.float32), lo, hi, 2, 0.37, quantum)
        assert np.array_equal(pos, op.astype(np.int64))
        seed = float(rng.random())
        nodes.append((off, u, d0, seed))
        refs.append((off, x, om, seed))
    err, pos, ch = dev.scan_wave(nodes, 1.5)
    for i, (off, x, om, seed) in enumerate(refs):
        lo, hi = valid_box(prob.dims, off)
        ov, op, _ = oracle.select_device(prob.dims, om.astype(np.float32), lo, hi, 3, seed, quantum)
        if list(pos[i]) != list(op[:2]):
            sc = prob.scale_map(off, x)
            for a, b in ((0, 1), (1, 2)):
                if float(om[op[a]] - om[op[b]]) <= MAP_RTOL * float(sc[op[a]] + sc[op[b]]):
                    break
            else:
                raise AssertionError(f"node {i}: {pos[i]} vs oracle {op[:2]}")
    dev.close()
    ti.close()
