"""GPU parity of the auto_qs calibration path (called through the C-ABI).

- qsmi_distribution_wave: the k best of each node's map with the verbatim
  disk excluded, against the oracle (FP64 direct map, the disk by its
  definition, the device selection rule): continuous maps (exact selection)
  give exactly the oracle's positions and values; with a tie quantum
  (FP32 selection) positions may differ only where the FP64 values are
  within quantum + 2 eps. No position inside the disk, and without
  exclusion exactly qsmi_scan_wave's positions and values. Generic 2-D /
  3-D kernels and the 1000 x 1000 fast path.
- qsmi_set_value_in_error_array_with_radius: bit-exact against the disk
  applied to the same map.
- qsmi_calibration against oracle.calibration (the CPU restatement, pinned in
  test_oracle_calib.py) on the same seed: the whole run is identical (draw
  log, positions, counts, means): continuous TIs run the exact selection,
  and categorical TIs with unit weights make every map value an exact
  multiple of the tie quantum.
"""
import sys

import numpy as np
import pytest

import g2s_amd
import oracle
from qs_helpers import grf2d, neighbourhood, valid_box
from test_gpu_scan import make_problem

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _disk(dims, o, r):
    S = int(np.prod(dims))
    c = np.array(np.unravel_index(np.arange(S), dims[::-1])[::-1]).T
    oc = np.array(np.unravel_index(o, dims[::-1])[::-1])
    d2 = ((c - oc).astype(np.float32) ** 2).sum(1)
    return d2 <= (np.float32(r) * np.float32(r) if r >= 1 else 0)


DCASES = [
    ("cont", (100, 60), 20, 9, 2),
    ("cat", (64, 48), 25, 7, 3),
    ("cont", (24, 20, 18), 30, 3, 2),
    ("cont", (1000, 1000), 80, 40, 2),
]


@pytest.mark.parametrize("kind,dims,nb,span,K", DCASES)
def test_distribution_wave_exclusion(kind, dims, nb, span, K):
    _log(f"distribution wave {kind} {dims} nb={nb} K={K}")
    rng = np.random.default_rng(sum(dims) + nb)
    prob = make_problem(kind, dims, rng)
    ti = g2s_amd.TrainingImage(prob.sub, prob.dims)
    dev = g2s_amd.AcceleratorDevice(ti, prob.alpha)
    quantum = 0.0
    if kind == "cat":
        quantum = float(2.0 ** (np.ceil(np.log2(nb)) - 16))
        dev.set_tie_quantum(quantum)
    S = int(np.prod(dims))
    nodes, refs, vpos, vrad = [], [], [], []
    radii = (0.0, 0.5, 1.0, 3.0, 12.0, 70.7)
    for i in range(12):
        off, x = neighbourhood(prob, rng, nb, span)
        u, d0 = prob.convert(x)
        seed = float(rng.random())
        lo, hi = valid_box(dims, off)
        # the disk around a valid position (where the exact match would be)
        o = int(np.ravel_multi_index(tuple(int(rng.integers(l, h + 1)) for l, h in zip(lo, hi))[::-1], dims[::-1]))
        r = radii[i % len(radii)]
        nodes.append((off, u, d0, seed))
        refs.append((off, x, u, d0, seed, o, r))
        vpos.append(o)
        vrad.append(r)
    err, pos = dev.distribution_wave(nodes, float(K), np.array(vpos, np.uint32), np.array(vrad, np.float32))
    for i, (off, x, u, d0, seed, o, r) in enumerate(refs):
        om = prob.oracle_map(off, u, d0)
        ex = _disk(dims, o, r)
        assert not ex[pos[i][pos[i] != 0xFFFFFFFF]].any(), f"node {i}: a position inside the disk"
        om[ex] = -np.inf
        lo, hi = valid_box(dims, off)
        ov, op, _ = oracle.select_device_f64(dims, om, lo, hi, K + 1, seed, quantum)
        if quantum == 0.0:
            assert list(pos[i]) == list(op[:K]), f"node {i}: {pos[i]} vs oracle {op[:K]}"
            assert np.array_equal(err[i], ov[:K].astype(np.float32))
        elif list(pos[i]) != list(op[:K]):
            a = next(j for j in range(K) if pos[i][j] != op[j])
            gap = float(om[op[a]] - om[op[a + 1]])
            bound = quantum + 2.0 * prob.eps_bound(u)
            assert gap <= bound, f"node {i}: {pos[i]} vs oracle {op[:K]} (gap {gap} > {bound})"
    # no exclusion: exactly the scan wave's k best (same kernels, fixed
    # summation order)
    e2, p2 = dev.distribution_wave(nodes, float(K), np.full(len(nodes), 0xFFFFFFFF, np.uint32),
                                   np.zeros(len(nodes), np.float32))
    e3, p3, _ = dev.scan_wave(nodes, float(K))
    assert np.array_equal(p2, p3) and np.array_equal(e2, e3)
    with pytest.raises(g2s_amd.QsmiError):
        dev.distribution_wave(nodes[:1], float(K), np.array([S], np.uint32), np.zeros(1, np.float32))
    dev.close()
    ti.close()


@pytest.mark.parametrize("dims", [(64, 48), (24, 20, 18), (1000, 1000)])
def test_set_value_with_radius(dims):
    _log(f"set value with radius {dims}")
    rng = np.random.default_rng(len(dims) * 7 + dims[0])
    prob = make_problem("cont", dims, rng)
    ti = g2s_amd.TrainingImage(prob.sub, prob.dims)
    dev = g2s_amd.AcceleratorDevice(ti, prob.alpha)
    for r in (0.0, 0.7, 1.0, 2.5, 9.0, 70.7):
        off, x = neighbourhood(prob, rng, 12, 4)
        u, d0 = prob.convert(x)
        assert dev.candidate_for_pattern(off, u, d0)
        before = dev.error_map()
        # the disk's centre is the index's position (cvtPositionToIndex
        # takes (position + min) modulo the extent in unsigned arithmetic as
        # the reference does, so not every position survives the round trip)
        idx = dev.cvt_position_to_index(int(rng.integers(0, before.size)))
        o = dev.cvt_index_to_position(idx)
        dev.set_value_in_error_array_with_radius(idx, -np.inf, r)
        ref = before.copy()
        ref[_disk(dims, o, r)] = -np.inf
        assert np.array_equal(dev.error_map(), ref), r
    dev.close()
    ti.close()


def _ti(kind, dims, seed):
    rng = np.random.default_rng(seed)
    g = _grf(dims, rng)
    if kind == "cont":
        return np.exp(g).astype(np.float32)
    return np.digitize(g, np.quantile(g, [0.5, 0.8])).astype(np.float32)


def _grf(dims, rng):
    return grf2d(tuple(reversed(dims)), rng, smooth=2.5).reshape(-1).astype(np.float64)


def _compare(got, ref, min_samples=20):
    n = ref["distributions"]
    assert n >= min_samples
    assert got["distributions"] == n
    assert np.array_equal(got["log_setting"], ref["log_setting"])
    assert np.array_equal(got["log_cell"], ref["log_cell"])
    assert np.array_equal(got["log_pos"], ref["log_pos"])
    assert np.array_equal(got["samples"], ref["samples"])
    assert np.array_equal(got["mean"], ref["mean"], equal_nan=True)
    assert np.array_equal(got["dev"], ref["dev"], equal_nan=True)


CCASES = {
    # name: (kind, TI dims, n TIs, configs, densities, K, power, iterations, min_iter, noise)
    "categorical_skip": ("cat", (48, 40), 1, [3, 6, 10], [0.5, 0.15], 2, 1.0, 10, 2, 0.0),
    "categorical_two_tis_noise": ("cat", (40, 36), 2, [4, 8], [0.3], 3, 1.0, 6, 10**6, 0.6),
    "continuous": ("cont", (48, 40), 1, [4, 8, 12], [0.3, 0.1], 2, 2.0, 8, 10**6, 0.0),
    "continuous_skip": ("cont", (48, 40), 1, [4, 8, 12], [0.3, 0.1], 2, 2.0, 8, 2, 0.0),
    "continuous_3d": ("cont", (16, 14, 12), 1, [6, 12], [0.4], 2, 2.0, 8, 10**6, 0.0),
}


@pytest.mark.parametrize("name", sorted(CCASES))
def test_calibration_matches_oracle(name):
    kind, dims, nti, cfgs, dens, K, power, iters, min_iter, noise = CCASES[name]
    _log(f"calibration {name}: TI {dims} x {nti}, configs {cfgs}, densities {dens}, K={K}, {iters} iterations")
    tis = [_ti(kind, dims, 40 + t) for t in range(nti)]
    D = len(dims)
    kd = [(9,) * D, (7,) * D]
    kernels = [np.ones(int(np.prod(kd[0])), np.float32), np.ones(int(np.prod(kd[1])), np.float32)]
    vt = 0 if kind == "cont" else 1
    nb = np.array(cfgs, np.uint32)[:, None]
    got = g2s_amd.calibration(tis, dims, [vt], kernels, kd, nb_configs=nb, densities=dens, max_k=K, power=power,
                              max_iter=iters, min_iter=min_iter, noise=noise, seed=11, log_cap=100000)
    # the product's tie quantum for categorical-only TIs: 2^(ceil(log2(max n x nVar x max|w|)) - 16)
    quantum = 0.0 if vt == 0 else float(2.0 ** (np.ceil(np.log2(max(cfgs))) - 16))
    ref = oracle.calibration(tis, dims, [vt], kernels, kd, nb, dens, K, power, iters, min_iter, noise, 11,
                             quantum=quantum, log_cap=100000)
    _compare(got, ref)
    _log(f"  {ref['distributions']} samples, {got['waves']} waves, identical to the oracle, "
         f"{got['ms']:.1f} ms ({got['scan_ms']:.1f} ms in the scans)")


def test_calibration_fast_path_1000():
    """The 1000 x 1000 fast kernels under the calibration loop (exclusion
    radius 70.7 around each sample's cell)."""
    _log("calibration on a 1000x1000 continuous TI, 3 iterations")
    dims = (1000, 1000)
    ti = _ti("cont", dims, 77)
    kd = [(41, 41)]
    kernels = [np.ones(41 * 41, np.float32)]
    nb = np.array([[20], [60]], np.uint32)
    got = g2s_amd.calibration([ti], dims, [0], kernels, kd, nb_configs=nb, densities=[0.5, 0.125], max_k=2,
                              power=2.0, max_iter=3, min_iter=10**6, seed=3, log_cap=1000)
    ref = oracle.calibration([ti], dims, [0], kernels, kd, nb, [0.5, 0.125], 2, 2.0, 3, 10**6, 0.0, 3, log_cap=1000)
    assert np.array_equal(got["log_cell"], ref["log_cell"])
    cells = got["log_cell"][:, 1]
    for i in range(got["distributions"]):
        for t, p in got["log_pos"][i]:
            if p == 0xFFFFFFFF:
                continue
            dy = int(p // 1000) - int(cells[i] // 1000)
            dx = int(p % 1000) - int(cells[i] % 1000)
            assert dx * dx + dy * dy > 70.7 ** 2
    agree = float(np.mean(np.all(got["log_pos"] == ref["log_pos"], axis=(1, 2))))
    _log(f"  {got['distributions']} samples, position agreement {agree:.3f}")
    _compare(got, ref, min_samples=10)
