"""The oracle's cross-measurement form (training images with NaN).

- the two coefficient layers and the conversion types equal the
  reference's generateCoefMatrix4Xcorr with forXMeasurement, and the
  sub-images (with their 0/1 mask images) equal convertInput4Xcorr with
  needCrossMeasurement: committed vectors (tests/golden/cross_vectors.npz,
  from oracle/_ref) and, when oracle/_ref is built here, live;
- the node map equals its definition evaluated by brute force on tiny
  inputs: at every valid TI position c, E = -sum_j m(c+d_j) (x_j - y(c+d_j))^2
  over the informed neighbours, C = m(c) * sum_j m(c+d_j), map = -|E / C^4|
  (-inf where C = 0) — a sign, a dropped mask or a wrong power fails it;
- a whole serial realisation never copies a NaN: every value is the TI
  value at its recorded position.
"""
import os

import numpy as np
import pytest

import oracle
from g2s_amd import synth

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "cross_vectors.npz"))
L = oracle.lib()
R = oracle.ref()
needs_ref = pytest.mark.skipif(R is None, reason="oracle/_ref not built (reference tree absent)")


def _coef_x(nv):
    a = np.zeros(64 * 64, np.float32)
    ax = np.zeros(64 * 64, np.float32)
    ct = np.zeros(64, np.int32)
    so = np.zeros(64, np.int32)
    n = L.orc_coef_matrix_x(nv, np.zeros(nv, np.int32), a, ax, ct, so)
    return a[:n * n], ax[:n * n], ct[:n]


def _sub_x(dims, nv, data):
    f = oracle.fft_size(dims)
    F = int(np.prod(f))
    out = np.zeros(3 * nv * F, np.float32)
    n = L.orc_ti_subimages_x(len(dims), oracle.u32(dims), f, nv, np.zeros(nv, np.int32), oracle.f32(data), 1, out)
    return out[:n * F]


def test_coef_layers_golden():
    for nv in (1, 2, 3):
        a, ax, ct = _coef_x(nv)
        assert np.array_equal(a, GOLD[f"cx_coef_{nv}_alpha"])
        assert np.array_equal(ax, GOLD[f"cx_coef_{nv}_alphaX"])
        assert np.array_equal(ct, GOLD[f"cx_coef_{nv}_conv"])


def test_subimages_golden():
    i = 0
    while f"cx_sub_{i}_out" in GOLD:
        dims = tuple(int(x) for x in GOLD[f"cx_sub_{i}_dims"])
        out = _sub_x(dims, int(GOLD[f"cx_sub_{i}_nvar"]), GOLD[f"cx_sub_{i}_data"])
        assert np.array_equal(out, GOLD[f"cx_sub_{i}_out"]), i
        i += 1
    assert i >= 4


@needs_ref
def test_subimages_live():
    rng = np.random.default_rng(8)
    for _ in range(10):
        D = int(rng.integers(1, 4))
        dims = tuple(int(x) for x in rng.integers(2, 7, D))
        nv = int(rng.integers(1, 3))
        data = rng.standard_normal(int(np.prod(dims)) * nv).astype(np.float32)
        data[rng.random(data.size) < 0.3] = np.nan
        f = oracle.fft_size(dims)
        F = int(np.prod(f))
        ref = np.zeros(3 * nv * F, np.float32)
        n = R.ref_convert_input4xcorr_x(D, oracle.u32(dims), f, nv, np.zeros(nv, np.int32), data,
                                        np.zeros(1, np.int32), np.zeros(1, np.float32), ref)
        assert np.array_equal(_sub_x(dims, nv, data), ref[:n * F])


def _tiny_problem(seed):
    g = synth.grf((9, 11), 3.0, seed=seed)
    ti = np.exp(g).astype(np.float32).reshape(-1)
    rng = np.random.default_rng(seed)
    ti[rng.random(ti.size) < 0.15] = np.nan
    dims = (11, 9)
    kd = [d // 2 + 1 for d in dims]
    offs = oracle.offsets(kd)
    sd = (7, 6)
    informed = np.zeros(int(np.prod(sd)), np.uint8)
    p, s, b = oracle.path(informed, seed)
    orc = oracle.Sim(ti, dims, [0], [], [], sd, kd, np.ones(int(np.prod(kd)), np.float32), offs, 6, 1.5, p, s, b,
                     mode=0, cross=True)
    return ti, dims, sd, offs, p, orc


@pytest.mark.parametrize("seed", [2, 3, 4])
def test_cross_map_brute_force(seed):
    ti, dims, sd, offs, p, orc = _tiny_problem(seed)
    s0, s1 = dims
    img = ti.reshape(s1, s0)
    rng = np.random.default_rng(seed + 100)
    di = np.exp(synth.grf((sd[1], sd[0]), 2.0, seed=seed + 7)).astype(np.float32).reshape(-1)
    for i in rng.choice(len(p), 6, replace=False):
        cell = int(p[i])
        idx = np.zeros(6, np.int32)
        nb = L.orc_sim_gather(orc.h, int(i), idx)
        m, sc, nb2 = orc.node_map(di, int(i))
        assert nb == nb2
        c0, c1 = cell % sd[0], cell // sd[0]
        d = offs[idx[:nb]]
        x = np.array([np.nan if (dd[0] == 0 and dd[1] == 0) else di[(c1 + dd[1]) * sd[0] + c0 + dd[0]] for dd in d])
        lo0, hi0 = -min(0, d[:, 0].min()), s0 - 1 - max(0, d[:, 0].max())
        lo1, hi1 = -min(0, d[:, 1].min()), s1 - 1 - max(0, d[:, 1].max())
        ref = np.full(s0 * s1, -np.inf)
        for q1 in range(lo1, hi1 + 1):
            for q0 in range(lo0, hi0 + 1):
                E = 0.0
                Cn = 0.0
                for (a0, a1), xv in zip(d, x):
                    if np.isnan(xv):
                        continue
                    y = float(img[q1 + a1, q0 + a0])
                    if np.isnan(y):
                        continue
                    E -= (float(xv) - y) ** 2
                    Cn += 1.0
                Cn *= 0.0 if np.isnan(img[q1, q0]) else 1.0
                ref[q1 * s0 + q0] = -np.inf if Cn == 0 else -abs(E / Cn ** 4)
        fin = np.isfinite(ref)
        assert np.array_equal(fin, np.isfinite(m)), i
        # the oracle rounds the images (y^2, y) and the converted values
        # (w, w x, w x^2) to float32 first: bounded by 1e-6 of the
        # magnitude of the cancelling terms (the oracle's scale output)
        assert np.all(np.abs(m[fin] - ref[fin]) <= 1e-6 * sc[fin]), (i, np.max(np.abs(m[fin] - ref[fin]) / sc[fin]))


def test_serial_realisation_no_nan():
    ti, dims, sd, offs, p, orc = _tiny_problem(5)
    di, idx = orc.run(np.full(int(np.prod(sd)), np.nan, np.float32))
    assert not np.isnan(di).any()
    assert not np.isnan(ti[idx]).any()
    assert np.array_equal(di, ti[idx])


def test_product_setup_matches_oracle():
    """qsmi_coef_matrix_x / qsmi_ti_subimages_x (the library's set-up, CPU)
    against the oracle's pinned restatement."""
    import ctypes as C

    from g2s_amd._lib import lib

    for nv in (1, 2, 3):
        a, ax, ct = _coef_x(nv)
        pa = np.zeros(256, np.float32)
        pax = np.zeros(256, np.float32)
        pct = np.zeros(16, np.int32)
        pso = np.zeros(16, np.int32)
        n = C.c_int(0)
        assert lib().qsmi_coef_matrix_x(nv, np.zeros(nv, np.int32), pa, pax, pct, pso, C.byref(n)) == 0
        assert n.value == 3 * nv
        assert np.array_equal(pa[:n.value ** 2], a) and np.array_equal(pax[:n.value ** 2], ax)
        assert np.array_equal(pct[:n.value], ct)
    rng = np.random.default_rng(12)
    for dims, nv in (((7, 5), 1), ((4, 3, 5), 2)):
        data = rng.standard_normal(int(np.prod(dims)) * nv).astype(np.float32)
        data[rng.random(data.size) < 0.25] = np.nan
        out = np.zeros(3 * nv * int(np.prod(dims)), np.float32)
        assert lib().qsmi_ti_subimages_x(len(dims), oracle.u32(dims), nv, np.zeros(nv, np.int32), data, out) == 0
        ref = np.zeros_like(out)
        L.orc_ti_subimages_x(len(dims), oracle.u32(dims), oracle.u32(dims), nv, np.zeros(nv, np.int32), data, 0, ref)
        assert np.array_equal(out, ref)
