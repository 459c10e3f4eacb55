"""The oracle's FFT and mismatch maps against what the mathematics fixes.

- FFT: numpy.fft (pocketfft, FP64) on every radix class the size rule
  produces, 1-D complex and 1/2/3-D real transforms.
- maps: the closed form (direct FP64 sum) against the reference dataflow
  (reflected spectra, index-space wrap mask, cvtIndexToPosition) and against a
  numpy brute force; invariants that catch a dropped term, a wrong sign or a
  transposed index: self-retrieval (E = 0 and maximal at the source cell), the
  valid-position count prod(s_i - (max_i - min_i)), linearity in alpha,
  translation of the TI.
- draw rank and index maps: closed forms and round trips.
"""
import numpy as np
import pytest

import oracle
from qs_helpers import Problem, grf2d, neighbourhood

L = oracle.lib()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 10, 11, 13, 12, 30, 64, 77, 99, 120, 143, 150, 169, 180, 250,
                               256, 343, 500, 1000, 1001])
def test_fft1d_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign, ref in ((-1, np.fft.fft(x)), (+1, np.fft.ifft(x) * n)):
        buf = np.empty(2 * n)
        buf[0::2], buf[1::2] = x.real, x.imag
        L.orc_fft1d_f64(n, buf, sign)
        y = buf[0::2] + 1j * buf[1::2]
        assert np.allclose(y, ref, rtol=1e-10, atol=1e-9 * np.sqrt(n))


@pytest.mark.parametrize("dims", [(16,), (15,), (12, 10), (9, 14), (10, 6, 4), (7, 5, 6)])
def test_rfftn_irfftn_vs_numpy(dims):
    rng = np.random.default_rng(len(dims))
    a = rng.standard_normal(tuple(reversed(dims)))  # C order, last axis = dim0
    H = (dims[0] // 2 + 1) * int(np.prod(dims[1:]))
    spec = np.zeros(2 * H)
    L.orc_rfftn_f64(len(dims), oracle.u32(dims), np.ascontiguousarray(a.reshape(-1)), spec)
    ref = np.fft.rfftn(a)
    got = (spec[0::2] + 1j * spec[1::2]).reshape(ref.shape)
    assert np.allclose(got, ref, atol=1e-9)
    back = np.zeros(a.size)
    L.orc_irfftn_f64(len(dims), oracle.u32(dims), spec, back)
    assert np.allclose(back.reshape(a.shape), a * a.size, atol=1e-8)


def brute_map(prob, off, u, d0):
    """numpy brute force of E(c) = d0 + sum_j sum_AB alpha_AB u_jB Y_A(c + d_j)."""
    dims = prob.dims
    shp = tuple(reversed(dims))
    Y = prob.sub.reshape((prob.nsub,) + shp).astype(np.float64)
    v = u.astype(np.float64) @ prob.alpha.T.astype(np.float64)  # v_jA = sum_B alpha_AB u_jB
    E = np.full(shp, float(d0))
    mn = np.minimum(off.min(0), 0)
    mx = np.maximum(off.max(0), 0)
    valid = np.ones(shp, bool)
    grids = np.meshgrid(*[np.arange(n) for n in shp], indexing="ij")  # C-order coords
    for i in range(len(dims)):
        ci = grids[len(dims) - 1 - i]
        valid &= (ci >= -mn[i]) & (ci <= dims[i] - 1 - mx[i])
    for j in range(off.shape[0]):
        for A in range(prob.nsub):
            if v[j, A] == 0:
                continue
            shifted = np.zeros(shp)
            idx = []
            for ax in range(len(shp)):
                i = len(dims) - 1 - ax
                d = off[j, i]
                idx.append(grids[ax] + d)
            ok = np.ones(shp, bool)
            for ax in range(len(shp)):
                ok &= (idx[ax] >= 0) & (idx[ax] < shp[ax])
            cl = tuple(np.clip(idx[ax], 0, shp[ax] - 1) for ax in range(len(shp)))
            shifted = np.where(ok, Y[A][cl], 0.0)
            E += v[j, A] * shifted
    return np.where(valid, E, -np.inf).reshape(-1)


def make(kind, dims, seed):
    rng = np.random.default_rng(seed)
    shp = tuple(reversed(dims))
    if kind == "cont":
        return Problem(np.exp(grf2d(shp, rng)).reshape(-1), dims, [0]), rng
    if kind == "cat":
        g = grf2d(shp, rng).reshape(-1)
        return Problem((g > 0.3).astype(np.float32) + (g > 1.0), dims, [1]), rng
    g1 = grf2d(shp, rng).reshape(-1)
    c = np.digitize(grf2d(shp, rng).reshape(-1), [-0.5, 0.5]).astype(np.float32)
    return Problem(np.stack([g1, c], 1), dims, [0, 1]), rng


@pytest.mark.parametrize("kind,dims", [("cont", (20, 16)), ("cat", (18, 14)), ("multi", (16, 12)),
                                       ("cont", (40,)), ("cat", (8, 7, 6))])
def test_map_direct_vs_brute_and_refflow(kind, dims):
    prob, rng = make(kind, dims, 3)
    for t in range(3):
        off, x = neighbourhood(prob, rng, 7, 2 if len(dims) > 1 else 6)
        u, d0 = prob.convert(x)
        m = prob.oracle_map(off, u, d0)
        b = brute_map(prob, off, u, d0)
        assert np.array_equal(np.isfinite(m), np.isfinite(b))
        ok = np.isfinite(m)
        assert np.allclose(m[ok], b[ok], rtol=1e-12, atol=1e-9)
        # reference dataflow in index space, FP64
        f = oracle.fft_size(prob.dims)
        refl = oracle.ti_subimages(prob.ti, prob.dims, prob.types, prob.ncat, prob.catvals, fdims=f, reflected=True)
        H = (int(f[0]) // 2 + 1) * int(np.prod(f[1:]))
        Yh = np.zeros(2 * H * prob.nsub)
        L.orc_ti_spectra_f64(len(f), oracle.u32(f), prob.nsub, refl.reshape(-1), Yh)
        mi = np.zeros(int(np.prod(f)), np.float32)
        L.orc_map_refflow_f64(len(f), oracle.u32(prob.dims), oracle.u32(f), prob.nsub, Yh,
                              oracle.f32(prob.alpha.T.reshape(-1)), off.shape[0], oracle.i32(off).reshape(-1),
                              oracle.f32(u).reshape(-1), d0, 0, 1, mi)
        mn = np.minimum(off.min(0), 0).astype(np.int32)
        pos = np.array([L.orc_index_to_position(len(f), oracle.u32(f), oracle.u32(prob.dims), mn, i)
                        for i in range(mi.size)])
        fin = np.isfinite(mi)
        assert fin.sum() == ok.sum()
        mp = np.full(m.size, -np.inf)
        mp[pos[fin]] = mi[fin]
        sc = prob.scale_map(off, x)
        assert np.all(np.abs(mp[ok] - m[ok]) <= 1e-6 * sc[ok] + 1e-6)


def test_self_retrieval_and_valid_count():
    prob, rng = make("cont", (30, 22), 8)
    dims = prob.dims
    offs = oracle.offsets([7, 7])
    for t in range(10):
        c0, c1 = int(rng.integers(3, dims[0] - 3)), int(rng.integers(3, dims[1] - 3))
        sel = offs[1:20]
        cells = (c0 + sel[:, 0]) + dims[0] * (c1 + sel[:, 1])
        x = prob.ti[cells]
        u, d0 = prob.convert(x)
        m = prob.oracle_map(sel, u, d0)
        src = c0 + dims[0] * c1
        # perfect match: E = 0 up to the float32 rounding of delta0 (accumulated in float as the reference)
        assert abs(m[src]) < 1e-6 * abs(d0) + 1e-9
        assert np.argmax(np.where(np.isfinite(m), m, -np.inf)) == src
        span = np.maximum(sel.max(0), 0) - np.minimum(sel.min(0), 0)
        assert np.isfinite(m).sum() == int(np.prod([dims[i] - span[i] for i in range(2)]))


def test_linearity_and_translation():
    prob, rng = make("cont", (24, 18), 9)
    off, x = neighbourhood(prob, rng, 9, 3)
    u, d0 = prob.convert(x)
    m = prob.oracle_map(off, u, d0)
    # linearity in alpha: (2*alpha) applied to u gives 2*(m - d0) + d0
    m2 = oracle.map_direct(prob.dims, prob.sub, 2 * prob.alpha, off, u, d0)
    ok = np.isfinite(m)
    assert np.allclose(m2[ok] - d0, 2 * (m[ok] - d0), rtol=1e-12, atol=1e-9)
    # translating every offset by +1 along dim0 moves the map by -1 along dim0
    offt = off.copy()
    offt[:, 0] += 1
    mt = oracle.map_direct(prob.dims, prob.sub, prob.alpha, offt, u, d0).reshape(tuple(reversed(prob.dims)))
    mm = m.reshape(tuple(reversed(prob.dims)))
    both = np.isfinite(mt[:, :-1]) & np.isfinite(mm[:, 1:])
    assert both.sum() > 0
    assert np.allclose(mt[:, :-1][both], mm[:, 1:][both], rtol=1e-12, atol=1e-9)


def test_draw_rank_closed_form():
    for k in (1.0, 1.2, 1.5, 2.0, 3.7):
        for s in np.linspace(0, 0.9999, 41).astype(np.float32):
            for nti in (1, 2, 3):
                expect = int(np.floor(np.float32(s) * np.float32(k) * np.float32(np.ceil(nti / k) / nti)))
                assert L.orc_draw_rank(float(s), k, nti) == expect


def test_index_maps_roundtrip():
    rng = np.random.default_rng(1)
    for t in range(50):
        D = int(rng.integers(1, 4))
        s = oracle.u32(rng.integers(4, 12, D))
        f = oracle.u32([L.orc_fft_size_1d(int(x) + int(rng.integers(0, 3))) for x in s])
        mn = oracle.i32(-rng.integers(0, 3, D))
        for q in range(20):
            pos = int(rng.integers(0, int(np.prod(s))))
            coords = np.unravel_index(pos, tuple(reversed(s)))[::-1]
            if np.any(np.asarray(coords) + mn < 0):
                continue
            idx = L.orc_position_to_index(D, f, s, mn, pos)
            assert L.orc_index_to_position(D, f, s, mn, idx) == pos
