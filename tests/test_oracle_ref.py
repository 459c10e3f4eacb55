"""The oracle against the reference's own compiled pieces.

Every check runs against the committed golden vectors (tests/golden/
ref_vectors.npz, written by tests/golden/make_golden.py from oracle/_ref) and,
when oracle/_ref was built in this tree, against the live reference library
on fresh random inputs as well.
"""
import os

import numpy as np
import pytest

import oracle

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))
L = oracle.lib()
R = oracle.ref()
needs_ref = pytest.mark.skipif(R is None, reason="oracle/_ref not built (reference tree absent)")


def test_size_rule_golden():
    out = np.array([L.orc_fft_size_1d(int(n)) for n in GOLD["size_rule_in"]], np.uint32)
    assert np.array_equal(out, GOLD["size_rule_out"])
    # the benchmark extents are their own FFT sizes
    for n in (1000, 250, 500, 180, 150, 120):
        assert L.orc_fft_size_1d(n) == n


def test_fkst_golden():
    t = 0
    while f"fkst_{t}_map" in GOLD:
        m = GOLD[f"fkst_{t}_map"]
        k = int(GOLD[f"fkst_{t}_k"])
        seed = float(GOLD[f"fkst_{t}_seed"])
        v, p = oracle.fkst(m, k, seed)
        assert np.array_equal(v, GOLD[f"fkst_{t}_val"]), t
        assert np.array_equal(p, GOLD[f"fkst_{t}_pos"]), t
        t += 1
    assert t >= 40


@needs_ref
def test_fkst_live_tie_heavy():
    rng = np.random.default_rng(2024)
    for t in range(600):
        N = int(rng.integers(1, 500))
        k = int(rng.integers(1, 12))
        m = rng.integers(-2, 3, N).astype(np.float32)
        if t % 2:
            m[: rng.integers(0, N + 1)] = -np.inf
        seed = float(rng.random())
        v1, p1 = oracle.fkst(m, k, seed)
        v2 = np.zeros(k, np.float32)
        p2 = np.zeros(k, np.uint32)
        R.ref_fkst(oracle.f32(m), N, k, seed, v2, p2)
        assert np.array_equal(v1, v2) and np.array_equal(p1, p2), t


def _conv_case(ci):
    g = {k[len(f"conv_{ci}_"):]: GOLD[k] for k in GOLD.files if k.startswith(f"conv_{ci}_")}
    return g


def test_convert_input4xcorr_golden():
    ci = 0
    while f"conv_{ci}_dims" in GOLD:
        g = _conv_case(ci)
        types = g["types"]
        ncat = g["ncat"] if np.any(types != 0) else []
        catv = g["catvals"] if np.any(types != 0) else []
        img = oracle.ti_subimages(g["data"], g["dims"], types, list(ncat), list(catv), fdims=g["fdims"],
                                  reflected=True)
        assert np.array_equal(img.reshape(-1), g["images"]), ci
        ci += 1
    assert ci >= 4


def test_coef_matrix_golden():
    ci = 0
    while f"coef_{ci}_alpha" in GOLD:
        types = GOLD[f"conv_{ci}_types"]
        ncat = list(GOLD[f"conv_{ci}_ncat"]) if np.any(types != 0) else []
        alpha, conv, sub_of = oracle.coef_matrix(types, ncat)
        n = alpha.shape[0]
        assert np.array_equal(alpha.T.reshape(-1), GOLD[f"coef_{ci}_alpha"]), ci  # [A + B*nSub]
        assert np.array_equal(conv, GOLD[f"coef_{ci}_conv"]), ci
        # constant term: P2 with -1 for continuous variables only
        ct = GOLD[f"coef_{ci}_consttype"]
        cc = GOLD[f"coef_{ci}_constcoef"]
        for v, t in enumerate(types):
            if t == 0:
                assert ct[v] == 2 and cc[v] == -1.0
            else:
                assert ct[v] == -9
        ci += 1


def test_kernel_generation_golden():
    ki = 0
    while f"kern_{ki}_dims" in GOLD:
        dims = GOLD[f"kern_{ki}_dims"]
        kt = GOLD[f"kern_{ki}_types"]
        out = np.zeros(GOLD[f"kern_{ki}_out"].size, np.float32)
        L.orc_kernel_generate(len(dims), dims, len(kt), kt, GOLD[f"kern_{ki}_w"], GOLD[f"kern_{ki}_a"], out)
        assert np.array_equal(out, GOLD[f"kern_{ki}_out"]), ki
        fac = GOLD[f"kexp_{ki}_factor"]
        ex = np.zeros(GOLD[f"kexp_{ki}_out"].size, np.float32)
        L.orc_kernel_expand(int(np.prod(dims)), len(kt), out, fac, ex)
        assert np.array_equal(ex, GOLD[f"kexp_{ki}_out"]), ki
        ki += 1
    assert ki == 3


def test_index_with_delta_golden():
    for row in GOLD["iwd"]:
        D = int(row[0])
        dims = oracle.u32(row[1:1 + D])
        pos = int(row[4])
        delta = oracle.i32(row[5:5 + D])
        loc = np.zeros(1, np.uint32)
        ok = L.orc_index_with_delta(D, dims, pos, delta, loc)
        assert int(loc[0]) == int(row[8]) and ok == int(row[9]), row


def test_complex_mac_golden():
    dst = GOLD["cadd_dst"].astype(np.float64)
    C = GOLD["cadd_C"].astype(np.float64)
    D = GOLD["cadd_D"].astype(np.float64)
    a = float(GOLD["cadd_alpha"])
    c = C[0::2] + 1j * C[1::2]
    d = D[0::2] + 1j * D[1::2]
    r = dst[0::2] + 1j * dst[1::2] + a * c * d  # dst += alpha*C*D, no conjugate
    out = GOLD["cadd_out"]
    assert np.allclose(out[0::2], r.real, rtol=1e-5, atol=1e-5)
    assert np.allclose(out[1::2], r.imag, rtol=1e-5, atol=1e-5)


@needs_ref
def test_convert_and_coef_live():
    rng = np.random.default_rng(77)
    for t in range(30):
        D = int(rng.integers(1, 4))
        dims = oracle.u32(rng.integers(2, 7, D))
        cells = int(np.prod(dims))
        types = oracle.i32([int(rng.integers(0, 2))] + ([0] if rng.random() < 0.5 else []))
        nvar = len(types)
        cols = []
        for tp in types:
            cols.append(rng.integers(0, 4, cells).astype(np.float32) if tp else rng.standard_normal(cells).astype(np.float32))
        data = np.stack(cols, 1).reshape(-1)
        data[rng.integers(0, data.size)] = np.nan
        ncat, catv = [], []
        for v in range(nvar):
            if types[v]:
                c = oracle.categories(data, nvar, v)
                ncat.append(len(c))
                catv.extend(c.tolist())
        fd = oracle.u32([L.orc_fft_size_1d(int(x) + int(rng.integers(0, 3))) for x in dims])
        nsub = 2 * int(np.sum(types == 0)) + int(sum(ncat))
        ref_img = np.zeros(nsub * int(np.prod(fd)), np.float32)
        R.ref_convert_input4xcorr(D, dims, fd, nvar, types, oracle.f32(data), oracle.i32(ncat or [0]),
                                  oracle.f32(catv or [0]), ref_img)
        img = oracle.ti_subimages(data, dims, types, ncat, catv, fdims=fd, reflected=True)
        assert np.array_equal(img.reshape(-1), ref_img), t


@needs_ref
def test_index_with_delta_live():
    rng = np.random.default_rng(3)
    for t in range(2000):
        D = int(rng.integers(1, 4))
        dims = oracle.u32(rng.integers(1, 12, D))
        pos = int(rng.integers(0, int(np.prod(dims))))
        delta = oracle.i32(rng.integers(-15, 16, D))
        a = np.zeros(1, np.uint32)
        b = np.zeros(1, np.uint32)
        assert L.orc_index_with_delta(D, dims, pos, delta, a) == R.ref_index_with_delta(D, dims, pos, delta, b)
        assert a[0] == b[0]


def test_product_kernel_generation_golden():
    """qsmi_kernel_generate (the library's CPU set-up) against the vectors the
    compiled reference's genearteKernel wrote."""
    import g2s_amd

    ki = 0
    while f"kern_{ki}_dims" in GOLD:
        dims = GOLD[f"kern_{ki}_dims"]
        out = g2s_amd.kernel_generate(dims, GOLD[f"kern_{ki}_types"], GOLD[f"kern_{ki}_w"], GOLD[f"kern_{ki}_a"])
        assert np.array_equal(out, GOLD[f"kern_{ki}_out"]), ki
        ki += 1
    assert ki >= 2


OFFS = np.load(os.path.join(os.path.dirname(__file__), "golden", "offsets_vectors.npz"))


def _offset_cases():
    i = 0
    while f"off_{i}_dims" in OFFS.files:
        kern = OFFS[f"off_{i}_kernel"] if f"off_{i}_kernel" in OFFS.files else None
        yield i, OFFS[f"off_{i}_dims"], kern, OFFS[f"off_{i}_out"].astype(np.int32)
        i += 1


def test_offset_order_golden():
    """The oracle's offset list and the product's (qsmi_offsets) in qs.cpp's
    exact order, ties included, against the order the compiled DataImage sort
    key gives (src/qs.cpp:1055-1137), 1-3 D, with and without NaN kernel
    cells, and the 501 x 501 kernel of the 1000^2 benchmark."""
    import g2s_amd

    n = 0
    for i, kd, kern, want in _offset_cases():
        want = want.reshape(-1, len(kd))
        got = oracle.offsets(kd, kern, 2)
        assert np.array_equal(got, want), (i, kd)
        prod = np.asarray(g2s_amd.offsets([int(d) for d in kd], kern, 2)).reshape(-1, len(kd))
        assert np.array_equal(prod, want), (i, kd)
        n += 1
    assert n >= 10


@needs_ref
def test_offset_order_live():
    rng = np.random.default_rng(202)
    for kd in ((13, 11), (6, 5, 4), (31,), (3, 3)):
        kern = np.ones(int(np.prod(kd)) * 2, np.float32)
        kern[rng.random(kern.size) < 0.25] = np.nan
        for k in (None, kern):
            cnt = int(np.prod([2 * ((d + 1) // 2) - 1 for d in kd]))
            res = np.zeros(cnt * len(kd), np.int32)
            m = R.ref_offsets(len(kd), oracle.u32(kd), 2, None if k is None else k.ctypes.data, res)
            assert np.array_equal(oracle.offsets(kd, k, 2), res[: m * len(kd)].reshape(m, len(kd)))
