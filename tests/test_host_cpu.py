"""The product's C++ host layer and C-ABI on the CPU (no device calls).

- libqsmi.so loads and exports every symbol include/qsmi.h declares;
- the host-layer restatements (size rule, categories, coefficient matrix,
  TI sub-images, offset list, path and seeds) equal the oracle's, which is
  pinned to the reference in test_oracle_ref.py;
- argument errors come back as error codes with a message, like the
  reference's argument checks, without touching a GPU.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import g2s_amd
import oracle
from g2s_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "qsmi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qsmi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = g2s_amd.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
    # and the Python bindings describe each of them
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_size_rule_matches_oracle():
    for n in list(range(1, 600)) + [1000, 1001, 1023, 2047, 4096]:
        assert g2s_amd.fft_size(n) == oracle.lib().orc_fft_size_1d(n)


def test_categories_coef_subimages_match_oracle():
    rng = np.random.default_rng(4)
    for types in ([0], [1], [0, 1], [1, 0]):
        dims = (7, 6)
        cells = 42
        cols = [rng.integers(0, 3, cells).astype(np.float32) * 1.5 if t else rng.standard_normal(cells).astype(np.float32)
                for t in types]
        ti = np.stack(cols, 1)
        ti[3, 0] = np.nan
        ncat, catv = [], []
        for v, t in enumerate(types):
            if t:
                a = g2s_amd.categories(ti, len(types), v)
                b = oracle.categories(ti, len(types), v)
                assert np.array_equal(a, b)
                ncat.append(len(a))
                catv.extend(a.tolist())
        a1, c1, s1 = g2s_amd.coef_matrix(types, ncat)
        a2, c2, s2 = oracle.coef_matrix(types, ncat)
        assert np.array_equal(a1, a2) and np.array_equal(c1, c2) and np.array_equal(s1, s2)
        im1 = g2s_amd.ti_subimages(ti, dims, types, ncat, catv)
        im2 = oracle.ti_subimages(ti, dims, types, ncat, catv)
        assert np.array_equal(im1, im2)


@pytest.mark.parametrize("kdims", [(7,), (21, 21), (126, 126), (9, 5, 3), (501, 501)])
def test_offset_list_matches_oracle(kdims):
    a = g2s_amd.offsets(kdims)
    b = oracle.offsets(kdims)
    assert np.array_equal(a, b)
    # sorted by distance, a permutation of the kernel box, centre first
    d2 = np.sum(a.astype(np.int64) ** 2, axis=1)
    assert np.all(np.diff(d2) >= 0) and d2[0] == 0
    assert a.shape[0] == np.prod([2 * ((k + 1) // 2) - 1 for k in kdims])


def test_offset_list_nan_kernel():
    k = np.ones(9 * 9, np.float32)
    k[::7] = np.nan
    a = g2s_amd.offsets((9, 9), kernel=k)
    b = oracle.offsets((9, 9), kernel=k)
    assert np.array_equal(a, b) and a.shape[0] < 81


@pytest.mark.parametrize("n,seed,frac", [(1000, 1, 0.0), (4096, 7, 0.1), (12345, 3, 0.5)])
def test_path_and_seeds_match_oracle(n, seed, frac):
    rng = np.random.default_rng(seed)
    informed = (rng.random(n) < frac).astype(np.uint8)
    p1, s1, b1 = g2s_amd.path(informed, seed)
    p2, s2, b2 = oracle.path(informed, seed)
    assert b1 == b2 == int(informed.sum())
    assert np.array_equal(p1, p2) and np.array_equal(s1, s2)
    assert np.array_equal(np.sort(p1), np.arange(n))
    assert np.all(informed[p1[:b1]] == 1) and np.all(informed[p1[b1:]] == 0)
    assert np.all((s1 >= 0) & (s1 < 1))


def test_argument_errors_without_gpu():
    L = g2s_amd.lib()
    h = C.c_void_p()
    dims = np.array([10, 10], np.uint32)
    img = np.zeros(200, np.float32)
    rc = L.qsmi_ti_create(0, 4, dims, None, 2, img, C.byref(h))
    assert rc == 3 and b"ndim" in L.qsmi_last_error()
    rc = L.qsmi_ti_create(0, 2, dims, None, 0, img, C.byref(h))
    assert rc == 3
    big = np.array([5000, 10], np.uint32)
    rc = L.qsmi_ti_create(0, 2, big, None, 1, np.zeros(50000, np.float32), C.byref(h))
    assert rc == 3 and b"4096" in L.qsmi_last_error()
    assert L.qsmi_device_destroy(None) == 0 and L.qsmi_ti_destroy(None) == 0
    with pytest.raises(g2s_amd.QsmiError):
        g2s_amd.offsets((3, 3, 3, 3))


def test_reflected_layout_golden():
    """The point-reflected FFT-grid sub-images the compiled reference's
    DataImage::convertInput4Xcorr writes (golden vectors; what
    SharedMemoryManager hands to allocAndInitSharedMemory) -> the TI-grid
    layout of qsmi_ti_create: equal to the oracle's unreflected images.
    TI-grid images placed without the reflection are rejected."""
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))
    ci = 0
    while f"conv_{ci}_dims" in gold:
        g = {k[len(f"conv_{ci}_"):]: gold[k] for k in gold.files if k.startswith(f"conv_{ci}_")}
        dims, fdims, types = g["dims"], g["fdims"], g["types"]
        ncat = list(g["ncat"]) if np.any(types != 0) else []
        catv = list(g["catvals"]) if np.any(types != 0) else []
        F = int(np.prod(fdims))
        imgs = g["images"].reshape(-1, F)
        got = g2s_amd.subimages_from_reflected(list(imgs), dims, fdims)
        ref = oracle.ti_subimages(g["data"], dims, types, ncat, catv)
        assert np.array_equal(got, ref), ci
        if types[0] == 0:
            # y^2 of cell 0 placed at FFT-grid index 0 is, read as reflected,
            # the last FFT-grid cell: outside the TI box when fftSize > srcSize
            bad = np.zeros_like(imgs)
            S = int(np.prod(dims))
            for q in range(S):
                c = np.unravel_index(q, tuple(reversed(dims)))[::-1]
                flat = int(np.ravel_multi_index(tuple(reversed(c)), tuple(reversed(fdims))))
                bad[:, flat] = ref[:, q]
            assert ref[0, 0] != 0
            with pytest.raises(g2s_amd.QsmiError):
                g2s_amd.subimages_from_reflected(list(bad), dims, fdims)
        ci += 1
    assert ci >= 4
