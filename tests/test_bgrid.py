""".bgrid files (qsmi_bgrid_*) against the reference's own serializer.

- serialized bytes equal DataImage::serialize's (include/DataImage.hpp:
  143-170) on the committed vectors (tests/golden/bgrid_vectors.npz, made by
  tests/golden/make_golden.py from oracle/_ref) and, when oracle/_ref is
  built here, on fresh random grids;
- the validity verdict equals DataImage::validateSerializedData's (:172-224)
  on every mutated buffer of the vectors (each header field, the size field
  and the length changed one at a time);
- parse inverts serialize; files round-trip plain and gzip-compressed, and
  the gzip form is ordinary gzip (Python's gzip module reads the same bytes).
"""
import gzip
import os

import numpy as np
import pytest

import oracle
from g2s_amd import bgrid

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "bgrid_vectors.npz"))
R = oracle.ref()
needs_ref = pytest.mark.skipif(R is None, reason="oracle/_ref not built (reference tree absent)")


def _cases():
    i = 0
    while f"grid_{i}_raw" in GOLD:
        yield i, GOLD[f"grid_{i}_dims"], GOLD[f"grid_{i}_types"], int(GOLD[f"grid_{i}_enc"]), GOLD[f"grid_{i}_data"]
        i += 1


def test_serialize_matches_reference_bytes():
    n = 0
    for i, dims, types, enc, data in _cases():
        raw = bgrid.serialize(data, dims, types, enc)
        assert raw == GOLD[f"grid_{i}_raw"].tobytes(), i
        n += 1
    assert n >= 5


def test_validation_matches_reference():
    checked = 0
    for k in GOLD.files:
        if k.startswith("mut_") and k.endswith("_raw"):
            raw = GOLD[k].tobytes()
            want = bool(GOLD[k[:-4] + "_valid"])
            assert bgrid.valid(raw) == want, k
            checked += 1
    assert checked >= 70
    assert not bgrid.valid(b"")


def test_parse_inverts_serialize():
    for i, dims, types, enc, data in _cases():
        d, dd, tt, ee = bgrid.parse(GOLD[f"grid_{i}_raw"].tobytes())
        assert dd == tuple(int(x) for x in dims) and tt == [int(t) for t in types] and ee == enc
        assert np.array_equal(d.reshape(-1), data, equal_nan=True)


@needs_ref
def test_serialize_matches_live_reference():
    rng = np.random.default_rng(5)
    for _ in range(20):
        nd = int(rng.integers(1, 4))
        dims = rng.integers(1, 9, nd).astype(np.uint32)
        nv = int(rng.integers(1, 4))
        types = rng.integers(0, 2, nv).astype(np.int32)
        enc = int(rng.integers(0, 3))
        data = rng.standard_normal(int(np.prod(dims)) * nv).astype(np.float32)
        buf = np.zeros(1 << 16, np.uint8)
        n = R.ref_bgrid_serialize(nd, dims, nv, types, enc, data, buf.ctypes.data, buf.size)
        assert bgrid.serialize(data, dims, types, enc) == buf[:n].tobytes()


def test_file_round_trip(tmp_path):
    rng = np.random.default_rng(9)
    data = rng.standard_normal((6 * 5 * 4, 2)).astype(np.float32)
    for gz in (False, True):
        p = tmp_path / ("g.bgrid.gz" if gz else "g.bgrid")
        bgrid.write(p, data, (6, 5, 4), [0, 1], encoding=0, gzip=gz)
        d, dims, types, enc = bgrid.read(p)
        assert dims == (6, 5, 4) and types == [0, 1] and enc == 0
        assert np.array_equal(d, data)
        raw = bgrid.serialize(data, (6, 5, 4), [0, 1])
        if gz:
            with gzip.open(p, "rb") as f:
                assert f.read() == raw
        else:
            assert p.read_bytes() == raw


def test_rejects_bad_input(tmp_path):
    from g2s_amd._lib import QsmiError

    with pytest.raises(QsmiError):
        bgrid.serialize(np.zeros(5, np.float32), (5,), [2])
    with pytest.raises(QsmiError):
        bgrid.serialize(np.zeros(5, np.float32), (5,), [0], encoding=3)
    p = tmp_path / "bad.bgrid"
    p.write_bytes(b"not a grid at all, far too short?")
    with pytest.raises(QsmiError):
        bgrid.read(p)
    with pytest.raises(QsmiError):
        bgrid.read(tmp_path / "missing.bgrid")
