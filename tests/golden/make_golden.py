"""Writes tests/golden/ref_vectors.npz from the reference's own header-only
QS pieces compiled in place (oracle/_ref/libqs_ref.so, built by
oracle/Makefile from /root/reference). Inputs and outputs only (data):

  size_rule        SharedMemoryManager size rule for 1..4096
  fkst_*           fKst::findKBiggest with the QSM RNG seeding on tie-heavy
                   maps with leading -inf runs (values, positions)
  conv_*           DataImage::convertInput4Xcorr sub-images (reflected f-grid)
  coef_*           DataImage::generateCoefMatrix4Xcorr matrices / conversions
  kern_*           DataImage::genearteKernel (uniform / exponential / gaussian)
  kexp_*           DataImage::offsetKernel4categories
  iwd_*            DataImage::indexWithDelta
  cadd_*           g2s::complexAddAlphaxCxD

and tests/golden/bgrid_vectors.npz:

  grid_*           DataImage::serialize bytes of small grids (1-3 D, one or
                   several variables, every variable type and encoding)
  mut_*            buffers derived from them with one header field, the
                   size field or the length changed, and
                   DataImage::validateSerializedData's verdict on each

and tests/golden/offsets_vectors.npz:

  off_*            the neighbour offsets of a kernel in qs.cpp's order (the
                   sort key image and its lookups from the compiled
                   DataImage, std::sort), with and without NaN kernel cells,
                   including the 501 x 501 kernel of the 1000^2 benchmark

and tests/golden/cross_vectors.npz (TIs with NaN, cross-measurement form):

  cx_coef_*        DataImage::generateCoefMatrix4Xcorr with forXMeasurement
                   (both layers, conversion types) for 1-3 continuous variables
  cx_sub_*         DataImage::convertInput4Xcorr with needCrossMeasurement
                   (reflected f-grid images incl. the 0/1 mask images)

Run: python tests/golden/make_golden.py  (needs oracle/_ref built).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(HERE, "ref_vectors.npz")
OUT_BGRID = os.path.join(HERE, "bgrid_vectors.npz")
OUT_CROSS = os.path.join(HERE, "cross_vectors.npz")
OUT_OFFSETS = os.path.join(HERE, "offsets_vectors.npz")


def ti_cases():
    rng = np.random.default_rng(1234)
    cases = []
    # (dims, types, data)
    d = rng.standard_normal(7 * 5).astype(np.float32)
    d[3] = np.nan
    cases.append(((7, 5), [0], d))
    c = rng.integers(0, 3, 11).astype(np.float32) * 2.5
    cases.append(((11,), [1], c))
    m = np.stack([rng.standard_normal(4 * 3 * 5), rng.integers(0, 4, 4 * 3 * 5)], 1).astype(np.float32).reshape(-1)
    cases.append(((4, 3, 5), [0, 1], m))
    # categorical before continuous. (Two categorical variables are left out:
    # convertInput4Xcorr never advances its category index,
    # include/DataImage.hpp:540-575, so the reference builds the second
    # variable's images from the first variable's classes.)
    m2 = np.stack([rng.integers(0, 2, 6 * 4), rng.standard_normal(6 * 4)], 1).astype(np.float32).reshape(-1)
    cases.append(((6, 4), [1, 0], m2))
    return cases


def main():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libqs_ref.so not built (run make -C oracle with the reference present)")
    out = {}
    # size rule
    sizes = np.arange(1, 4097, dtype=np.uint32)
    fs = np.zeros_like(sizes)
    for i, n in enumerate(sizes):
        o = np.zeros(1, np.uint32)
        R.ref_fft_size(1, oracle.u32([n]), o)
        fs[i] = o[0]
    out["size_rule_in"] = sizes
    out["size_rule_out"] = fs
    # fKst on tie-heavy maps
    rng = np.random.default_rng(99)
    for t in range(40):
        N = int(rng.integers(1, 400))
        k = int(rng.integers(1, 9))
        m = rng.integers(-3, 4, N).astype(np.float32)
        if t % 3 == 0:
            m[: rng.integers(0, N + 1)] = -np.inf
        if t % 5 == 0:
            m = rng.standard_normal(N).astype(np.float32)
        seed = np.float32(rng.random()) if t % 7 else np.float32(0.0)
        v = np.zeros(k, np.float32)
        p = np.zeros(k, np.uint32)
        R.ref_fkst(m, N, k, float(seed), v, p)
        out[f"fkst_{t}_map"] = m
        out[f"fkst_{t}_k"] = np.int32(k)
        out[f"fkst_{t}_seed"] = seed
        out[f"fkst_{t}_val"] = v
        out[f"fkst_{t}_pos"] = p
    # convertInput4Xcorr / generateCoefMatrix4Xcorr
    for ci, (dims, types, data) in enumerate(ti_cases()):
        dims = oracle.u32(dims)
        types = oracle.i32(types)
        nvar = len(types)
        ncat, catvals = [], []
        for v in range(nvar):
            if types[v]:
                cv = oracle.categories(data, nvar, v)
                ncat.append(len(cv))
                catvals.extend(cv.tolist())
        fd = oracle.u32([oracle.lib().orc_fft_size_1d(int(x) + 3) for x in dims])  # padded f-grid
        nsub = 2 * int(np.sum(types == 0)) + int(np.sum(ncat))
        img = np.zeros(nsub * int(np.prod(fd)), np.float32)
        cnt = R.ref_convert_input4xcorr(len(dims), dims, fd, nvar, types, oracle.f32(data),
                                        oracle.i32(ncat or [0]), oracle.f32(catvals or [0]), img)
        assert cnt == nsub
        alpha = np.zeros(nsub * nsub, np.float32)
        conv = np.zeros(nsub, np.int32)
        ctypes_ = np.zeros(nvar, np.int32)
        ccoef = np.zeros(nvar, np.float32)
        ns = R.ref_coef_matrix(nvar, types, oracle.i32(ncat or [0]), oracle.f32(catvals or [0]), alpha, conv, ctypes_,
                               ccoef)
        assert ns == nsub
        out[f"conv_{ci}_dims"] = dims
        out[f"conv_{ci}_fdims"] = fd
        out[f"conv_{ci}_types"] = types
        out[f"conv_{ci}_data"] = oracle.f32(data)
        out[f"conv_{ci}_ncat"] = oracle.i32(ncat or [0])
        out[f"conv_{ci}_catvals"] = oracle.f32(catvals or [0])
        out[f"conv_{ci}_images"] = img
        out[f"coef_{ci}_alpha"] = alpha
        out[f"coef_{ci}_conv"] = conv
        out[f"coef_{ci}_consttype"] = ctypes_
        out[f"coef_{ci}_constcoef"] = ccoef
    # kernels
    for ki, (dims, kt, w, a) in enumerate([((5, 7), [0, 2], [1.0, 0.5], [0.0, 0.3]),
                                          ((9,), [7], [1.0], [0.4]),
                                          ((3, 4, 5), [0, 7, 2], [2.0, 1.0, 1.0], [0.0, 0.2, 1.1])]):
        dims = oracle.u32(dims)
        kern = np.zeros(int(np.prod(dims)) * len(kt), np.float32)
        R.ref_generate_kernel(len(dims), dims, len(kt), oracle.i32(kt), oracle.f32(w), oracle.f32(a), kern)
        out[f"kern_{ki}_dims"] = dims
        out[f"kern_{ki}_types"] = oracle.i32(kt)
        out[f"kern_{ki}_w"] = oracle.f32(w)
        out[f"kern_{ki}_a"] = oracle.f32(a)
        out[f"kern_{ki}_out"] = kern
        fac = oracle.i32([1 + (i % 3) for i in range(len(kt))])
        ex = np.zeros(int(np.prod(dims)) * int(fac.sum()), np.float32)
        R.ref_kernel_expand(len(dims), dims, len(kt), kern, fac, ex)
        out[f"kexp_{ki}_factor"] = fac
        out[f"kexp_{ki}_out"] = ex
    # indexWithDelta
    rng = np.random.default_rng(5)
    iw = []
    for t in range(200):
        D = int(rng.integers(1, 4))
        dims = oracle.u32(rng.integers(1, 9, D))
        pos = int(rng.integers(0, int(np.prod(dims))))
        delta = oracle.i32(rng.integers(-10, 11, D))
        loc = np.zeros(1, np.uint32)
        ok = R.ref_index_with_delta(D, dims, pos, delta, loc)
        iw.append([D] + list(dims) + [0] * (3 - D) + [pos] + list(delta) + [0] * (3 - D) + [int(loc[0]), ok])
    out["iwd"] = np.array(iw, np.int64)
    # complexAddAlphaxCxD
    rng = np.random.default_rng(6)
    n = 37
    dst = rng.standard_normal(2 * n).astype(np.float32)
    Cm = rng.standard_normal(2 * n).astype(np.float32)
    Dm = rng.standard_normal(2 * n).astype(np.float32)
    res = dst.copy()
    R.ref_complex_add(res, Cm, Dm, np.float32(-1.5), n)
    out.update(cadd_dst=dst, cadd_C=Cm, cadd_D=Dm, cadd_alpha=np.float32(-1.5), cadd_out=res)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


def bgrid_cases():
    """(dims, types, encoding, data) of the serialized-grid vectors."""
    rng = np.random.default_rng(77)
    cases = [((5,), [0], 0), ((4, 3), [1], 0), ((3, 2, 2), [0, 1], 1), ((2, 3), [1, 0, 1], 2), ((1, 1, 1, 2), [0], 0)]
    out = []
    for dims, types, enc in cases:
        n = int(np.prod(dims)) * len(types)
        data = rng.standard_normal(n).astype(np.float32)
        data[::3] = np.round(data[::3] * 3)
        if n > 4:
            data[4] = np.nan
        out.append((dims, types, enc, data))
    return out


def bgrid_mutations(raw, nd, nv):
    """(name, buffer) variants of one serialized grid: every header field,
    the size field and the length changed one at a time."""
    muts = []
    b = bytearray(raw)

    def put(buf, off, v):
        buf = bytearray(buf)
        buf[off:off + 4] = np.uint32(v).tobytes()
        return bytes(buf)

    sz = 8
    muts.append(("ok", bytes(b)))
    muts.append(("short", bytes(b[:-4])))
    muts.append(("long", bytes(b) + b"\0\0\0\0"))
    muts.append(("size_field", bytes(np.uint64(len(b) + 4).tobytes()) + bytes(b[8:])))
    muts.append(("ndim0", put(b, sz, 0)))
    muts.append(("ndim17", put(b, sz, 17)))
    muts.append(("ndim_plus1", put(b, sz, nd + 1)))
    muts.append(("dim0", put(b, sz + 4, 0)))
    muts.append(("dim_plus1", put(b, sz + 4, int(np.frombuffer(b[sz + 4:sz + 8], np.uint32)[0]) + 1)))
    vo = sz + 4 + 4 * nd
    muts.append(("nvar0", put(b, vo, 0)))
    muts.append(("nvar4097", put(b, vo, 4097)))
    muts.append(("type2", put(b, vo + 4, 2)))
    muts.append(("type1", put(b, vo + 4, 1)))
    eo = vo + 4 + 4 * nv
    muts.append(("enc3", put(b, eo, 3)))
    muts.append(("enc2", put(b, eo, 2)))
    muts.append(("tiny", bytes(b[:12])))
    return muts


def write_bgrid():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref is not built: run make -C oracle with the reference tree present")
    out = {}
    for i, (dims, types, enc, data) in enumerate(bgrid_cases()):
        cap = 1 << 16
        buf = np.zeros(cap, np.uint8)
        n = R.ref_bgrid_serialize(len(dims), oracle.u32(dims), len(types), oracle.i32(types), enc, data,
                                  buf.ctypes.data, cap)
        raw = buf[:n].tobytes()
        out[f"grid_{i}_dims"] = np.array(dims, np.uint32)
        out[f"grid_{i}_types"] = np.array(types, np.int32)
        out[f"grid_{i}_enc"] = np.int32(enc)
        out[f"grid_{i}_data"] = data
        out[f"grid_{i}_raw"] = np.frombuffer(raw, np.uint8)
        for name, m in bgrid_mutations(raw, len(dims), len(types)):
            mb = np.frombuffer(m, np.uint8).copy()
            out[f"mut_{i}_{name}_raw"] = mb
            out[f"mut_{i}_{name}_valid"] = np.int32(R.ref_bgrid_valid(mb.ctypes.data if mb.size else None, mb.size))
    np.savez_compressed(OUT_BGRID, **out)
    print(f"wrote {OUT_BGRID} ({os.path.getsize(OUT_BGRID)} bytes)")


def write_cross():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref is not built: run make -C oracle with the reference tree present")
    out = {}
    for nv in (1, 2, 3):
        t = np.zeros(nv, np.int32)
        a = np.zeros(64 * 64, np.float32)
        ax = np.zeros(64 * 64, np.float32)
        ct = np.zeros(64, np.int32)
        n = R.ref_coef_matrix_x(nv, t, np.zeros(1, np.int32), np.zeros(1, np.float32), a, ax, ct)
        out[f"cx_coef_{nv}_alpha"] = a[:n * n].copy()
        out[f"cx_coef_{nv}_alphaX"] = ax[:n * n].copy()
        out[f"cx_coef_{nv}_conv"] = ct[:n].copy()
    rng = np.random.default_rng(4321)
    cases = [((7, 5), 1), ((9,), 2), ((4, 3, 5), 1), ((6, 4), 3)]
    for i, (dims, nv) in enumerate(cases):
        cells = int(np.prod(dims))
        data = rng.standard_normal(cells * nv).astype(np.float32)
        data[rng.random(cells * nv) < 0.2] = np.nan
        f = oracle.fft_size(dims)
        F = int(np.prod(f))
        buf = np.zeros(3 * nv * F, np.float32)
        n = R.ref_convert_input4xcorr_x(len(dims), oracle.u32(dims), f, nv, np.zeros(nv, np.int32), data,
                                        np.zeros(1, np.int32), np.zeros(1, np.float32), buf)
        out[f"cx_sub_{i}_dims"] = np.array(dims, np.uint32)
        out[f"cx_sub_{i}_nvar"] = np.int32(nv)
        out[f"cx_sub_{i}_data"] = data
        out[f"cx_sub_{i}_out"] = buf[:n * F].copy()
    np.savez_compressed(OUT_CROSS, **out)
    print(f"wrote {OUT_CROSS} ({os.path.getsize(OUT_CROSS)} bytes)")


OFFSET_CASES = [((5, 5), False), ((7, 4), False), ((9,), False), ((5, 6, 3), False), ((11, 9, 7), False),
                ((41, 41), False), ((501, 501), False), ((5, 5), True), ((7, 4), True), ((9,), True),
                ((5, 6, 3), True), ((41, 41), True)]


def write_offsets():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref is not built: run make -C oracle with the reference tree present")
    rng = np.random.default_rng(97)
    out = {}
    for i, (kd, nan) in enumerate(OFFSET_CASES):
        cells = int(np.prod(kd))
        kern = None
        if nan:
            kern = np.ones(cells * 2, np.float32)
            kern[rng.random(kern.size) < 0.3] = np.nan
        cnt = int(np.prod([2 * ((d + 1) // 2) - 1 for d in kd]))
        res = np.zeros(cnt * len(kd), np.int32)
        n = R.ref_offsets(len(kd), oracle.u32(kd), 2, kern.ctypes.data if nan else None, res)
        out[f"off_{i}_dims"] = np.array(kd, np.uint32)
        if nan:
            out[f"off_{i}_kernel"] = kern
        out[f"off_{i}_out"] = res[: n * len(kd)].astype(np.int16)
    np.savez_compressed(OUT_OFFSETS, **out)
    print(f"wrote {OUT_OFFSETS} ({os.path.getsize(OUT_OFFSETS)} bytes)")


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "bgrid":
        write_bgrid()
    elif len(sys.argv) > 1 and sys.argv[1] == "cross":
        write_cross()
    elif len(sys.argv) > 1 and sys.argv[1] == "offsets":
        write_offsets()
    else:
        main()
        write_bgrid()
        write_cross()
        write_offsets()
