"""ctypes bindings to the QuickSampling CPU oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this package, and only as the checker. The product (``g2s_amd``) never
imports it.

``lib()`` loads ``oracle/_build/liboracle.so`` (the restatement in
qs_oracle.cpp / qs_oracle_sim.cpp); ``ref()`` loads ``oracle/_ref/libqs_ref.so``
(the reference's own header-only pieces compiled in place) when it was built.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build(force=False):
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    target = os.path.join(HERE, "_build", "liboracle.so")
    if force or not os.path.exists(target):
        subprocess.run(["make", "-s", "-C", HERE], check=True)


def _sig(lib, name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
    return fn


class OrcSim(C.Structure):
    _fields_ = [
        ("D", C.c_int),
        ("tiDims", C.c_uint * 3),
        ("nVar", C.c_int),
        ("types", C.c_void_p),
        ("ti", C.c_void_p),
        ("ncat", C.c_void_p),
        ("catvals", C.c_void_p),
        ("simDims", C.c_uint * 3),
        ("kdims", C.c_uint * 3),
        ("kernel", C.c_void_p),
        ("nOffsets", C.c_int),
        ("offsets", C.c_void_p),
        ("n", C.c_int),
        ("k", C.c_float),
        ("path", C.c_void_p),
        ("seeds", C.c_void_p),
        ("beginPath", C.c_uint),
        ("pathLen", C.c_uint),
        ("mode", C.c_int),
        ("quantum", C.c_float),
        ("cross", C.c_int),
        ("noVerbatim", C.c_int),
        ("nTI", C.c_int),
        ("nImage", C.c_void_p),
        ("kImage", C.c_void_p),
        ("circularSim", C.c_int),
        ("circularTI", C.c_int),
        ("tiImage", C.c_void_p),
        ("tiDimsAll", C.c_void_p),
        ("di0", C.c_void_p),
    ]


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(HERE, "_build", "liboracle.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    _sig(L, "orc_fft_size_1d", C.c_uint, [C.c_uint])
    _sig(L, "orc_fft_size", None, [C.c_int, u32p, u32p])
    _sig(L, "orc_categories", C.c_int, [C.c_size_t, C.c_int, f32p, C.c_int, f32p, C.c_int])
    _sig(L, "orc_ti_subimages", C.c_int, [C.c_int, u32p, u32p, C.c_int, i32p, f32p, i32p, f32p, C.c_int, f32p])
    _sig(L, "orc_coef_matrix", C.c_int, [C.c_int, i32p, i32p, f32p, i32p, i32p])
    _sig(L, "orc_coef_matrix_x", C.c_int, [C.c_int, i32p, f32p, f32p, i32p, i32p])
    _sig(L, "orc_ti_subimages_x", C.c_int, [C.c_int, u32p, u32p, C.c_int, i32p, f32p, C.c_int, f32p])
    _sig(L, "orc_kernel_generate", None, [C.c_int, u32p, C.c_int, i32p, f32p, f32p, f32p])
    _sig(L, "orc_kernel_expand", C.c_int, [C.c_size_t, C.c_int, f32p, i32p, f32p])
    _sig(L, "orc_offsets", C.c_int, [C.c_int, u32p, C.c_int, C.c_void_p, i32p])
    _sig(L, "orc_path", C.c_uint, [C.c_size_t, u8p, C.c_uint, u32p, f32p])
    _sig(L, "orc_fkst", None, [f32p, C.c_uint, C.c_ushort, C.c_float, f32p, u32p])
    _sig(L, "orc_seed_from_unit_float", C.c_uint, [C.c_float])
    _sig(L, "orc_tie_hash", C.c_uint32, [C.c_uint32, C.c_uint32])
    _sig(L, "orc_select_device", C.c_int, [C.c_int, u32p, f32p, i32p, i32p, C.c_int, C.c_float, C.c_float, f32p, u32p])
    _sig(L, "orc_select_device_f64", C.c_int,
         [C.c_int, u32p, f64p, i32p, i32p, C.c_int, C.c_float, C.c_float, f64p, u32p])
    _sig(L, "orc_draw_rank", C.c_uint, [C.c_float, C.c_float, C.c_uint])
    _sig(L, "orc_index_to_position", C.c_uint, [C.c_int, u32p, u32p, i32p, C.c_uint])
    _sig(L, "orc_position_to_index", C.c_uint, [C.c_int, u32p, u32p, i32p, C.c_uint])
    _sig(L, "orc_bbox", None, [C.c_int, C.c_int, i32p, i32p, i32p])
    _sig(L, "orc_map_direct_f64", None, [C.c_int, u32p, C.c_int, f32p, f32p, C.c_int, i32p, f32p, C.c_double, C.c_int, f64p])
    _sig(L, "orc_map_refflow_f64", None, [C.c_int, u32p, u32p, C.c_int, f64p, f32p, C.c_int, i32p, f32p, C.c_float, C.c_int, C.c_int, f32p])
    _sig(L, "orc_map_refflow_f32", None, [C.c_int, u32p, u32p, C.c_int, f32p, f32p, C.c_int, i32p, f32p, C.c_float, C.c_int, C.c_int, f32p])
    _sig(L, "orc_ti_spectra_f64", None, [C.c_int, u32p, C.c_int, f32p, f64p])
    _sig(L, "orc_ti_spectra_f32", None, [C.c_int, u32p, C.c_int, f32p, f32p])
    _sig(L, "orc_fft1d_f64", None, [C.c_size_t, f64p, C.c_int])
    _sig(L, "orc_rfftn_f64", None, [C.c_int, u32p, f64p, f64p])
    _sig(L, "orc_irfftn_f64", None, [C.c_int, u32p, f64p, f64p])
    _sig(L, "orc_convert_neighbors", C.c_float, [C.c_int, C.c_int, i32p, C.c_int, f32p, C.c_int, i32p, i32p, C.c_int, i32p, u32p, f32p, f32p])
    _sig(L, "orc_index_with_delta", C.c_int, [C.c_int, u32p, C.c_uint, i32p, u32p])
    _sig(L, "orc_sim_create", C.c_void_p, [C.POINTER(OrcSim)])
    _sig(L, "orc_sim_free", None, [C.c_void_p])
    _sig(L, "orc_sim_posterior", None, [C.c_void_p, u32p])
    _sig(L, "orc_sim_nsub", C.c_int, [C.c_void_p])
    _sig(L, "orc_sim_gather", C.c_int, [C.c_void_p, C.c_uint, i32p])
    _sig(L, "orc_sim_node", C.c_uint, [C.c_void_p, f32p, C.c_uint, C.c_int, f32p, u32p, C.POINTER(C.c_int)])
    _sig(L, "orc_sim_run", None, [C.c_void_p, f32p, u32p])
    _sig(L, "orc_sim_node_map", C.c_int, [C.c_void_p, f32p, C.c_uint, f64p, C.c_void_p])
    _sig(L, "orc_sim_node_norm", C.c_double, [C.c_void_p, f32p, C.c_uint])
    _sig(L, "orc_sim_set_index", None, [C.c_void_p, C.c_void_p])
    _sig(L, "orc_calib_offsets", C.c_int, [C.c_int, u32p, i32p, C.c_int])
    _sig(L, "orc_exclude_disk", None, [C.c_int, u32p, f32p, C.c_uint, C.c_float])
    _sig(L, "orc_marginal_cell", C.c_uint, [C.c_void_p, C.c_size_t, C.c_int, C.c_float])
    _sig(L, "orc_calibration", C.c_int, [C.c_int, C.c_int, u32p, C.c_void_p, C.c_int, i32p, C.c_int, u32p, C.c_void_p,
                                         C.c_int, C.c_int, u32p, C.c_int, f32p, C.c_int, C.c_float, C.c_uint,
                                         C.c_uint, C.c_float, C.c_uint, C.c_float, f32p, f32p, u32p, C.c_int, i32p,
                                         u32p, u32p, C.POINTER(C.c_longlong)])
    _LIB = L
    return L


def ref():
    """The reference's compiled header-only pieces, or None when not built."""
    global _REF
    if _REF is not None:
        return _REF
    path = os.path.join(HERE, "_ref", "libqs_ref.so")
    if not os.path.exists(path):
        return None
    R = C.CDLL(path)
    _sig(R, "ref_fft_size", None, [C.c_int, u32p, u32p])
    _sig(R, "ref_fkst", None, [f32p, C.c_uint, C.c_ushort, C.c_float, f32p, u32p])
    _sig(R, "ref_convert_input4xcorr", C.c_int, [C.c_int, u32p, u32p, C.c_int, i32p, f32p, i32p, f32p, f32p])
    _sig(R, "ref_coef_matrix", C.c_int, [C.c_int, i32p, i32p, f32p, f32p, i32p, i32p, f32p])
    _sig(R, "ref_generate_kernel", None, [C.c_int, u32p, C.c_int, i32p, f32p, f32p, f32p])
    _sig(R, "ref_kernel_expand", C.c_int, [C.c_int, u32p, C.c_int, f32p, i32p, f32p])
    _sig(R, "ref_offsets", C.c_int, [C.c_int, u32p, C.c_int, C.c_void_p, i32p])
    _sig(R, "ref_index_with_delta", C.c_int, [C.c_int, u32p, C.c_uint, i32p, u32p])
    _sig(R, "ref_complex_add", None, [f32p, f32p, f32p, C.c_float, C.c_uint])
    _sig(R, "ref_bgrid_serialize", C.c_size_t, [C.c_int, u32p, C.c_int, i32p, C.c_int, f32p, C.c_void_p, C.c_size_t])
    _sig(R, "ref_bgrid_valid", C.c_int, [C.c_void_p, C.c_size_t])
    _sig(R, "ref_convert_input4xcorr_x", C.c_int, [C.c_int, u32p, u32p, C.c_int, i32p, f32p, i32p, f32p, f32p])
    _sig(R, "ref_coef_matrix_x", C.c_int, [C.c_int, i32p, i32p, f32p, f32p, f32p, i32p])
    _REF = R
    return R


# ----------------------------------------------------------------------------
# Convenience wrappers (numpy in, numpy out)
# ----------------------------------------------------------------------------
def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def fft_size(dims):
    dims = u32(dims)
    out = np.zeros_like(dims)
    lib().orc_fft_size(len(dims), dims, out)
    return out


def fkst(data, k, seed):
    data = f32(data)
    out = np.zeros(k, np.float32)
    pos = np.zeros(k, np.uint32)
    lib().orc_fkst(data, data.size, k, seed, out, pos)
    return out, pos


def categories(ti, nvar, var, maxc=256):
    ti = f32(ti).reshape(-1)
    out = np.zeros(maxc, np.float32)
    n = lib().orc_categories(ti.size // nvar, nvar, ti, var, out, maxc)
    return out[:n].copy()


def coef_matrix(types, ncat):
    types = i32(types)
    ncat = i32(ncat if len(ncat) else [0])
    alpha = np.zeros(64 * 64, np.float32)
    conv = np.zeros(64, np.int32)
    sub_of = np.zeros(64, np.int32)
    nsub = lib().orc_coef_matrix(len(types), types, ncat, alpha, conv, sub_of)
    # alpha[A + B*nSub] -> matrix [A, B]
    return alpha[: nsub * nsub].reshape(nsub, nsub).T.copy(), conv[:nsub].copy(), sub_of[:nsub].copy()


def ti_subimages(ti, dims, types, ncat, catvals, fdims=None, reflected=False):
    """ti: array (cells, nVar) in dim0-fastest order; returns (nSub, F) f-grid images."""
    dims = u32(dims)
    fdims = dims if fdims is None else u32(fdims)
    types = i32(types)
    ti = f32(ti).reshape(-1)
    nsub = sum(2 if t == 0 else 0 for t in types) + int(np.sum(ncat))
    F = int(np.prod(fdims))
    out = np.zeros(nsub * F, np.float32)
    lib().orc_ti_subimages(len(dims), dims, fdims, len(types), types, ti, i32(ncat if len(ncat) else [0]),
                           f32(catvals if len(catvals) else [0]), int(reflected), out)
    return out.reshape(nsub, F)


def offsets(kdims, kernel=None, knvar=1):
    kdims = u32(kdims)
    cnt = int(np.prod([2 * ((int(d) + 1) // 2) - 1 for d in kdims]))
    out = np.zeros(cnt * len(kdims), np.int32)
    kp = None
    if kernel is not None:
        kernel = f32(kernel)
        kp = kernel.ctypes.data
    n = lib().orc_offsets(len(kdims), kdims, knvar, kp, out)
    return out[: n * len(kdims)].reshape(n, len(kdims)).copy()


def marginal_cell(complete, nvar, seed):
    """The marginal draw's TI cell for `seed` (one TI, complete: 1 per cell
    without NaN)."""
    c = np.ascontiguousarray(complete, np.uint8)
    return int(lib().orc_marginal_cell(c.ctypes.data, c.size, int(nvar), float(seed)))


def path(informed, seed):
    informed = np.ascontiguousarray(informed, dtype=np.uint8)
    p = np.zeros(informed.size, np.uint32)
    s = np.zeros(informed.size, np.float32)
    begin = lib().orc_path(informed.size, informed, seed, p, s)
    return p, s, int(begin)


def map_direct(dims, Y, alpha, off, u, delta0, circular=False):
    dims = u32(dims)
    off = i32(off)
    u = f32(u)
    Y = f32(Y)
    alpha = f32(np.asarray(alpha, np.float32).T.reshape(-1))  # alpha[A + B*nSub] layout
    S = int(np.prod(dims))
    out = np.zeros(S, np.float64)
    nsub = Y.shape[0]
    lib().orc_map_direct_f64(len(dims), dims, nsub, Y.reshape(-1), alpha, off.shape[0], off.reshape(-1),
                             u.reshape(-1), float(delta0), int(circular), out)
    return out


def select_device(dims, m, lo, hi, K, seed, quantum=0.0):
    dims = u32(dims)
    m = f32(m)
    outv = np.zeros(K, np.float32)
    outp = np.zeros(K, np.uint32)
    n = lib().orc_select_device(len(dims), dims, m, i32(lo), i32(hi), K, seed, quantum, outv, outp)
    return outv, outp, n


def select_device_f64(dims, m, lo, hi, K, seed, quantum=0.0):
    """The device rule on an FP64 map (the exact selection's result)."""
    dims = u32(dims)
    m = np.ascontiguousarray(m, np.float64)
    outv = np.zeros(K, np.float64)
    outp = np.zeros(K, np.uint32)
    n = lib().orc_select_device_f64(len(dims), dims, m, i32(lo), i32(hi), K, seed, quantum, outv, outp)
    return outv, outp, n


class Sim:
    """Serial QS simulation / per-node lockstep checker over the oracle."""

    def __init__(self, ti, tidims, types, ncat, catvals, simdims, kdims, kernel_exp, offs, n, k, path, seeds,
                 begin, mode=0, quantum=0.0, cross=False, no_verbatim=False, n_ti=1, n_image=None,
                 k_image=None, circular_sim=False, circular_ti=False, ti_image=None, ti_dims_all=None, di0=None):
        """di0: the simulation grid as given, when it holds partially informed
        cells (a node's own values when it is reached)."""
        self._keep = []

        def keep(a):
            self._keep.append(a)
            return a.ctypes.data

        D = len(tidims)
        s = OrcSim()
        s.D = D
        for i in range(3):
            s.tiDims[i] = int(tidims[i]) if i < D else 1
            s.simDims[i] = int(simdims[i]) if i < D else 1
            s.kdims[i] = int(kdims[i]) if i < D else 1
        types = i32(types)
        s.nVar = len(types)
        s.types = keep(types)
        s.ti = keep(f32(ti).reshape(-1))
        s.ncat = keep(i32(ncat if len(ncat) else [0]))
        s.catvals = keep(f32(catvals if len(catvals) else [0]))
        s.kernel = keep(f32(kernel_exp).reshape(-1))
        offs = i32(offs)
        s.nOffsets = offs.shape[0]
        s.offsets = keep(offs.reshape(-1))
        s.n = int(n)
        s.k = float(k)
        path = u32(path)
        s.path = keep(path)
        s.seeds = keep(f32(seeds))
        s.beginPath = int(begin)
        s.pathLen = path.size
        s.mode = int(mode)
        s.quantum = float(quantum)
        s.cross = int(bool(cross))
        s.noVerbatim = int(bool(no_verbatim))
        s.nTI = int(n_ti)  # ti then holds the n_ti images one after the other
        s.circularSim = int(bool(circular_sim))
        s.circularTI = int(bool(circular_ti))
        if ti_image is not None:
            s.tiImage = keep(f32(ti_image).reshape(-1))
        if ti_dims_all is not None:  # several TIs of their own extents, nTI x D
            tda = np.ones((len(ti_dims_all), 3), np.uint32)
            tda[:, :D] = np.asarray(ti_dims_all, np.uint32)
            s.tiDimsAll = keep(tda.reshape(-1))
        if n_image is not None:  # per cell n; n must be the largest value
            s.nImage = keep(f32(n_image).reshape(-1))
        if k_image is not None:
            s.kImage = keep(f32(k_image).reshape(-1))
        if di0 is not None:
            s.di0 = keep(f32(di0).reshape(-1))
        self.spec = s
        self._index = None
        self.nvar = len(types)
        self.K = int(np.ceil(k))
        self.h = lib().orc_sim_create(C.byref(s))
        self.ncells = int(np.prod(simdims))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_free(self.h)
            self.h = None

    def posterior(self):
        out = np.zeros(self.ncells, np.uint32)
        lib().orc_sim_posterior(self.h, out)
        return out

    def gather(self, i):
        out = np.zeros(max(self.spec.n, 1), np.int32)
        c = lib().orc_sim_gather(self.h, i, out)
        return out[:c].copy()

    def set_index(self, idx):
        """The importDataIndex state -nV reads (kept alive here)."""
        self._index = np.ascontiguousarray(idx, np.uint32)
        lib().orc_sim_set_index(self.h, self._index.ctypes.data)

    def node(self, di, i):
        v = np.zeros(self.K, np.float32)
        p = np.zeros(self.K, np.uint32)
        nb = C.c_int(0)
        chosen = lib().orc_sim_node(self.h, f32(di).reshape(-1), i, self.K, v, p, C.byref(nb))
        return chosen, v, p, nb.value

    def node_map(self, di, i, with_scale=True):
        """The node's exact FP64 map and s(p) (position space)."""
        m = np.zeros(int(np.prod(self.spec.tiDims[: self.spec.D])), np.float64)
        sc = np.zeros_like(m) if with_scale else None
        nb = lib().orc_sim_node_map(self.h, f32(di).reshape(-1), i, m, None if sc is None else sc.ctypes.data)
        return m, sc, nb

    def node_norm(self, di, i):
        """sum_A rms(Y_A) sum_j |v_jA| of node i: the scale of its FP32 map error."""
        return float(lib().orc_sim_node_norm(self.h, f32(di).reshape(-1), i))

    def run(self, di):
        di = f32(di).reshape(-1).copy()
        idx = np.zeros(self.ncells, np.uint32)  # hard data keep 0, as the reference's empty index image
        lib().orc_sim_run(self.h, di, idx)
        return di, idx


def calib_offsets(kdims):
    """Offsets in auto_qs's calibration order (src/auto_qs.cpp:449-492)."""
    kd = u32(kdims)
    D = len(kdims)
    cap = int(np.prod([2 * ((k + 1) // 2) - 1 for k in kdims]))
    out = np.zeros(cap * D, np.int32)
    n = lib().orc_calib_offsets(D, kd, out, cap)
    return out[:n * D].reshape(n, D)


def calibration(tis, ti_dims, types, kernels, kernel_dims, nb_configs, densities, max_k, power, max_iter, min_iter,
                noise, seed, quantum=0.0, log_cap=0):
    """The calibration loop restated on the CPU (brute-force maps); arrays
    shaped (nDensity, nKernels, nConfigs, max_k) as the product returns."""
    D = len(ti_dims)
    tis = [f32(t) for t in tis]
    kernels = [f32(k) for k in kernels]
    nb = np.ascontiguousarray(np.atleast_2d(np.asarray(nb_configs, np.uint32)))
    dens = f32(densities)
    tsz = np.ones((len(tis), 3), np.uint32)
    tsz[:, :D] = ti_dims
    ksz = np.ones((len(kernels), 3), np.uint32)
    ksz[:, :D] = kernel_dims
    tptr = (C.c_void_p * len(tis))(*[t.ctypes.data for t in tis])
    kptr = (C.c_void_p * len(kernels))(*[k.ctypes.data for k in kernels])
    n_set = nb.shape[0] * len(kernels) * dens.size
    mean = np.zeros(n_set * max_k, np.float32)
    dev = np.zeros(n_set * max_k, np.float32)
    cnt = np.zeros(n_set * max_k, np.uint32)
    lc = max(log_cap, 1)
    lset = np.zeros(lc, np.int32)
    lcell = np.zeros(lc * 2, np.uint32)
    lpos = np.zeros(lc * max_k * 2, np.uint32)
    nd = C.c_longlong(0)
    lib().orc_calibration(D, len(tis), tsz.reshape(-1), C.cast(tptr, C.c_void_p), len(types), i32(types),
                          len(kernels), ksz.reshape(-1), C.cast(kptr, C.c_void_p), nb.shape[0], nb.shape[1],
                          nb.reshape(-1), dens.size, dens, max_k, power, max_iter, min_iter, noise, seed, quantum,
                          mean, dev, cnt, log_cap, lset, lcell, lpos, C.byref(nd))
    shape = (dens.size, len(kernels), nb.shape[0], max_k)
    out = dict(mean=mean.reshape(shape), dev=dev.reshape(shape), samples=cnt.reshape(shape), distributions=nd.value)
    if log_cap > 0:
        n = min(log_cap, nd.value)
        out.update(log_setting=lset[:n], log_cell=lcell.reshape(-1, 2)[:n],
                   log_pos=lpos.reshape(-1, max_k, 2)[:n])
    return out
