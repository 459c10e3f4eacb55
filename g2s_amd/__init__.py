"""g2s_amd — MI355X-native QuickSampling exact scan behind G2S's device interface.

Python mirror of the host side. The compute path is libqsmi.so (HIP for
gfx950, include/qsmi.h); this module only marshals arrays.

  * :class:`TrainingImage` / :class:`AcceleratorDevice` mirror G2S's
    SharedMemoryManager + AcceleratorDevice (include/AcceleratorDevice.hpp:37-90):
    ``candidate_for_pattern``, ``search_k_biggest``, ``cvt_index_to_position``...
  * :meth:`AcceleratorDevice.scan_wave` runs a batch of independent path
    nodes (scan + select + draw) in one launch sequence.
  * :func:`qs` mirrors ``g2s('-a','qs', ...)`` / ``g2s::qs::runInMemory``
    (include/qsCore.hpp:17-58) for one TI.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import QsmiError, check, lib

__all__ = [
    "QsmiError",
    "lib",
    "fft_size",
    "device_count",
    "categories",
    "coef_matrix",
    "ti_subimages",
    "offsets",
    "kernel_generate",
    "path",
    "TrainingImage",
    "set_default_tuning",
    "AcceleratorDevice",
    "qs",
    "calibration",
]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def fft_size(n):
    """SharedMemoryManager size rule (src/sharedMemoryManager.cpp:26-58)."""
    return int(lib().qsmi_fft_size(int(n)))


def device_count():
    c = C.c_int(0)
    check(lib().qsmi_get_device_count(C.byref(c)))
    return c.value


def categories(ti, nvar, var):
    ti = _f32(ti).reshape(-1)
    out = np.zeros(4096, np.float32)
    cnt = C.c_int(0)
    check(lib().qsmi_categories(ti.size // nvar, nvar, ti, var, out, 4096, C.byref(cnt)))
    return out[: cnt.value].copy()


def coef_matrix(types, ncat=()):
    """(alpha[A,B], convTypes, subVarOf) of generateCoefMatrix4Xcorr."""
    types = _i32(types)
    ncat = _i32(list(ncat) or [0])
    alpha = np.zeros(256, np.float32)
    conv = np.zeros(16, np.int32)
    sub_of = np.zeros(16, np.int32)
    ns = C.c_int(0)
    check(lib().qsmi_coef_matrix(len(types), types, ncat, alpha, conv, sub_of, C.byref(ns)))
    n = ns.value
    # alpha[A + B*nSub] -> matrix [A, B]
    return alpha[: n * n].reshape(n, n).T.copy(), conv[:n].copy(), sub_of[:n].copy()


def ti_subimages(ti, dims, types, ncat=(), catvals=()):
    """convertInput4Xcorr images, unreflected, shape (nSub, prod(dims))."""
    dims = _u32(dims)
    types = _i32(types)
    ti = _f32(ti).reshape(-1)
    nsub = 2 * int(np.sum(types == 0)) + int(np.sum(ncat))
    out = np.zeros(nsub * int(np.prod(dims)), np.float32)
    check(lib().qsmi_ti_subimages(len(dims), dims, len(types), types, ti, _i32(list(ncat) or [0]),
                                  _f32(list(catvals) or [0]), out))
    return out.reshape(nsub, -1)


def offsets(kernel_size, kernel=None, knvar=1):
    """Sorted neighbour offsets of qs.cpp:1055-1137, shape (count, ndim)."""
    ks = _u32(kernel_size)
    cap = int(np.prod([2 * ((int(d) + 1) // 2) - 1 for d in ks]))
    out = np.zeros(cap * len(ks), np.int32)
    cnt = C.c_int(0)
    kp = None
    if kernel is not None:
        kernel = _f32(kernel)
        kp = kernel.ctypes.data
    check(lib().qsmi_offsets(len(ks), ks, knvar, kp, out, cap, C.byref(cnt)))
    return out[: cnt.value * len(ks)].reshape(cnt.value, len(ks)).copy()


def kernel_generate(dims, types, weights, alphas):
    """DataImage::genearteKernel: prod(dims) x nVar kernel (types 0 uniform,
    2 exponential, 7 gaussian per variable)."""
    types = _i32(types)
    out = np.zeros(int(np.prod(dims)) * types.size, np.float32)
    check(lib().qsmi_kernel_generate(len(dims), _u32(dims), types.size, types, _f32(weights), _f32(alphas), out))
    return out


def path(informed, seed):
    """(path, seeds, beginPath) of qs.cpp:1159-1258."""
    informed = np.ascontiguousarray(informed, dtype=np.uint8).reshape(-1)
    p = np.zeros(informed.size, np.uint32)
    s = np.zeros(informed.size, np.float32)
    b = C.c_uint(0)
    check(lib().qsmi_path(informed.size, informed, seed, p, s, C.byref(b)))
    return p, s, int(b.value)


class TrainingImage:
    """TI spectra resident on one GPU (SharedMemoryManager + allocAndInitSharedMemory)."""

    def __init__(self, sub_images, dims, gpu=0, fft_dims=None):
        dims = _u32(dims)
        sub = _f32(sub_images)
        self.dims = tuple(int(d) for d in dims)
        self.nsub = sub.shape[0]
        self.ndim = len(dims)
        h = C.c_void_p()
        fd = None if fft_dims is None else _u32(fft_dims)
        check(lib().qsmi_ti_create(gpu, len(dims), dims, None if fd is None else fd.ctypes.data, self.nsub,
                                   sub.reshape(-1), C.byref(h)))
        self._fd = fd
        self.h = h

    @classmethod
    def from_reflected(cls, images, dims, fft_dims, gpu=0):
        """From the sub-images as G2S's SharedMemoryManager hands them to
        allocAndInitSharedMemory: one point-reflected FFT-grid image each
        (qsmi_ti_create_from_reflected)."""
        dims = _u32(dims)
        fd = _u32(fft_dims)
        imgs = [_f32(im).reshape(-1) for im in images]
        ptrs = (C.c_void_p * len(imgs))(*[im.ctypes.data for im in imgs])
        self = cls.__new__(cls)
        self.dims = tuple(int(d) for d in dims)
        self.nsub = len(imgs)
        self.ndim = len(dims)
        h = C.c_void_p()
        check(lib().qsmi_ti_create_from_reflected(gpu, len(dims), dims, fd, len(imgs), C.cast(ptrs, C.c_void_p),
                                                  C.byref(h)))
        self._fd = fd
        self.h = h
        return self

    def set_tuning(self, key, value):
        """Launch parameter of the scan kernels (qsmi_ti_set_tuning)."""
        check(lib().qsmi_ti_set_tuning(self.h, key.encode(), int(value)))

    def get_tuning(self, key):
        v = C.c_int(0)
        check(lib().qsmi_ti_get_tuning(self.h, key.encode(), C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            lib().qsmi_ti_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def set_default_tuning(key, value):
    """Launch parameter every TI created afterwards starts with
    (qsmi_set_default_tuning); keys as TrainingImage.set_tuning."""
    check(lib().qsmi_set_default_tuning(key.encode(), int(value)))


def subimages_from_reflected(images, dims, fft_dims):
    """Unreflected TI-grid sub-images (nSub, prod(dims)) from G2S's reflected
    FFT-grid layout (qsmi_subimages_from_reflected, host only)."""
    dims = _u32(dims)
    fd = _u32(fft_dims)
    imgs = [_f32(im).reshape(-1) for im in images]
    ptrs = (C.c_void_p * len(imgs))(*[im.ctypes.data for im in imgs])
    out = np.zeros(len(imgs) * int(np.prod(dims)), np.float32)
    check(lib().qsmi_subimages_from_reflected(len(dims), dims, fd, len(imgs), C.cast(ptrs, C.c_void_p), out))
    return out.reshape(len(imgs), -1)


class AcceleratorDevice:
    """Mirror of G2S's AcceleratorDevice contract over the MI355X scan."""

    def __init__(self, ti, alpha, circular=False):
        alpha = np.asarray(alpha, np.float32)
        self.ti = ti
        self.nsub = alpha.shape[0]
        a = _f32(alpha.T.reshape(-1))  # [A + B*nSub]
        h = C.c_void_p()
        check(lib().qsmi_device_create(ti.h, a, int(circular), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().qsmi_device_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_true_mismatch(self, v):
        check(lib().qsmi_set_true_mismatch(self.h, int(bool(v))))

    def set_tie_quantum(self, q):
        check(lib().qsmi_set_tie_quantum(self.h, float(q)))

    def set_stage_timing(self, on):
        check(lib().qsmi_set_stage_timing(self.h, int(bool(on))))

    def last_stage_ms(self):
        out = np.zeros(4, np.float32)
        check(lib().qsmi_last_stage_ms(self.h, out))
        return out

    def candidate_for_pattern(self, offsets, values, delta0):
        """offsets (nb, ndim) relative to the cell; values (nb, nSub) converted."""
        offsets = _i32(offsets).reshape(-1)
        values = _f32(values).reshape(-1)
        d0 = _f32([delta0])
        comp = C.c_int(0)
        nb = values.size // self.nsub if self.nsub else 0
        check(lib().qsmi_candidate_for_pattern(self.h, nb, offsets, values, d0, C.byref(comp)))
        return bool(comp.value)

    def search_k_biggest(self, k, seed):
        err = np.zeros(k, np.float32)
        pos = np.zeros(k, np.uint32)
        check(lib().qsmi_search_k_biggest(self.h, err, pos, k, float(seed)))
        return err, pos

    def search_k_biggest_noalloc(self, k, seed):
        """search_k_biggest without any device allocation (the library's
        per-GPU scratch array; calls on one GPU must not overlap)."""
        err = np.zeros(k, np.float32)
        pos = np.zeros(k, np.uint32)
        check(lib().qsmi_search_k_biggest_noalloc(self.h, err, pos, k, float(seed)))
        return err, pos

    def cvt_index_to_position(self, index):
        out = C.c_uint(0)
        check(lib().qsmi_cvt_index_to_position(self.h, int(index), C.byref(out)))
        return out.value

    def cvt_position_to_index(self, position):
        out = C.c_uint(0)
        check(lib().qsmi_cvt_position_to_index(self.h, int(position), C.byref(out)))
        return out.value

    def array_size(self):
        out = C.c_uint(0)
        check(lib().qsmi_get_array_size(self.h, C.byref(out)))
        return out.value

    def error_map(self):
        out = np.zeros(int(np.prod(self.ti.dims)), np.float32)
        check(lib().qsmi_get_error_map(self.h, out))
        return out

    def set_value_in_error_array(self, index, value):
        check(lib().qsmi_set_value_in_error_array(self.h, int(index), float(value)))

    def set_cross_layer(self, alpha_x):
        """The count layer of a TI with NaN (an AcceleratorDevice built
        withCrossMeasurement): alpha_x (nSub, nSub), the second coefficient
        matrix; the TI holds the cross sub-images (ti_subimages_x)."""
        a = _f32(np.asarray(alpha_x, np.float32).T.reshape(-1))  # [A + B*nSub]
        check(lib().qsmi_device_set_cross_layer(self.h, a))

    def mask_layer_with_variable(self, layer, variable):
        """AcceleratorDevice::maskLayerWithVariable: layer 0 (map) or 1
        (count layer) times the mask image of ``variable``."""
        check(lib().qsmi_mask_layer_with_variable(self.h, int(layer), int(variable)))

    def compensate_missing_data(self):
        """AcceleratorDevice::compensateMissingData: map = -|E / C^4|, -inf
        where the count is zero."""
        check(lib().qsmi_compensate_missing_data(self.h))

    def cross_map(self):
        """The count layer in position space (getCrossErrorAtPosition)."""
        out = np.zeros(int(np.prod(self.ti.dims)), np.float32)
        check(lib().qsmi_get_cross_map(self.h, out))
        return out

    def set_value_in_error_array_with_radius(self, index, value, radius):
        """AcceleratorDevice::setValueInErrorArrayWithRadius: the Euclidean
        disk of ``radius`` around the position of ``index``."""
        check(lib().qsmi_set_value_in_error_array_with_radius(self.h, int(index), float(value), float(radius)))

    def synchronize(self):
        check(lib().qsmi_synchronize(self.h))

    @staticmethod
    def _pack(nodes):
        """nodes: list of (offsets (nb,ndim), values (nb,nSub), delta0, seed)."""
        starts = [0]
        offs, vals, d0, sd = [], [], [], []
        for o, v, d, s in nodes:
            o = np.asarray(o, np.int32)
            v = np.asarray(v, np.float32)
            starts.append(starts[-1] + o.shape[0])
            offs.append(o.reshape(-1))
            vals.append(v.reshape(-1))
            d0.append(d)
            sd.append(s)
        cat = lambda xs, dt: np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros(1, dt), dt)
        return (_i32(starts), cat(offs, np.int32), cat(vals, np.float32), _f32(d0), _f32(sd))

    def scan_wave(self, nodes, k):
        """Scan + select + draw for a wave of independent nodes.

        Returns (errors (n,K), positions (n,K), chosen (n,)) in TI positions."""
        starts, offs, vals, d0, sd = self._pack(nodes)
        n = len(nodes)
        K = int(np.ceil(k))
        err = np.zeros(n * K, np.float32)
        pos = np.zeros(n * K, np.uint32)
        ch = np.zeros(n, np.uint32)
        check(lib().qsmi_scan_wave(self.h, n, starts, offs, vals, d0, sd, float(k), err, pos, ch.ctypes.data))
        return err.reshape(n, K), pos.reshape(n, K), ch

    def distribution_wave(self, nodes, k, verbatim_pos, verbatim_radius):
        """QuantileSamplingModule::distribution for a wave of nodes on this TI:
        the ceil(k) best positions, best first, with no candidate within
        ``verbatim_radius[i]`` of ``verbatim_pos[i]`` (UINT32_MAX: none).

        Returns (errors (n,K), positions (n,K))."""
        starts, offs, vals, d0, sd = self._pack(nodes)
        n = len(nodes)
        K = int(np.ceil(k))
        err = np.zeros(n * K, np.float32)
        pos = np.zeros(n * K, np.uint32)
        check(lib().qsmi_distribution_wave(self.h, n, starts, offs, vals, d0, sd, float(k), _u32(verbatim_pos),
                                           _f32(verbatim_radius), err, pos))
        return err.reshape(n, K), pos.reshape(n, K)

    def wave(self, nodes, k):
        return Wave(self, nodes, k)


class Wave:
    """A path wave resident in HBM (for repeated timed launches)."""

    def __init__(self, dev, nodes, k):
        starts, offs, vals, d0, sd = AcceleratorDevice._pack(nodes)
        self.dev = dev
        self.n = len(nodes)
        self.K = int(np.ceil(k))
        h = C.c_void_p()
        check(lib().qsmi_wave_create(dev.h, self.n, starts, offs, vals, d0, sd, float(k), C.byref(h)))
        self.h = h

    def run(self):
        check(lib().qsmi_wave_run(self.h))

    def read(self):
        err = np.zeros(self.n * self.K, np.float32)
        pos = np.zeros(self.n * self.K, np.uint32)
        ch = np.zeros(self.n, np.uint32)
        check(lib().qsmi_wave_read(self.h, err.ctypes.data, pos.ctypes.data, ch.ctypes.data))
        return err.reshape(self.n, self.K), pos.reshape(self.n, self.K), ch

    def close(self):
        if getattr(self, "h", None):
            lib().qsmi_wave_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def qs(ti, ti_dims, di, sim_dims, var_types, n, k, seed, kernel=None, kernel_dims=None, gpu=0,
            wave_size=0, tie_quantum=-1.0, stage_timing=False, record=0, lp=None, lp_shortlist=0,
            lp_fallback_every=0, lp_fallback_mse=0.0, lp_audit=False, lp_min_cover=0.0, no_verbatim=False,
            n_image=None, k_image=None, kernels=None, kernels_dims=None, kernel_index_image=None,
            circular_sim=False, circular_ti=False, ti_index_image=None, ti_dims_all=None, progress=0):
    """In-memory QuickSampling (qsmi_qs_run_in_memory).

    Arrays are flat in G2S DataImage order (dimension 0 fastest, variable
    fastest of all): ti has prod(ti_dims)*nVar values, di prod(sim_dims)*nVar
    (NaN = simulate). Returns (simulation, index, stats dict).

    progress > 0: a progress line on stderr about every `progress` seconds.
    stage_timing: True (or 1) times every wave's stages with HIP events,
    k > 1 every k-th wave (stage totals scaled to all waves), False off.
    record > 0: stats["pairs"] = (offsets [m,80,2], values [m,80], best [m])
    of the first `record` path nodes (2-D, one continuous variable).
    lp: a LearnedProposal with keys built for this TI (learned-proposal
    form; the stats then carry its counters).
    ti may be a list of arrays of one extent (several training images, qs
    -ti a b ...); index is then position * nTI + TI. n_image / k_image: per
    simulation cell n and k (qs -ni / -kvi). kernels / kernels_dims /
    kernel_index_image: several kernels, one chosen per cell (qs -ki ... -kii)."""
    tis = None
    if isinstance(ti, (list, tuple)):
        tis = [_f32(t).reshape(-1) for t in ti]
        ti = tis[0]
        if len(tis) == 1:
            tis = None
    ti = _f32(ti).reshape(-1)
    di = _f32(di).reshape(-1)
    var_types = _i32(var_types)
    nvar = len(var_types)
    D = len(ti_dims)
    rq = _lib.QsRequest()
    rq.gpu = gpu
    rq.ndim = D
    for i in range(3):
        rq.tiSize[i] = int(ti_dims[i]) if i < D else 1
        rq.simSize[i] = int(sim_dims[i]) if i < D else 1
        rq.kernelSize[i] = (int(kernel_dims[i]) if (kernel_dims is not None and i < D) else 0)
    rq.nVar = nvar
    rq.varTypes = var_types.ctypes.data
    rq.ti = ti.ctypes.data
    rq.di = di.ctypes.data
    kern = None
    if kernel is not None:
        kern = _f32(kernel).reshape(-1)
        rq.kernel = kern.ctypes.data
    rq.n = int(n)
    rq.k = float(k)
    rq.seed = int(seed)
    rq.waveSize = int(wave_size)
    rq.tieQuantum = float(tie_quantum)
    rq.stageTiming = int(stage_timing)
    rq.progressSeconds = int(progress)
    rq.noVerbatim = int(bool(no_verbatim))
    rq.circularSim = int(bool(circular_sim))
    rq.circularTI = int(bool(circular_ti))
    tiimg = None
    if ti_index_image is not None:
        tiimg = _f32(ti_index_image).reshape(-1)
        rq.tiIndexImage = tiimg.ctypes.data
    nimg = kimg = None
    if n_image is not None:
        nimg = _f32(n_image).reshape(-1)
        rq.nImage = nimg.ctypes.data
    if k_image is not None:
        kimg = _f32(k_image).reshape(-1)
        rq.kImage = kimg.ctypes.data
    kptr = kimgi = ksz = None
    if kernels is not None and len(kernels) > 1:
        kernels = [_f32(x).reshape(-1) for x in kernels]
        kptr = (C.c_void_p * len(kernels))(*[x.ctypes.data for x in kernels])
        ksz = np.ones((len(kernels), 3), np.uint32)
        ksz[:, :D] = np.asarray(kernels_dims, np.uint32)
        kimgi = _f32(kernel_index_image).reshape(-1)
        rq.nKernels = len(kernels)
        rq.kernels = C.cast(kptr, C.c_void_p)
        rq.kernelSizes = ksz.ctypes.data
        rq.kernelIndexImage = kimgi.ctypes.data
    tptr = tsz = None
    if tis is not None:
        tptr = (C.c_void_p * len(tis))(*[t.ctypes.data for t in tis])
        rq.nTI = len(tis)
        rq.tis = C.cast(tptr, C.c_void_p)
        if ti_dims_all is not None:
            tsz = np.ones((len(tis), 3), np.uint32)
            tsz[:, :D] = np.asarray(ti_dims_all, np.uint32)
            rq.tiSizes = tsz.ctypes.data
    rec = None
    if record:
        rec = (np.zeros((int(record), 80, 2), np.int32), np.zeros((int(record), 80), np.float32),
               np.zeros(int(record), np.uint32))
        rq.recordCap = int(record)
        rq.recOffsets = rec[0].ctypes.data
        rq.recValues = rec[1].ctypes.data
        rq.recBest = rec[2].ctypes.data
    if lp is not None:
        rq.lp = lp.h
        rq.lpShortlist = int(lp_shortlist)
        rq.lpFallbackEvery = int(lp_fallback_every)
        rq.lpFallbackMse = float(lp_fallback_mse)
        rq.lpAudit = int(bool(lp_audit))
        rq.lpMinCover = float(lp_min_cover)
    sim = np.empty_like(di)
    idx = np.zeros(int(np.prod(sim_dims)), np.uint32)
    rs = _lib.QsResult()
    rs.simulation = sim.ctypes.data
    rs.index = idx.ctypes.data
    check(lib().qsmi_qs_run_in_memory(C.byref(rq), C.byref(rs)))
    stats = dict(setup_ms=rs.setupMs, path_loop_ms=rs.pathLoopMs, scan_ms=rs.scanMs, spectrum_ms=rs.stageMs[0],
                 gather_ms=rs.stageMs[1], rows_ms=rs.stageMs[2], merge_ms=rs.stageMs[3], nodes=rs.nodes,
                 scanned_nodes=rs.scannedNodes, waves=rs.waves, search_ms=rs.phaseMs[0], levels_ms=rs.phaseMs[1],
                 upload_ms=rs.phaseMs[2], device_loop_ms=rs.phaseMs[3], exact_nodes=rs.exactNodes,
                 band_positions=rs.bandPositions, second_pass_blocks=rs.secondPassBlocks,
                 full_rescore_nodes=rs.fullRescoreNodes)
    if rec is not None:
        m = min(int(record), int(rs.nodes))
        stats["pairs"] = (rec[0][:m], rec[1][:m], rec[2][:m])
    if lp is not None:
        stats.update(lp_proposed=rs.lpProposed, lp_fallbacks=rs.lpFallbacks, lp_overflows=rs.lpOverflows,
                     lp_shortlist_sum=rs.lpShortlistSum, lp_audit_nodes=rs.lpAuditNodes,
                     lp_audit_hits=rs.lpAuditHits, lp_audit_best_hits=rs.lpAuditBestHits)
    return sim, idx, stats


# auto_qs's defaults (src/auto_qs.cpp:115-121, 642-666)
DEFAULT_DENSITIES = (0.0078, 0.0156, 0.0312, 0.0625, 0.1250, 0.2500, 0.5000)


def default_neighbour_configs(max_nb=(80,)):
    """auto_qs's list of neighbour counts for ``-mnn`` values ``max_nb``:
    configurations i = 1 .. max_nb[0]-1, entry j = round(i*max_nb[j]/max_nb[0])
    with the integer division first, as src/auto_qs.cpp:654-664 computes it."""
    max_nb = [int(m) for m in max_nb]
    return np.array([[(i * m) // max_nb[0] for m in max_nb] for i in range(1, max_nb[0])], np.uint32)


def calibration(tis, ti_dims, var_types, kernels, kernel_dims, nb_configs=None, densities=DEFAULT_DENSITIES,
                max_k=1, power=2.0, max_iter=25000, min_iter=1000, max_seconds=0.0, noise=0.0, seed=0, gpu=0,
                wave_size=0, log_cap=0, stage_timing=False):
    """auto_qs calibration (src/auto_qs.cpp, include/calibration.hpp:30-253)
    with the scans on the GPU.

    tis: list of arrays prod(ti_dims) x nVar (one shape for all); kernels:
    list of arrays prod(kernel_dims) x nVar (one shape for all, kernel 0's
    extent sets the offsets). nb_configs: (nConfigs, width) neighbour counts
    (default: :func:`default_neighbour_configs`). Returns a dict with mean,
    dev, samples shaped (nDensity, nKernels, nConfigs, max_k) — auto_qs's
    im_1 / im_2 / im_3 images — and the run's statistics; with log_cap > 0
    also the first log_cap samples (setting, (TI, cell), max_k x (TI, pos))."""
    D = len(ti_dims)
    tis = [_f32(t) for t in tis]
    nvar = len(var_types)
    kernels = [_f32(k) for k in kernels]
    if nb_configs is None:
        nb_configs = default_neighbour_configs()
    nb = np.ascontiguousarray(np.atleast_2d(np.asarray(nb_configs, np.uint32)))
    dens = _f32(densities)
    tsz = np.ones((len(tis), 3), np.uint32)
    tsz[:, :D] = ti_dims
    ksz = np.ones((len(kernels), 3), np.uint32)
    ksz[:, :D] = kernel_dims
    tptr = (C.c_void_p * len(tis))(*[t.ctypes.data for t in tis])
    kptr = (C.c_void_p * len(kernels))(*[k.ctypes.data for k in kernels])
    types = _i32(var_types)
    n_set = nb.shape[0] * len(kernels) * dens.size
    mean = np.zeros(n_set * max_k, np.float32)
    dev = np.zeros(n_set * max_k, np.float32)
    cnt = np.zeros(n_set * max_k, np.uint32)
    lset = np.zeros(max(log_cap, 1), np.int32)
    lcell = np.zeros((max(log_cap, 1), 2), np.uint32)
    lpos = np.zeros((max(log_cap, 1), max_k, 2), np.uint32)
    rq = _lib.CalibRequest(gpu=gpu, ndim=D, nTI=len(tis), tiSize=tsz.ctypes.data, ti=C.cast(tptr, C.c_void_p),
                           nVar=nvar, varTypes=types.ctypes.data, nKernels=len(kernels), kernelSize=ksz.ctypes.data,
                           kernels=C.cast(kptr, C.c_void_p), nConfigs=nb.shape[0], nbWidth=nb.shape[1],
                           nbList=nb.ctypes.data, nDensity=dens.size, density=dens.ctypes.data, maxK=max_k,
                           power=power, maxIter=max_iter, minIter=min_iter, maxSeconds=max_seconds, noise=noise,
                           seed=seed, waveSize=wave_size, logCap=log_cap, logSetting=lset.ctypes.data,
                           logCell=lcell.ctypes.data, logPos=lpos.ctypes.data, stageTiming=int(stage_timing))
    rs = _lib.CalibResult(meanError=mean.ctypes.data, devError=dev.ctypes.data, samples=cnt.ctypes.data)
    check(lib().qsmi_calibration(C.byref(rq), C.byref(rs)))
    shape = (dens.size, len(kernels), nb.shape[0], max_k)
    out = dict(mean=mean.reshape(shape), dev=dev.reshape(shape), samples=cnt.reshape(shape),
               distributions=rs.distributions, iterations=rs.iterations, waves=rs.waves, ms=rs.ms,
               scan_ms=rs.scanMs, draw_ms=rs.drawMs, setup_ms=rs.setupMs, spectrum_ms=rs.stageMs[0],
               rows_ms=rs.stageMs[2], merge_ms=rs.stageMs[3])
    if log_cap > 0:
        n = min(log_cap, rs.distributions)
        out.update(log_setting=lset[:n], log_cell=lcell[:n], log_pos=lpos[:n])
    return out
