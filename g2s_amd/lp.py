"""Learned proposal for QuickSampling (qsmi_lp_* of include/qsmi.h).

Pattern encoders propose, for every path node, a shortlist of TI positions
that is rescored exactly; the FFT scan is the fallback and the judge. The
network (two MLP towers, bf16 MFMA products with fp32 accumulation) is
trained on the GPU by the library's own forward and backward kernels on
(neighbourhood, exact best position) pairs that plain QS runs record.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

NB = 80
IN = 3 * NB
EMB = 128


def canonical_offsets(ti_dims):
    """The 80 nearest non-centre offsets of the sorted offset list of the
    default kernel (src/qs.cpp:1055-1137 order): the key encoder's stencil."""
    from . import offsets

    offs = offsets([d // 2 + 1 for d in ti_dims])
    return np.ascontiguousarray(offs[1:NB + 1], np.int32)


def param_count():
    n = C.c_longlong(0)
    check(lib().qsmi_lp_param_count(C.byref(n)))
    return n.value


class LearnedProposal:
    def __init__(self, gpu=0, seed=1):
        h = C.c_void_p()
        check(lib().qsmi_lp_create(int(gpu), int(seed), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().qsmi_lp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # parameters -------------------------------------------------------
    def get_params(self):
        out = np.zeros(param_count(), np.float32)
        check(lib().qsmi_lp_get_params(self.h, out))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == param_count()
        check(lib().qsmi_lp_set_params(self.h, p))

    # primitives (tests) ------------------------------------------------
    def gemm(self, A, B, trans_a=False, trans_b=False):
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        M = A.shape[1] if trans_a else A.shape[0]
        K = A.shape[0] if trans_a else A.shape[1]
        N = B.shape[0] if trans_b else B.shape[1]
        out = np.zeros((M, N), np.float32)
        check(lib().qsmi_lp_gemm(self.h, int(trans_a), int(trans_b), M, N, K, A, B, out))
        return out

    def encode(self, tower, feats):
        feats = np.ascontiguousarray(feats, np.float32).reshape(-1, IN)
        out = np.zeros((feats.shape[0], EMB), np.float32)
        check(lib().qsmi_lp_encode(self.h, int(tower), feats.shape[0], feats, out))
        return out

    def loss_grad(self, qfeat, kfeat, inv_tau=10.0):
        qfeat = np.ascontiguousarray(qfeat, np.float32).reshape(-1, IN)
        kfeat = np.ascontiguousarray(kfeat, np.float32).reshape(-1, IN)
        grad = np.zeros(param_count(), np.float32)
        loss = C.c_float(0)
        check(lib().qsmi_lp_loss_grad(self.h, qfeat.shape[0], kfeat.shape[0], qfeat, kfeat, float(inv_tau),
                                      C.byref(loss), grad.ctypes.data))
        return loss.value, grad

    # training and keys ------------------------------------------------
    def train(self, ti, ti_dims, pairs, steps, batch=256, negatives=768, lr=1e-3, inv_tau=10.0, seed=1):
        """Adam steps on (neighbourhood, exact best) pairs recorded on one TI.
        Returns the mean loss of each block of 50 steps."""
        offs, vals, best = pairs
        ti = np.ascontiguousarray(ti, np.float32).reshape(-1)
        offs = np.ascontiguousarray(offs, np.int32)
        vals = np.ascontiguousarray(vals, np.float32)
        best = np.ascontiguousarray(best, np.uint32)
        canon = canonical_offsets(ti_dims)
        ts = _lib.LpTrainSet()
        ts.tiSize[0], ts.tiSize[1] = int(ti_dims[0]), int(ti_dims[1])
        ts.ti = ti.ctypes.data
        ts.canon = canon.ctypes.data
        ts.nPairs = int(best.size)
        ts.offsets = offs.ctypes.data
        ts.values = vals.ctypes.data
        ts.best = best.ctypes.data
        op = _lib.LpTrainOpts()
        op.steps, op.batch, op.negatives = int(steps), int(batch), int(negatives)
        op.lr, op.invTau, op.seed = float(lr), float(inv_tau), int(seed)
        loss = np.zeros(max(1, (int(steps) + 49) // 50), np.float32)
        check(lib().qsmi_lp_train(self.h, C.byref(ts), C.byref(op), loss.ctypes.data))
        return loss

    def build_keys(self, ti, ti_dims):
        ti = np.ascontiguousarray(ti, np.float32).reshape(-1)
        dims = np.array([int(ti_dims[0]), int(ti_dims[1])], np.uint32)
        check(lib().qsmi_lp_build_keys(self.h, dims, ti, canonical_offsets(ti_dims)))
