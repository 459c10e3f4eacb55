"""Shared helpers of the parity tests: small seeded QS problems, neighbourhoods,
the QSM conversion, and the map tolerance of the parity bar.

The oracle (``oracle``) is the checker; ``g2s_amd`` is the product under test.
"""
import numpy as np

import oracle

# The library's statement of its FP32 map error, restated: every value of a
# node's map is within eps = ERR_C * sum_a rms(Y_a) * sum_j |v_ja| of the exact
# (FP64) value, rms over the FFT grid of TI sub-image a, v_ja = sum_B
# alpha_aB u_jB the neighbours' alpha-combined converted values. The FFT
# rounding error of a correlation scales with the norms of its two operands,
# not with the local magnitude of the mismatch.
ERR_C = 4e-5

# Relative size of the map error against s(p) = sum_j w_j (x_j^2 +
# y(p+d_j)^2) (continuous) or 2 sum_j w_j (categorical), the magnitude of the
# cancelling terms: reported, and the bound of the cross-measurement form
# (TIs with NaN), whose selection stays in FP32.
MAP_RTOL = 1e-4


def grf2d(shape, rng, smooth=3.0):
    """Cheap seeded smooth field (box-blurred white noise), float32."""
    a = rng.standard_normal(shape)
    f = np.fft.rfftn(a)
    ky = np.fft.fftfreq(shape[0])[:, None] if len(shape) > 1 else None
    if len(shape) == 1:
        k2 = np.fft.rfftfreq(shape[0]) ** 2
    elif len(shape) == 2:
        k2 = ky ** 2 + np.fft.rfftfreq(shape[1])[None, :] ** 2
    else:
        k2 = (np.fft.fftfreq(shape[0])[:, None, None] ** 2 + np.fft.fftfreq(shape[1])[None, :, None] ** 2
              + np.fft.rfftfreq(shape[2])[None, None, :] ** 2)
    f *= np.exp(-k2 * (smooth * 2 * np.pi) ** 2 / 2)
    out = np.fft.irfftn(f, s=shape, axes=tuple(range(len(shape))))
    out = (out - out.mean()) / (out.std() + 1e-12)
    return out.astype(np.float32)


def g2s_dims(arr):
    """numpy C-order array (z, y, x) -> G2S dims (x, y, z) with dim0 fastest."""
    return tuple(reversed(arr.shape))


class Problem:
    """One TI in G2S order (dims[0] fastest) with its sub-images and alpha.

    ti: (cells, nvar) values; dims: G2S dims; types: per variable."""

    def __init__(self, ti, dims, types):
        self.types = np.asarray(types, np.int32)
        self.nvar = len(self.types)
        self.dims = tuple(int(d) for d in dims)
        self.D = len(self.dims)
        self.ti = np.ascontiguousarray(np.asarray(ti, np.float32).reshape(-1, self.nvar))
        self.ncat = []
        self.catvals = []
        for v in range(self.nvar):
            if self.types[v] != 0:
                c = oracle.categories(self.ti, self.nvar, v)
                self.ncat.append(len(c))
                self.catvals.extend(c.tolist())
        self.alpha, self.conv, self.subof = oracle.coef_matrix(self.types, self.ncat)
        self.nsub = self.alpha.shape[0]
        self.sub = oracle.ti_subimages(self.ti, self.dims, self.types, self.ncat, self.catvals)
        self.nexp = sum(self.ncat_iter())

    def ncat_iter(self):
        it = iter(self.ncat)
        for t in self.types:
            yield 1 if t == 0 else next(it)

    def expand(self, x):
        """one-hot expansion of raw neighbour values (nb, nvar) -> (nb, nexp)."""
        cols = []
        ci = 0
        cv = list(self.catvals)
        pos = 0
        for v, t in enumerate(self.types):
            if t == 0:
                cols.append(x[:, v])
            else:
                n = self.ncat[ci]
                for q in range(n):
                    cols.append((x[:, v] == cv[pos + q]).astype(np.float32))
                pos += n
                ci += 1
        return np.stack(cols, axis=1).astype(np.float32)

    def convert(self, x, w=None):
        """QSM conversion of raw values (nb, nvar) with weights w (nb,) -> (u, delta0)."""
        nb = x.shape[0]
        w = np.ones(nb, np.float32) if w is None else np.asarray(w, np.float32)
        xe = self.expand(x)
        u = np.zeros((nb, self.nsub), np.float32)
        for B in range(self.nsub):
            e = self.subof[B]
            xv = xe[:, e]
            ok = ~np.isnan(xv)
            u[:, B] = np.where(ok, w * (1.0 if self.conv[B] == 0 else xv), 0.0)
        d0 = np.float32(0)
        for v, t in enumerate(self.types):
            if t == 0:
                xv = xe[:, v]
                for j in range(nb):
                    if not np.isnan(xv[j]):
                        d0 = np.float32(d0 + np.float32(-1.0) * np.float32(w[j] * xv[j] * xv[j]))
        return u, float(d0)

    def sub_rms(self):
        """rms of every TI sub-image over the FFT grid (zero padding included)."""
        n_fft = int(np.prod([oracle.lib().orc_fft_size_1d(d) for d in self.dims]))
        y = self.sub.reshape(self.nsub, -1).astype(np.float64)
        return np.sqrt((y * y).sum(axis=1) / n_fft)

    def eps_bound(self, u):
        """The FP32 map error bound of a node with converted values u (nb, nsub)."""
        v = u.astype(np.float64) @ self.alpha.astype(np.float64).T
        return ERR_C * float(np.sum(self.sub_rms() * np.abs(v).sum(axis=0)))

    def oracle_map(self, off, u, d0, circular=False):
        return oracle.map_direct(self.dims, self.sub, self.alpha, off, u, d0, circular)

    def scale_map(self, off, x, w=None):
        """s(p) of the tolerance, FP64: per continuous variable
        sum_j |w_j| (x_j^2 + y(p+d_j)^2), per categorical 2 sum_j |w_j|."""
        nb = x.shape[0]
        w = np.ones(nb) if w is None else np.asarray(w, np.float64)
        total = np.zeros(int(np.prod(self.dims)))
        b = 0
        ci = 0
        for v, t in enumerate(self.types):
            if t == 0:
                alpha = np.zeros_like(self.alpha)
                alpha[b, b] = 1.0
                u = np.zeros((nb, self.nsub), np.float32)
                u[:, b] = np.abs(w)
                xv = x[:, v].astype(np.float64)
                const = float(np.nansum(np.abs(w) * xv * xv))
                total += oracle.map_direct(self.dims, self.sub, alpha, off, u, const)
                b += 2
            else:
                total += 2.0 * float(np.sum(np.abs(w)))
                b += self.ncat[ci]
                ci += 1
        return total


def neighbourhood(problem, rng, nb, radius, centre=True):
    """Random neighbourhood: offsets (nb, D) within radius (first = centre),
    raw values drawn from the TI marginal (nb, nvar), centre NaN."""
    D = problem.D
    offs = oracle.offsets([2 * radius + 1] * D)
    sel = np.sort(rng.choice(np.arange(1, offs.shape[0]), size=nb - 1, replace=False))
    o = np.concatenate([offs[:1], offs[sel]], axis=0) if centre else offs[sel]
    cells = rng.integers(0, problem.ti.shape[0], size=o.shape[0])
    x = problem.ti[cells].copy()
    if centre:
        x[0, :] = np.nan
    return o.astype(np.int32), x.astype(np.float32)


def neighbourhood_low(problem, rng, nb, q_lo=0.02, q_hi=0.10, keep=0.9, noise=0.05, var=0):
    """The hard case of an FP32 map on a heavy-tailed TI: a dense
    neighbourhood (the nb-1 nearest offsets, each kept with probability
    `keep`, then the next ones) copied, with multiplicative noise, from a TI
    cell whose value lies between the q_lo and q_hi quantiles of variable
    `var`. Its best matches sit in low-valued TI regions, where the FP32
    error set by the TI's global norms is largest against the mismatch gaps.
    Returns offsets (nb, D) with the centre first and raw values (centre NaN)."""
    D = problem.D
    dims = np.asarray(problem.dims)
    vals = problem.ti[:, var]
    lo, hi = np.quantile(vals, [q_lo, q_hi])
    offs = oracle.offsets([2 * 9 + 1] * D)
    rad = np.abs(offs).max(axis=1)
    for _ in range(1000):
        cands = np.nonzero((vals >= lo) & (vals <= hi))[0]
        src = int(cands[rng.integers(len(cands))])
        c = np.array(np.unravel_index(src, tuple(reversed(problem.dims)))[::-1])
        inside = np.all((c[None, :] + offs >= 0) & (c[None, :] + offs < dims[None, :]), axis=1)
        sel = [0]
        for j in range(1, offs.shape[0]):
            if len(sel) == nb:
                break
            if inside[j] and rng.random() < keep:
                sel.append(j)
        if len(sel) == nb and rad[sel].max() < 9:
            break
    o = offs[sel].astype(np.int32)
    t = c[None, :] + o
    flat = np.ravel_multi_index(tuple(t[:, ::-1].T), tuple(reversed(problem.dims)))
    x = problem.ti[flat].astype(np.float64)
    x *= 1.0 + noise * rng.standard_normal(x.shape)
    x = x.astype(np.float32)
    x[0, :] = np.nan
    return o, x


def valid_box(dims, off, circular=False):
    D = len(dims)
    mn = np.minimum(off.min(axis=0), 0)
    mx = np.maximum(off.max(axis=0), 0)
    if circular:
        return [0] * D, [d - 1 for d in dims]
    return [int(-mn[i]) for i in range(D)], [int(dims[i] - 1 - mx[i]) for i in range(D)]
