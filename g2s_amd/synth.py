"""Seeded synthetic training images for the benchmark workloads.

The reference ships no generator (its TIs are user data), so the workloads
are built here from Gaussian random fields: FP64 spectral synthesis by
circulant embedding of the covariance on a grid padded to twice the extent
(negative eigenvalues clipped), cropped, then standardised.

Arrays are returned in numpy C order (last axis = G2S dimension 0); use
:func:`to_g2s` for the flat G2S order (dim0 fastest, variable fastest).
"""
import numpy as np


def grf(shape, rng_range, seed, anis=(1.0,), angle_deg=0.0, model="exp"):
    """Standardised Gaussian random field of `shape` (C order).

    rng_range: range scalar or per axis (C order), the length scale of the
    exponential covariance exp(-h / range) ('exp') or of exp(-(h/range)^2)
    ('gauss'). anis/angle_deg (2-D only): anisotropy ratio
    major:minor with the major axis at angle_deg from axis x (last axis).
    """
    shape = tuple(int(s) for s in shape)
    D = len(shape)
    rngs = np.broadcast_to(np.asarray(rng_range, np.float64), (D,)).astype(np.float64)
    pad = tuple(2 * s for s in shape)
    coords = np.meshgrid(*[np.fft.fftfreq(p) * p for p in pad], indexing="ij")  # lag per axis (C order)
    lag = [c / r for c, r in zip(coords, rngs)]
    if D == 2 and (anis[0] != 1.0 or angle_deg != 0.0):
        a = np.deg2rad(angle_deg)
        x = coords[1] / rngs[1]
        y = coords[0] / rngs[0]
        u = x * np.cos(a) + y * np.sin(a)
        v = -x * np.sin(a) + y * np.cos(a)
        h = np.sqrt(u ** 2 + (v * anis[0]) ** 2)
    else:
        h = np.sqrt(sum(l ** 2 for l in lag))
    cov = np.exp(-h) if model == "exp" else np.exp(-(h ** 2))
    lam = np.fft.rfftn(cov).real
    lam = np.clip(lam, 0.0, None)
    rng = np.random.default_rng(seed)
    white = rng.standard_normal(pad)
    field = np.fft.irfftn(np.fft.rfftn(white) * np.sqrt(lam), s=pad, axes=tuple(range(D)))
    field = field[tuple(slice(0, s) for s in shape)]
    field = (field - field.mean()) / field.std()
    return field


def to_g2s(*fields):
    """C-order fields of equal shape -> flat G2S array (cells, nVar) float32 and G2S dims."""
    stack = np.stack([np.asarray(f, np.float64).reshape(-1) for f in fields], axis=1)
    dims = tuple(reversed(np.asarray(fields[0]).shape))
    return np.ascontiguousarray(stack, np.float32), dims


# ---------------------------------------------------------------------------
# the five workloads (GRF seeds fixed)
# ---------------------------------------------------------------------------
def config1_categorical_2d():
    """TI 250x250 binary facies: exp. cov. range 15, anisotropy 3:1 at 30 deg,
    class 1 above the 72nd percentile; sim 200x200, n=50, k=1.2."""
    g = grf((250, 250), 15.0, seed=101, anis=(3.0,), angle_deg=30.0)
    ti = (g > np.quantile(g, 0.72)).astype(np.float32)
    ti, dims = to_g2s(ti)
    return dict(ti=ti, ti_dims=dims, sim_dims=(200, 200), types=[1], n=50, k=1.2, seed=1)


def config2_continuous_2d(size=1000):
    """TI 1000x1000 log-normal: exp(GRF), exp. cov. range 20; sim 1000x1000,
    n=80, k=1.5 (the benchmark workload)."""
    g = grf((size, size), 20.0, seed=102)
    ti, dims = to_g2s(np.exp(g))
    return dict(ti=ti, ti_dims=dims, sim_dims=(size, size), types=[0], n=80, k=1.5, seed=2)


def config3_categorical_3d():
    """TI 180x150x120 (G2S dims) with 3 facies from two thresholded
    anisotropic GRFs (ranges 25/25/6 voxels), proportions 0.5/0.3/0.2;
    sim 200x200x100, n=100, k=1.5."""
    shape = (120, 150, 180)  # C order (z, y, x)
    ranges = (6.0, 25.0, 25.0)
    g1 = grf(shape, ranges, seed=1031)
    g2 = grf(shape, ranges, seed=1032)
    fac = np.zeros(shape, np.float32)
    fac[g1 > np.quantile(g1, 0.5)] = 1.0  # 50% stays 0
    sel = fac == 1.0
    thr = np.quantile(g2[sel], 0.4)  # of the upper half: 0.3 -> class 1, 0.2 -> class 2
    fac[sel & (g2 > thr)] = 2.0
    ti, dims = to_g2s(fac)
    return dict(ti=ti, ti_dims=dims, sim_dims=(200, 200, 100), types=[1], n=100, k=1.5, seed=3)


def config4_multivariate_2d():
    """TI 500x500 with a continuous GRF (range 12) and a 4-class map from a
    GRF correlated rho=0.7, thresholded at its quartiles; sim 500x500 with 10%
    seeded hard data; n=60, k=1.5."""
    a = grf((500, 500), 12.0, seed=1041)
    b = grf((500, 500), 12.0, seed=1042)
    c = 0.7 * a + np.sqrt(1 - 0.49) * b
    cat = np.digitize(c, np.quantile(c, [0.25, 0.5, 0.75])).astype(np.float32)
    ti, dims = to_g2s(a, cat)
    # hard data: 10% of the cells, values from an independent realisation
    ra = grf((500, 500), 12.0, seed=1043)
    rb = grf((500, 500), 12.0, seed=1044)
    rc = 0.7 * ra + np.sqrt(1 - 0.49) * rb
    rcat = np.digitize(rc, np.quantile(rc, [0.25, 0.5, 0.75])).astype(np.float32)
    hard, _ = to_g2s(ra, rcat)
    rng = np.random.default_rng(1045)
    di = np.full_like(hard, np.nan)
    pick = rng.choice(hard.shape[0], size=hard.shape[0] // 10, replace=False)
    di[pick] = hard[pick]
    return dict(ti=ti, ti_dims=dims, sim_dims=(500, 500), types=[0, 1], n=60, k=1.5, seed=4, di=di)
