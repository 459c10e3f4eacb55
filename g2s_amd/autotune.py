"""Launch-parameter autotuner of the QuickSampling scan.

For one training-image geometry (extents, variable types, neighbour count)
the autotuner times short QS runs under candidate values of the scan's
tuning keys (``qsmi_set_default_tuning``) and keeps the fastest:

* 2-D TIs of extents up to 1024: the 1024 x 1024 grid against the TI's own
  FFT grid (``scan.grid1024_min``);
* on the own grid: the largest FFT radix of the plans (``fft.radix_max``),
  the transforms per FFT block (``fft.lognv_max``) and the path nodes per
  pencil block (``pencil.npb``), one key after the other (coordinate
  descent from the defaults);
* on the 1024 x 1024 grid: launched kernels against one captured HIP graph
  per wave step (``loop.graph``).

Choices are cached in a JSON file keyed by the geometry; a later call with
the same cache reads the choice instead of measuring again. The choice is
applied as the process-wide default for the TIs created afterwards.
``fft.radix_max`` / ``fft.lognv_max`` change the FP32 rounding of the maps:
with a continuous variable the exact selection makes the realisation
independent of them; categorical-only TIs keep the FP32 selection, whose ties
within the quantum may then resolve differently.
"""
import json
import os
import tempfile

import numpy as np

from . import qs, set_default_tuning

DEFAULTS = {"scan.grid1024_min": 320, "fft.radix_max": 13, "fft.lognv_max": 6, "pencil.npb": 1, "loop.graph": 0}
OWN_GRID_CANDIDATES = (("fft.radix_max", (13, 8, 5)), ("fft.lognv_max", (6, 4, 3)), ("pencil.npb", (1, 2, 4)))


def geometry_key(ti_dims, var_types, n):
    """Cache key of a TI geometry: dimensionality, extents, types, n."""
    return "D%d:%s:t%s:n%d" % (len(ti_dims), "x".join(str(int(d)) for d in ti_dims),
                                "".join(str(int(t)) for t in var_types), int(n))


def load_cache(path):
    """The cache's {geometry key: {tuning key: value}} (empty if absent)."""
    if not path or not os.path.exists(path):
        return {}
    with open(path) as fh:
        data = json.load(fh)
    if not isinstance(data, dict):
        raise ValueError(f"autotune cache {path}: not a JSON object")
    return data


def save_cache(path, cache):
    """Writes the cache atomically (temporary file + rename)."""
    d = os.path.dirname(os.path.abspath(path))
    fd, tmp = tempfile.mkstemp(prefix=".autotune", dir=d)
    with os.fdopen(fd, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(settings):
    """Makes `settings` ({tuning key: value}) the process-wide defaults."""
    for key, value in settings.items():
        set_default_tuning(key, int(value))


def reset_defaults():
    """Restores the library's default launch parameters."""
    apply(DEFAULTS)


def _probe_run(ti, ti_dims, var_types, n, k, sim_dims, seed):
    nvar = len(var_types)
    di = np.full(int(np.prod(sim_dims)) * nvar, np.nan, np.float32)
    _, _, st = qs(ti, ti_dims, di, sim_dims, var_types, n, k, seed)
    return st["path_loop_ms"]


def _time(settings, ti, ti_dims, var_types, n, k, sim_dims, repeats, seed):
    apply(settings)
    _probe_run(ti, ti_dims, var_types, n, k, sim_dims, seed)  # kernels loaded, buffers sized
    return float(np.median([_probe_run(ti, ti_dims, var_types, n, k, sim_dims, seed + 1 + r)
                            for r in range(repeats)]))


def autotune(ti, ti_dims, var_types, n, k, cache=None, sim_dims=None, repeats=3, seed=7):
    """Chooses the launch parameters for this TI geometry (from `cache` when
    it holds them, else by timing) and applies them. Returns (settings,
    measured) where measured is {description: path loop ms} (empty when the
    choice came from the cache)."""
    key = geometry_key(ti_dims, var_types, n)
    data = load_cache(cache)
    if key in data:
        settings = {k_: int(v) for k_, v in data[key].items()}
        apply(settings)
        return settings, {}
    if sim_dims is None:  # a short probe: a few thousand path nodes on the full TI
        side = 48 if len(ti_dims) <= 2 else 16
        sim_dims = tuple(min(int(d), side) for d in ti_dims)
    best = dict(DEFAULTS)
    measured = {}

    def trial(label, cand):
        t = _time(cand, ti, ti_dims, var_types, n, k, sim_dims, repeats, seed)
        measured[label] = t
        return t

    own_grid = True
    if len(ti_dims) == 2 and max(ti_dims) <= 1024:
        t1024 = trial("scan.grid1024_min=2", dict(best, **{"scan.grid1024_min": 2}))
        town = trial("scan.grid1024_min=1025", dict(best, **{"scan.grid1024_min": 1025}))
        own_grid = town <= t1024
        best["scan.grid1024_min"] = 1025 if own_grid else 2
    if not own_grid:
        times = {v: trial(f"loop.graph={v}", dict(best, **{"loop.graph": v})) for v in (0, 1)}
        best["loop.graph"] = min(times, key=times.get)
    if own_grid:
        for tkey, values in OWN_GRID_CANDIDATES:
            times = {v: trial(f"{tkey}={v}", dict(best, **{tkey: v})) for v in values}
            best[tkey] = min(times, key=times.get)
    apply(best)
    if cache:
        data = load_cache(cache)
        data[key] = best
        save_cache(cache, data)
    return best, measured
