"""Autotuner on the GPU: a choice is measured once and cached, and the
tuned launch parameters leave a continuous realisation unchanged (the exact
selection decides on FP64 values)."""
import json
import sys

import numpy as np
import pytest

import g2s_amd
from g2s_amd import autotune, synth

pytestmark = pytest.mark.gpu


def _log(m):
    print(m, flush=True)
    sys.stdout.flush()


def test_autotune_cache_and_invariance(tmp_path):
    g = synth.grf((90, 70), 5.0, seed=31)
    ti, dims = synth.to_g2s(np.exp(g))
    cache = str(tmp_path / "tune.json")
    try:
        s1, m1 = autotune.autotune(ti, dims, [0], 20, 1.5, cache=cache, sim_dims=(24, 24), repeats=1)
        _log(f"autotune {dims}: {s1} from {m1}")
        assert m1 and set(s1) == set(autotune.DEFAULTS)
        with open(cache) as fh:
            assert autotune.geometry_key(dims, [0], 20) in json.load(fh)
        s2, m2 = autotune.autotune(ti, dims, [0], 20, 1.5, cache=cache, sim_dims=(24, 24), repeats=1)
        assert m2 == {} and s2 == s1
        di = np.full(40 * 36, np.nan, np.float32)
        tuned = []
        for settings in (s1, {"scan.grid1024_min": 1025, "fft.radix_max": 5, "fft.lognv_max": 3, "pencil.npb": 4},
                         autotune.DEFAULTS):
            autotune.apply(settings)
            s, i, _ = g2s_amd.qs(ti, dims, di, (40, 36), [0], 20, 1.5, 9)
            tuned.append((s.tobytes(), i.tobytes()))
        assert tuned[0] == tuned[1] == tuned[2]
    finally:
        autotune.reset_defaults()
