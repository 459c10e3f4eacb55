"""qsmi_search_k_biggest_noalloc (the allocation-free variant of
qsmi_search_k_biggest, through the C-ABI) against qsmi_search_k_biggest on
the same maps: identical positions and values, on two calls in a row, the
second on a smaller TI with a smaller K (the scratch array keeps the first
call's contents), continuous (exact selection) and categorical (tie
quantum) maps, and the size limit."""
import numpy as np
import pytest

import g2s_amd
from qs_helpers import neighbourhood, neighbourhood_low
from test_gpu_scan import make_problem

pytestmark = pytest.mark.gpu


def _device(kind, dims, rng, nb):
    prob = make_problem(kind, dims, rng)
    ti = g2s_amd.TrainingImage(prob.sub, prob.dims)
    dev = g2s_amd.AcceleratorDevice(ti, prob.alpha)
    if kind == "cat":
        dev.set_tie_quantum(float(2.0 ** (np.ceil(np.log2(nb)) - 16)))
    return prob, ti, dev


@pytest.mark.parametrize("kind", ["cont", "cat"])
def test_noalloc_matches_search_k_biggest(kind):
    rng = np.random.default_rng(31)
    calls = [((160, 120), 40, 16), ((40, 30), 12, 2)]  # the second call smaller
    for dims, nb, K in calls:
        prob, ti, dev = _device(kind, dims, rng, nb)
        for trial in range(3):
            if kind == "cont" and trial == 2:
                off, x = neighbourhood_low(prob, rng, nb)
            else:
                off, x = neighbourhood(prob, rng, nb, 5)
            u, d0 = prob.convert(x)
            assert dev.candidate_for_pattern(off, u, d0)
            seed = float(rng.random())
            e1, i1 = dev.search_k_biggest(K, seed)
            e2, i2 = dev.search_k_biggest_noalloc(K, seed)
            assert np.array_equal(i1, i2), (dims, K, i1, i2)
            assert np.array_equal(e1, e2), (dims, K, e1, e2)
        with pytest.raises(g2s_amd.QsmiError):
            dev.search_k_biggest_noalloc(17, 0.5)
        dev.close()
        ti.close()
