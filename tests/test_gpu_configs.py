"""The four exact-scan workloads at their full sizes on the GPU, each
checked node by node against the oracle on a seeded sample of path nodes.

Every node of the sample is recomputed by the oracle from the GPU's final
realisation restricted to what the serial loop would have seen (FP64 direct
map, device selection rule, the reference's draw). Configs 2 and 4 run the
exact selection: the GPU's choice must be the oracle's at every sampled node.
Configs 1 and 3 (categorical-only TIs, tie quantum) keep the FP32 selection:
a differing node is accepted only when the two FP64 values are within
quantum + 2 eps (see test_gpu_sim.lockstep). Sample sizes keep the oracle's
CPU time to about a minute per workload.

  config 1: TI 250x250 binary facies, sim 200x200, n=50, k=1.2
  config 2: TI 1000x1000 exp(GRF), sim 1000x1000, n=80, k=1.5 (the benchmark)
  config 3: TI 180x150x120 three facies, n=100, k=1.5, on a 60x60x30
            simulation (the full 200x200x100 grid takes minutes)
  config 4: TI 500x500 continuous + 4-class, sim 500x500 with 10 % hard
            data, n=60, k=1.5, per-variable weights 0.5 / 0.5
"""
import sys

import numpy as np
import pytest

from g2s_amd import synth
from test_gpu_sim import lockstep

pytestmark = pytest.mark.gpu


def _log(m):
    print(m, flush=True)
    sys.stdout.flush()


def _sample(n_nodes, count, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n_nodes, size=min(count, n_nodes), replace=False))


def _quantum(n, wmax=1.0):
    # the automatic tie quantum of categorical-only TIs: 2^(ceil(log2(n max|w|)) - 16)
    return float(2.0 ** (np.ceil(np.log2(n * wmax)) - 16))


def test_config1_categorical_2d_full():
    cfg = synth.config1_categorical_2d()
    sd = cfg["sim_dims"]
    di = np.full(int(np.prod(sd)), np.nan, np.float32)
    _log("config 1: TI 250x250 categorical, sim 200x200, n=50, k=1.2, waves of 256")
    nodes = _sample(int(np.prod(sd)), 10000, 11)
    sim, idx, st = lockstep(cfg["ti"], cfg["ti_dims"], di, sd, cfg["types"], cfg["n"], cfg["k"], cfg["seed"],
                            _quantum(cfg["n"]), nodes=nodes, wave_size=256, threads=16)
    _log(f"  facies 1 proportion: realisation {sim.mean():.3f}, TI {cfg['ti'].mean():.3f}")


def test_config2_continuous_2d_full():
    cfg = synth.config2_continuous_2d(1000)
    sd = cfg["sim_dims"]
    di = np.full(int(np.prod(sd)), np.nan, np.float32)
    _log("config 2: TI 1000x1000 continuous, sim 1000x1000, n=80, k=1.5, waves of 256")
    nodes = _sample(int(np.prod(sd)), 800, 12)
    lockstep(cfg["ti"], cfg["ti_dims"], di, sd, cfg["types"], cfg["n"], cfg["k"], cfg["seed"], 0.0, nodes=nodes,
             wave_size=256, threads=16)


def test_config3_categorical_3d_full_ti():
    cfg = synth.config3_categorical_3d()
    sd = (60, 60, 30)
    di = np.full(int(np.prod(sd)), np.nan, np.float32)
    _log("config 3: TI 180x150x120 three facies, sim 60x60x30, n=100, k=1.5, waves of 256")
    nodes = _sample(int(np.prod(sd)), 160, 13)
    lockstep(cfg["ti"], cfg["ti_dims"], di, sd, cfg["types"], cfg["n"], cfg["k"], cfg["seed"], _quantum(cfg["n"]),
             nodes=nodes, wave_size=256, threads=16)


def test_config4_multivariate_hard_data_full():
    cfg = synth.config4_multivariate_2d()
    sd = cfg["sim_dims"]
    _log("config 4: TI 500x500 continuous + categorical, sim 500x500, 10% hard data, n=60, k=1.5, weights 0.5/0.5")
    free = int(np.isnan(cfg["di"].reshape(-1, 2)).any(axis=1).sum())
    nodes = _sample(free, 2000, 14)
    lockstep(cfg["ti"], cfg["ti_dims"], cfg["di"], sd, cfg["types"], cfg["n"], cfg["k"], cfg["seed"], 0.0,
             nodes=nodes, weights=[0.5, 0.5], wave_size=256, threads=16)
