"""Ensemble statistics of long stochastic runs, GPU against the oracle.

Trajectories with noise diverge between any two implementations once rounding
differences are amplified, so long runs are compared by a statistic of the
ensemble: the demixing index of the two-species sorting system (the mean, over
agents with at least one neighbour within r, of the fraction of neighbours of
the agent's own species; 1/2 for a random mixture, 1 for complete sorting).
No such index exists in the reference; this one is defined here.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from cellular_raza_amd import MechanicsBackend, workloads
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu


def demixing_index(pos, species, r):
    pairs = cKDTree(pos).query_pairs(r, output_type="ndarray")
    same = (species[pairs[:, 0]] == species[pairs[:, 1]]).astype(float)
    n = np.bincount(pairs.ravel(), minlength=len(pos)).astype(float)
    s = np.bincount(pairs[:, 0], weights=same, minlength=len(pos)) + np.bincount(pairs[:, 1], weights=same,
                                                                                minlength=len(pos))
    ok = n > 0
    return float((s[ok] / n[ok]).mean())


def test_sorting_demixing_index_statistics():
    """3000-agent sorting system (the benchmark's interaction and Langevin parameters),
    800 steps from one initial state, 8 Philox seeds on the GPU (mixed and FP64) and 8
    on the oracle: the mean indices agree within 4 standard errors of their
    difference, and both show sorting (the index rises above the initial one)."""
    wl = workloads.c2_sorting_3d(n=3000, seed=31, steps=800)
    r = 2.5
    initial = demixing_index(wl.pos, wl.species, r)
    seeds = range(1, 9)

    def ensemble(make):
        out = []
        for seed in seeds:
            sim = make(seed)
            sim.set_agents(wl.pos, wl.vel, wl.species)
            sim.step(800)
            out.append(demixing_index(sim.get_state()[0], wl.species, r))
            sim.close()
        return np.array(out)

    orc = ensemble(lambda seed: OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=2, noise="philox",
                                          seed=seed))
    for precision in ("fp64", "mixed"):
        gpu = ensemble(lambda seed: MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=2,
                                                     noise="philox", seed=1000 + seed, precision=precision))
        se = np.sqrt(gpu.var(ddof=1) / len(gpu) + orc.var(ddof=1) / len(orc))
        assert abs(gpu.mean() - orc.mean()) <= 4.0 * se + 1e-12, (precision, gpu, orc)
        assert gpu.mean() > initial + 3.0 * gpu.std(ddof=1) / np.sqrt(len(gpu)), (precision, initial, gpu)
    assert orc.mean() > initial
