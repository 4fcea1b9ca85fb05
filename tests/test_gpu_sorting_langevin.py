"""GPU parity of the Langevin cell-sorting interaction
(cellular_raza-examples/cell_sorting/src/main-langevin.rs:44-92) against the oracle,
on the example's own configuration (1600 cells, 200^3, 5^3 voxels, Langevin3D with
Philox noise shared by both sides).

Tolerances: FP64 forces and one step <= 1e-12 relative; trajectories <= 1e-9 over
30 steps (the example is chaotic: the oracle against itself with a 1e-15 relative
input perturbation differs by 8e-12 after 50 steps and by 2e-6 after 100); mixed
precision forces <= 1e-5.
"""
import numpy as np
import pytest

from cellular_raza_amd import CellSpecificInteractionLangevin, InstanceSpec, MechanicsBackend, workloads
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
TRAJ_TOL = 1e-9
MIXED_TOL = 1e-5


def rel_err(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return np.linalg.norm(a - b, axis=1).max() / np.linalg.norm(b, axis=1).max()


def example(asymmetric=False):
    wl = workloads.langevin_sorting_example(steps=30)
    if asymmetric:  # the non-symmetric path evaluates both sides of each full-weight pair
        spec = wl.instances[0]
        pot = CellSpecificInteractionLangevin(cell_radius=6.0, attraction_rr=0, attraction_bb=0, attraction_rb=0,
                                              repulsion=5.0, cutoff=20.0,
                                              attraction_matrix=[[1.5, 0.015], [0.4, 0.075]])
        wl.instances = [InstanceSpec(spec.domain, pot, spec.mechanics)]
    return wl


def pair(wl, **kw):
    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, **kw)
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order,
                  **{k: v for k, v in kw.items() if k in ("weighting", "noise", "seed")})
    g.set_agents(wl.pos, wl.vel, wl.species)
    o.set_agents(wl.pos, wl.vel, wl.species)
    return g, o


@pytest.mark.parametrize("asymmetric", [False, True])
@pytest.mark.parametrize("weighting", ["reference", "physical"])
def test_forces_match_oracle(asymmetric, weighting):
    wl = example(asymmetric)
    g, o = pair(wl, weighting=weighting)
    fg, fo = g.compute_forces(), o.compute_forces()
    assert np.abs(fo).max() > 0
    assert rel_err(fg, fo) <= FORCE_TOL, rel_err(fg, fo)
    gm = MechanicsBackend(wl.instances, dt=wl.dt, precision="mixed", weighting=weighting)
    gm.set_agents(wl.pos, wl.vel, wl.species)
    assert rel_err(gm.compute_forces(), fo) <= MIXED_TOL, rel_err(gm.compute_forces(), fo)
    for h in (g, o, gm):
        h.close()


def test_trajectory_matches_oracle():
    wl = example()
    g, o = pair(wl, noise="philox", seed=3)
    g.step(1)
    o.step(1)
    assert rel_err(g.get_state()[0], o.get_state()[0]) <= FORCE_TOL
    g.step(29)
    o.step(29)
    pg, vg = g.get_state()
    po, vo = o.get_state()
    assert rel_err(pg, po) <= TRAJ_TOL, rel_err(pg, po)
    assert rel_err(vg, vo) <= TRAJ_TOL * 100, rel_err(vg, vo)
    assert np.array_equal(g.voxel_order(), o.get_voxels()[2].astype(np.uint32))
    g.close()
    o.close()
