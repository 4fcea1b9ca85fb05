"""crm_remove_agents_noalloc (static device scratch, no hipMalloc) against
crm_remove_agents on a twin context, through the C ABI: two removals in a row,
the second on fewer agents, each followed by steps; state, ring history, voxel
lists, identifiers and per-step results must be bitwise equal (integer and index
work is exact, and both run the same kernels on the same data)."""
import ctypes as C

import numpy as np
import pytest

from cellular_raza_amd import MechanicsBackend, SimulationError, _abi, workloads

pytestmark = pytest.mark.gpu


def twins(wl, **kw):
    out = []
    for _ in range(2):
        g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, **kw)
        g.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
        out.append(g)
    return out


def remove(g, mask, noalloc):
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    fn = g._lib.crm_remove_agents_noalloc if noalloc else g._lib.crm_remove_agents
    g._check(fn(g._h, m.ctypes.data_as(_abi.U8P)))
    g.n_agents -= int(np.count_nonzero(m))


def assert_same(a, b):
    assert a.n_agents == b.n_agents
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y)
    ha, hb = a.get_history(), b.get_history()
    assert ha[2] == hb[2]
    for i in range(ha[2]):
        assert np.array_equal(ha[0][i], hb[0][i]) and np.array_equal(ha[1][i], hb[1][i])
    for x, y in zip(a.get_voxels(), b.get_voxels()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.voxel_order(), b.voxel_order())
    ia, ib = a.identifiers(), b.identifiers()
    for k in ia:
        assert np.array_equal(ia[k], ib[k])


@pytest.mark.parametrize("case", ["sorting3d_mixed", "ensemble2d_fp64"])
def test_noalloc_removal_matches_allocating_removal(case):
    if case == "sorting3d_mixed":
        wl = workloads.c2_sorting_3d(n=20000, seed=51, steps=10)
        ref, na = twins(wl, precision="mixed", noise="philox", seed=4)
    else:
        wl = workloads.c4_ensemble_2d(n_instances=8, seed=52, steps=10, n_min=200, n_max=3000)
        ref, na = twins(wl)
    rng = np.random.default_rng(53)
    for frac in (0.2, 0.05):  # the second call sees fewer agents
        for g in (ref, na):
            g.step(3)
        mask = rng.random(ref.n_agents) < frac
        remove(ref, mask, noalloc=False)
        remove(na, mask, noalloc=True)
        assert_same(ref, na)
        for g in (ref, na):
            g.step(2)
        assert_same(ref, na)
    ref.close()
    na.close()


def test_noalloc_removal_rejects_too_many_agents():
    from cellular_raza_amd import CartesianCuboid, InstanceSpec, NewtonDamped, NoInteraction

    n = (1 << 22) + 1
    dom = CartesianCuboid.from_boundaries_and_n_voxels([0.0, 0.0], [100.0, 100.0], [1024, 1024])
    g = MechanicsBackend([InstanceSpec(dom, NoInteraction(), NewtonDamped(1.0, 1.0))], dt=0.01)
    pos = np.random.default_rng(1).uniform(1e-3, 99.9, (n, 2))
    g.set_agents(pos)
    mask = np.zeros(n, dtype=bool)
    mask[::7] = True
    with pytest.raises(SimulationError, match="at most 4194304 agents"):
        remove(g, mask, noalloc=True)
    assert g.n_agents == n
    assert np.array_equal(g.get_state()[0], pos)
    g.close()
