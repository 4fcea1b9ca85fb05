"""CycleEvent::Remove on the GPU (crm_remove_agents; chili/update_cycle.rs:41-55)
against the oracle, through the C ABI: removals between steps, then more steps.

Bars: agent counts and voxel assignment (by agent index, and the voxel start
table) bit-exact; FP64 states within the trajectory bars of the plain path (one
step 1e-12, tens of steps 1e-9 relative); identifiers and parents of the rods
cycle bit-exact."""
import numpy as np
import pytest

from cellular_raza_amd import (
    Brownian,
    CartesianCuboid,
    InstanceSpec,
    Langevin,
    MechanicsBackend,
    MorsePotential,
    RodGrowthCycle,
    RodInteraction,
    RodMechanics,
    SimulationError,
    workloads,
)
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

TRAJ_TOL = 1e-9


def rel_err(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    num = np.linalg.norm(a - b, axis=1).max()
    den = np.linalg.norm(b, axis=1).max()
    return num / den if den > 0 else num


def pair(wl, **kw):
    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, precision="fp64", **kw)
    okw = {k: v for k, v in kw.items() if k in ("noise", "seed")}
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, **okw)
    g.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    o.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    return g, o


def assert_same(g, o, tag, tol=TRAJ_TOL):
    assert g.n_agents == o.n_agents, tag
    vg, sg = g.get_voxels()
    vo, so, order = o.get_voxels()
    assert np.array_equal(vg, vo), tag
    assert np.array_equal(sg, so), tag
    assert np.array_equal(g.voxel_order(), order.astype(np.uint32)), tag
    pg, vlg = g.get_state()
    po, vlo = o.get_state()
    assert rel_err(pg, po) <= tol, (tag, rel_err(pg, po))
    assert rel_err(vlg, vlo) <= tol * 100 or np.abs(vlo).max() == 0.0, (tag, rel_err(vlg, vlo))


def remove_both(g, o, rm):
    g.remove_agents(rm)
    o.remove_agents(rm)


@pytest.mark.parametrize("order", [0, 1, 2])
def test_remove_points_newton_damped(order):
    """3-D sorting (NewtonDamped, keys binned ahead by the gather), every solver order:
    three removals of random subsets, the state compared right after each removal
    (the survivors' rows, bitwise) and after the following steps."""
    wl = workloads.c2_sorting_3d(n=5000, seed=31, steps=10)
    wl.solver_order = order
    g, o = pair(wl)
    rng = np.random.default_rng(5)
    for k in range(3):
        g.step(4)
        o.step(4)
        p_before = g.get_state()[0]
        rm = rng.random(g.n_agents) < 0.15
        remove_both(g, o, rm)
        assert np.array_equal(g.get_state()[0], p_before[~rm]), k
        assert_same(g, o, f"after removal {k}")
    g.step(6)
    o.step(6)
    assert_same(g, o, "end")


def test_remove_points_mixed_precision_records():
    """Mixed precision: after a removal the fixed-point force records are rebuilt for
    the compacted order; forces against the oracle set to the GPU's surviving state
    (same species) within the mixed bar, 1e-5 relative."""
    wl = workloads.c2_sorting_3d(n=6000, seed=38, steps=10)
    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, precision="mixed")
    g.set_agents(wl.pos, wl.vel, wl.species)
    g.step(5)
    rm = np.random.default_rng(9).random(g.n_agents) < 0.3
    g.remove_agents(rm)
    pos, vel = g.get_state()
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order)
    o.set_agents(pos, vel, wl.species[~rm])
    fo = o.compute_forces()
    assert np.abs(fo).max() > 0
    assert rel_err(g.compute_forces(), fo) <= 1e-5
    g.step(3)


def test_remove_points_langevin_philox_and_graphs():
    """Langevin with Philox noise (streams keyed by the compacted agent index), steps
    replayed as captured graphs before and after the removal."""
    wl = workloads.c2_sorting_3d(n=4000, seed=32, steps=10)
    wl.instances = [InstanceSpec(wl.instances[0].domain, wl.instances[0].interaction,
                                 Langevin(damping=1.0, mass=1.0, kb_temperature=0.1))]
    g, o = pair(wl, noise="philox", seed=9)
    g.step(6)
    o.step(6)
    rm = np.zeros(g.n_agents, dtype=bool)
    rm[::7] = True
    remove_both(g, o, rm)
    g.step(8)
    o.step(8)
    assert_same(g, o, "langevin")


def test_remove_points_brownian():
    """Brownian agents (their keys come from a key kernel after the forces)."""
    wl = workloads.c2_sorting_3d(n=3000, seed=33, steps=10)
    wl.instances = [InstanceSpec(wl.instances[0].domain, wl.instances[0].interaction,
                                 Brownian(diffusion_constant=0.05, kb_temperature=1.0))]
    wl.solver_order = 1
    g, o = pair(wl, noise="philox", seed=4)
    g.step(5)
    o.step(5)
    rm = np.random.default_rng(6).random(g.n_agents) < 0.4
    remove_both(g, o, rm)
    g.step(5)
    o.step(5)
    assert_same(g, o, "brownian")


def test_remove_in_an_ensemble():
    """Several instances: per-instance agent counts shrink by the agents removed
    from each (one instance loses all but one agent)."""
    wl = workloads.c4_ensemble_2d(n_instances=6, seed=34, steps=10, n_min=40, n_max=400)
    g, o = pair(wl)
    g.step(3)
    o.step(3)
    rm = np.random.default_rng(7).random(g.n_agents) < 0.25
    first = int(wl.counts[0])
    vox = g.get_voxels()[0]
    # instance 0's agents: those whose voxel lies in instance 0's (first) voxel range
    nv0 = int(np.prod(wl.instances[0].domain.get_n_voxels()))
    in_first = vox < nv0
    assert in_first.sum() == first
    rm[in_first] = True
    rm[np.flatnonzero(in_first)[0]] = False
    remove_both(g, o, rm)
    g.step(5)
    o.step(5)
    assert_same(g, o, "ensemble")


def test_remove_rods_plain():
    """2-D rods without the cycle, AB3 history carried by the survivors."""
    wl = workloads.c3_rods_2d(n=1500, seed=35, steps=10)
    g, o = pair(wl)
    g.step(5)
    o.step(5)
    rm = np.random.default_rng(8).random(g.n_agents) < 0.3
    remove_both(g, o, rm)
    g.step(10)
    o.step(10)
    assert_same(g, o, "rods")


def test_remove_rods_with_the_cycle():
    """A dividing colony: rods removed twice while it grows; agent counts,
    identifiers, parents (bit-exact) and states after every phase.  Daughters after
    a removal are numbered from the compacted count."""
    rng = np.random.default_rng(36)
    n, nv, L = 30, 5, 40.0
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [L, L], 4.5)
    spec = [InstanceSpec(dom, RodInteraction(MorsePotential(radius=0.5, potential_stiffness=2.0, cutoff=1.5,
                                                            strength=0.1)),
                         RodMechanics(n_vertices=nv, diffusion_constant=0.0, spring_tension=10.0, rigidity=1.0,
                                      spring_length=0.5, damping=1.5))]
    pos = workloads.straight_rods(rng, n, 2, L, nv, 0.5).reshape(n, -1)
    spring = rng.uniform(0.45, 0.6, n)
    cyc = RodGrowthCycle(growth_base_rate=0.2, division_threshold=0.62, division_radius=0.5, sensing_radius=1.0,
                         max_agents=512)
    g = MechanicsBackend(spec, dt=0.05, mechanics_solver_order=2)
    g.set_rod_cycle(cyc)
    g.set_agents(pos)
    g.set_rod_spring_lengths(spring)
    o = OracleSim(spec, dt=0.05, mechanics_solver_order=2)
    o.set_agents(pos)
    o.set_rod_cycle(cyc)
    o.set_spring_lengths(spring)
    for phase in range(3):
        g.step(10)
        o.step(10)
        assert g.n_agents == o.n_agents
        if phase < 2:
            rm = rng.random(g.n_agents) < 0.2
            remove_both(g, o, rm)
        sg, so = g.cycle_state(), o.cycle_state()
        assert np.array_equal(sg["identifier"], so["identifier"]), phase
        assert np.array_equal(sg["parent"], so["parent"]), phase
        assert rel_err(sg["spring_length"][:, None], so["spring_length"][:, None]) <= 1e-12
        pg, po = g.get_state()[0], o.get_state()[0]
        assert rel_err(pg, po) <= 1e-7, (phase, rel_err(pg, po))
    assert (g.cycle_state()["identifier"][:, 0] == 1).any()  # divisions happened


def test_remove_errors():
    wl = workloads.c2_sorting_3d(n=200, seed=37, steps=10)
    g = MechanicsBackend(wl.instances, dt=wl.dt)
    with pytest.raises(SimulationError, match="InvalidArgument"):  # before set_agents
        g.remove_agents(np.zeros(0, dtype=bool))
    g.set_agents(wl.pos, wl.vel, wl.species)
    with pytest.raises(ValueError):
        g.remove_agents(np.zeros(10, dtype=bool))
    with pytest.raises(SimulationError, match="InvalidArgument"):
        g.remove_agents(np.ones(200, dtype=bool))
    g.remove_agents(np.zeros(200, dtype=bool))  # nothing to do
    assert g.n_agents == 200
    g.step(2)
    g.close()


def test_run_simulation_with_removals():
    """run_simulation(remove=...): the first agent is removed after every third step;
    the saved positions against the oracle doing the same between single steps."""
    from cellular_raza_amd import FixedStepsize, Settings, run_simulation

    wl = workloads.c1_morse_2d(n=400, seed=39, steps=10)
    spec = wl.instances[0]
    time = FixedStepsize.from_partial_save_steps(0.0, wl.dt, 12, 4)

    def rm(be, it):
        return (np.arange(be.n_agents) == 0) if it % 3 == 0 else None

    out = run_simulation(spec.domain, spec.interaction, spec.mechanics, wl.pos, Settings(time=time),
                         velocities=wl.vel, mechanics_solver_order=wl.solver_order, remove=rm, precision="fp64")
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order)
    o.set_agents(wl.pos, wl.vel)
    assert np.array_equal(out[0], wl.pos)
    for it in range(1, 13):
        o.step(1)
        if it % 3 == 0:
            o.remove_agents(np.arange(o.n_agents) == 0)
        if it % 4 == 0:
            assert out[it].shape[0] == o.n_agents == 400 - it // 3
            assert rel_err(out[it], o.get_state()[0]) <= TRAJ_TOL, it


def test_remove_at_full_size_invariants():
    """1 M agents (the 3-D sorting workload, mixed precision): 10 % removed between
    steps.  The survivors' rows are unchanged by the removal; after more steps the
    voxel lists are a permutation of the agents, each list holds the agents whose
    positions bin to that voxel by trunc((p - min) / dx), and the starts add up."""
    wl = workloads.c2_sorting_3d(n=1_000_000, seed=40, steps=10)
    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, precision="mixed",
                         noise="philox", seed=2)
    g.set_agents(wl.pos, wl.vel, wl.species)
    g.step(4)
    rm = np.random.default_rng(41).random(g.n_agents) < 0.1
    p_before = g.get_state()[0]
    g.remove_agents(rm)
    assert g.n_agents == int((~rm).sum())
    assert np.array_equal(g.get_state()[0], p_before[~rm])
    g.step(4)
    vox, start = g.get_voxels()
    order = g.voxel_order()
    assert start[0] == 0 and start[-1] == g.n_agents and np.all(np.diff(start.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(order), np.arange(g.n_agents, dtype=np.uint32))
    counts = np.diff(start.astype(np.int64))
    assert np.array_equal(vox[order], np.repeat(np.arange(len(counts)), counts))
    dom = wl.instances[0].domain
    n, dx, pos = dom.get_n_voxels(), dom.get_dx(), g.get_state()[0]
    plain = np.zeros(g.n_agents, dtype=np.int64)
    for k in range(3):
        plain = plain * n[k] + np.trunc((pos[:, k] - dom.min[k]) / dx[k]).astype(np.int64)
    assert np.array_equal(vox.astype(np.int64), plain)
    g.close()


def test_remove_every_agent_of_one_instance():
    """An instance of an ensemble emptied by a removal keeps stepping with no agents
    while the others go on (work blocks and voxel lists without it)."""
    wl = workloads.c4_ensemble_2d(n_instances=4, seed=42, steps=10, n_min=60, n_max=300)
    g, o = pair(wl)
    g.step(2)
    o.step(2)
    vox = g.get_voxels()[0]
    nv0 = int(np.prod(wl.instances[0].domain.get_n_voxels()))
    nv1 = int(np.prod(wl.instances[1].domain.get_n_voxels()))
    rm = (vox >= nv0) & (vox < nv0 + nv1)  # every agent of instance 1
    assert rm.sum() == int(wl.counts[1])
    remove_both(g, o, rm)
    g.step(6)
    o.step(6)
    assert_same(g, o, "empty instance")
