"""The rods example's Cycle + NeighborSensing on the GPU (aspects [Mechanics,
Interaction, Cycle, NeighborSensing], bacterial_rods/src/main.rs:48-111; division
bookkeeping update_cycle.rs:10-69) against the oracle, through the C ABI.

Bars: the agent count, every identifier and parent, bit-exact; spring lengths and
growth rates <= 1e-12 relative; positions and velocities <= 1e-7 relative over 60
steps (FP64; the stiff springs of freshly divided rods amplify the summation-order
rounding of the forces, which stays below 1e-9 over 30 steps)."""
import numpy as np
import pytest

from cellular_raza_amd import (
    CartesianCuboid,
    InstanceSpec,
    MechanicsBackend,
    MorsePotential,
    NoInteraction,
    RodGrowthCycle,
    RodInteraction,
    RodMechanics,
    SimulationError,
    workloads,
)
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    num = np.linalg.norm(a - b, axis=1).max()
    den = np.linalg.norm(b, axis=1).max()
    return num / den if den > 0 else num


def pair(spec, cycle, pos, vel=None, spring=None, dt=0.1, order=2):
    g = MechanicsBackend(spec, dt=dt, mechanics_solver_order=order)
    g.set_rod_cycle(cycle)
    g.set_agents(pos, vel)
    o = OracleSim(spec, dt=dt, mechanics_solver_order=order)
    o.set_agents(pos, vel)
    o.set_rod_cycle(cycle)
    if spring is not None:
        g.set_rod_spring_lengths(spring)
        o.set_spring_lengths(spring)
    return g, o


def assert_same(g, o, tag, tol=1e-7):
    assert g.n_agents == o.n_agents, tag
    # the voxels' lists (the order that numbers the mothers and appends daughters)
    vo, so, order = o.get_voxels()
    assert np.array_equal(g.get_voxels()[1], so), tag
    assert np.array_equal(g.voxel_order(), order.astype(np.uint32)), tag
    sg, so = g.cycle_state(), o.cycle_state()
    assert np.array_equal(sg["identifier"], so["identifier"]), tag
    assert np.array_equal(sg["parent"], so["parent"]), tag
    assert rel_err(sg["spring_length"][:, None], so["spring_length"][:, None]) <= 1e-12, tag
    assert rel_err(sg["growth_rate"][:, None], so["growth_rate"][:, None]) <= 1e-12, tag
    pg, vg = g.get_state()
    po, vo = o.get_state()
    ep, ev = rel_err(pg, po), rel_err(vg, vo)
    assert ep <= tol and (ev <= tol or np.abs(vo).max() == 0.0), (tag, ep, ev)


def test_divide_doctest_on_the_gpu():
    """The divide doctest (bacterial_rods.rs:560-585) as one step of a resting rod."""
    pos = np.zeros((1, 7, 2))
    pos[0, :, 0] = 10.0 + 0.5 * np.arange(7)
    pos[0, :, 1] = 10.0
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [40.0, 40.0], 5.0)
    spec = [InstanceSpec(dom, RodInteraction(NoInteraction()),
                         RodMechanics(n_vertices=7, diffusion_constant=0.0, spring_tension=0.1, rigidity=0.05,
                                      spring_length=0.5, damping=0.0))]
    cyc = RodGrowthCycle(growth_base_rate=0.0, division_threshold=0.4, division_radius=0.25, sensing_radius=0.5,
                         max_agents=8)
    g, o = pair(spec, cyc, pos.reshape(1, -1))
    g.step(1)
    o.step(1)
    assert g.n_agents == 2
    p = g.get_state()[0].reshape(2, 7, 2)
    assert abs(np.linalg.norm(p[0, 6] - p[1, 0]) - 0.5) < 1e-3
    assert_same(g, o, "doctest")
    st = g.cycle_state()
    assert st["identifier"][:, 0].tolist() == [1, 1] and st["identifier"][:, 2].tolist() == [2, 3]
    assert st["parent"].tolist() == [[0, 0, 0], [0, 0, 0]]


@pytest.mark.parametrize("dim", [2, 3])
def test_colony_against_the_oracle(dim):
    """A colony of rods with Morse interactions that divides several times: agent
    count, identifiers and parents bit-exact, state within the FP64 trajectory bar,
    every few steps; spring lengths start per rod (crm_set_rod_spring_lengths)."""
    rng = np.random.default_rng(40 + dim)
    n, nv = 30, 5
    L = 40.0 if dim == 2 else 16.0
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * dim, [L] * dim, 4.5)
    spec = [InstanceSpec(dom, RodInteraction(MorsePotential(radius=0.5, potential_stiffness=2.0, cutoff=1.5,
                                                            strength=0.1)),
                         RodMechanics(n_vertices=nv, diffusion_constant=0.0, spring_tension=10.0, rigidity=1.0,
                                      spring_length=0.5, damping=1.5))]
    pos = workloads.straight_rods(rng, n, dim, L, nv, 0.5).reshape(n, -1)
    spring = rng.uniform(0.45, 0.6, n)
    cyc = RodGrowthCycle(growth_base_rate=0.2, division_threshold=0.62, division_radius=0.5, sensing_radius=1.0,
                         max_agents=512)
    g, o = pair(spec, cyc, pos, spring=spring, dt=0.05)
    assert_same(g, o, "start")
    for k in range(12):
        g.step(5)
        o.step(5)
        assert_same(g, o, f"after {5 * (k + 1)} steps")
    assert g.n_agents >= 1.8 * n  # most rods divided (then crowding stops the growth)
    kinds = g.cycle_state()["identifier"][:, 0]
    assert (kinds == 1).sum() > n


def test_cycle_errors():
    rng = np.random.default_rng(3)
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [30.0, 30.0], 4.0)
    spec = [InstanceSpec(dom, RodInteraction(NoInteraction()),
                         RodMechanics(n_vertices=4, diffusion_constant=0.0, spring_tension=1.0, rigidity=1.0,
                                      spring_length=0.5, damping=1.0))]
    pos = workloads.straight_rods(rng, 4, 2, 30.0, 4, 0.5).reshape(4, -1)
    g = MechanicsBackend(spec, dt=0.1)
    g.set_agents(pos)
    with pytest.raises(SimulationError, match="InvalidArgument"):  # after set_agents
        g.set_rod_cycle(RodGrowthCycle(0.1, 1.0, 0.25, 0.5))
    g.close()
    g = MechanicsBackend(spec, dt=0.1)
    g.set_rod_cycle(RodGrowthCycle(growth_base_rate=1.0, division_threshold=0.4, division_radius=0.1,
                                   sensing_radius=0.5, max_agents=5))
    g.set_agents(pos)
    with pytest.raises(SimulationError, match="InvalidArgument"):  # 8 rods > max_agents 5
        g.step(1)
    g.close()
    wl = workloads.c2_sorting_3d(n=100, seed=3)
    p = MechanicsBackend(wl.instances, dt=wl.dt)
    with pytest.raises(SimulationError, match="InvalidArgument"):  # points
        p.set_rod_cycle(RodGrowthCycle(0.1, 1.0, 0.25, 0.5))
    p.close()


def test_bacterial_rods_example_end_to_end(tmp_path):
    """The rods example (main.rs:113-209) for its first 6.7 h (4000 steps) through
    run_simulation with JSON storage, read back the way its plotter reads the cells
    (plotter.py:22-42); every 20th save point's rod count and identifiers (exact)
    against the oracle's run, positions within 1e-6 of the domain length through step
    3000.  The crowded colony is chaotic: rounding differences of 1e-13 m at step 3500
    reach 1e-9 m by step 4000, so later positions are not comparable while the
    divisions (counts, identifiers) still are."""
    import glob
    import json
    import os

    from cellular_raza_amd import FixedStepsize, Settings, StorageBuilder, run_simulation
    from cellular_raza_amd import storage as st

    ex = workloads.bacterial_rods_example()
    wl = ex.workload
    spec = wl.instances[0]
    steps = 4000
    time = FixedStepsize.from_partial_save_steps(0.0, wl.dt, steps, 25)
    builder = StorageBuilder().location(tmp_path).add_date(False)
    access = run_simulation(spec.domain, spec.interaction, spec.mechanics, wl.pos.reshape(5, -1), Settings(
        time=time, storage=builder), velocities=wl.vel.reshape(5, -1), mechanics_solver_order=2, cycle=ex.cycle,
        spring_lengths=ex.spring_length, cell_extra={"growth_rate": st.GROWTH_RATE,
                                                     "spring_length_threshold": 6e-6})
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=2)
    o.set_agents(wl.pos.reshape(5, -1), wl.vel.reshape(5, -1))
    o.set_rod_cycle(ex.cycle)
    o.set_spring_lengths(ex.spring_length)
    done = 0
    for it in range(0, steps + 1, 500):
        if it > done:
            o.step(it - done)
            done = it
        files = glob.glob(os.path.join(tmp_path, "cells", "json", f"{it:020d}", "*"))
        cells = []
        for f in files:
            with open(f) as fh:
                cells.extend(d["element"][0] for d in json.load(fh)["data"])
        assert len(cells) == o.n_agents, it
        st_o = o.cycle_state()
        want = {st.identifier_json(r).__repr__(): a for a, r in enumerate(st_o["identifier"])}
        po = o.get_state()[0].reshape(-1, 5, 3)
        for c in cells:
            a = want[c["identifier"].__repr__()]
            assert c["cell"]["spring_length_threshold"] == 6e-6
            if it > 3000:
                continue
            p = np.array(c["cell"]["mechanics"]["pos"][0], dtype=float).reshape(3, -1).T
            # 1e-6 of the domain's length: FP64 summation-order rounding accumulated
            # over thousands of steps of a crowded colony
            assert np.allclose(p, po[a], rtol=0, atol=1e-6 * 200e-6), (it, a, np.abs(p - po[a]).max())
            assert c["cell"]["growth_rate"] == st_o["growth_rate"][a]
    assert o.n_agents > 40  # the colony grew (5 -> 44 rods in 6.7 h)
    assert sorted(access.cells.get_all_iterations()) == list(range(0, steps + 1, 25))
