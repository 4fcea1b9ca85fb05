"""Parity of the gfx950 library (through the C ABI) against the CPU oracle.

Tolerances (FP64 mode), from the specification of this backend:
  per-step forces and positions   <= 1e-12 relative (max over agents of |dF| / max |F|)
  100-step trajectories           <= 1e-9 relative
  voxel assignment, neighbour sets bit-exact
Mixed precision: <= 1e-5 relative force error against the FP64 oracle.
"""
import math

import numpy as np
import pytest

from cellular_raza_amd import (
    BoundLennardJones,
    Brownian,
    CartesianCuboid,
    CellSpecificInteraction,
    InstanceSpec,
    Langevin,
    MechanicsBackend,
    MiePotential,
    MorsePotential,
    NewtonDamped,
    NoInteraction,
    SimulationError,
    SpeciesMorsePotential,
    workloads,
)
from oracle.oracle import OracleError, OracleSim

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
TRAJ_TOL = 1e-9
MIXED_TOL = 1e-5


def rel_err(a, b):
    """max_i |a_i - b_i| / max_i |b_i| over agent vectors (norm-wise)."""
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    num = np.linalg.norm(a - b, axis=1).max()
    den = np.linalg.norm(b, axis=1).max()
    return num / den if den > 0 else num


def pair(wl, **kw):
    opts = dict(mechanics_solver_order=wl.solver_order)
    g_kw = dict(opts)
    g_kw.update({k: v for k, v in kw.items()})
    o_kw = dict(opts)
    o_kw.update({k: v for k, v in kw.items() if k in ("weighting", "history_order", "noise", "seed")})
    g = MechanicsBackend(wl.instances, dt=wl.dt, **g_kw)
    o = OracleSim(wl.instances, dt=wl.dt, **o_kw)
    g.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    o.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    return g, o


def voxel_reference(domain, pos):
    """Voxel plain index by the reference formula trunc((p - min) / dx), computed in numpy."""
    n = domain.get_n_voxels()
    dx = domain.get_dx()
    idx = [np.trunc((pos[:, k] - domain.min[k]) / dx[k]).astype(np.int64) for k in range(domain.dim)]
    plain = np.zeros(len(pos), dtype=np.int64)
    for k in range(domain.dim):
        plain = plain * n[k] + idx[k]
    return plain


def small_workloads():
    rng = np.random.default_rng(11)
    out = {}
    out["morse2d"] = workloads.c1_morse_2d(n=1500, seed=3, steps=50)
    out["species3d"] = workloads.c2_sorting_3d(n=4000, seed=4, steps=50)
    out["mie3d"] = workloads.c5_mie_3d(n=3000, seed=5, steps=50)
    # bounded Lennard-Jones on a jittered lattice
    w = workloads.c5_mie_3d(n=2000, seed=6, steps=20)
    w.instances = [InstanceSpec(w.instances[0].domain, BoundLennardJones(epsilon=0.5, sigma=1.6, bound=5.0,
                                                                         cutoff=2.5), w.instances[0].mechanics)]
    out["blj3d"] = w
    # cell-sorting interaction of the benchmark (sign +1) and of the example (sign -1)
    for variant in ("benchmark", "example"):
        n = 3000
        L = 60.0
        dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [L] * 3, 1.5 * 6.0 * 2.0 / 4)
        pot = CellSpecificInteraction(cell_radius=1.5, potential_strength=2.0, relative_interaction_range=1.5,
                                      variant=variant)
        pos = rng.uniform(1e-9, L, size=(n, 3))
        sp = (rng.random(n) < 0.5).astype(np.int32)
        out["sorting_" + variant] = workloads.Workload(
            "cell sorting " + variant, [InstanceSpec(dom, pot, NewtonDamped(damping_constant=2.0, mass=1.0))],
            0.01, 20, 2, pos, np.zeros_like(pos), sp, [n])
    return out


WL = small_workloads()


@pytest.mark.parametrize("name", sorted(WL))
@pytest.mark.parametrize("weighting", ["reference", "physical"])
def test_forces_match_oracle(name, weighting):
    wl = WL[name]
    g, o = pair(wl, weighting=weighting)
    fg, fo = g.compute_forces(), o.compute_forces()
    assert np.abs(fo).max() > 0
    assert rel_err(fg, fo) <= FORCE_TOL, (name, rel_err(fg, fo))


@pytest.mark.parametrize("name", sorted(WL))
@pytest.mark.parametrize("weighting", ["reference", "physical"])
def test_mixed_forces_match_oracle(name, weighting):
    """The float pair path of every potential (the Morse, Mie and cell-sorting paths
    take 1/r from v_rsq_f32) within the mixed-precision bound of the FP64 oracle."""
    wl = WL[name]
    g, o = pair(wl, weighting=weighting, precision="mixed")
    fg, fo = g.compute_forces(), o.compute_forces()
    assert np.abs(fo).max() > 0
    assert rel_err(fg, fo) <= MIXED_TOL, (name, rel_err(fg, fo))
    g.close()


@pytest.mark.parametrize("name", ["species3d", "mie3d"])
def test_step_and_trajectory(name):
    wl = WL[name]
    noise = "philox" if wl.noise == "philox" else "none"
    g, o = pair(wl, noise=noise, seed=7)
    g.step(1)
    o.step(1)
    pg, vg = g.get_state()
    po, vo = o.get_state()
    assert rel_err(pg, po) <= FORCE_TOL
    assert rel_err(vg, vo) <= FORCE_TOL * 10
    g.step(99)
    o.step(99)
    pg, vg = g.get_state()
    po, vo = o.get_state()
    assert rel_err(pg, po) <= TRAJ_TOL, rel_err(pg, po)
    assert rel_err(vg, vo) <= TRAJ_TOL * 100, rel_err(vg, vo)


def test_c1_per_step_parity_along_trajectory():
    """The 2-D Morse configuration (area fraction 1.8 at R = 6) is chaotic: the oracle
    against itself with a 1e-15 input perturbation already differs by ~1e-3 after 20
    steps, so a 100-step trajectory comparison is ill-posed.  Euler keeps no history,
    so every step along the GPU trajectory is checked by restarting the oracle from the
    GPU state.  The initial overlaps throw agents against the walls hard enough that a
    reflection can leave the box (BoundaryError); the oracle must then fail at the
    same step."""
    wl = workloads.c1_morse_2d(n=1000, seed=1, steps=100)
    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=0)
    g.set_agents(wl.pos, wl.vel)
    for k in range(100):
        p0, v0 = g.get_state()
        o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=0)
        o.set_agents(p0, v0)
        try:
            g.step(1)
        except SimulationError as e:
            with pytest.raises(OracleError) as eo:
                o.step(1)
            assert e.kind == eo.value.kind == "BoundaryError", (k, e, eo.value)
            return
        o.step(1)
        pg, vg = g.get_state()
        po, vo = o.get_state()
        assert rel_err(pg, po) <= FORCE_TOL, (k, rel_err(pg, po))
        assert rel_err(vg, vo) <= FORCE_TOL, (k, rel_err(vg, vo))


def test_noise_buffer_matches_oracle():
    wl = workloads.c2_sorting_3d(n=3000, seed=8, steps=10)
    g, o = pair(wl, noise="buffer")
    rng = np.random.default_rng(1)
    nz = rng.standard_normal((10, wl.n_agents, 3))
    g.step(10, nz)
    o.step(10, nz)
    assert rel_err(g.get_state()[0], o.get_state()[0]) <= TRAJ_TOL


@pytest.mark.parametrize("order,hist", [(0, "oldest_first"), (1, "oldest_first"), (2, "oldest_first"),
                                        (2, "newest_first")])
def test_solver_orders_and_history(order, hist):
    wl = workloads.c5_mie_3d(n=1500, seed=9, steps=20)
    wl.solver_order = order
    g, o = pair(wl, history_order=hist)
    for k in range(4):
        dg, vg_, ng = g.get_history()
        do, vo_, no = o.get_history()
        assert ng == no == min(k, order)
        for s in range(ng):
            assert rel_err(dg[s], do[s]) <= FORCE_TOL and rel_err(vg_[s], vo_[s]) <= FORCE_TOL
        g.step(1)
        o.step(1)
        assert rel_err(g.get_state()[0], o.get_state()[0]) <= FORCE_TOL


def test_voxels_and_neighbour_sets_bit_exact():
    wl = workloads.c2_sorting_3d(n=5000, seed=12, steps=10)
    g, o = pair(wl, noise="philox", seed=3)
    dom = wl.instances[0].domain
    for _ in range(3):
        vg, sg = g.get_voxels()
        vo, so, order = o.get_voxels()
        pg = g.get_state()[0]
        # binning of the GPU positions by the reference formula, computed independently
        assert np.array_equal(vg.astype(np.int64), voxel_reference(dom, pg))
        # same agent lists per voxel as the oracle, in the reference's order: the
        # agents that stayed keep their order, the arrivals follow by old voxel
        assert np.array_equal(g.voxel_order(), order.astype(np.uint32))
        assert np.array_equal(vg, vo)
        assert np.array_equal(sg, so)
        assert sg[-1] == wl.n_agents and np.all(np.diff(sg.astype(np.int64)) >= 0)
        g.step(3)
        o.step(3)


@pytest.mark.parametrize("mech", ["brownian", "langevin"])
@pytest.mark.parametrize("graph", [True, False])
def test_crowded_arrivals_keep_previous_slot_order(mech, graph):
    """Many agents per voxel and large moves: every step brings several agents into
    the same voxel, which the gather must put after the voxel's stayers in order of
    old voxel, then old position (sort_cells_in_voxels_step_2,
    update_mechanics.rs:470-513).  Voxel lists bit-exact with the oracle every step,
    Brownian (binned after the forces) and Langevin (binned ahead)."""
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [20.0, 20.0], 2.0)
    pot = MorsePotential(1.0, 0.2, 2.0, 0.05)
    m = Brownian(2.0, 1.0) if mech == "brownian" else Langevin(1.0, 0.5, 20.0)
    spec = [InstanceSpec(dom, pot, m)]
    rng = np.random.default_rng(61)
    pos = rng.uniform(0.5, 19.5, size=(1500, 2))
    vel = rng.normal(0.0, 4.0, size=(1500, 2))
    g = MechanicsBackend(spec, dt=0.05, mechanics_solver_order=2, noise="philox", seed=9, graph=graph)
    o = OracleSim(spec, dt=0.05, mechanics_solver_order=2, noise="philox", seed=9)
    g.set_agents(pos, vel)
    o.set_agents(pos, vel)
    most = 0
    for _ in range(6):
        before = g.get_voxels()[0]
        g.step(2)
        o.step(2)
        after = g.get_voxels()[0]
        moved = after[before != after]
        if moved.size:
            most = max(most, int(np.bincount(moved).max()))
        assert np.array_equal(g.voxel_order(), o.get_voxels()[2].astype(np.uint32))
        pg, po = g.get_state()[0], o.get_state()[0]
        assert rel_err(pg, po) <= TRAJ_TOL
    assert most >= 3  # voxels received several agents in the same two steps
    g.close()


def test_mixed_precision_force_error():
    wl = workloads.c2_sorting_3d(n=20000, seed=13, steps=10)
    g, o = pair(wl, precision="mixed")
    err = rel_err(g.compute_forces(), o.compute_forces())
    assert err <= MIXED_TOL, err
    wl5 = workloads.c5_mie_3d(n=8000, seed=14, steps=10)
    g5, o5 = pair(wl5, precision="mixed")
    err5 = rel_err(g5.compute_forces(), o5.compute_forces())
    assert err5 <= MIXED_TOL, err5


@pytest.mark.parametrize("potential", ["morse", "mie", "blj", "sorting"])
def test_mixed_precision_cutoff_band_and_close_pairs(potential):
    """The forces jump at the cutoff, so pairs at the cutoff radius +- 1e-9 must get
    the FP64 decision in the mixed path too; pairs 1e-4 apart must keep their
    direction (fixed-point voxel-relative coordinates)."""
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [30.0] * 3, 2.5)
    pots = {
        "morse": (MorsePotential(1.0, 1.0, 2.5, 0.5), 2.5),
        "mie": (MiePotential(1.0, 1.0, 1e3, 2.5, 12.0, 6.0), 2.5),
        "blj": (BoundLennardJones(0.5, 1.6, 5.0, 2.5), 2.5),
        "sorting": (CellSpecificInteraction(0.8, 1.0, 1.5), 1.5 * 1.6),
    }
    pot, rc = pots[potential]
    rng = np.random.default_rng(21)
    pos = []
    for k in range(60):
        c = rng.uniform(3, 27, 3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        eps = (-1e-9, 0.0, 1e-9, 1e-4)[k % 4] if k % 5 else 0.0
        pos += [c, c + u * (rc + eps)]
    for k in range(20):
        c = rng.uniform(3, 27, 3)
        u = rng.normal(size=3)
        pos += [c, c + u / np.linalg.norm(u) * 0.9]
    pos = np.array(pos)
    spec = [InstanceSpec(dom, pot, NewtonDamped(1.0, 1.0))]
    sp = np.zeros(len(pos), dtype=np.int32)
    g = MechanicsBackend(spec, dt=0.01, precision="mixed")
    o = OracleSim(spec, dt=0.01)
    g.set_agents(pos, species=sp)
    o.set_agents(pos, species=sp)
    fg, fo = g.compute_forces(), o.compute_forces()
    # the same pairs are inside the cutoff
    assert np.array_equal(np.linalg.norm(fg, axis=1) > 0, np.linalg.norm(fo, axis=1) > 0)
    assert rel_err(fg, fo) <= MIXED_TOL, rel_err(fg, fo)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lds_tiles_equal_global_path_and_crowded_tiles(precision):
    """The force kernel stages each tile's neighbourhood in LDS; a tile whose
    neighbourhood overflows LDS (a crowded cluster in a sparse domain) reads from
    global memory.  Both paths do the same arithmetic in the same order."""
    rng = np.random.default_rng(31)
    L = 200.0
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [L] * 3, 2.5)
    sparse = rng.uniform(1e-6, L, (20000, 3))
    cluster = rng.normal(100.0, 1.5, (6000, 3))
    pos = np.concatenate([sparse, cluster])
    sp = (rng.random(len(pos)) < 0.5).astype(np.int32)
    pot = SpeciesMorsePotential(1.0, 1.0, 2.5, 0.5, [[1.0, 0.2], [0.2, 1.0]])
    spec = [InstanceSpec(dom, pot, Langevin(1.0, 2.0, 0.01))]
    F = []
    for lds in (True, False):
        g = MechanicsBackend(spec, dt=0.005, precision=precision, lds=lds)
        g.set_agents(pos, species=sp)
        F.append(g.compute_forces())
        g.close()
    assert np.array_equal(F[0], F[1])
    o = OracleSim(spec, dt=0.005)
    o.set_agents(pos, species=sp)
    tol = FORCE_TOL if precision == "fp64" else MIXED_TOL
    assert rel_err(F[0], o.compute_forces()) <= tol


def test_ensemble_matches_oracle():
    wl = workloads.c4_ensemble_2d(n_instances=12, seed=15, steps=20, n_min=50, n_max=800)
    g, o = pair(wl)
    assert rel_err(g.compute_forces(), o.compute_forces()) <= FORCE_TOL
    g.step(20)
    o.step(20)
    assert rel_err(g.get_state()[0], o.get_state()[0]) <= TRAJ_TOL


def test_graph_and_eager_bitwise_equal_and_deterministic():
    wl = workloads.c2_sorting_3d(n=6000, seed=16, steps=10)
    runs = []
    for graph in (True, False, True):
        g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=2, noise="philox", seed=5,
                             precision="mixed", graph=graph)
        g.set_agents(wl.pos, wl.vel, wl.species)
        g.step(7)
        g.step(37)  # two long replays, two two-step replays and a single step
        runs.append(g.get_state())
        g.close()
    for r in runs[1:]:
        assert np.array_equal(r[0], runs[0][0]) and np.array_equal(r[1], runs[0][1])


def test_kernel_timing_leaves_results_alone():
    """CRM_FLAG_TIMING brackets every launch with HIP events (read back in batches of
    64 pairs): one launch per class and step, positive times, the same state as an
    untimed eager run, and crm_reset_timing clears the counts."""
    wl = workloads.c2_sorting_3d(n=6000, seed=17, steps=30)
    states, tim = [], None
    for timing in (False, True):
        g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=2, noise="philox", seed=5,
                             precision="mixed", graph=False, timing=timing)
        g.set_agents(wl.pos, wl.vel, wl.species)
        if timing:
            g.reset_timing()
            assert all(v["launches"] == 0 for v in g.timing().values())
        g.step(30)  # 120 timed launches: the 64-pair pool is read back mid-run
        if timing:
            tim = g.timing()
            g.reset_timing()
            assert all(v["launches"] == 0 and v["total_ms"] == 0.0 for v in g.timing().values())
        states.append(g.get_state())
        g.close()
    for cls in ("force", "scan", "scatter", "gather"):
        assert tim[cls]["launches"] == 30, (cls, tim[cls])
        assert tim[cls]["total_ms"] > 0.0, (cls, tim[cls])
    assert np.array_equal(states[0][0], states[1][0]) and np.array_equal(states[0][1], states[1][1])


def test_initial_placement_errors():
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [10.0, 10.0], 2.0)
    spec = [InstanceSpec(dom, MorsePotential(1.0, 1.0, 2.0, 1.0), NewtonDamped(1.0, 1.0))]
    g = MechanicsBackend(spec, dt=0.1)
    with pytest.raises(SimulationError) as e:  # p == min fails the strict check
        g.set_agents(np.array([[0.0, 5.0], [3.0, 3.0]]))
    assert e.value.kind == "BoundaryError"
    with pytest.raises(SimulationError) as e:  # p == max has no voxel
        g.set_agents(np.array([[10.0, 5.0], [3.0, 3.0]]))
    assert e.value.kind == "IndexError"
    g.set_agents(np.array([[1.0, 5.0], [3.0, 3.0]]))


def test_boundary_error_same_step_as_oracle():
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [10.0, 10.0], 2.0)
    spec = [InstanceSpec(dom, MorsePotential(1.0, 1.0, 2.0, 1.0), NewtonDamped(0.0, 1.0))]
    pos = np.array([[5.0, 5.0], [1.0, 1.0]])
    vel = np.array([[0.0, 0.0], [-300.0, 0.0]])  # x = 1 - 30 = -29, reflected to 29 > 10
    g = MechanicsBackend(spec, dt=0.1, mechanics_solver_order=0)
    o = OracleSim(spec, dt=0.1, mechanics_solver_order=0)
    g.set_agents(pos, vel)
    o.set_agents(pos, vel)
    with pytest.raises(SimulationError) as e:
        g.step(3)
    with pytest.raises(OracleError) as eo:
        o.step(3)
    assert e.value.kind == eo.value.kind == "BoundaryError"
    inf = g.info()
    assert inf.error_step_valid and inf.error_step == 0 and inf.error_agent == 1
    assert inf.step == 1  # the two steps after the failing one did not run


@pytest.mark.parametrize("graph,n_steps", [(True, 10), (True, 40), (False, 10)])
@pytest.mark.parametrize("case", ["newton", "brownian"])
def test_error_halts_at_the_failing_step(case, graph, n_steps):
    """The step loop breaks at the first step that fails (run_sim.rs:380-403): a
    10- or 40-step call whose step 4 (NewtonDamped, anti-damped so the speed grows;
    agents binned ahead by the gather) or step 3 (Brownian, binned after the forces)
    leaves the domain runs no later step, in graph replays of two or sixteen steps
    as in single launches.  The failing step is the oracle's, the step count stops at its end, and
    the state equals that of a context stepped one step at a time up to the failure."""
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [10.0, 10.0], 2.0)
    rng = np.random.default_rng(1)
    if case == "newton":
        spec = [InstanceSpec(dom, MorsePotential(1.0, 1.0, 2.0, 1.0), NewtonDamped(-5.0, 1.0))]
        pos, vel = rng.uniform(1, 9, (20, 2)), rng.normal(0, 3, (20, 2))
        kw, t_err = {}, 4
    else:
        spec = [InstanceSpec(dom, NoInteraction(), Brownian(100.0, 1.0))]
        pos, vel = rng.uniform(1, 9, (50, 2)), None
        kw, t_err = dict(noise="philox", seed=5), 3
    o = OracleSim(spec, dt=0.1, mechanics_solver_order=2, **kw)
    o.set_agents(pos, vel)
    o.step(t_err)
    with pytest.raises(OracleError) as eo:
        o.step(1)
    g = MechanicsBackend(spec, dt=0.1, mechanics_solver_order=2, graph=graph, **kw)
    g.set_agents(pos, vel)
    with pytest.raises(SimulationError) as e:
        g.step(n_steps)
    assert e.value.kind == eo.value.kind == "BoundaryError"
    inf = g.info()
    assert inf.error_step_valid and inf.error_step == t_err and inf.step == t_err + 1
    one = MechanicsBackend(spec, dt=0.1, mechanics_solver_order=2, graph=False, **kw)
    one.set_agents(pos, vel)
    for _ in range(t_err):
        one.step(1)
    with pytest.raises(SimulationError):
        one.step(1)
    assert one.info().step == t_err + 1
    for a, b in zip(g.get_state(), one.get_state()):
        assert np.array_equal(a, b)
    ha, hb = g.get_history(), one.get_history()
    assert ha[2] == hb[2] == 2
    assert np.array_equal(ha[0], hb[0]) and np.array_equal(ha[1], hb[1])
    assert np.array_equal(g.voxel_order(), one.voxel_order())
    g.close()
    one.close()


def test_mie_identical_positions_calc_error():
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [10.0] * 3, 2.5)
    spec = [InstanceSpec(dom, MiePotential(1.0, 1.0, 100.0, 2.5, 12.0, 6.0), NewtonDamped(1.0, 1.0))]
    g = MechanicsBackend(spec, dt=0.01)
    g.set_agents(np.array([[5.0, 5.0, 5.0], [5.0, 5.0, 5.0], [1.0, 1.0, 1.0]]))
    with pytest.raises(SimulationError) as e:
        g.compute_forces()
    assert e.value.kind == "CalcError"


def _msd(pos, x0):
    d2 = ((pos - x0) ** 2).sum(axis=1)
    return d2.mean(), d2.std()


@pytest.mark.parametrize("dim,damping,seed", [(3, 10.0, 1), (3, 1.0, 4), (3, 0.1, 5), (2, 10.0, 6)])
def test_langevin_msd_statistics(dim, damping, seed):
    """brownian_diffusion_constant_approx.rs:260-283: 160 particles from the centre,
    dt 1e-3, 200 steps, MSD within 3.5 standard deviations of the closed form."""
    L, n, dt, steps, D = 200.0, 160, 1e-3, 200, 1.0
    dom = CartesianCuboid.from_boundaries_and_n_voxels([0.0] * dim, [L] * dim, [3] * dim)
    kT = D * damping * 1.0
    spec = [InstanceSpec(dom, NoInteraction(), Langevin(mass=1.0, damping=damping, kb_temperature=kT))]
    g = MechanicsBackend(spec, dt=dt, noise="philox", seed=seed)
    x0 = np.full(dim, L / 2)
    g.set_agents(np.tile(x0, (n, 1)))
    for it in range(1, steps + 1):
        g.step(1)
        if it > 1 and it % 20 == 0:
            t = it * dt
            expected = (-dim * kT / damping ** 2 * (1 - math.exp(-damping * t)) * (3 - math.exp(-damping * t))
                        + 2 * dim * kT * t / damping)
            mean, sd = _msd(g.get_state()[0], x0)
            assert abs(expected - mean) <= 3.5 * sd


@pytest.mark.parametrize("dim,D,seed", [(3, 1.0, 1), (3, 0.5, 2), (2, 0.25, 3)])
def test_brownian_msd_statistics(dim, D, seed):
    """brownian_diffusion_constant_approx.rs:109-123: MSD = 2 d D t within 3.5 sigma."""
    L, n, dt, steps = 200.0, 160, 1e-3, 200
    dom = CartesianCuboid.from_boundaries_and_n_voxels([0.0] * dim, [L] * dim, [3] * dim)
    spec = [InstanceSpec(dom, NoInteraction(), Brownian(diffusion_constant=D, kb_temperature=1.0))]
    g = MechanicsBackend(spec, dt=dt, noise="philox", seed=seed)
    x0 = np.full(dim, L / 2)
    g.set_agents(np.tile(x0, (n, 1)))
    for it in range(20, steps + 1, 20):
        g.step(20)
        mean, sd = _msd(g.get_state()[0], x0)
        assert abs(2 * dim * D * it * dt - mean) <= 3.5 * sd


def test_free_drift_closed_form():
    """chili/mod.rs:134-156: no forces, no damping: x = x0 + k * dt * v0."""
    dom = CartesianCuboid.from_boundaries_and_interaction_range([-100.0] * 3, [100.0] * 3, 10.0)
    spec = [InstanceSpec(dom, NoInteraction(), NewtonDamped(damping_constant=0.0, mass=1.0))]
    g = MechanicsBackend(spec, dt=0.1)
    x0 = np.array([[1.0, 2.0, 3.0], [-4.0, 5.0, -6.0]])
    v0 = np.array([[0.5, -0.25, 1.0], [2.0, 1.0, -0.5]])
    g.set_agents(x0, v0)
    g.step(10)
    assert np.abs(g.get_state()[0] - (x0 + 10 * 0.1 * v0)).max() < 1e-5


def test_full_size_c2_invariants():
    """1M-agent cell-sorting configuration: size-independent checks after a few steps."""
    wl = workloads.c2_sorting_3d(n=1_000_000, seed=2, steps=3)
    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=2, noise="philox", seed=1,
                         precision="mixed")
    g.set_agents(wl.pos, wl.vel, wl.species)
    g.step(3)
    vox, start = g.get_voxels()
    pos = g.get_state()[0]
    assert start[-1] == wl.n_agents
    assert np.array_equal(np.bincount(vox, minlength=len(start) - 1), np.diff(start.astype(np.int64)))
    assert np.array_equal(vox.astype(np.int64), voxel_reference(wl.instances[0].domain, pos))
    assert np.isfinite(pos).all()


def test_autotune_keeps_results_and_caches(tmp_path):
    """The tiling autotuner changes only the work decomposition.  FP64 forces and
    states are bitwise those of the default tiling under a one-lane policy, and
    within 1e-14 under a policy with several lanes per agent (the partial sums'
    association; states within 1e-12 after 5 steps).  Mixed precision quantises
    positions with the tile region's fixed-point step, so its forces move with the
    tiling at that step's level (far inside the mixed bar); for one tiling they are
    bitwise reproducible.  The choice is written to and read from the JSON cache."""
    wl = workloads.c2_sorting_3d(n=60000, seed=40, steps=5)
    for precision in ("mixed", "fp64"):
        ref = MechanicsBackend(wl.instances, dt=wl.dt, precision=precision, noise="philox", seed=2)
        ref.set_agents(wl.pos, wl.vel, wl.species)
        path = tmp_path / f"tiling_{precision}.json"
        tuned = MechanicsBackend(wl.instances, dt=wl.dt, precision=precision, noise="philox", seed=2,
                                 autotune=str(path))
        tuned.set_agents(wl.pos, wl.vel, wl.species)
        assert tuned.tiling_policy is not None and path.exists()
        assert f"precision={1 if precision == 'mixed' else 0}" in path.read_text()
        t = tuned.tiling()
        assert t["rows"] >= 1 and t["cols"] >= 1 and t["tiles"] >= 1 and t["lanes_per_agent"] in (1, 2, 4, 8)
        one_lane = t["lanes_per_agent"] == 1
        f_tuned, f_ref = tuned.compute_forces(), ref.compute_forces()
        if precision == "fp64":
            assert np.array_equal(f_tuned, f_ref) if one_lane else rel_err(f_tuned, f_ref) <= 1e-14
        else:
            assert rel_err(f_tuned, f_ref) <= 1e-6
        again = MechanicsBackend(wl.instances, dt=wl.dt, precision=precision, noise="philox", seed=2,
                                 autotune=str(path))
        again.set_agents(wl.pos, wl.vel, wl.species)
        assert again.tiling_policy == tuned.tiling_policy
        assert np.array_equal(again.compute_forces(), f_tuned)
        tuned.step(5)
        ref.step(5)
        again.step(5)
        assert np.array_equal(again.get_state()[0], tuned.get_state()[0])
        if precision == "fp64":
            if one_lane:
                assert np.array_equal(tuned.get_state()[0], ref.get_state()[0])
            else:
                assert rel_err(tuned.get_state()[0], ref.get_state()[0]) <= 1e-12
        for h in (ref, tuned, again):
            h.close()
