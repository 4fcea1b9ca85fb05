"""Parity of the gfx950 rod path (RodMechanics + RodInteraction,
cell_building_blocks/bacterial_rods.rs:82-555) against the CPU oracle, through the C ABI.

Tolerances (FP64):
  forces, one-step positions            <= 1e-12 relative (max over rods of |dF| / max |F|)
  10-30 step trajectories               <= 1e-9 relative
  rod voxels (centroid binning)         bit-exact, same voxel start table
"""
import numpy as np
import pytest

from cellular_raza_amd import (
    BoundLennardJones,
    CartesianCuboid,
    InstanceSpec,
    MechanicsBackend,
    MiePotential,
    MorsePotential,
    NoInteraction,
    RodInteraction,
    RodMechanics,
    SimulationError,
    SpeciesMorsePotential,
    workloads,
)
from oracle.oracle import OracleError, OracleSim

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
TRAJ_TOL = 1e-9


def rel_err(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    num = np.linalg.norm(a - b, axis=1).max()
    den = np.linalg.norm(b, axis=1).max()
    return num / den if den > 0 else num


POTENTIALS = {
    "morse": lambda: MorsePotential(radius=0.5, potential_stiffness=2.0, cutoff=1.5, strength=0.1),
    "mie": lambda: MiePotential(radius=0.5, strength=0.05, bound=5.0, cutoff=1.5, en=12.0, em=6.0),
    "blj": lambda: BoundLennardJones(epsilon=0.2, sigma=0.8, bound=5.0, cutoff=1.5),
    "species": lambda: SpeciesMorsePotential(radius=0.5, potential_stiffness=2.0, cutoff=1.5, strength=0.1,
                                             species_matrix=[[1.0, 0.3], [0.6, 1.0]]),
    "none": lambda: NoInteraction(),
}


def rods_case(n=1200, L=60.0, nv=8, dim=2, seed=1, pot="morse", diffusion=0.0, jitter=0.05, vscale=0.1,
              irange=None, mech=None):
    """Straight rods, jittered vertices (so springs and bending act), random velocities."""
    rng = np.random.default_rng(seed)
    irange = irange if irange is not None else 0.5 * (nv - 1) + 1.5
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * dim, [L] * dim, irange)
    mech = mech or RodMechanics(n_vertices=nv, diffusion_constant=diffusion, spring_tension=10.0, rigidity=2.0,
                                spring_length=0.5, damping=1.5)
    pos = workloads.straight_rods(rng, n, dim, L, nv, 0.5)
    pos += rng.uniform(-jitter, jitter, size=pos.shape)
    vel = rng.normal(scale=vscale, size=pos.shape)
    species = (rng.random(n) < 0.5).astype(np.int32) if pot == "species" else None
    return [InstanceSpec(dom, RodInteraction(POTENTIALS[pot]()), mech)], pos, vel, species


def make_pair(spec, pos, vel, species, dt=0.002, order=2, **kw):
    g = MechanicsBackend(spec, dt=dt, mechanics_solver_order=order, precision="fp64", **kw)
    okw = {k: v for k, v in kw.items() if k in ("weighting", "history_order", "noise", "seed")}
    o = OracleSim(spec, dt=dt, mechanics_solver_order=order, **okw)
    g.set_agents(pos, vel, species)
    o.set_agents(pos, vel, species)
    return g, o


@pytest.mark.parametrize("weighting", ["reference", "physical"])
@pytest.mark.parametrize("pot", ["morse", "mie", "blj", "species", "none"])
def test_rod_forces(pot, weighting):
    spec, pos, vel, sp = rods_case(pot=pot, seed=3)
    g, o = make_pair(spec, pos, vel, sp, weighting=weighting)
    fo = o.compute_forces()
    fg = g.compute_forces()
    if pot == "none":
        assert np.all(fg == 0.0) and np.all(fo == 0.0)
        return
    assert np.abs(fo).max() > 0.0
    assert rel_err(fg, fo) <= FORCE_TOL


def test_rod_forces_3d_and_vertex_counts():
    for nv, dim in [(2, 2), (3, 2), (5, 3), (16, 2), (8, 3)]:
        n = 800 if dim == 2 else 1500
        L = 50.0 if dim == 2 else 22.0
        spec, pos, vel, sp = rods_case(n=n, L=L, nv=nv, dim=dim, seed=nv + dim)
        g, o = make_pair(spec, pos, vel, sp)
        fo = o.compute_forces()
        assert np.abs(fo).max() > 0.0, (nv, dim)
        assert rel_err(g.compute_forces(), fo) <= FORCE_TOL, (nv, dim)
        g.step(1)
        o.step(1)
        assert rel_err(g.get_state()[0], o.get_state()[0]) <= FORCE_TOL, (nv, dim)


def test_rod_one_step_and_trajectory():
    spec, pos, vel, sp = rods_case(seed=5)
    g, o = make_pair(spec, pos, vel, sp)
    g.step(1)
    o.step(1)
    assert rel_err(g.get_state()[0], o.get_state()[0]) <= FORCE_TOL
    g.step(29)
    o.step(29)
    pg, vg = g.get_state()
    po, vo = o.get_state()
    assert rel_err(pg, po) <= TRAJ_TOL
    assert rel_err(vg, vo) <= TRAJ_TOL
    assert np.abs(pg - pos.reshape(pg.shape)).max() > 1e-4  # rods moved


@pytest.mark.parametrize("order", [0, 1, 2])
def test_rod_solver_orders_and_history(order):
    spec, pos, vel, sp = rods_case(n=600, L=45.0, seed=7)
    g, o = make_pair(spec, pos, vel, sp, order=order)
    for _ in range(4):
        g.step(3)
        o.step(3)
        assert rel_err(g.get_state()[0], o.get_state()[0]) <= TRAJ_TOL
    dxg, dvg, ng = g.get_history()
    dxo, dvo, no = o.get_history()
    assert ng == no
    for i in range(ng):
        assert rel_err(dxg[i], dxo[i]) <= TRAJ_TOL
        assert rel_err(dvg[i], dvo[i]) <= TRAJ_TOL


def test_rod_voxels_centroid_binning_bit_exact():
    # fast rods so that many centroids change voxel
    spec, pos, vel, sp = rods_case(n=1500, L=60.0, seed=9, vscale=40.0)
    g, o = make_pair(spec, pos, vel, sp, dt=0.01)
    dom = spec[0].domain
    n = dom.get_n_voxels()
    dx = dom.get_dx()
    moved = 0
    v_prev = g.get_voxels()[0].copy()
    for _ in range(4):
        g.step(2)
        o.step(2)
        vg, sg = g.get_voxels()
        vo, so, order = o.get_voxels()
        assert np.array_equal(vg, vo)
        assert np.array_equal(sg, so)
        # the voxels' lists in the reference's order (stayers, then arrivals)
        assert np.array_equal(g.voxel_order(), order.astype(np.uint32))
        # independent check: centroid voxel by trunc((c - min) / dx)
        c = g.get_state()[0].reshape(len(pos), -1, 2).sum(axis=1) / pos.shape[1]
        idx = [np.trunc((c[:, k] - dom.min[k]) / dx[k]).astype(np.int64) for k in range(2)]
        plain = idx[0] * n[1] + idx[1]
        assert np.array_equal(vg.astype(np.int64), plain)
        moved += int((vg != v_prev).sum())
        v_prev = vg.copy()
    assert moved > 50


def test_rod_noise_philox_and_buffer():
    spec, pos, vel, sp = rods_case(n=800, L=50.0, seed=13, diffusion=0.5)
    g, o = make_pair(spec, pos, vel, sp, noise="philox", seed=77)
    g.step(10)
    o.step(10)
    assert rel_err(g.get_state()[0], o.get_state()[0]) <= TRAJ_TOL
    # caller-supplied normals: (steps, rods, vertices * dim)
    g2, o2 = make_pair(spec, pos, vel, sp, noise="buffer")
    z = np.random.default_rng(5).normal(size=(6, len(pos), pos.shape[1] * pos.shape[2]))
    g2.step(6, z)
    o2.step(6, z)
    p2 = g2.get_state()[0]
    assert rel_err(p2, o2.get_state()[0]) <= TRAJ_TOL
    g3, _ = make_pair(spec, pos, vel, sp)
    g3.step(6)
    assert np.abs(p2 - g3.get_state()[0]).max() > 1e-3  # the noise entered


def test_rod_graph_and_eager_identical():
    spec, pos, vel, sp = rods_case(n=1000, seed=15, diffusion=0.2)
    a = MechanicsBackend(spec, dt=0.002, precision="fp64", noise="philox", seed=4, graph=True)
    b = MechanicsBackend(spec, dt=0.002, precision="fp64", noise="philox", seed=4, graph=False)
    a.set_agents(pos, vel)
    b.set_agents(pos, vel)
    a.step(11)
    b.step(11)
    assert np.array_equal(a.get_state()[0], b.get_state()[0])
    assert np.array_equal(a.get_state()[1], b.get_state()[1])


def test_rod_boundary_error_same_step_as_oracle():
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [40.0, 40.0], 5.0)
    mech = RodMechanics(n_vertices=4, diffusion_constant=0.0, spring_tension=10.0, rigidity=2.0,
                        spring_length=0.5, damping=0.0)
    spec = [InstanceSpec(dom, RodInteraction(POTENTIALS["morse"]()), mech)]
    pos = np.array([[[10.0 + 0.5 * v, 10.0] for v in range(4)], [[20.0 + 0.5 * v, 20.0] for v in range(4)]])
    vel = np.zeros_like(pos)
    vel[1, :, 0] = -3000.0  # x = 20 - 300 -> reflected to far beyond 40
    g, o = make_pair(spec, pos, vel, None, dt=0.1, order=0)
    with pytest.raises(SimulationError) as e:
        g.step(3)
    with pytest.raises(OracleError) as eo:
        o.step(3)
    assert e.value.kind == eo.value.kind == "BoundaryError"
    inf = g.info()
    assert inf.error_step_valid and inf.error_step == 0 and inf.error_agent == 1
    assert inf.step == 1  # the two steps after the failing one did not run


@pytest.mark.parametrize("pot", ["morse", "mie", "blj", "species"])
def test_rod_mixed_precision(pot):
    """Mixed mode (float vertex-segment forces, FP64 state): forces within 1e-5 of the
    FP64 oracle, and for the soft Morse cases a 10-step trajectory within 1e-6 (the
    Mie / BoundLJ overlaps of this crowded case amplify any difference within a few
    steps)."""
    spec, pos, vel, sp = rods_case(pot=pot, seed=21)
    g = MechanicsBackend(spec, dt=0.002, precision="mixed")
    g.set_agents(pos, vel, sp)
    o = OracleSim(spec, dt=0.002)
    o.set_agents(pos, vel, sp)
    assert rel_err(g.compute_forces(), o.compute_forces()) <= 1e-5
    if pot in ("morse", "species"):
        g.step(10)
        o.step(10)
        assert rel_err(g.get_state()[0], o.get_state()[0]) <= 1e-6


def test_rod_create_rejects_bad_configs():
    spec, pos, vel, sp = rods_case(n=10, L=30.0)
    dom = spec[0].domain
    bad = RodMechanics(n_vertices=17, diffusion_constant=0.0, spring_tension=1.0, rigidity=1.0, spring_length=0.5,
                       damping=1.0)
    with pytest.raises(SimulationError):
        MechanicsBackend([InstanceSpec(dom, spec[0].interaction, bad)], dt=0.002)


def test_c3_scaled_invariants():
    """C3 rods at 20k rods: GPU vs oracle per-step forces and 5-step positions."""
    wl = workloads.c3_rods_2d(n=20_000, seed=3, steps=5)
    g, o = make_pair(wl.instances, wl.pos, wl.vel, None, dt=wl.dt, order=wl.solver_order)
    assert rel_err(g.compute_forces(), o.compute_forces()) <= FORCE_TOL
    g.step(5)
    o.step(5)
    assert rel_err(g.get_state()[0], o.get_state()[0]) <= TRAJ_TOL
    vg, sg = g.get_voxels()
    vo, so, order = o.get_voxels()
    assert np.array_equal(vg, vo) and np.array_equal(sg, so)
    assert np.array_equal(g.voxel_order(), order.astype(np.uint32))


def _rods_with_cluster(n, L, nv, dim, seed, n_cluster, spread):
    """Straight rods at C3's density plus a crowded cluster, whose blocks' stencil
    rows exceed the force kernel's LDS budget (they read global memory)."""
    rng = np.random.default_rng(seed)
    irange = max(0.5 * (nv - 1) + 1.5, 5.0 if dim == 2 else 3.5)
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * dim, [L] * dim, irange)
    mech = RodMechanics(n_vertices=nv, diffusion_constant=0.0, spring_tension=10.0, rigidity=2.0, spring_length=0.5,
                        damping=1.5)
    pos = workloads.straight_rods(rng, n, dim, L, nv, 0.5)
    centre = rng.normal(0.5 * L, spread, size=(n_cluster, dim))
    u = rng.normal(size=(n_cluster, dim))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    off = (np.arange(nv) - 0.5 * (nv - 1)) * 0.5
    cl = centre[:, None, :] + off[None, :, None] * u[:, None, :]
    pos = np.concatenate([pos, np.clip(cl, 0.3, L - 0.3)])
    pos += rng.uniform(-0.05, 0.05, size=pos.shape)
    sp = (rng.random(len(pos)) < 0.5).astype(np.int32)
    pot = SpeciesMorsePotential(radius=0.5, potential_stiffness=2.0, cutoff=1.5, strength=0.1,
                                species_matrix=[[1.0, 0.3], [0.6, 1.0]])
    return [InstanceSpec(dom, RodInteraction(pot), mech)], pos, sp


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("weighting", ["reference", "physical"])
def test_rod_lds_staging_equals_global_path(precision, weighting):
    """The rod force kernel copies each block's candidate rows (records, bounding
    and segment balls, species) to LDS when they fit and reads global memory
    otherwise; order and arithmetic are the same, so forces and neighbour counts
    are bitwise those with staging switched off.  Sparse and crowded blocks in
    one run, 2-D with 8 vertices and 3-D with 5; FP64 also against the oracle."""
    for n, L, nv, dim, n_cl, spread in [(6000, 350.0, 8, 2, 1500, 8.0), (3000, 90.0, 5, 3, 800, 4.0)]:
        spec, pos, sp = _rods_with_cluster(n, L, nv, dim, 7 + dim, n_cl, spread)
        out = []
        for lds in (True, False):
            g = MechanicsBackend(spec, dt=0.002, precision=precision, weighting=weighting, lds=lds)
            g.set_agents(pos, species=sp)
            g.set_neighbor_sensing(1.0)
            f = g.compute_forces()
            out.append((f, g.neighbor_counts()))
            g.close()
        assert np.abs(out[0][0]).max() > 0.0
        assert np.array_equal(out[0][0], out[1][0]), (dim, nv)
        assert np.array_equal(out[0][1], out[1][1]), (dim, nv)
        if precision == "fp64":
            o = OracleSim(spec, dt=0.002, weighting=weighting)
            o.set_agents(pos, species=sp)
            assert rel_err(out[0][0], o.compute_forces()) <= FORCE_TOL, (dim, nv)
