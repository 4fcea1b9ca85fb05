"""The DomainForce aspect on the GPU, through the C ABI: the reference's spring test
(cellular_raza/tests/domain_force_spring.rs) end to end with Memory storage, and
forces / trajectories against the oracle (FP64 <= 1e-12 per step, <= 1e-9 over a
trajectory; mixed <= 1e-5) for points and rods, spring and rod gel."""
import math

import numpy as np
import pytest

from cellular_raza_amd import (
    CartesianCuboid,
    DomainSpring,
    FixedStepsize,
    InstanceSpec,
    MechanicsBackend,
    NewtonDamped,
    NoInteraction,
    RodGelForce,
    Settings,
    SimulationError,
    StorageBuilder,
    StorageOption,
    run_simulation,
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


def test_spring_single_particle(tmp_path):
    """domain_force_spring.rs:71-112 as written: Memory storage, histories loaded
    with load_all_element_histories, |x - x0 cos(omega t)| < 1e-3 at every save point;
    and the same trajectory as the oracle's."""
    k, m, dt, x0 = 0.0032, 0.5, 0.001, 10.0
    dom = CartesianCuboid.from_boundaries_and_n_voxels([-77.0, -77.0], [77.0, 77.0], [3, 3])
    time = FixedStepsize.from_partial_save_steps(0.0, dt, 10000, 100)
    storage = StorageBuilder().priority([StorageOption.Memory]).location(tmp_path)
    access = run_simulation(dom, NoInteraction(), NewtonDamped(damping_constant=0.0, mass=m), np.array([[x0, 0.0]]),
                            Settings(time=time, storage=storage), velocities=np.zeros((1, 2)),
                            domain_force=DomainSpring(strength=[k, 0.0]))
    hists = access.cells.load_all_element_histories()
    (_, history), = hists.items()
    omega = math.sqrt(k / m)
    assert sorted(history) == list(range(0, 10001, 100))
    for it, (cbox, _aux) in history.items():
        x = cbox["cell"]["mechanics"]["pos"][0]
        assert abs(x - x0 * math.cos(omega * it * dt)) < 1e-3, it
    spec = [InstanceSpec(dom, NoInteraction(), NewtonDamped(damping_constant=0.0, mass=m), DomainSpring([k, 0.0]))]
    o = OracleSim(spec, dt=dt, mechanics_solver_order=2)
    o.set_agents(np.array([[x0, 0.0]]), np.zeros((1, 2)))
    o.step(10000)
    assert abs(history[10000][0]["cell"]["mechanics"]["pos"][0] - o.get_state()[0][0, 0]) <= 1e-9 * x0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_points_spring_with_interaction(precision):
    wl = workloads.c2_sorting_3d(n=3000, seed=31)
    L = wl.instances[0].domain.max[0]
    force = DomainSpring(strength=[0.01, 0.02, 0.005], center=[0.5 * L, 0.4 * L, 0.6 * L])
    spec = [InstanceSpec(wl.instances[0].domain, wl.instances[0].interaction, wl.instances[0].mechanics, force)]
    g = MechanicsBackend(spec, dt=wl.dt, mechanics_solver_order=2, precision=precision)
    o = OracleSim(spec, dt=wl.dt, mechanics_solver_order=2)
    g.set_agents(wl.pos, wl.vel, wl.species)
    o.set_agents(wl.pos, wl.vel, wl.species)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert rel_err(g.compute_forces(), o.compute_forces()) <= tol
    g.step(20)
    o.step(20)
    assert rel_err(g.get_state()[0], o.get_state()[0]) <= (1e-9 if precision == "fp64" else 1e-6)
    # removing it restores the interaction-only forces
    g.set_domain_force(None)
    o.set_domain_force(None)
    assert rel_err(g.compute_forces(), o.compute_forces()) <= tol
    g.close()
    o.close()


@pytest.mark.parametrize("kind,dim", [("spring", 2), ("spring", 3), ("gel", 3)])
def test_rods_domain_force(kind, dim):
    from test_gpu_rods import rods_case

    spec, pos, vel, sp = rods_case(n=600, L=24.0 if dim == 3 else 50.0, nv=6, dim=dim, seed=32)
    L = spec[0].domain.max[0]
    force = (DomainSpring(strength=[0.03] * dim, center=[0.5 * L] * dim) if kind == "spring" else
             RodGelForce(gel_pressure=0.4, surface_friction=0.6, surface_friction_distance=3.0))
    spec = [InstanceSpec(spec[0].domain, spec[0].interaction, spec[0].mechanics, force)]
    g = MechanicsBackend(spec, dt=0.002, mechanics_solver_order=2)
    o = OracleSim(spec, dt=0.002, mechanics_solver_order=2)
    g.set_agents(pos, vel, sp)
    o.set_agents(pos, vel, sp)
    assert rel_err(g.compute_forces(), o.compute_forces()) <= 1e-12
    g.step(15)
    o.step(15)
    assert rel_err(g.get_state()[0], o.get_state()[0]) <= 1e-9
    g.close()
    o.close()


def test_gel_needs_3d_rods():
    wl = workloads.c2_sorting_3d(n=200, seed=33)
    g = MechanicsBackend(wl.instances, dt=wl.dt)
    with pytest.raises(SimulationError, match="InvalidArgument"):
        g.set_domain_force(RodGelForce(gel_pressure=1.0, surface_friction=1.0, surface_friction_distance=1.0))
    g.close()
