"""Lanes per agent in the force kernel, against the CPU oracle, through the C ABI.

A tiling policy with 2 or 4 lanes per agent splits each agent's candidate stream
(update_mechanics.rs:153-238: the 3^D voxel stencil in the reference's order) into
equal parts, one per lane, and adds the parts' force sums with wavefront shuffles.
The pairs and their arithmetic are those of one lane per agent; only the sum's
association changes.  Bars: FP64 forces within 1e-12 of the oracle and within
1e-14 of the one-lane result; mixed precision within 1e-5 of the FP64 oracle;
neighbour counts bit-exact; a 20-step FP64 trajectory within 1e-9 of the oracle.
The policy is pinned through the tuning cache file (crm_autotune reads it first).
"""
import json

import numpy as np
import pytest

from cellular_raza_amd import MechanicsBackend, workloads
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

# policy index -> lanes per agent (the library's table: 0-7 one lane, 8-9 two, 10-11 four,
# 12-13 one lane with up to 1.5 or 1.9 own agents per lane, 14-15 eight)
POLICIES = {0: 1, 8: 2, 9: 2, 10: 4, 11: 4, 14: 8, 15: 8}
CROWDED = (12, 13)


def rel_err(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    den = np.linalg.norm(b, axis=1).max()
    num = np.linalg.norm(a - b, axis=1).max()
    return num / den if den > 0 else num


KEYS = {}  # (workload, precision) -> the tuning cache key the library wrote for it


def pinned(name, wl, precision, policy, path, **kw):
    """A backend whose tiling policy is `policy`: the first autotune run of a
    workload writes its cache key (later runs in this process find it in the
    process cache and write nothing), and a cache file holding that key with
    `policy` is read by the backend's own autotune call."""
    if (name, precision) not in KEYS:
        first = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order,
                                 precision=precision, autotune=str(path), **kw)
        first.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
        first.close()
        KEYS[(name, precision)] = list(json.loads(path.read_text()))
    path.write_text(json.dumps({k: policy for k in KEYS[(name, precision)]}))
    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, precision=precision,
                         autotune=str(path), **kw)
    g.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    assert g.tiling_policy == policy
    return g


WORKLOADS = {
    "sorting_3d": lambda: workloads.c2_sorting_3d(n=20000, seed=31, steps=20),
    "morse_2d": lambda: workloads.c1_morse_2d(n=3000, seed=32, steps=20),
    "ensemble_2d": lambda: workloads.c4_ensemble_2d(n_instances=5, seed=33, n_min=300, n_max=900, steps=20),
    "mie_3d": lambda: workloads.c5_mie_3d(n=8000, seed=34, steps=20),
}


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_lanes_per_agent_forces_match_oracle(name, tmp_path):
    wl = WORKLOADS[name]()
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order)
    o.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    f_o = o.compute_forces()
    one = {}
    for policy, lanes in POLICIES.items():
        for precision in ("fp64", "mixed"):
            g = pinned(name, wl, precision, policy, tmp_path / f"p{policy}_{precision}.json")
            assert g.tiling()["lanes_per_agent"] == lanes
            f = g.compute_forces()
            if precision == "fp64":
                assert rel_err(f, f_o) <= 1e-12, (policy, rel_err(f, f_o))
                if lanes == 1:
                    one[name] = f
                else:
                    assert rel_err(f, one[name]) <= 1e-14, (policy, rel_err(f, one[name]))
            else:
                assert rel_err(f, f_o) <= 1e-5, (policy, rel_err(f, f_o))
            g.close()


@pytest.mark.parametrize("lanes_policy", [8, 10, 14])
def test_lanes_per_agent_counts_and_trajectory(lanes_policy, tmp_path):
    """Neighbour counts (integers, summed over the lanes of a group) equal the
    oracle's; the FP64 trajectory follows the oracle for 20 steps."""
    wl = WORKLOADS["sorting_3d"]()
    for precision in ("fp64", "mixed"):
        g = pinned("sorting_3d", wl, precision, lanes_policy, tmp_path / f"c_{precision}.json")
        for radius in (1.2, 2.5):  # inside the cutoff (2.5) and at it
            g.set_neighbor_sensing(radius)
            g.compute_forces()
            o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order)
            o.set_agents(wl.pos, None, wl.species, wl.counts)
            assert np.array_equal(g.neighbor_counts(), o.neighbor_counts(radius))
            o.close()
        g.set_neighbor_sensing(None)
        g.close()
    g = pinned("sorting_3d", wl, "fp64", lanes_policy, tmp_path / "traj.json", noise="philox", seed=4)
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order, noise="philox", seed=4)
    o.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    g.step(20)
    o.step(20)
    assert rel_err(g.get_state()[0], o.get_state()[0]) <= 1e-9
    g.close()


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_tiles_with_more_agents_than_lanes(name, tmp_path):
    """Policies whose tiles hold more own agents than the workgroup has lanes (a
    lane walks several agents in turn): FP64 forces bitwise those of policy 0
    (the candidate order of an agent does not depend on the tiling), mixed within
    the mixed bar of the oracle."""
    wl = WORKLOADS[name]()
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=wl.solver_order)
    o.set_agents(wl.pos, wl.vel, wl.species, wl.counts)
    f_o = o.compute_forces()
    base = pinned(name, wl, "fp64", 0, tmp_path / "p0_fp64.json")
    f0 = base.compute_forces()
    base.close()
    for policy in CROWDED:
        for precision in ("fp64", "mixed"):
            g = pinned(name, wl, precision, policy, tmp_path / f"p{policy}_{precision}.json")
            assert g.tiling()["lanes_per_agent"] == 1
            f = g.compute_forces()
            if precision == "fp64":
                assert np.array_equal(f, f0), policy
            else:
                assert rel_err(f, f_o) <= 1e-5, (policy, rel_err(f, f_o))
            g.close()
