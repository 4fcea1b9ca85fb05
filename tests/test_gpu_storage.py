"""Snapshots and the JSON storage boundary on the GPU, through the C ABI.

The device gather (crm_get_snapshot) must reproduce crm_get_state / crm_get_history
bitwise; the JSON written by the library (crm_save_cells_json) must hold the same
numbers bitwise (shortest round-trip decimal), in the reference's layout, readable
the way the reference's plot scripts read it (cell_sorting/src/plot.py:27-35,
bacterial_rods/plotter.py:22-42); run_simulation with a StorageBuilder stores the
trajectory that run_simulation returns without one.
"""
import glob
import json
import os

import numpy as np
import pytest

from cellular_raza_amd import (
    CartesianCuboid,
    CellSpecificInteraction,
    FixedStepsize,
    InstanceSpec,
    Langevin,
    MechanicsBackend,
    MorsePotential,
    NewtonDamped,
    Settings,
    SimulationError,
    StorageBuilder,
    StorageOption,
    run_simulation,
    workloads,
)
from cellular_raza_amd import storage as st

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_snapshot_matches_state_and_history(order, precision):
    wl = workloads.c2_sorting_3d(n=5000, seed=21)
    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=order, precision=precision,
                         noise="philox", seed=3)
    g.set_agents(wl.pos, wl.vel, wl.species)
    for steps in (0, 1, 4):
        if steps:
            g.step(steps)
        s = g.snapshot()
        pos, vel = g.get_state()
        hdx, hdv, n_stored = g.get_history()
        assert np.array_equal(s["pos"], pos) and np.array_equal(s["vel"], vel)
        assert s["dx"].shape[1] == n_stored
        for i in range(n_stored):
            assert np.array_equal(s["dx"][:, i], hdx[i]) and np.array_equal(s["dv"][:, i], hdv[i])
        assert np.array_equal(s["species"], wl.species)
    g.close()


def load_cells(path, iteration):
    """As the reference's cell-sorting plot script reads a save point."""
    cells = []
    for batch in glob.glob(os.path.join(path, f"{iteration:020d}", "*")):
        with open(batch) as f:
            cells.extend(json.load(f)["data"])
    return cells


def test_json_cells_hold_the_state_bitwise(tmp_path):
    wl = workloads.c2_sorting_3d(n=4000, seed=22)
    inter = CellSpecificInteraction(cell_radius=1.0, potential_strength=0.5, relative_interaction_range=2.5)
    mech = NewtonDamped(damping_constant=2.0, mass=1.0)
    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=2, noise="philox", seed=4)
    g.set_agents(wl.pos, wl.vel, wl.species)
    g.step(3)
    templates = st.cell_templates(mech, inter, n_species=2)
    path = tmp_path / "cells" / "json"
    g.save_cells_json(str(path), 3, templates)
    g.save_cells_json(str(path), 3, templates)  # a second batch at the same iteration
    names = sorted(os.listdir(path / f"{3:020d}"))
    assert names == [f"batch_{0:020d}_{0:020d}.json", f"batch_{0:020d}_{1:020d}.json"]
    with open(path / f"{3:020d}" / names[0]) as f:
        data = json.load(f)["data"]
    pos, vel = g.get_state()
    hdx, hdv, n_stored = g.get_history()
    assert len(data) == g.n_agents
    for a, d in enumerate(data):
        assert d["identifier"] == {"Initial": a}
        box, aux = d["element"]
        assert box["identifier"] == {"Initial": a} and box["parent"] is None
        c = box["cell"]
        assert np.array_equal(c["mechanics"]["pos"], pos[a]) and np.array_equal(c["mechanics"]["vel"], vel[a])
        assert c["mechanics"]["damping_constant"] == 2.0
        assert c["interaction"]["species"] == ["RedCell", "BlueCell"][wl.species[a]]
        m = aux["mechanics"]
        assert len(m["positions"]) == n_stored
        for i in range(n_stored):
            assert np.array_equal(m["positions"][i], hdx[i, a]) and np.array_equal(m["velocities"][i], hdv[i, a])
        assert m["current_force"] == [0.0, 0.0, 0.0] == m["zero_force"]
    # the reference's reading path: element[0].cell.mechanics.pos
    cells = load_cells(str(path), 3)
    assert len(cells) == 2 * g.n_agents
    g.close()


def test_json_rods_matrix_layout(tmp_path):
    """Rods: pos / vel as nalgebra's dynamic matrix; the rods example's plotter
    (x[0] reshaped to (dim, n_vertices)) recovers vertex j as column j."""
    from test_gpu_rods import rods_case

    spec, pos0, vel0, sp = rods_case(n=300, L=30.0, nv=5, dim=3, seed=5)
    g = MechanicsBackend(spec, dt=0.002, mechanics_solver_order=2)
    g.set_agents(pos0, vel0, sp)
    g.step(2)
    t = st.cell_templates(spec[0].mechanics, spec[0].interaction, extra={"growth_rate": 0.1})
    g.save_cells_json(str(tmp_path), 2, t)
    pos, vel = g.get_state()
    cells = [d["element"][0] for d in load_cells(str(tmp_path), 2)]
    for a, c in enumerate(cells):
        p = np.array(c["cell"]["mechanics"]["pos"][0], dtype=float).reshape(3, -1)
        assert np.array_equal(p.T, pos[a].reshape(5, 3))
        assert c["cell"]["mechanics"]["pos"][1:] == [5, 3]
        assert c["cell"]["growth_rate"] == 0.1
        assert np.array_equal(np.array(c["cell"]["mechanics"]["vel"][0]).reshape(3, -1).T, vel[a].reshape(5, 3))
    g.close()


@pytest.mark.parametrize("option", [StorageOption.SerdeJson, StorageOption.Memory])
def test_run_simulation_stores_the_trajectory(tmp_path, option):
    rng = np.random.default_rng(6)
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [60.0, 60.0], 4.0)
    inter = MorsePotential(radius=1.0, potential_stiffness=1.0, cutoff=4.0, strength=0.2)
    mech = Langevin(mass=1.0, damping=1.0, kb_temperature=0.1)
    pos = rng.uniform(2.0, 58.0, size=(600, 2))
    time = FixedStepsize.from_partial_save_steps(0.0, 0.01, 40, 10)
    ref = run_simulation(dom, inter, mech, pos, Settings(time=time), noise="philox", seed=9)
    builder = StorageBuilder().location(tmp_path).priority([option]).add_date(False)
    access = run_simulation(dom, inter, mech, pos, Settings(time=time, storage=builder), noise="philox", seed=9)
    assert sorted(access.cells.get_all_iterations()) == sorted(ref) == [0, 10, 20, 30, 40]
    for it, p in ref.items():
        cells = access.cells.load_all_elements_at_iteration(it)
        got = np.array([cells[("Initial", a)][0]["cell"]["mechanics"]["pos"] for a in range(len(pos))])
        assert np.array_equal(got, p), it
        sub = access.subdomains.load_single_element(it, 0)
        assert sub["domain_min"] == [0.0, 0.0] and sub["domain_max"] == [60.0, 60.0]
    if option is StorageOption.SerdeJson:
        assert os.path.isdir(tmp_path / "cells" / "json" / f"{40:020d}")
        assert os.path.isdir(tmp_path / "subdomains" / "json" / f"{40:020d}")
        hist = access.cells.load_element_history(("Initial", 7))
        assert sorted(hist) == [0, 10, 20, 30, 40]


def test_save_errors(tmp_path):
    wl = workloads.c2_sorting_3d(n=300, seed=23)
    g = MechanicsBackend(wl.instances, dt=wl.dt)
    g.set_agents(wl.pos, wl.vel, wl.species)
    blocker = tmp_path / "file"
    blocker.write_text("not a folder")
    with pytest.raises(SimulationError, match="StorageError"):
        g.save_cells_json(str(blocker / "json"), 0, ["{}"])
    with pytest.raises(SimulationError, match="InvalidArgument"):
        g.save_cells_json(str(tmp_path), 0, [])
    g.close()


def rel_err(a, b):
    a = np.asarray(a, dtype=float).reshape(len(a), -1)
    b = np.asarray(b, dtype=float).reshape(len(b), -1)
    den = np.abs(b).max()
    return np.abs(a - b).max() / den if den > 0 else np.abs(a - b).max()


@pytest.mark.parametrize("mech", ["newton", "langevin"])
def test_snapshot_and_json_match_the_oracle(tmp_path, mech):
    """The stored save point against the oracle's state and AuxStorageMechanics after
    the same steps (AB3, so the ring is full): snapshot records and the JSON's cell
    pos / vel and ring entries (oldest first) within 1e-12; the element in serde's
    field order: CellBox {identifier, parent, cell} (cell.rs:128-137) and
    AuxStorageMechanics {positions, velocities, current_force, zero_force}
    (aux_storage.rs:41-46), inside CombinedSaveFormat {identifier, element}
    (storage/concepts.rs:176-190) in a {"data": [...]} batch (serde_json.rs:19-74)."""
    from oracle.oracle import OracleSim

    wl = workloads.c2_sorting_3d(n=3000, seed=24)
    inst = wl.instances[0]
    noise = "philox"
    if mech == "newton":
        mechanics = NewtonDamped(damping_constant=2.0, mass=1.0)
        wl.instances = [InstanceSpec(inst.domain, inst.interaction, mechanics)]
        noise = "none"
    g = MechanicsBackend(wl.instances, dt=wl.dt, mechanics_solver_order=2, noise=noise, seed=8)
    o = OracleSim(wl.instances, dt=wl.dt, mechanics_solver_order=2, noise=noise, seed=8)
    g.set_agents(wl.pos, wl.vel, wl.species)
    o.set_agents(wl.pos, wl.vel, wl.species)
    g.step(5)
    o.step(5)
    po, vo = o.get_state()
    dxo, dvo, no = o.get_history()
    s = g.snapshot()
    assert s["dx"].shape[1] == no == 2
    assert rel_err(s["pos"], po) <= 1e-12 and rel_err(s["vel"], vo) <= 1e-12
    for i in range(no):
        assert rel_err(s["dx"][:, i], dxo[i]) <= 1e-12 and rel_err(s["dv"][:, i], dvo[i]) <= 1e-12
    templates = st.cell_templates(wl.instances[0].mechanics, wl.instances[0].interaction, n_species=2)
    g.save_cells_json(str(tmp_path), 5, templates)
    data = load_cells(str(tmp_path), 5)
    assert len(data) == wl.n_agents
    P, V = np.empty_like(po), np.empty_like(vo)
    HX, HV = np.empty((no,) + po.shape), np.empty((no,) + vo.shape)
    for d in data:
        assert list(d) == ["identifier", "element"]
        box, aux = d["element"]
        assert list(box) == ["identifier", "parent", "cell"] and box["identifier"] == d["identifier"]
        assert list(aux) == ["mechanics"]
        assert list(aux["mechanics"]) == ["positions", "velocities", "current_force", "zero_force"]
        a = d["identifier"]["Initial"]
        P[a], V[a] = box["cell"]["mechanics"]["pos"], box["cell"]["mechanics"]["vel"]
        HX[:, a], HV[:, a] = aux["mechanics"]["positions"], aux["mechanics"]["velocities"]
    assert rel_err(P, po) <= 1e-12 and rel_err(V, vo) <= 1e-12
    for i in range(no):
        assert rel_err(HX[i], dxo[i]) <= 1e-12 and rel_err(HV[i], dvo[i]) <= 1e-12
    g.close()
    o.close()


@pytest.mark.parametrize("option", [StorageOption.SerdeJson, StorageOption.Memory])
def test_removal_keeps_identifiers_in_storage(tmp_path, option):
    """CycleEvent::Remove keeps every survivor's CellIdentifier (update_cycle.rs:41-55):
    after removals the stored cells are keyed by the survivors' original indices, and
    the cell stored under each identifier is the oracle's agent with that identifier
    (the oracle removes the same agents)."""
    from oracle.oracle import OracleSim

    rng = np.random.default_rng(26)
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [60.0, 60.0], 4.0)
    inter = MorsePotential(radius=1.0, potential_stiffness=1.0, cutoff=4.0, strength=0.2)
    mech = NewtonDamped(damping_constant=1.0, mass=1.0)
    n = 500
    pos = rng.uniform(2.0, 58.0, size=(n, 2))
    time = FixedStepsize.from_partial_save_steps(0.0, 0.01, 12, 4)
    masks = {}

    def remove(be, it):
        if it % 3:
            return None
        m = np.random.default_rng(it).random(be.n_agents) < 0.1
        masks[it] = m
        return m

    builder = StorageBuilder().location(tmp_path).priority([option]).add_date(False)
    access = run_simulation(dom, inter, mech, pos, Settings(time=time, storage=builder), remove=remove,
                            mechanics_solver_order=2)
    # the same run on the oracle, tracking the original index of every agent
    o = OracleSim([InstanceSpec(dom, inter, mech)], dt=0.01, mechanics_solver_order=2)
    o.set_agents(pos)
    orig = np.arange(n)
    for it in range(1, 13):
        o.step(1)
        if it in masks:
            o.remove_agents(masks[it])
            orig = orig[~masks[it]]
        if it % 4 == 0:
            cells = access.cells.load_all_elements_at_iteration(it)
            assert sorted(cells) == sorted(("Initial", int(a)) for a in orig)
            po = o.get_state()[0]
            got = np.array([cells[("Initial", int(a))][0]["cell"]["mechanics"]["pos"] for a in orig])
            assert rel_err(got, po) <= 1e-12, it
            for a in orig[:5]:
                assert cells[("Initial", int(a))][0]["identifier"] == {"Initial": int(a)}
                assert cells[("Initial", int(a))][0]["parent"] is None
    assert sum(int(m.sum()) for m in masks.values()) > 0
    o.close()
