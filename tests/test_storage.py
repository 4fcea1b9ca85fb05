"""Storage layer (StorageBuilder / StorageManager, SerdeJson and Memory options) on the
CPU: the reference's own storage tests (cellular_raza-core/src/storage/test.rs), its
file naming (storage/concepts.rs:702-740), the history semantics documented at
concepts.rs:872-898, and the cell text built from the building blocks."""
import json
import os

import numpy as np
import pytest

from cellular_raza_amd import (
    CartesianCuboid,
    CellSpecificInteraction,
    MorsePotential,
    NewtonDamped,
    RodInteraction,
    RodMechanics,
    StorageBuilder,
    StorageError,
    StorageManager,
    StorageOption,
)
from cellular_raza_amd import storage as st

OPTIONS = [StorageOption.SerdeJson, StorageOption.Memory]


def manager(tmp_path, option, instance=0):
    b = StorageBuilder().location(tmp_path / "tempdir").priority([option]).init()
    return StorageManager.open_or_create(b, instance)


def test_builder_defaults_and_paths(tmp_path):
    b = StorageBuilder()
    assert b.get_location() == "./out"
    assert b.get_priority() == StorageOption.default_priority() == [StorageOption.SerdeJson]
    assert b.get_suffix() == "" and b.get_add_date()
    b = b.location(tmp_path).suffix("cells").init_with_date("2024-01-01-T00-00-00")
    assert b.get_full_path() == os.path.join(str(tmp_path), "2024-01-01-T00-00-00", "cells")
    assert StorageBuilder().location(tmp_path).add_date(False).init().get_full_path() == str(tmp_path)
    with pytest.raises(StorageError):
        b.location("elsewhere")
    with pytest.raises(StorageError):
        StorageBuilder().get_full_path()
    assert StorageBuilder().priority([StorageOption.Memory, StorageOption.Memory]).get_priority() == [
        StorageOption.Memory]


def test_store_load_json(tmp_path):
    """store_load_json (storage/test.rs:1-25), with the load the reference leaves as TODO."""
    m = manager(tmp_path, StorageOption.SerdeJson)
    m.store_single_element(0, 22, {"data": [{"identifier": 22, "element": 3.0}]})
    assert m.load_single_element(0, 22)["data"][0]["element"] == 3.0


@pytest.mark.parametrize("option", OPTIONS)
def test_store_multiple_load_multiple_2_interfaces(tmp_path, option):
    """storage/test.rs:60-96: two instances store batches at one iteration; each loads
    the union."""
    m0 = manager(tmp_path, option, 0)
    m1 = m0.clone_to_new_instance(1)
    e0 = {i: float(i) for i in range(0, 10)}
    e1 = {i: float(i) for i in range(20, 30)}
    m0.store_batch_elements(100, e0.items())
    m1.store_batch_elements(100, e1.items())
    both = dict(e0)
    both.update(e1)
    assert m0.load_all_elements_at_iteration(100) == both
    assert m1.load_all_elements_at_iteration(100) == both


@pytest.mark.parametrize("option", OPTIONS)
def test_store_single_load_single_element(tmp_path, option):
    """storage/test.rs:98-116."""
    m = manager(tmp_path, option)
    for identifier, element in [(30, "This is sparta!"), (40, "This is not"), (35, "Something else")]:
        m.store_single_element(350, identifier, element)
        assert m.load_single_element(350, identifier) == element


@pytest.mark.parametrize("option", OPTIONS)
def test_store_single_load_multiple_elements(tmp_path, option):
    """storage/test.rs:118-133."""
    m = manager(tmp_path, option)
    items = [("hasta", [3, 12]), ("la", [3, 11]), ("vista", [45, 33])]
    for i, e in items:
        m.store_single_element(1001, i, e)
    loaded = m.load_all_elements_at_iteration(1001)
    for i, e in items:
        assert loaded[i] == e


def test_json_file_layout(tmp_path):
    """<path>/json/<iteration:020>/{batch,single}_<instance:020>_<n:020>.json, n the
    first free counter; {"data": [...]} and {"identifier", "element"} documents."""
    m = manager(tmp_path, StorageOption.SerdeJson, instance=3)
    m.store_batch_elements(7, [(1, 1.5)])
    m.store_batch_elements(7, [(2, 2.5)])
    m.store_single_element(7, 9, [1, 2])
    d = os.path.join(m.json_path, f"{7:020d}")
    names = sorted(os.listdir(d))
    assert names == [f"batch_{3:020d}_{0:020d}.json", f"batch_{3:020d}_{1:020d}.json",
                     f"single_{3:020d}_{0:020d}.json"]
    with open(os.path.join(d, names[1])) as f:
        assert json.load(f) == {"data": [{"identifier": 2, "element": 2.5}]}
    with open(os.path.join(d, names[2])) as f:
        assert json.load(f) == {"identifier": 9, "element": [1, 2]}
    assert m.get_all_iterations() == [7]
    assert m.load_all_elements_at_iteration(8) == {}


@pytest.mark.parametrize("option", OPTIONS)
def test_element_history(tmp_path, option):
    """The valid state of concepts.rs:878-889: each element's history is its
    contiguous run of iterations; tagged identifiers load as tuples."""
    m = manager(tmp_path, option)
    state = {0: ["E1", "E2", "E3"], 1: ["E1", "E2", "E3", "E4"], 2: ["E1", "E2", "E3", "E4"],
             3: ["E1", "E2", "E4"], 4: ["E2", "E4"], 5: ["E2", "E4", "E5"], 6: ["E4", "E5"]}
    for it, names in state.items():
        m.store_batch_elements(it, [(("Initial", int(n[1:])), f"{n}@{it}") for n in names])
    assert m.load_element_history(("Initial", 3)) == {0: "E3@0", 1: "E3@1", 2: "E3@2"}
    assert sorted(m.load_element_history(("Initial", 2))) == [0, 1, 2, 3, 4, 5]
    hist = m.load_all_element_histories()
    assert sorted(hist[("Initial", 4)]) == [1, 2, 3, 4, 5, 6]
    assert sorted(m.load_all_elements()) == list(range(7))


def test_cell_templates_are_json_with_the_reference_field_order():
    t = st.cell_templates(NewtonDamped(damping_constant=2.0, mass=1.0),
                          CellSpecificInteraction(cell_radius=6.0, potential_strength=2.0,
                                                  relative_interaction_range=1.5), n_species=2)
    assert len(t) == 2
    cells = [st.fill_template(x, np.array([1.0, 2.0, 3.0]), np.zeros(3), False) for x in t]
    assert list(cells[0]) == ["mechanics", "interaction"]
    assert list(cells[0]["mechanics"]) == ["pos", "vel", "damping_constant", "mass"]
    assert cells[0]["mechanics"]["pos"] == [1.0, 2.0, 3.0]
    assert [c["interaction"]["species"] for c in cells] == ["RedCell", "BlueCell"]
    assert list(cells[1]["interaction"]) == ["species", "cell_radius", "potential_strength",
                                             "relative_interaction_range"]
    rod = st.cell_templates(RodMechanics(n_vertices=3, diffusion_constant=0.0, spring_tension=1.0, rigidity=1.0,
                                         spring_length=0.5, damping=1.0),
                            RodInteraction(MorsePotential(radius=1.0, potential_stiffness=1.0, cutoff=2.0,
                                                          strength=0.1)))[0]
    p = np.arange(6.0).reshape(3, 2)
    c = st.fill_template(rod, p, p, True)
    # nalgebra dynamic matrix: [column-major data, rows, columns]; the rods example's
    # plotter reads data.reshape(dim, -1)[:, j] as vertex j (plotter.py:36-60)
    assert c["mechanics"]["pos"][1:] == [3, 2]
    assert np.array_equal(np.array(c["mechanics"]["pos"][0]).reshape(2, -1).T, p)
    assert list(c["interaction"]) == ["radius", "potential_stiffness", "cutoff", "strength"]


def test_subdomain_element():
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [10.0, 7.0], 2.0)
    e = st.subdomain_element(dom)
    assert e["domain_n_voxels"] == [5, 3] and e["domain_min"] == [0.0, 0.0] and e["domain_max"] == [10.0, 7.0]
    assert e["max"] == [10.0, 0.0 + 3 * (7.0 / 3)]
    assert e["voxels"][:4] == [[0, 0], [0, 1], [0, 2], [1, 0]] and len(e["voxels"]) == 15
