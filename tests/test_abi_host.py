"""CPU checks of the C ABI library and of the host layer (no GPU compute calls)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from cellular_raza_amd import (
    CartesianCuboid,
    FixedStepsize,
    InstanceSpec,
    MorsePotential,
    NewtonDamped,
    SpeciesMorsePotential,
    _abi,
    _native,
    workloads,
)
from cellular_raza_amd.building_blocks import instance_struct
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "crmech.h")).read()
    return sorted(set(re.findall(r"^\s*(?:crm_status|size_t|void)\s+(crm_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    declared = header_functions()
    assert declared, "no declarations parsed"
    lib = C.CDLL(_native.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    # the ctypes signature table covers exactly the header
    assert sorted(_abi.SIGNATURES) == declared


def test_struct_layouts_match_header_sizes():
    # sizes implied by the C declarations (x86-64 alignment rules)
    assert C.sizeof(_abi.Domain) == 72
    assert C.sizeof(_abi.Mechanics) == 64
    assert C.sizeof(_abi.Config) == 56
    assert C.sizeof(_abi.Potential) == 8 + 11 * 8 + 16 * 8


def test_crm_grid_matches_oracle_and_doctest():
    lib = _native.load()
    rng = np.random.default_rng(3)
    cases = [([2.0, 3.0, 1.0], [10.0, 10.0, 20.0], 2.0)]
    for _ in range(50):
        lo = rng.uniform(-10, 10, 3)
        cases.append((lo, lo + rng.uniform(0.5, 50, 3), rng.uniform(0.1, 10)))
    for lo, hi, r in cases:
        d = CartesianCuboid.from_boundaries_and_interaction_range(lo, hi, r)
        n = (C.c_uint64 * 3)()
        dx = (C.c_double * 3)()
        assert lib.crm_grid(C.byref(d.to_struct()), 3, n, dx) == 0
        on, odx = O.grid(d, 3)
        assert list(n) == on and list(dx) == odx
        assert list(n) == d.get_n_voxels() and list(dx) == d.get_dx()
    # invalid boundaries: BoundaryError
    bad = _abi.Domain()
    bad.min[0], bad.max[0] = 1.0, 1.0
    bad.max[1] = 1.0
    bad.interaction_range = 0.1
    assert lib.crm_grid(C.byref(bad), 2, (C.c_uint64 * 3)(), (C.c_double * 3)()) == _abi.CRM_BOUNDARY_ERROR


def test_create_without_device_reports_error():
    lib = _native.load()
    wl = workloads.c1_morse_2d(n=10)
    cfg = _abi.Config()
    cfg.dim = 2
    cfg.dt = 0.01
    cfg.device = 0
    arr = (_abi.Instance * 1)(instance_struct(wl.instances[0].domain, wl.instances[0].interaction,
                                              wl.instances[0].mechanics))
    h = C.c_void_p()
    cfg.solver_order = 7  # invalid: rejected before any device call
    assert lib.crm_create(C.byref(cfg), arr, 1, C.byref(h)) == _abi.CRM_INVALID_ARGUMENT
    assert not h.value
    buf = C.create_string_buffer(256)
    lib.crm_last_error(None, buf, 256)
    assert b"solver_order" in buf.value


def test_fixed_stepsize_matches_reference_semantics():
    """time.rs:93-111: n_steps with save_interval -> save iterations; the cell-sorting
    benchmark's save_interval = n_steps + 1 executes n_steps + 1 iterations."""
    t = FixedStepsize.from_partial_save_steps(0.0, 0.1, 100, 10)
    assert t.save_iterations == list(range(0, 101, 10)) and t.maximum_iterations == 100
    t = FixedStepsize.from_partial_save_steps(0.0, 0.1, 10, 11)
    assert t.maximum_iterations == 11
    with pytest.raises(ValueError):
        FixedStepsize.from_partial_save_points(1.0, 0.1, [0.5])


def test_building_blocks_fill_structs():
    d = CartesianCuboid.from_boundaries_and_n_voxels([0.0, 0.0], [10.0, 20.0], [5, 4])
    assert d.get_n_voxels() == [5, 4] and d.get_dx() == [2.0, 5.0]
    inst = instance_struct(d, SpeciesMorsePotential(1.0, 2.0, 3.0, 0.5, [[1, 2], [3, 4]]),
                           NewtonDamped(damping_constant=1.5, mass=2.0))
    assert inst.potential.kind == _abi.POTENTIAL_MORSE_SPECIES and inst.potential.n_species == 2
    assert inst.potential.species_matrix[0 * 4 + 1] == 2 and inst.potential.species_matrix[1 * 4 + 0] == 3
    assert inst.mechanics.damping == 1.5 and inst.mechanics.mass == 2.0
    assert list(inst.domain.n_voxels)[:2] == [5, 4]
    with pytest.raises(ValueError):
        CartesianCuboid([0.0, 1.0], [0.0, 2.0], interaction_range=1.0)


def test_workload_geometry():
    """Derived geometry of the configurations (cell sorting: L for volume fraction
    0.3 and a 96^3 grid at N = 1e6; Mie lattice spacing 2; ensemble counts)."""
    L = (1e6 * 4 * math.pi / 3 / 0.3) ** (1 / 3)
    assert abs(L - 240.7996) < 1e-3
    d = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [L] * 3, 2.5)
    assert d.get_n_voxels() == [96, 96, 96]
    wl = workloads.c5_mie_3d(n=1000)
    assert wl.pos.min() > 0 and wl.pos.max() < wl.instances[0].domain.max[0]
    e = workloads.c4_ensemble_2d(n_instances=16, n_min=500, n_max=20000)
    assert len(e.counts) == 16 and all(480 <= c <= 20000 for c in e.counts)
    assert e.pos.shape == (sum(e.counts), 2)
