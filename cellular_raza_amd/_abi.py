"""ctypes mirror of include/crmech.h (structs and enums only, no library loading)."""
from __future__ import annotations

import ctypes as C

CRM_MAX_SPECIES = 4
CRM_MAX_ROD_VERTICES = 16

# crm_status
CRM_OK = 0
CRM_CALC_ERROR = 1
CRM_BOUNDARY_ERROR = 2
CRM_INDEX_ERROR = 3
CRM_RNG_ERROR = 4
CRM_SETUP_ERROR = 5
CRM_DEVICE_ERROR = 6
CRM_INVALID_ARGUMENT = 7
CRM_STORAGE_ERROR = 8
CRM_DIVISION_ERROR = 9

STATUS_NAMES = {
    CRM_OK: "Ok",
    CRM_CALC_ERROR: "CalcError",
    CRM_BOUNDARY_ERROR: "BoundaryError",
    CRM_INDEX_ERROR: "IndexError",
    CRM_RNG_ERROR: "RngError",
    CRM_SETUP_ERROR: "SetupError",
    CRM_DEVICE_ERROR: "DeviceError",
    CRM_INVALID_ARGUMENT: "InvalidArgument",
    CRM_STORAGE_ERROR: "StorageError",
    CRM_DIVISION_ERROR: "DivisionError",
}

# crm_potential_kind
POTENTIAL_MORSE = 0
POTENTIAL_MIE = 1
POTENTIAL_BOUND_LENNARD_JONES = 2
POTENTIAL_MORSE_SPECIES = 3
POTENTIAL_CELL_SORTING = 4
POTENTIAL_NONE = 5
POTENTIAL_CELL_SORTING_LANGEVIN = 6

# crm_mechanics_kind
MECHANICS_NEWTON_DAMPED = 0
MECHANICS_LANGEVIN = 1
MECHANICS_BROWNIAN = 2
MECHANICS_ROD = 3

PRECISION = {"fp64": 0, "mixed": 1}
WEIGHTING = {"reference": 0, "physical": 1}
HISTORY_ORDER = {"oldest_first": 0, "newest_first": 1}
NOISE = {"none": 0, "buffer": 1, "philox": 2}

FLAG_GRAPH = 1
FLAG_TIMING = 2
FLAG_NO_LDS = 4

KERNEL_CLASSES = ("force", "integrate", "scan", "scatter", "fixup", "gather")


class Domain(C.Structure):
    _fields_ = [
        ("min", C.c_double * 3),
        ("max", C.c_double * 3),
        ("interaction_range", C.c_double),
        ("n_voxels", C.c_uint32 * 3),
    ]


class Potential(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("n_species", C.c_uint32),
        ("radius", C.c_double),
        ("potential_stiffness", C.c_double),
        ("cutoff", C.c_double),
        ("strength", C.c_double),
        ("bound", C.c_double),
        ("en", C.c_double),
        ("em", C.c_double),
        ("epsilon", C.c_double),
        ("sigma", C.c_double),
        ("relative_interaction_range", C.c_double),
        ("sign", C.c_double),
        ("species_matrix", C.c_double * (CRM_MAX_SPECIES * CRM_MAX_SPECIES)),
    ]


class Mechanics(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("n_vertices", C.c_uint32),
        ("mass", C.c_double),
        ("damping", C.c_double),
        ("kb_temperature", C.c_double),
        ("diffusion_constant", C.c_double),
        ("spring_tension", C.c_double),
        ("rigidity", C.c_double),
        ("spring_length", C.c_double),
    ]


class Instance(C.Structure):
    _fields_ = [("domain", Domain), ("potential", Potential), ("mechanics", Mechanics)]


class RodCycle(C.Structure):
    _fields_ = [
        ("growth_base_rate", C.c_double),
        ("division_threshold", C.c_double),
        ("division_radius", C.c_double),
        ("sensing_radius", C.c_double),
        ("neighbor_saturation", C.c_double),
        ("max_agents", C.c_uint64),
    ]


DOMAIN_FORCE_NONE = 0
DOMAIN_FORCE_SPRING = 1
DOMAIN_FORCE_ROD_GEL = 2


class DomainForce(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("pad", C.c_uint32),
        ("strength", C.c_double * 3),
        ("center", C.c_double * 3),
        ("gel_pressure", C.c_double),
        ("surface_friction", C.c_double),
        ("surface_friction_distance", C.c_double),
    ]


class Config(C.Structure):
    _fields_ = [
        ("dim", C.c_uint32),
        ("solver_order", C.c_uint32),
        ("precision", C.c_uint32),
        ("weighting", C.c_uint32),
        ("history_order", C.c_uint32),
        ("noise_mode", C.c_uint32),
        ("seed", C.c_uint64),
        ("dt", C.c_double),
        ("device", C.c_int32),
        ("flags", C.c_uint32),
        ("n_subdomains", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class Info(C.Structure):
    _fields_ = [
        ("n_agents", C.c_uint64),
        ("n_voxels_total", C.c_uint64),
        ("step", C.c_uint64),
        ("n_instances", C.c_uint32),
        ("n_vertices", C.c_uint32),
        ("dim", C.c_uint32),
        ("error_step_valid", C.c_uint32),
        ("error_step", C.c_uint64),
        ("error_agent", C.c_uint64),
    ]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double * 6), ("launches", C.c_uint64 * 6)]


class JsonCells(C.Structure):
    _fields_ = [
        ("cell_templates", C.POINTER(C.c_char_p)),
        ("n_templates", C.c_uint32),
        ("storage_instance", C.c_uint32),
    ]


# Every symbol include/crmech.h declares, with (restype, argtypes).
P = C.c_void_p
DP = C.POINTER(C.c_double)
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
U8P = C.POINTER(C.c_uint8)
I32P = C.POINTER(C.c_int32)

SIGNATURES = {
    "crm_grid": (C.c_int, [C.POINTER(Domain), C.c_uint32, U64P, DP]),
    "crm_create": (C.c_int, [C.POINTER(Config), C.POINTER(Instance), C.c_uint32, C.POINTER(P)]),
    "crm_set_agents": (C.c_int, [P, U64P, DP, DP, I32P]),
    "crm_step": (C.c_int, [P, C.c_uint32, DP]),
    "crm_compute_forces": (C.c_int, [P, DP]),
    "crm_get_state": (C.c_int, [P, DP, DP]),
    "crm_get_history": (C.c_int, [P, DP, DP, U32P]),
    "crm_get_voxels": (C.c_int, [P, U32P, U32P]),
    "crm_get_voxel_order": (C.c_int, [P, U32P]),
    "crm_get_info": (C.c_int, [P, C.POINTER(Info)]),
    "crm_autotune": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_uint32)]),
    "crm_get_tiling": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "crm_set_neighbor_sensing": (C.c_int, [P, C.c_uint32, C.c_double]),
    "crm_get_neighbor_counts": (C.c_int, [P, U32P]),
    "crm_get_snapshot": (C.c_int, [P, DP, C.POINTER(C.c_int32), U32P]),
    "crm_set_domain_force": (C.c_int, [P, C.c_uint32, C.POINTER(DomainForce)]),
    "crm_set_rod_cycle": (C.c_int, [P, C.POINTER(RodCycle)]),
    "crm_set_rod_spring_lengths": (C.c_int, [P, DP]),
    "crm_get_cycle_state": (C.c_int, [P, DP, DP, U64P, U64P, U32P]),
    "crm_get_identifiers": (C.c_int, [P, U64P, U64P, U32P]),
    "crm_remove_agents": (C.c_int, [P, U8P]),
    "crm_remove_agents_noalloc": (C.c_int, [P, U8P]),
    "crm_insert_agents": (C.c_int, [P, C.c_uint64, U64P, DP, DP, DP, DP, I32P]),
    "crm_save_cells_json": (C.c_int, [P, C.c_char_p, C.c_uint64, C.POINTER(JsonCells)]),
    "crm_get_timing": (C.c_int, [P, C.POINTER(Timing)]),
    "crm_reset_timing": (C.c_int, [P]),
    "crm_last_error": (C.c_size_t, [P, C.c_char_p, C.c_size_t]),
    "crm_destroy": (None, [P]),
}
