"""Storage of simulation results in the reference's layout (cellular_raza-core/src/storage).

Mirrors StorageOption / StorageBuilder / StorageManager (storage/concepts.rs:125-650)
for the SerdeJson and Memory options, with the reference's JSON folder and file
naming (concepts.rs:655-790: <location>/<date>/<suffix>/json/<iteration:020>/
batch_<instance:020>_<n:020>.json and single_<instance:020>_<n:020>.json) and its
save formats (CombinedSaveFormat {identifier, element}, BatchSaveFormat {data}).

Cells of a GPU simulation are written by the library (crm_save_cells_json): one
device gather into agent order, one copy, C++ formatting on host threads.  Their
JSON element is [CellBox {identifier, parent, cell}, AuxStorage {mechanics}] as
run_simulation! stores it (datastructures.rs:371-396, aux_storage.rs:41-46), with
the cell text built from the building blocks' serde fields by `cell_templates`.

Identifiers that JSON cannot key a Python dict with (CellIdentifier's
externally tagged enum, {"Initial": 3}) are returned as tuples ("Initial", 3).
"""
from __future__ import annotations

import copy
import datetime
import enum
import glob
import json
import os
from dataclasses import dataclass
from typing import Any, Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import building_blocks as bb


class StorageError(RuntimeError):
    """StorageError (storage/concepts.rs:16-42)."""


class StorageOption(enum.Enum):
    """StorageOption (storage/concepts.rs:125-150); Sled, SledTemp and Ron are not provided."""

    SerdeJson = "SerdeJson"
    Memory = "Memory"

    @staticmethod
    def default_priority() -> List["StorageOption"]:
        return [StorageOption.SerdeJson]


def _unique(options: Iterable[StorageOption]) -> List[StorageOption]:
    out: List[StorageOption] = []
    for o in options:
        if not isinstance(o, StorageOption):
            raise StorageError(f"unsupported storage option {o!r} (SerdeJson and Memory are provided)")
        if o not in out:
            out.append(o)
    return out


class StorageBuilder:
    """StorageBuilder (storage/concepts.rs:176-300): location ./out, priority
    [SerdeJson], empty suffix, date folder added by default; `init` fixes the date."""

    def __init__(self):
        self._location = "./out"
        self._priority = [StorageOption.SerdeJson]
        self._suffix = ""
        self._add_date = True
        self._date: Optional[str] = None

    def _with(self, **kw) -> "StorageBuilder":
        b = copy.copy(self)
        for k, v in kw.items():
            setattr(b, k, v)
        return b

    @property
    def initialized(self) -> bool:
        return self._date is not None

    def priority(self, priority: Iterable[StorageOption]) -> "StorageBuilder":
        return self._with(_priority=_unique(priority))

    def get_priority(self) -> List[StorageOption]:
        return list(self._priority)

    def suffix(self, suffix) -> "StorageBuilder":
        return self._with(_suffix=str(suffix))

    def get_suffix(self) -> str:
        return self._suffix

    def add_date(self, add_date: bool) -> "StorageBuilder":
        return self._with(_add_date=bool(add_date))

    def get_add_date(self) -> bool:
        return self._add_date

    def location(self, location) -> "StorageBuilder":
        if self.initialized:
            raise StorageError("the location cannot change after init")
        return self._with(_location=str(location))

    def get_location(self) -> str:
        return self._location

    def init(self) -> "StorageBuilder":
        date = datetime.datetime.now().strftime("%Y-%m-%d-T%H-%M-%S") if self._add_date else ""
        return self.init_with_date(date)

    def init_with_date(self, date) -> "StorageBuilder":
        return self._with(_date=str(date))

    def get_full_path(self) -> str:
        if not self.initialized:
            raise StorageError("StorageBuilder.init must be called first")
        parts = [self._location]
        if self._add_date and self._date:
            parts.append(self._date)
        if self._suffix:
            parts.append(self._suffix)
        return os.path.join(*parts)


def _to_key(x):
    """JSON identifier -> hashable key (dicts of one tag become (tag, *values))."""
    if isinstance(x, dict) and len(x) == 1:
        (tag, val), = x.items()
        return (tag,) + (tuple(_to_key(v) for v in val) if isinstance(val, list) else (_to_key(val),))
    if isinstance(x, list):
        return tuple(_to_key(v) for v in x)
    return x


def _to_json_id(x):
    """Inverse of _to_key for identifiers given as ("Tag", values...)."""
    if isinstance(x, tuple) and x and isinstance(x[0], str):
        vals = [_to_json_id(v) for v in x[1:]]
        return {x[0]: vals[0] if len(vals) == 1 else vals}
    if isinstance(x, tuple):
        return [_to_json_id(v) for v in x]
    return x


def _jsonable(x):
    if isinstance(x, np.ndarray):
        return x.tolist()
    if isinstance(x, (np.floating, np.integer)):
        return x.item()
    if isinstance(x, dict):
        return {k: _jsonable(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_jsonable(v) for v in x]
    return x


class _JsonStore:
    """JsonStorageInterface wrapped in StorageWrapper (serde_json.rs, concepts.rs:655-1110)."""

    def __init__(self, path: str, instance: int):
        self.path = path
        self.instance = instance
        os.makedirs(path, exist_ok=True)

    def iteration_path(self, iteration: int) -> str:
        return os.path.join(self.path, f"{iteration:020d}")

    def new_file(self, iteration: int, mode: str) -> str:
        d = self.iteration_path(iteration)
        os.makedirs(d, exist_ok=True)
        n = 0
        while True:
            p = os.path.join(d, f"{mode}_{self.instance:020d}_{n:020d}.json")
            if not os.path.exists(p):
                return p
            n += 1

    def store_single_element(self, iteration, identifier, element):
        with open(self.new_file(iteration, "single"), "w") as f:
            json.dump({"identifier": _jsonable(_to_json_id(identifier)), "element": _jsonable(element)}, f,
                      indent=2)

    def store_batch_elements(self, iteration, pairs):
        data = [{"identifier": _jsonable(_to_json_id(i)), "element": _jsonable(e)} for i, e in pairs]
        with open(self.new_file(iteration, "batch"), "w") as f:
            json.dump({"data": data}, f, indent=2)

    def _entries(self, iteration):
        d = self.iteration_path(iteration)
        if not os.path.isdir(d):
            return
        for p in sorted(glob.glob(os.path.join(d, "*.json"))):
            head = os.path.basename(p).split("_")[0]
            with open(p) as f:
                content = json.load(f)
            if head == "batch":
                for e in content["data"]:
                    yield _to_key(e["identifier"]), e["element"]
            elif head == "single":
                yield _to_key(content["identifier"]), content["element"]

    def load_single_element(self, iteration, identifier):
        key = _to_key(_to_json_id(identifier)) if isinstance(identifier, tuple) else identifier
        for i, e in self._entries(iteration):
            if i == key:
                return e
        return None

    def load_all_elements_at_iteration(self, iteration) -> Dict[Any, Any]:
        return dict(self._entries(iteration))

    def get_all_iterations(self) -> List[int]:
        out = []
        for p in os.listdir(self.path):
            try:
                out.append(int(p))
            except ValueError:
                pass
        return out


class _MemoryStore:
    """MemoryStorageInterface (storage/memory_storage.rs): one map shared by all instances."""

    def __init__(self, shared: Optional[Dict[int, Dict[Any, Any]]] = None):
        self.map: Dict[int, Dict[Any, Any]] = {} if shared is None else shared

    def store_single_element(self, iteration, identifier, element):
        self.map.setdefault(iteration, {}).setdefault(identifier, copy.deepcopy(element))

    def store_batch_elements(self, iteration, pairs):
        self.map.setdefault(iteration, {}).update({i: copy.deepcopy(e) for i, e in pairs})

    def load_single_element(self, iteration, identifier):
        return copy.deepcopy(self.map.get(iteration, {}).get(identifier))

    def load_all_elements_at_iteration(self, iteration):
        return copy.deepcopy(self.map.get(iteration, {}))

    def get_all_iterations(self):
        return list(self.map.keys())


class StorageManager:
    """StorageManager (storage/concepts.rs:150-650): stores to every option of the
    priority list, loads from the first."""

    def __init__(self, builder: StorageBuilder, instance: int, stores: Dict[StorageOption, Any]):
        self.builder = builder
        self.instance = instance
        self._stores = stores

    @classmethod
    def open_or_create(cls, builder: StorageBuilder, instance: int = 0) -> "StorageManager":
        if not builder.initialized:
            builder = builder.init()
        full = builder.get_full_path()
        stores: Dict[StorageOption, Any] = {}
        for opt in builder.get_priority():
            if opt is StorageOption.SerdeJson:
                stores[opt] = _JsonStore(os.path.join(full, "json"), instance)
            elif opt is StorageOption.Memory:
                stores[opt] = _MemoryStore()
        return cls(builder, instance, stores)

    def clone_to_new_instance(self, instance: int) -> "StorageManager":
        stores: Dict[StorageOption, Any] = {}
        for opt, st in self._stores.items():
            stores[opt] = _JsonStore(st.path, instance) if opt is StorageOption.SerdeJson else _MemoryStore(st.map)
        return StorageManager(self.builder, instance, stores)

    @property
    def json_path(self) -> Optional[str]:
        st = self._stores.get(StorageOption.SerdeJson)
        return None if st is None else st.path

    def _first(self):
        for opt in self.builder.get_priority():
            return self._stores[opt]
        return None

    def store_single_element(self, iteration: int, identifier, element) -> None:
        for st in self._stores.values():
            st.store_single_element(int(iteration), identifier, element)

    def store_batch_elements(self, iteration: int, pairs: Iterable[Tuple[Any, Any]]) -> None:
        pairs = list(pairs)
        for st in self._stores.values():
            st.store_batch_elements(int(iteration), pairs)

    def load_single_element(self, iteration: int, identifier):
        st = self._first()
        return None if st is None else st.load_single_element(int(iteration), identifier)

    def load_all_elements_at_iteration(self, iteration: int) -> Dict[Any, Any]:
        st = self._first()
        return {} if st is None else st.load_all_elements_at_iteration(int(iteration))

    def get_all_iterations(self) -> List[int]:
        st = self._first()
        return [] if st is None else st.get_all_iterations()

    def load_element_history(self, identifier) -> Dict[int, Any]:
        """Every occurrence from the first one until the element is missing
        (concepts.rs:872-930)."""
        out: Dict[int, Any] = {}
        started = False
        for it in sorted(self.get_all_iterations()):
            e = self.load_single_element(it, identifier)
            if e is not None:
                started = True
                out[it] = e
            elif started:
                break
        return out

    def load_all_elements(self) -> Dict[int, Dict[Any, Any]]:
        return {it: self.load_all_elements_at_iteration(it) for it in sorted(self.get_all_iterations())}

    def load_all_element_histories(self) -> Dict[Any, Dict[int, Any]]:
        out: Dict[Any, Dict[int, Any]] = {}
        for it, elements in self.load_all_elements().items():
            for i, e in elements.items():
                out.setdefault(i, {})[it] = e
        return out


@dataclass
class StorageAccess:
    """chili::StorageAccess (run_sim.rs:402-405): the cells' and subdomains' managers."""

    cells: StorageManager
    subdomains: StorageManager


# --------------------------------------------------------------------------
# Cell text from the building blocks' serde fields
# --------------------------------------------------------------------------
_POS, _VEL = "__crm_pos__", "__crm_vel__"
# values of extra fields that the writer fills per rod
SPRING_LENGTH, GROWTH_RATE = "__crm_spring_length__", "__crm_growth_rate__"
_HOLES = {_POS: "{pos}", _VEL: "{vel}", SPRING_LENGTH: "{spring_length}", GROWTH_RATE: "{growth_rate}"}


def mechanics_fields(m) -> Dict[str, Any]:
    """Serde fields of the mechanics struct in declaration order (mechanics.rs:50-66,
    252-261, 410-426; bacterial_rods.rs:82-107), state as placeholders."""
    if isinstance(m, bb.NewtonDamped):
        return {"pos": _POS, "vel": _VEL, "damping_constant": m.damping_constant, "mass": m.mass}
    if isinstance(m, bb.Langevin):
        return {"pos": _POS, "vel": _VEL, "mass": m.mass, "damping": m.damping, "kb_temperature": m.kb_temperature}
    if isinstance(m, bb.Brownian):
        return {"pos": _POS, "diffusion_constant": m.diffusion_constant, "kb_temperature": m.kb_temperature}
    if isinstance(m, bb.RodMechanics):  # spring_length per rod
        return {"pos": _POS, "vel": _VEL, "diffusion_constant": m.diffusion_constant,
                "spring_tension": m.spring_tension, "rigidity": m.rigidity, "spring_length": SPRING_LENGTH,
                "damping": m.damping}
    raise StorageError(f"no serde layout for mechanics {type(m).__name__}")


def interaction_fields(i, species: int, species_names: Sequence[str]):
    """Serde value of the interaction (interaction.rs field orders; the cell-sorting
    benchmark's CellSpecificInteraction, cell_sorting.rs:23-29)."""
    if isinstance(i, bb.RodInteraction):  # newtype struct: the inner value
        return interaction_fields(i.interaction, species, species_names)
    if isinstance(i, bb.MorsePotential):
        return {"radius": i.radius, "potential_stiffness": i.potential_stiffness, "cutoff": i.cutoff,
                "strength": i.strength}
    if isinstance(i, bb.MiePotential):
        return {"radius": i.radius, "strength": i.strength, "bound": i.bound, "cutoff": i.cutoff, "en": i.en,
                "em": i.em}
    if isinstance(i, bb.BoundLennardJones):
        return {"epsilon": i.epsilon, "sigma": i.sigma, "bound": i.bound, "cutoff": i.cutoff}
    if isinstance(i, bb.CellSpecificInteraction):
        return {"species": species_names[species], "cell_radius": i.cell_radius,
                "potential_strength": i.potential_strength,
                "relative_interaction_range": i.relative_interaction_range}
    if isinstance(i, bb.CellSpecificInteractionLangevin):  # main-langevin.rs:35-43
        m = i.matrix()
        rb = m[0][1] if len(m) > 1 else i.attraction_rb
        bb_ = m[1][1] if len(m) > 1 else i.attraction_bb
        return {"species": species_names[species], "cell_radius": i.cell_radius, "attraction_rr": m[0][0],
                "attraction_bb": bb_, "attraction_rb": rb, "repulsion": i.repulsion, "cutoff": i.cutoff}
    if isinstance(i, bb.SpeciesMorsePotential):  # composite of this library: Morse fields + species index
        return {"radius": i.radius, "potential_stiffness": i.potential_stiffness, "cutoff": i.cutoff,
                "strength": i.strength, "species": species}
    if isinstance(i, bb.NoInteraction):
        return None
    raise StorageError(f"no serde layout for interaction {type(i).__name__}")


def cell_templates(mechanics, interaction, n_species: int = 1,
                   species_names: Sequence[str] = ("RedCell", "BlueCell", "Species2", "Species3"),
                   mechanics_field: str = "mechanics", interaction_field: str = "interaction",
                   extra: Optional[Dict[str, Any]] = None) -> List[str]:
    """JSON text of the agent struct per species, {pos} / {vel} where the state goes.
    The default field names are those of the benchmarks' and examples' agents
    (`Cell { mechanics, interaction }`, cell_sorting.rs:96-102); `extra` appends
    further fields; the values SPRING_LENGTH and GROWTH_RATE stand for the rod's own
    (e.g. the rods example's `growth_rate: GROWTH_RATE, spring_length_threshold: 6e-6`)."""
    out = []
    for s in range(max(1, n_species)):
        cell = {mechanics_field: mechanics_fields(mechanics),
                interaction_field: interaction_fields(interaction, s, species_names)}
        if extra:
            cell.update(_jsonable(extra))
        text = json.dumps(cell, indent=2).replace("\n", "\n          ")
        for key, hole in _HOLES.items():
            text = text.replace(f'"{key}"', hole)
        out.append(text)
    return out


def fill_template(template: str, pos: np.ndarray, vel: np.ndarray, matrix: bool, spring_length: float = 0.0,
                  growth_rate: float = 0.0) -> Any:
    """The cell object of a template with its state filled in (the Memory option)."""
    def vec(v):
        if not matrix:
            return json.dumps([float(x) for x in v])
        nv, d = v.shape
        return json.dumps([[float(x) for x in v.T.ravel()], nv, d])
    text = template.replace("{pos}", vec(pos)).replace("{vel}", vec(vel))
    text = text.replace("{spring_length}", json.dumps(float(spring_length)))
    return json.loads(text.replace("{growth_rate}", json.dumps(float(growth_rate))))


def identifier_json(row) -> Any:
    """(kind, a, b) -> CellIdentifier as serde writes it; kind 3 -> None."""
    kind, a, b = (int(x) for x in row)
    if kind == 0:
        return {"Initial": a}
    if kind == 1:
        return {"Division": [a, b]}
    if kind == 2:
        return {"Inserted": [a, b]}
    return None


def subdomain_element(domain: bb.CartesianCuboid) -> Dict[str, Any]:
    """CartesianSubDomain of a one-subdomain decomposition (cartesian_cuboid_n.rs:415-423,
    531-580): every voxel, min = domain min, max = min + n dx; voxels in plain-index order."""
    n = [int(x) for x in domain._n]
    dx = [float(x) for x in domain._dx]
    mn = [float(x) for x in domain.min]
    mx = [mn[i] + n[i] * dx[i] for i in range(len(n))]
    voxels = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), axis=-1).reshape(-1, len(n))
    return {"min": mn, "max": mx, "dx": dx, "voxels": voxels.tolist(), "domain_min": mn,
            "domain_max": [float(x) for x in domain.max], "domain_n_voxels": n}
