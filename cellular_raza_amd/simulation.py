"""Host-side mirror of the reference's simulation entry points for the mechanics path.

`MechanicsBackend` is one context of the C ABI (include/crmech.h): it owns the
device buffers of one or more instances and advances them with the gfx950
kernels.  `run_simulation` mirrors chili's run_simulation! for the
[Mechanics, Interaction] aspects (cellular_raza-core-proc-macro/src/run_sim.rs:220-409)
with `FixedStepsize` save points (cellular_raza-core/src/time.rs:93-265).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _abi
from ._native import load
from .building_blocks import CartesianCuboid, instance_struct


class SimulationError(RuntimeError):
    """A non-Ok status of the C ABI; `kind` is the reference error name."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.kind = _abi.STATUS_NAMES.get(status, f"status {status}")
        super().__init__(f"{self.kind}: {message}")


def _dp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_abi.DP)


@dataclass
class InstanceSpec:
    domain: CartesianCuboid
    interaction: object
    mechanics: object
    domain_force: Optional[object] = None  # DomainSpring / RodGelForce (the DomainForce aspect)


class MechanicsBackend:
    """One simulation context on one GPU (ensemble of >= 1 instances)."""

    def __init__(
        self,
        instances: Sequence[InstanceSpec],
        dt: float,
        mechanics_solver_order: int = 2,
        precision: str = "fp64",
        weighting: str = "reference",
        history_order: str = "oldest_first",
        noise: str = "none",
        seed: int = 0,
        device: int = 0,
        graph: bool = True,
        timing: bool = False,
        lds: bool = True,
        autotune=False,
        n_subdomains: int = 1,
    ):
        """n_subdomains: chili's thread count (Settings::n_threads) for the reference
        weighting -- pairs across the borders of its lexicographic voxel chunks get the
        full force (point agents)."""
        self._lib = load()
        if not instances:
            raise ValueError("at least one instance")
        dim = instances[0].domain.dim
        self.dim = dim
        self.n_instances = len(instances)
        self.n_vertices = getattr(instances[0].mechanics, "n_vertices", 1)
        cfg = _abi.Config()
        cfg.dim = dim
        cfg.solver_order = mechanics_solver_order
        cfg.precision = _abi.PRECISION[precision]
        cfg.weighting = _abi.WEIGHTING[weighting]
        cfg.history_order = _abi.HISTORY_ORDER[history_order]
        cfg.noise_mode = _abi.NOISE[noise]
        cfg.seed = seed
        cfg.dt = dt
        cfg.device = device
        cfg.flags = ((_abi.FLAG_GRAPH if graph else 0) | (_abi.FLAG_TIMING if timing else 0)
                     | (0 if lds else _abi.FLAG_NO_LDS))
        cfg.n_subdomains = int(n_subdomains)
        arr = (_abi.Instance * len(instances))()
        for i, spec in enumerate(instances):
            if spec.domain.dim != dim:
                raise ValueError("all instances must have the same dimension")
            arr[i] = instance_struct(spec.domain, spec.interaction, spec.mechanics)
        self._cfg = cfg
        handle = C.c_void_p()
        st = self._lib.crm_create(C.byref(cfg), arr, len(instances), C.byref(handle))
        if st != _abi.CRM_OK:
            raise SimulationError(st, self._message(None))
        self._h = handle
        self.n_agents = 0
        self.counts: List[int] = []
        for i, spec in enumerate(instances):
            if spec.domain_force is not None:
                self.set_domain_force(spec.domain_force, i)
        # False: default tiling; True: tune in memory after each set_agents; a path:
        # also read / update that JSON cache of earlier choices
        self.autotune = autotune
        self.tiling_policy: Optional[int] = None
        self.cycle = None

    # -- helpers -----------------------------------------------------------
    def _message(self, h) -> str:
        buf = C.create_string_buffer(1024)
        self._lib.crm_last_error(h, buf, len(buf))
        return buf.value.decode(errors="replace")

    def _check(self, st: int) -> None:
        if st != _abi.CRM_OK:
            raise SimulationError(st, self._message(self._h))

    @property
    def K(self) -> int:
        return self.n_vertices * self.dim

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.crm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- API ---------------------------------------------------------------
    def set_agents(self, pos, vel=None, species=None, counts=None) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = pos.shape[0]
        pos = pos.reshape(n, self.K)
        if counts is None:
            if self.n_instances != 1:
                raise ValueError("counts is required for several instances")
            counts = [n]
        counts_a = np.ascontiguousarray(counts, dtype=np.uint64)
        if int(counts_a.sum()) != n:
            raise ValueError("sum(counts) must equal the number of agents")
        vel_a = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(n, self.K)
        sp_a = None if species is None else np.ascontiguousarray(species, dtype=np.int32).reshape(n)
        st = self._lib.crm_set_agents(
            self._h,
            counts_a.ctypes.data_as(_abi.U64P),
            _dp(pos),
            _dp(vel_a),
            None if sp_a is None else sp_a.ctypes.data_as(_abi.I32P),
        )
        self._check(st)
        if self.autotune:
            path = None if self.autotune is True else str(self.autotune).encode()
            chosen = C.c_uint32(0)
            self._check(self._lib.crm_autotune(self._h, path, C.byref(chosen)))
            self.tiling_policy = int(chosen.value)
        self.n_agents = n
        self.counts = [int(c) for c in counts_a]

    def step(self, n_steps: int = 1, noise: Optional[np.ndarray] = None) -> None:
        nz = None
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64)
            if nz.size != n_steps * self.n_agents * self.K:
                raise ValueError("noise must hold n_steps * n_agents * n_vertices * dim values")
        st = self._lib.crm_step(self._h, n_steps, _dp(nz))
        if self.cycle is not None:  # divisions add rods
            inf = _abi.Info()
            if self._lib.crm_get_info(self._h, C.byref(inf)) == _abi.CRM_OK:
                self.n_agents = int(inf.n_agents)
        self._check(st)

    def set_rod_cycle(self, cycle) -> None:
        """The rods example's Cycle + NeighborSensing (crm_set_rod_cycle, a
        RodGrowthCycle); call before set_agents, which sizes the context for
        cycle.max_agents rods."""
        self._check(self._lib.crm_set_rod_cycle(self._h, C.byref(cycle.to_struct())))
        self.cycle = cycle

    def set_rod_spring_lengths(self, spring_length) -> None:
        sl = np.ascontiguousarray(spring_length, dtype=np.float64)
        if sl.shape != (self.n_agents,):
            raise ValueError("one spring length per rod")
        self._check(self._lib.crm_set_rod_spring_lengths(self._h, _dp(sl)))

    def cycle_state(self) -> Dict[str, np.ndarray]:
        """Per rod by agent index: spring_length, growth_rate, identifier and parent
        as (kind, a, b) rows (kind 0 Initial(a), 1 Division(a = voxel, b = counter),
        3 no parent)."""
        n = self.n_agents
        sl, gr = np.empty(n), np.empty(n)
        ident, parent = np.empty((n, 3), dtype=np.uint64), np.empty((n, 3), dtype=np.uint64)
        held = np.empty(n, dtype=np.uint32)
        self._check(self._lib.crm_get_cycle_state(self._h, _dp(sl), _dp(gr), ident.ctypes.data_as(_abi.U64P),
                                                  parent.ctypes.data_as(_abi.U64P),
                                                  held.ctypes.data_as(_abi.U32P)))
        return {"spring_length": sl, "growth_rate": gr, "identifier": ident, "parent": parent,
                "history_held": held}

    def identifiers(self) -> Dict[str, np.ndarray]:
        """CellIdentifier and parent of every agent as (kind, a, b) rows (kind 0
        Initial(a), 1 Division(a = voxel, b = counter), 3 no parent) and the history
        entries each holds (crm_get_identifiers); they survive removals."""
        n = self.n_agents
        ident, parent = np.empty((n, 3), dtype=np.uint64), np.empty((n, 3), dtype=np.uint64)
        held = np.empty(n, dtype=np.uint32)
        self._check(self._lib.crm_get_identifiers(self._h, ident.ctypes.data_as(_abi.U64P),
                                                  parent.ctypes.data_as(_abi.U64P),
                                                  held.ctypes.data_as(_abi.U32P)))
        return {"identifier": ident, "parent": parent, "history_held": held}

    def remove_agents(self, remove) -> None:
        """CycleEvent::Remove for the agents with remove[a] true (a boolean per
        agent index; crm_remove_agents): they leave their voxels, the others keep
        their order, state, history and identifiers, and agent indices are
        compacted in order (update_cycle.rs:41-55)."""
        rm = np.ascontiguousarray(remove, dtype=np.uint8)
        if rm.shape != (self.n_agents,):
            raise ValueError("one flag per agent")
        st = self._lib.crm_remove_agents(self._h, rm.ctypes.data_as(_abi.U8P))
        self._check(st)
        self.n_agents -= int(np.count_nonzero(rm))

    def insert_agents(self, mother, daughter_pos, daughter_vel=None, mother_pos=None, mother_vel=None,
                      daughter_species=None) -> None:
        """CycleEvent::Division for point agents (crm_insert_agents,
        update_cycle.rs:28-66): mother[i] takes the state C::divide left it
        (mother_pos / mother_vel, None = unchanged) and a Division identifier;
        daughter i becomes agent n_agents + i with an empty history."""
        m = np.ascontiguousarray(mother, dtype=np.uint64).reshape(-1)
        n = m.shape[0]

        def arr(x):
            return None if x is None else np.ascontiguousarray(x, dtype=np.float64).reshape(n, self.K)

        dp, dv, mp, mv = arr(daughter_pos), arr(daughter_vel), arr(mother_pos), arr(mother_vel)
        sp = None if daughter_species is None else np.ascontiguousarray(daughter_species, dtype=np.int32).reshape(n)
        st = self._lib.crm_insert_agents(self._h, n, m.ctypes.data_as(_abi.U64P), _dp(mp), _dp(mv), _dp(dp),
                                         _dp(dv), None if sp is None else sp.ctypes.data_as(_abi.I32P))
        self._check(st)
        self.n_agents += n

    def compute_forces(self) -> np.ndarray:
        out = np.empty((self.n_agents, self.K), dtype=np.float64)
        self._check(self._lib.crm_compute_forces(self._h, _dp(out)))
        return out

    def get_state(self):
        pos = np.empty((self.n_agents, self.K), dtype=np.float64)
        vel = np.empty((self.n_agents, self.K), dtype=np.float64)
        self._check(self._lib.crm_get_state(self._h, _dp(pos), _dp(vel)))
        return pos, vel

    def get_history(self):
        dx = np.empty((2, self.n_agents, self.K), dtype=np.float64)
        dv = np.empty((2, self.n_agents, self.K), dtype=np.float64)
        n = C.c_uint32()
        self._check(self._lib.crm_get_history(self._h, _dp(dx), _dp(dv), C.byref(n)))
        return dx, dv, int(n.value)

    def info(self) -> _abi.Info:
        inf = _abi.Info()
        self._check(self._lib.crm_get_info(self._h, C.byref(inf)))
        return inf

    def get_voxels(self):
        inf = self.info()
        vox = np.empty(self.n_agents, dtype=np.uint32)
        start = np.empty(int(inf.n_voxels_total) + 1, dtype=np.uint32)
        self._check(
            self._lib.crm_get_voxels(self._h, vox.ctypes.data_as(_abi.U32P), start.ctypes.data_as(_abi.U32P))
        )
        return vox, start

    def voxel_order(self) -> np.ndarray:
        """Agent indices in voxel-list order (crm_get_voxel_order): with the starts
        of get_voxels, the cell list of every voxel."""
        out = np.empty(self.n_agents, dtype=np.uint32)
        self._check(self._lib.crm_get_voxel_order(self._h, out.ctypes.data_as(_abi.U32P)))
        return out

    def tiling(self, inst: int = 0) -> Dict[str, int]:
        """Force tiling of instance `inst`: voxel rows and columns per tile, tile count,
        lanes per agent."""
        r, c, n, lanes = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(self._lib.crm_get_tiling(self._h, inst, C.byref(r), C.byref(c), C.byref(n), C.byref(lanes)))
        return {"rows": int(r.value), "cols": int(c.value), "tiles": int(n.value),
                "lanes_per_agent": int(lanes.value)}

    def set_domain_force(self, force, inst: int = 0) -> None:
        """Domain force of instance `inst` (crm_set_domain_force); None removes it."""
        f = _abi.DomainForce() if force is None else force.to_struct()
        self._check(self._lib.crm_set_domain_force(self._h, inst, C.byref(f)))

    def set_neighbor_sensing(self, radius: Optional[float]) -> None:
        """Count, in every force evaluation, each agent's (point or vertex) pairs
        with the agents of its 3^D voxel stencil closer than `radius` (the
        accumulator of bacterial_rods/src/main.rs:48-78, NeighborSensing in
        concepts interaction.rs:64-79); None switches it off."""
        if radius is None:
            self._check(self._lib.crm_set_neighbor_sensing(self._h, 0, 0.0))
        else:
            self._check(self._lib.crm_set_neighbor_sensing(self._h, 1, float(radius)))

    def neighbor_counts(self) -> np.ndarray:
        """Counts of the last force evaluation by agent index (uint32)."""
        out = np.empty(self.n_agents, dtype=np.uint32)
        self._check(self._lib.crm_get_neighbor_counts(self._h, out.ctypes.data_as(_abi.U32P)))
        return out

    def snapshot(self) -> Dict[str, np.ndarray]:
        """Agent-order snapshot gathered on the device (crm_get_snapshot): pos and vel
        (n_agents, n_vertices * dim), dx / dv history (n_agents, n_hist, n_vertices * dim)
        oldest first, species."""
        nh = C.c_uint32()
        self._check(self._lib.crm_get_snapshot(self._h, None, None, C.byref(nh)))
        n, K, h = self.n_agents, self.K, int(nh.value)
        rec = np.empty((n, K * (2 + 2 * h)), dtype=np.float64)
        sp = np.empty(n, dtype=np.int32)
        self._check(self._lib.crm_get_snapshot(self._h, rec.ctypes.data_as(_abi.DP),
                                               sp.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nh)))
        return {"pos": rec[:, :K], "vel": rec[:, K:2 * K], "dx": rec[:, 2 * K:(2 + h) * K].reshape(n, h, K),
                "dv": rec[:, (2 + h) * K:].reshape(n, h, K), "species": sp}

    def save_cells_json(self, path: str, iteration: int, templates: Sequence[str], instance: int = 0) -> None:
        """Writes the cells in the reference's JSON batch layout (crm_save_cells_json)."""
        enc = [t.encode() for t in templates]
        arr = (C.c_char_p * len(enc))(*enc)
        spec = _abi.JsonCells(C.cast(arr, C.POINTER(C.c_char_p)), len(enc), instance)
        self._check(self._lib.crm_save_cells_json(self._h, str(path).encode(), int(iteration), C.byref(spec)))

    def timing(self) -> Dict[str, Dict[str, float]]:
        t = _abi.Timing()
        self._check(self._lib.crm_get_timing(self._h, C.byref(t)))
        return {
            name: {"total_ms": t.total_ms[i], "launches": int(t.launches[i])}
            for i, name in enumerate(_abi.KERNEL_CLASSES)
        }

    def reset_timing(self) -> None:
        self._check(self._lib.crm_reset_timing(self._h))


class FixedStepsize:
    """FixedStepsize::from_partial_save_steps (cellular_raza-core/src/time.rs:93-111,174-235)."""

    def __init__(self, t0: float, dt: float, save_iterations: List[int], maximum_iterations: int):
        self.t0 = t0
        self.dt = dt
        self.save_iterations = save_iterations
        self.maximum_iterations = maximum_iterations

    @classmethod
    def from_partial_save_steps(cls, t0: float, dt: float, n_steps: int, save_interval: int):
        max_save_points = -(-n_steps // save_interval)
        points = [t0 + float(save_interval * n) * dt for n in range(max_save_points + 1)]
        return cls.from_partial_save_points(t0, dt, points)

    @classmethod
    def from_partial_save_points(cls, t0: float, dt: float, save_points: Sequence[float]):
        pts = sorted(save_points)
        if any(t0 > x for x in pts):
            raise ValueError("TimeError: Evaluation time point is before starting time point.")
        if not pts:
            raise ValueError("TimeError: No savepoints specified.")
        maximum = int(round((pts[-1] - t0) / dt))
        iters = sorted({int(round((t - t0) / dt)) for t in pts})
        return cls(t0, dt, iters, maximum)


@dataclass
class Settings:
    """chili::Settings (cellular_raza-core/src/backend/chili/setup.rs:17-27).  n_threads
    is chili's subdomain count; for point agents it selects the same pair weighting
    (pairs across the borders of the n_threads lexicographic voxel chunks get the full
    force), for rods the single-subdomain weighting is used."""

    time: FixedStepsize
    n_threads: int = 1
    progressbar: Optional[str] = None
    storage: Optional[object] = None  # storage.StorageBuilder: store cells and subdomains


def run_simulation(
    domain: CartesianCuboid,
    interaction,
    mechanics,
    positions,
    settings: Settings,
    velocities=None,
    species=None,
    mechanics_solver_order: int = 2,
    domain_force=None,
    cycle=None,
    spring_lengths=None,
    cell_extra=None,
    remove=None,
    **backend_kwargs,
) -> Dict[int, np.ndarray]:
    """Runs the [Mechanics, Interaction] aspects.  Without settings.storage it returns
    {iteration: positions} at the save iterations (iteration 0 is the initial state,
    run_sim.rs:363-367); with a StorageBuilder it stores the cells and the subdomain
    at those iterations as run_simulation! does (run_sim.rs:375-378,397-398,430-445)
    and returns the StorageAccess.

    remove(backend, iteration) -> mask or None, called after every step, stands for
    the agents' Cycle returning CycleEvent::Remove in that step: the agents of a
    returned boolean mask (by agent index) are removed (MechanicsBackend.remove_agents,
    update_cycle.rs:41-55) before the next step or save point."""
    kw = dict(backend_kwargs)
    if getattr(mechanics, "n_vertices", 1) == 1 and kw.get("weighting", "reference") == "reference":
        kw.setdefault("n_subdomains", settings.n_threads)
    be = MechanicsBackend(
        [InstanceSpec(domain, interaction, mechanics, domain_force)],
        dt=settings.time.dt,
        mechanics_solver_order=mechanics_solver_order,
        **kw,
    )
    try:
        if cycle is not None:  # the aspects Cycle + NeighborSensing of the rods example
            be.set_rod_cycle(cycle)
        be.set_agents(positions, velocities, species)
        if spring_lengths is not None:
            be.set_rod_spring_lengths(spring_lengths)
        out: Dict[int, np.ndarray] = {}
        access = None
        if settings.storage is not None:
            from . import storage as st

            builder = settings.storage if settings.storage.initialized else settings.storage.init()
            base = builder.get_suffix()
            access = st.StorageAccess(
                cells=st.StorageManager.open_or_create(builder.suffix(os.path.join(base, "cells")), 0),
                subdomains=st.StorageManager.open_or_create(builder.suffix(os.path.join(base, "subdomains")), 0),
            )
            n_sp = int(np.max(species)) + 1 if species is not None and len(species) else 1
            templates = st.cell_templates(mechanics, interaction, n_species=n_sp, extra=cell_extra)
            sub = st.subdomain_element(domain)
        it = 0
        for target in settings.time.save_iterations:
            target = min(target, settings.time.maximum_iterations)
            if target > it and remove is None:
                be.step(target - it)
                it = target
            while it < target:  # one step at a time, each followed by the removals
                be.step(1)
                it += 1
                mask = remove(be, it)
                if mask is not None and np.any(mask):
                    be.remove_agents(mask)
            if access is None:
                out[it] = be.get_state()[0]
                continue
            _store_cells(be, access.cells, it, templates)
            access.subdomains.store_single_element(it, 0, sub)
        return out if access is None else access
    finally:
        be.close()


def _store_cells(be: MechanicsBackend, cells, iteration: int, templates: Sequence[str]) -> None:
    """Cells of one save point into every option of the manager: JSON through the
    library's writer, Memory as Python objects of the same structure."""
    from . import storage as st

    for opt, store in cells._stores.items():
        if opt is st.StorageOption.SerdeJson:
            be.save_cells_json(store.path, iteration, templates, cells.instance)
            continue
        snap = be.snapshot()
        nv, dim = be.n_vertices, be.dim
        matrix = nv > 1

        def shaped(x):
            return x.reshape(nv, dim) if matrix else x

        def vec(x):  # SVector: [x, y, z]; dynamic matrix: [column-major data, rows, columns]
            return [x.reshape(nv, dim).T.ravel().tolist(), nv, dim] if matrix else x.tolist()

        zero = vec(np.zeros(be.K))
        cyc = be.cycle_state() if matrix else None
        ids = be.identifiers()
        n_hist = snap["dx"].shape[1]
        held = np.minimum(ids["history_held"].astype(np.int64), n_hist)  # agents born later hold fewer
        pairs = []
        for a in range(be.n_agents):
            t = templates[int(snap["species"][a]) if len(templates) > 1 else 0]
            ident = st.identifier_json(ids["identifier"][a])
            parent = st.identifier_json(ids["parent"][a])
            sl = 0.0 if cyc is None else cyc["spring_length"][a]
            gr = 0.0 if cyc is None else cyc["growth_rate"][a]
            cellbox = {"identifier": ident, "parent": parent,
                       "cell": st.fill_template(t, shaped(snap["pos"][a]), shaped(snap["vel"][a]), matrix, sl, gr)}
            first = n_hist - int(held[a])
            aux = {"mechanics": {"positions": [vec(x) for x in snap["dx"][a][first:]],
                                 "velocities": [vec(x) for x in snap["dv"][a][first:]],
                                 "current_force": zero, "zero_force": zero}}
            pairs.append((st._to_key(ident), [cellbox, aux]))
        store.store_batch_elements(iteration, pairs)
