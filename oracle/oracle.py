"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / CPU baseline.  The product path
(cellular_raza_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from cellular_raza_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "-B"], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    DP = _abi.DP
    sig = {
        "orc_grid": (C.c_int, [C.POINTER(_abi.Domain), C.c_uint32, _abi.U64P, DP]),
        "orc_voxel_index": (C.c_int, [C.POINTER(_abi.Domain), C.c_uint32, DP, C.c_int, _abi.U64P]),
        "orc_neighbors": (C.c_uint32, [_abi.U64P, C.c_uint32, C.c_uint64, _abi.U64P]),
        "orc_decomp_res": (C.c_int, [C.c_uint64, C.c_uint64, _abi.U64P]),
        "orc_pair_force": (C.c_int, [C.POINTER(_abi.Potential), C.c_uint32, DP, DP, C.c_int32, C.c_int32, DP, DP]),
        "orc_rod_pair_force": (C.c_int, [C.POINTER(_abi.Potential), C.c_uint32, C.c_uint32, DP, DP, DP, DP]),
        "orc_nearest_point": (None, [C.c_uint32, DP, DP, DP, DP, DP, DP]),
        "orc_rod_increment": (None, [C.POINTER(_abi.Mechanics), C.c_uint32, C.c_uint32, DP, DP, DP, DP, DP]),
        "orc_philox4x32_10": (None, [_abi.U32P, _abi.U32P, _abi.U32P]),
        "orc_gaussians": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, DP]),
        "orc_create": (C.c_int, [C.POINTER(_abi.Config), C.POINTER(_abi.Instance), C.c_uint32, C.c_uint32,
                                 C.POINTER(P)]),
        "orc_set_agents": (C.c_int, [P, _abi.U64P, DP, DP, _abi.I32P]),
        "orc_step": (C.c_int, [P, C.c_uint32, DP]),
        "orc_step_timed": (C.c_int, [P, C.c_uint32, DP]),
        "orc_compute_forces": (C.c_int, [P, DP]),
        "orc_get_state": (C.c_int, [P, DP, DP]),
        "orc_get_history": (C.c_int, [P, DP, DP, _abi.U32P]),
        "orc_get_voxels": (C.c_int, [P, _abi.U32P, _abi.U32P, _abi.U64P]),
        "orc_neighbor_counts": (C.c_int, [P, C.c_double, _abi.U32P]),
        "orc_set_domain_force": (C.c_int, [P, C.c_uint32, C.POINTER(_abi.DomainForce)]),
        "orc_set_rod_cycle": (C.c_int, [P, C.POINTER(_abi.RodCycle)]),
        "orc_set_spring_lengths": (C.c_int, [P, _abi.DP]),
        "orc_cycle_out": (C.c_int, [P, _abi.DP, _abi.DP, _abi.U64P, _abi.U64P]),
        "orc_remove_agents": (C.c_int, [P, _abi.U8P]),
        "orc_insert_agents": (C.c_int, [P, C.c_uint64, _abi.U64P, _abi.DP, _abi.DP, _abi.DP, _abi.DP, _abi.I32P]),
        "orc_n_voxels_total": (C.c_uint64, [P]),
        "orc_n_agents": (C.c_uint64, [P]),
        "orc_last_error": (C.c_size_t, [P, C.c_char_p, C.c_size_t]),
        "orc_destroy": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_abi.DP)


class OracleError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        self.kind = _abi.STATUS_NAMES.get(status, str(status))
        super().__init__(f"{self.kind}: {message}")


class OracleSim:
    """Same surface as cellular_raza_amd.MechanicsBackend, on the CPU restatement.

    n_subdomains = 1 is the parity mode (reference tests run with n_threads: 1);
    n_subdomains > 1 runs chili's thread-per-subdomain structure for timing.
    """

    def __init__(self, instances, dt, mechanics_solver_order=2, weighting="reference",
                 history_order="oldest_first", noise="none", seed=0, n_subdomains=1):
        from cellular_raza_amd.building_blocks import instance_struct

        self._lib = load()
        self.dim = instances[0].domain.dim
        self.n_vertices = getattr(instances[0].mechanics, "n_vertices", 1)
        self.n_instances = len(instances)
        cfg = _abi.Config()
        cfg.dim = self.dim
        cfg.solver_order = mechanics_solver_order
        cfg.precision = 0
        cfg.weighting = _abi.WEIGHTING[weighting]
        cfg.history_order = _abi.HISTORY_ORDER[history_order]
        cfg.noise_mode = _abi.NOISE[noise]
        cfg.seed = seed
        cfg.dt = dt
        arr = (_abi.Instance * len(instances))()
        for i, spec in enumerate(instances):
            arr[i] = instance_struct(spec.domain, spec.interaction, spec.mechanics)
        h = C.c_void_p()
        st = self._lib.orc_create(C.byref(cfg), arr, len(instances), n_subdomains, C.byref(h))
        if st:
            raise OracleError(st, "orc_create failed")
        self._h = h
        self.n_agents = 0
        for i, spec in enumerate(instances):
            if getattr(spec, "domain_force", None) is not None:
                self.set_domain_force(spec.domain_force, i)

    def set_domain_force(self, force, inst=0):
        f = _abi.DomainForce() if force is None else force.to_struct()
        self._check(self._lib.orc_set_domain_force(self._h, inst, C.byref(f)))

    @property
    def K(self):
        return self.n_vertices * self.dim

    def _check(self, st):
        if st:
            buf = C.create_string_buffer(1024)
            self._lib.orc_last_error(self._h, buf, len(buf))
            raise OracleError(st, buf.value.decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_agents(self, pos, vel=None, species=None, counts=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = pos.shape[0]
        pos = pos.reshape(n, self.K)
        counts = np.ascontiguousarray([n] if counts is None else counts, dtype=np.uint64)
        vel_a = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(n, self.K)
        sp = None if species is None else np.ascontiguousarray(species, dtype=np.int32)
        self._check(self._lib.orc_set_agents(self._h, counts.ctypes.data_as(_abi.U64P), _dp(pos), _dp(vel_a),
                                             None if sp is None else sp.ctypes.data_as(_abi.I32P)))
        self.n_agents = n

    def step(self, n_steps=1, noise=None):
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
        st = self._lib.orc_step(self._h, n_steps, _dp(nz))
        self.n_agents = int(self._lib.orc_n_agents(self._h))  # division adds agents
        self._check(st)

    def set_rod_cycle(self, cycle):
        self._check(self._lib.orc_set_rod_cycle(self._h, C.byref(cycle.to_struct())))

    def set_spring_lengths(self, spring_length):
        sl = np.ascontiguousarray(spring_length, dtype=np.float64)
        self._check(self._lib.orc_set_spring_lengths(self._h, _dp(sl)))

    def remove_agents(self, remove):
        rm = np.ascontiguousarray(remove, dtype=np.uint8)
        self._check(self._lib.orc_remove_agents(self._h, rm.ctypes.data_as(_abi.U8P)))
        self.n_agents = int(self._lib.orc_n_agents(self._h))

    def insert_agents(self, mother, daughter_pos, daughter_vel=None, mother_pos=None, mother_vel=None,
                      daughter_species=None):
        m = np.ascontiguousarray(mother, dtype=np.uint64).reshape(-1)
        n = m.shape[0]

        def arr(x):
            return None if x is None else np.ascontiguousarray(x, dtype=np.float64).reshape(n, self.K)

        dp, dv, mp, mv = arr(daughter_pos), arr(daughter_vel), arr(mother_pos), arr(mother_vel)
        sp = None if daughter_species is None else np.ascontiguousarray(daughter_species, dtype=np.int32).reshape(n)
        self._check(self._lib.orc_insert_agents(self._h, n, m.ctypes.data_as(_abi.U64P), _dp(mp), _dp(mv), _dp(dp),
                                                _dp(dv), None if sp is None else sp.ctypes.data_as(_abi.I32P)))
        self.n_agents += n

    def cycle_state(self):
        n = self.n_agents
        sl, gr = np.empty(n), np.empty(n)
        ident, parent = np.empty((n, 3), dtype=np.uint64), np.empty((n, 3), dtype=np.uint64)
        self._check(self._lib.orc_cycle_out(self._h, _dp(sl), _dp(gr), ident.ctypes.data_as(_abi.U64P),
                                            parent.ctypes.data_as(_abi.U64P)))
        return {"spring_length": sl, "growth_rate": gr, "identifier": ident, "parent": parent}

    def step_timed(self, n_steps):
        sec = C.c_double()
        st = self._lib.orc_step_timed(self._h, n_steps, C.byref(sec))
        self.n_agents = int(self._lib.orc_n_agents(self._h))
        self._check(st)
        return sec.value

    def compute_forces(self):
        out = np.empty((self.n_agents, self.K))
        self._check(self._lib.orc_compute_forces(self._h, _dp(out)))
        return out

    def get_state(self):
        pos = np.empty((self.n_agents, self.K))
        vel = np.empty((self.n_agents, self.K))
        self._check(self._lib.orc_get_state(self._h, _dp(pos), _dp(vel)))
        return pos, vel

    def get_history(self):
        dx = np.empty((2, self.n_agents, self.K))
        dv = np.empty((2, self.n_agents, self.K))
        n = C.c_uint32()
        self._check(self._lib.orc_get_history(self._h, _dp(dx), _dp(dv), C.byref(n)))
        return dx, dv, int(n.value)

    def get_voxels(self):
        vt = int(self._lib.orc_n_voxels_total(self._h))
        vox = np.empty(self.n_agents, dtype=np.uint32)
        start = np.empty(vt + 1, dtype=np.uint32)
        order = np.empty(self.n_agents, dtype=np.uint64)
        self._check(self._lib.orc_get_voxels(self._h, vox.ctypes.data_as(_abi.U32P),
                                             start.ctypes.data_as(_abi.U32P), order.ctypes.data_as(_abi.U64P)))
        return vox, start, order

    def neighbor_counts(self, radius):
        """Stencil neighbour count per agent: (point or vertex) pairs closer than radius."""
        out = np.empty(self.n_agents, dtype=np.uint32)
        self._check(self._lib.orc_neighbor_counts(self._h, float(radius), out.ctypes.data_as(_abi.U32P)))
        return out


# -- pure helpers ---------------------------------------------------------
def grid(domain, dim):
    n = (C.c_uint64 * 3)()
    dx = (C.c_double * 3)()
    st = load().orc_grid(C.byref(domain.to_struct()), dim, n, dx)
    if st:
        raise OracleError(st, "grid")
    return list(n)[:dim], list(dx)[:dim]


def voxel_index(domain, p, initial=False):
    p = np.ascontiguousarray(p, dtype=np.float64)
    idx = (C.c_uint64 * 3)()
    st = load().orc_voxel_index(C.byref(domain.to_struct()), len(p), _dp(p), int(initial), idx)
    return st, list(idx)[:len(p)]


def neighbors(n_voxels, plain):
    n = (C.c_uint64 * 3)(*(list(n_voxels) + [1] * (3 - len(n_voxels))))
    out = (C.c_uint64 * 27)()
    k = load().orc_neighbors(n, len(n_voxels), plain, out)
    return list(out)[:k]


def decomp_res(n_voxel, n_regions):
    out = (C.c_uint64 * 3)()
    st = load().orc_decomp_res(n_voxel, n_regions, out)
    return None if st else tuple(out)


def pair_force(potential, own, ext, sp_own=0, sp_ext=0):
    """(status, force on own, force on ext) of one Interaction::calculate_force_between call."""
    p = _abi.Potential()
    potential.fill(p)
    own = np.ascontiguousarray(own, dtype=np.float64)
    ext = np.ascontiguousarray(ext, dtype=np.float64)
    fo = np.empty_like(own)
    fe = np.empty_like(own)
    st = load().orc_pair_force(C.byref(p), len(own), _dp(own), _dp(ext), sp_own, sp_ext, _dp(fo), _dp(fe))
    return st, fo, fe


def rod_pair_force(potential, own, ext):
    p = _abi.Potential()
    potential.fill(p)
    own = np.ascontiguousarray(own, dtype=np.float64)
    ext = np.ascontiguousarray(ext, dtype=np.float64)
    nv, dim = own.shape
    fo = np.empty_like(own)
    fe = np.empty_like(own)
    st = load().orc_rod_pair_force(C.byref(p), dim, nv, _dp(own), _dp(ext), _dp(fo), _dp(fe))
    return st, fo, fe


def nearest_point(p, a, b):
    p, a, b = (np.ascontiguousarray(x, dtype=np.float64) for x in (p, a, b))
    d = C.c_double()
    t = C.c_double()
    x = np.empty_like(p)
    load().orc_nearest_point(len(p), _dp(p), _dp(a), _dp(b), C.byref(d), _dp(x), C.byref(t))
    return d.value, x, t.value


def rod_increment(mechanics, pos, vel, force):
    m = _abi.Mechanics()
    mechanics.fill(m)
    pos, vel, force = (np.ascontiguousarray(x, dtype=np.float64) for x in (pos, vel, force))
    nv, dim = pos.shape
    dx = np.empty_like(pos)
    dv = np.empty_like(pos)
    load().orc_rod_increment(C.byref(m), dim, nv, _dp(pos), _dp(vel), _dp(force), _dp(dx), _dp(dv))
    return dx, dv


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    out = (C.c_uint32 * 4)()
    load().orc_philox4x32_10(c, k, out)
    return list(out)


def gaussians(seed, agent, step, n):
    out = np.empty(n)
    load().orc_gaussians(seed, agent, step, n, _dp(out))
    return out
