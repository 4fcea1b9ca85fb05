"""Seeded synthetic populations for the five benchmark configurations.

Every generator takes a scale argument so that the same construction serves the
full-size benchmark and the small parity cases.  Positions are uniform on the
half-open box [0, L) like the reference examples' random_range(0.0..L)
(cellular_raza-examples/cell_sorting/src/main.rs:119-123), nudged off the
domain minimum because the initial placement requires min < p strictly
(cartesian_cuboid_n.rs:149-162,320-337).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .building_blocks import (
    CartesianCuboid,
    CellSpecificInteraction,
    CellSpecificInteractionLangevin,
    Langevin,
    MiePotential,
    MorsePotential,
    NewtonDamped,
    RodInteraction,
    RodMechanics,
    SpeciesMorsePotential,
)
from .simulation import InstanceSpec


@dataclass
class Workload:
    name: str
    instances: List[InstanceSpec]
    dt: float
    steps: int
    solver_order: int
    pos: np.ndarray
    vel: Optional[np.ndarray]
    species: Optional[np.ndarray]
    counts: List[int]
    noise: str = "none"

    @property
    def n_agents(self) -> int:
        return int(sum(self.counts))

    @property
    def dim(self) -> int:
        return self.instances[0].domain.dim


def _uniform_open(rng, n, dim, L):
    p = rng.uniform(0.0, L, size=(n, dim))
    tiny = np.nextafter(0.0, 1.0)
    return np.where(p <= 0.0, tiny, p)


def c1_morse_2d(n: int = 1000, seed: int = 1, steps: int = 2000) -> Workload:
    """2-D damped-Newton point cells, Morse (R 6, lambda 1/1.5, s 0.1, cutoff 12), Euler, dt 0.01.
    Box side scales with sqrt(n) to keep the density of the 1000-agent / 250^2 case."""
    L = 250.0 * math.sqrt(n / 1000.0)
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [L, L], 12.0)
    pot = MorsePotential(radius=6.0, potential_stiffness=1.0 / 1.5, cutoff=12.0, strength=0.1)
    mech = NewtonDamped(damping_constant=1.0, mass=1.0)
    rng = np.random.default_rng(seed)
    pos = _uniform_open(rng, n, 2, L)
    return Workload("C1 2-D Morse NewtonDamped", [InstanceSpec(dom, pot, mech)], 0.01, steps, 0, pos,
                    np.zeros_like(pos), None, [n])


def c2_sorting_3d(n: int = 1_000_000, seed: int = 2, steps: int = 1000) -> Workload:
    """3-D two-species cell sorting, Langevin (gamma 2, kT 0.01, m 1), species-Morse
    (R 1, lambda 1, cutoff 2.5, strength 0.5 x [[1, .2], [.2, 1]]), volume fraction 0.3, AB3."""
    L = (n * 4.0 * math.pi / 3.0 / 0.3) ** (1.0 / 3.0)
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [L] * 3, 2.5)
    pot = SpeciesMorsePotential(radius=1.0, potential_stiffness=1.0, cutoff=2.5, strength=0.5,
                                species_matrix=[[1.0, 0.2], [0.2, 1.0]])
    mech = Langevin(mass=1.0, damping=2.0, kb_temperature=0.01)
    rng = np.random.default_rng(seed)
    pos = _uniform_open(rng, n, 3, L)
    species = (rng.random(n) < 0.5).astype(np.int32)
    return Workload("C2 3-D cell sorting Langevin", [InstanceSpec(dom, pot, mech)], 0.005, steps, 2, pos,
                    np.zeros_like(pos), species, [n], noise="philox")


def straight_rods(rng, n: int, dim: int, L: float, n_vertices: int, spacing: float) -> np.ndarray:
    """n straight rods of n_vertices vertices `spacing` apart: centres uniform in
    [margin, L - margin)^dim (margin = half length + 0.25), directions uniform
    (angle in 2-D, unit sphere in 3-D).  Returns (n, n_vertices, dim)."""
    half = 0.5 * spacing * (n_vertices - 1)
    margin = half + 0.25
    centre = rng.uniform(margin, L - margin, size=(n, dim))
    if dim == 2:
        ang = rng.uniform(0.0, 2.0 * math.pi, n)
        u = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    else:
        u = rng.normal(size=(n, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
    off = (np.arange(n_vertices) - 0.5 * (n_vertices - 1)) * spacing
    return centre[:, None, :] + off[None, :, None] * u[:, None, :]


def c3_rods_2d(n: int = 200_000, seed: int = 3, steps: int = 500, n_vertices: int = 8,
               diffusion_constant: float = 0.0) -> Workload:
    """2-D bacterial rods: n rods of 8 vertices 0.5 apart (length 3.5), RodMechanics
    (spring tension 10, rigidity 2, spring length 0.5, damping 1.5), RodInteraction of
    Morse (R 0.5, lambda 2, strength 0.1, cutoff 1.5), dt 0.002, AB3.  Box side
    2000 * sqrt(n / 200000), voxel edge 5.0 (>= rod length + cutoff, centroid binning).
    Morse stiffness and strength are not specified for this workload; lambda 2 / s 0.1
    give a soft contact."""
    L = 2000.0 * math.sqrt(n / 200_000.0)
    L = max(L, 20.0)
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [L, L], 5.0)
    pot = RodInteraction(MorsePotential(radius=0.5, potential_stiffness=2.0, cutoff=1.5, strength=0.1))
    mech = RodMechanics(n_vertices=n_vertices, diffusion_constant=diffusion_constant, spring_tension=10.0,
                        rigidity=2.0, spring_length=0.5, damping=1.5)
    rng = np.random.default_rng(seed)
    pos = straight_rods(rng, n, 2, L, n_vertices, 0.5)
    return Workload("C3 2-D bacterial rods", [InstanceSpec(dom, pot, mech)], 0.002, steps, 2, pos,
                    np.zeros_like(pos), None, [n], noise="philox" if diffusion_constant > 0 else "none")


@dataclass
class RodsExample:
    """The bacterial rods example: workload, per-rod spring lengths and the cycle."""

    workload: Workload
    spring_length: np.ndarray
    cycle: object


def bacterial_rods_example(seed: int = 5, steps: int = 15000, n_rods: int = 5,
                           max_agents: int = 1 << 16) -> RodsExample:
    """cellular_raza-examples/bacterial_rods/src/main.rs:113-209 in SI units: 5 rods of
    5 vertices, RodMechanics (spring tension 10 /min^2, rigidity 1 um/min^2, damping
    1.5 /min, spring length 3 um), RodInteraction(Morse R 3 um, lambda 0.5 /um, strength
    0.1 um^2/min^2, cutoff 5 um), growth 0.1 um/min slowed by neighbours within 2 R
    ((8 - n) / 8), division at a spring length of 6 um; each rod's spring length
    U[1.5, 2.5] um, first vertex U[dx, 2 dx] x U[dx/3, 2 dx/3] x U[dx, 2 dx] (dx = 15
    um), later vertices 3 um apart at angles theta + U[-pi/8, pi/8] in the x-y plane;
    domain [0, 200] x [0, 15] x [0, 45] um in 8 x 1 x 2 voxels, no gel force; dt 0.1
    min, 25 h (15000 steps), AB3.  The example seeds ChaCha8; this draws the same
    distributions from numpy's generator."""
    from .building_blocks import RodGrowthCycle

    um, minute = 1e-6, 1.0
    nv, sl0 = 5, 3.0 * um
    dxr = sl0 * nv
    rng = np.random.default_rng(seed)
    spring = rng.uniform(1.5, 2.5, n_rods) * um
    pos = np.zeros((n_rods, nv, 3))
    for a in range(n_rods):
        pos[a, 0] = [rng.uniform(dxr, 2 * dxr), rng.uniform(dxr / 3, 2 * dxr / 3), rng.uniform(dxr, 2 * dxr)]
        theta = rng.uniform(0.0, 2 * math.pi)
        for i in range(1, nv):
            phi = theta + rng.uniform(-math.pi / 8, math.pi / 8)
            pos[a, i] = pos[a, i - 1] + sl0 * np.array([math.cos(phi), math.sin(phi), 0.0])
    dom = CartesianCuboid.from_boundaries_and_n_voxels([0.0] * 3, [4 * 50.0 * um, dxr, 3 * dxr], [8, 1, 2])
    pot = RodInteraction(MorsePotential(radius=3.0 * um, potential_stiffness=0.5 / um, cutoff=5.0 * um,
                                        strength=0.1 * um ** 2 / minute ** 2))
    mech = RodMechanics(n_vertices=nv, diffusion_constant=0.0, spring_tension=10.0 / minute ** 2,
                        rigidity=1.0 * um / minute ** 2, spring_length=sl0, damping=1.5 / minute)
    cycle = RodGrowthCycle(growth_base_rate=0.1 * um / minute, division_threshold=6.0 * um,
                           division_radius=3.0 * um, sensing_radius=2.0 * 3.0 * um, neighbor_saturation=8.0,
                           max_agents=max_agents)
    wl = Workload("bacterial rods example (Cycle + NeighborSensing)", [InstanceSpec(dom, pot, mech)],
                  0.1 * minute, steps, 2, pos, np.zeros_like(pos), None, [n_rods])
    return RodsExample(wl, spring, cycle)


def langevin_sorting_example(seed: int = 6, steps: int = 50_000, n_blue: int = 600, n_red: int = 1000,
                             domain_size: float = 200.0) -> Workload:
    """cellular_raza-examples/cell_sorting/src/main-langevin.rs:9-146: 1600 cells
    (indices 0..n_blue BlueCell -- the example's `n <= N_CELLS_BLUE` makes n_blue + 1
    of them -- the rest RedCell) uniform in [0, 200)^3, Langevin3D (mass 10, damping 2,
    kT 1), CellSpecificInteraction (cell radius 6, attraction rr 1.5, bb 0.075, rb
    0.015, repulsion 5, cutoff 20), domain interaction range 2 x cutoff, dt 0.25,
    50000 steps.  Species 0 = RedCell, 1 = BlueCell.  The example seeds ChaCha8;
    this draws the same distributions from numpy's generator; the solver order is the
    run_simulation! default (AB3)."""
    n = n_blue + n_red
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [domain_size] * 3, 2.0 * 20.0)
    pot = CellSpecificInteractionLangevin(cell_radius=6.0, attraction_rr=1.5, attraction_bb=0.05 * 1.5,
                                          attraction_rb=0.01 * 1.5, repulsion=5.0, cutoff=20.0)
    mech = Langevin(mass=10.0, damping=2.0, kb_temperature=1.0)
    rng = np.random.default_rng(seed)
    pos = _uniform_open(rng, n, 3, domain_size)
    species = np.where(np.arange(n) <= n_blue, 1, 0).astype(np.int32)
    return Workload("Langevin cell sorting example", [InstanceSpec(dom, pot, mech)], 0.25, steps, 2, pos,
                    np.zeros_like(pos), species, [n], noise="philox")


def cell_sorting_benchmark(size_index: int = 8, seed: int = 7, steps: int = 10) -> Workload:
    """The reference's own CPU benchmark (cellular_raza-benchmarks/src/cell_sorting.rs:
    116-177, sim-size sweep 280-313): 2 x 10 * 4^k cells (indices 0..n1 BlueCell --
    `n <= n_cells_1` makes n1 + 1 of them -- the rest RedCell) uniform in [0, L)^3 with
    L = round(360 * 4^(k/3)) / 10, NewtonDamped3D (damping 2, mass 1), the benchmark's
    CellSpecificInteraction (cell radius 6, strength 2, relative range 1.5), voxels from
    the interaction range 18, dt 0.1, the default solver (AB3).  k = 8: 1,310,720 cells
    in L = 1451.4, the largest size published for all three machines.  The benchmark
    seeds ChaCha8; this draws the same distributions from numpy's generator."""
    n1 = 10 * 4 ** size_index
    n = 2 * n1
    L = round(360.0 * 4.0 ** (size_index / 3.0)) * 0.1
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [L] * 3, 1.5 * 6.0 * 2.0)
    pot = CellSpecificInteraction(cell_radius=6.0, potential_strength=2.0, relative_interaction_range=1.5,
                                  variant="benchmark")
    mech = NewtonDamped(damping_constant=2.0, mass=1.0)
    rng = np.random.default_rng(seed)
    pos = _uniform_open(rng, n, 3, L)
    species = np.where(np.arange(n) <= n1, 1, 0).astype(np.int32)
    return Workload("reference cell-sorting benchmark", [InstanceSpec(dom, pot, mech)], 0.1, steps, 2, pos,
                    np.zeros_like(pos), species, [n])


def c4_ensemble_2d(n_instances: int = 512, seed: int = 4, steps: int = 1000, n_min: int = 500,
                   n_max: int = 20000) -> Workload:
    """Ensemble of 2-D Morse runs: N_i log-uniform in [n_min, n_max], strength U[0.05, 0.5],
    damping U[0.5, 2.0], box side sqrt(N_i / 0.02); Morse R 2, lambda 1/1.5, cutoff 12; Euler.
    The radius is not specified for this ensemble; R = 2 keeps the area fraction at 0.25
    (R = 6 at density 0.02 would overlap 2.3-fold and throw agents out of the box)."""
    rng = np.random.default_rng(seed)
    counts = np.round(np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_instances))).astype(int)
    strengths = rng.uniform(0.05, 0.5, n_instances)
    dampings = rng.uniform(0.5, 2.0, n_instances)
    inst, pos = [], []
    for i in range(n_instances):
        L = math.sqrt(counts[i] / 0.02)
        dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [L, L], 12.0)
        pot = MorsePotential(radius=2.0, potential_stiffness=1.0 / 1.5, cutoff=12.0, strength=float(strengths[i]))
        mech = NewtonDamped(damping_constant=float(dampings[i]), mass=1.0)
        inst.append(InstanceSpec(dom, pot, mech))
        pos.append(_uniform_open(rng, int(counts[i]), 2, L))
    P = np.concatenate(pos)
    return Workload("C4 2-D Morse ensemble", inst, 0.01, steps, 0, P, np.zeros_like(P), None,
                    [int(c) for c in counts])


def c5_mie_3d(n: int = 4_000_000, seed: int = 5, steps: int = 500, spacing: float = 2.0,
              bound: float = 1000.0) -> Workload:
    """3-D Mie (n 12, m 6, R 1, eps 1, cutoff 2.5) on a jittered cubic lattice (jitter
    U[-0.2a, 0.2a]), NewtonDamped (gamma 1, m 1), dt 0.001, AB3."""
    side = int(math.ceil(n ** (1.0 / 3.0)))
    a = spacing
    idx = np.arange(n)
    ijk = np.stack([idx // (side * side), (idx // side) % side, idx % side], axis=1).astype(np.float64)
    rng = np.random.default_rng(seed)
    pos = (ijk + 0.5) * a + rng.uniform(-0.2 * a, 0.2 * a, size=(n, 3))
    L = side * a
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [L] * 3, 2.5)
    pot = MiePotential(radius=1.0, strength=1.0, bound=bound, cutoff=2.5, en=12.0, em=6.0)
    mech = NewtonDamped(damping_constant=1.0, mass=1.0)
    return Workload("C5 3-D Mie lattice", [InstanceSpec(dom, pot, mech)], 0.001, steps, 2, pos,
                    np.zeros_like(pos), None, [n])
