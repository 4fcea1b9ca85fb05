"""Building blocks with the reference's names and fields.

Each class mirrors a cellular_raza building block (paths relative to the
cellular_raza workspace) and knows how to fill the C-ABI structs of
include/crmech.h.  Positions and velocities are not stored in these objects:
agents are passed as arrays (one row per agent), which is how a GPU backend
receives them.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Sequence

from . import _abi


# --------------------------------------------------------------------------
# Domain
# --------------------------------------------------------------------------
class CartesianCuboid:
    """CartesianCuboid (cellular_raza-building-blocks/src/domains/cartesian_cuboid_n.rs:65-256)."""

    def __init__(self, min, max, interaction_range=None, n_voxels=None):
        self.min = [float(x) for x in min]
        self.max = [float(x) for x in max]
        if len(self.min) != len(self.max) or len(self.min) not in (2, 3):
            raise ValueError("min and max must both have 2 or 3 components")
        self.interaction_range = None if interaction_range is None else float(interaction_range)
        self.n_voxels_requested = None if n_voxels is None else [int(n) for n in n_voxels]
        if (self.interaction_range is None) == (self.n_voxels_requested is None):
            raise ValueError("give exactly one of interaction_range or n_voxels")
        for a, b in zip(self.min, self.max):
            if a >= b:
                raise ValueError(f"BoundaryError: Min {self.min} must be smaller than Max {self.max}")
        self._dx, self._n = self._grid()

    @classmethod
    def from_boundaries_and_interaction_range(cls, min, max, interaction_range):
        return cls(min, max, interaction_range=interaction_range)

    @classmethod
    def from_boundaries_and_n_voxels(cls, min, max, n_voxels):
        return cls(min, max, n_voxels=n_voxels)

    @property
    def dim(self) -> int:
        return len(self.min)

    def _grid(self):
        # cartesian_cuboid_n.rs:203-256: n = max(floor(L / range), 1), dx = L / floor(L / range)
        dx, n = [], []
        for a, b in zip(self.min, self.max):
            if self.interaction_range is not None:
                nf = math.floor((b - a) / self.interaction_range)
                n.append(max(int(nf), 1))
                dx.append((b - a) / nf if nf != 0 else math.inf)
            else:
                k = self.n_voxels_requested[len(n)]
                n.append(k)
                dx.append((b - a) / k)
        return dx, n

    def get_min(self):
        return list(self.min)

    def get_max(self):
        return list(self.max)

    def get_dx(self):
        return list(self._dx)

    def get_n_voxels(self):
        return list(self._n)

    def to_struct(self) -> _abi.Domain:
        d = _abi.Domain()
        for i in range(self.dim):
            d.min[i] = self.min[i]
            d.max[i] = self.max[i]
        if self.interaction_range is not None:
            d.interaction_range = self.interaction_range
        else:
            for i in range(self.dim):
                d.n_voxels[i] = self.n_voxels_requested[i]
        return d


# --------------------------------------------------------------------------
# Interactions
# --------------------------------------------------------------------------
@dataclass
class MorsePotential:
    """MorsePotential (cell_building_blocks/interaction.rs:184-299)."""

    radius: float
    potential_stiffness: float
    cutoff: float
    strength: float

    def fill(self, p: _abi.Potential) -> None:
        p.kind = _abi.POTENTIAL_MORSE
        p.radius = self.radius
        p.potential_stiffness = self.potential_stiffness
        p.cutoff = self.cutoff
        p.strength = self.strength

    symmetric = True


@dataclass
class MiePotential:
    """MiePotential (interaction.rs:301-434)."""

    radius: float
    strength: float
    bound: float
    cutoff: float
    en: float
    em: float

    def fill(self, p: _abi.Potential) -> None:
        p.kind = _abi.POTENTIAL_MIE
        p.radius = self.radius
        p.strength = self.strength
        p.bound = self.bound
        p.cutoff = self.cutoff
        p.en = self.en
        p.em = self.em


@dataclass
class BoundLennardJones:
    """BoundLennardJones (interaction.rs:84-147)."""

    epsilon: float
    sigma: float
    bound: float
    cutoff: float

    def fill(self, p: _abi.Potential) -> None:
        p.kind = _abi.POTENTIAL_BOUND_LENNARD_JONES
        p.epsilon = self.epsilon
        p.sigma = self.sigma
        p.bound = self.bound
        p.cutoff = self.cutoff


@dataclass
class SpeciesMorsePotential:
    """Morse potential whose strength is scaled by species_matrix[own][ext].

    Not a reference type: the two-species cell-sorting workload composes the
    reference MorsePotential (interaction.rs:151-182) with a species matrix.
    """

    radius: float
    potential_stiffness: float
    cutoff: float
    strength: float
    species_matrix: Sequence[Sequence[float]] = field(default_factory=lambda: [[1.0]])

    def fill(self, p: _abi.Potential) -> None:
        p.kind = _abi.POTENTIAL_MORSE_SPECIES
        p.radius = self.radius
        p.potential_stiffness = self.potential_stiffness
        p.cutoff = self.cutoff
        p.strength = self.strength
        n = len(self.species_matrix)
        if n < 1 or n > _abi.CRM_MAX_SPECIES:
            raise ValueError("species_matrix must be 1x1 .. 4x4")
        p.n_species = n
        for a in range(n):
            for b in range(n):
                p.species_matrix[a * _abi.CRM_MAX_SPECIES + b] = float(self.species_matrix[a][b])


@dataclass
class CellSpecificInteraction:
    """CellSpecificInteraction of the cell-sorting benchmark / example.

    variant="benchmark": cellular_raza-benchmarks/src/cell_sorting.rs:31-94
    variant="example":   cellular_raza-examples/cell_sorting/src/main.rs:32-103
    (the example returns the opposite sign).  Species are integers 0..n_species-1.
    """

    cell_radius: float
    potential_strength: float
    relative_interaction_range: float
    n_species: int = 2
    variant: str = "benchmark"

    def fill(self, p: _abi.Potential) -> None:
        p.kind = _abi.POTENTIAL_CELL_SORTING
        p.radius = self.cell_radius
        p.strength = self.potential_strength
        p.relative_interaction_range = self.relative_interaction_range
        p.n_species = self.n_species
        if self.variant not in ("benchmark", "example"):
            raise ValueError("variant must be 'benchmark' or 'example'")
        p.sign = 1.0 if self.variant == "benchmark" else -1.0


@dataclass
class CellSpecificInteractionLangevin:
    """CellSpecificInteraction of the Langevin cell-sorting example
    (cellular_raza-examples/cell_sorting/src/main-langevin.rs:35-92).

    R = 2 cell_radius, r = |own - ext|, dir = (own - ext) / r (0 at r = 0):
    force = repulsion (R - r) dir for r <= R, -a (r - R)(cutoff - r) / cutoff
    exp(-(r / R - 1)^2) dir for R < r <= cutoff, else 0, where a is the species
    pair's attraction; the interaction returns (-force, force).  Species 0 is
    RedCell and 1 BlueCell (the example's enum order), so the attraction matrix is
    [[attraction_rr, attraction_rb], [attraction_rb, attraction_bb]];
    `attraction_matrix` (row = own species, up to 4 x 4) replaces it when given.
    """

    cell_radius: float
    attraction_rr: float
    attraction_bb: float
    attraction_rb: float
    repulsion: float
    cutoff: float
    attraction_matrix: Sequence[Sequence[float]] = None

    def matrix(self):
        if self.attraction_matrix is not None:
            return [[float(x) for x in row] for row in self.attraction_matrix]
        return [[self.attraction_rr, self.attraction_rb], [self.attraction_rb, self.attraction_bb]]

    def fill(self, p: _abi.Potential) -> None:
        p.kind = _abi.POTENTIAL_CELL_SORTING_LANGEVIN
        p.radius = self.cell_radius
        p.strength = self.repulsion
        p.cutoff = self.cutoff
        m = self.matrix()
        n = len(m)
        if n < 1 or n > _abi.CRM_MAX_SPECIES or any(len(row) != n for row in m):
            raise ValueError("the attraction matrix must be square, 1x1 .. 4x4")
        p.n_species = n
        for a in range(n):
            for b in range(n):
                p.species_matrix[a * _abi.CRM_MAX_SPECIES + b] = float(m[a][b])


@dataclass
class NoInteraction:
    """NoInteraction (interaction.rs:12-35): the Mechanics-only aspect set."""

    def fill(self, p: _abi.Potential) -> None:
        p.kind = _abi.POTENTIAL_NONE


@dataclass
class RodInteraction:
    """RodInteraction(I) (cell_building_blocks/bacterial_rods.rs:267-337)."""

    interaction: object

    def fill(self, p: _abi.Potential) -> None:
        self.interaction.fill(p)


# --------------------------------------------------------------------------
# Mechanics
# --------------------------------------------------------------------------
@dataclass
class NewtonDamped:
    """NewtonDamped{2,3}D (cell_building_blocks/mechanics.rs:14-204)."""

    damping_constant: float
    mass: float = 1.0

    def fill(self, m: _abi.Mechanics) -> None:
        m.kind = _abi.MECHANICS_NEWTON_DAMPED
        m.n_vertices = 1
        m.damping = self.damping_constant
        m.mass = self.mass


@dataclass
class Langevin:
    """Langevin{2,3}D (mechanics.rs:389-571)."""

    mass: float
    damping: float
    kb_temperature: float

    def fill(self, m: _abi.Mechanics) -> None:
        m.kind = _abi.MECHANICS_LANGEVIN
        m.n_vertices = 1
        m.mass = self.mass
        m.damping = self.damping
        m.kb_temperature = self.kb_temperature


@dataclass
class Brownian:
    """Brownian{2,3}D (mechanics.rs:228-387)."""

    diffusion_constant: float
    kb_temperature: float

    def fill(self, m: _abi.Mechanics) -> None:
        m.kind = _abi.MECHANICS_BROWNIAN
        m.n_vertices = 1
        m.diffusion_constant = self.diffusion_constant
        m.kb_temperature = self.kb_temperature


@dataclass
class RodMechanics:
    """RodMechanics (bacterial_rods.rs:82-265); n_vertices vertices per rod."""

    n_vertices: int
    diffusion_constant: float
    spring_tension: float
    rigidity: float
    spring_length: float
    damping: float

    def fill(self, m: _abi.Mechanics) -> None:
        m.kind = _abi.MECHANICS_ROD
        m.n_vertices = self.n_vertices
        m.diffusion_constant = self.diffusion_constant
        m.spring_tension = self.spring_tension
        m.rigidity = self.rigidity
        m.spring_length = self.spring_length
        m.damping = self.damping


# Dimension-named aliases, as in the reference.
NewtonDamped2D = NewtonDamped3D = NewtonDamped
Langevin2D = Langevin3D = Langevin
Brownian2D = Brownian3D = Brownian


# --------------------------------------------------------------------------
# Domain forces (the DomainForce aspect, SubDomainForce)
# --------------------------------------------------------------------------
@dataclass
class DomainSpring:
    """A linear spring of the domain on every point / rod vertex,
    f_k = (-strength_k) * (x_k - center_k): the test domain of
    cellular_raza/tests/domain_force_spring.rs:56-67, per axis."""

    strength: Sequence[float]
    center: Sequence[float] = (0.0, 0.0, 0.0)

    def to_struct(self) -> _abi.DomainForce:
        f = _abi.DomainForce()
        f.kind = _abi.DOMAIN_FORCE_SPRING
        for k, v in enumerate(self.strength):
            f.strength[k] = float(v)
        for k, v in enumerate(self.center):
            f.center[k] = float(v)
        return f


@dataclass
class RodGelForce:
    """CartesianSubDomainRods' gel pressure and surface friction on 3-D rods
    (cell_building_blocks/bacterial_rods.rs:425-479)."""

    gel_pressure: float
    surface_friction: float
    surface_friction_distance: float

    def to_struct(self) -> _abi.DomainForce:
        f = _abi.DomainForce()
        f.kind = _abi.DOMAIN_FORCE_ROD_GEL
        f.gel_pressure = self.gel_pressure
        f.surface_friction = self.surface_friction
        f.surface_friction_distance = self.surface_friction_distance
        return f


@dataclass
class RodGrowthCycle:
    """The Cycle + NeighborSensing of the rods example (bacterial_rods/src/main.rs:48-111):
    growth rate from the neighbour count, spring-length growth, division of the rod
    (RodMechanics::divide) at a threshold.  The example uses sensing_radius = 2 R,
    division_radius = R (R the interaction radius) and neighbor_saturation = 8."""

    growth_base_rate: float
    division_threshold: float
    division_radius: float
    sensing_radius: float
    neighbor_saturation: float = 8.0
    max_agents: int = 1 << 20

    def to_struct(self) -> _abi.RodCycle:
        y = _abi.RodCycle()
        y.growth_base_rate = self.growth_base_rate
        y.division_threshold = self.division_threshold
        y.division_radius = self.division_radius
        y.sensing_radius = self.sensing_radius
        y.neighbor_saturation = self.neighbor_saturation
        y.max_agents = int(self.max_agents)
        return y


def instance_struct(domain: CartesianCuboid, interaction, mechanics) -> _abi.Instance:
    inst = _abi.Instance()
    inst.domain = domain.to_struct()
    interaction.fill(inst.potential)
    mechanics.fill(inst.mechanics)
    return inst
