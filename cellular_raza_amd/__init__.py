"""MI355X (gfx950) backend for cellular_raza's per-step mechanics hot path.

The product is the C-ABI library `libcrmech.so` (declared in include/crmech.h);
this package binds it with ctypes and mirrors the reference's building-block
names so that host code reads like a cellular_raza simulation.
"""
from .building_blocks import (  # noqa: F401
    BoundLennardJones,
    Brownian,
    Brownian2D,
    Brownian3D,
    CartesianCuboid,
    CellSpecificInteraction,
    CellSpecificInteractionLangevin,
    Langevin,
    Langevin2D,
    Langevin3D,
    MiePotential,
    MorsePotential,
    NewtonDamped,
    NewtonDamped2D,
    NewtonDamped3D,
    NoInteraction,
    RodInteraction,
    RodGelForce,
    RodGrowthCycle,
    RodMechanics,
    DomainSpring,
    SpeciesMorsePotential,
)
from .simulation import (  # noqa: F401
    FixedStepsize,
    InstanceSpec,
    MechanicsBackend,
    Settings,
    SimulationError,
    run_simulation,
)
from .storage import (  # noqa: F401
    StorageAccess,
    StorageBuilder,
    StorageError,
    StorageManager,
    StorageOption,
)
