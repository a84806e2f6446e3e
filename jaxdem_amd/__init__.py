"""jaxdem_amd: the DEM time step of cdelv/JaxDEM as hand-written gfx950 HIP
kernels behind JaxDEM's own plugin API.

Public names follow the reference package (jaxdem/__init__.py): ``State``,
``System``, ``Collider``, ``ForceModel``, ``ForceManager``,
``LinearIntegrator``, ``RotationIntegrator``, ``Domain``, ``Material``,
``MaterialTable``, ``MaterialMatchmaker`` and the ``utils`` helpers.  The
kernels live in ``libjdem_hip.so`` (C ABI: include/jdem.h).
"""

from .config import config
from .colliders import Collider, DynamicCellList, NaiveSimulator, NeighborList, refresh_collider
from .domains import Domain, FreeDomain, PeriodicDomain, ReflectDomain
from .factory import Factory
from .forces import (
    CundallStrackForce,
    ForceManager,
    ForceModel,
    ForceRouter,
    HertzFrictionForce,
    HertzianForce,
    LawCombiner,
    SpringForce,
    SpringFrictionForce,
)
from .integrators import (
    DirectEuler,
    LinearIntegrator,
    RotationIntegrator,
    Spiral,
    VelocityVerlet,
    VelocityVerletSpiral,
)
from .materials import Material, MaterialMatchmaker, MaterialTable
from .state import Quaternion, State
from .system import System, save_state
from . import batching, minimizers, precision, serving, utils, writers
from .writers import VTKBaseWriter, VTKWriter
from .batching import Batch
from .minimizers import fire, minimize

__version__ = "0.1.0"

__all__ = [
    "config",
    "Collider",
    "DynamicCellList",
    "NaiveSimulator",
    "NeighborList",
    "refresh_collider",
    "Domain",
    "FreeDomain",
    "PeriodicDomain",
    "ReflectDomain",
    "Factory",
    "ForceModel",
    "ForceManager",
    "SpringForce",
    "HertzianForce",
    "CundallStrackForce",
    "SpringFrictionForce",
    "HertzFrictionForce",
    "LawCombiner",
    "ForceRouter",
    "LinearIntegrator",
    "RotationIntegrator",
    "VelocityVerlet",
    "DirectEuler",
    "VelocityVerletSpiral",
    "Spiral",
    "Material",
    "MaterialMatchmaker",
    "MaterialTable",
    "State",
    "Quaternion",
    "System",
    "save_state",
    "VTKWriter",
    "VTKBaseWriter",
    "Batch",
    "batching",
    "serving",
    "precision",
    "fire",
    "minimize",
    "minimizers",
    "utils",
]
