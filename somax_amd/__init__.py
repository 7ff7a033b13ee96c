"""somax_amd: MI355X-native (gfx950 HIP) QG / shallow-water time step of somax.

Public names mirror ``somax.models`` and ``somax.core``
(somax/models/__init__.py, somax/core/__init__.py) for the models this
library accelerates.
"""

from .core import (
    ArakawaCGrid2D,
    Diagnostics,
    ModalTransform,
    Params,
    PhysConsts,
    Solution,
    SomaxModel,
    State,
    StratificationProfile,
)
from .qg import (
    BaroclinicQG,
    BaroclinicQGDiagnostics,
    BaroclinicQGParams,
    BaroclinicQGPhysConsts,
    BaroclinicQGState,
    BarotropicQG,
    BarotropicQGDiagnostics,
    BarotropicQGParams,
    BarotropicQGPhysConsts,
    BarotropicQGState,
)
from .ensemble import QGEnsemble
from .swm import (
    MultilayerShallowWater2D,
    MultilayerSW2DDiagnostics,
    MultilayerSW2DParams,
    MultilayerSW2DPhysConsts,
    MultilayerSW2DState,
    NonlinearShallowWater2D,
    NonlinearSW2DDiagnostics,
    NonlinearSW2DParams,
    NonlinearSW2DPhysConsts,
    NonlinearSW2DState,
)
from .reparam import ReparameterizedQG, ReparamQGDiagnostics
from .testcases import doublegyre_baroclinic_qg, doublegyre_qg, doublegyre_reparameterized_qg
from .elliptic import MaskedHelmholtz, compute_capacitance_matrices, irregular_boundary
from .run import IntegrationDivergedError, assert_finite_state, chunked_integrate, state_diagnostics

__all__ = [
    "ArakawaCGrid2D", "Diagnostics", "ModalTransform", "Params", "PhysConsts", "Solution", "SomaxModel",
    "State", "StratificationProfile", "BaroclinicQG", "BaroclinicQGDiagnostics", "BaroclinicQGParams",
    "BaroclinicQGPhysConsts", "BaroclinicQGState", "BarotropicQG", "BarotropicQGDiagnostics",
    "BarotropicQGParams", "BarotropicQGPhysConsts", "BarotropicQGState", "doublegyre_baroclinic_qg",
    "doublegyre_qg", "QGEnsemble", "MultilayerShallowWater2D", "MultilayerSW2DDiagnostics", "MultilayerSW2DParams",
    "MultilayerSW2DPhysConsts", "MultilayerSW2DState", "NonlinearShallowWater2D", "NonlinearSW2DDiagnostics",
    "NonlinearSW2DParams", "NonlinearSW2DPhysConsts", "NonlinearSW2DState", "ReparameterizedQG",
    "ReparamQGDiagnostics", "doublegyre_reparameterized_qg", "IntegrationDivergedError", "assert_finite_state",
    "chunked_integrate", "state_diagnostics", "MaskedHelmholtz", "compute_capacitance_matrices",
    "irregular_boundary",
]
