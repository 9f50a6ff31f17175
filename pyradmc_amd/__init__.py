"""pyradmc_amd: MI355X-native transport-and-scoring engine for pyRadMC workloads.

``HipEngine`` drops in for pyRadMC's ``WarpEngine``: the same constructor, ``run`` and
``run_dij`` signatures and result types, backed by hand-written HIP kernels for gfx950
through the C ABI in ``include/hipmc.h``.
"""

from __future__ import annotations

from pyradmc_amd.constants import (
    DIJ_TRUNCATION_RELATIVE,
    ECUT_MEV,
    ELECTRON_MASS_MEV,
    GY_PER_MEV_PER_G,
    PCUT_MEV,
    VERSION as __version__,
)
from pyradmc_amd.data.analytic import AnalyticCrossSections
from pyradmc_amd.data.materials import MATERIALS, WATER
from pyradmc_amd.data.tabulated import TabulatedCrossSections, TabulatedData
from pyradmc_amd.collimation import (
    MLC,
    BeamFrame,
    BeamLimitingStack,
    CollimatedBeamletSource,
    CollimatedSource,
    JawPair,
    TransmissionMaskBeamletSource,
    TransmissionMaskSource,
    project_between_planes,
)
from pyradmc_amd.geometry import (
    BeamletGridSource,
    GaussianSpotBeamletSource,
    GaussianSpotBeamSource,
    ParallelBeamSource,
    PencilBeamSource,
    SpectralBeamletSource,
    SpectralBeamSource,
    SpectralParallelBeamSource,
    Spectrum,
    VoxelGrid,
    bremsstrahlung_6mv_like,
)
from pyradmc_amd.head import AirColumn, presolve_head
from pyradmc_amd.phasespace import (
    DevicePhaseSpace,
    IAEAPhaseSpace,
    InMemoryPhaseSpaceSource,
    PhaseSpaceSource,
    read_iaea_header,
    write_iaea_phase_space,
)
from pyradmc_amd.progress import ProgressEvent
from pyradmc_amd.scoring import DijResult, RunProvenance, ScoringGrid, TransportResult


def __getattr__(name: str):
    # The engine loads the native library; import it lazily so the host-side types
    # stay importable on a machine without the built library or a GPU.
    if name == "HipEngine":
        from pyradmc_amd.engine import HipEngine

        return HipEngine
    if name == "DeviceDij":
        from pyradmc_amd.spmm import DeviceDij

        return DeviceDij
    raise AttributeError(name)


__all__ = [
    "AirColumn",
    "AnalyticCrossSections",
    "BeamFrame",
    "BeamLimitingStack",
    "BeamletGridSource",
    "CollimatedBeamletSource",
    "CollimatedSource",
    "DIJ_TRUNCATION_RELATIVE",
    "DeviceDij",
    "DevicePhaseSpace",
    "DijResult",
    "ECUT_MEV",
    "ELECTRON_MASS_MEV",
    "GY_PER_MEV_PER_G",
    "GaussianSpotBeamSource",
    "GaussianSpotBeamletSource",
    "HipEngine",
    "IAEAPhaseSpace",
    "InMemoryPhaseSpaceSource",
    "JawPair",
    "MLC",
    "MATERIALS",
    "PCUT_MEV",
    "ParallelBeamSource",
    "PhaseSpaceSource",
    "PencilBeamSource",
    "ProgressEvent",
    "RunProvenance",
    "ScoringGrid",
    "SpectralBeamSource",
    "SpectralBeamletSource",
    "SpectralParallelBeamSource",
    "Spectrum",
    "TabulatedCrossSections",
    "TabulatedData",
    "TransmissionMaskBeamletSource",
    "TransmissionMaskSource",
    "TransportResult",
    "VoxelGrid",
    "WATER",
    "bremsstrahlung_6mv_like",
    "presolve_head",
    "project_between_planes",
    "read_iaea_header",
    "write_iaea_phase_space",
]
