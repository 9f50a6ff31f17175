"""Accuracy-defining constants of the transport-and-scoring path.

Each value is the one the reference engine compiles into its kernels; the HIP kernels
receive the same numbers through their launch parameters or compile them in
(``csrc/hipmc_device.h``). Changing one here changes the physics, so they live in one
place. Reference: ``pyradmc/__init__.py:22-101``, ``pyradmc/transport/electron.py:86-112``,
``pyradmc/backends/warp/kernels.py:90-103``, ``pyradmc/data/interface.py:38-65``.
"""

from __future__ import annotations

ECUT_MEV: float = 0.200
"""Electron transport and production cutoff (kinetic, MeV)."""

PCUT_MEV: float = 0.050
"""Photon transport cutoff (MeV)."""

DIJ_TRUNCATION_RELATIVE: float = 1.0e-3
"""Default Dij truncation relative to each column's maximum."""

PHOTON_ROULETTE_MEV: float = 0.5
PHOTON_ROULETTE_SURVIVAL: float = 0.5
PHOTON_ROULETTE_WEIGHT_CAP: float = 4.0
PHOTON_SPLIT_N: int = 1

ELECTRON_MASS_MEV: float = 0.510_998_950_69
"""Electron rest energy (CODATA 2022), used by the transport samplers."""

GY_PER_MEV_PER_G: float = 1.602_176_634e-10

ENERGY_QUANTUM_MEV: float = 1.0e-9
"""Fixed-point tally quantum: one int64 count is 1e-9 MeV."""

ENTRY_NUDGE_CM: float = 1.0e-4
"""float32 nudge past the grid surface after a vacuum flight."""

BOUNDARY_NUDGE_CM: float = 1.0e-4
"""Over-step past a voxel face on a boundary-limited electron substep."""

STEP_ENERGY_FRACTION: float = 0.05
"""Substep fraction of the CSDA range under the Gaussian hinge."""

GS_STEP_ENERGY_FRACTION: float = 0.20
"""Substep fraction of the CSDA range under Goudsmit-Saunderson (the default)."""

STEP_HINGE_THETA2_MAX: float = 0.10
"""Gaussian-hinge mean-square angle cap per substep (rad^2)."""

SUBSTEP_DEPOSIT_CAP: int = 1024

GS_BINS_PER_LOG: float = 4.0
GS_TABLE_NODES: int = 512
GS_THETA2_MIN: float = 1.0e-12

TABLE_POINTS: int = 2048
COHERENT_POINTS: int = 512
MAJORANT_HEADROOM: float = 1.0 + 1.0e-6
E_MAX_MARGIN: float = 1.0 + 1.0e-6

RAYLEIGH_MOMENTUM_TRANSFER_PER_MEV: float = 80.65543

VERSION: str = "0.1.0"


def default_step_energy_fraction(msc_model: str) -> float:
    """Per-model substep fraction default (``transport/electron.py:115-124``)."""
    return GS_STEP_ENERGY_FRACTION if msc_model == "gs" else STEP_ENERGY_FRACTION
