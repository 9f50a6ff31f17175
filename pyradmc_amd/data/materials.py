"""Material registry entries the analytic water source needs.

Material indices cross unchanged into the kernels (uint8 voxel map). The registry
size matters for grid validation (an index beyond it is refused), and the water row
carries every number the analytic source evaluates. The values are physical data
(ICRU-37 I-value, Sternheimer-Berger-Seltzer density-effect row, NIST ESTAR radiative
anchors, IUPAC 2021 atomic weights) and are the same values the reference registry
holds (``pyradmc/data/materials.py:46-71, 148-226``) because they are data, not design.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import NamedTuple

AVOGADRO: float = 6.022_140_76e23

STANDARD_ATOMIC_WEIGHT: dict[int, float] = {1: 1.008, 8: 15.999}

WATER: int = 0
AIR: int = 1
LUNG: int = 2
ADIPOSE: int = 3
CORTICAL_BONE: int = 4
TUNGSTEN: int = 5
TUNGSTEN_ALLOY: int = 6

REGISTRY_NAMES: tuple[str, ...] = (
    "water",
    "air",
    "lung",
    "adipose",
    "cortical_bone",
    "tungsten",
    "tungsten_alloy",
)
"""Registry order (``pyradmc/data/materials.py:148-186``); only water has data here."""

N_REGISTRY: int = len(REGISTRY_NAMES)


class SternheimerParameters(NamedTuple):
    a: float
    m: float
    x0: float
    x1: float
    cbar: float
    delta0: float


@dataclass(frozen=True)
class MaterialData:
    name: str
    density: float
    electrons_per_gram: float
    composition: tuple[tuple[int, float], ...]
    mean_excitation_mev: float
    sternheimer: SternheimerParameters
    radiative_anchors: tuple[tuple[float, float], ...]


_WATER_MOLAR_MASS = 2.0 * STANDARD_ATOMIC_WEIGHT[1] + STANDARD_ATOMIC_WEIGHT[8]

WATER_DATA = MaterialData(
    name="water",
    density=1.0,
    # 10 electrons per H2O molecule over the molecular standard weight 18.01528 g/mol.
    electrons_per_gram=10.0 * AVOGADRO / 18.01528,
    composition=(
        (1, 2.0 * STANDARD_ATOMIC_WEIGHT[1] / _WATER_MOLAR_MASS),
        (8, STANDARD_ATOMIC_WEIGHT[8] / _WATER_MOLAR_MASS),
    ),
    mean_excitation_mev=75.0e-6,
    sternheimer=SternheimerParameters(
        a=0.09116, m=3.4773, x0=0.2400, x1=2.8004, cbar=3.5017, delta0=0.097
    ),
    radiative_anchors=((1.0, 0.0128), (10.0, 0.1813), (20.0, 0.4008)),
)
