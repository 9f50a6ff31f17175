"""Material registry: the physical data of every registry material.

Material indices cross unchanged into the kernels (uint8 voxel map); a compiled
table's rows follow this order. The analytic source evaluates the water row; a
tabulated source (:mod:`pyradmc_amd.data.tabulated`) reads each material's elemental
composition for its elastic screening. The values are physical data (ICRU-37
I-values, Sternheimer-Berger-Seltzer density-effect rows, NIST ESTAR radiative
anchors, PDG 2024 ICRP compositions, IUPAC 2021 atomic weights) and are the values
the reference registry holds (``pyradmc/data/materials.py:46-71, 148-343``) because
they are data, not design.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import NamedTuple

AVOGADRO: float = 6.022_140_76e23

STANDARD_ATOMIC_WEIGHT: dict[int, float] = {
    1: 1.008, 6: 12.011, 7: 14.007, 8: 15.999, 11: 22.98976928, 12: 24.305, 13: 26.9815384,
    15: 30.973761998, 16: 32.06, 17: 35.45, 18: 39.95, 19: 39.0983, 20: 40.078, 26: 55.845,
    28: 58.693, 29: 63.546, 30: 65.38, 74: 183.84,
}

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
"""Registry order (``pyradmc/data/materials.py:148-186``)."""

N_REGISTRY: int = len(REGISTRY_NAMES)


def electrons_per_gram_from_composition(composition: tuple[tuple[int, float], ...]) -> float:
    """``N_A * sum_i w_i Z_i / A_i`` over ``((Z, mass fraction), ...)``, in 1/g."""
    if not composition:
        raise ValueError("empty composition")
    return AVOGADRO * sum(w * z / STANDARD_ATOMIC_WEIGHT[z] for z, w in composition)


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

# ICRP media as published (PDG 2024 atomic and nuclear properties; trace elements kept).
_AIR = ((6, 0.000124), (7, 0.755267), (8, 0.231781), (18, 0.012827))
_LUNG = ((1, 0.101278), (6, 0.102310), (7, 0.028650), (8, 0.757072), (11, 0.001840),
         (12, 0.000730), (15, 0.000800), (16, 0.002250), (17, 0.002660), (19, 0.001940),
         (20, 0.000090), (26, 0.000370), (30, 0.000010))
_ADIPOSE = ((1, 0.119477), (6, 0.637240), (7, 0.007970), (8, 0.232333), (11, 0.000500),
            (12, 0.000020), (15, 0.000160), (16, 0.000730), (17, 0.001190), (19, 0.000320),
            (20, 0.000020), (26, 0.000020), (30, 0.000020))
_CORTICAL_BONE = ((1, 0.047234), (6, 0.144330), (7, 0.041990), (8, 0.446096), (12, 0.002200),
                  (15, 0.104970), (16, 0.003150), (20, 0.209930), (30, 0.000100))
_TUNGSTEN_ALLOY = ((28, 0.035), (29, 0.015), (74, 0.95))


def _medium(name, density, composition, i_ev, sternheimer, anchors) -> MaterialData:
    return MaterialData(name=name, density=density,
                        electrons_per_gram=electrons_per_gram_from_composition(composition),
                        composition=composition, mean_excitation_mev=i_ev * 1.0e-6,
                        sternheimer=SternheimerParameters(*sternheimer), radiative_anchors=anchors)


MATERIALS: tuple[MaterialData, ...] = (
    WATER_DATA,
    _medium("air", 1.205e-3, _AIR, 85.7, (0.1091, 3.3994, 1.7418, 4.2759, 10.5961, 0.0),
            ((1.0, 1.271e-2), (10.0, 1.795e-1), (20.0, 4.042e-1))),
    _medium("lung", 1.050, _LUNG, 75.3, (0.0859, 3.5353, 0.2261, 2.8001, 3.4708, 0.0),
            ((1.0, 1.265e-2), (10.0, 1.793e-1), (20.0, 4.038e-1))),
    _medium("adipose", 0.920, _ADIPOSE, 63.2, (0.1028, 3.4817, 0.1827, 2.6530, 3.2367, 0.0),
            ((1.0, 1.070e-2), (10.0, 1.542e-1), (20.0, 3.485e-1))),
    _medium("cortical_bone", 1.850, _CORTICAL_BONE, 106.4,
            (0.0620, 3.5919, 0.1161, 3.0919, 3.6488, 0.0),
            ((1.0, 1.824e-2), (10.0, 2.476e-1), (20.0, 5.525e-1))),
    _medium("tungsten", 19.30, ((74, 1.0),), 727.0, (0.1551, 2.8447, 0.2167, 3.4960, 5.4059, 0.14),
            ((1.0, 1.159e-1), (10.0, 1.132), (20.0, 2.406))),
    _medium("tungsten_alloy", 18.0, _TUNGSTEN_ALLOY, 692.5,
            (0.1564, 2.8413, 0.2323, 3.4878, 5.3699, 0.14),
            ((1.0, 1.124e-1), (10.0, 1.104), (20.0, 2.348))),
)
"""Every registry material, indexed like ``REGISTRY_NAMES``."""
