"""CT volumes to voxel grids: the reference's CT adapter (pyradmc/adapters/ct.py:1-244).

A Hounsfield calibration maps each CT number to a mass density (piecewise linear in HU
between control points, held flat beyond the end points) and to a registry material
(HU bins between ascending thresholds, a bin closed on the left). ``grid_from_hu``
builds the grid from a HU array already in the engine's (x, y, z) order, in cm.
``grid_from_array`` does the frame conversion an ITK image needs (array in (z, y, x)
order, millimetres, origin at the centre of the first voxel, a diagonal direction
matrix whose negative entries flip an axis); ``grid_from_image`` and ``read_ct`` feed
it from SimpleITK, which is optional and only imported there.

``DEFAULT_CALIBRATION`` holds the reference's illustrative Schneider-style curve
(pyradmc/adapters/ct.py:54-66): air and water anchors, a bone segment reaching about
1.85 g/cm3 near 1400 HU; bins into air, lung, adipose, water and cortical bone.
``WATER_CALIBRATION`` keeps the same density ramp and maps every bin to water: the
form this build's analytic (water-only) cross sections can transport, with the CT's
heterogeneity carried by density alone.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from pyradmc_amd.data.materials import ADIPOSE, AIR, CORTICAL_BONE, LUNG, N_REGISTRY, WATER
from pyradmc_amd.geometry import VoxelGrid

MM_PER_CM = 10.0


@dataclass(frozen=True)
class HounsfieldCalibration:
    """HU -> density (g/cm3) and HU -> registry material index.

    ``density_hu`` (strictly ascending) and ``density_values`` (positive) are the ramp's
    control points; ``material_thresholds`` (strictly ascending, n of them) cut the HU
    axis into n + 1 bins, bin i taking ``material_indices[i]``.
    """

    density_hu: tuple
    density_values: tuple
    material_thresholds: tuple
    material_indices: tuple
    _nodes: np.ndarray = field(init=False, repr=False, compare=False)
    _values: np.ndarray = field(init=False, repr=False, compare=False)
    _cuts: np.ndarray = field(init=False, repr=False, compare=False)
    _mats: np.ndarray = field(init=False, repr=False, compare=False)

    def __post_init__(self) -> None:
        nodes = np.asarray(self.density_hu, dtype=np.float64)
        values = np.asarray(self.density_values, dtype=np.float64)
        cuts = np.asarray(self.material_thresholds, dtype=np.float64)
        mats = np.asarray(self.material_indices, dtype=np.int32)
        if nodes.ndim != 1 or nodes.size < 2 or values.shape != nodes.shape:
            raise ValueError("density_hu and density_values need the same length, at least 2")
        if np.any(np.diff(nodes) <= 0.0):
            raise ValueError("density_hu must be strictly ascending")
        if np.any(values <= 0.0):
            raise ValueError("density_values must be positive (g/cm3)")
        if cuts.ndim != 1 or np.any(np.diff(cuts) <= 0.0):
            raise ValueError("material_thresholds must be strictly ascending")
        if mats.shape != (cuts.size + 1,):
            raise ValueError(f"{cuts.size} thresholds make {cuts.size + 1} bins; got {mats.size} material_indices")
        if np.any((mats < 0) | (mats >= N_REGISTRY)):
            raise ValueError("material_indices must be registry indices")
        object.__setattr__(self, "_nodes", nodes)
        object.__setattr__(self, "_values", values)
        object.__setattr__(self, "_cuts", cuts)
        object.__setattr__(self, "_mats", mats)

    def density(self, hu) -> np.ndarray:
        """Density (g/cm3) at the given HU, linear between nodes, flat beyond them."""
        return np.interp(np.asarray(hu, dtype=np.float64), self._nodes, self._values)

    def material(self, hu) -> np.ndarray:
        """Registry index of the bin holding each HU (a threshold belongs to the bin above)."""
        return self._mats[np.searchsorted(self._cuts, np.asarray(hu, dtype=np.float64), side="right")]


_RAMP_HU = (-1000.0, 0.0, 1000.0, 3000.0)
_RAMP_DENSITY = (0.00121, 1.000, 1.600, 2.800)
_BIN_CUTS = (-950.0, -120.0, -20.0, 125.0)

DEFAULT_CALIBRATION = HounsfieldCalibration(_RAMP_HU, _RAMP_DENSITY, _BIN_CUTS,
                                            (AIR, LUNG, ADIPOSE, WATER, CORTICAL_BONE))
WATER_CALIBRATION = HounsfieldCalibration(_RAMP_HU, _RAMP_DENSITY, _BIN_CUTS, (WATER,) * 5)


def grid_from_hu(hu, spacing, *, calibration: HounsfieldCalibration = DEFAULT_CALIBRATION,
                 origin=(0.0, 0.0, 0.0)) -> VoxelGrid:
    """VoxelGrid from a HU volume in (x, y, z) order; ``spacing`` and ``origin`` in cm,
    the origin at the lower corner of voxel (0, 0, 0)."""
    hu = np.asarray(hu, dtype=np.float64)
    if hu.ndim != 3:
        raise ValueError(f"expected a 3D HU volume, got {hu.ndim} dimensions")
    return VoxelGrid(shape=hu.shape, spacing=tuple(spacing), origin=tuple(origin),
                     density=calibration.density(hu), material=calibration.material(hu).astype(np.int32))


def grid_from_array(hu_zyx, spacing_mm, origin_mm, direction, *,
                    calibration: HounsfieldCalibration = DEFAULT_CALIBRATION) -> VoxelGrid:
    """VoxelGrid from an ITK-frame volume: ``hu_zyx`` indexed (z, y, x); ``spacing_mm``,
    ``origin_mm`` (centre of the first voxel) and the 3x3 ``direction`` matrix in (x, y, z)
    order. The matrix must be diagonal; a negative entry means that axis's coordinate
    falls with the index, so the axis is reversed to store coordinates ascending."""
    vol = np.asarray(hu_zyx, dtype=np.float64)
    if vol.ndim != 3:
        raise ValueError(f"expected a 3D image, got {vol.ndim} dimensions")
    d = np.asarray(direction, dtype=np.float64).reshape(3, 3)
    if np.max(np.abs(d - np.diag(np.diag(d)))) > 1e-6:
        raise ValueError("oblique CT orientation is not supported; resample to an axis-aligned grid first")
    hu = np.transpose(vol, (2, 1, 0))  # (x, y, z)
    spacing = np.asarray(spacing_mm, dtype=np.float64)
    first = np.asarray(origin_mm, dtype=np.float64).copy()
    for axis in range(3):
        if d[axis, axis] < 0.0:
            hu = np.flip(hu, axis=axis)
            first[axis] -= spacing[axis] * (hu.shape[axis] - 1)  # the last voxel has the lowest coordinate
    corner_cm = (first - 0.5 * spacing) / MM_PER_CM
    return grid_from_hu(np.ascontiguousarray(hu), tuple(spacing / MM_PER_CM), calibration=calibration,
                        origin=tuple(corner_cm))


def grid_from_image(image, *, calibration: HounsfieldCalibration = DEFAULT_CALIBRATION) -> VoxelGrid:
    """VoxelGrid from a 3D SimpleITK image whose values are HU (NIfTI, NRRD, or a DICOM
    series with its rescale applied)."""
    import SimpleITK as sitk

    if image.GetDimension() != 3:
        raise ValueError(f"expected a 3D CT image, got {image.GetDimension()}D")
    return grid_from_array(sitk.GetArrayFromImage(image), image.GetSpacing(), image.GetOrigin(),
                           image.GetDirection(), calibration=calibration)


def read_ct(path, *, calibration: HounsfieldCalibration = DEFAULT_CALIBRATION) -> VoxelGrid:
    """Read a single-volume CT file with SimpleITK (optional dependency)."""
    try:
        import SimpleITK as sitk
    except ImportError as exc:
        raise ImportError("read_ct needs SimpleITK (the reference's optional CT extra)") from exc
    return grid_from_image(sitk.ReadImage(str(path)), calibration=calibration)
