"""Beam-limiting devices (jaws, a rounded-tip MLC) as ray-attenuation geometry, and the
sources that see a field through them (pyradmc/geometry/collimation.py).

Geometry. Devices live in a beam frame: origin at the focal spot, ``w`` the beam axis
(downstream), ``u`` the leaf-travel axis, ``v`` the leaf-width axis, a right-handed
orthonormal triad given in the engine frame. Every device is a slab
``z_top <= w <= z_bottom``; edges and leaf tips are physical coordinates on the
device mid-plane ``z_mid``. A jaw pair holds a block ``q >= edge_pos`` and a block
``q <= edge_neg`` (``q`` the local ``u`` or ``v``); a focused edge is the plane through
the focal spot and the mid-plane edge, ``z_mid q = edge w``. An MLC tiles ``v`` with
leaf strips (parallel or focused sides); each leaf end is the union of the flat body
``u >= tip + R`` (mirrored for the other bank) and a cylinder of radius ``R`` about
``(tip + R, z_mid)`` in the ``(u, w)`` plane, so a leaf's chord is
``|S & H| + |S & D| - |S & H & D|`` over the slab-and-strip ``S``, body ``H`` and disc
``D``. Chords are measured along the whole line (no ``t >= 0`` clamp) unless
``from_origin``, because a source may sit upstream or downstream of the devices.

Intervals of the ray parameter are ``(lo, hi)`` array pairs, an empty one
``(+inf, -inf)``; a ray trapped inside a block parallel to the slab has an infinite
chord (zero transmission).

Sources. ``CollimatedSource`` / ``CollimatedBeamletSource`` multiply each primary's
weight by ``exp(-sum_i mu_i(E) rho_i t_i)`` (narrow-beam Beer-Lambert with the total
attenuation coefficient, log-log tables of ``mu/rho`` per device material);
``TransmissionMaskSource`` / ``TransmissionMaskBeamletSource`` by a 0..1 mask read
bilinearly where the ray crosses a plane rectangle. Wrapping a built-in source keeps
it on the in-kernel generator, which evaluates the same factor per primary on the
device; wrapping a host-sampled source applies the factor to its sampled columns.
Neither draws a random number.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from pyradmc_amd.geometry import BeamletSource, Source

__all__ = [
    "MLC",
    "BeamFrame",
    "BeamLimitingStack",
    "CollimatedBeamletSource",
    "CollimatedSource",
    "CompiledStack",
    "JawPair",
    "MaskTransmission",
    "StackTransmission",
    "TransmissionMaskBeamletSource",
    "TransmissionMaskSource",
    "project_between_planes",
]

_INF = math.inf


def _unit(v, name: str) -> np.ndarray:
    a = np.asarray(v, dtype=np.float64).reshape(3)
    n = float(np.linalg.norm(a))
    if n <= 0.0:
        raise ValueError(f"{name} must be a nonzero vector")
    return a / n


@dataclass(frozen=True)
class BeamFrame:
    """The beam frame (pyradmc/geometry/collimation.py:85-124): ``origin`` the focal
    spot, axes normalised on construction, required orthonormal and right-handed
    (``u x v = w``) to 1e-9, since a mirrored frame would swap the leaf banks."""

    origin: tuple[float, float, float]
    u_axis: tuple[float, float, float] = (1.0, 0.0, 0.0)
    v_axis: tuple[float, float, float] = (0.0, 1.0, 0.0)
    w_axis: tuple[float, float, float] = (0.0, 0.0, 1.0)

    def __post_init__(self) -> None:
        u, v, w = _unit(self.u_axis, "u_axis"), _unit(self.v_axis, "v_axis"), _unit(self.w_axis, "w_axis")
        basis = np.stack([u, v, w])
        if not np.allclose(basis @ basis.T, np.eye(3), atol=1e-9):
            raise ValueError("beam frame axes must be orthonormal")
        if float(np.dot(np.cross(u, v), w)) < 1.0 - 1e-9:
            raise ValueError("beam frame axes must be right-handed (u x v = w)")
        object.__setattr__(self, "origin", tuple(float(c) for c in self.origin))
        for name, a in (("u_axis", u), ("v_axis", v), ("w_axis", w)):
            object.__setattr__(self, name, tuple(float(c) for c in a))

    @property
    def basis(self) -> np.ndarray:
        """Rows u, v, w (engine frame): local = basis @ (p - origin)."""
        return np.array([self.u_axis, self.v_axis, self.w_axis], dtype=np.float64)

    def to_local(self, points, directions) -> tuple[np.ndarray, np.ndarray]:
        """(n, 3) engine-frame points and directions in the beam frame."""
        b = self.basis
        p = (np.asarray(points, dtype=np.float64) - np.asarray(self.origin)) @ b.T
        d = np.asarray(directions, dtype=np.float64) @ b.T
        return p, d


def project_between_planes(coordinate, z_from: float, z_to: float):
    """A transverse coordinate on the plane ``w = z_from`` carried along the focal-spot
    ray to the plane ``w = z_to`` (scale ``z_to / z_from``), e.g. isocentre settings to
    device mid-plane positions (pyradmc/geometry/collimation.py:127-140)."""
    if z_from == 0.0:
        raise ValueError("the source plane may not contain the focal point (z_from = 0)")
    if isinstance(coordinate, np.ndarray):
        return coordinate * (z_to / z_from)
    return float(coordinate) * (z_to / z_from)


# ------------------------------------------------------------ interval algebra


def _affine(a, b, above: bool):
    """Ray parameters where ``a + t b >= 0`` (``<= 0`` unless ``above``)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if not above:
        a, b = -a, -b
    safe = np.where(b == 0.0, 1.0, b)
    root = -a / safe
    lo = np.where(b > 0.0, root, -_INF)
    hi = np.where(b < 0.0, root, _INF)
    flat_in = a >= 0.0
    lo = np.where(b == 0.0, np.where(flat_in, -_INF, _INF), lo)
    hi = np.where(b == 0.0, np.where(flat_in, _INF, -_INF), hi)
    return lo, hi


def _band(p, d, lo: float, hi: float):
    """Ray parameters where ``lo <= p + t d <= hi``."""
    return _meet(_affine(p - lo, d, True), _affine(p - hi, d, False))


def _meet(x, y):
    return np.maximum(x[0], y[0]), np.minimum(x[1], y[1])


def _span(x) -> np.ndarray:
    lo, hi = x
    with np.errstate(invalid="ignore"):
        return np.where(hi > lo, hi - lo, 0.0)


def _disc(p_u, d_u, p_w, d_w, cu: float, cw: float, r: float):
    """Ray parameters inside the disc of radius ``r`` about ``(cu, cw)`` in the (u, w)
    plane; a tangent ray misses, a ray along the cylinder axis is all in or all out."""
    ou, ow = p_u - cu, p_w - cw
    a = d_u * d_u + d_w * d_w
    b = 2.0 * (d_u * ou + d_w * ow)
    c = ou * ou + ow * ow - r * r
    disc = b * b - 4.0 * a * c
    hit = (a > 0.0) & (disc > 0.0)
    root = np.sqrt(np.where(hit, disc, 0.0))
    two_a = np.where(a > 0.0, 2.0 * a, 1.0)
    axis_in = (a == 0.0) & (c <= 0.0)
    lo = np.where(hit, (-b - root) / two_a, np.where(axis_in, -_INF, _INF))
    hi = np.where(hit, (-b + root) / two_a, np.where(axis_in, _INF, -_INF))
    return lo, hi


def _slab(p_w, d_w, z_top: float, z_bottom: float, from_origin: bool):
    s = _band(p_w, d_w, z_top, z_bottom)
    if from_origin:
        s = (np.maximum(s[0], 0.0), s[1])
    return s


def _check_slab(z_top: float, z_bottom: float, material: int, density: float) -> None:
    if z_top <= 0.0:
        raise ValueError(f"z_top must be downstream of the focal spot, got {z_top}")
    if z_bottom <= z_top:
        raise ValueError(f"z_bottom must exceed z_top, got {z_bottom} <= {z_top}")
    if material < 0:
        raise ValueError(f"material index must be non-negative, got {material}")
    if density <= 0.0:
        raise ValueError(f"density must be positive, got {density}")


# ------------------------------------------------------------------ devices


@dataclass(frozen=True)
class JawPair:
    """A pair of jaw blocks limiting local ``u`` or ``v`` (pyradmc/geometry/
    collimation.py:492-584). ``edge_neg <= edge_pos``; ``focused`` pivots both edge
    faces through the focal spot."""

    axis: str
    z_top: float
    z_bottom: float
    edge_neg: float
    edge_pos: float
    material: int
    density: float
    focused: bool = False

    def __post_init__(self) -> None:
        if self.axis not in ("u", "v"):
            raise ValueError(f"axis must be 'u' or 'v', got {self.axis!r}")
        _check_slab(self.z_top, self.z_bottom, self.material, self.density)
        if self.edge_neg > self.edge_pos:
            raise ValueError(f"edges cross: edge_neg {self.edge_neg} > edge_pos {self.edge_pos}")

    @property
    def z_mid(self) -> float:
        return 0.5 * (self.z_top + self.z_bottom)

    def path_lengths(self, p_local, d_local, *, from_origin: bool = False) -> np.ndarray:
        """Chord (cm) of each (n, 3) local-frame ray through both blocks."""
        p = np.asarray(p_local, dtype=np.float64)
        d = np.asarray(d_local, dtype=np.float64)
        k = 0 if self.axis == "u" else 1
        s = _slab(p[:, 2], d[:, 2], self.z_top, self.z_bottom, from_origin)
        if self.focused:
            zm = self.z_mid
            pos = _affine(zm * p[:, k] - self.edge_pos * p[:, 2], zm * d[:, k] - self.edge_pos * d[:, 2], True)
            neg = _affine(zm * p[:, k] - self.edge_neg * p[:, 2], zm * d[:, k] - self.edge_neg * d[:, 2], False)
        else:
            pos = _affine(p[:, k] - self.edge_pos, d[:, k], True)
            neg = _affine(p[:, k] - self.edge_neg, d[:, k], False)
        return _span(_meet(s, pos)) + _span(_meet(s, neg))


@dataclass(frozen=True)
class MLC:
    """Paired leaf banks with rounded tips (pyradmc/geometry/collimation.py:587-727).

    ``leaf_edges_v``: ``n_pairs + 1`` strictly increasing mid-plane strip edges; pair
    ``i``'s ``-u`` tip apex at ``tips_neg[i]`` and ``+u`` tip apex at ``tips_pos[i]``
    (``tips_neg[i] <= tips_pos[i]``); ``tip_radius`` at least half the leaf height so
    the arc spans the leaf face. No tongue-and-groove and no interleaf gap."""

    z_top: float
    z_bottom: float
    leaf_edges_v: tuple[float, ...]
    tips_neg: tuple[float, ...]
    tips_pos: tuple[float, ...]
    tip_radius: float
    material: int
    density: float
    focused_sides: bool = False

    def __post_init__(self) -> None:
        _check_slab(self.z_top, self.z_bottom, self.material, self.density)
        half = 0.5 * (self.z_bottom - self.z_top)
        if self.tip_radius < half:
            raise ValueError(f"tip_radius {self.tip_radius} must cover the half-height {half} "
                             "or the arc does not span the leaf face")
        edges = tuple(float(e) for e in self.leaf_edges_v)
        if len(edges) < 2:
            raise ValueError("leaf_edges_v needs at least two edges (one pair)")
        if not all(b > a for a, b in zip(edges, edges[1:])):
            raise ValueError("leaf_edges_v must be strictly increasing")
        n = len(edges) - 1
        if len(self.tips_neg) != n:
            raise ValueError(f"tips_neg has {len(self.tips_neg)} entries for {n} pairs")
        if len(self.tips_pos) != n:
            raise ValueError(f"tips_pos has {len(self.tips_pos)} entries for {n} pairs")
        for i, (a, b) in enumerate(zip(self.tips_neg, self.tips_pos)):
            if a > b:
                raise ValueError(f"pair {i}: tip positions cross ({a} > {b})")
        object.__setattr__(self, "leaf_edges_v", edges)
        object.__setattr__(self, "tips_neg", tuple(float(t) for t in self.tips_neg))
        object.__setattr__(self, "tips_pos", tuple(float(t) for t in self.tips_pos))

    @property
    def n_pairs(self) -> int:
        return len(self.leaf_edges_v) - 1

    @property
    def z_mid(self) -> float:
        return 0.5 * (self.z_top + self.z_bottom)

    def path_lengths(self, p_local, d_local, *, from_origin: bool = False) -> np.ndarray:
        """Chord (cm) of each (n, 3) local-frame ray through both banks."""
        p = np.asarray(p_local, dtype=np.float64)
        d = np.asarray(d_local, dtype=np.float64)
        pu, pv, pw = p[:, 0], p[:, 1], p[:, 2]
        du, dv, dw = d[:, 0], d[:, 1], d[:, 2]
        zm, r = self.z_mid, self.tip_radius
        s = _slab(pw, dw, self.z_top, self.z_bottom, from_origin)
        total = np.zeros(p.shape[0])
        for i in range(self.n_pairs):
            e0, e1 = self.leaf_edges_v[i], self.leaf_edges_v[i + 1]
            if self.focused_sides:
                strip = _meet(_affine(zm * pv - e0 * pw, zm * dv - e0 * dw, True),
                              _affine(zm * pv - e1 * pw, zm * dv - e1 * dw, False))
            else:
                strip = _band(pv, dv, e0, e1)
            inside = _meet(s, strip)
            for apex, side in ((self.tips_pos[i], 1.0), (self.tips_neg[i], -1.0)):
                cu = apex + side * r
                body = _affine(pu - cu, du, side > 0.0)
                disc = _disc(pu, du, pw, dw, cu, zm, r)
                in_body = _meet(inside, body)
                total += _span(in_body) + _span(_meet(inside, disc)) - _span(_meet(in_body, disc))
        return total


@dataclass(frozen=True)
class CompiledStack:
    """A stack as flat arrays for the device (one row per device; jaws read the jaw
    columns, MLCs the leaf arrays from their offsets)."""

    origin: np.ndarray       # [3]
    basis: np.ndarray        # [3][3] rows u, v, w
    kind: np.ndarray         # int32, 0 jaw pair, 1 MLC
    flags: np.ndarray        # int32: jaw axis (0 u, 1 v) | focused << 1 (jaw) / focused sides << 1 (MLC)
    z_top: np.ndarray
    z_bottom: np.ndarray
    edge_neg: np.ndarray     # jaw edges; MLC: tip radius in edge_neg
    edge_pos: np.ndarray
    n_pairs: np.ndarray      # int32
    offset: np.ndarray       # int32: first strip edge of the MLC in leaf_edges, first pair in tips
    pair_offset: np.ndarray  # int32
    leaf_edges: np.ndarray
    tips_neg: np.ndarray
    tips_pos: np.ndarray
    material: np.ndarray     # int32
    density: np.ndarray


@dataclass(frozen=True)
class BeamLimitingStack:
    """Devices stacked along ``w`` in one frame, ascending in ``z_top`` and disjoint
    (pyradmc/geometry/collimation.py:730-867)."""

    frame: BeamFrame
    devices: tuple

    def __post_init__(self) -> None:
        devices = tuple(self.devices)
        if not devices:
            raise ValueError("a stack needs at least one device")
        tops = [dv.z_top for dv in devices]
        if tops != sorted(tops):
            raise ValueError("devices must be ordered ascending in z_top")
        for up, down in zip(devices, devices[1:]):
            if down.z_top < up.z_bottom:
                raise ValueError(f"devices overlap in z: [{up.z_top}, {up.z_bottom}] and "
                                 f"[{down.z_top}, {down.z_bottom}]")
        object.__setattr__(self, "devices", devices)

    @property
    def materials(self) -> tuple[int, ...]:
        return tuple(int(dv.material) for dv in self.devices)

    @property
    def densities(self) -> tuple[float, ...]:
        return tuple(float(dv.density) for dv in self.devices)

    @property
    def exit_z(self) -> float:
        """The downstream face of the last device (local w, cm)."""
        return float(self.devices[-1].z_bottom)

    def path_lengths(self, origins, directions, *, from_origin: bool = False) -> np.ndarray:
        """(n, n_devices) chords of engine-frame rays, one column per device."""
        p, d = self.frame.to_local(np.asarray(origins, dtype=np.float64).reshape(-1, 3),
                                   np.asarray(directions, dtype=np.float64).reshape(-1, 3))
        return np.column_stack([dv.path_lengths(p, d, from_origin=from_origin) for dv in self.devices])

    def compile(self) -> CompiledStack:
        """The stack as flat arrays (cf. BeamLimitingStack.flatten,
        pyradmc/geometry/collimation.py:795-867)."""
        n = len(self.devices)
        kind = np.zeros(n, np.int32)
        flags = np.zeros(n, np.int32)
        e_neg = np.zeros(n)
        e_pos = np.zeros(n)
        n_pairs = np.zeros(n, np.int32)
        off = np.zeros(n, np.int32)
        pair_off = np.zeros(n, np.int32)
        edges: list[float] = []
        tn: list[float] = []
        tp: list[float] = []
        for j, dv in enumerate(self.devices):
            if isinstance(dv, JawPair):
                flags[j] = (0 if dv.axis == "u" else 1) | (2 if dv.focused else 0)
                e_neg[j], e_pos[j] = dv.edge_neg, dv.edge_pos
            else:
                kind[j] = 1
                flags[j] = 2 if dv.focused_sides else 0
                e_neg[j] = dv.tip_radius
                n_pairs[j] = dv.n_pairs
                off[j] = len(edges)
                pair_off[j] = len(tp)
                edges.extend(dv.leaf_edges_v)
                tn.extend(dv.tips_neg)
                tp.extend(dv.tips_pos)
        return CompiledStack(
            origin=np.asarray(self.frame.origin, dtype=np.float64), basis=self.frame.basis, kind=kind,
            flags=flags, z_top=np.array([dv.z_top for dv in self.devices], dtype=np.float64),
            z_bottom=np.array([dv.z_bottom for dv in self.devices], dtype=np.float64), edge_neg=e_neg,
            edge_pos=e_pos, n_pairs=n_pairs, offset=off, pair_offset=pair_off,
            leaf_edges=np.asarray(edges or [0.0], dtype=np.float64), tips_neg=np.asarray(tn or [0.0]),
            tips_pos=np.asarray(tp or [0.0]), material=np.asarray(self.materials, np.int32),
            density=np.asarray(self.densities, np.float64))


# ------------------------------------------------------------- weight models


class StackTransmission:
    """``exp(-sum_i mu_i(E) rho_i t_i)`` along each ray (pyradmc/geometry/
    collimation.py:927-984): per-device log-log tables of the total ``mu/rho`` on
    ``n_table`` geometric energies in ``[e_min, e_max]`` from the cross-section
    source, read with linear interpolation in ``log E`` (clamped at the table ends)."""

    def __init__(self, stack: BeamLimitingStack, cross_sections, e_max: float, e_min: float = 0.010,
                 n_table: int = 512) -> None:
        if max(stack.materials) >= int(cross_sections.n_materials):
            raise ValueError(f"stack material index {max(stack.materials)} is beyond this cross-section "
                             f"source ({cross_sections.n_materials} materials); device materials like "
                             "tungsten need the tabulated backend")
        if not 0.0 < e_min < e_max:
            raise ValueError(f"need 0 < e_min < e_max, got {e_min}, {e_max}")
        if n_table < 2:
            raise ValueError("n_table must be at least 2")
        self.stack = stack
        energies = np.geomspace(e_min, e_max, int(n_table))
        self.log_energies = np.log(energies)
        rows = []
        for m in stack.materials:
            rows.append(np.log(np.asarray(cross_sections.mu_total_bulk(energies, m), dtype=np.float64)))
        self.log_mu = np.stack(rows)
        self.densities = np.asarray(stack.densities, dtype=np.float64)

    def mu(self, energies) -> np.ndarray:
        """(n_devices, n) macroscopic total attenuation (1/cm)."""
        le = np.log(np.asarray(energies, dtype=np.float64))
        return np.exp(np.stack([np.interp(le, self.log_energies, row) for row in self.log_mu])) * \
            self.densities[:, None]

    def factors(self, energies, origins, directions) -> np.ndarray:
        t = self.stack.path_lengths(origins, directions)
        return np.exp(-np.sum(self.mu(energies).T * t, axis=1))


class MaskTransmission:
    """A 0..1 transmission mask on a plane rectangle (pyradmc/geometry/collimation.py:
    987-1068). ``mask[i, j]``: pixel ``i`` along ``u_axis``, ``j`` along ``v_axis``,
    centres tiling the ``width_u x width_v`` rectangle about ``plane_center``; read
    bilinearly between pixel centres in the incremental form (equal neighbours give the
    pixel value exactly), clamped in the half-pixel border. A ray crossing the plane
    outside the rectangle, or parallel to it, gets 0. Energy plays no role."""

    def __init__(self, mask, plane_center, width_u: float, width_v: float, u_axis=(1.0, 0.0, 0.0),
                 v_axis=(0.0, 1.0, 0.0)) -> None:
        m = np.asarray(mask, dtype=np.float64)
        if m.ndim != 2 or m.size == 0:
            raise ValueError(f"the mask must be a non-empty 2D array, got shape {m.shape}")
        if not np.all(np.isfinite(m)) or m.min() < 0.0 or m.max() > 1.0:
            raise ValueError("mask values must be finite and within 0..1")
        if width_u <= 0.0 or width_v <= 0.0:
            raise ValueError(f"widths must be positive, got {width_u}, {width_v}")
        u, v = _unit(u_axis, "u_axis"), _unit(v_axis, "v_axis")
        if abs(float(np.dot(u, v))) > 1e-9:
            raise ValueError("u_axis and v_axis must be orthogonal")
        self.mask = m
        self.center = np.asarray(plane_center, dtype=np.float64).reshape(3)
        self.widths = (float(width_u), float(width_v))
        self.u_axis, self.v_axis = u, v
        self.normal = np.cross(u, v)

    def _axis(self, c, k: int):
        n = self.mask.shape[k]
        w = self.widths[k]
        x = np.clip((c + 0.5 * w) / (w / n) - 0.5, 0.0, float(n - 1))
        i0 = np.minimum(x.astype(np.int64), max(n - 2, 0))
        return i0, x - i0

    def factors(self, energies, origins, directions) -> np.ndarray:
        p = np.asarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.asarray(directions, dtype=np.float64).reshape(-1, 3)
        den = d @ self.normal
        t = ((self.center - p) @ self.normal) / np.where(den == 0.0, 1.0, den)
        off = p + t[:, None] * d - self.center
        cu, cv = off @ self.u_axis, off @ self.v_axis
        inside = (den != 0.0) & (np.abs(cu) <= 0.5 * self.widths[0]) & (np.abs(cv) <= 0.5 * self.widths[1])
        i0, fu = self._axis(cu, 0)
        j0, fv = self._axis(cv, 1)
        i1 = np.minimum(i0 + 1, self.mask.shape[0] - 1)
        j1 = np.minimum(j0 + 1, self.mask.shape[1] - 1)
        m = self.mask
        a, b, c, e = m[i0, j0], m[i1, j0], m[i0, j1], m[i1, j1]
        value = a + fu * (b - a) + fv * (c - a) + fu * fv * (e - b - c + a)
        return np.where(inside, value, 0.0)


def _weighted(columns: dict, model) -> dict:
    """Sampled columns with the weight multiplied by the model's ray factors."""
    o = np.stack([columns["x"], columns["y"], columns["z"]], axis=1).astype(np.float64)
    d = np.stack([columns["ux"], columns["uy"], columns["uz"]], axis=1).astype(np.float64)
    f = model.factors(np.asarray(columns["energy"], dtype=np.float64), o, d)
    out = dict(columns)
    out["weight"] = (np.asarray(columns["weight"], dtype=np.float64) * f).astype(np.float32)
    return out


class _Weighted(Source):
    """An open-field source whose primaries carry a deterministic ray factor."""

    def __init__(self, inner, model) -> None:
        self.inner = inner
        self.model = model

    @property
    def max_energy(self) -> float:
        return float(self.inner.max_energy)

    def sample_batch(self, seed: int, history_offset: int, n: int) -> dict:
        """The inner source's host-sampled columns with the weights attenuated (only
        for an inner source that is host-sampled; built-in inner sources are
        generated and attenuated on the device)."""
        if not callable(getattr(self.inner, "sample_batch", None)):
            raise TypeError(f"{type(self.inner).__name__} is generated on the device and has no sample_batch")
        return _weighted(self.inner.sample_batch(seed, history_offset, n), self.model)


class _WeightedBeamlets(BeamletSource):
    """A beamlet source whose primaries carry a deterministic ray factor."""

    def __init__(self, inner, model) -> None:
        self.inner = inner
        self.model = model

    @property
    def max_energy(self) -> float:
        return float(self.inner.max_energy)

    @property
    def n_beamlets(self) -> int:
        return int(self.inner.n_beamlets)

    def sample_beamlet_batch(self, seed: int, history_offset: int, n: int, beamlet: int) -> dict:
        if not callable(getattr(self.inner, "sample_beamlet_batch", None)):
            raise TypeError(f"{type(self.inner).__name__} is generated on the device and has no "
                            "sample_beamlet_batch")
        return _weighted(self.inner.sample_beamlet_batch(seed, history_offset, n, beamlet), self.model)


class CollimatedSource(_Weighted):
    """``inner`` seen through a beam-limiting stack (pyradmc/geometry/collimation.py:
    1154-1180): each primary's weight times the stack transmission along its ray."""

    def __init__(self, inner, stack: BeamLimitingStack, cross_sections, *, e_min: float = 0.010,
                 n_table: int = 512) -> None:
        super().__init__(inner, StackTransmission(stack, cross_sections, float(inner.max_energy), e_min, n_table))
        self.stack = stack


class CollimatedBeamletSource(_WeightedBeamlets):
    """A beamlet source seen through a beam-limiting stack (pyradmc/geometry/
    collimation.py:1183-1202)."""

    def __init__(self, inner, stack: BeamLimitingStack, cross_sections, *, e_min: float = 0.010,
                 n_table: int = 512) -> None:
        super().__init__(inner, StackTransmission(stack, cross_sections, float(inner.max_energy), e_min, n_table))
        self.stack = stack


class TransmissionMaskSource(_Weighted):
    """``inner`` shaped by a 0..1 mask on a plane rectangle (pyradmc/geometry/
    collimation.py:1205-1261)."""

    def __init__(self, inner, *, mask, plane_center, width_u: float, width_v: float,
                 u_axis=(1.0, 0.0, 0.0), v_axis=(0.0, 1.0, 0.0)) -> None:
        super().__init__(inner, MaskTransmission(mask, plane_center, width_u, width_v, u_axis, v_axis))

    @property
    def mask(self) -> np.ndarray:
        return self.model.mask


class TransmissionMaskBeamletSource(_WeightedBeamlets):
    """A beamlet source shaped by a 0..1 mask (pyradmc/geometry/collimation.py:
    1264-1284)."""

    def __init__(self, inner, *, mask, plane_center, width_u: float, width_v: float,
                 u_axis=(1.0, 0.0, 0.0), v_axis=(0.0, 1.0, 0.0)) -> None:
        super().__init__(inner, MaskTransmission(mask, plane_center, width_u, width_v, u_axis, v_axis))
