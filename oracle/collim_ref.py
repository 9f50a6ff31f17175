"""ORACLE (test infrastructure only): scalar restatement of the beam-limiting geometry
and the weight models of collimated and mask-shaped sources.

Only tests/ use this module, as the checker of the vectorised host geometry
(``pyradmc_amd.collimation``) and of the device evaluation of the same factors. One
ray at a time, in float64, with plain ``math``:

  jaw_chord        pyradmc/geometry/collimation.py:280-323   (straight or focused edges)
  leaf_end_chord   pyradmc/geometry/collimation.py:326-390   (body | tip disc, incl.-excl.)
  mlc_chord        pyradmc/geometry/collimation.py:393-489   (every strip, both banks)
  stack_factor     pyradmc/geometry/collimation.py:927-984   (log-log mu table, Beer-Lambert)
  mask_factor      pyradmc/geometry/collimation.py:987-1068  (bilinear mask on a plane)

Pinned in tests/test_collimation.py by closed forms: straight and slanted chords of a
slab, the full-or-nothing chord of a focused edge for focal-spot rays, the circular
chord of a rounded leaf end, exact transmission at table nodes, and pixel values of a
mask at pixel centres.
"""

from __future__ import annotations

import math

INF = math.inf


def _length(lo: float, hi: float) -> float:
    return hi - lo if hi > lo else 0.0


def _clip(s_lo: float, s_hi: float, a: float, b: float, above: bool) -> float:
    """|[s_lo, s_hi] & {a + t b >= 0}| (``<= 0`` unless above)."""
    if b == 0.0:
        ok = a >= 0.0 if above else a <= 0.0
        return _length(s_lo, s_hi) if ok else 0.0
    root = -a / b
    if (b > 0.0) == above:
        return _length(max(s_lo, root), s_hi)
    return _length(s_lo, min(s_hi, root))


def _slab(p_w: float, d_w: float, z_top: float, z_bottom: float, from_origin: bool):
    if d_w == 0.0:
        if p_w < z_top or p_w > z_bottom:
            return INF, -INF
        lo, hi = -INF, INF
    else:
        t1, t2 = (z_top - p_w) / d_w, (z_bottom - p_w) / d_w
        lo, hi = min(t1, t2), max(t1, t2)
    if from_origin:
        lo = max(lo, 0.0)
    return lo, hi


def jaw_chord(p_q, d_q, p_w, d_w, z_top, z_bottom, edge_neg, edge_pos, focused, from_origin=False) -> float:
    s_lo, s_hi = _slab(p_w, d_w, z_top, z_bottom, from_origin)
    if s_hi <= s_lo:
        return 0.0
    if focused:
        zm = 0.5 * (z_top + z_bottom)
        ap, bp = zm * p_q - edge_pos * p_w, zm * d_q - edge_pos * d_w
        an, bn = zm * p_q - edge_neg * p_w, zm * d_q - edge_neg * d_w
    else:
        ap, bp, an, bn = p_q - edge_pos, d_q, p_q - edge_neg, d_q
    return _clip(s_lo, s_hi, ap, bp, True) + _clip(s_lo, s_hi, an, bn, False)


def leaf_end_chord(i_lo, i_hi, p_u, d_u, p_w, d_w, cu, zm, r, above) -> float:
    """|I & (H | D)| over the strip interval I = [i_lo, i_hi] for the body half-space
    H (u >= cu, or <= for the other bank) and the tip disc D about (cu, zm)."""
    if d_u == 0.0:
        inside = p_u >= cu if above else p_u <= cu
        h_lo, h_hi = (-INF, INF) if inside else (INF, -INF)
    else:
        root = (cu - p_u) / d_u
        if (d_u > 0.0) == above:
            h_lo, h_hi = root, INF
        else:
            h_lo, h_hi = -INF, root
    ou, ow = p_u - cu, p_w - zm
    a = d_u * d_u + d_w * d_w
    b = 2.0 * (d_u * ou + d_w * ow)
    c = ou * ou + ow * ow - r * r
    if a > 0.0:
        disc = b * b - 4.0 * a * c
        if disc > 0.0:
            q = math.sqrt(disc)
            c_lo, c_hi = (-b - q) / (2.0 * a), (-b + q) / (2.0 * a)
        else:
            c_lo, c_hi = INF, -INF
    else:
        c_lo, c_hi = (-INF, INF) if c <= 0.0 else (INF, -INF)
    body = _length(max(i_lo, h_lo), min(i_hi, h_hi))
    tip = _length(max(i_lo, c_lo), min(i_hi, c_hi))
    both = _length(max(i_lo, h_lo, c_lo), min(i_hi, h_hi, c_hi))
    return body + tip - both


def _strip(k, p_v, d_v, p_w, d_w, edge_lo, edge_hi):
    """Ray interval inside the strip between two (focused if k != 1) side planes
    k v - edge w_p >= 0 and k v - edge w_p <= 0."""
    if k is None:  # parallel sides
        a_lo, b_lo, a_hi, b_hi = p_v - edge_lo, d_v, p_v - edge_hi, d_v
    else:
        a_lo, b_lo = k * p_v - edge_lo * p_w, k * d_v - edge_lo * d_w
        a_hi, b_hi = k * p_v - edge_hi * p_w, k * d_v - edge_hi * d_w
    lo, hi = -INF, INF
    if b_lo == 0.0:
        if a_lo < 0.0:
            return INF, -INF
    elif b_lo > 0.0:
        lo = max(lo, -a_lo / b_lo)
    else:
        hi = min(hi, -a_lo / b_lo)
    if b_hi == 0.0:
        if a_hi > 0.0:
            return INF, -INF
    elif b_hi < 0.0:
        lo = max(lo, -a_hi / b_hi)
    else:
        hi = min(hi, -a_hi / b_hi)
    return lo, hi


def mlc_chord(p, d, z_top, z_bottom, edges, tips_neg, tips_pos, r, focused_sides, from_origin=False) -> float:
    p_u, p_v, p_w = p
    d_u, d_v, d_w = d
    s_lo, s_hi = _slab(p_w, d_w, z_top, z_bottom, from_origin)
    zm = 0.5 * (z_top + z_bottom)
    total = 0.0
    for i in range(len(edges) - 1):
        st_lo, st_hi = _strip(zm if focused_sides else None, p_v, d_v, p_w, d_w, edges[i], edges[i + 1])
        i_lo, i_hi = max(s_lo, st_lo), min(s_hi, st_hi)
        total += leaf_end_chord(i_lo, i_hi, p_u, d_u, p_w, d_w, tips_pos[i] + r, zm, r, True)
        total += leaf_end_chord(i_lo, i_hi, p_u, d_u, p_w, d_w, tips_neg[i] - r, zm, r, False)
    return total


def to_local(origin, basis, p, d):
    """Engine-frame point and direction in the beam frame (basis rows u, v, w)."""
    q = [p[k] - origin[k] for k in range(3)]
    return ([sum(basis[i][k] * q[k] for k in range(3)) for i in range(3)],
            [sum(basis[i][k] * d[k] for k in range(3)) for i in range(3)])


def device_chords(devices, origin, basis, p, d, from_origin=False) -> list[float]:
    """Per-device chords; ``devices`` are dicts: kind 'jaw' (axis, z_top, z_bottom,
    edge_neg, edge_pos, focused) or 'mlc' (z_top, z_bottom, edges, tips_neg, tips_pos,
    r, focused_sides)."""
    pl, dl = to_local(origin, basis, p, d)
    out = []
    for dv in devices:
        if dv["kind"] == "jaw":
            k = 0 if dv["axis"] == "u" else 1
            out.append(jaw_chord(pl[k], dl[k], pl[2], dl[2], dv["z_top"], dv["z_bottom"], dv["edge_neg"],
                                 dv["edge_pos"], dv["focused"], from_origin))
        else:
            out.append(mlc_chord(pl, dl, dv["z_top"], dv["z_bottom"], dv["edges"], dv["tips_neg"],
                                 dv["tips_pos"], dv["r"], dv["focused_sides"], from_origin))
    return out


def interp(x: float, xs, ys) -> float:
    """Linear interpolation clamped to the end values (numpy.interp's rule)."""
    if x <= xs[0]:
        return ys[0]
    if x >= xs[-1]:
        return ys[-1]
    lo, hi = 0, len(xs) - 1
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if xs[mid] <= x:
            lo = mid
        else:
            hi = mid
    f = (x - xs[lo]) / (xs[lo + 1] - xs[lo])
    return ys[lo] + f * (ys[lo + 1] - ys[lo])


def stack_factor(devices, origin, basis, log_e, log_mu_rows, densities, energy, p, d) -> float:
    """exp(-sum_i rho_i exp(interp(log E, log_e, log_mu_i)) t_i)."""
    t = device_chords(devices, origin, basis, p, d)
    le = math.log(energy)
    tau = 0.0
    for i, ti in enumerate(t):
        if ti > 0.0:
            tau += densities[i] * math.exp(interp(le, log_e, log_mu_rows[i])) * ti
    return math.exp(-tau)


def mask_factor(mask, center, widths, u_axis, v_axis, p, d) -> float:
    """Bilinear mask value where the ray's line crosses the plane; 0 outside the
    rectangle or for a ray parallel to the plane."""
    n = [u_axis[1] * v_axis[2] - u_axis[2] * v_axis[1], u_axis[2] * v_axis[0] - u_axis[0] * v_axis[2],
         u_axis[0] * v_axis[1] - u_axis[1] * v_axis[0]]
    den = sum(d[k] * n[k] for k in range(3))
    if den == 0.0:
        return 0.0
    t = sum((center[k] - p[k]) * n[k] for k in range(3)) / den
    off = [p[k] + t * d[k] - center[k] for k in range(3)]
    cu = sum(off[k] * u_axis[k] for k in range(3))
    cv = sum(off[k] * v_axis[k] for k in range(3))
    if abs(cu) > 0.5 * widths[0] or abs(cv) > 0.5 * widths[1]:
        return 0.0
    nu, nv = len(mask), len(mask[0])

    def axis(c, n_pix, w):
        x = min(max((c + 0.5 * w) / (w / n_pix) - 0.5, 0.0), float(n_pix - 1))
        i0 = min(int(x), max(n_pix - 2, 0))
        return i0, x - i0

    i0, fu = axis(cu, nu, widths[0])
    j0, fv = axis(cv, nv, widths[1])
    i1, j1 = min(i0 + 1, nu - 1), min(j0 + 1, nv - 1)
    a, b, c, e = mask[i0][j0], mask[i1][j0], mask[i0][j1], mask[i1][j1]
    return a + fu * (b - a) + fv * (c - a) + fu * fv * (e - b - c + a)
