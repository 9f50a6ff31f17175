"""Beam-limiting geometry and collimated / mask-shaped sources on the host (CPU).

The oracle's scalar chords (oracle/collim_ref.py) are pinned by closed forms, then the
product's vectorised geometry must agree with them on random rays through every device
variety (straight and focused jaws, MLCs with parallel and focused sides, rays from
either side, from_origin)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import collim_ref as R
from pyradmc_amd.collimation import (
    MLC,
    BeamFrame,
    BeamLimitingStack,
    CollimatedSource,
    JawPair,
    MaskTransmission,
    StackTransmission,
    TransmissionMaskSource,
    project_between_planes,
)

H = 8.0  # device height (cm)


# ------------------------------------------------------- oracle, closed forms


def test_straight_jaw_closed_forms():
    # vertical rays: full thickness inside either block, nothing in the opening
    assert R.jaw_chord(5.0, 0.0, 0.0, 1.0, 20.0, 28.0, -3.0, 4.0, False) == pytest.approx(8.0)
    assert R.jaw_chord(-5.0, 0.0, 0.0, 1.0, 20.0, 28.0, -3.0, 4.0, False) == pytest.approx(8.0)
    assert R.jaw_chord(0.5, 0.0, 0.0, 1.0, 20.0, 28.0, -3.0, 4.0, False) == 0.0
    # a slanted ray wholly inside a block crosses H / cos(theta)
    th = 0.2
    ch = R.jaw_chord(10.0, math.sin(th), 0.0, math.cos(th), 20.0, 28.0, -3.0, 4.0, False)
    assert ch == pytest.approx(8.0 / math.cos(th), rel=1e-14)
    # a vertical ray on the edge plane: the block is closed (q >= edge)
    assert R.jaw_chord(4.0, 0.0, 0.0, 1.0, 20.0, 28.0, -3.0, 4.0, False) == pytest.approx(8.0)
    # parallel to the slab inside a block: infinite; outside the slab: nothing
    assert R.jaw_chord(5.0, 1.0, 24.0, 0.0, 20.0, 28.0, -3.0, 4.0, False) == math.inf
    assert R.jaw_chord(5.0, 1.0, 30.0, 0.0, 20.0, 28.0, -3.0, 4.0, False) == 0.0
    # from_origin: a ray starting at the mid-plane sees half the slab
    assert R.jaw_chord(5.0, 0.0, 24.0, 1.0, 20.0, 28.0, -3.0, 4.0, False, True) == pytest.approx(4.0)


def test_focused_jaw_is_all_or_nothing_for_focal_rays():
    # A ray from the focal spot crossing the mid-plane at q: the focused block either
    # covers its whole slant chord H / w or none of it.
    zt, zb, e = 20.0, 28.0, 4.0
    zm = 0.5 * (zt + zb)
    for q_mid, full in ((4.5, True), (3.9, False), (-3.2, True), (-2.9, False)):
        d = np.array([q_mid, 0.0, zm])
        d /= np.linalg.norm(d)
        ch = R.jaw_chord(0.0, d[0], 0.0, d[2], zt, zb, -3.0, e, True)
        assert ch == pytest.approx(H / d[2] if full else 0.0, abs=1e-12)
    # the straight edge at the same crossing cuts only part of the chord
    d = np.array([4.05, 0.0, zm])
    d /= np.linalg.norm(d)
    part = R.jaw_chord(0.0, d[0], 0.0, d[2], zt, zb, -3.0, e, False)
    assert 0.0 < part < H / d[2]


def test_rounded_leaf_end_circular_chord():
    # A vertical ray at distance delta from the tip cylinder's axis (inside the arc,
    # outside the flat body) crosses 2 min(sqrt(R^2 - delta^2), H/2).
    zt, zb, r = 20.0, 28.0, 12.0
    apex = 1.0  # +u leaf tip apex; axis at u = apex + r
    edges, tn, tp = [-1.0, 1.0], [-5.0], [apex]
    for u in (1.2, 2.0, 4.0, 8.0, 12.9):
        delta = apex + r - u
        expect = 2.0 * min(math.sqrt(r * r - delta * delta), 0.5 * (zb - zt))
        got = R.mlc_chord([u, 0.0, 0.0], [0.0, 0.0, 1.0], zt, zb, edges, tn, tp, r, False)
        assert got == pytest.approx(expect, rel=1e-13)
    # behind the axis the body covers the whole height; in the gap nothing
    assert R.mlc_chord([apex + r + 0.5, 0.0, 0.0], [0, 0, 1.0], zt, zb, edges, tn, tp, r, False) == pytest.approx(8.0)
    assert R.mlc_chord([-2.0, 0.0, 0.0], [0, 0, 1.0], zt, zb, edges, tn, tp, r, False) == 0.0
    # outside every strip: the bank ends at its outer edges
    assert R.mlc_chord([apex + r + 0.5, 3.0, 0.0], [0, 0, 1.0], zt, zb, edges, tn, tp, r, False) == 0.0


def test_mask_oracle_at_pixel_centres():
    mask = [[0.1, 0.2, 0.3], [0.4, 0.5, 0.6]]
    center, widths = (0.0, 0.0, 10.0), (4.0, 6.0)
    for i in range(2):
        for j in range(3):
            cu = -2.0 + (i + 0.5) * 2.0
            cv = -3.0 + (j + 0.5) * 2.0
            got = R.mask_factor(mask, center, widths, (1, 0, 0), (0, 1, 0), (cu, cv, 0.0), (0, 0, 1.0))
            assert got == pytest.approx(mask[i][j], abs=1e-15)
    assert R.mask_factor(mask, center, widths, (1, 0, 0), (0, 1, 0), (2.5, 0, 0), (0, 0, 1.0)) == 0.0
    assert R.mask_factor(mask, center, widths, (1, 0, 0), (0, 1, 0), (0, 0, 0), (1.0, 0, 0)) == 0.0


# --------------------------------------------- product geometry against oracle


def _rays(n, rng, spread=12.0, z0=(-5.0, 5.0), cone=0.6):
    p = np.column_stack([rng.uniform(-spread, spread, n), rng.uniform(-spread, spread, n),
                         rng.uniform(*z0, n)])
    th = rng.uniform(0.0, cone, n)
    ph = rng.uniform(0.0, 2 * np.pi, n)
    d = np.column_stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    flip = rng.random(n) < 0.2
    d[flip] *= -1.0  # some rays run upstream
    d[:5] = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [0.6, 0, 0.8], [0, -0.6, 0.8]]  # axis-aligned cases
    p[:3, 2] = [24.0, 34.0, 0.0]
    return p, d


def _devices():
    return (
        JawPair("v", 20.0, 28.0, -5.0, 6.0, 0, 19.3),
        JawPair("u", 29.0, 36.0, -4.5, 3.5, 0, 19.3, focused=True),
        MLC(38.0, 45.0, tuple(np.linspace(-6.0, 6.0, 13)), tuple(np.linspace(-4.0, -1.0, 12)),
            tuple(np.linspace(-0.5, 3.0, 12)), 9.0, 0, 18.0),
        MLC(46.0, 52.0, (-3.0, -1.0, 0.5, 2.0), (-2.0, -2.5, -1.0), (1.0, 0.0, 2.5), 4.0, 0, 18.0,
            focused_sides=True),
    )


def _oracle_devices(devices):
    out = []
    for dv in devices:
        if isinstance(dv, JawPair):
            out.append(dict(kind="jaw", axis=dv.axis, z_top=dv.z_top, z_bottom=dv.z_bottom,
                            edge_neg=dv.edge_neg, edge_pos=dv.edge_pos, focused=dv.focused))
        else:
            out.append(dict(kind="mlc", z_top=dv.z_top, z_bottom=dv.z_bottom, edges=list(dv.leaf_edges_v),
                            tips_neg=list(dv.tips_neg), tips_pos=list(dv.tips_pos), r=dv.tip_radius,
                            focused_sides=dv.focused_sides))
    return out


@pytest.mark.parametrize("from_origin", [False, True])
def test_vectorised_chords_match_the_oracle(from_origin):
    rng = np.random.default_rng(5)
    frame = BeamFrame((0.3, -0.2, 0.0))
    stack = BeamLimitingStack(frame, _devices())
    p, d = _rays(600, rng, z0=(-5.0, 50.0) if from_origin else (-5.0, 5.0))
    got = stack.path_lengths(p, d, from_origin=from_origin)
    ora = np.array([R.device_chords(_oracle_devices(stack.devices), frame.origin, frame.basis.tolist(), p[i],
                                    d[i], from_origin) for i in range(len(p))])
    both_inf = np.isinf(got) & np.isinf(ora)
    np.testing.assert_allclose(np.where(both_inf, 0, got), np.where(both_inf, 0, ora), rtol=1e-12, atol=1e-12)
    assert (got > 0).sum() > 300  # the rays do hit the devices
    assert np.count_nonzero((got > 0) & (got < np.inf)) > 300


def test_rotating_frame_and_rays_together_changes_nothing():
    rng = np.random.default_rng(8)
    c, s = math.cos(0.7), math.sin(0.7)
    rot = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])  # rotation about y
    base = BeamFrame((1.0, 2.0, 3.0))
    turned = BeamFrame(tuple(rot @ np.array(base.origin)), tuple(rot @ [1, 0, 0]), tuple(rot @ [0, 1, 0]),
                       tuple(rot @ [0, 0, 1]))
    p, d = _rays(200, rng)
    p, d = p[5:] + np.array(base.origin), d[5:]  # rays parallel to a slab stay so only exactly
    a = BeamLimitingStack(base, _devices()).path_lengths(p, d)
    b = BeamLimitingStack(turned, _devices()).path_lengths(p @ rot.T, d @ rot.T)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-10)


def test_compiled_stack_layout():
    stack = BeamLimitingStack(BeamFrame((0, 0, 0)), _devices())
    cs = stack.compile()
    assert cs.kind.tolist() == [0, 0, 1, 1]
    assert cs.flags.tolist() == [1, 2, 0, 2]
    assert cs.n_pairs.tolist() == [0, 0, 12, 3]
    assert cs.offset.tolist() == [0, 0, 0, 13] and cs.pair_offset.tolist() == [0, 0, 0, 12]
    np.testing.assert_array_equal(cs.leaf_edges[13:], [-3.0, -1.0, 0.5, 2.0])
    assert cs.edge_neg[2] == 9.0 and cs.density.tolist() == [19.3, 19.3, 18.0, 18.0]


# ------------------------------------------------------------ weight models


def test_stack_transmission_at_table_nodes_and_against_the_oracle():
    from pyradmc_amd import AnalyticCrossSections

    xs = AnalyticCrossSections()
    stack = BeamLimitingStack(BeamFrame((0, 0, 0)), (JawPair("u", 20.0, 28.0, -3.0, 4.0, 0, 19.3),))
    model = StackTransmission(stack, xs, 6.0, n_table=64)
    nodes = np.exp(model.log_energies[[3, 20, 63]])
    p = np.array([[10.0, 0.0, 0.0]] * 3)
    d = np.array([[0.0, 0.0, 1.0]] * 3)
    expect = np.exp(-19.3 * xs.mu_total_bulk(nodes, 0) * 8.0)
    np.testing.assert_allclose(model.factors(nodes, p, d), expect, rtol=1e-12)
    # open rays transmit exactly 1
    assert model.factors(np.array([2.0]), np.array([[0.5, 0, 0]]), np.array([[0, 0, 1.0]]))[0] == 1.0
    rng = np.random.default_rng(2)
    full = BeamLimitingStack(BeamFrame((0.1, 0.0, 0.0)), _devices())
    m = StackTransmission(full, xs, 6.0)
    pr, dr = _rays(200, rng)
    e = rng.uniform(0.005, 7.0, 200)  # past both table ends: clamped
    f = m.factors(e, pr, dr)
    ora = [R.stack_factor(_oracle_devices(full.devices), full.frame.origin, full.frame.basis.tolist(),
                          m.log_energies.tolist(), m.log_mu.tolist(), m.densities.tolist(), e[i], pr[i], dr[i])
           for i in range(200)]
    np.testing.assert_allclose(f, ora, rtol=1e-12, atol=1e-300)
    assert f.max() == 1.0 and np.count_nonzero((f > 0.0) & (f < 0.01)) > 50


def test_mask_model_against_the_oracle():
    rng = np.random.default_rng(3)
    mask = rng.random((7, 5))
    mask[2:4, 1:3] = 0.25  # a plateau
    m = MaskTransmission(mask, (0.5, -0.5, 30.0), 7.0, 5.0, (1, 1, 0), (-1, 1, 0))
    p, d = _rays(300, rng, spread=6.0)
    got = m.factors(None, p, d)
    ora = [R.mask_factor(mask.tolist(), m.center.tolist(), m.widths, m.u_axis.tolist(), m.v_axis.tolist(),
                         p[i], d[i]) for i in range(300)]
    np.testing.assert_allclose(got, ora, rtol=1e-13, atol=1e-15)
    ones = MaskTransmission(np.ones((4, 4)), (0, 0, 30.0), 8.0, 8.0)
    f = ones.factors(None, np.array([[0.3, -1.7, 0], [9.0, 0, 0]]), np.array([[0, 0, 1.0], [0, 0, 1.0]]))
    assert f.tolist() == [1.0, 0.0]


def test_host_sampled_inner_gets_attenuated_weights():
    from pyradmc_amd import AnalyticCrossSections

    class Inner:
        max_energy = 2.0

        def sample_batch(self, seed, offset, n):
            rng = np.random.default_rng([seed, offset])
            return dict(particle_type=np.ones(n, np.int32), energy=np.full(n, 2.0, np.float32),
                        x=rng.uniform(-8, 8, n).astype(np.float32), y=np.zeros(n, np.float32),
                        z=np.zeros(n, np.float32), ux=np.zeros(n, np.float32), uy=np.zeros(n, np.float32),
                        uz=np.ones(n, np.float32), weight=np.full(n, 0.5, np.float32))

    stack = BeamLimitingStack(BeamFrame((0, 0, 0)), (JawPair("u", 20.0, 28.0, -3.0, 4.0, 0, 19.3),))
    src = CollimatedSource(Inner(), stack, AnalyticCrossSections())
    b = src.sample_batch(1, 0, 1000)
    x = b["x"]
    open_ = (x > -3.0) & (x < 4.0)
    assert np.all(b["weight"][open_] == np.float32(0.5))
    assert np.all(b["weight"][~open_] < 0.5 * 1e-3)
    masked = TransmissionMaskSource(Inner(), mask=np.full((2, 2), 0.5), plane_center=(0, 0, 10.0), width_u=4.0,
                                    width_v=4.0)
    mb = masked.sample_batch(1, 0, 1000)
    inside = np.abs(mb["x"]) <= 2.0
    assert np.all(mb["weight"][inside] == np.float32(0.25)) and np.all(mb["weight"][~inside] == 0.0)


@pytest.mark.parametrize("make,match", [
    (lambda: BeamFrame((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, -1)), "right-handed"),
    (lambda: BeamFrame((0, 0, 0), (1, 0, 0), (0.5, 1, 0)), "orthonormal"),
    (lambda: JawPair("w", 1, 2, 0, 1, 0, 1), "axis"),
    (lambda: JawPair("u", 1, 2, 1, 0, 0, 1), "edges cross"),
    (lambda: JawPair("u", 0, 2, 0, 1, 0, 1), "downstream"),
    (lambda: MLC(1, 9, (0, 1), (0,), (1,), 3.0, 0, 1), "half-height"),
    (lambda: MLC(1, 3, (1, 0), (0,), (1,), 3.0, 0, 1), "increasing"),
    (lambda: MLC(1, 3, (0, 1), (2,), (1,), 3.0, 0, 1), "cross"),
    (lambda: BeamLimitingStack(BeamFrame((0, 0, 0)), (JawPair("u", 5, 9, 0, 1, 0, 1),
                                                      JawPair("v", 8, 10, 0, 1, 0, 1))), "overlap"),
    (lambda: MaskTransmission(np.full((2, 2), 1.5), (0, 0, 0), 1, 1), "within 0..1"),
])
def test_validation(make, match):
    with pytest.raises(ValueError, match=match):
        make()


def test_projection_between_planes():
    assert project_between_planes(5.0, 100.0, 30.0) == pytest.approx(1.5)
    np.testing.assert_allclose(project_between_planes(np.array([2.0, -4.0]), 100.0, 50.0), [1.0, -2.0])
    with pytest.raises(ValueError):
        project_between_planes(1.0, 0.0, 1.0)


def test_engine_routing_of_weighted_sources():
    from pyradmc_amd import AnalyticCrossSections, BeamletGridSource, ParallelBeamSource
    from pyradmc_amd.collimation import CollimatedBeamletSource, TransmissionMaskBeamletSource
    from pyradmc_amd.engine import _in_kernel

    xs = AnalyticCrossSections()
    stack = BeamLimitingStack(BeamFrame((0, 0, -100.0)), (JawPair("u", 20.0, 28.0, -3.0, 4.0, 0, 19.3),))
    beam = ParallelBeamSource(2.0, -0.1, (0.0, 4.0), (0.0, 4.0))
    lattice = BeamletGridSource(2.0, -0.1, (0.0, 4.0), (0.0, 4.0), 2, 2)
    assert _in_kernel(CollimatedSource(beam, stack, xs))
    assert _in_kernel(CollimatedBeamletSource(lattice, stack, xs))
    assert _in_kernel(TransmissionMaskBeamletSource(lattice, mask=np.ones((2, 2)), plane_center=(2, 2, 0),
                                                    width_u=4, width_v=4))
    nested = TransmissionMaskSource(CollimatedSource(beam, stack, xs), mask=np.ones((2, 2)),
                                    plane_center=(2, 2, 0), width_u=4, width_v=4)
    assert not _in_kernel(nested)  # one device model per launch: nesting is host-sampled

    class Host:
        max_energy = 2.0

        def sample_batch(self, seed, offset, n):
            raise AssertionError

    assert not _in_kernel(CollimatedSource(Host(), stack, xs))
    with pytest.raises(TypeError, match="generated on the device"):
        nested.sample_batch(1, 0, 10)
