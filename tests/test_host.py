"""Host-side helpers of the drivers (no GPU): the Ritter dam-break reference
(dam_break.py:179-230) and the DamBreakRun derived values (dam_break.py:
271-284).

The Ritter checks are the solution's own landmarks, as the reference's
tests/test_ritter_reference.py pins them: front at 2 sqrt(gH) t, depth H at
the rarefaction tail, 4H/9 at the dam and 0 at the front, the Riemann
invariant u + 2 sqrt(g h) = 2 sqrt(g H) across the fan, self-similarity, the
water area ahead of the dam (4/27) H x_front, and Z/a = 2 tau for H = 2a.
"""
import math

import pytest
import torch

from tensorsph_amd.dam_break import (DamBreakConfig, DamBreakRun, ritter_front, ritter_front_over_a,
                                     ritter_height_profile)

G, A = 9.81, 0.2
H = 2.0 * A


@pytest.fixture(scope="module")
def cfg():
    return DamBreakConfig(dx=0.02, column_width=A, column_height=H, pool_width=2.0, gravity=G)


def test_front_travels_at_twice_the_wave_speed(cfg):
    c = math.sqrt(G * H)
    for t in (0.05, 0.1, 0.25, 0.5):
        assert ritter_front(t, cfg) == pytest.approx(2.0 * c * t, rel=1e-15)
        assert ritter_height_profile(ritter_front(t, cfg), t, cfg) == pytest.approx(0.0, abs=1e-15)


def test_landmarks_and_riemann_invariant(cfg):
    c = math.sqrt(G * H)
    t = 0.3
    assert ritter_height_profile(-c * t, t, cfg) == pytest.approx(H, rel=1e-14)
    assert ritter_height_profile(0.0, t, cfg) == pytest.approx(4.0 * H / 9.0, rel=1e-14)
    assert ritter_height_profile(-2.0 * c * t, t, cfg) == H          # undisturbed reservoir
    assert ritter_height_profile(3.0 * c * t, t, cfg) == 0.0          # dry bed
    assert ritter_height_profile(0.1, 0.0, cfg) is None               # before release
    for k in range(11):
        x = -c * t + k * (3.0 * c * t) / 10.0
        h = ritter_height_profile(x, t, cfg)
        u = (2.0 / 3.0) * (c + x / t)
        assert u + 2.0 * math.sqrt(G * h) == pytest.approx(2.0 * c, rel=1e-12)


def test_self_similar_and_area_ahead_of_dam(cfg):
    c = math.sqrt(G * H)
    for lam in (0.5, 2.0, 3.7):
        for x, t in ((0.05, 0.1), (-0.1, 0.2), (0.3, 0.25)):
            assert ritter_height_profile(lam * x, lam * t, cfg) == pytest.approx(
                ritter_height_profile(x, t, cfg), rel=1e-13)
    t = 0.2
    front = 2.0 * c * t
    n = 4000                                   # Simpson over [0, front]
    hstep = front / n
    area = ritter_height_profile(0.0, t, cfg) + ritter_height_profile(front, t, cfg)
    for k in range(1, n):
        area += (4 if k % 2 else 2) * ritter_height_profile(k * hstep, t, cfg)
    area *= hstep / 3.0
    assert area == pytest.approx(4.0 / 27.0 * H * front, rel=1e-12)


def test_z_over_a_is_two_tau(cfg):
    for tau in (0.5, 1.0, 2.5):
        t = tau * math.sqrt(A / (2.0 * G))
        assert ritter_front(t, cfg) / A == pytest.approx(ritter_front_over_a(tau), rel=1e-14)
        assert ritter_front_over_a(tau) == 2.0 * tau


def test_run_record_derived_values(cfg):
    run = DamBreakRun(times=[0.0, 0.1, 0.2], front=[A, 0.3, 0.5], kinetic_energy=[0.0] * 3,
                      potential_energy=[0.0] * 3, mass=4.0, escaped=0, rebuilds=1, max_speed=0.0,
                      max_density_compression=0.0, compression_before_wall_impact=0.0,
                      wall_impact_time=None, initial_mass=4.0, final_mass=4.0,
                      final_positions=torch.zeros((0, 2)))
    assert run.mass_drift == 0.0
    tau, za = run.dimensionless(cfg)
    rate = math.sqrt(2.0 * G / A)
    assert tau == pytest.approx([0.0, 0.1 * rate, 0.2 * rate], rel=1e-15)
    assert za == pytest.approx([1.0, 1.5, 2.5], rel=1e-15)


def _prefilter_band(rc2, dim, extent, box=None):
    """The undecided band of the FP32 list prefilter (build_slot_lists in
    csrc/tsph_cells.hpp): twice the bound on |d2_fp32 - d2|."""
    rc = math.sqrt(rc2)
    e = 2.0 * math.ldexp(extent, -23) + math.ldexp(1.01 * rc, -24)
    if box is not None:
        e += math.ldexp(box, -22)
    return 2.0 * (2.0 * math.sqrt(dim) * rc * e + dim * e * e + 4.4 * math.ldexp(rc2, -24))


@pytest.mark.parametrize("dim,extent,rc,box", [(3, 3.22, 0.015, None), (2, 4.0, 0.0033, None),
                                                (3, 1.0, 3.25 / 128, 1.0), (3, 60.0, 0.02, None),
                                                (2, 1.0e3, 0.5, None)])
def test_list_prefilter_band_holds_the_fp32_error(dim, extent, rc, box):
    """Outside the band the FP32 test decides as the FP64 one does.  Pairs are
    placed around the build radius (within a few band widths and far from it)
    anywhere in the extent, the FP32 distance formed as the kernel forms it."""
    import numpy as np
    rng = np.random.default_rng(20261204)
    rc2 = rc * rc
    band = _prefilter_band(rc2, dim, extent, box)
    lo2 = np.nextafter(np.float32(rc2 - band), np.float32(0))
    hi2 = np.nextafter(np.float32(rc2 + band), np.float32(np.inf))
    n = 200_000
    x0 = rng.uniform(-extent, extent, dim) * (0.0 if box else 1.0)
    lo, hi = (0.0, box) if box else (x0, x0 + extent)
    xi = rng.uniform(lo, hi, (n, dim))
    u = rng.normal(size=(n, dim))
    u /= np.linalg.norm(u, axis=1)[:, None]
    r = rc * (1.0 + rng.choice([-1, 1], n) * np.concatenate(
        [rng.uniform(0, 8 * band / rc2, n // 2), rng.uniform(0, 0.5, n - n // 2)]))
    xj = xi + u * r[:, None]
    if box:
        xj = np.mod(xj, box)
    else:
        xj = np.clip(xj, lo, hi)
    fi = (xi - xi[0]).astype(np.float32)          # FP32 rows relative to row 0
    fj = (xj - xi[0]).astype(np.float32)
    df = fi - fj
    d = xi - xj
    if box:
        bf, ibf = np.float32(box), np.float32(1.0) / np.float32(box)
        df = (df - bf * np.rint(df * ibf)).astype(np.float32)
        d = d - box * np.rint(d / box)
    d2f = (df * df).sum(axis=1, dtype=np.float32)
    d2 = (d * d).sum(axis=1)
    assert band < 0.01 * rc2
    assert not np.any((d2f < lo2) & (d2 > rc2))
    assert not np.any((d2f > hi2) & (d2 <= rc2))
    # the band is thin: of the pairs placed within 8 band widths of the build
    # radius (half of them) about one in eight falls inside it
    assert np.mean((d2f >= lo2) & (d2f <= hi2)) < 0.1


def test_case_builders_match_reference_lattices(golden):
    """cases.py against the reference's own lattices (lattices.npz: make_particles,
    wall_lattice with and without top=, make_column, make_column_3d,
    wall_lattice_3d, obstacle_lattice_3d, lattice_mass_unit and the 3-D
    lattice sum), element for element."""
    import numpy as np
    from tensorsph_amd import cases
    from tensorsph_amd.config import Config, DamBreakConfig, DamBreak3DConfig, Kleefsman3DConfig
    g = golden("lattices")
    cfg1 = Config(dx=0.005, pool_width=1.0, pool_depth=0.2, drop_radius=0.05, drop_center_y=0.35,
                  viscosity=1e-6, sound_speed=10.0 * math.sqrt(2 * 9.81 * 0.2), dt=9.84e-5)
    pos, mass, n_pool = cases.make_particles(cfg1)
    assert np.array_equal(pos, g["impact_pos"])
    assert np.abs(mass - g["impact_mass"]).max() <= 1e-15 * g["impact_mass"].max()
    assert np.array_equal(cases.wall_lattice(cfg1, top=0.6), g["impact_wall_top06"])
    assert np.array_equal(cases.wall_lattice(cfg1), g["impact_wall_default"])
    assert cases.lattice_mass_unit(cfg1) == pytest.approx(float(g["impact_mass_unit"]), rel=1e-15)
    assert cases.lattice_kernel_sum(1.3 * 0.005, 0.005) == pytest.approx(float(g["lattice_sum_13"]),
                                                                          rel=1e-14)
    cfg2 = DamBreakConfig(dx=0.05, tank_height=3.0, column_width=1.0, column_height=2.0,
                          pool_width=4.0, sound_speed=10.0 * math.sqrt(2 * 9.81 * 2.0))
    pos, mass = cases.make_column(cfg2)
    assert np.array_equal(pos, g["dam_pos"])
    assert np.abs(mass - g["dam_mass"]).max() <= 1e-15 * g["dam_mass"].max()
    assert np.array_equal(cases.wall_lattice(cfg2, top=cfg2.tank_top()), g["dam_wall"])
    cfg3 = DamBreak3DConfig(dx=0.05, tank_height=0.6, tank_depth=0.3, column_width=0.15,
                            column_depth=0.15, column_height=0.3, pool_width=1.0)
    pos, mass = cases.make_column_3d(cfg3)
    assert np.array_equal(pos, g["dam3_pos"])
    assert np.abs(mass - g["dam3_mass"]).max() <= 1e-15 * g["dam3_mass"].max()
    assert np.array_equal(cases.wall_lattice_3d(cfg3), g["dam3_wall"])
    assert cases.lattice_kernel_sum_3d(0.1, 0.13) == pytest.approx(float(g["k3_lattice_sum"]),
                                                                    rel=1e-14)
    assert np.array_equal(cases.obstacle_lattice_3d(Kleefsman3DConfig(dx=0.04)), g["kl_obstacle"])
