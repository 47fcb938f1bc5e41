"""GPU parity of the Kleefsman gauges, pressure taps, run record and
checkpoints (kleefsman_obstacle_3d.py:273-726, marin_probes.py:15-80).

Bars: the device gauges against the FP64 oracle (counts and maxima exact,
sums 1e-12) and the reference's values on a seeded cloud; FP32-compensated
within 1e-5; run_case in FP64 against the reference's own run at dx = 0.05
to t = 0.62 s (sample times exact, front / H2 / taps / loads to 1e-7 of
their scale, band counts within one particle at no more than three samples);
a run interrupted after a checkpoint and resumed ends with the uninterrupted
run's record.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from tests.gpu_util import n, scaled_err, t
from tests.test_oracle import gauge_cfg, gauge_kw

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tag", ["point", "disk"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_device_gauges_match_oracle_and_reference(golden, tag, dtype):
    from tensorsph_amd import kleefsman as kl
    g = golden("kleefsman_gauges")
    cfg = gauge_cfg()
    pos = t(g["cloud_pos"])
    vol, rho, p = (t(g[k], dtype) for k in ("cloud_vol", "cloud_rho", "cloud_p"))
    nodes = g[f"cloud_{tag}_nodes"]
    pv, sv = kl.tap_node_pressures(pos, p, rho, vol, cfg, nodes)
    f = lambda a: np.asarray(a).astype(np.float32 if dtype == torch.float32 else np.float64)
    o = orc.kleefsman_gauges(g["cloud_pos"], f(g["cloud_vol"]), f(g["cloud_rho"]), f(g["cloud_p"]),
                             nodes=nodes, **gauge_kw(cfg))
    bar = 1e-12 if dtype == torch.float64 else 1e-6
    assert scaled_err(pv, o["tap_p"]) < bar
    assert np.array_equal(sv, o["tap_support"])
    assert kl.slab_height(pos, cfg) == pytest.approx(o["slab"], rel=1e-15)
    assert kl.peak_height(pos, cfg) == o["peak"] and kl.bypass_count(pos, cfg) == o["bypass"]
    assert kl.centreline_height(pos, vol, cfg) == pytest.approx(o["h2"], rel=bar)
    if dtype == torch.float64:
        assert scaled_err(pv, g[f"cloud_{tag}_p"]) < 1e-12
        assert kl.centreline_height(pos, vol, cfg) == pytest.approx(float(g["cloud_h2"]), rel=1e-12)
        assert kl.slab_height(pos, cfg) == pytest.approx(float(g["cloud_slab"]), rel=1e-15)
    if tag == "disk":
        assert scaled_err(kl.disk_average(pv), g["cloud_disk_avg"] if dtype == torch.float64
                          else kl.disk_average(o["tap_p"])) < bar
        assert np.allclose(kl.tap_disk_nodes(cfg, kl.sensor_positions(cfg), 0.03), nodes,
                           rtol=0, atol=1e-15)


def _run(golden, **kw):
    from tensorsph_amd import kleefsman as kl
    samples = []
    res = kl.run_case(gauge_cfg(), on_sample=lambda i, s: samples.append(s), **kw)
    return res, samples


def test_run_case_matches_reference_run(golden):
    """The band gauges count particles, so a particle within round-off of a
    band edge can land on the other side: the slab and bypass counts may
    differ by one particle at no more than three samples (the peak height
    only at those samples); every other value is held to 1e-7 of its scale."""
    g = golden("kleefsman_gauges")
    cfg = gauge_cfg()
    res, samples = _run(golden)
    keys = list(g["run_row_keys"])
    rows = np.array([[float(s[k]) for k in keys] for s in samples])
    ref = g["run_rows"]
    assert rows.shape == ref.shape
    assert np.array_equal(rows[:, 0], ref[:, 0])                      # sample times
    bad = []
    one = cfg.dx ** 3 / (cfg.gauge_aperture * cfg.obstacle_width)
    c = keys.index("slab_height_m")
    off = np.abs(rows[:, c] - ref[:, c]) > 1e-12
    if off.sum() > 3 or np.abs(rows[:, c] - ref[:, c]).max() > one * (1 + 1e-9):
        bad.append(("slab", np.nonzero(off)[0].tolist()))
    c = keys.index("n_bypass")
    boff = rows[:, c] != ref[:, c]
    if boff.sum() > 3 or np.abs(rows[:, c] - ref[:, c]).max() > 1:
        bad.append(("bypass", np.nonzero(boff)[0].tolist()))
    c = keys.index("height_operator_m")
    hoff = np.abs(rows[:, c] - ref[:, c]) > 1e-7 * np.abs(ref[:, c]).max()
    if np.any(hoff & ~off):
        bad.append(("peak height", np.nonzero(hoff & ~off)[0].tolist()))
    for k in ("front_x_m", "H2_local_m"):
        c = keys.index(k)
        e = np.abs(rows[:, c] - ref[:, c]).max() / max(np.abs(ref[:, c]).max(), 1e-3)
        if e > 1e-7:
            bad.append((k, e))
    for name, key in (("obstacle_sph_force_N", "run_force"), ("obstacle_sph_impulse_Ns", "run_impulse"),
                      ("obstacle_sph_moment_Nm", "run_moment"), ("tap_pressure_Pa", "run_tap_p"),
                      ("disk_pressure_Pa", "run_disk_p")):
        got = np.array([s[name] for s in samples])
        e = np.abs(got - g[key]).max() / max(np.abs(g[key]).max(), 1.0)
        if e > 1e-7:
            bad.append((name, e))
    sup = np.array([s["tap_support"] for s in samples])
    if np.abs(sup - g["run_tap_support"]).max() > 1:
        bad.append(("tap support", np.abs(sup - g["run_tap_support"]).max()))
    d = res.to_dict()
    for k in ("steps", "boundary_escapes", "obstacle_clamps"):
        if d[k] != float(g["summary_" + k]):
            bad.append((k, d[k], float(g["summary_" + k])))
    if abs(d["max_bypass_particles"] - float(g["summary_max_bypass_particles"])) > 1:
        bad.append(("max_bypass_particles", d["max_bypass_particles"]))
    for k in ("simulated_time_s", "peak_height_m", "peak_time_s", "peak_height_operator_m",
              "max_speed", "max_density_compression", "local_h2_peak_m", "arrival_time_s",
              "local_h2_arrival_s"):
        want = float(g["summary_" + k])
        got = d[k]
        if (got is None) != np.isnan(want) or (got is not None and abs(got - want) > 1e-7 * abs(want) + 1e-12):
            bad.append((k, got, want))
    for k, key in (("obstacle_sph_impulse_Ns", "summary_sph_impulse"),
                   ("obstacle_clamp_impulse_Ns", "summary_clamp_impulse")):
        if not np.allclose(d[k], g[key], rtol=1e-7, atol=1e-9):
            bad.append((k, d[k], g[key].tolist()))
    assert not bad, bad
    assert res.operator_consistency()["height_column_admissible"]


def test_run_case_resumes_from_its_checkpoint(tmp_path, golden):
    """A run that dies after a checkpoint resumes from it and ends with the
    uninterrupted run's record; the file is gone once t_end is reached."""
    import dataclasses
    from tensorsph_amd import kleefsman as kl
    cfg = dataclasses.replace(gauge_cfg(), t_end=0.15)
    full = kl.run_case(cfg)
    ck = tmp_path / "state.pt"

    class Crash(Exception):
        pass

    def die(i, s):
        if s["t_s"] > 0.1:
            raise Crash()

    with pytest.raises(Crash):
        kl.run_case(cfg, on_sample=die, checkpoint_path=ck, checkpoint_every=100)
    assert ck.exists()
    part = kl.load_checkpoint(ck, cfg)
    assert part["step"] % 100 == 0
    res = kl.run_case(cfg, checkpoint_path=ck, checkpoint_every=100)
    assert res.resumed and not ck.exists()
    assert res.times == full.times and res.steps == full.steps
    for a, b in ((res.front, full.front), (res.local_h2, full.local_h2), (res.slab, full.slab)):
        assert np.abs(np.array(a) - np.array(b)).max() <= 1e-9
    assert res.obstacle_clamps == full.obstacle_clamps
    assert res.boundary_escapes == full.boundary_escapes
    assert np.allclose(res.sph_impulse, full.sph_impulse, rtol=1e-9, atol=1e-12)
    assert res.max_speed == pytest.approx(full.max_speed, rel=1e-9)


def test_gauges_need_a_step_and_valid_nodes():
    from tensorsph_amd import kleefsman as kl
    cfg = gauge_cfg()
    sim, _, _ = kl.make_simulation(cfg, "fp64")
    try:
        spec = kl.gauge_spec(cfg, kl.sensor_positions(cfg)[:, None, :])
        with pytest.raises(ValueError):
            sim.gauges(spec)                       # no evaluated step yet
        sim.step(1)
        row = sim.gauges(spec)
        assert row[2] > 0.0                        # the front
    finally:
        sim.close()
    with pytest.raises(ValueError):
        kl.gauge_spec(cfg, np.zeros((8, 3, 3)))
    with pytest.raises(ValueError):
        kl.tap_disk_nodes(cfg, kl.sensor_positions(cfg), 0.2)   # beyond the face


def _adaptive_cfg():
    import dataclasses
    return dataclasses.replace(gauge_cfg(), t_end=0.45, pressure_disk_diameter=0.0, adaptive_dt=True,
                               max_courant=0.085, max_steps=4000)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_run_case_adaptive_step_matches_reference_run(golden, precision):
    """run_case with adaptive_dt (kleefsman_obstacle_3d.py:644-713): the step
    shrinks with the flow speed and is not shortened onto t_end; every gauge
    column, the loads, the clamp record and the run summary against the
    reference's own run (FP64: 1e-7 of the scale; FP32-compensated: 1e-5, the
    run's integrated impulse 2e-5; band counts within one particle)."""
    from tensorsph_amd import kleefsman as kl
    g = golden("kleefsman_adaptive")
    cfg = _adaptive_cfg()
    samples = []
    res = kl.run_case(cfg, on_sample=lambda i, s: samples.append(s), precision=precision)
    bar = 1e-7 if precision == "fp64" else 1e-5
    keys = list(g["row_keys"])
    rows = np.array([[float(s[k]) for k in keys] for s in samples])
    assert rows.shape == g["rows"].shape
    assert np.allclose(rows[:, 0], g["rows"][:, 0], rtol=1e-12 if precision == "fp64" else 1e-6,
                       atol=0)                                               # sample times
    for k in ("front_x_m", "H2_local_m", "slab_height_m", "height_operator_m"):
        c = keys.index(k)
        err = np.abs(rows[:, c] - g["rows"][:, c])
        scale = max(np.abs(g["rows"][:, c]).max(), 1e-3)
        if k in ("front_x_m", "H2_local_m"):
            assert err.max() <= bar * scale, k
        else:
            # band gauges count particles: a particle at a band edge may fall
            # either side, so allow a lattice spacing at no more than 3 samples
            assert (err > bar * scale).sum() <= 3 and err.max() <= cfg.dx + bar * scale, k
    c = keys.index("n_bypass")
    assert np.abs(rows[:, c] - g["rows"][:, c]).max() <= 1
    for name, key in (("obstacle_sph_force_N", "force"), ("obstacle_sph_impulse_Ns", "impulse")):
        got = np.array([s[name] for s in samples])
        assert np.abs(got - g[key]).max() <= bar * max(np.abs(g[key]).max(), 1.0), name
    d = res.to_dict()
    assert d["steps"] == int(g["steps"])
    assert d["simulated_time_s"] == pytest.approx(float(g["simulated_time"]), rel=1e-12 if bar < 1e-6 else 1e-6)
    # the impulse integrates 2,059 steps of force: FP32-compensated rounding of
    # the forces (each within 1e-5 above) and of the speeds that set each step
    # size adds up to 1.5e-5 of it
    assert scaled_err(d["obstacle_sph_impulse_Ns"], g["sph_impulse"]) <= (bar if precision == "fp64"
                                                                          else 2e-5)
    assert d["max_speed"] == pytest.approx(float(g["max_speed"]), rel=bar)
    assert d["max_density_compression"] == pytest.approx(float(g["compression"]), rel=10 * bar)
    assert d["obstacle_clamps"] == int(g["clamps"])
    assert np.abs(np.array(d["obstacle_clamp_impulse_Ns"]) - g["clamp_impulse"]).max() <= \
        bar * max(np.abs(g["clamp_impulse"]).max(), 1.0)
    assert d["boundary_escapes"] == int(g["escapes"])


def test_run_case_adaptive_resumes_from_its_checkpoint(tmp_path):
    """An adaptive run that dies after a checkpoint resumes from the saved
    positions, velocities, time and step-size variable and ends with the
    uninterrupted run's record: the same steps and sample count, times to
    1e-12, gauges, impulses and peaks to 1e-9."""
    from tensorsph_amd import kleefsman as kl
    cfg = _adaptive_cfg()
    full = kl.run_case(cfg)
    ck = tmp_path / "adaptive.pt"

    class Crash(Exception):
        pass

    def die(i, s):
        if s["t_s"] > 0.3:
            raise Crash()

    with pytest.raises(Crash):
        kl.run_case(cfg, on_sample=die, checkpoint_path=ck, checkpoint_every=150)
    assert ck.exists()
    part = kl.load_checkpoint(ck, cfg)
    assert part["step"] % 150 == 0 and part["step"] > 0
    assert 0.0 < part["dt"] <= cfg.dt
    res = kl.run_case(cfg, checkpoint_path=ck, checkpoint_every=150)
    assert res.resumed and not ck.exists()
    assert res.steps == full.steps and len(res.times) == len(full.times)
    assert np.allclose(res.times, full.times, rtol=1e-12, atol=0)
    for a, b in ((res.front, full.front), (res.local_h2, full.local_h2), (res.slab, full.slab)):
        assert np.abs(np.array(a) - np.array(b)).max() <= 1e-9
    assert res.obstacle_clamps == full.obstacle_clamps
    assert res.boundary_escapes == full.boundary_escapes
    assert np.allclose(res.sph_impulse, full.sph_impulse, rtol=1e-9, atol=1e-12)
    assert res.max_speed == pytest.approx(full.max_speed, rel=1e-9)
    assert res.simulated_time == pytest.approx(full.simulated_time, rel=1e-12)
