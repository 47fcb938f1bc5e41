"""GPU tests of the served δ⁺ batch engine.

Bars: a run in the batch engine with index-ordered lists is bitwise the
single-step API (delta_plus_step with fresh pair lists and the same adaptive
dt) over 20 steps.  With cell-ordered lists (the serving default) the sums
run in another order; δ⁺'s thresholds (lambda bands, the empty-cap test) then
flip for a few particles and the difference spreads to their neighbours, so
the bar is a median within 1e-11 and a maximum within 1e-6 in position.  A
run's final state is bitwise the
same alone and in a mix admitted at different times; a served stream with
too little capacity for all requests completes every one, each equal to its
alone run.
"""
import numpy as np
import pytest
import torch

from tests.gpu_util import n, scaled_err

pytestmark = pytest.mark.gpu


def _small(kind, n_target, steps, **kw):
    from tensorsph_amd.serving import drop_case, impact_case
    if kind == "drop":
        return drop_case(n_target, radius=kw.get("radius", 0.1), strain=kw.get("strain", 5.0),
                         viscosity=kw.get("nu", 1e-4), c_mult=kw.get("mult", 12.0), steps=steps)
    return impact_case(n_target, radius=kw.get("radius", 0.06), speed=kw.get("speed", 1.0),
                       viscosity=kw.get("nu", 1e-5), c_mult=kw.get("mult", 15.0), steps=steps)


def _alone(case, sorted_lists=False):
    from tensorsph_amd.serving import BatchEngine
    eng = BatchEngine(case.n + 16, case.n_wall + 16, sims=2, sorted_lists=sorted_lists)
    slot = eng.admit(case)
    while not eng.poll():
        eng.step(1)
    out = eng.read(slot)
    eng.close()
    return out


@pytest.mark.parametrize("kind", ["drop", "impact"])
def test_batch_tracks_single_step_api(kind):
    from tensorsph_amd.delta_plus import delta_plus_step
    case = _small(kind, 1500 if kind == "drop" else 3000, steps=20)
    got = _alone(case, sorted_lists=True)
    dev = torch.device("cuda:0")
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to(dev)
    pos, vel, rho, mass = f(case.pos), f(case.vel), f(case.rho), f(case.mass)
    wall, wv, wn = f(case.wall), f(case.wall_vel), f(case.wall_normal)
    for _ in range(case.steps):
        vmax = float(torch.sqrt((vel * vel).sum(1)).max())
        dt = case.courant * 2.0 * case.dx / (case.sound_speed + vmax)
        st = delta_plus_step(pos, vel, rho, mass, dx=case.dx, sound_speed=case.sound_speed,
                             reference_speed=case.reference_speed, dt=dt, rho0=case.rho0,
                             gravity=case.gravity, viscosity=case.viscosity, diffusion=case.diffusion,
                             alpha=case.alpha, wall=wall if case.n_wall else None,
                             wall_velocity=wv if case.n_wall else None,
                             wall_normal=wn if case.n_wall else None)
        pos, vel, rho = st.position, st.velocity, st.density
    assert got["steps"] == case.steps
    # same lists in the same order, same adaptive dt: bitwise the same run
    assert torch.equal(got["position"], pos)
    assert torch.equal(got["velocity"], vel)
    assert torch.equal(got["density"], rho)
    # cell-ordered lists sum in another order: a few particles may take the
    # other side of a δ⁺ threshold; nearly all must agree to round-off
    unsorted = _alone(case)
    assert unsorted["steps"] == case.steps
    d = np.abs(n(unsorted["position"]) - n(pos)).max(1)
    assert np.median(d) < 1e-11
    assert d.max() < 1e-6


def test_single_case_api_loop_matches_batch():
    """The reference-style driver loop (adaptive dt, pair cache with a skin,
    delta_plus_step) gives the batch engine's run bit for bit: skin pairs fail
    the strict support test and the sums keep ascending neighbour order."""
    from tensorsph_amd.serving import run_single_case_api
    case = _small("impact", 2000, steps=30)
    got = run_single_case_api(case)
    ref = _alone(case, sorted_lists=True)
    assert ref["steps"] == case.steps and got["rebuilds"] >= 1
    for key in ("position", "velocity", "density"):
        assert torch.equal(got[key], ref[key]), key


def test_alone_and_in_a_mix_are_bitwise_equal():
    from tensorsph_amd.serving import BatchEngine
    a = _small("impact", 2500, steps=25)
    b = _small("drop", 1800, steps=40, radius=0.08, strain=3.0)
    c = _small("impact", 1200, steps=10, speed=0.7)
    d = _small("drop", 900, steps=60)
    alone = _alone(a)
    assert alone["steps"] == a.steps
    eng = BatchEngine(a.n + b.n + c.n + d.n + 64, a.n_wall + c.n_wall + 64, sims=8)
    sb = eng.admit(b)
    sc = eng.admit(c)
    eng.step(3)
    sa = eng.admit(a)                     # joins a running mix
    eng.step(4)
    sd = eng.admit(d)
    finished = {}
    while sa not in finished:
        eng.step(1)
        for slot, state in eng.poll():
            finished[slot] = (state, eng.read(slot))
            eng.release(slot)
    state, mixed = finished[sa]
    assert state == 2 and mixed["steps"] == a.steps
    for key in ("position", "velocity", "density"):
        assert torch.equal(mixed[key], alone[key]), key
    eng.close()


def test_served_stream_with_queueing_matches_alone_runs():
    from tensorsph_amd.serving import BatchEngine, build_case, request_stream, serve
    reqs = request_stream(count=10, seed=7, steps=15, n_min=800, n_max=4000)
    cases = [build_case(k, kw) for k, kw in reqs]
    # room for about a third of the stream at once: requests must wait
    fluid = sum(c.n for c in cases) // 3 + max(c.n for c in cases)
    wall = max(sum(c.n_wall for c in cases) // 3, max(c.n_wall for c in cases)) + 16
    eng = BatchEngine(fluid, wall, sims=6)
    results = {}
    rep = serve(eng, reqs, rate_per_s=1e4, on_result=lambda i, s, r: results.__setitem__(i, (s, r)))
    eng.close()
    summary = rep.summary()
    assert summary["sims"] == 10 and summary["failed"] == 0 and len(results) == 10
    for i in (0, 5, 9):
        state, r = results[i]
        ref = _alone(cases[i])
        assert state == 2 and r["steps"] == cases[i].steps == ref["steps"]
        for key in ("position", "velocity", "density"):
            assert torch.equal(r[key], ref[key]), (i, key)


def test_admission_limits():
    from tensorsph_amd.serving import BatchEngine
    case = _small("drop", 1500, steps=5)
    eng = BatchEngine(case.n - 1, 8, sims=2)
    assert eng.admit(case) is None             # does not fit: the scheduler waits
    eng.close()
    eng = BatchEngine(case.n * 2, 8, sims=1)
    assert eng.admit(case) is not None
    assert eng.admit(case) is None             # no free slot
    case.courant = 0.5
    eng2 = BatchEngine(case.n * 2, 8, sims=2)
    with pytest.raises(ValueError):
        eng2.admit(case)
    eng.close()
    eng2.close()


def _mix_final(launch):
    """Three runs admitted together, stepped to completion on one engine with
    the given launch shape; final states by admission order."""
    from tensorsph_amd.serving import BatchEngine
    cases = [_small("drop", 1200, steps=12), _small("impact", 2500, steps=10),
             _small("drop", 800, steps=15, radius=0.08)]
    eng = BatchEngine(sum(c.n for c in cases) + 64, sum(c.n_wall for c in cases) + 64, sims=4)
    try:
        if launch is not None:
            eng.set_launch(*launch)
        slots = [eng.admit(c) for c in cases]
        out, left = {}, set(slots)
        while left:
            eng.step(1)
            for s, _ in eng.poll():
                if s in left:
                    out[s] = eng.read(s)
                    eng.release(s)
                    left.discard(s)
        return [out[s] for s in slots]
    finally:
        eng.close()


@pytest.mark.parametrize("launch", [(128, 256, True), (64, 128, False), (192, 64, True)])
def test_launch_shapes_give_the_same_bits(launch):
    """tsph_batch_launch changes workgroup sizes and the block-to-XCD mapping
    only: every run's final state is bitwise the default shape's."""
    print(f"launch shape {launch}", flush=True)
    ref = _mix_final(None)
    got = _mix_final(launch)
    for r, g in zip(ref, got):
        assert r["steps"] == g["steps"]
        for k in ("position", "velocity", "density"):
            assert torch.equal(r[k], g[k]), k


def test_autotune_keeps_the_fastest_candidate():
    from tensorsph_amd.serving import BatchEngine
    cases = [_small("drop", 1500, steps=200), _small("impact", 3000, steps=200)]
    eng = BatchEngine(sum(c.n for c in cases) + 64, sum(c.n_wall for c in cases) + 64, sims=4)
    try:
        for c in cases:
            assert eng.admit(c) is not None
        cands = [(256, 256, True), (128, 128, False)]
        res = eng.autotune(candidates=cands, steps=2)
        assert res["best"] in cands and eng.launch == res["best"]
        assert len(res["ms"]) == 2 and all(v > 0 for v in res["ms"].values())
        assert res["steps"] == 2 * 3 and eng.stats()["steps"] == 6
    finally:
        eng.close()


def test_launch_shape_rejects_other_sizes():
    from tensorsph_amd.serving import BatchEngine
    eng = BatchEngine(4096, 512, sims=2)
    try:
        with pytest.raises(ValueError):
            eng.set_launch(96, 256, True)
        with pytest.raises(ValueError):
            eng.set_launch(256, 512, True)
    finally:
        eng.close()
