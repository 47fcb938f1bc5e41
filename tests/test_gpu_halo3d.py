"""GPU checks of the 3-D halo layout (tsph_sim_configure layout 2 on the 3-D
resident stepper): the density and force sweeps of each 256-row workgroup
evaluated from its neighbourhood rows (the 27 cells around the tile's cells)
staged in LDS, the lists built from the same staged positions.  The lists
name the same rows in the same order as the cell walk's, so every run must
give the same bits as the plain slot-list layout: state, box loads, clamp
and meter rows, across list rebuilds, in FP64 and FP32-compensated.
"""
import numpy as np
import pytest

from tests.gpu_util import n

pytestmark = pytest.mark.gpu

KEYS = ("position", "velocity", "density", "pressure", "acceleration")


def _cfg(**kw):
    from tensorsph_amd.config import Kleefsman3DConfig
    base = dict(dx=0.04, obstacle_x=1.25, sound_speed=20.0, dt=2.5e-4, viscosity=1e-6)
    base.update(kw)
    return Kleefsman3DConfig(**base)


def _run(cfg, precision, halo, steps, meter_every=10, skin_factor=None):
    import dataclasses
    from tensorsph_amd.kleefsman import make_simulation
    if skin_factor is not None:
        cfg = dataclasses.replace(cfg, pair_skin_factor=skin_factor)
    sim, nf, nw = make_simulation(cfg, precision)
    try:
        sim.configure(tiles=False, halo=halo)
        sim.step(steps, meter_every=meter_every)
        return {"state": sim.read(), "loads": sim.loads(), "meters": sim.meters(),
                "layout": sim.layout(), "stats": sim.stats(), "nf": nf}
    finally:
        sim.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_halo3d_gives_the_bits_of_the_lists(precision):
    cfg = _cfg()
    print(f"3-D halo vs lists: {precision}, dx {cfg.dx}", flush=True)
    a = _run(cfg, precision, False, 150, skin_factor=0.1)
    b = _run(cfg, precision, True, 150, skin_factor=0.1)
    lay = b["layout"]
    assert a["layout"] == {"layout": 0, "halo_tiles": 0, "tiles": 0, "halo_max": 0}
    assert lay["layout"] == 2 and lay["tiles"] == (b["nf"] + 255) // 256
    print(f"  layout {lay}", flush=True)
    # a small tank: most tiles reach the walls, whose rows join their halos
    assert lay["halo_tiles"] > 0
    assert 256 < lay["halo_max"] <= 2816
    assert b["stats"]["rebuilds"] >= 3 and b["stats"]["rebuilds"] == a["stats"]["rebuilds"]
    for k in KEYS:
        assert np.array_equal(n(a["state"][k]), n(b["state"][k])), k
    assert np.array_equal(a["loads"], b["loads"])
    assert np.array_equal(a["meters"], b["meters"])


def test_halo3d_with_density_diffusion():
    cfg = _cfg(density_diffusion=0.1)
    a = _run(cfg, "fp32", False, 60)
    b = _run(cfg, "fp32", True, 60)
    for k in KEYS:
        assert np.array_equal(n(a["state"][k]), n(b["state"][k])), k
