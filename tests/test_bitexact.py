"""The device's FP32 arithmetic is fixed by the source: contraction is off
everywhere and every fused multiply-add is an explicit fmaf, so the HIP
kernels -- the fused step, the two sweep kernels and the batched ensemble --
and the CPU emulator of the same node arithmetic (tests/native/line_emul.cpp)
must agree bit for bit, also where the flow is chaotic at rounding level
(frictionless thin films on a steep coast, where any change of rounding
changes the run-up after ~1,000 steps)."""
import numpy as np
import pytest

import synth
from emul import EmulGrid

pytestmark = pytest.mark.gpu

NAMES = ("cel", "xvel", "yvel", "hmax", "velmax")


def _cases():
    chaotic = synth.shelf_coast(nx=300, ny=220, seed=131, dx=400.0, source_amp=7.9,
                                source_len=60e3, steps=1500)
    fric = synth.shelf_coast(nx=260, ny=200, seed=3, dx=600.0, source_amp=5.0, source_len=80e3,
                             steps=600, ground=synth.ENSEMBLE_GROUND, crough=synth.ENSEMBLE_CROUGH)
    sph = synth.coastal(nx=250, ny=190, seed=4, cartesian=False, crough=0.0025, steps=300)
    return [("chaotic", chaotic), ("friction", fric), ("spherical", sph)]


def _state(g):
    return g.download() + g.download_max()


def _same(a, b, what):
    """Bit-identical, except that a NaN equals any NaN: the host's invalid
    operations return the negative default NaN, the GPU's the positive one
    (the spherical case's explicit friction blows up in thin films after
    ~270 steps, on both, in the same cells)."""
    for x, y, name in zip(a, b, NAMES):
        nan = np.isnan(x) & np.isnan(y)
        bad = int(((x.view(np.int64) != y.view(np.int64)) & ~nan).sum())
        assert bad == 0, f"{what} {name}: {bad} cells differ"


@pytest.mark.parametrize("name,case", _cases(), ids=[c[0] for c in _cases()])
def test_device_equals_emulator(gpu_lib, name, case):
    s1, s2, zeta = case.metrics()
    kw = dict(dt=case.dt, ground=case.ground, crough=case.crough)
    em = EmulGrid(case.dep, s1, s2, zeta, **kw)
    em.upload(case.cel, case.xvel, case.yvel)
    em.run(case.steps)
    want = _state(em)
    grids = []
    for mode in ("fused", "unfused", "ensemble"):
        g = gpu_lib.Grid(case.dep, s1, s2, zeta, **kw)
        g.upload(case.cel, case.xvel, case.yvel)
        if mode == "unfused":
            g.set_fused(False)
        if mode == "ensemble":
            e = gpu_lib.Ensemble([g])
            e.run(case.steps)
            e.sync()
            e.close()
        else:
            g.run(case.steps)
        _same(_state(g), want, f"{name} {mode}")
        grids.append(g)
    if name == "chaotic":
        ru = synth.max_runup(want[3], case.dep, em.celmin)
        assert ru > 5.0, "the chaotic case must flood land"
    for g in grids:
        g.close()
