"""cliffs_run_gauges_noalloc (the gauge-series entry point that allocates no
device memory: its gauge table and series sit in a library-wide device
scratch) against cliffs_run_gauges on twin grids: the series and the final
states bit-identical, on two calls in a row -- the second one smaller, so it
runs over what the first left in the scratch -- and a request beyond the
scratch refused before any step."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def _grid(gpu_lib, case):
    s1, s2, zeta = case.metrics()
    g = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.upload(case.cel, case.xvel, case.yvel)
    return g


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint32 if x.dtype == np.float32 else np.uint64)


def test_noalloc_gauges_match_allocating_entry_point(gpu_lib):
    case = synth.coastal(nx=260, ny=200, seed=21, crough=0.0025)
    a, b = _grid(gpu_lib, case), _grid(gpu_lib, case)
    rng = np.random.default_rng(5)
    for n, every, nsamples in ((300, 3, 40), (17, 2, 9)):
        I = rng.integers(1, case.dep.shape[1] + 1, n)
        J = rng.integers(1, case.dep.shape[0] + 1, n)
        sa = a.run_gauges(every, nsamples, I, J)
        sb = b.run_gauges(every, nsamples, I, J, noalloc=True)
        assert sa.shape == sb.shape == (nsamples, n)
        assert np.array_equal(_bits(sa), _bits(sb))
        assert np.abs(sa).max() > 0.0          # the series carries the wave
        for x, y in zip(a.download() + a.download_max(), b.download() + b.download_max()):
            assert np.array_equal(_bits(x), _bits(y))
    # beyond the 2^20-word scratch: refused, state untouched
    before = b.download()
    I = np.ones(4096, dtype=np.int32)
    with pytest.raises(gpu_lib.CliffsError) as ei:
        b.run_gauges(1, 300, I, I, noalloc=True)
    assert ei.value.code == -1 and "scratch" in str(ei.value)
    for x, y in zip(before, b.download()):
        assert np.array_equal(_bits(x), _bits(y))
    a.close()
    b.close()
