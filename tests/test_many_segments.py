"""Lines with more than 99 wet segments.

The reference keeps the ghost nodes of a line's wet segments in fixed
arrays of 99 (pqvdL, pqvdR, cliffs.f:26): a line with more segments writes
past them, which is undefined behaviour.  The HIP path has no such limit,
and the oracle's ghost storage grows with the line.  Segments separated by
five or more dry nodes do not interact (expansion, ghost and limiter rules
reach at most two nodes into the gap, and no gluing occurs), so a line of
150 segments must give, segment by segment, what the same line gives when
split into two halves of at most 99 segments each, where the reference's
behaviour is defined.  That pins the oracle past the limit; the device
path is then checked against it with the bars of the other parity tests.
"""
import numpy as np
import pytest

import oracle as O
import synth

GROUND = 0.01
GAP = 5


def _segment_rows(rng, nx, ny):
    """Rows of wet segments (3..8 nodes) separated by GAP dry land nodes,
    starting and ending on land; -> dep, cel, u, v, per-row segment spans."""
    dep = np.full((ny, nx), -1.0)
    spans = []
    for j in range(ny):
        row, i = [], GAP
        while True:
            n = int(rng.integers(3, 9))
            if i + n + GAP > nx:
                break
            dep[j, i:i + n] = rng.uniform(0.5, 3.0, n)
            dep[j, i + n:i + n + GAP] = rng.uniform(-1.0, -0.2, GAP)
            row.append((i, i + n))
            i += n + GAP
        spans.append(row)
    eta = np.where(dep > 0, rng.normal(0, 0.2, dep.shape), 0.0)
    cel = synth.depth_to_celerity(dep + eta, GROUND)
    u = np.where(cel > 0, rng.normal(0, 0.5, dep.shape), 0.0)
    v = np.where(cel > 0, rng.normal(0, 0.5, dep.shape), 0.0)
    return dep, cel, u, v, spans


def _without(dep, cel, u, v, spans, keep):
    """The same rows with the segments whose index is not in `keep` turned
    into dry land."""
    d, c, uu, vv = dep.copy(), cel.copy(), u.copy(), v.copy()
    for j, row in enumerate(spans):
        for k, (a, b) in enumerate(row):
            if k not in keep:
                d[j, a:b] = -0.5
                c[j, a:b] = 0.0
                uu[j, a:b] = 0.0
                vv[j, a:b] = 0.0
    return d, c, uu, vv


def _sweep_oracle(dep, cel, u, v, x, y, dt=0.05):
    s1, s2, zeta = synth.metrics(x, y, True)
    orc = O.Oracle(dep, s1, s2, zeta, dt=dt, ground=GROUND)
    orc.set_state(cel, u, v)
    over = orc.sweep(1, 1, dep.shape[0])
    return over, orc.get_state()[:3], orc


def _case(seed, nx=1700, ny=4):
    rng = np.random.default_rng(seed)
    dep, cel, u, v, spans = _segment_rows(rng, nx, ny)
    x = np.cumsum(rng.uniform(0.5, 1.5, nx))
    y = np.cumsum(rng.uniform(0.5, 1.5, ny))
    return dep, cel, u, v, spans, x, y


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_past_99_segments_equals_split_lines(seed):
    dep, cel, u, v, spans, x, y = _case(seed)
    ns = min(len(r) for r in spans)
    assert ns > 140
    over, full, _ = _sweep_oracle(dep, cel, u, v, x, y)
    assert over == dep.shape[0]          # every row is past the 99-segment storage
    half = ns // 2
    first = set(range(half))
    over1, out1, _ = _sweep_oracle(*_without(dep, cel, u, v, spans, first), x, y)
    over2, out2, _ = _sweep_oracle(*_without(dep, cel, u, v, spans,
                                             set(range(half, max(len(r) for r in spans)))), x, y)
    assert over1 == 0 and over2 == 0     # both halves within the reference's storage
    for j, row in enumerate(spans):
        cut = row[half][0] - GAP // 2 - 1   # middle of the gap between the halves
        for f, a, b in zip(full, out1, out2):
            assert np.array_equal(f[j, :cut], a[j, :cut]), f"row {j}: first half differs"
            assert np.array_equal(f[j, cut:], b[j, cut:]), f"row {j}: second half differs"


@pytest.fixture(params=["emul", pytest.param("gpu", marks=pytest.mark.gpu)])
def backend(request):
    if request.param == "gpu":
        return request.getfixturevalue("gpu_lib").Grid
    import emul
    return emul.EmulGrid


@pytest.mark.parametrize("seed", [0, 1])
def test_device_past_99_segments(backend, seed):
    dep, cel, u, v, spans, x, y = _case(seed)
    over, (c_o, u_o, v_o), orc = _sweep_oracle(dep, cel, u, v, x, y)
    assert over == dep.shape[0]
    s1, s2, zeta = synth.metrics(x, y, True)
    g = backend(dep, s1, s2, zeta, dt=0.05, ground=GROUND)
    g.upload(cel, u, v)
    mx = max(dep.shape)
    z = np.zeros((mx, 3))
    g.sweep(1, 1, dep.shape[0], z, z)
    c_g, u_g, v_g = g.download()
    # expansion / zeroing decisions exact, values to the single-sweep bars
    assert np.array_equal(c_o == 0.0, c_g == 0.0)
    assert np.array_equal(c_o == orc.celmin, np.isclose(c_g, orc.celmin, rtol=1e-6, atol=0))
    assert np.abs(c_o - c_g).max() <= 2e-5 * (1.0 + np.abs(c_o)).max()
    assert np.abs(u_o - u_g).max() <= 2e-4
    assert np.abs(v_o - v_g).max() <= 2e-4
    if hasattr(g, "close"):
        g.close()
