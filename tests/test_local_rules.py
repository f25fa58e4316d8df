"""The wet/dry front logic: sequential segment scan vs. local rules.

The reference decides shorelines with a sequential scan over the line
(cliffs.f:65-123, ghosts cliffs.f:137-164, limiter cliffs.f:264-285, edge
gating cliffs.f:289-290, 310-311).  The HIP kernels use per-node rules of
radius <= 3 instead.  This test restates the scan at the level of its
decisions (which node is newland, expanded, interior, which neighbour is a
ghost, which limiter case fires, which edge is integrated) and compares it
with the local rules over every land mask of length n <= 10.  The
data-dependent conditions (flood > ground, velocity sign) enter both forms
only as the conjunctions fr[i] & ur[i] (right expansion into land node i
after water) and fl[i] & ul[i] (left expansion into land node i before
water); every combination of these conjunctions at the nodes
where they can matter is enumerated (a false conjunction with either member
false, the other conditions random).
"""
import itertools

import numpy as np
import pytest


def sequential(land, fr, ur, fl, ul):
    """Decisions of cliffs.f:58-164/264-311 for one line (0-based nodes).

    fr[i]: (h(i-1)^2/g - d(i-1) + d(i) > ground) for a right expansion into i
    ur[i]: u(i-1) >= 0
    fl[i]: (h(i+1)^2/g - d(i+1) + d(i) > ground) for a left expansion into i
    ul[i]: u(i+1) <= 0
    """
    n = len(land)
    newland = list(land)
    lg, rg = {}, {}
    RE, LE = [False] * n, [False] * n
    water = not land[0]
    kseg = 1 if water else 0
    if water:
        lg[1] = 0           # reference lghost(1) = 1 (1-based) -> node index 0
    k1 = kseg + 1
    for i in range(1, n):
        if water:
            if land[i]:
                if fr[i] and (not land[i - 1]) and ur[i]:
                    newland[i] = False
                    RE[i] = True
                else:
                    water = False
                    rg[kseg] = i
        elif not land[i]:
            water = True
            kseg += 1
            if fl[i - 1] and ul[i - 1]:
                lg[kseg] = max(i - 2, 0)
                newland[i - 1] = False
                LE[i - 1] = True
            else:
                lg[kseg] = i - 1
            if kseg > 1 and lg[kseg] < rg.get(kseg - 1, -1):
                del lg[kseg]
                kseg -= 1
                rg[kseg] = -1
    if kseg == 0:
        return None
    if rg.get(kseg, -1) == -1:
        rg[kseg] = n - 1
    k2 = kseg - 1
    if newland[rg[kseg]]:
        k2 = kseg
    interior = {}
    for k in range(1, kseg + 1):
        i1, i2 = lg[k] + 1, rg[k] - 1
        for i in range(i1, i2 + 1):
            # neighbour kinds: 'g' ghost, 'r' real
            left = "g" if (i == i1 and k >= k1) else "r"
            right = "g" if (i == i2 and k <= k2) else "r"
            if i == i1 and k < k1:
                left = "r"        # pqvdL(:,1) = real node 1 (cliffs.f:154-158)
            if i == i1 and lg[k] == 0 and not newland[0]:
                left = "r"        # node 1 flooded: cliffs.f:154-158 overrides the ghost
            if i == i2 and k == kseg and not newland[n - 1]:
                right = "r"       # cliffs.f:160-164
            interior[i] = (left, right)
    lim_left, lim_right = set(), set()
    if k1 <= kseg:
        for k in range(k1, kseg + 1):
            i = lg[k] + 1
            if land[i]:
                lim_left.add(i)
    if k2 >= 1:
        for k in range(1, k2 + 1):
            i = rg[k] - 1
            if land[i]:
                lim_right.add(i)
    edge_first = (not newland[0]) and (not newland[1])
    edge_last = (not newland[n - 2]) and (not newland[n - 1])
    return dict(newland=newland, RE=RE, LE=LE, interior=interior, lim_left=lim_left,
                lim_right=lim_right, edge_first=edge_first, edge_last=edge_last)


def local(land, fr, ur, fl, ul):
    """The kernels' local rules (cliffs-src_amd/csrc/sweeps.hip header)."""
    n = len(land)
    if all(land):
        return None
    RE = [i >= 1 and land[i] and not land[i - 1] and fr[i] and ur[i] for i in range(n)]
    start = [(not land[i]) and (i == 0 or (land[i - 1] and not RE[i - 1])) for i in range(n)]
    LE = [i + 1 < n and start[i + 1] and fl[i] and ul[i] for i in range(n)]
    newland = [land[i] and not RE[i] and not LE[i] for i in range(n)]
    interior = {}
    for i in range(1, n - 1):
        if not newland[i]:
            interior[i] = ("g" if newland[i - 1] else "r", "g" if newland[i + 1] else "r")
    lim_left = {i for i in range(1, n - 1) if LE[i] and newland[i - 1]}
    lim_right = {i for i in range(1, n - 1) if RE[i] and newland[i + 1]}
    return dict(newland=newland, RE=RE, LE=LE, interior=interior, lim_left=lim_left,
                lim_right=lim_right,
                edge_first=(not newland[0]) and (not newland[1]),
                edge_last=(not newland[n - 2]) and (not newland[n - 1]))


def _check(land, fr, ur, fl, ul):
    a = sequential(land, fr, ur, fl, ul)
    b = local(land, fr, ur, fl, ul)
    assert (a is None) == (b is None), (land, fr, ur, fl, ul)
    if a is None:
        return
    for key in a:
        assert a[key] == b[key], (key, land, fr, ur, fl, ul, a[key], b[key])


@pytest.mark.parametrize("n", range(3, 11))
def test_local_rules_exhaustive(n):
    """Every land mask and every combination of the expansion conditions
    where they can matter: right expansion into land node i after a water
    node (land[i] and not land[i-1]), left expansion into land node i before
    a water node (land[i] and not land[i+1])."""
    rng = np.random.default_rng(100 + n)
    cases = 0
    for mask in itertools.product([False, True], repeat=n):
        land = list(mask)
        R = [i for i in range(1, n) if land[i] and not land[i - 1]]
        L = [i for i in range(n - 1) if land[i] and not land[i + 1]]
        for bits in itertools.product([False, True], repeat=len(R) + len(L)):
            fr, ur, fl, ul = (list(rng.integers(0, 2, n).astype(bool)) for _ in range(4))
            for (i, b), (x, y) in zip(zip(R + L, bits), [(fr, ur)] * len(R) + [(fl, ul)] * len(L)):
                if b:
                    x[i] = y[i] = True
                elif rng.integers(0, 2):
                    x[i] = False
                else:
                    y[i] = False
            _check(land, fr, ur, fl, ul)
            cases += 1
    assert cases >= 2 ** n


def test_local_rules_long_random_lines():
    rng = np.random.default_rng(1)
    for _ in range(300):
        n = int(rng.integers(11, 400))
        p = rng.uniform(0.05, 0.95)
        land = list(rng.uniform(size=n) < p)
        fr, ur, fl, ul = (list(rng.integers(0, 2, n).astype(bool)) for _ in range(4))
        a = sequential(land, fr, ur, fl, ul)
        b = local(land, fr, ur, fl, ul)
        assert (a is None) == (b is None)
        if a is not None:
            for key in a:
                assert a[key] == b[key], key
