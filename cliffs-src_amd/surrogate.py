"""Learned nest surrogate: Python side of include/cliffs_surrogate.h.

The network (hand-written BF16 MFMA kernels, csrc/surrogate.hip) maps the
surface-elevation series of 64 gauges on a nested grid's wet boundary (512
samples) to the nested grid's 512 x 512 max-inundation map (per pixel, one
child cell: the ground height the water reached, 0 where it did not) and
to its maximum run-up (the ground height of the highest flooded cell).
This module defines the nested site, generates data with the solver
(parents stepped as an ensemble with device gauge recording for the inputs,
the coupled parent/child nested solve for the targets), permutes maps to
and from the network's quadtree order, trains, and gates predictions:
several independently trained networks must agree on the run-up, otherwise
the scenario is solved with the full nested solve (the fallback, and the
answer its error is measured against).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import cliffs_gpu
import synth

G, T, MAP = 64, 512, 512
NOUT = MAP * MAP
LEVELS = 6                   # 2 x 2 sub-pixel levels below the 8 x 8 base
_fp = C.POINTER(C.c_float)

_sur = None


def _lib():
    global _sur
    if _sur is not None:
        return _sur
    lib = cliffs_gpu.load_library()
    H = C.c_void_p
    sig = {
        "cliffs_sur_last_error": ([], C.c_char_p),
        "cliffs_sur_dims": ([C.POINTER(C.c_int)] * 4, C.c_int),
        "cliffs_sur_create": ([C.c_int, C.c_int, C.c_uint, C.POINTER(H)], C.c_int),
        "cliffs_sur_destroy": ([H], C.c_int),
        "cliffs_sur_nparams": ([H, C.POINTER(C.c_long)], C.c_int),
        "cliffs_sur_get_params": ([H, _fp], C.c_int),
        "cliffs_sur_set_params": ([H, _fp], C.c_int),
        "cliffs_sur_forward": ([H, C.c_int, _fp, _fp, _fp], C.c_int),
        "cliffs_sur_loss_grad": ([H, C.c_int, _fp, _fp, _fp, C.c_float, C.c_float, _fp, _fp], C.c_int),
        "cliffs_sur_set_data": ([H, C.c_long, _fp, _fp, _fp], C.c_int),
        "cliffs_sur_train": ([H, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_uint, _fp],
                             C.c_int),
        "cliffs_sur_gemm_bf16": ([C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_long, C.c_long, _fp,
                                  C.c_long, C.c_long, _fp], C.c_int),
        "cliffs_sur_set_ground": ([H, _fp], C.c_int),
        "cliffs_sur_predict": ([C.POINTER(H), C.c_int, C.c_int, _fp, _fp, _fp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _sur = lib
    return lib


class SurrogateError(RuntimeError):
    pass


def _chk(rc):
    if rc != 0:
        raise SurrogateError(f"{rc}: {_lib().cliffs_sur_last_error().decode()}")


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected {shape}, got {a.shape}")
    return a


def _p(a):
    return a.ctypes.data_as(_fp)


# -- layout ------------------------------------------------------------------

def quadtree_order(size=MAP, levels=LEVELS):
    """q[iy*size + ix] = position of pixel (iy, ix) in the network's output:
    the 8 x 8 base pixel, then `levels` 2 x 2 sub-pixel levels."""
    iy, ix = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
    q = (iy >> levels) * 8 + (ix >> levels)
    for k in range(1, levels + 1):
        q = q * 4 + 2 * ((iy >> (levels - k)) & 1) + ((ix >> (levels - k)) & 1)
    return q.ravel()


_Q = quadtree_order()


def to_quadtree(maps):
    maps = np.asarray(maps, dtype=np.float32).reshape(-1, NOUT)
    out = np.empty_like(maps)
    out[:, _Q] = maps
    return out


def from_quadtree(y):
    y = np.asarray(y, dtype=np.float32).reshape(-1, NOUT)
    return y[:, _Q].reshape(-1, MAP, MAP)


# -- the network ---------------------------------------------------------------

W_MAP, W_RUNUP = 10.0, 1.0   # loss weights of the map and the run-up


class Surrogate:
    def __init__(self, batch_max=64, seed=0, device=0):
        h = C.c_void_p()
        _chk(_lib().cliffs_sur_create(int(device), int(batch_max), int(seed), C.byref(h)))
        self._h = h
        self.batch_max = batch_max
        self.has_ground = False
        n = C.c_long()
        _chk(_lib().cliffs_sur_nparams(h, C.byref(n)))
        self.nparams = n.value

    def close(self):
        if getattr(self, "_h", None):
            _lib().cliffs_sur_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        _chk(_lib().cliffs_sur_get_params(self._h, _p(p)))
        return p

    def set_params(self, p):
        p = _f32(p, (self.nparams,))
        _chk(_lib().cliffs_sur_set_params(self._h, _p(p)))

    def forward(self, x, maps=True):
        """x (B, 512, 64) -> (quadtree maps (B, NOUT) or None, run-ups (B,))."""
        x = _f32(x)
        B = x.shape[0]
        out = np.empty((B, NOUT), dtype=np.float32) if maps else None
        ru = np.empty(B, dtype=np.float32)
        for b0 in range(0, B, self.batch_max):
            b1 = min(B, b0 + self.batch_max)
            xb = np.ascontiguousarray(x[b0:b1])
            rb = np.empty(b1 - b0, dtype=np.float32)
            if maps:
                yb = np.empty((b1 - b0, NOUT), dtype=np.float32)
                _chk(_lib().cliffs_sur_forward(self._h, b1 - b0, _p(xb), _p(yb), _p(rb)))
                out[b0:b1] = yb
            else:
                _chk(_lib().cliffs_sur_forward(self._h, b1 - b0, _p(xb), None, _p(rb)))
            ru[b0:b1] = rb
        return out, ru

    def loss_grad(self, x, y, yr, w_map=W_MAP, w_runup=W_RUNUP):
        x, y, yr = _f32(x), _f32(y), _f32(yr)
        loss = C.c_float()
        g = np.empty(self.nparams, dtype=np.float32)
        _chk(_lib().cliffs_sur_loss_grad(self._h, x.shape[0], _p(x), _p(y), _p(yr), float(w_map),
                                         float(w_runup), C.byref(loss), _p(g)))
        return loss.value, g

    def set_ground(self, ground):
        """Ground heights (MAP x MAP, raster; inf off land) for predict_all's maps."""
        g = _f32(np.where(np.isfinite(ground), ground, np.inf), (MAP, MAP))
        _chk(_lib().cliffs_sur_set_ground(self._h, _p(g)))
        self.has_ground = True

    def set_data(self, x, y, yr):
        """The training set, kept on the device for train()."""
        x, y, yr = _f32(x), _f32(y), _f32(yr)
        _chk(_lib().cliffs_sur_set_data(self._h, x.shape[0], _p(x), _p(y), _p(yr)))

    def train(self, batch=32, steps=1000, lr=1e-3, seed=0, w_map=W_MAP, w_runup=W_RUNUP):
        losses = np.empty(steps, dtype=np.float32)
        _chk(_lib().cliffs_sur_train(self._h, int(batch), int(steps), float(lr), float(w_map),
                                     float(w_runup), int(seed), _p(losses)))
        return losses


def predict_all(nets, x, maps=True):
    """The networks' answer for inputs x (cliffs_sur_predict): every
    network's run-up (k, n) in m, clipped at 0, and, with maps, the
    inundation maps (n, MAP, MAP, raster) of their mean flood levels against
    the site's ground (built on the device)."""
    x = _f32(x)
    n, k = x.shape[0], len(nets)
    bm = min(net.batch_max for net in nets)
    if maps and not nets[0].has_ground:
        nets[0].set_ground(site().ground_win)
    hs = (C.c_void_p * k)(*[net._h for net in nets])
    ru = np.empty((k, n), dtype=np.float32)
    out = np.empty((n, MAP, MAP), dtype=np.float32) if maps else None
    for b0 in range(0, n, bm):
        b1 = min(n, b0 + bm)
        xb = np.ascontiguousarray(x[b0:b1])
        rb = np.empty((k, b1 - b0), dtype=np.float32)
        mb = np.empty((b1 - b0, MAP, MAP), dtype=np.float32) if maps else None
        _chk(_lib().cliffs_sur_predict(hs, k, b1 - b0, _p(xb), _p(rb), _p(mb) if maps else None))
        ru[:, b0:b1] = rb
        if maps:
            out[b0:b1] = mb
    return np.maximum(ru.astype(np.float64), 0.0), out


def gemm_bf16(A, B, device=0):
    """C = A B with BF16 operands and FP32 accumulation on the MFMA units."""
    A, B = _f32(A), _f32(B)
    M, K = A.shape
    N = B.shape[1]
    out = np.empty((M, N), dtype=np.float32)
    _chk(_lib().cliffs_sur_gemm_bf16(device, M, N, K, _p(A), K, 1, _p(B), N, 1, _p(out)))
    return out


# -- the nested site and its scenarios ------------------------------------------
#
# One site: a regional parent grid (128 x 128 cells of 1350 m, shelf and
# coastline of synth.shelf_coast with terrain seed SITE_SEED) and a coastal
# child grid at 15x refinement (90 m, 526 x 526 cells) over 35 x 35 parent
# cells around the coast, fed one-way by the parent (cliffs_nest_*).  The
# child steps CSTEPS times per parent step (dt/3: Courant number about 0.6
# in its deepest water, 150 m).  Minimal flow depth and friction as the
# reference sets them for its 30" and 10" Okushiri grids (paramsOK30s.txt,
# paramsOK10s.txt: 1 m and 0.5 m, Manning n^2 0.000625).  Scenarios differ
# only in the source.  The network's input is the parent's surface
# elevation at 64 gauges along the child's wet boundary (its offshore edge
# and the wet parts of its two lateral edges), every EVERY parent steps; its
# output is the child's max-inundation map over the 512 x 512-cell window
# that leaves 7 cells at each of the child's edges.

SITE_SEED = 5
PNX = PNY = 128
PDX = 1350.0
REFINE = 15
CSTEPS = 3                   # child steps per parent step
CI0, CI1 = 81, 116           # child spans parent columns CI0..CI1 (0-based)
CJ0, CJ1 = 46, 81            # and parent rows CJ0..CJ1
EVERY = 2                    # parent steps between gauge samples and feed records
CUKE = 1.0e-3                # arrival threshold of the feed (m)
MAP_I0, MAP_J0 = 7, 7        # map window: child columns / rows MAP_I0.. / MAP_J0..
WIN = MAP                    # window cells per side (one map pixel per cell)
AMP = (0.5, 8.0)             # source amplitude (m)
SRC_LEN = (20e3, 60e3)       # source length (m)
SRC_X = (0.15, 0.5)          # source centre, fraction of the parent's width


def plain(d):
    """Ground height (m) at distance d (m) inland: a coastal plain rising
    0.1 m per km for 25 km, then 10 m per km.  On the child's 90 m cells the
    plain rises 9 mm per cell, so the solver's run-up (the ground height of
    the highest flooded cell) moves in steps of under 1 % of a metre."""
    return np.where(d < 25e3, 1e-4 * d, 2.5 + 1e-2 * (d - 25e3))


P_GROUND = synth.ENSEMBLE_GROUND   # minimal flow depth of the parent (m)
C_GROUND = 0.5                     # and of the child
CROUGH = synth.ENSEMBLE_CROUGH     # Manning n^2 of both

X_SCALE = 2.0     # gauge elevation (m) -> network input
RUNUP_MIN = 1.0   # run-ups below this (m) are compared in absolute terms

_site = None


def _perimeter_gauges(wet):
    """G parent cells (1-based I, J) evenly spaced along the child's wet
    boundary: its south edge from the coast to the offshore corner, the
    offshore (west) edge, the north edge back to the coast."""
    cells = [(i, CJ0) for i in range(CI1, CI0, -1)] + [(CI0, j) for j in range(CJ0, CJ1)] + \
            [(i, CJ1) for i in range(CI0, CI1 + 1)]
    cells = [(i, j) for i, j in cells if wet[j, i]]
    pick = np.round(np.linspace(0, len(cells) - 1, G)).astype(int)
    I = np.array([cells[k][0] + 1 for k in pick], dtype=np.int32)
    J = np.array([cells[k][1] + 1 for k in pick], dtype=np.int32)
    return I, J


class Site:
    """Terrain, child grid and nest maps of the site (built once)."""

    def __init__(self):
        base = synth.shelf_coast(PNX, PNY, SITE_SEED, PDX, AMP[1], SRC_LEN[0], ground=P_GROUND,
                                 crough=CROUGH, land=plain)
        self.dt = base.dt                           # CFL step for the largest source amplitude
        self.dt_child = self.dt / CSTEPS
        self.dep, self.ground = base.dep, base.ground
        self.x, self.y = base.xcrd, base.ycrd
        self.xc = np.linspace(self.x[CI0], self.x[CI1], REFINE * (CI1 - CI0) + 1)
        self.yc = np.linspace(self.y[CJ0], self.y[CJ1], REFINE * (CJ1 - CJ0) + 1)
        self.depc = _bilinear(self.dep, self.x, self.y, self.xc, self.yc)
        self.cmetrics = synth.metrics(self.xc, self.yc, True)
        self.maps = cliffs_gpu.nest_maps(self.x, self.y, self.xc, self.yc)
        self.Xc, self.Yc = np.meshgrid(self.xc, self.yc)
        self.rows = None                            # tuned tile heights (parent, child)
        self.gI, self.gJ = _perimeter_gauges(self.dep > 5.0)
        # ground rise from one land cell to the next across the coast (median
        # over the window): the step in which the solver's run-up moves
        wg = _window(np.where(self.depc < 0, -self.depc, np.nan))
        self.land_step = float(np.nanmedian(np.abs(np.diff(wg, axis=1))))
        # ground height (FP32, as the maps hold it) of the window's land cells,
        # inf at sea
        self.ground_win = _window(np.where(self.depc < 0, -self.depc, np.inf)).astype(np.float32)


def site():
    global _site
    if _site is None:
        _site = Site()
    return _site


def _bilinear(f, x, y, xq, yq):
    """f (ny, nx) on the grid (y, x) at the points yq x xq."""
    fi = (np.interp(xq, x, np.arange(len(x))))
    fj = (np.interp(yq, y, np.arange(len(y))))
    i0 = np.minimum(fi.astype(int), len(x) - 2)
    j0 = np.minimum(fj.astype(int), len(y) - 2)
    wx, wy = (fi - i0)[None, :], (fj - j0)[:, None]
    a = f[j0][:, i0]
    b = f[j0][:, i0 + 1]
    c = f[j0 + 1][:, i0]
    d = f[j0 + 1][:, i0 + 1]
    return (1 - wy) * ((1 - wx) * a + wx * b) + wy * ((1 - wx) * c + wx * d)


def _window(a):
    return a[MAP_J0:MAP_J0 + WIN, MAP_I0:MAP_I0 + WIN]


def scenario(seed):
    """The site with one source: amplitude AMP, length SRC_LEN, seeded
    strike and place (x in SRC_X of the parent's width)."""
    S = site()
    rng = np.random.default_rng(50_000 + seed)
    return synth.shelf_coast(PNX, PNY, SITE_SEED, PDX, float(rng.uniform(*AMP)),
                             float(rng.uniform(*SRC_LEN)), steps=EVERY * T,
                             source_seed=60_000 + seed, source_x=SRC_X, dt=S.dt,
                             ground=P_GROUND, crough=CROUGH, land=plain)


def _parent(case, device):
    s1, s2, zeta = case.metrics()
    g = cliffs_gpu.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground,
                        crough=case.crough, device=device)
    g.upload(case.cel, case.xvel, case.yvel)
    return g


def flood_levels(maps):
    """The network's map target: per row of the map window, the flood level
    (the highest ground a flooded cell of that row has, 0 where none),
    broadcast along the row -> (n, MAP, MAP).  The ground rises inland along
    every row, so the map is recovered from its levels (map_from_levels)."""
    m = np.asarray(maps, dtype=np.float32).reshape(-1, MAP, MAP)
    return np.broadcast_to(m.max(axis=2, keepdims=True), m.shape).astype(np.float32)


def map_from_levels(levels):
    """Inundation-height maps (raster or quadtree order, like `levels`) from
    flood-level fields: the ground height of every land cell at or below the
    level, 0 elsewhere."""
    lv = np.asarray(levels, dtype=np.float32)
    g = site().ground_win if lv.shape[-1] == MAP else _ground_q()
    return np.where(g <= lv, g, 0.0).astype(np.float32)


def _ground_q():
    S = site()
    if getattr(S, "ground_q", None) is None:
        S.ground_q = to_quadtree(S.ground_win)[0]
    return S.ground_q


def inundation_map(hmax, depc, celmin):
    """Inundation height over the map window, one pixel per child cell: the
    ground height (-dep) of land cells the water reached (celerity maximum
    above celmin, as synth.max_runup counts them), 0 elsewhere.  The map's
    maximum is the run-up over the window."""
    e = np.where((hmax > celmin) & (depc < 0.0), -depc, 0.0)
    return _window(e).astype(np.float32)


def runup_from_map(maps):
    """Run-up read from inundation-height maps: the maximum over pixels
    (exact on the solver's own map)."""
    m = np.asarray(maps, dtype=np.float64).reshape(-1, NOUT)
    return np.maximum(m.max(axis=1), 0.0)


def _tuned_rows(grid):
    """Fastest fused-step tile height for this grid (cliffs_ensemble_tune on
    an ensemble of one; the state is not advanced)."""
    e = cliffs_gpu.Ensemble([grid])
    rows = e.tune(3)
    e.close()
    return rows


class NestedSolve:
    """The full nested solve of one scenario, parent and child on the device
    (cliffs_nest_couple, cliffs_nest_run), EVERY * T parent steps; tile
    heights tuned once per site.  advance(k) runs k record intervals."""

    def __init__(self, case, device=0):
        S = site()
        self.case = case
        self.g = _parent(case, device)
        self.c = cliffs_gpu.Grid(S.depc, *S.cmetrics, dt=S.dt_child, ground=C_GROUND,
                                 crough=case.crough, device=device)
        z = np.zeros_like(S.depc)
        src = case.meta["source"]
        self.c.upload(synth.depth_to_celerity(S.depc + synth.dipole(S.Xc, S.Yc, S.depc, src), C_GROUND),
                      z, z)
        if S.rows is None:
            S.rows = (_tuned_rows(self.g), _tuned_rows(self.c))
        self.g.set_tile_rows(S.rows[0])
        self.c.set_tile_rows(S.rows[1])
        self.n = cliffs_gpu.Nest(self.g, S.maps)
        self.n.couple(self.c, capacity=4)
        self.iarv = False
        self.done = 0

    def advance(self, k):
        S = site()
        self.iarv = self.n.run(k, EVERY, S.dt, EVERY * CSTEPS, S.dt_child, CUKE, self.case.ground,
                               self.iarv)
        self.done += k

    def hmax(self):
        return self.c.download_max()[0]

    def close(self):
        self.n.close()
        self.c.close()
        self.g.close()


def nested_hmax(case, device=0):
    """The full nested solve of one scenario -> (child hmax, child celmin)."""
    ns = NestedSolve(case, device)
    ns.advance(T)
    hmax, cm = ns.hmax(), ns.c.celmin
    ns.close()
    return hmax, cm


def nested_solve(case, device=0):
    """-> (child inundation map (512, 512), exact max run-up over the map
    window (m)) of the full nested solve (nested_hmax)."""
    S = site()
    hmax, cm = nested_hmax(case, device)
    return inundation_map(hmax, S.depc, cm), synth.max_runup(_window(hmax), _window(S.depc), cm)


_pools = {}


def parent_gauges(cases, device=0):
    """Parents of all `cases` stepped together (cliffs_ensemble_*), the
    gauge series recorded on the device: -> (n, T, G) in m.  The parents of
    the site differ only in their initial state, so the grids and their
    ensemble (tile shape tuned once, cliffs_ensemble_tune) are kept per
    batch size and device, and each batch uploads its states and resets the
    maxima."""
    S = site()
    key = (len(cases), device)
    if key not in _pools:
        grids = [_parent(c, device) for c in cases]
        ens = cliffs_gpu.Ensemble(grids)
        ens.tune(3)
        _pools[key] = (grids, ens)
    else:
        grids, ens = _pools[key]
        for g, c in zip(grids, cases):
            g.upload(c.cel, c.xvel, c.yvel)
            g.reset_max()
    return ens.run_gauges(EVERY, T, S.gI, S.gJ)


def release_parents():
    """Free the kept parent grids and ensembles."""
    for grids, ens in _pools.values():
        ens.close()
        for g in grids:
            g.close()
    _pools.clear()


def dataset(seeds, batch=64, solvers=8, device=0, log=None):
    """For the scenarios `seeds`: network inputs x (N, T, G), quadtree
    level-field targets y (N, NOUT, m; flood_levels), exact run-ups ru (N,),
    flooded cells area (N,) and, per scenario, the cells where the map
    rebuilt from its own levels differs from the solver's map (rec).  The
    parents of each batch are stepped as one ensemble, the nested solves run
    `solvers` at a time from host threads (each on its own grids and
    streams; the library's calls release the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    seeds = list(seeds)
    n = len(seeds)
    x = np.empty((n, T, G), dtype=np.float32)
    yq = np.empty((n, NOUT), dtype=np.float32)
    ru = np.empty(n)
    area = np.empty(n, dtype=np.int64)
    rec = np.empty(n, dtype=np.int64)
    for b0 in range(0, n, batch):
        cases = [scenario(int(s)) for s in seeds[b0:b0 + batch]]
        x[b0:b0 + len(cases)] = parent_gauges(cases, device) * X_SCALE
        with ThreadPoolExecutor(max_workers=solvers) as ex:
            out = list(ex.map(lambda c: nested_solve(c, device), cases))
        for k, (m, r) in enumerate(out):
            lv = flood_levels(m)
            yq[b0 + k] = to_quadtree(lv)[0]
            ru[b0 + k] = r
            area[b0 + k] = int((m > 0).sum())
            rec[b0 + k] = int((map_from_levels(lv)[0] != m).sum())
        if log:
            log(b0 + len(cases), n)
    return dict(x=x, y=yq, ru=ru, area=area, rec=rec)


def predict_runups(nets, x):
    """Each network's run-up (m, clipped at 0) for inputs x: the gate's input."""
    return list(predict_all(nets, x, maps=False)[0])


def predict_levels(nets, x):
    """The networks' mean flood-level field (quadtree order) for inputs x."""
    acc = None
    for net in nets:
        m, _ = net.forward(x)
        acc = m if acc is None else acc + m
    return acc / len(nets)


def answer(cases, nets, tol, lo=-np.inf, hi=np.inf, device=0):
    """Max-inundation maps and run-ups of `cases` at the site: the parents
    stepped as one ensemble give the gauge series; where the gate accepts
    the networks' run-ups (agreement within tol, mean in [lo, hi]) the
    answer is their mean run-up and the map of their mean flood levels
    (built on the device), elsewhere the full nested solve runs.  -> (maps
    (n, 512, 512), run-ups (n,), answered by the surrogate (n,))."""
    x = parent_gauges(cases, device) * X_SCALE
    rus, _ = predict_all(nets, x, maps=False)
    ok = gate(list(rus), tol=tol, lo=lo, hi=hi)
    runups = rus.mean(axis=0)
    maps = np.zeros((len(cases), MAP, MAP), dtype=np.float32)
    if ok.any():
        maps[ok] = predict_all(nets, x[ok])[1]
    for k in np.flatnonzero(~ok):
        maps[k], runups[k] = nested_solve(cases[k], device)
    return maps, runups, ok


def flooded_cells(qlevels):
    """Flooded land cells of the maps of quadtree-ordered flood-level fields."""
    return (_ground_q() <= np.asarray(qlevels)).sum(axis=-1)


def gate_score(runups):
    """Disagreement of the networks' run-ups (list of (n,) arrays): their
    spread (largest minus smallest) relative to their mean, or to RUNUP_MIN
    where the mean is smaller."""
    r = np.stack(runups)
    return (r.max(axis=0) - r.min(axis=0)) / np.maximum(r.mean(axis=0), RUNUP_MIN)


def gate(runups, tol, lo=-np.inf, hi=np.inf):
    """Confidence gate: the networks' run-ups agree within tol and their
    mean lies in [lo, hi] (m), the range where they were found reliable
    (calibrate_gate)."""
    m = np.mean(np.stack(runups), axis=0)
    return (gate_score(runups) <= tol) & (m >= lo) & (m <= hi)


def calibrate_gate(runups, true, budget, q=0.95):
    """The gate (tol, lo, hi) that accepts the most of the calibration runs
    whose networks' run-ups are `runups` (list of (n,)) while keeping the
    q-quantile of the accepted run-up errors (relative, absolute / RUNUP_MIN
    below RUNUP_MIN) within `budget`: lo and hi among quantiles of the
    predicted run-ups, for each pair the loosest tolerance.  -> (tol, lo,
    hi), or None when no gate meets the budget."""
    r = np.stack(runups)
    m = r.mean(axis=0)
    err = np.abs(m - true) / np.maximum(true, RUNUP_MIN)
    sc = gate_score(runups)
    los = [-np.inf] + list(np.quantile(m, [0.01, 0.02, 0.03, 0.05, 0.075, 0.1]))
    his = [np.inf] + list(np.quantile(m, [0.99, 0.98, 0.97, 0.95, 0.925, 0.9]))
    best, best_n = None, 0
    for lo in los:
        for hi in his:
            inr = (m >= lo) & (m <= hi)
            for t in np.sort(sc[inr])[::-1]:
                ok = inr & (sc <= t)
                if np.quantile(err[ok], q) <= budget:
                    if ok.sum() > best_n:
                        best, best_n = (float(t), float(lo), float(hi)), int(ok.sum())
                    break
    return best
