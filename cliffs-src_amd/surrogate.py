"""Learned nest surrogate: Python side of include/cliffs_surrogate.h.

The network (hand-written BF16 MFMA kernels, csrc/surrogate.hip) maps the
surface-elevation series of 64 gauges on a nested grid's offshore boundary
(256 samples) to the nested grid's 64 x 64 max-inundation map (per pixel the
highest ground the water reached, so the map's maximum is the run-up).  This module
defines the nested site, generates data with the solver (parents stepped as
an ensemble with device gauge recording for the inputs, the coupled
parent/child nested solve for the targets), permutes maps to and from the
network's quadtree order, trains, and gates predictions: two independently
trained networks must agree, otherwise the scenario is solved with the full
nested solve (the fallback, and the answer its error is measured against).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import cliffs_gpu
import synth

G, T, MAP = 64, 256, 64
NOUT = MAP * MAP
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
        "cliffs_sur_forward": ([H, C.c_int, _fp, _fp], C.c_int),
        "cliffs_sur_loss_grad": ([H, C.c_int, _fp, _fp, _fp, _fp], C.c_int),
        "cliffs_sur_train": ([H, C.c_long, _fp, _fp, C.c_int, C.c_int, C.c_float, C.c_uint, _fp],
                             C.c_int),
        "cliffs_sur_gemm_bf16": ([C.c_int, C.c_int, C.c_int, C.c_int, _fp, C.c_long, C.c_long, _fp,
                                  C.c_long, C.c_long, _fp], C.c_int),
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

def quadtree_order():
    """q[iy*64 + ix] = position of pixel (iy, ix) in the network's output
    (8 x 8 base pixel, then three 2 x 2 sub-pixel levels)."""
    iy, ix = np.meshgrid(np.arange(MAP), np.arange(MAP), indexing="ij")
    p0 = (iy >> 3) * 8 + (ix >> 3)
    s = [2 * ((iy >> (3 - k)) & 1) + ((ix >> (3 - k)) & 1) for k in (1, 2, 3)]
    return (((p0 * 4 + s[0]) * 4 + s[1]) * 4 + s[2]).ravel()


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

class Surrogate:
    def __init__(self, batch_max=64, seed=0, device=0):
        h = C.c_void_p()
        _chk(_lib().cliffs_sur_create(int(device), int(batch_max), int(seed), C.byref(h)))
        self._h = h
        self.batch_max = batch_max
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

    def forward(self, x):
        """x (B, 256, 64) -> (B, 4096) quadtree-ordered outputs."""
        x = _f32(x)
        B = x.shape[0]
        out = np.empty((B, NOUT), dtype=np.float32)
        for b0 in range(0, B, self.batch_max):
            b1 = min(B, b0 + self.batch_max)
            xb = np.ascontiguousarray(x[b0:b1])
            yb = np.empty((b1 - b0, NOUT), dtype=np.float32)
            _chk(_lib().cliffs_sur_forward(self._h, b1 - b0, _p(xb), _p(yb)))
            out[b0:b1] = yb
        return out

    def loss_grad(self, x, y):
        x, y = _f32(x), _f32(y)
        loss = C.c_float()
        g = np.empty(self.nparams, dtype=np.float32)
        _chk(_lib().cliffs_sur_loss_grad(self._h, x.shape[0], _p(x), _p(y), C.byref(loss), _p(g)))
        return loss.value, g

    def train(self, x, y, batch=32, steps=1000, lr=1e-3, seed=0):
        x, y = _f32(x), _f32(y)
        losses = np.empty(steps, dtype=np.float32)
        _chk(_lib().cliffs_sur_train(self._h, x.shape[0], _p(x), _p(y), int(batch), int(steps),
                                     float(lr), int(seed), _p(losses)))
        return losses


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
# One site: a parent grid (192 x 192, 900 m, shelf and coastline of
# synth.shelf_coast with terrain seed SITE_SEED) and a child grid at 5x
# refinement (180 m, dt/5) over the coastal strip, fed one-way by the parent
# (cliffs_nest_*).  Minimal flow depth and friction as the reference sets
# them for its 30" and 10" Okushiri grids (paramsOK30s.txt, paramsOK10s.txt:
# 1 m and 0.5 m, Manning n^2 0.000625), at which run-up is a well-posed
# function of the source.  Scenarios differ only in the source.  The
# network's input is the parent's surface elevation at 64 gauges along the
# child's offshore (west) edge, every EVERY parent steps; its output is the
# child's max-inundation map over the coastal window.

SITE_SEED = 5
PNX = PNY = 192
PDX = 900.0
REFINE = 5
CI0, CI1 = 124, 164          # child spans parent columns CI0..CI1 (0-based)
CJ0, CJ1 = 32, 161           # and parent rows CJ0..CJ1
EVERY = 4                    # parent steps between gauge samples and feed records
CUKE = 1.0e-3                # arrival threshold of the feed (m)
MAP_I0, MAP_W = 56, 128      # map window: child columns MAP_I0.. (2 per pixel),
MAP_J0, MAP_H = 3, 640        # rows MAP_J0.. (10 per pixel), clear of the child's edges

def plain(d):
    """Ground height (m) at distance d (m) inland: a coastal plain rising
    0.1 m per km for 25 km, then 10 m per km.  On the child's 180 m cells
    the plain rises 1.8 cm per cell, so the solver's run-up (the ground
    height of the highest flooded cell) moves in steps of ~1 % of a metre."""
    return np.where(d < 25e3, 1e-4 * d, 2.5 + 1e-2 * (d - 25e3))


P_GROUND = synth.ENSEMBLE_GROUND   # minimal flow depth of the parent (m)
C_GROUND = 0.5                     # and of the child
CROUGH = synth.ENSEMBLE_CROUGH     # Manning n^2 of both

X_SCALE = 2.0     # gauge elevation (m) -> network input
Y_SCALE = 0.125   # inundation height (m) -> network output
RUNUP_MIN = 1.0   # run-ups below this (m) are compared in absolute terms

_site = None


class Site:
    """Terrain, child grid and nest maps of the site (built once)."""

    def __init__(self):
        base = synth.shelf_coast(PNX, PNY, SITE_SEED, PDX, 4.0, 10e3, ground=P_GROUND,
                                 crough=CROUGH, land=plain)
        self.dt = base.dt                           # CFL step for the largest source amplitude
        self.dep, self.ground = base.dep, base.ground
        self.x, self.y = base.xcrd, base.ycrd
        self.xc = np.linspace(self.x[CI0], self.x[CI1], REFINE * (CI1 - CI0) + 1)
        self.yc = np.linspace(self.y[CJ0], self.y[CJ1], REFINE * (CJ1 - CJ0) + 1)
        self.depc = _bilinear(self.dep, self.x, self.y, self.xc, self.yc)
        self.cmetrics = synth.metrics(self.xc, self.yc, True)
        self.maps = cliffs_gpu.nest_maps(self.x, self.y, self.xc, self.yc)
        self.Xc, self.Yc = np.meshgrid(self.xc, self.yc)
        self.rows = None                            # tuned tile heights (parent, child)
        # gauges: parent cells on the child's west edge (1-based)
        self.gI = np.full(G, CI0 + 1, dtype=np.int32)
        self.gJ = np.round(np.linspace(CJ0 + 1, CJ1 + 1, G)).astype(np.int32)
        # ground rise from one land cell to the next across the coast (median
        # over the window): the step in which the solver's run-up moves
        wg = _window(np.where(self.depc < 0, -self.depc, np.nan))
        self.land_step = float(np.nanmedian(np.abs(np.diff(wg, axis=1))))
        # lowest land ground height per map pixel (inf where the pixel is all sea)
        w = _window(np.where(self.depc < 0, -self.depc, np.inf))
        self.pix_ground_min = w.reshape(MAP, MAP_H // MAP, MAP, MAP_W // MAP).min(axis=(1, 3))


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
    return a[MAP_J0:MAP_J0 + MAP_H, MAP_I0:MAP_I0 + MAP_W]


def scenario(seed):
    """The site with one source: amplitude 0.5-4 m, length 8-30 km, seeded
    strike and place (x in 0.3-0.5 of the parent's width)."""
    S = site()
    rng = np.random.default_rng(50_000 + seed)
    return synth.shelf_coast(PNX, PNY, SITE_SEED, PDX, float(rng.uniform(0.5, 4.0)),
                             float(rng.uniform(8e3, 30e3)), steps=EVERY * T,
                             source_seed=60_000 + seed, source_x=(0.3, 0.5), dt=S.dt,
                             ground=P_GROUND, crough=CROUGH, land=plain)


def _parent(case, device):
    s1, s2, zeta = case.metrics()
    g = cliffs_gpu.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground,
                        crough=case.crough, device=device)
    g.upload(case.cel, case.xvel, case.yvel)
    return g


def inundation_map(hmax, depc, celmin):
    """Inundation height: the ground height (-dep) of land cells the water
    reached (celerity maximum above celmin, as synth.max_runup counts them),
    0 elsewhere; block maximum over 10 x 2 child cells of the coastal window
    -> 64 x 64.  The map's maximum is the run-up over the window."""
    e = np.where((hmax > celmin) & (depc < 0.0), -depc, 0.0)
    e = _window(e)
    return e.reshape(MAP, MAP_H // MAP, MAP, MAP_W // MAP).max(axis=(1, 3))


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


def nested_hmax(case, device=0):
    """The full nested solve of one scenario: parent and child on the device
    (cliffs_nest_couple/record/child_step), EVERY * T parent steps, tile
    heights tuned once per site.  -> (child hmax, child celmin)."""
    S = site()
    g = _parent(case, device)
    c = cliffs_gpu.Grid(S.depc, *S.cmetrics, dt=case.dt / REFINE, ground=C_GROUND,
                        crough=case.crough, device=device)
    src = case.meta["source"]
    z = np.zeros_like(S.depc)
    c.upload(synth.depth_to_celerity(S.depc + synth.dipole(S.Xc, S.Yc, S.depc, src), C_GROUND), z, z)
    if S.rows is None:
        S.rows = (_tuned_rows(g), _tuned_rows(c))
    g.set_tile_rows(S.rows[0])
    c.set_tile_rows(S.rows[1])
    n = cliffs_gpu.Nest(g, S.maps)
    n.couple(c, capacity=4)
    dtc, per = case.dt / REFINE, REFINE * EVERY      # child steps per record
    trec, iarv, kc = [], False, 0
    for k in range(T):
        g.run(EVERY)
        tp = (k + 1) * EVERY * case.dt
        iarv, nrec = n.record(CUKE, case.ground, tp, iarv)
        if nrec > len(trec):
            trec.append(tp)
        while kc + 1 < per * (len(trec) - 1):
            kc += 1
            j = kc // per
            n.child_step(trec[j] + (kc - j * per) * dtc)
    hmax, _ = c.download_max()
    cm = c.celmin
    n.close()
    c.close()
    g.close()
    return hmax, cm


def nested_solve(case, device=0):
    """-> (child inundation map (64, 64), exact max run-up over the map
    window (m)) of the full nested solve (nested_hmax)."""
    S = site()
    hmax, cm = nested_hmax(case, device)
    return inundation_map(hmax, S.depc, cm), synth.max_runup(_window(hmax), _window(S.depc), cm)


def parent_gauges(cases, device=0):
    """Parents of all `cases` stepped together (cliffs_ensemble_*), the
    gauge series recorded on the device: -> (n, T, G) in m."""
    S = site()
    grids = [_parent(c, device) for c in cases]
    ens = cliffs_gpu.Ensemble(grids)
    ens.tune(3)
    out = ens.run_gauges(EVERY, T, S.gI, S.gJ)
    ens.close()
    for g in grids:
        g.close()
    return out


def dataset(seeds, batch=64, device=0):
    """Network inputs (N, T, G), quadtree targets (N, 4096), maps and exact
    run-ups, for the scenarios `seeds`."""
    xs, maps, runups = [], [], []
    seeds = list(seeds)
    for b0 in range(0, len(seeds), batch):
        cases = [scenario(int(s)) for s in seeds[b0:b0 + batch]]
        xs.append(parent_gauges(cases, device) * X_SCALE)
        for c in cases:
            m, r = nested_solve(c, device)
            maps.append(m)
            runups.append(r)
    maps = np.array(maps, dtype=np.float32)
    return (np.concatenate(xs).astype(np.float32), to_quadtree(maps * Y_SCALE), maps,
            np.array(runups))


def predict_maps(net, x):
    return np.maximum(from_quadtree(net.forward(x)) / Y_SCALE, 0.0)


def answer(cases, nets, tol, device=0):
    """Max-inundation maps of `cases` at the site: the parents stepped as one
    ensemble give the gauge series; where the networks' run-ups agree within
    tol (gate) the mean of their maps is the answer, elsewhere the full
    nested solve runs.  -> (maps (n, 64, 64), answered by the surrogate (n,))."""
    x = parent_gauges(cases, device) * X_SCALE
    preds = [predict_maps(n, x) for n in nets]
    ok = gate(*preds, tol=tol)
    maps = np.mean(preds, axis=0).astype(np.float64)
    for k in np.flatnonzero(~ok):
        maps[k] = nested_solve(cases[k], device)[0]
    return maps, ok


def flooded_area(maps):
    """Pixels with flooded land: inundation height at or above the pixel's
    lowest land cell."""
    S = site()
    return (np.asarray(maps) >= S.pix_ground_min).sum(axis=(-2, -1))


def gate_score(*maps):
    """Disagreement of networks' run-ups (map maxima): their spread (largest
    minus smallest) relative to their mean, or to RUNUP_MIN where the mean
    is smaller."""
    r = np.stack([runup_from_map(m) for m in maps])
    return (r.max(axis=0) - r.min(axis=0)) / np.maximum(r.mean(axis=0), RUNUP_MIN)


def gate(*maps, tol):
    """Confidence gate: the networks' run-ups agree within tol."""
    return gate_score(*maps) <= tol
