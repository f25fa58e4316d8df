"""Learned nest surrogate: Python side of include/cliffs_surrogate.h.

The network (hand-written BF16 MFMA kernels, csrc/surrogate.hip) maps the
surface-elevation series of 64 gauges on a nested grid's offshore boundary
(256 samples) to the grid's 64 x 64 max-inundation map.  This module
generates training data with the solver (cliffs_gpu), permutes maps to and
from the network's quadtree order, trains, and gates predictions: two
independently trained networks must agree, otherwise the scenario is solved
with the full GPU solve (the fallback and the judge of the surrogate).
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


# -- training data from the solver ---------------------------------------------

X_SCALE = 0.5     # gauge elevation (m) -> network input
Y_SCALE = 0.25    # inundation depth (m) -> network output


def scenario(seed, nx=320, ny=256, dx=150.0):
    """One nested-grid scenario: shelf and coastline (synth.shelf_coast),
    dipole source of length 8-30 km and amplitude 0.5-4 m at a seeded place
    and strike."""
    rng = np.random.default_rng(50_000 + seed)
    return synth.shelf_coast(nx, ny, seed, dx, float(rng.uniform(0.5, 4.0)),
                             float(rng.uniform(8e3, 30e3)), steps=1536)


def gauge_line(case):
    """64 gauges along the offshore line x = 0.58 Lx (1-based indices)."""
    ny, nx = case.dep.shape
    I = np.full(G, int(0.58 * nx) + 1, dtype=np.int32)
    J = np.round(np.linspace(1, ny, G)).astype(np.int32)
    return I, J


def inundation_map(hmax, dep, celmin, cols=128):
    """Max water depth on flooded land (H = hmax^2 / g where hmax > celmin,
    dep < 0) over the last `cols` columns, block-averaged to 64 x 64."""
    ny, nx = dep.shape
    h = np.where((hmax > celmin) & (dep < 0.0), hmax * hmax / synth.GRAV, 0.0)[:, nx - cols:]
    return h.reshape(MAP, ny // MAP, MAP, cols // MAP).mean(axis=(1, 3))


def solve(case, every=6, device=0):
    """The full GPU solve of one scenario: -> (gauge series (256, 64) in m,
    inundation map (64, 64) in m, max run-up in m)."""
    s1, s2, zeta = case.metrics()
    g = cliffs_gpu.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, device=device)
    g.upload(case.cel, case.xvel, case.yvel)
    I, J = gauge_line(case)
    series = g.run_gauges(every, T, I, J)
    hmax, _ = g.download_max()
    cm = g.celmin
    g.close()
    return series, inundation_map(hmax, case.dep, cm), synth.max_runup(hmax, case.dep, cm)


def dataset(seeds, device=0):
    """Network inputs (N, 256, 64), quadtree targets (N, 4096), maps and run-ups."""
    xs, maps, runups = [], [], []
    for s in seeds:
        series, m, r = solve(scenario(int(s)), device=device)
        xs.append(series * X_SCALE)
        maps.append(m)
        runups.append(r)
    maps = np.array(maps, dtype=np.float32)
    return (np.array(xs, dtype=np.float32), to_quadtree(maps * Y_SCALE), maps,
            np.array(runups))


def predict_maps(net, x):
    return np.maximum(from_quadtree(net.forward(x)) / Y_SCALE, 0.0)


def flooded_area(maps, thr=0.05):
    return (np.asarray(maps) > thr).sum(axis=(-2, -1))


def gate(maps_a, maps_b, tol):
    """Confidence gate: the two networks' maps agree within tol (relative L2
    of their difference to their mean)."""
    a, b = np.asarray(maps_a), np.asarray(maps_b)
    num = np.sqrt(((a - b) ** 2).sum(axis=(-2, -1)))
    den = np.sqrt((((a + b) * 0.5) ** 2).sum(axis=(-2, -1))) + 1e-3
    return num / den <= tol
