"""Python binding of the cliffs_gpu C ABI (include/cliffs_gpu.h) via ctypes.

Mirrors the reference's call surface for the hot path: a grid handle owning
the MASTER/PARAMETERS state (global_modules.f:16-55), ``sweep`` for one
OpenMP loop of ``cliffs(iu,irw,n)`` calls (cliffs.f:17, cliffs_main.f:325-329,
340-344), ``maxwave`` (cliffs_main.f:347-355) and ``step``/``run`` for whole
time steps.  Arrays are numpy float64 of shape (ny, nx) -- the reference's
Fortran (nx, ny) bytes.

There is no CPU fallback: constructing a ``Grid`` without the built HIP
library, or on a machine without a HIP device, raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcliffs_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cliffs_gpu.h")

OK = 0
ERR = {-1: "CLIFFS_ERR_ARG", -2: "CLIFFS_ERR_HIP", -3: "CLIFFS_ERR_NOMEM",
       -4: "CLIFFS_ERR_METRIC", -5: "CLIFFS_ERR_NODEVICE", -6: "CLIFFS_ERR_STATE"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class CliffsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR.get(code, code)}: {msg}")
        self.code = code


class _Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dep", _dp), ("s1", _dp), ("s2", _dp),
                ("zeta", _dp), ("dt", C.c_double), ("ground", C.c_double),
                ("crough", C.c_double), ("device", C.c_int)]


_lib = None


def load_library():
    """Load libcliffs_gpu.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with cliffs-src_amd/build.sh "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    H = C.c_void_p
    sig = {
        "cliffs_create": ([C.POINTER(_Desc), C.POINTER(H)], C.c_int),
        "cliffs_destroy": ([H], C.c_int),
        "cliffs_upload_state": ([H, _dp, _dp, _dp], C.c_int),
        "cliffs_download_state": ([H, _dp, _dp, _dp], C.c_int),
        "cliffs_download_max": ([H, _dp, _dp], C.c_int),
        "cliffs_reset_max": ([H], C.c_int),
        "cliffs_sweep": ([H, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int], C.c_int),
        "cliffs_maxwave": ([H], C.c_int),
        "cliffs_step": ([H, _dp, _dp, _dp, _dp, C.c_int], C.c_int),
        "cliffs_run": ([H, C.c_int], C.c_int),
        "cliffs_run_timed": ([H, C.c_int, _dp, _dp, _dp], C.c_int),
        "cliffs_set_fused": ([H, C.c_int], C.c_int),
        "cliffs_set_tile_rows": ([H, C.c_int], C.c_int),
        "cliffs_set_tile_width": ([H, C.c_int], C.c_int),
        "cliffs_ensemble_shape": ([H, _ip, _ip], C.c_int),
        "cliffs_run_gauges": ([H, C.c_int, C.c_int, C.c_int, _ip, _ip, C.POINTER(C.c_float)],
                              C.c_int),
        "cliffs_gauges": ([H, C.c_int, _ip, _ip, _dp, _dp, _dp], C.c_int),
        "cliffs_run_gauges_noalloc": ([H, C.c_int, C.c_int, C.c_int, _ip, _ip, C.POINTER(C.c_float)],
                                      C.c_int),
        "cliffs_health": ([H, _dp, C.POINTER(C.c_longlong)], C.c_int),
        "cliffs_ensemble_create": ([C.POINTER(H), C.c_int, C.POINTER(H)], C.c_int),
        "cliffs_ensemble_destroy": ([H], C.c_int),
        "cliffs_ensemble_run": ([H, C.c_int], C.c_int),
        "cliffs_ensemble_run_timed": ([H, C.c_int, _dp], C.c_int),
        "cliffs_ensemble_tune": ([H, C.c_int, _ip], C.c_int),
        "cliffs_ensemble_sync": ([H], C.c_int),
        "cliffs_ensemble_run_gauges": ([H, C.c_int, C.c_int, C.c_int, _ip, _ip, C.POINTER(C.c_float)],
                                       C.c_int),
        "cliffs_sync": ([H], C.c_int),
        "cliffs_stream": ([H, C.POINTER(C.c_void_p)], C.c_int),
        "cliffs_info": ([H, _dp, _dp, _ip, _ip, _ip, _ip], C.c_int),
        "cliffs_last_error": ([], C.c_char_p),
        "cliffs_nest_maps": ([C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp,
                              _ip, _dp, _ip, _dp, _ip, _dp, _ip, _dp, _ip, _dp, _ip, _dp],
                             C.c_int),
        "cliffs_nest_create": ([H, C.c_int, C.c_int, _ip, _dp, _ip, _dp, C.c_int, C.c_double,
                                C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.c_double,
                                C.POINTER(H)], C.c_int),
        "cliffs_nest_destroy": ([H], C.c_int),
        "cliffs_nest_feed": ([H, C.c_double, C.c_double, _ip, _ip, _dp, _dp, _dp, _dp], C.c_int),
        "cliffs_nest_couple": ([H, H, C.c_int], C.c_int),
        "cliffs_nest_record": ([H, C.c_double, C.c_double, C.c_double, _ip, _ip], C.c_int),
        "cliffs_nest_end_feed": ([H], C.c_int),
        "cliffs_nest_child_step": ([H, C.c_double], C.c_int),
        "cliffs_nest_child_start": ([H, C.c_double, _dp], C.c_int),
        "cliffs_nest_run": ([H, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                             C.c_double, C.c_double, _ip], C.c_int),
        # FP64 line mode (include/cliffs_fp64.h)
        "cliffs64_create": ([C.POINTER(_Desc), C.POINTER(H)], C.c_int),
        "cliffs64_destroy": ([H], C.c_int),
        "cliffs64_upload_state": ([H, _dp, _dp, _dp], C.c_int),
        "cliffs64_download_state": ([H, _dp, _dp, _dp], C.c_int),
        "cliffs64_download_max": ([H, _dp, _dp], C.c_int),
        "cliffs64_sweep": ([H, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int], C.c_int),
        "cliffs64_maxwave": ([H], C.c_int),
        "cliffs64_run": ([H, C.c_int], C.c_int),
        "cliffs64_gauges": ([H, C.c_int, _ip, _ip, _dp, _dp, _dp], C.c_int),
        "cliffs64_health": ([H, _dp, C.POINTER(C.c_longlong)], C.c_int),
        "cliffs64_nest_create": ([H, C.c_int, C.c_int, _ip, _dp, _ip, _dp, C.c_int, C.c_double,
                                  C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.c_double,
                                  C.POINTER(H)], C.c_int),
        "cliffs64_nest_feed": ([H, C.c_double, C.c_double, _ip, _ip, _dp, _dp, _dp, _dp], C.c_int),
        "cliffs64_nest_destroy": ([H], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def header_symbols(path: str = HEADER):
    """Names of the functions declared in include/cliffs_gpu.h."""
    import re
    txt = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(cliffs_\w+)\s*\(", txt, re.M)))


def _chk(rc: int):
    if rc != OK:
        raise CliffsError(rc, load_library().cliffs_last_error().decode())


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Grid:
    """One grid resident on a GPU (the reference's MASTER state)."""

    def __init__(self, dep, s1=None, s2=None, zeta=None, *, dt, ground, crough=0.0, device=0):
        lib = load_library()
        dep = _f64(dep)
        ny, nx = dep.shape
        self.nx, self.ny = nx, ny
        keep = [dep]
        s1a = _f64(s1) if s1 is not None and nx > 2 else None
        s2a = _f64(s2) if s2 is not None and ny > 2 else None
        za = _f64(zeta) if zeta is not None and ny > 2 else None
        keep += [s1a, s2a, za]
        d = _Desc(nx, ny, _p(dep), _p(s1a), _p(s2a), _p(za), float(dt), float(ground),
                  float(crough), int(device))
        h = C.c_void_p()
        _chk(lib.cliffs_create(C.byref(d), C.byref(h)))
        self._h = h
        self._keep = keep
        # nests and ensembles using this grid, closed before it (strong
        # references: a weak one is already cleared when the garbage
        # collector finalises a grid and its nest together)
        self._deps = set()
        g, cm = C.c_double(), C.c_double()
        r = [C.c_int() for _ in range(4)]
        _chk(lib.cliffs_info(h, C.byref(g), C.byref(cm), *[C.byref(x) for x in r]))
        self.grav, self.celmin = g.value, cm.value
        self.ii1, self.ii2, self.jj1, self.jj2 = [x.value for x in r]
        self.dep = dep
        self.mx = max(ny if nx > 2 else 1, nx if ny > 2 else 1)

    def close(self):
        for d in list(getattr(self, "_deps", ())):
            d.close()           # they hold this grid's device pointers
        if getattr(self, "_h", None):
            load_library().cliffs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state --------------------------------------------------------------
    def upload(self, cel, xvel, yvel):
        shp = (self.ny, self.nx)
        c, u, v = _f64(cel, shp), _f64(xvel, shp), _f64(yvel, shp)
        _chk(load_library().cliffs_upload_state(self._h, _p(c), _p(u), _p(v)))

    def download(self):
        out = [np.empty((self.ny, self.nx)) for _ in range(3)]
        _chk(load_library().cliffs_download_state(self._h, *[_p(a) for a in out]))
        return tuple(out)

    def download_max(self):
        out = [np.empty((self.ny, self.nx)) for _ in range(2)]
        _chk(load_library().cliffs_download_max(self._h, *[_p(a) for a in out]))
        return tuple(out)

    def reset_max(self):
        _chk(load_library().cliffs_reset_max(self._h))

    # -- compute ------------------------------------------------------------
    @staticmethod
    def _edge(e, mx):
        if e is None:
            return None
        e = np.asarray(e, dtype=np.float64)
        if e.shape[1] != 3 or e.shape[0] < mx:
            raise ValueError("edge arrays are (mx, 3)")
        return np.ascontiguousarray(e.T)   # Fortran (mx,3) column-major bytes

    def sweep(self, iu, first, last, edge1=None, edge2=None):
        e1, e2 = self._edge(edge1, self.mx), self._edge(edge2, self.mx)
        mx = self.mx if (e1 is not None or e2 is not None) else 0
        if e1 is not None:
            mx = e1.shape[1]
        elif e2 is not None:
            mx = e2.shape[1]
        _chk(load_library().cliffs_sweep(self._h, iu, first, last, _p(e1), _p(e2), mx))

    def maxwave(self):
        _chk(load_library().cliffs_maxwave(self._h))

    def step(self, xedge1=None, xedge2=None, yedge1=None, yedge2=None):
        es = [self._edge(e, self.mx) for e in (xedge1, xedge2, yedge1, yedge2)]
        mx = next((e.shape[1] for e in es if e is not None), self.mx)
        _chk(load_library().cliffs_step(self._h, *[_p(e) for e in es], mx))

    def run(self, nsteps: int):
        _chk(load_library().cliffs_run(self._h, int(nsteps)))

    def set_fused(self, on: bool):
        """2-D steps as one fused kernel (default) or as two sweep kernels."""
        _chk(load_library().cliffs_set_fused(self._h, 1 if on else 0))

    def set_tile_width(self, width: int):
        """Tile width of the fused step: 248 (default) or 120 columns."""
        _chk(load_library().cliffs_set_tile_width(self._h, int(width)))

    def set_tile_rows(self, rows: int):
        """Tile height of the fused step (8..256, 0 = automatic)."""
        _chk(load_library().cliffs_set_tile_rows(self._h, int(rows)))

    def run_timed(self, nsteps: int):
        """-> (ms_total, ms_a, ms_b), HIP events on the handle's stream: mean
        durations of the x- and y-sweep kernels, or of the fused step kernel
        and its flag-check kernel when the step runs fused, over every 8th
        step of the run (the first included)."""
        t, x, y = C.c_double(), C.c_double(), C.c_double()
        _chk(load_library().cliffs_run_timed(self._h, int(nsteps), C.byref(t), C.byref(x),
                                             C.byref(y)))
        return t.value, x.value, y.value

    def run_gauges(self, every, nsamples, I, J, noalloc=False):
        """nsamples x every steps; -> (nsamples, n) FP32 surface elevation at
        the gauges (1-based I, J) after every `every` steps (0 where dry).
        noalloc: through cliffs_run_gauges_noalloc (no device allocation)."""
        I = np.ascontiguousarray(I, dtype=np.int32)
        J = np.ascontiguousarray(J, dtype=np.int32)
        out = np.empty((int(nsamples), len(I)), dtype=np.float32)
        f = load_library().cliffs_run_gauges_noalloc if noalloc else load_library().cliffs_run_gauges
        _chk(f(self._h, int(every), int(nsamples), len(I), I.ctypes.data_as(_ip),
               J.ctypes.data_as(_ip), out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def gauges(self, I, J):
        I = np.ascontiguousarray(I, dtype=np.int32)
        J = np.ascontiguousarray(J, dtype=np.int32)
        n = len(I)
        eta, u, v = np.empty(n), np.empty(n), np.empty(n)
        _chk(load_library().cliffs_gauges(self._h, n, I.ctypes.data_as(_ip), J.ctypes.data_as(_ip),
                                          _p(eta), _p(u), _p(v)))
        return eta, u, v

    def health(self):
        """-> (largest Courant number over wet cells, count of cells with a
        non-finite c, u or v) of the current state (cliffs_health)."""
        cour, bad = C.c_double(), C.c_longlong()
        _chk(load_library().cliffs_health(self._h, C.byref(cour), C.byref(bad)))
        return cour.value, bad.value

    def sync(self):
        _chk(load_library().cliffs_sync(self._h))

    def stream(self) -> int:
        s = C.c_void_p()
        _chk(load_library().cliffs_stream(self._h, C.byref(s)))
        return s.value or 0


class Grid64:
    """A grid in FP64 line mode (include/cliffs_fp64.h): the reference's
    state and line procedure in double precision on the GPU, equal to the
    reference's FP64 arithmetic bit for bit; one wave per line with a
    sequential segment scan, so slower than Grid."""

    def __init__(self, dep, s1=None, s2=None, zeta=None, *, dt, ground, crough=0.0, device=0):
        lib = load_library()
        dep = _f64(dep)
        ny, nx = dep.shape
        self.nx, self.ny = nx, ny
        s1a = _f64(s1) if s1 is not None and nx > 2 else None
        s2a = _f64(s2) if s2 is not None and ny > 2 else None
        za = _f64(zeta) if zeta is not None and ny > 2 else None
        d = _Desc(nx, ny, _p(dep), _p(s1a), _p(s2a), _p(za), float(dt), float(ground),
                  float(crough), int(device))
        h = C.c_void_p()
        _chk(lib.cliffs64_create(C.byref(d), C.byref(h)))
        self._h = h
        self._deps = set()     # nests of this grid, closed before it
        self.mx = max(ny if nx > 2 else 1, nx if ny > 2 else 1)

    def close(self):
        for d in list(getattr(self, "_deps", ())):
            d.close()
        if getattr(self, "_h", None):
            load_library().cliffs64_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, cel, xvel, yvel):
        shp = (self.ny, self.nx)
        c, u, v = _f64(cel, shp), _f64(xvel, shp), _f64(yvel, shp)
        _chk(load_library().cliffs64_upload_state(self._h, _p(c), _p(u), _p(v)))

    def download(self):
        out = [np.empty((self.ny, self.nx)) for _ in range(3)]
        _chk(load_library().cliffs64_download_state(self._h, *[_p(a) for a in out]))
        return tuple(out)

    def download_max(self):
        out = [np.empty((self.ny, self.nx)) for _ in range(2)]
        _chk(load_library().cliffs64_download_max(self._h, *[_p(a) for a in out]))
        return tuple(out)

    def sweep(self, iu, first, last, edge1=None, edge2=None):
        e1, e2 = Grid._edge(edge1, self.mx), Grid._edge(edge2, self.mx)
        mx = next((e.shape[1] for e in (e1, e2) if e is not None), 0)
        _chk(load_library().cliffs64_sweep(self._h, iu, first, last, _p(e1), _p(e2), mx))

    def maxwave(self):
        _chk(load_library().cliffs64_maxwave(self._h))

    def run(self, nsteps: int):
        _chk(load_library().cliffs64_run(self._h, int(nsteps)))

    def gauges(self, I, J):
        """(eta, u, v) at 1-based gauge indices I, J (output_gages.f:18-29)."""
        I = np.ascontiguousarray(I, dtype=np.int32)
        J = np.ascontiguousarray(J, dtype=np.int32)
        out = [np.empty(len(I)) for _ in range(3)]
        _chk(load_library().cliffs64_gauges(self._h, len(I), I.ctypes.data_as(_ip), J.ctypes.data_as(_ip),
                                            *[_p(a) for a in out]))
        return tuple(out)

    def health(self):
        """(largest Courant number over wet cells, count of non-finite cells)."""
        c, n = C.c_double(), C.c_longlong()
        _chk(load_library().cliffs64_health(self._h, C.byref(c), C.byref(n)))
        return c.value, n.value


class Nest64:
    """One child grid's boundary feed from a Grid64 parent (feed_children_grid
    in the reference's FP64 arithmetic)."""

    def __init__(self, parent: Grid64, maps: dict):
        lib = load_library()
        self.parent = parent
        self.nyc, self.nxc = len(maps["we_ny"]), len(maps["sn_nx"])
        we_ny = np.ascontiguousarray(maps["we_ny"], dtype=np.int32)
        sn_nx = np.ascontiguousarray(maps["sn_nx"], dtype=np.int32)
        we_cy, sn_cx = _f64(maps["we_cy"]), _f64(maps["sn_cx"])
        h = C.c_void_p()
        _chk(lib.cliffs64_nest_create(parent._h, self.nxc, self.nyc, we_ny.ctypes.data_as(_ip),
                                      _p(we_cy), sn_nx.ctypes.data_as(_ip), _p(sn_cx),
                                      int(maps["w_nx"]), float(maps["w_cx"]), int(maps["e_nx"]),
                                      float(maps["e_cx"]), int(maps["s_ny"]), float(maps["s_cy"]),
                                      int(maps["n_ny"]), float(maps["n_cy"]), C.byref(h)))
        self._h = h
        parent._deps.add(self)

    def close(self):
        if getattr(self, "_h", None):
            load_library().cliffs64_nest_destroy(self._h)
            self._h = None
            self.parent._deps.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def feed(self, cuke, hmin, iarv=False):
        """-> (arrived, fed, uvq_w (nyc,3), uvq_e, uvq_s (nxc,3), uvq_n)."""
        a, f = C.c_int(1 if iarv else 0), C.c_int(0)
        w, e = np.zeros((3, self.nyc)), np.zeros((3, self.nyc))
        s, n = np.zeros((3, self.nxc)), np.zeros((3, self.nxc))
        _chk(load_library().cliffs64_nest_feed(self._h, float(cuke), float(hmin), C.byref(a),
                                               C.byref(f), _p(w), _p(e), _p(s), _p(n)))
        return bool(a.value), bool(f.value), w.T.copy(), e.T.copy(), s.T.copy(), n.T.copy()


class Ensemble:
    """Grids stepped together: one fused launch per step for all members,
    replayed from a HIP graph (cliffs_ensemble_*).  Members stay usable."""

    def __init__(self, grids):
        self.grids = list(grids)
        arr = (C.c_void_p * len(self.grids))(*[g._h.value for g in self.grids])
        h = C.c_void_p()
        _chk(load_library().cliffs_ensemble_create(arr, len(self.grids), C.byref(h)))
        self._h = h
        for g in self.grids:
            g._deps.add(self)

    def run(self, nsteps: int):
        _chk(load_library().cliffs_ensemble_run(self._h, int(nsteps)))

    def run_timed(self, nsteps: int) -> float:
        t = C.c_double()
        _chk(load_library().cliffs_ensemble_run_timed(self._h, int(nsteps), C.byref(t)))
        return t.value

    def tune(self, trial_steps: int = 3) -> int:
        r = C.c_int()
        _chk(load_library().cliffs_ensemble_tune(self._h, int(trial_steps), C.byref(r)))
        return r.value

    def shape(self):
        """-> (tile width, tile height) the ensemble steps with."""
        w, r = C.c_int(), C.c_int()
        _chk(load_library().cliffs_ensemble_shape(self._h, C.byref(w), C.byref(r)))
        return w.value, r.value

    def sync(self):
        _chk(load_library().cliffs_ensemble_sync(self._h))

    def run_gauges(self, every, nsamples, I, J):
        """nsamples x every steps (every even); -> (members, nsamples, n) FP32
        surface elevation at each member's gauges (1-based I, J of shape (n,)
        for all members, or (members, n)) after every `every` steps."""
        m = len(self.grids)
        I = np.ascontiguousarray(np.broadcast_to(np.asarray(I, dtype=np.int32), (m, np.shape(I)[-1])))
        J = np.ascontiguousarray(np.broadcast_to(np.asarray(J, dtype=np.int32), I.shape))
        out = np.empty((m, int(nsamples), I.shape[1]), dtype=np.float32)
        _chk(load_library().cliffs_ensemble_run_gauges(self._h, int(every), int(nsamples), I.shape[1],
                                                       I.ctypes.data_as(_ip), J.ctypes.data_as(_ip),
                                                       out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def close(self):
        if getattr(self, "_h", None):
            load_library().cliffs_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nest_maps(xp, yp, xc, yc):
    """Parent index maps of one child grid (NestedGrid.f:104-119), computed by
    the library's host code.  Returns a dict of 1-based maps, as the NESTED
    module holds them for one child."""
    lib = load_library()
    xp, yp, xc, yc = (_f64(a) for a in (xp, yp, xc, yc))
    nxc, nyc = len(xc), len(yc)
    we_ny = np.zeros(nyc, dtype=np.int32)
    sn_nx = np.zeros(nxc, dtype=np.int32)
    we_cy = np.zeros(nyc)
    sn_cx = np.zeros(nxc)
    sc = [C.c_int(), C.c_double(), C.c_int(), C.c_double(), C.c_int(), C.c_double(),
          C.c_int(), C.c_double()]
    _chk(lib.cliffs_nest_maps(len(xp), _p(xp), len(yp), _p(yp), nxc, _p(xc), nyc, _p(yc),
                              we_ny.ctypes.data_as(_ip), _p(we_cy), sn_nx.ctypes.data_as(_ip),
                              _p(sn_cx), *[C.byref(s) for s in sc]))
    return dict(we_ny=we_ny, we_cy=we_cy, sn_nx=sn_nx, sn_cx=sn_cx,
                w_nx=sc[0].value, w_cx=sc[1].value, e_nx=sc[2].value, e_cx=sc[3].value,
                s_ny=sc[4].value, s_cy=sc[5].value, n_ny=sc[6].value, n_cy=sc[7].value)


class Nest:
    """One child grid's boundary feed from a parent Grid (feed_children_grid)."""

    def __init__(self, parent: Grid, maps: dict):
        lib = load_library()
        self.parent = parent
        self.nyc, self.nxc = len(maps["we_ny"]), len(maps["sn_nx"])
        we_ny = np.ascontiguousarray(maps["we_ny"], dtype=np.int32)
        sn_nx = np.ascontiguousarray(maps["sn_nx"], dtype=np.int32)
        we_cy, sn_cx = _f64(maps["we_cy"]), _f64(maps["sn_cx"])
        h = C.c_void_p()
        _chk(lib.cliffs_nest_create(parent._h, self.nxc, self.nyc, we_ny.ctypes.data_as(_ip),
                                    _p(we_cy), sn_nx.ctypes.data_as(_ip), _p(sn_cx),
                                    int(maps["w_nx"]), float(maps["w_cx"]), int(maps["e_nx"]),
                                    float(maps["e_cx"]), int(maps["s_ny"]), float(maps["s_cy"]),
                                    int(maps["n_ny"]), float(maps["n_cy"]), C.byref(h)))
        self._h = h
        parent._deps.add(self)

    def close(self):
        if getattr(self, "_h", None):
            load_library().cliffs_nest_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def feed(self, cuke, hmin, iarv=False):
        """-> (arrived, fed, uvq_w (nyc,3), uvq_e, uvq_s (nxc,3), uvq_n)."""
        a, f = C.c_int(1 if iarv else 0), C.c_int(0)
        w, e = np.zeros((3, self.nyc)), np.zeros((3, self.nyc))
        s, n = np.zeros((3, self.nxc)), np.zeros((3, self.nxc))
        _chk(load_library().cliffs_nest_feed(self._h, float(cuke), float(hmin), C.byref(a),
                                             C.byref(f), _p(w), _p(e), _p(s), _p(n)))
        return bool(a.value), bool(f.value), w.T.copy(), e.T.copy(), s.T.copy(), n.T.copy()

    # -- parent and child on one device ------------------------------------
    def couple(self, child: Grid, capacity: int = 4):
        """Keep feed records on the device and step `child` from them."""
        _chk(load_library().cliffs_nest_couple(self._h, child._h, int(capacity)))
        self.child = child
        child._deps.add(self)

    def record(self, cuke, hmin, t, iarv=False):
        """-> (arrived, records so far)."""
        a, k = C.c_int(1 if iarv else 0), C.c_int()
        _chk(load_library().cliffs_nest_record(self._h, float(cuke), float(hmin), float(t),
                                               C.byref(a), C.byref(k)))
        return bool(a.value), k.value

    def end_feed(self):
        _chk(load_library().cliffs_nest_end_feed(self._h))

    def child_start(self, cuke):
        """The child's start record under its threshold cuke
        (get_boundary_feed.f:144-167) -> that record's time t0."""
        t0 = C.c_double()
        _chk(load_library().cliffs_nest_child_start(self._h, float(cuke), C.byref(t0)))
        return t0.value

    def child_step(self, t):
        _chk(load_library().cliffs_nest_child_step(self._h, float(t)))

    def run(self, nrec, every, dt, per, dt_child, cuke, hmin, iarv=False, t0=0.0):
        """`nrec` record intervals of the coupled solve in one call
        (cliffs_nest_run): per interval `every` parent steps of dt, a record,
        and the child's steps of dt_child (`per` per interval) as far as the
        records allow.  -> arrival latch."""
        a = C.c_int(1 if iarv else 0)
        _chk(load_library().cliffs_nest_run(self._h, int(nrec), int(every), float(t0), float(dt),
                                            int(per), float(dt_child), float(cuke), float(hmin),
                                            C.byref(a)))
        return bool(a.value)


NAN32 = np.float32(-1.0e32)   # the reference's real*4 NAN (global_modules.f:18)


def maxwave_fields(hmax, velmax, dep, grav, celmin):
    """MaxE and MaxV as the reference writes them (output_maxwave.f:14-24):
    MaxE = hmax**2/g - dep and MaxV = sqrt(velmax) where hmax > celmin, else
    the real*4 NAN (-1e32); FP64 arithmetic, FP32 results."""
    hmax = np.asarray(hmax, dtype=np.float64)
    velmax = np.asarray(velmax, dtype=np.float64)
    wet = hmax > celmin
    e = np.where(wet, hmax * hmax / grav - np.asarray(dep, dtype=np.float64), 0.0)
    v = np.where(wet, np.sqrt(np.maximum(velmax, 0.0)), 0.0)
    return (np.where(wet, e.astype(np.float32), NAN32),
            np.where(wet, v.astype(np.float32), NAN32))
