"""ctypes access to the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Two interchangeable backends with one interface:

* ``Oracle``  -- ``libcliffs_oracle.so``, the plain-C FP64 restatement of
  cliffs.f / cliffs_main.f:317-355 / feed_children_grid.f / NestedGrid.f
  pinpoint / MasterGrid.f metrics (see cliffs_oracle.h for the citations).
* ``RefCore`` -- ``_ref/libcliffs_ref.so``, the reference's own
  global_modules.f + cliffs.f compiled in place by this directory's
  Makefile, driven through ref_harness.f90.  Only present where the
  reference tree was available at build time.

Only tests/, bench.py's cpu_baseline leg and __graft_entry__.smoke() import
this module, and only as the checker of the HIP path.

All 2-D arrays are numpy float64 of shape (ny, nx) -- C order with x
fastest, which is byte-for-byte the reference's Fortran (nx, ny) layout.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libcliffs_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcliffs_ref.so")
REFERENCE_TREE = "/root/reference"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(with_ref: bool = True) -> None:
    """Compile the C oracle and, when the reference tree is present, the
    reference core (oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if with_ref and os.path.isdir(REFERENCE_TREE):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class _OrcGrid(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("dep", _dp), ("cel", _dp), ("xvel", _dp), ("yvel", _dp),
        ("hmax", _dp), ("velmax", _dp),
        ("s1", _dp), ("s2", _dp), ("zeta", _dp),
        ("edge1", _dp), ("edge2", _dp),
        ("mx", C.c_int),
        ("dt", C.c_double), ("ground", C.c_double), ("crough", C.c_double),
        ("celmin", C.c_double),
    ]


def line_ranges(nx: int, ny: int):
    """cliffs_main.f:140-164 -> (ii1, ii2, jj1, jj2, xrun, yrun, mx)."""
    if nx > 2:
        ii1, ii2, xrun, nc = 2, nx - 1, True, ny
    else:
        ii1, ii2, xrun, nc = 1, 1, False, 1
    if ny > 2:
        jj1, jj2, yrun, mc = 2, ny - 1, True, nx
    else:
        jj1, jj2, yrun, mc = 1, 1, False, 1
    return ii1, ii2, jj1, jj2, xrun, yrun, max(nc, mc)


class _Base:
    """Common state handling.  ``s1`` (ny, nx-1), ``s2`` (ny-1,), ``zeta`` (ny,)."""

    def _arrays(self, dep, s1, s2, zeta):
        ny, nx = dep.shape
        self.nx, self.ny = nx, ny
        self.dep = np.ascontiguousarray(dep, dtype=np.float64)
        self.s1 = np.ascontiguousarray(
            s1 if s1 is not None else np.zeros((ny, max(nx - 1, 1))), dtype=np.float64)
        self.s2 = np.ascontiguousarray(
            s2 if s2 is not None else np.zeros(max(ny - 1, 1)), dtype=np.float64)
        self.zeta = np.ascontiguousarray(
            zeta if zeta is not None else np.zeros(ny), dtype=np.float64)
        self.mx = line_ranges(nx, ny)[6]


class Oracle(_Base):
    """FP64 C restatement.  Holds its own copy of the state."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_SO):
                build(with_ref=False)
            lib = C.CDLL(ORACLE_SO)
            lib.orc_grav.restype = C.c_double
            lib.orc_pi.restype = C.c_double
            lib.orc_cliffs_line.argtypes = [C.POINTER(_OrcGrid), C.c_int, C.c_int]
            lib.orc_cliffs_line.restype = C.c_int
            lib.orc_sweep.argtypes = [C.POINTER(_OrcGrid), C.c_int, C.c_int, C.c_int]
            lib.orc_sweep.restype = C.c_int
            lib.orc_maxwave.argtypes = [C.POINTER(_OrcGrid)]
            lib.orc_steps.argtypes = [C.POINTER(_OrcGrid), C.c_int]
            lib.orc_steps.restype = C.c_int
            lib.orc_metrics.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp]
            lib.orc_pinpoint.argtypes = [_dp, C.c_int, C.c_double, _ip, _dp]
            lib.orc_nest_maps.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp,
                                          _ip, _dp, _ip, _dp, _ip, _dp, _ip, _dp, _ip, _dp, _ip, _dp]
            lib.orc_feed_child.argtypes = [C.POINTER(_OrcGrid), C.c_double, C.c_double,
                                           C.c_int, C.c_int, _ip, _dp, _ip, _dp,
                                           C.c_int, C.c_double, C.c_int, C.c_double,
                                           C.c_int, C.c_double, C.c_int, C.c_double,
                                           _ip, _dp, _dp, _dp, _dp]
            lib.orc_feed_child.restype = C.c_int
            lib.orc_apply_initial_conditions.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp,
                                                         C.c_int, C.c_int, _dp, _dp, _dp]
            cls._lib = lib
        return cls._lib

    def __init__(self, dep, s1=None, s2=None, zeta=None, *, dt, ground, crough=0.0):
        self._arrays(dep, s1, s2, zeta)
        lib = self.lib()
        self.grav = lib.orc_grav()
        self.dt, self.ground, self.crough = float(dt), float(ground), float(crough)
        self.celmin = float(np.sqrt(self.grav * self.ground))
        shp = self.dep.shape
        self.cel = np.zeros(shp)
        self.xvel = np.zeros(shp)
        self.yvel = np.zeros(shp)
        self.hmax = np.zeros(shp)
        self.velmax = np.zeros(shp)
        self.edge1 = np.zeros((3, self.mx))   # Fortran (mx,3) column-major
        self.edge2 = np.zeros((3, self.mx))
        self._g = _OrcGrid(self.nx, self.ny, _ptr(self.dep), _ptr(self.cel), _ptr(self.xvel),
                           _ptr(self.yvel), _ptr(self.hmax), _ptr(self.velmax),
                           _ptr(self.s1), _ptr(self.s2), _ptr(self.zeta),
                           _ptr(self.edge1), _ptr(self.edge2), self.mx,
                           self.dt, self.ground, self.crough, self.celmin)

    # state ---------------------------------------------------------------
    def set_state(self, cel, xvel, yvel):
        self.cel[...] = cel
        self.xvel[...] = xvel
        self.yvel[...] = yvel

    def get_state(self):
        return (self.cel.copy(), self.xvel.copy(), self.yvel.copy(),
                self.hmax.copy(), self.velmax.copy())

    def set_edges(self, edge1, edge2):
        """edge1/edge2 given as (mx, 3) arrays (reference index order)."""
        self.edge1[...] = np.asarray(edge1, dtype=np.float64).T
        self.edge2[...] = np.asarray(edge2, dtype=np.float64).T

    # compute -------------------------------------------------------------
    def line(self, iu: int, irw: int) -> int:
        return self.lib().orc_cliffs_line(C.byref(self._g), iu, irw)

    def sweep(self, iu: int, first: int, last: int) -> int:
        return self.lib().orc_sweep(C.byref(self._g), iu, first, last)

    def maxwave(self):
        self.lib().orc_maxwave(C.byref(self._g))

    def steps(self, n: int) -> int:
        return self.lib().orc_steps(C.byref(self._g), n)


class RefCore(_Base):
    """The reference's own cliffs.f (oracle/_ref).  Module-global state: only
    one instance is live at a time."""

    _lib = None

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(REF_SO)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            lib = C.CDLL(REF_SO)
            lib.ref_init.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                     C.c_double, C.c_double, C.c_double]
            lib.ref_grav.restype = C.c_double
            lib.ref_celmin.restype = C.c_double
            lib.ref_mx.restype = C.c_int
            lib.ref_set_state.argtypes = [_dp, _dp, _dp]
            lib.ref_get_state.argtypes = [_dp, _dp, _dp, _dp, _dp]
            lib.ref_set_edges.argtypes = [_dp, _dp]
            lib.ref_line.argtypes = [C.c_int, C.c_int]
            lib.ref_sweep.argtypes = [C.c_int, C.c_int, C.c_int]
            lib.ref_maxwave.argtypes = []
            lib.ref_steps.argtypes = [C.c_int]
            cls._lib = lib
        return cls._lib

    def __init__(self, dep, s1=None, s2=None, zeta=None, *, dt, ground, crough=0.0):
        self._arrays(dep, s1, s2, zeta)
        lib = self.lib()
        lib.ref_init(self.nx, self.ny, _ptr(self.dep), _ptr(self.s1), _ptr(self.s2),
                     _ptr(self.zeta), float(dt), float(ground), float(crough))
        self.grav = lib.ref_grav()
        self.celmin = lib.ref_celmin()
        assert lib.ref_mx() == self.mx

    def set_state(self, cel, xvel, yvel):
        c = np.ascontiguousarray(cel, dtype=np.float64)
        u = np.ascontiguousarray(xvel, dtype=np.float64)
        v = np.ascontiguousarray(yvel, dtype=np.float64)
        self.lib().ref_set_state(_ptr(c), _ptr(u), _ptr(v))

    def get_state(self):
        out = [np.empty((self.ny, self.nx)) for _ in range(5)]
        self.lib().ref_get_state(*[_ptr(a) for a in out])
        return tuple(out)

    def set_edges(self, edge1, edge2):
        e1 = np.ascontiguousarray(np.asarray(edge1, dtype=np.float64).T)
        e2 = np.ascontiguousarray(np.asarray(edge2, dtype=np.float64).T)
        self.lib().ref_set_edges(_ptr(e1), _ptr(e2))

    def line(self, iu: int, irw: int) -> int:
        self.lib().ref_line(iu, irw)
        return 0

    def sweep(self, iu: int, first: int, last: int) -> int:
        self.lib().ref_sweep(iu, first, last)
        return 0

    def maxwave(self):
        self.lib().ref_maxwave()

    def steps(self, n: int) -> int:
        self.lib().ref_steps(n)
        return 0


# -------------------------------------------------------------------------
# stand-alone helpers of the oracle library
def metrics(xcrd, ycrd, cartesian: bool):
    """MasterGrid.f:85-116 -> (s1 (ny,nx-1), s2 (ny-1,), zeta (ny,))."""
    x = np.ascontiguousarray(xcrd, dtype=np.float64)
    y = np.ascontiguousarray(ycrd, dtype=np.float64)
    nx, ny = len(x), len(y)
    s1 = np.zeros((ny, max(nx - 1, 1)))
    s2 = np.zeros(max(ny - 1, 1))
    zeta = np.zeros(ny)
    Oracle.lib().orc_metrics(nx, ny, _ptr(x), _ptr(y), 1 if cartesian else 0,
                             _ptr(s1), _ptr(s2), _ptr(zeta))
    return s1, s2, zeta


def pinpoint(xx, xp):
    """NestedGrid.f:129-148 -> (ip (1-based), cp)."""
    x = np.ascontiguousarray(xx, dtype=np.float64)
    ip = C.c_int(0)
    cp = C.c_double(0)
    Oracle.lib().orc_pinpoint(_ptr(x), len(x), float(xp), C.byref(ip), C.byref(cp))
    return ip.value, cp.value


def nest_maps(xp, yp, xc, yc):
    """NestedGrid.f:104-119 for one child -> dict of 1-based maps."""
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    yp = np.ascontiguousarray(yp, dtype=np.float64)
    xc = np.ascontiguousarray(xc, dtype=np.float64)
    yc = np.ascontiguousarray(yc, dtype=np.float64)
    nxc, nyc = len(xc), len(yc)
    we_ny = np.zeros(nyc, dtype=np.int32)
    we_cy = np.zeros(nyc)
    sn_nx = np.zeros(nxc, dtype=np.int32)
    sn_cx = np.zeros(nxc)
    sc = [C.c_int(0), C.c_double(0), C.c_int(0), C.c_double(0),
          C.c_int(0), C.c_double(0), C.c_int(0), C.c_double(0)]
    Oracle.lib().orc_nest_maps(len(xp), _ptr(xp), len(yp), _ptr(yp), nxc, _ptr(xc), nyc, _ptr(yc),
                               we_ny.ctypes.data_as(_ip), _ptr(we_cy),
                               sn_nx.ctypes.data_as(_ip), _ptr(sn_cx),
                               *[C.byref(s) for s in sc])
    return dict(we_ny=we_ny, we_cy=we_cy, sn_nx=sn_nx, sn_cx=sn_cx,
                w_nx=sc[0].value, w_cx=sc[1].value, e_nx=sc[2].value, e_cx=sc[3].value,
                s_ny=sc[4].value, s_cy=sc[5].value, n_ny=sc[6].value, n_cy=sc[7].value)


def feed_child(orc: Oracle, maps: dict, cuke: float, hmin: float, iarv: bool):
    """feed_children_grid.f:10-179 on the oracle's current state.

    Returns (arrived, uvq_w (nyc,3), uvq_e, uvq_s (nxc,3), uvq_n)."""
    nyc, nxc = len(maps["we_ny"]), len(maps["sn_nx"])
    w = np.zeros((3, nyc)); e = np.zeros((3, nyc))
    s = np.zeros((3, nxc)); n = np.zeros((3, nxc))
    flag = C.c_int(1 if iarv else 0)
    we_ny = np.ascontiguousarray(maps["we_ny"], dtype=np.int32)
    sn_nx = np.ascontiguousarray(maps["sn_nx"], dtype=np.int32)
    we_cy = np.ascontiguousarray(maps["we_cy"], dtype=np.float64)
    sn_cx = np.ascontiguousarray(maps["sn_cx"], dtype=np.float64)
    rc = Oracle.lib().orc_feed_child(
        C.byref(orc._g), float(cuke), float(hmin), nxc, nyc,
        we_ny.ctypes.data_as(_ip), _ptr(we_cy), sn_nx.ctypes.data_as(_ip), _ptr(sn_cx),
        maps["w_nx"], maps["w_cx"], maps["e_nx"], maps["e_cx"],
        maps["s_ny"], maps["s_cy"], maps["n_ny"], maps["n_cy"],
        C.byref(flag), _ptr(w), _ptr(e), _ptr(s), _ptr(n))
    if rc < 0:
        raise MemoryError("orc_feed_child")
    return bool(flag.value), w.T.copy(), e.T.copy(), s.T.copy(), n.T.copy()


def apply_initial_conditions(xq, yq, qq, xp, yp, pp):
    """apply_initial_conditions.f:7-70.  qq (nyq, nxq), pp (nyp, nxp) numpy
    (Fortran (nx,ny) bytes); returns pp + interpolated qq."""
    xq = np.ascontiguousarray(xq, dtype=np.float64)
    yq = np.ascontiguousarray(yq, dtype=np.float64)
    qq = np.ascontiguousarray(qq, dtype=np.float64)
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    yp = np.ascontiguousarray(yp, dtype=np.float64)
    out = np.array(pp, dtype=np.float64, order="C", copy=True)
    Oracle.lib().orc_apply_initial_conditions(len(xq), len(yq), _ptr(xq), _ptr(yq), _ptr(qq),
                                              len(xp), len(yp), _ptr(xp), _ptr(yp), _ptr(out))
    return out
