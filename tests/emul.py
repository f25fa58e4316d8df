"""ctypes wrapper of the test-only CPU emulator (tests/native/line_emul.cpp)
with the same interface as ``cliffs_gpu.Grid``, so that parity tests run
unchanged against the device arithmetic on a machine without a GPU."""
import ctypes as C

import numpy as np

from conftest import build_emulator

_dp = C.POINTER(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(build_emulator())
        L.emul_create.restype = C.c_void_p
        L.emul_create.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                  C.c_double]
        L.emul_destroy.argtypes = [C.c_void_p]
        L.emul_upload.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.emul_download.argtypes = [C.c_void_p] + [_dp] * 5
        L.emul_sweep.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.emul_maxwave.argtypes = [C.c_void_p]
        L.emul_steps.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


class EmulGrid:
    def __init__(self, dep, s1=None, s2=None, zeta=None, *, dt, ground, crough=0.0, device=0):
        self.dep = np.ascontiguousarray(dep, dtype=np.float64)
        ny, nx = self.dep.shape
        self.nx, self.ny = nx, ny
        self._s1 = np.ascontiguousarray(s1 if s1 is not None else np.zeros((ny, max(nx - 1, 1))))
        self._s2 = np.ascontiguousarray(s2 if s2 is not None else np.zeros(max(ny - 1, 1)))
        self._z = np.ascontiguousarray(zeta if zeta is not None else np.zeros(ny))
        self._h = lib().emul_create(nx, ny, _p(self.dep), _p(self._s1), _p(self._s2), _p(self._z),
                                    float(dt), float(ground), float(crough))
        self.mx = max(ny if nx > 2 else 1, nx if ny > 2 else 1)
        self.grav = float(np.float32(9.8))
        self.celmin = float(np.sqrt(self.grav * ground))
        self._state = None

    def __del__(self):
        if getattr(self, "_h", None):
            lib().emul_destroy(self._h)
            self._h = None

    def upload(self, cel, xvel, yvel):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (cel, xvel, yvel)]
        lib().emul_upload(self._h, *[_p(x) for x in a])

    def _all(self):
        out = [np.empty((self.ny, self.nx)) for _ in range(5)]
        lib().emul_download(self._h, *[_p(x) for x in out])
        return out

    def download(self):
        return tuple(self._all()[:3])

    def download_max(self):
        return tuple(self._all()[3:])

    @staticmethod
    def _edge(e):
        return None if e is None else np.ascontiguousarray(np.asarray(e, dtype=np.float64).T)

    def sweep(self, iu, first, last, edge1=None, edge2=None):
        e1, e2 = self._edge(edge1), self._edge(edge2)
        mx = (e1 if e1 is not None else e2).shape[1] if (e1 is not None or e2 is not None) else 0
        lib().emul_sweep(self._h, iu, first, last, mx, _p(e1), _p(e2))

    def maxwave(self):
        lib().emul_maxwave(self._h)

    def step(self, xedge1=None, xedge2=None, yedge1=None, yedge2=None):
        nx, ny = self.nx, self.ny
        if nx > 2:
            self.sweep(1, 2 if ny > 2 else 1, ny - 1 if ny > 2 else 1, xedge1, xedge2)
        if ny > 2:
            self.sweep(2, 2 if nx > 2 else 1, nx - 1 if nx > 2 else 1, yedge1, yedge2)
        self.maxwave()

    def run(self, nsteps):
        lib().emul_steps(self._h, int(nsteps), 0, None, None, None, None)

    def sync(self):
        pass

    def gauges(self, I, J):
        c, u, v = self.download()
        eta, uu, vv = [], [], []
        for i, j in zip(I, J):
            cc = c[j - 1, i - 1]
            eta.append(cc * cc / self.grav - self.dep[j - 1, i - 1] if cc > self.celmin
                       else float(np.float32(-1e32)))
            uu.append(u[j - 1, i - 1])
            vv.append(v[j - 1, i - 1])
        return np.array(eta), np.array(uu), np.array(vv)
