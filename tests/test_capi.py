"""C ABI checks that need no GPU: the library builds and loads, exports every
function include/cliffs_gpu.h declares, and the host-side conversions keep
the step-1 wet/dry predicates of the FP64 reference exactly."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import cliffs_gpu
import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    lib = cliffs_gpu.load_library()
    names = cliffs_gpu.header_symbols()
    assert len(names) == 40
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", cliffs_gpu.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported


def test_header_is_plain_c():
    """The boundary header compiles as C with no HIP or torch types."""
    hdr = os.path.join(ROOT, "include", "cliffs_gpu.h")
    src = '#include "%s"\nint main(void){return 0;}\n' % hdr
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-x", "c", "-", "-o", os.devnull],
                   input=src, text=True, check=True)
    # declarations only (comments stripped): plain C types, no HIP/torch types
    import re
    code = re.sub(r"/\*.*?\*/", "", open(hdr).read(), flags=re.S)
    for banned in ("hipstream", "hiperror", "hip_runtime", "torch", "float4", "__device__",
                   "std::", "#include"):
        assert banned not in code.lower(), banned


def test_argument_errors_without_compute():
    lib = cliffs_gpu.load_library()
    h = C.c_void_p()
    assert lib.cliffs_create(None, C.byref(h)) == -1
    assert b"NULL" in lib.cliffs_last_error()
    assert lib.cliffs_sync(None) == -1
    assert lib.cliffs_run(None, 1) == -1
    cour, bad = C.c_double(), C.c_longlong()
    assert lib.cliffs_health(None, C.byref(cour), C.byref(bad)) == -1
    assert lib.cliffs_destroy(None) == 0


def test_bad_descriptor_rejected_before_device_use():
    lib = cliffs_gpu.load_library()
    dep = np.ones((4, 5))
    d = cliffs_gpu._Desc(5, 4, dep.ctypes.data_as(cliffs_gpu._dp), None, None, None,
                         0.1, 0.0, 0.0, 0)
    h = C.c_void_p()
    assert lib.cliffs_create(C.byref(d), C.byref(h)) == -1      # ground must be > 0
    assert b"ground" in lib.cliffs_last_error()
    d.ground = 0.01
    assert lib.cliffs_create(C.byref(d), C.byref(h)) == -1      # s1 needed for nx > 2
    assert b"s1" in lib.cliffs_last_error()


def test_grid_without_gpu_raises_not_falls_back():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("GPU present")
    case = synth.open_ocean(n=16, seed=0)
    s1, s2, zeta = case.metrics()
    with pytest.raises(cliffs_gpu.CliffsError) as e:
        cliffs_gpu.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground)
    assert e.value.code == -5


def test_conversion_keeps_fp64_predicates():
    """Upload/download through the same host conversion the library uses
    (compiled into the test emulator): celerities on either side of c_min
    and depths on either side of `ground`, to the last bit."""
    import emul
    ground = 0.01
    g = synth.GRAV
    cm = np.sqrt(g * ground)
    cs = np.array([cm, np.nextafter(cm, 0), np.nextafter(cm, 1), 0.0, -0.01, cm * 1.0000001,
                   3.0, 198.1, 0.5 * cm])
    ds = np.array([ground, np.nextafter(ground, 0), np.nextafter(ground, 1), -2.0, 0.0, 4000.0,
                   1.0, 3999.7, 0.002])
    C_, D_ = np.meshgrid(cs, ds)
    dep = D_.copy()
    cel = C_.copy()
    e = emul.EmulGrid(dep, np.ones((dep.shape[0], dep.shape[1] - 1)), np.ones(dep.shape[0] - 1),
                      np.zeros(dep.shape[0]), dt=0.01, ground=ground)
    e.upload(cel, np.zeros_like(cel), np.zeros_like(cel))
    c2, _, _ = e.download()
    assert np.array_equal(c2 <= cm, cel <= cm)
    wet = cel > cm
    # wet cells keep their surface elevation to FP32 precision of the perturbation
    eta0 = cel ** 2 / g - dep
    eta1 = c2 ** 2 / g - dep
    assert np.abs(eta0 - eta1)[wet].max() < 1e-6 * (1 + np.abs(eta0[wet]).max())


def test_fortran_module_binds_every_header_function():
    """cliffs_gpu_mod.f90 declares a bind(C) interface for every function of
    include/cliffs_gpu.h, under the same name."""
    import re
    path = os.path.join(os.path.dirname(cliffs_gpu.HEADER), "..", "cliffs-src_amd", "fortran",
                        "cliffs_gpu_mod.f90")
    with open(path) as f:
        txt = f.read()
    bound = set(re.findall(r'bind\(C, name="(cliffs_\w+)"\)', txt))
    assert set(cliffs_gpu.header_symbols()) == bound
    # and of include/cliffs_fp64.h (the FP64 line mode)
    hdr64 = os.path.join(os.path.dirname(cliffs_gpu.HEADER), "cliffs_fp64.h")
    code = re.sub(r"/\*.*?\*/", "", open(hdr64).read(), flags=re.S)
    want64 = set(re.findall(r"^\s*int\s+(cliffs64_\w+)\s*\(", code, re.M))
    assert len(want64) == 13 and want64 == set(re.findall(r'bind\(C, name="(cliffs64_\w+)"\)', txt))


def test_surrogate_header_exported_and_plain_c():
    """include/cliffs_surrogate.h: plain C, every declared function exported."""
    hdr = os.path.join(ROOT, "include", "cliffs_surrogate.h")
    names = cliffs_gpu.header_symbols(hdr)
    assert len(names) == 14 and "cliffs_sur_predict" in names and "cliffs_sur_set_data" in names
    out = subprocess.run(["nm", "-D", "--defined-only", cliffs_gpu.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported
    src = '#include "%s"\nint main(void){return 0;}\n' % hdr
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-x", "c", "-", "-o", os.devnull],
                   input=src, text=True, check=True)
