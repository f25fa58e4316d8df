"""Shared test setup.

Markers:
  gpu -- needs a HIP device (MI355X); compares the HIP path against the
         CPU oracle (oracle/), which is only ever the checker.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cliffs-src_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)

NATIVE = os.path.join(ROOT, "tests", "native")
EMUL_SO = os.path.join(NATIVE, "libline_emul.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


def build_emulator():
    src = os.path.join(NATIVE, "line_emul.cpp")
    if (not os.path.exists(EMUL_SO)) or os.path.getmtime(EMUL_SO) < max(
            os.path.getmtime(src),
            os.path.getmtime(os.path.join(PKG, "csrc", "vtcs_node.h")),
            os.path.getmtime(os.path.join(PKG, "csrc", "convert.h"))):
        # contraction off, as the library is built: the emulator rounds every
        # operation as the device does
        subprocess.run(["g++", "-O2", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC", "-shared",
                        "-std=c++17", src, "-o", EMUL_SO], check=True)
    return EMUL_SO


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    if not os.path.exists(oracle.ORACLE_SO):
        oracle.build(with_ref=False)
    return oracle


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, loaded; fails loudly when it is missing on a machine with a GPU."""
    if not gpu_available():
        pytest.skip("no GPU")
    import cliffs_gpu
    cliffs_gpu.load_library()
    return cliffs_gpu


def eta_of(cel, dep, grav):
    return cel * cel / grav - dep
