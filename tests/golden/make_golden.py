"""Write the committed golden fixtures under tests/golden/.

Run in the build container (needs /root/reference and oracle/_ref built):
    python tests/golden/make_golden.py

* slpbeach.npz -- data files of the reference's own 1-D run-up example
  (Examples/slpbeach.zip): grid slpbeach1D.nc, initial fields
  solitary.0185_{h,u}.nc and solitary.30_{h,u}.nc, the analytic profiles
  canonical_profiles.txt and the run parameters of slpbeach_params.txt.
* refcore_cases.npz -- inputs and outputs of the reference's own cliffs.f
  (compiled in place by oracle/Makefile into oracle/_ref) on seeded cases:
  single lines in both directions, dense wet/dry patterns, edges,
  friction, spherical term, 1-D and 2-D steps.
"""
import io
import os
import sys
import zipfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "cliffs-src_amd"))

REF_ZIP = "/root/reference/Examples/slpbeach.zip"


def slpbeach():
    from scipy.io import netcdf_file
    z = zipfile.ZipFile(REF_ZIP)

    def nc(name):
        f = netcdf_file(io.BytesIO(z.read("slpbeach/" + name)), "r", mmap=False)
        out = {k: np.array(v.data) for k, v in f.variables.items()}
        f.close()
        return out

    g = nc("slpbeach1D.nc")
    h1, u1 = nc("solitary.0185_h.nc"), nc("solitary.0185_u.nc")
    h3, u3 = nc("solitary.30_h.nc"), nc("solitary.30_u.nc")
    rows = []
    for line in z.read("slpbeach/canonical_profiles.txt").decode().splitlines():
        parts = line.split()
        if parts:
            rows.append([float(p) for p in parts])
    canon = np.array(rows)
    np.savez_compressed(
        os.path.join(HERE, "slpbeach.npz"),
        x=g["xxx"].astype(np.float64), y=g["yyy"].astype(np.float64),
        bathy=g["bathy"].astype(np.float64),            # FP32 in the file, widened as MasterGrid.f:75
        xq=h1["xxx"], h0185=h1["ha"][0, 0], u0185=u1["ua"][0, 0],
        xq30=h3["xxx"], h30=h3["ha"][0, 0], u30=u3["ua"][0, 0],
        canonical=canon,
        dt=0.03, ground=0.003, crough=0.0, steps=2000)


def random_case(rng, nx, ny, spherical=False):
    import synth
    dep = rng.uniform(-1.5, 3.0, (ny, nx))
    eta = rng.normal(0, 0.3, (ny, nx))
    ground = 0.01
    cel = synth.depth_to_celerity(dep + eta, ground)
    # a few cells exactly at celmin and a few negative celerities
    cm = np.sqrt(synth.GRAV * ground)
    k = rng.integers(0, nx * ny, 5)
    cel.flat[k[:3]] = cm
    cel.flat[k[3:]] = -0.01
    u = rng.normal(0, 0.5, (ny, nx))
    v = rng.normal(0, 0.5, (ny, nx))
    if spherical:
        x = 140 + np.cumsum(rng.uniform(0.5, 1.5, nx)) / 120.0
        y = 38 + np.cumsum(rng.uniform(0.5, 1.5, ny)) / 120.0
    else:
        x = np.cumsum(rng.uniform(0.5, 1.5, nx))
        y = np.cumsum(rng.uniform(0.5, 1.5, ny))
    s1, s2, zeta = synth.metrics(x, y, not spherical)
    return dep, cel, u, v, s1, s2, zeta


def refcore_cases():
    import oracle as O
    rng = np.random.default_rng(20261203)
    out = {}
    specs = [(37, 23, False, 0.0, "lines"), (29, 41, True, 0.003, "lines"),
             (64, 1, False, 0.0, "steps"), (1, 53, False, 0.002, "steps"),
             (33, 27, False, 0.0, "steps"), (26, 31, True, 0.003, "steps")]
    for n, (nx, ny, sph, crough, kind) in enumerate(specs):
        dep, cel, u, v, s1, s2, zeta = random_case(rng, nx, ny, sph)
        dt = 0.02 if not sph else 0.02
        mx = O.line_ranges(nx, ny)[6]
        e1 = rng.normal(0, 0.2, (mx, 3))
        e2 = rng.normal(0, 0.2, (mx, 3))
        r = O.RefCore(dep, s1, s2, zeta, dt=dt, ground=0.01, crough=crough)
        r.set_state(cel, u, v)
        r.set_edges(e1, e2)
        if kind == "lines":
            for iu, nl in ((1, ny), (2, nx)):
                for irw in range(1, nl + 1):
                    r.line(iu, irw)
        else:
            r.steps(3)
        c2, u2, v2, hm, vm = r.get_state()
        p = f"c{n}_"
        out.update({p + "dep": dep, p + "cel": cel, p + "u": u, p + "v": v, p + "s1": s1,
                    p + "s2": s2, p + "zeta": zeta, p + "e1": e1, p + "e2": e2,
                    p + "out_cel": c2, p + "out_u": u2, p + "out_v": v2, p + "out_hmax": hm,
                    p + "out_velmax": vm,
                    p + "meta": np.array([nx, ny, int(sph), crough, dt, 0.01,
                                          1 if kind == "lines" else 0])})
    np.savez_compressed(os.path.join(HERE, "refcore_cases.npz"), ncases=len(specs), **out)


if __name__ == "__main__":
    slpbeach()
    refcore_cases()
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))
