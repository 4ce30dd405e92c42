"""Write the golden fixtures of tests/golden/ from the reference's own code.

Runs only on a machine that holds the reference's sources under /root/reference: it builds
oracle/_ref/libbirp_ref.so from the reference's src/DataCube.cpp, src/Cmem.cpp and src/Helper.cpp
(oracle/Makefile) and records inputs and the reference's outputs as JSON:

  datacube_samples.json  a seeded 12x9x10 cube written in the reference's .dat format and loaded
                         by DataCube::Load (DataCube.cpp:28-138); GetSample (trilinear and nearest)
                         at random, grid, boundary and out-of-bounds points (DataCube.cpp:203-277)
  cmem_samples.json      CMEM::GetSample (Cmem.cpp:117-186) after CMEM::Init() and after
                         several Init(...) parameter sets, at random points
  helper_matrix.json     Helper::MatrixMultiply and Helper::GetInverse (Helper.cpp:70-264) on
                         rotation products built like Camera::Integrate's, plus non-rotations
  helper_pointing.json   the pointing utilities (Helper.cpp:266-585): Gei2GseTransforms,
                         RaDecToGSI and RaDecGSEConversion at random and quadrant-boundary
                         angles, TwoMin with random masks, XYToRaDec on a polar sky grid (10 deg,
                         across the 0/360 seam), Interp1 and findNearestNeighbourIndex on ragged
                         grids, XYZToSpherical(Alt) at random, axis and zero points

Floats are stored as their IEEE bit patterns (uint32) so the fixtures are exact.
Usage: python tests/golden/make_golden.py
"""
import ctypes as C
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32).tolist()


def write_dat(path, xs, ys, zs, data):
    """The .dat layout DataCube::Load parses: metadata line 1, dimensions 'nx ny nz' on line 10,
    then the x, y and z coordinates (one line each), then one line per x-row of values, a blank
    line after each z-slice (DataCube.cpp:49-125)."""
    nz, ny, nx = data.shape
    with open(path, "w") as f:
        f.write("synthetic cube written by tests/golden/make_golden.py\n")
        for _ in range(8):
            f.write("#\n")
        f.write(f"{nx} {ny} {nz}\n")
        for c in (xs, ys, zs):
            f.write(" ".join("%.9g" % v for v in c) + "\n")
        for z in range(nz):
            for y in range(ny):
                f.write(" ".join("%.9g" % v for v in data[z, y]) + "\n")
            f.write("\n")


def datacube(R, rng):
    nx, ny, nz = 12, 9, 10
    xs = (np.float32(-10) + np.arange(nx, dtype=np.float32) * np.float32(30 / (nx - 1))).astype(np.float32)
    ys = (np.float32(-15) + np.arange(ny, dtype=np.float32) * np.float32(30 / (ny - 1))).astype(np.float32)
    zs = (np.float32(-15) + np.arange(nz, dtype=np.float32) * np.float32(30 / (nz - 1))).astype(np.float32)
    data = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    data[rng.random(data.shape) < 0.1] = 0  # zeros and mixed signs exercise std::lerp's branches
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "golden.dat")
        write_dat(path, xs, ys, zs, data)
        dat_text = open(path).read()
        out = {"nx": nx, "ny": ny, "nz": nz, "xs": bits(xs), "ys": bits(ys), "zs": bits(zs), "values": bits(data),
               "dat_text": dat_text}
        n = 4000
        pts = np.column_stack([rng.uniform(-12, 22, n), rng.uniform(-17, 17, n), rng.uniform(-17, 17, n)]).astype(np.float32)
        grid = np.array([[xs[i], ys[j], zs[k]] for i, j, k in [(0, 0, 0), (3, 4, 5), (10, 7, 8), (11, 0, 0), (0, 8, 0), (5, 5, 9)]],
                        np.float32)
        edge = np.array([[xs[-1], 0, 0], [np.nextafter(xs[-1], np.float32(0)), 0, 0], [xs[0], ys[0], zs[0]],
                         [np.nextafter(xs[0], np.float32(-20)), 0, 0], [0, ys[-1], 0], [0, 0, zs[-1]]], np.float32)
        pts = np.vstack([pts, grid, edge]).astype(np.float32)
        out["points"] = bits(pts)
        for tri in (1, 0):
            h = R.ref_cube_load(path.encode(), tri)
            assert h, "DataCube::Load failed"
            s3 = (C.c_int * 3)()
            sp = (C.c_float * 3)()
            org = (C.c_float * 3)()
            R.ref_cube_info(h, s3, sp, org)
            out["size"] = list(s3)
            out["spacing"] = bits(list(sp))
            out["origin"] = bits(list(org))
            loaded = np.zeros(nx * ny * nz, np.float32)
            assert R.ref_cube_values(h, loaded.ctypes.data_as(C.POINTER(C.c_float)), loaded.size) == 0
            assert np.array_equal(loaded.view(np.uint32), data.ravel().view(np.uint32)), "loader round trip"
            err = C.c_int(0)
            vals = []
            for p in pts:
                v = R.ref_cube_sample(h, float(p[0]), float(p[1]), float(p[2]), C.byref(err))
                assert err.value == 0
                vals.append(v)
            out["trilinear" if tri else "nearest"] = bits(vals)
            R.ref_cube_free(h)
    return out


CMEM_PARAM_SETS = [
    None,  # CMEM::Init()
    # tests/CMEM_test.cpp:63 binds its arguments shifted by one (p1 -> p0_passed, ...), kept as is
    dict(v=[400, 0, 0], b=[0, 0, 4], dipole=0.0, p0=1, p1=4, p2=6, p3=3, B=3.6, alpha=-1.0),
    dict(v=[550, 20, -10], b=[1, -2, -3], dipole=0.2, p0=0.9, p1=2, p2=2, p3=5, B=2.5, alpha=2.0, beta=-1.2, bs=13.0,
         A1=8e-6, A2=3e-6, density=7.5),
    dict(v=[350, 0, 0], b=[0, 0, 8], dipole=-0.15, p0=1.1, p1=1, p2=4, p3=3, ay_bs=0.7, az_bs=0.8),
    # the CLI's own parameter Init (main.cpp:285-294 with p = {1,1,3,4} and B, alpha, beta = NULL)
    dict(dipole=0.0, p0=1.0, p1=1, p2=3, p3=4, B=0.0, alpha=0.0, beta=0.0, bs=0.0, A1=0.0, A2=0.0),
]


def cmem(R, rng):
    n = 1500
    pts = np.vstack([rng.uniform(-25, 25, (n, 3)),
                     np.array([[8, 4, 4], [4, 12, 10], [3, 1.5, 1.2], [8, 10, 10], [8, -10, -10], [8, 10, -10],
                               [12, 12, 10], [12, 10, 10], [0, 5, 5], [5, 0, 3], [5, 3, 0], [-20, 0.5, 0.5]])]).astype(np.float32)
    out = {"points": bits(pts), "sets": []}
    for params in CMEM_PARAM_SETS:
        h = O.ref_cmem(params)
        vals = [R.ref_cmem_sample(h, float(p[0]), float(p[1]), float(p[2])) for p in pts]
        R.ref_cmem_free(h)
        out["sets"].append({"params": params, "values": bits(vals)})
    return out


def helper(R, rng):
    mats = []
    for _ in range(300):
        th, ph, ro = rng.uniform(0, 2 * np.pi, 3).astype(np.float32)
        rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]], np.float32)
        rx = np.array([[1, 0, 0], [0, np.cos(ph), -np.sin(ph)], [0, np.sin(ph), np.cos(ph)]], np.float32)
        rz2 = np.array([[np.cos(ro), -np.sin(ro), 0], [np.sin(ro), np.cos(ro), 0], [0, 0, 1]], np.float32)
        mats.append((rz2, rx, rz))
    gen = [rng.standard_normal((3, 3)).astype(np.float32) for _ in range(100)]
    out = {"products": [], "general": []}
    fp = C.POINTER(C.c_float)
    for rz2, rx, rz in mats:
        a = np.zeros(9, np.float32)
        R.ref_matmul3(rz2.ravel().ctypes.data_as(fp), rx.ravel().ctypes.data_as(fp), a.ctypes.data_as(fp))
        b = np.zeros(9, np.float32)
        R.ref_matmul3(a.ctypes.data_as(fp), rz.ravel().ctypes.data_as(fp), b.ctypes.data_as(fp))
        inv = np.zeros(9, np.float32)
        rc = R.ref_get_inverse3(b.ctypes.data_as(fp), inv.ctypes.data_as(fp))
        out["products"].append({"rz2": bits(rz2.ravel()), "rx": bits(rx.ravel()), "rz": bits(rz.ravel()), "ab": bits(a),
                                "abc": bits(b), "inverse": bits(inv), "rc": rc})
    for g in gen + [np.array([[1, 2, 3], [2, 4, 6], [1, 0, 1]], np.float32)]:
        inv = np.zeros(9, np.float32)
        rc = R.ref_get_inverse3(g.ravel().copy().ctypes.data_as(fp), inv.ctypes.data_as(fp))
        out["general"].append({"m": bits(g.ravel()), "inverse": bits(inv), "rc": rc})
    return out


def pointing(R, rng):
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int)
    f = np.float32
    out = {}
    m = np.zeros(16, f)
    R.ref_gei2gse(m.ctypes.data_as(fp))
    out["gei2gse"] = bits(m)

    ra = rng.uniform(0, 360, 300).astype(f).tolist() + [0, 45, 89.5, 90, 90.5, 135, 180, 225, 270, 315, 359.5, 360]
    dec = rng.uniform(-90, 90, 300).astype(f).tolist() + [0, 30, -30, 90, -90, 60, 0, -60, 45, 10, -89.9, 89.9]
    out["radec"] = []
    for a, d in zip(ra, dec):
        g = np.zeros(3, f)
        e = np.zeros(3, f)
        R.ref_radec_to_gsi(a, d, g.ctypes.data_as(fp))
        R.ref_radec_gse(a, d, m.ctypes.data_as(fp), e.ctypes.data_as(fp))
        out["radec"].append({"ra": bits([a])[0], "dec": bits([d])[0], "gsi": bits(g), "gse": bits(e)})

    out["two_min"] = []
    for _ in range(150):
        n = int(rng.integers(3, 24))
        arr = rng.standard_normal(n).astype(f)
        if rng.random() < 0.3:
            arr = np.round(arr).astype(f)  # ties
        mask = np.sort(rng.choice(n + 2, size=int(rng.integers(0, n + 2)), replace=False)).astype(np.int32)
        r2 = np.zeros(2, f)
        R.ref_two_min(arr.ctypes.data_as(fp), n, mask.ctypes.data_as(ip), len(mask), r2.ctypes.data_as(fp))
        out["two_min"].append({"arr": bits(arr), "mask": mask.tolist(), "out": bits(r2)})

    # a polar sky grid: RA 0..350 and Dec -60..80 in 10 degree steps, image position (90 - dec) at angle ra
    skysep = f(10)
    ras, decs = np.meshgrid(np.arange(0, 360, 10, dtype=f), np.arange(-60, 90, 10, dtype=f))
    sky1, sky2 = ras.ravel().astype(f), decs.ravel().astype(f)
    rad = np.deg2rad(sky1.astype(np.float64))
    xsky = ((90 - sky2) * np.cos(rad)).astype(f)
    ysky = ((90 - sky2) * np.sin(rad)).astype(f)
    out["sky"] = {"skysep": bits([skysep])[0], "xsky": bits(xsky), "ysky": bits(ysky), "sky1": bits(sky1), "sky2": bits(sky2),
                  "points": []}
    r = rng.uniform(15, 140, 300)
    th = rng.uniform(0, 2 * np.pi, 300)
    th[:20] = rng.uniform(-0.05, 0.05, 20)  # near the 0/360 seam
    for x, y in zip((r * np.cos(th)).astype(f), (r * np.sin(th)).astype(f)):
        r2 = np.zeros(2, f)
        R.ref_xy_to_radec(x, y, xsky.ctypes.data_as(fp), ysky.ctypes.data_as(fp), sky1.ctypes.data_as(fp),
                          sky2.ctypes.data_as(fp), len(sky1), skysep, r2.ctypes.data_as(fp))
        out["sky"]["points"].append({"x": bits([x])[0], "y": bits([y])[0], "out": bits(r2)})

    out["interp1"] = []
    for _ in range(60):
        n = int(rng.integers(2, 30))
        x = np.sort(rng.uniform(-10, 10, n)).astype(f)
        x = np.unique(x)
        y = rng.standard_normal(len(x)).astype(f)
        k = int(rng.integers(1, 40))
        xn = rng.uniform(float(x[0]), float(x[-1]) + 2, k).astype(f)
        xn = xn[xn > x[0]]
        if len(x) < 2 or len(xn) == 0:
            continue
        yn = np.zeros(len(xn), f)
        R.ref_interp1(x.ctypes.data_as(fp), y.ctypes.data_as(fp), len(x), xn.ctypes.data_as(fp), len(xn),
                      yn.ctypes.data_as(fp))
        out["interp1"].append({"x": bits(x), "y": bits(y), "xn": bits(xn), "yn": bits(yn)})

    out["nearest"] = []
    for _ in range(100):
        n = int(rng.integers(1, 20))
        x = rng.uniform(-5, 5, n).astype(f)
        if rng.random() < 0.3:
            x = np.round(x).astype(f)
        v = f(rng.uniform(-6, 6)) if rng.random() < 0.8 else x[int(rng.integers(0, n))]
        out["nearest"].append({"x": bits(x), "value": bits([v])[0],
                               "index": R.ref_nearest_below(v, x.ctypes.data_as(fp), n)})

    pts = rng.standard_normal((200, 3)).astype(f) * f(20)
    pts = np.concatenate([pts, np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1],
                                         [-0.0, 0, 0], [0, -0.0, -2], [3, -4, 0]], f)])
    out["spherical"] = []
    for p in pts:
        p = np.ascontiguousarray(p)
        a = np.zeros(3, f)
        b = np.zeros(3, f)
        R.ref_xyz_to_spherical(p.ctypes.data_as(fp), 0, a.ctypes.data_as(fp))
        R.ref_xyz_to_spherical(p.ctypes.data_as(fp), 1, b.ctypes.data_as(fp))
        out["spherical"].append({"xyz": bits(p), "sph": bits(a), "alt": bits(b)})
    return out


def main():
    O.build()
    if not O.ref_available():
        sys.exit("oracle/_ref/libbirp_ref.so is not available (the reference sources are needed)")
    R = O.ref()
    rng = np.random.default_rng(20240601)
    for name, fn in (("datacube_samples.json", datacube), ("cmem_samples.json", cmem), ("helper_matrix.json", helper)):
        data = fn(R, rng)
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print("wrote", name, os.path.getsize(os.path.join(HERE, name)), "bytes")
    data = pointing(R, np.random.default_rng(20240602))
    with open(os.path.join(HERE, "helper_pointing.json"), "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print("wrote helper_pointing.json", os.path.getsize(os.path.join(HERE, "helper_pointing.json")), "bytes")


if __name__ == "__main__":
    main()
