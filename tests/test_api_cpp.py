"""The C++ API (include/DataCube.h, Cmem.h, Camera.h, Helper.h -> libbirp_api.so), driven through
tests/cpp/api_tests.cpp: host-side parts against the reference's golden vectors and the oracle,
GPU renders against the oracle bit for bit (CMEM renders by the rule in tests/parity.py)."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from parity import CMEM_EDGE_BAND, cmem_render_parity

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
LIB = os.path.join(ROOT, "birp_cpp_amd", "lib")


def f32(b):
    return np.asarray(b, np.uint32).view(np.float32)


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    """The C++ API driver, prebuilt into tests/_build by `make -C tests` (__graft_entry__.build());
    compiled here only when that build is missing or older than its source."""
    if not os.path.exists(os.path.join(LIB, "libbirp_api.so")):
        pytest.fail("birp_cpp_amd/lib/libbirp_api.so is missing: run `make -C birp_cpp_amd`")
    src = os.path.join(HERE, "cpp", "api_tests.cpp")
    built = os.path.join(HERE, "_build", "api_tests")
    inc = os.path.join(ROOT, "include")
    deps = [src, os.path.join(LIB, "libbirp_api.so")] + [os.path.join(inc, h) for h in os.listdir(inc)]
    if os.path.exists(built) and os.path.getmtime(built) >= max(os.path.getmtime(d) for d in deps):
        return built
    out = tmp_path_factory.mktemp("api") / "api_tests"
    subprocess.run(["g++", "-O2", "-std=c++20", "-I", os.path.join(ROOT, "include"), src,
                    "-o", str(out), "-L", LIB, "-lbirp_api", "-lbirp_hip", f"-Wl,-rpath,{LIB}"], check=True)
    return str(out)


def run(exe, *args, stdin=""):
    r = subprocess.run([exe, *map(str, args)], input=stdin, capture_output=True, text=True)
    return r.returncode, r.stdout, r.stderr


@pytest.fixture(scope="module")
def golden_dat(tmp_path_factory):
    g = json.load(open(os.path.join(GOLDEN, "datacube_samples.json")))
    p = tmp_path_factory.mktemp("dat") / "golden.dat"
    p.write_text(g["dat_text"])
    return str(p), g


# ---------------------------------------------------------------------------------- host side
def test_helper_inverse_matches_reference(exe):
    g = json.load(open(os.path.join(GOLDEN, "helper_matrix.json")))
    mats = [p["abc"] for p in g["products"]] + [p["m"] for p in g["general"]]
    want = [p["inverse"] for p in g["products"]] + [p["inverse"] if p["rc"] == 0 else None for p in g["general"]]
    rc, out, err = run(exe, "inverse", stdin="\n".join(" ".join(map(str, m)) for m in mats) + "\n")
    assert rc == 0, err
    lines = out.strip().split("\n")
    assert len(lines) == len(mats)
    for line, w in zip(lines, want):
        if w is None:
            assert line.startswith("throw Determinant is 0")
        else:
            assert [int(x) for x in line.split()] == w


def test_helper_pointing_matches_reference(exe):
    """Gei2GseTransforms, RaDecToGSI, RaDecGSEConversion, TwoMin, XYToRaDec, Interp1,
    findNearestNeighbourIndex and XYZToSpherical(Alt) against the reference's own outputs
    (tests/golden/helper_pointing.json), bit for bit."""
    g = json.load(open(os.path.join(GOLDEN, "helper_pointing.json")))
    j = lambda v: " ".join(map(str, v))  # noqa: E731
    lines, want = ["gei"], [g["gei2gse"]]
    for c in g["radec"]:
        lines += [f"gsi {c['ra']} {c['dec']}", f"gse {c['ra']} {c['dec']} {j(g['gei2gse'])}"]
        want += [c["gsi"], c["gse"]]
    for c in g["two_min"]:
        lines.append(f"twomin {len(c['arr'])} {j(c['arr'])} {len(c['mask'])} {j(c['mask'])}")
        want.append(c["out"])
    s = g["sky"]
    lines.append(f"sky {len(s['xsky'])} {s['skysep']} {j(s['xsky'])} {j(s['ysky'])} {j(s['sky1'])} {j(s['sky2'])}")
    for c in s["points"]:
        lines.append(f"xyradec {c['x']} {c['y']}")
        want.append(c["out"])
    for c in g["interp1"]:
        lines.append(f"interp {len(c['x'])} {len(c['xn'])} {j(c['x'])} {j(c['y'])} {j(c['xn'])}")
        want.append(c["yn"])
    for c in g["nearest"]:
        lines.append(f"nearest {len(c['x'])} {c['value']} {j(c['x'])}")
        want.append([c["index"]])
    for c in g["spherical"]:
        lines.append(f"sph {j(c['xyz'])}")
        want.append(c["sph"] + c["alt"])
    # below every grid point: the reference reads slope[-1]; this API throws instead
    one = int(np.float32(1).view(np.uint32))
    lines.append(f"interp 2 1 {one} {int(np.float32(2).view(np.uint32))} 0 0 {one}")
    rc, out, err = run(exe, "pointing", stdin="\n".join(lines) + "\n")
    assert rc == 0, err
    got = out.strip().split("\n")
    assert len(got) == len(want) + 1
    for k, (line, w) in enumerate(zip(got, want)):
        assert [int(x) for x in line.split()] == w, k
    assert got[-1] == "throw"
    # sanity of the fixture itself: the GEI->GSE matrix is a rotation about x then z (orthonormal)
    m = f32(g["gei2gse"]).reshape(4, 4).astype(np.float64)
    assert np.allclose(m @ m.T, np.eye(4), atol=1e-6)
    # RA/Dec -> unit vectors of the right quadrant
    for c in g["radec"][:300]:
        ra, dec = float(f32([c["ra"]])[0]), float(f32([c["dec"]])[0])
        v = f32(c["gsi"]).astype(np.float64)
        assert abs(np.linalg.norm(v) - 1) < 1e-5
        e = np.array([np.cos(np.radians(dec)) * np.cos(np.radians(ra)), np.cos(np.radians(dec)) * np.sin(np.radians(ra)),
                      np.sin(np.radians(dec))])
        if 0 < ra < 270 and abs(ra - 90) > 1 and abs(ra - 180) > 1:
            assert np.allclose(v, e, atol=2e-4), (ra, dec, v, e)


@pytest.mark.parametrize("which,params", [("default", None),
                                          ("test_params", dict(v=[400, 0, 0], b=[0, 0, 4], dipole=0.0, p0=1, p1=4, p2=6, p3=3,
                                                               B=3.6, alpha=-1.0)),
                                          ("cli_params", dict(dipole=0.0, p0=1.0, p1=1, p2=3, p3=4, B=0.0, alpha=0.0, beta=0.0,
                                                              bs=0.0, A1=0.0, A2=0.0))])
def test_cmem_init_state_matches_oracle(exe, oracle, which, params):
    rc, out, err = run(exe, "cmem_state", which)
    assert rc == 0, err
    got = [int(x) for x in out.split()]
    st = O.OracleCmem(params).state()
    names = ["p0", "p1", "p2", "p3", "bs", "A1", "A2", "B", "alpha", "beta", "ay_bs", "az_bs"]
    want = [st[n] for n in names] + [st["a"][5], st["a"][21]] + list(st["beta_c"]) + \
           [st[n] for n in ("charlie", "dn", "ds", "theta_n", "theta_s", "r0_lin")]
    assert got == np.asarray(want, np.float32).view(np.uint32).tolist()


def test_load_matches_reference_loader(exe, golden_dat):
    path, g = golden_dat
    rc, out, err = run(exe, "load", path)
    assert rc == 0, err
    lines = out.split("\n")
    assert [int(x) for x in lines[0].split()] == g["size"]
    assert [int(x) for x in lines[1].split()] == g["spacing"]
    assert [int(x) for x in lines[2].split()] == g["origin"]
    vals = [int(x) for x in lines[3:] if x]
    assert vals == np.asarray(g["values"]).ravel().tolist()


def test_parallel_load_of_a_large_varied_file(exe, tmp_path):
    """DataCube::Load splits the numbers after line 10 over threads by runs of whole lines: a
    70 x 60 x 50 cube (3.6 MB of text) written with mixed number forms ('%.9g', '%.8e', a leading
    '+', subnormals, zeros, negatives), blanks (runs of spaces, tabs) and CRLF line ends loads
    with every value bitwise equal to the float32 it was printed from."""
    rng = np.random.default_rng(8)
    nz, ny, nx = 50, 60, 70
    vals = (rng.standard_normal((nz, ny, nx)) * np.exp(rng.uniform(-30, 30, (nz, ny, nx)))).astype(np.float32)
    vals[rng.random(vals.shape) < 0.05] = 0.0
    vals.ravel()[::997] = np.float32(1e-40)  # a subnormal
    fmts = ["%.9g", "%.8e", "+%.9g"]
    seps = [" ", "  ", "\t", " \t "]
    p = tmp_path / "varied.dat"
    with open(p, "w", newline="") as f:
        f.write("cube\r\n" + "#\r\n" * 8 + f"{nx} {ny} {nz}\r\n")
        for n in (nx, ny, nz):
            f.write(" ".join("%.9g" % x for x in np.arange(n, dtype=np.float32) * np.float32(0.5)) + "\r\n")
        k = 0
        for z in range(nz):
            for y in range(ny):
                toks = []
                for x in vals[z, y]:
                    fmt = fmts[k % 3]
                    toks.append((fmt if x >= 0 or fmt[0] != "+" else fmt[1:]) % x)
                    toks.append(seps[k % 4])
                    k += 1
                f.write("".join(toks[:-1]) + ("\r\n" if (z + y) % 2 else "\n"))
    assert os.path.getsize(p) > 3_000_000
    rc, out, err = run(exe, "load", p)
    assert rc == 0, err
    lines = out.split("\n")
    assert [int(x) for x in lines[0].split()] == [nx, ny, nz]
    got = np.array([int(x) for x in lines[3:] if x], np.uint32)
    assert np.array_equal(got, vals.ravel().view(np.uint32))


def test_save_then_load_roundtrip(exe, golden_dat, tmp_path):
    rc, out, err = run(exe, "save_load", golden_dat[0], tmp_path / "copy.dat")
    assert rc == 0 and out.strip() == "same", err


def test_bad_files_throw_invalid_argument(exe, golden_dat, tmp_path):
    rc, out, _ = run(exe, "bad_load", tmp_path / "missing.dat")
    assert rc == 0 and out.startswith("invalid_argument")
    trunc = tmp_path / "trunc.dat"
    trunc.write_text(golden_dat[1]["dat_text"][: len(golden_dat[1]["dat_text"]) // 2])
    rc, out, _ = run(exe, "bad_load", trunc)
    assert rc == 0 and out.startswith("invalid_argument")


def test_todat_and_tofits_formats(exe, tmp_path):
    base = tmp_path / "img"
    rc, out, err = run(exe, "write_outputs", base)
    assert rc == 0, err
    vals = out.split()
    W, H = int(vals[0]), int(vals[1])
    img = f32([int(x) for x in vals[2:]]).reshape(H, W)
    # ToDat: std::to_string (6 decimals) then ',' per value, one row per line (Camera.cpp:235-250)
    rows = open(str(base) + ".dat").read().split("\n")
    assert rows[-1] == "" and len(rows) == H + 1
    for j in range(H):
        assert rows[j] == "".join("%f," % v for v in img[j])
    # ToFITS: primary HDU, BITPIX -32, NAXIS1 = W, NAXIS2 = H, header cards, row 0 first, big-endian
    assert os.path.getsize(str(base) + ".fits") % 2880 == 0
    cards, data = _read_fits(str(base) + ".fits")
    assert cards["SIMPLE"] == "T" and cards["BITPIX"] == "-32" and cards["NAXIS"] == "2"
    assert int(cards["NAXIS1"]) == W and int(cards["NAXIS2"]) == H
    assert float(cards["CRVAL1"]) == -1.5 and float(cards["CRVAL2"]) == -1.0
    assert float(cards["CDELT1"]) == 0.25 and float(cards["POS_Z"]) == 3.25 and float(cards["AIM_X"]) == 7.0
    assert cards["RHO"] == "'0       '"
    assert np.array_equal(data, img)


# ---------------------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_getsample_matches_reference_golden(exe, golden_dat, device):
    path, g = golden_dat
    pts = np.asarray(g["points"], np.uint32).reshape(-1, 3)
    stdin = "\n".join(" ".join(map(str, p)) for p in pts) + "\n"
    for tri, key in ((1, "trilinear"), (0, "nearest")):
        rc, out, err = run(exe, "sample", path, tri, stdin=stdin)
        assert rc == 0, err
        assert [int(x) for x in out.split()] == g[key]


def _read_image(out):
    vals = out.split()
    W, H = int(vals[0]), int(vals[1])
    return f32([int(x) for x in vals[2:]]).reshape(H, W)


def _write_dat(path, vals, xs, ys, zs):
    """The .dat cube format DataCube::Load reads (DataCube.cpp:28-138): a title line, 8 comment
    lines, "nx ny nz", the x, y and z coordinates, then the values x fastest, one row per line."""
    nz, ny, nx = vals.shape
    with open(path, "w") as f:
        f.write("cube\n" + "#\n" * 8 + f"{nx} {ny} {nz}\n")
        for c in (xs, ys, zs):
            f.write(" ".join("%.9g" % v for v in c) + "\n")
        for z in range(nz):
            for y in range(ny):
                f.write(" ".join("%.9g" % v for v in vals[z, y]) + "\n")


def _read_fits(path):
    """(header cards, H x W float32 image) of a primary-HDU FITS file, rows in file order."""
    raw = open(path, "rb").read()
    cards, k = {}, 0
    while True:
        c = raw[k:k + 80].decode()
        k += 80
        if c.startswith("END"):
            break
        key, _, rest = c.partition("=")
        cards[key.strip()] = rest.split("/")[0].strip()
    hdr = (k + 2879) // 2880 * 2880
    W, H = int(cards["NAXIS1"]), int(cards["NAXIS2"])
    return cards, np.frombuffer(raw[hdr:hdr + 4 * W * H], dtype=">f4").reshape(H, W).astype(np.float32)


def _shell_dat(path, n, seed):
    from birp_cpp_amd import capi
    vals, org, spc, _ = capi.synth_shell(n, (8.0, 11.0), 0.05, seed=seed)
    xs = (np.float32(org[0]) + np.float32(spc[0]) * np.arange(n, dtype=np.float32)).astype(np.float32)
    ys = (np.float32(org[1]) + np.float32(spc[1]) * np.arange(n, dtype=np.float32)).astype(np.float32)
    _write_dat(path, vals, xs, ys, ys)
    # the coordinates the loader reads back define the cube's metadata
    return O.OracleCube(vals, (float(xs[0]), float(ys[0]), float(ys[0])),
                        (float(abs(xs[0] - xs[1])), float(abs(ys[0] - ys[1])), float(abs(ys[0] - ys[1]))))


@pytest.mark.gpu
@pytest.mark.parametrize("semantics", [0, 1])
def test_camera_render_bitwise(exe, device, tmp_path, semantics):
    """Camera::Render through the C++ API on a .dat cube equals the oracle bit for bit."""
    p = tmp_path / "c.dat"
    oc = _shell_dat(p, 64, 11)
    pos, aim = (12.0, 4.0, 6.0), 1.0
    rc, out, err = run(exe, "render", p, 1, 0.25, 0, 36.0, 36.0, *pos, aim, 44, 200, semantics)
    assert rc == 0, err
    img = _read_image(out)
    v = O.make_view(144, 144, 36.0, 36.0, list(pos), [aim, 0, 0], 44, 200)
    ref, _ = O.render_f32(oc, v, semantics)
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    rc, out, err = run(exe, "render_twice", p, 1, 0.25, 0, 36.0, 36.0, *pos, aim, 44, 200, semantics)
    twice = _read_image(out)
    ref2, _ = O.render_f32(oc, v, semantics, init=ref)
    assert np.array_equal(twice.view(np.uint32), ref2.view(np.uint32))


@pytest.mark.gpu
def test_cmem_api_known_answers_and_render(exe, device):
    pts = np.array([[8, 4, 4], [4, 12, 10], [3, 1.5, 1.2], [8, 10, 10], [8, -10, -10], [8, 10, -10]], np.float32)
    rc, out, err = run(exe, "cmem_sample", "default", stdin="\n".join(" ".join(map(str, p.view(np.uint32))) for p in pts) + "\n")
    assert rc == 0, err
    got = f32([int(x) for x in out.split()])
    want = np.array([0, 4.46853346e-06, 0, 8.15947317e-07, 8.15947317e-07, 8.15947317e-07], np.float32)
    np.testing.assert_allclose(got, want, rtol=4e-6, atol=0)  # device libm vs glibc: a few ulp
    rc, out, err = run(exe, "cmem_render", "default")
    assert rc == 0, err
    img = _read_image(out)
    v = O.make_view(144, 144, 36.0, 36.0, [1.0, 1.0, 1.0], [1.0, 0, 0], 44, 200)
    oc = O.OracleCmem()
    ref, _ = O.render_f32(oc, v, O.SEM_REFERENCE, nthreads=16)
    ref64, _ = O.render_f64(oc, v, O.SEM_REFERENCE, nthreads=16)
    cmem_render_parity(img, ref, ref64, *oc.edge_bound(v, CMEM_EDGE_BAND, nthreads=16))


def _fp32_prefix_sums(s1, n):
    """P[m] = the FP32 sum, in order, of m copies of s1 (P[0] = 0)."""
    P = np.zeros(n + 1, np.float32)
    acc = np.float32(0.0)
    for m in range(1, n + 1):
        acc = np.float32(acc + s1)
        P[m] = acc
    return P


@pytest.mark.gpu
def test_cubefake_renders_its_constant_field(exe, device, tmp_path):
    """The reference's own Camera tests (tests/Camera_test.cpp:25-73): a CubeFake (GetSample == 1)
    at (1,1,1) aiming at x = 1, 0.25 deg pixels over 36 x 36 deg, rendered and written with ToDat
    and ToFITS. Every pixel is the FP32 sum of 200 copies of the scaled sample s1 = f32(f64(w) *
    637810000 / (4/pi) / 1e4) (closed form; w = 44 * 1.5/200 - 44 * 0.5/200 in FP32), and a second
    Render adds 200 more; the files hold the image."""
    rc, out, err = run(exe, "space_render", "fake", tmp_path, 0.25, 36.0, 36.0, 1.0, 1.0, 1.0, 1.0, 44.0, 200, 1)
    assert rc == 0, err
    img = _read_image(out)
    t, w = O.ray_dist(44.0, 200)
    s1 = np.float32(O.scale_sample(1.0, w))
    P = _fp32_prefix_sums(s1, 400)
    assert img.shape == (144, 144) and np.array_equal(img.view(np.uint32), np.full_like(img, P[200]).view(np.uint32))
    rows = open(str(tmp_path / "testout.dat")).read().split("\n")
    assert len(rows) == 145 and rows[0] == ("%f," % P[200]) * 144
    cards, data = _read_fits(str(tmp_path / "testout.fits"))
    assert np.array_equal(data, img) and float(cards["POS_X"]) == 1.0
    rc, out, err = run(exe, "space_render", "fake", "-", 0.25, 36.0, 36.0, 1.0, 1.0, 1.0, 1.0, 44.0, 200, 2)
    assert rc == 0, err
    twice = _read_image(out)
    acc = P[200]
    for _ in range(200):
        acc = np.float32(acc + s1)
    assert np.array_equal(twice.view(np.uint32), np.full_like(twice, acc).view(np.uint32))


@pytest.mark.gpu
def test_user_space_with_negative_values_stops_rays(exe, device):
    """A user Space that is 1 for x > 0.5 and -1 elsewhere, seen from (3, -10, 0) along +y (the FOV
    spans rays that cross x = 0.5 within the ray length and rays that do not): each ray adds s1 for
    its samples up to its first point with x <= 0.5 (FP32 points d t_k + P from the oracle's rays)
    and stops there (Camera.cpp:207-210). The host calls GetSample up to and including that sample
    and no further, as the reference's loop does."""
    pos, aim, L, n = (3.0, -10.0, 0.0), 3.0, 20.0, 300
    rc, out, err = run(exe, "space_render", "halfspace", "-", 0.5, 30.0, 40.0, *pos, aim, L, n, 1)
    assert rc == 0, err
    img = _read_image(out)
    H, W = img.shape
    v = O.make_view(W, H, 40.0, 30.0, list(pos), [aim, 0, 0], L, n)
    d = O.ray_dirs(v).reshape(H, W, 3)
    t, w = O.ray_dist(L, n)
    x = (d[:, :, 0:1] * t[None, None, :].astype(np.float32)).astype(np.float32) + np.float32(pos[0])
    neg = x <= np.float32(0.5)
    m = np.where(neg.any(axis=2), neg.argmax(axis=2), n)  # samples added before the stop
    P = _fp32_prefix_sums(np.float32(O.scale_sample(1.0, w)), n)
    assert 0 < (m < n).mean() < 1 and m.min() > 0
    assert np.array_equal(img.view(np.uint32), P[m].view(np.uint32))
    calls = int(err.split("calls")[1].split()[0])
    assert calls == int(np.minimum(m + 1, n).sum())


@pytest.mark.gpu
def test_render_all_equals_renders_one_by_one(exe, device, tmp_path):
    """Camera::RenderAll (one exact batch over three .dat cubes and a CMEM model, twice) equals
    Render() twice on each camera bit for bit, cube images equal the oracle's FP32 renders
    (accumulated twice), and the in-cube counts agree."""
    paths, ocs = [], []
    for k, (n, seed) in enumerate(((48, 31), (40, 32), (56, 33))):
        p = tmp_path / f"c{k}.dat"
        ocs.append(_shell_dat(p, n, seed))
        paths.append(str(p))
    for sem in (0, 1):
        rc, out, err = run(exe, "render_all", 30.0, 300, sem, *paths)
        assert rc == 0, err
        vals = out.split()
        imgs, at = [], 0
        for _ in range(2 * (len(paths) + 1)):
            W, H = int(vals[at]), int(vals[at + 1])
            imgs.append(f32([int(x) for x in vals[at + 2:at + 2 + W * H]]).reshape(H, W))
            at += 2 + W * H
        counts = [int(x) for x in vals[at:]]
        half = len(paths) + 1
        for k in range(half):
            assert np.array_equal(imgs[k].view(np.uint32), imgs[half + k].view(np.uint32)), (sem, k)
        for k, oc in enumerate(ocs):
            H, W = imgs[k].shape
            v = O.make_view(W, H, 30.0, 20.0, [0.0 + k, -20.0 + 3.0 * k, 10.0], [10.0, 0, 0], 30.0, 300)
            once, c1 = O.render_f32(oc, v, sem, nthreads=16)
            twice, _ = O.render_f32(oc, v, sem, nthreads=16, init=once)
            assert np.array_equal(imgs[k].view(np.uint32), twice.view(np.uint32)), (sem, k)
            assert counts[k] == counts[half + k] == c1, (sem, k)
        assert imgs[half - 1].any()  # the CMEM camera rendered


# ---------------------------------------------------------------------------------- the CLIs
# oracle/_ref/birp_main is the reference's own src/main.cpp, compiled unchanged against include/ and
# linked to libbirp_api.so (oracle/Makefile; built only in the development container, where the
# reference sources are; it never reaches the GPU). birp_cpp_amd/bin/birp is this repository's own
# CLI with the same flags (birp_cpp_amd/cli/birp_cli.cpp).
CLI = os.path.join(ROOT, "oracle", "_ref", "birp_main")
needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="oracle/_ref/birp_main not built (no reference sources)")
BIRP = os.path.join(ROOT, "birp_cpp_amd", "bin", "birp")


def run_birp(*args, cwd=None):
    r = subprocess.run([BIRP, *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=120)
    return r.returncode, r.stdout, r.stderr


@needs_cli
def test_reference_cli_links_this_library():
    """The reference's main.cpp builds unchanged against the C++ API and loads libbirp_api.so;
    without a GPU it fails loudly instead of falling back to a CPU render."""
    out = subprocess.run(["ldd", CLI], capture_output=True, text=True).stdout
    lib = [l for l in out.split("\n") if "libbirp_api.so" in l]
    assert lib and os.path.realpath(LIB) in os.path.realpath(lib[0].split("=>")[1].split("(")[0].strip())
    if not os.path.exists("/dev/kfd"):
        r = subprocess.run([CLI, "-c", "-o", "/nonexistent/"], capture_output=True, text=True, timeout=120)
        assert r.returncode != 0 and "no GPU" in r.stderr


def test_birp_cli_arguments_and_no_gpu():
    """The CLI refuses a missing or unusable input like main.cpp (exit 3), and without a GPU a render
    fails loudly (no CPU fallback)."""
    assert os.path.exists(BIRP), "birp_cpp_amd/bin/birp is missing: run `make -C birp_cpp_amd`"
    rc, so, err = run_birp()
    assert rc == 3 and "Please provide valid .dat" in so + err
    rc, so, err = run_birp("-i", "cube.txt")
    assert rc == 3
    if not os.path.exists("/dev/kfd"):
        rc, so, err = run_birp("-c", "-o", "/nonexistent/")
        assert rc != 0 and "no GPU" in err


@pytest.mark.gpu
def test_birp_cli_datacube_render(device, tmp_path):
    """main.cpp:266-305 single-cube mode: -i cube.dat -t, pose flags, default 0.25 deg pixels and
    36 x 36 deg FOV; the FITS output equals the oracle's FP32 render bit for bit."""
    oc = _shell_dat(tmp_path / "c.dat", 64, 12)
    out = str(tmp_path) + "/"
    rc, so, err = run_birp("-i", tmp_path / "c.dat", "-t", "-x", 12, "-y", 4, "-z", 6, "-a", 1, "-o", out)
    assert rc == 0, so + err
    cards, img = _read_fits(out + "c.dat.fits")
    v = O.make_view(144, 144, 36.0, 36.0, [12.0, 4.0, 6.0], [1.0, 0, 0], 44, 200)
    ref, _ = O.render_f32(oc, v, O.SEM_REFERENCE, nthreads=16)
    assert float(cards["POS_X"]) == 12.0 and float(cards["AIM_X"]) == 1.0
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("trilinear", [False, True])
def test_birp_cli_batch_directory(device, tmp_path, trilinear):
    """main.cpp:242-264 batch mode: every file of the input directory is rendered (nearest
    sampling without -t) and written as <name>.fits, here all in one batched launch; each equals
    the oracle bit for bit."""
    d = tmp_path / "cubes"
    d.mkdir()
    ocs = {name: _shell_dat(d / name, n, seed) for name, n, seed in (("a.dat", 48, 21), ("b.dat", 40, 22), ("c.dat", 64, 23))}
    for oc in ocs.values():
        oc.c.trilinear = int(trilinear)
    out = tmp_path / "out"
    out.mkdir()
    args = ["-i", str(d) + "/", "-x", 9, "-y", 2, "-z", 3, "-a", 0.5, "-s", 0.5, "-o", str(out) + "/"] + (["-t"] if trilinear else [])
    rc, so, err = run_birp(*args)
    assert rc == 0, so + err
    v = O.make_view(72, 72, 36.0, 36.0, [9.0, 2.0, 3.0], [0.5, 0, 0], 44, 200)
    for name, oc in ocs.items():
        _, img = _read_fits(str(out / name) + ".fits")
        ref, _ = O.render_f32(oc, v, O.SEM_REFERENCE, nthreads=16)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), name


@pytest.mark.gpu
@pytest.mark.parametrize("flags,params", [
    ([], dict(dipole=0.0, p0=1.0, p1=1, p2=3, p3=4, B=0.0, alpha=0.0, beta=0.0, bs=0.0, A1=0.0, A2=0.0)),
    (["-p", "0.9,2,3,4", "-q", "2", "-r", "2.5", "-f", "-1.6"],
     dict(dipole=0.0, p0=0.9, p1=2, p2=3, p3=4, B=2.0, alpha=2.5, beta=-1.6, bs=0.0, A1=0.0, A2=0.0))])
def test_birp_cli_cmem_render(device, tmp_path, flags, params):
    """main.cpp:276-305 with -c: CMEM through the CLI's parameter Init (with no flags p = {1,1,3,4}
    and the NULL floats 0), the default pose (5.91813, 6.74116, 17.7067) aiming at 7.81768, 0.25 deg
    pixels over 36 x 36 deg. Checked by the CMEM zone-edge rule."""
    out = str(tmp_path) + "/"
    rc, so, err = run_birp("-c", *flags, "-o", out)
    assert rc == 0, so + err
    _, img = _read_fits(out + "cmembirp.fits")
    oc = O.OracleCmem(params)
    v = O.make_view(144, 144, 36.0, 36.0, [5.91813, 6.74116, 17.7067], [7.81768, 0, 0], 44, 200)
    ref, _ = O.render_f32(oc, v, O.SEM_REFERENCE, nthreads=16)
    ref64, _ = O.render_f64(oc, v, O.SEM_REFERENCE, nthreads=16)
    cmem_render_parity(img, ref, ref64, *oc.edge_bound(v, CMEM_EDGE_BAND, nthreads=16))
