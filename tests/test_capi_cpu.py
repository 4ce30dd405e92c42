"""CPU tests of the C-ABI library (no GPU compute): it loads, it exports every entry point
include/birp.h declares, its host-side pose math equals the oracle's bit for bit, its synthetic
workloads are deterministic and have the stated geometry, and it reports the absence of a GPU
as an error instead of crashing."""
import os
import re

import numpy as np
import pytest

from birp_cpp_amd import capi
from birp_cpp_amd import workloads as W
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "birp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(birp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes signature table covers exactly the header
    assert sorted(n for n, _, _ in capi.SIGNATURES) == names


def test_version_and_strerror():
    assert capi.lib().birp_version() >= 10000
    assert capi.lib().birp_strerror(capi.ERR_SINGULAR) == b"Determinant is 0"
    assert capi.lib().birp_strerror(0) == b"ok"


def test_null_arguments_are_rejected_before_any_device_work():
    L, C = capi.lib(), capi.C
    blank = C.create_string_buffer(4096)  # stands for a handle; never read, the checks come first
    h = C.cast(blank, C.c_void_p)
    v = capi.make_view(16, 16, 26.5, 15.5, (15.0, 5.0, 8.0), (0.0, 0.0, 0.0), 44.0, 200)
    img = np.zeros(256, np.float32)
    n = C.c_longlong()
    inv = capi.ERR_INVALID_ARGUMENT
    assert L.birp_render(None, C.byref(v), capi.fptr(img), None) == inv
    assert L.birp_render(h, C.byref(v), None, None) == inv
    assert L.birp_count_samples(h, None, C.byref(n)) == inv
    assert L.birp_count_samples(None, C.byref(v), C.byref(n)) == inv
    assert L.birp_render_f64(None, C.byref(v), None) == inv
    assert L.birp_volume_sample(None, None, 1, None) == inv
    assert L.birp_volume_sample(h, capi.fptr(img), -1, capi.fptr(img)) == inv
    assert L.birp_volume_sample(h, capi.fptr(img), (1 << 32) + 1, capi.fptr(img)) == inv
    assert L.birp_render_async(None, C.byref(v), None) == inv
    assert L.birp_batch_create(None, C.byref(v), 1, None) == inv
    assert L.birp_server_submit(None, None, None, None) == inv
    assert L.birp_server_wait(None, 1, 0.0, None) == inv
    assert L.birp_device_synchronize(None) == inv
    assert L.birp_batch_set_output(None, None) == inv
    p = C.c_void_p()
    assert L.birp_host_alloc(0, C.byref(p)) == inv and not p.value
    assert L.birp_host_free(None) == capi.OK
    assert L.birp_host_free(h) == capi.ERR_NOT_FOUND  # not from birp_host_alloc: refused, not freed


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_no_gpu_is_an_error_not_a_crash():
    h = capi.C.c_void_p()
    assert capi.lib().birp_device_open(0, capi.C.byref(h)) == capi.ERR_NO_DEVICE
    with pytest.raises(capi.BirpError) as e:
        capi.Device(0)
    assert e.value.status == capi.ERR_NO_DEVICE


def test_view_basis_matches_oracle_bitwise(oracle):
    """The library's FP32 Orient/basis (Camera.cpp:85-137, 309-508) equals the oracle's."""
    rng = np.random.default_rng(1)
    for k in range(2000):
        pos = rng.uniform(-30, 30, 3).astype(np.float32)
        aim = (float(rng.uniform(-15, 25)), 0.0, 0.0)
        if k % 50 == 0:
            aim = (float(rng.uniform(-15, 25)), float(rng.uniform(-3, 3)), float(rng.uniform(-3, 3)))  # Orient uses all of aim
        v = capi.make_view(144, 84, 26.5, 15.5, pos, aim, 80, 160)
        st, a, rot, rinv = capi.view_basis(v)
        ov = O.make_view(144, 84, 26.5, 15.5, list(map(float, pos)), list(aim), 80, 160)
        rc, orot, orinv = O.view_basis(ov)
        assert st == rc == 0
        assert np.float32(a).tobytes() == np.float32(O.orient(pos, np.float32(aim))).tobytes()
        assert rot.tobytes() == orot.tobytes()
        assert rinv.tobytes() == orinv.tobytes()


@pytest.mark.parametrize("kw,oracle_params", [
    ({}, None),
    (dict(v=[400, 0, 0], b=[0, 0, 4], dipole=0.0, p0=1.0, p1=4, p2=6, p3=3, B=3.6, alpha=-1.0),
     dict(v=[400, 0, 0], b=[0, 0, 4], dipole=0.0, p0=1, p1=4, p2=6, p3=3, B=3.6, alpha=-1.0)),
    (dict(dipole=0.0, p0=1.0, p1=1, p2=3, p3=4, B=0.0, alpha=0.0, beta=0.0, bs=0.0, A1=0.0, A2=0.0),
     dict(dipole=0.0, p0=1.0, p1=1, p2=3, p3=4, B=0.0, alpha=0.0, beta=0.0, bs=0.0, A1=0.0, A2=0.0)),
    (dict(v=[350, -20, 10], b=[2, -3, -4], dipole=0.3, density=7.5, ay_bs=0.8, az_bs=0.7),
     dict(v=[350, -20, 10], b=[2, -3, -4], dipole=0.3, density=7.5, ay_bs=0.8, az_bs=0.7)),
])
def test_cmem_init_matches_oracle_bitwise(oracle, kw, oracle_params):
    """birp_cmem_init (CMEM::Init, Cmem.cpp:35-115) gives the oracle's post-Init state bit for bit:
    the defaults, the reference test's parameter binding (tests/CMEM_test.cpp:51-63), the CLI's
    NULL-float call (main.cpp:285-294) and a general set."""
    got = capi.cmem_init(capi.cmem_params(**kw))
    want = capi.cmem_state_from(oracle.OracleCmem(oracle_params).state())
    names = [n for n, _ in capi.CmemState._fields_]
    for n in names:
        a, b = getattr(got, n), getattr(want, n)
        if n == "beta_c":
            assert np.array(a[:], np.float32).tobytes() == np.array(b[:], np.float32).tobytes(), n
        else:
            assert np.float32(a).tobytes() == np.float32(b).tobytes(), (n, a, b)


def test_view_basis_rejects_nonfinite_pose():
    v = capi.make_view(16, 16, 26.5, 15.5, (float("nan"), 0, 0), (10, 0, 0), 80, 160)
    st, *_ = capi.view_basis(v)
    assert st == capi.ERR_INVALID_ARGUMENT


def test_view_basis_on_the_sun_earth_line(oracle):
    """On the x axis the basis is finite looking at the Sun and NaN looking away from it (the Sun
    projects onto the view axis, Camera.cpp:440-444), in the library as in the oracle."""
    for pos, aim, finite in (((-10.0, 0, 0), (20.0, 0, 0), True), ((20.0, 0, 0), (-10.0, 0, 0), False)):
        v = capi.make_view(16, 16, 26.5, 15.5, pos, aim, 80, 160)
        st, a, rot, rinv = capi.view_basis(v)
        want = O.orient(pos, aim)
        assert st == 0
        assert np.isfinite(a) == np.isfinite(want) == finite
        if finite:
            assert np.float32(a) == np.float32(want) == 270.0
            rc, orot, orinv = O.view_basis(O.make_view(16, 16, 26.5, 15.5, list(pos), list(aim), 80, 160))
            assert rot.tobytes() == orot.tobytes() and rinv.tobytes() == orinv.tobytes()


def test_shell_generator_deterministic_and_shaped():
    a, org, spc, r0 = capi.synth_shell(48, (8.0, 11.0), 0.05, seed=1)
    b, *_ = capi.synth_shell(48, (8.0, 11.0), 0.05, seed=1)
    c, *_ = capi.synth_shell(48, (8.0, 11.0), 0.05, seed=2)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert 8.0 <= r0 <= 11.0
    assert org == (-10.0, -15.0, -15.0)
    assert spc[0] == pytest.approx(30 / 47, rel=1e-6) and spc[1] == spc[2] == pytest.approx(30 / 47, rel=1e-6)
    assert np.all(a > 0) and np.all(np.isfinite(a))
    # shell vs background: the magnetopause at the subsolar point is r0; its value there is
    # 1e-5 (r0/10)^-3, the background 1e-7 (times lognormal noise with sigma 0.05)
    g = -10 + spc[0] * np.arange(48)
    gy = -15 + spc[1] * np.arange(48)
    iy = int(np.argmin(np.abs(gy)))
    sub = a[iy, iy, :]  # the x axis (y = z ~ 0)
    xin = [i for i in range(48) if r0 * 1.02 < g[i] < r0 * 1.28]
    xout = [i for i in range(48) if g[i] > r0 * 1.35 or (0 < g[i] < r0 * 0.95)]
    assert np.median(sub[xin]) > 30 * np.median(sub[xout])
    assert np.median(sub[xout]) == pytest.approx(1e-7, rel=0.1)


def test_shell_noise_is_lognormal():
    a, *_ = capi.synth_shell(40, (100.0, 101.0), 0.1, seed=7)  # magnetopause beyond the cube: background only
    z = np.log(a.astype(np.float64) / 1e-7)
    assert abs(z.mean()) < 0.01 and z.std() == pytest.approx(0.1, rel=0.05)


def test_orbit_geometry():
    p = capi.synth_orbit(2000, seed=3).astype(np.float64)
    r = np.linalg.norm(p, axis=1)
    assert r.min() >= 1.8 - 1e-4 and r.max() <= 19.0 + 1e-4
    assert r.min() < 2.5 and r.max() > 18.5
    # the orbit plane contains the x axis and is tilted 70 degrees from the x-y plane
    normal = np.array([0.0, -np.sin(np.radians(70)), np.cos(np.radians(70))])
    assert np.abs(p @ normal).max() < 1e-4
    assert np.array_equal(p, capi.synth_orbit(2000, seed=3).astype(np.float64))


def test_workload_definitions():
    v1, v2 = W.view_144x84(), W.view_576x336()
    assert (v1.width, v1.height, v1.samples) == (144, 84, 160)
    assert (v2.width, v2.height, v2.samples) == (576, 336, 800)
    assert O.image_dims(0.1845, 15.5, 26.5) == (144, 84)  # the reference's FOV rule (Camera.cpp:38-39)
    vs = W.orbit_views_mixed(256)
    assert len(vs) == 256 and {(v.width, v.height) for v in vs} == set(W.SIZES)
    assert {v.samples for v in vs} == {800, 320}
    assert all(v.math == capi.MATH_FAST for v in vs)
    assert W.orbit_views_576x336_n1600(4)[0].samples == 1600
    cubes = W.serving_cubes()
    assert len(cubes) == 8 and all(n in (128, 256, 384, 512) for n, _ in cubes)
    t, which, views = W.serving_requests(1000)
    rate = (len(t) - 1) / (t[-1] - t[0])
    assert rate == pytest.approx(200, rel=0.15) and which.max() < 8


STRUCTS = {"VolumeDesc": "birp_volume_desc", "CmemState": "birp_cmem_state", "CmemParams": "birp_cmem_params",
           "View": "birp_view", "RenderStats": "birp_render_stats", "Tuning": "birp_tuning",
           "ServerConfig": "birp_server_config", "Request": "birp_request", "ResultInfo": "birp_result_info",
           "ShellParams": "birp_shell_params"}


def test_ctypes_structs_match_header_layout(tmp_path):
    """Every ctypes mirror in capi.py has the size, field names and field offsets that gcc gives
    the C struct of include/birp.h (an appended or reordered field breaks the ABI)."""
    import subprocess
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "birp.h"', "int main(void) {"]
    for py, c in STRUCTS.items():
        lines.append(f'    printf("{py} size %zu\\n", sizeof({c}));')
        for name, _ in getattr(capi, py)._fields_:
            lines.append(f'    printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n"):
        if line:
            py, name, val = line.split()
            got[(py, name)] = int(val)
    for py in STRUCTS:
        cls = getattr(capi, py)
        assert got[(py, "size")] == capi.C.sizeof(cls), py
        for name, _ in cls._fields_:
            assert got[(py, name)] == getattr(cls, name).offset, (py, name)


def test_tuning_table_load_validates_entries(tmp_path):
    """birp_autotune_load keeps only entries that birp_batch_set_tuning would accept (tile_w a
    power of two, tile_w * tile_h * lanes_per_ray = 64, lanes and unroll in {1, 2, 4}, a known
    dtype), and fails when none is left. (Host-only: no GPU needed.)"""
    row = ('{{"dtype": {}, "size_class": 2, "samples_class": 2, "tile_w": {}, "tile_h": {}, "lanes_per_ray": {}, '
           '"unroll": {}, "ms": 1.0}}')
    bad = [row.format(1, 8, 1, 8, 4), row.format(1, 6, 4, 1, 4), row.format(1, 8, 8, 1, 3), row.format(7, 8, 8, 1, 4),
           row.format(1, 16, 8, 1, 4), "not json"]
    p = tmp_path / "bad.jsonl"
    p.write_text("\n".join(bad) + "\n")
    assert capi.lib().birp_autotune_load(str(p).encode()) == capi.ERR_NOT_FOUND
    # the committed table's own entry for BF16 / 576x336 / 1600 samples, so later tests see no change
    good = tmp_path / "good.jsonl"
    good.write_text("\n".join(bad + [row.format(1, 8, 8, 1, 4)]) + "\n")
    assert capi.lib().birp_autotune_load(str(good).encode()) == capi.OK
    t = capi.Tuning()
    assert capi.lib().birp_autotune_lookup(1, 576, 336, 1600, capi.C.byref(t)) == capi.OK
    assert (t.tile_w, t.tile_h, t.lanes_per_ray, t.unroll) == (8, 8, 1, 4)
