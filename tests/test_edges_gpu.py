"""GPU edge cases of the render path: ragged and degenerate cube shapes (each brick order and the
linear fallback), the largest exact ray, tiny and empty images, rays that miss the cube, and the
argument checks of the C ABI. Exact renders are compared with the oracle bit for bit; fast
renders with the FP64 render at the FP32-storage tolerance (1e-5)."""
import os

import numpy as np
import pytest

from birp_cpp_amd import capi
from birp_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def oview(O, v):
    return O.make_view(v.width, v.height, v.fov_w, v.fov_h, list(v.position), list(v.aim), v.ray_length, v.samples)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def positive_cube(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.lognormal(-13.0, 0.4, shape)).astype(np.float32)


# (nz, ny, nx), origin, spacing, camera inside the box
RAGGED = [
    ((60, 50, 57), (-10.0, -15.0, -12.0), (0.3, 0.6, 0.5), (4.0, -1.0, 2.0)),    # odd sizes, Morton bricks
    ((12, 130, 9), (-4.0, -40.0, -5.0), (1.0, 0.6, 0.9), (0.0, 5.0, 1.0)),       # elongated: x-fastest brick order
    ((2, 2, 2), (-1.0, -1.0, -1.0), (2.0, 2.0, 2.0), (0.2, 0.1, -0.3)),          # a single cell
    ((2, 2, 8200), (-40.0, -15.0, -15.0), (0.01, 30.0, 30.0), (1.0, 2.0, 3.0)),  # nx+ny+nz > 8192: linear layout
]


@pytest.mark.parametrize("shape,org,spc,pos", RAGGED)
@pytest.mark.parametrize("semantics", [0, 1])
def test_ragged_cubes_exact_bitwise(device, oracle, shape, org, spc, pos, semantics):
    vals = positive_cube(shape, seed=sum(shape))
    vol = capi.Volume.upload(device, vals, org, spc)
    try:
        oc = oracle.OracleCube(vals, org, spc)
        for aim in ((pos[0] + 3.0, 0.0, 0.0), (pos[0] - 5.0, 0.0, 0.0)):
            v = capi.make_view(48, 40, 30.0, 20.0, pos, aim, 20.0, 300, semantics, capi.MATH_EXACT)
            img, st = vol.render(v)
            ref, cnt = oracle.render_f32(oc, oview(oracle, v), semantics, nthreads=NT)
            assert st.samples_in_cube == cnt
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (shape, aim, semantics)
            assert cnt > 0
    finally:
        vol.free()


@pytest.mark.parametrize("shape,org,spc,pos", RAGGED[:3])
def test_ragged_cubes_fast_paths(device, shape, org, spc, pos):
    vals = positive_cube(shape, seed=sum(shape) + 1)
    v32 = capi.Volume.upload(device, vals, org, spc)
    v16 = capi.Volume.upload(device, vals, org, spc, dtype=capi.DTYPE_F16)
    try:
        v = capi.make_view(48, 40, 30.0, 20.0, pos, (pos[0] + 3.0, 0.0, 0.0), 20.0, 300, capi.SEM_EXTENDED, capi.MATH_FAST)
        ref = v32.render_f64(v)
        f32, _ = v32.render(v)
        f16, _ = v16.render(v)
        assert rel_l2(f32, ref) <= 1e-5
        assert rel_l2(f16, ref) <= 1e-4
    finally:
        v32.free()
        v16.free()


def test_fast_march_on_faces(device):
    """Cameras on a face, an edge and a corner of the cube, and rays along a face plane: the fast
    path's per-ray march fit keeps every sample inside, and the images meet the FP32-storage
    tolerance (1e-5) and the F16 one (1e-4) against the FP64 render."""
    n = 64
    vals = positive_cube((n, n, n), seed=77)
    org, spc = (-10.0, -15.0, -15.0), (30.0 / (n - 1),) * 3
    hi = tuple(o + (n - 1) * s for o, s in zip(org, spc))
    v32 = capi.Volume.upload(device, vals, org, spc)
    v16 = capi.Volume.upload(device, vals, org, spc, dtype=capi.DTYPE_F16)
    poses = [((org[0], 0.0, 0.0), (hi[0], 0.0, 0.0)),            # on the x axis, looking at the Sun
             ((org[0], 2.0, 1.0), (hi[0], 0.0, 0.0)),            # on the x = min face, looking along +x
             ((hi[0], 1.0, -2.0), (org[0], 0.0, 0.0)),           # on the far face, looking back
             ((org[0], org[1], 0.0), (hi[0], 0.0, 0.0)),         # on an edge
             ((org[0], org[1], org[2]), (hi[0], 0.0, 0.0)),      # on a corner
             ((0.0, org[1], 0.0), (12.0, 0.0, 0.0))]             # on the y = min face
    try:
        for pos, aim in poses:
            for L, n_s in ((50.0, 500), (80.0, 1600)):
                v = capi.make_view(64, 48, 40.0, 30.0, pos, aim, L, n_s, capi.SEM_EXTENDED, capi.MATH_FAST)
                ref = v32.render_f64(v)
                f32, _ = v32.render(v)
                f16, _ = v16.render(v)
                assert np.isfinite(f32).all() and np.isfinite(f16).all()
                assert rel_l2(f32, ref) <= 1e-5, (pos, L)
                assert rel_l2(f16, ref) <= 1e-4, (pos, L)
    finally:
        v32.free()
        v16.free()


@pytest.mark.parametrize("interp", ["trilinear", "nearest"])
def test_random_views_of_signed_cube_bitwise(device, oracle, interp):
    """A seeded sweep of 24 random views (cameras inside and outside, image sizes 1..32 px a side,
    2..900 samples, both semantics) of a signed 40^3 cube with 15% zeros, trilinear (the general
    brick kernel: std::lerp's straddle branch, early stops on negative samples) and nearest: every
    image is bitwise the oracle's and every in-cube sample count equals its count."""
    rng = np.random.default_rng(11)
    n = 40
    vals = (rng.standard_normal((n, n, n)) * 1e-5 + 2e-6).astype(np.float32)
    vals[rng.random(vals.shape) < 0.15] = 0.0
    org, spc = (-10.0, -15.0, -15.0), (30.0 / (n - 1),) * 3
    nearest = interp == "nearest"
    vol = capi.Volume.upload(device, vals, org, spc, interp=capi.INTERP_NEAREST if nearest else capi.INTERP_TRILINEAR)
    try:
        oc = oracle.OracleCube(vals, org, spc, trilinear=not nearest)
        done = total = 0
        while done < 24:
            inside = rng.random() < 0.7  # else anywhere within 15 units of the cube
            pos = rng.uniform((-10.0, -15.0, -15.0), (20.0, 15.0, 15.0)) if inside else rng.uniform((-25.0, -30.0, -30.0), (35.0, 30.0, 30.0))
            aim = rng.uniform((-10.0, -15.0, -15.0), (20.0, 15.0, 15.0))
            if np.linalg.norm(pos - aim) < 0.5 or not np.isfinite(oracle.orient(tuple(pos), tuple(aim))):
                continue
            w, h = (int(x) for x in rng.integers(1, 33, 2))
            sem = done % 2
            v = capi.make_view(w, h, float(rng.uniform(5.0, 70.0)), float(rng.uniform(5.0, 70.0)), tuple(pos), tuple(aim),
                               float(rng.uniform(5.0, 80.0)), int(rng.integers(2, 901)), sem, capi.MATH_EXACT)
            img, st = vol.render(v)
            ref, cnt = oracle.render_f32(oc, oview(oracle, v), sem, nthreads=NT)
            assert st.samples_in_cube == cnt, (done, pos, aim, w, h, sem)
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (done, pos, aim, w, h, sem)
            done += 1
            total += cnt
        assert total > 20_000  # the sweep reaches the cube (negative samples end rays early)
    finally:
        vol.free()


def test_sunward_axis_pose_refused_on_both_paths(cfg_small, oracle):
    """A camera on the Sun-Earth line looking away from the Sun: the Sun projects onto the view
    axis, the reference's PointToPlane divides 0 by 0 (Camera.cpp:440-444) and its angle and basis are
    NaN. Exact and fast renders refuse the pose alike, with BIRP_ERR_INVALID_ARGUMENT."""
    vol, _ = cfg_small
    assert np.isnan(oracle.orient((20.0, 0.0, 0.0), (-10.0, 0.0, 0.0)))
    buf = np.zeros(16 * 16, np.float32)
    for math in (capi.MATH_EXACT, capi.MATH_FAST):
        v = capi.make_view(16, 16, 26.5, 15.5, (20.0, 0.0, 0.0), (-10.0, 0.0, 0.0), 44.0, 200, capi.SEM_EXTENDED, math)
        assert capi.lib().birp_render(vol.h, capi.C.byref(v), capi.fptr(buf), None) == capi.ERR_INVALID_ARGUMENT
    assert not buf.any()


def test_longest_exact_ray_and_beyond(device, oracle):
    """n = 8192 is the exact path's largest ray (its t table lives in LDS): bitwise there, and a
    clear BIRP_ERR_UNSUPPORTED one sample beyond; the fast path takes longer rays."""
    vals, org, spc, _ = W.cube_128()
    vol = capi.Volume.upload(device, vals, org, spc)
    try:
        oc = oracle.OracleCube(vals, org, spc)
        v = capi.make_view(24, 16, 26.5, 15.5, (15.0, 5.0, 8.0), (0.0, 0.0, 0.0), 80.0, 8192)
        img, st = vol.render(v)
        ref, cnt = oracle.render_f32(oc, oview(oracle, v), oracle.SEM_EXTENDED, nthreads=NT)
        assert st.samples_in_cube == cnt
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
        v.samples = 8193
        with pytest.raises(capi.BirpError) as e:
            vol.render(v)
        assert e.value.status == capi.ERR_UNSUPPORTED
        v.math = capi.MATH_FAST
        fast, _ = vol.render(v)
        assert rel_l2(fast, vol.render_f64(v)) <= 1e-5
    finally:
        vol.free()


def test_tiny_images_and_rays_that_miss(cfg_small, oracle):
    vol, oc = cfg_small
    # 1 x 1 and 1 x N images
    for wh in ((1, 1), (1, 7), (9, 1)):
        v = capi.make_view(wh[0], wh[1], 10.0, 10.0, (15.0, 5.0, 8.0), (0.0, 0.0, 0.0), 44.0, 200)
        img, _ = vol.render(v)
        ref, _ = oracle.render_f32(oc, oview(oracle, v), oracle.SEM_EXTENDED)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    # a camera outside the cube looking away: every ray misses, the image is all zeros
    v = capi.make_view(32, 32, 20.0, 20.0, (60.0, 60.0, 60.0), (200.0, 0.0, 0.0), 44.0, 200)
    img, st = vol.render(v)
    assert st.samples_in_cube == 0 and not img.any()
    v.math = capi.MATH_FAST
    fast, _ = vol.render(v)
    assert not fast.any()


def test_argument_checks(cfg_small):
    vol, _ = cfg_small
    bad = [dict(width=0), dict(height=-3), dict(samples=1), dict(ray_length=0.0), dict(ray_length=float("inf")),
           dict(fov_w=float("nan")), dict(semantics=7), dict(math=5)]
    buf = np.zeros(16 * 16, np.float32)  # never written: the view is rejected first
    for b in bad:
        v = capi.make_view(16, 16, 26.5, 15.5, (15.0, 5.0, 8.0), (0.0, 0.0, 0.0), 44.0, 200)
        for k, val in b.items():
            setattr(v, k, val)
        st = capi.lib().birp_render(vol.h, capi.C.byref(v), capi.fptr(buf), None)
        assert st == capi.ERR_INVALID_ARGUMENT, b
    assert not buf.any()
    # uploads: too small, bad spacing, unknown dtype, 16-bit nearest
    tiny = np.ones((1, 4, 4), np.float32)
    with pytest.raises(capi.BirpError) as e:
        capi.Volume.upload(vol.dev, tiny, (0, 0, 0), (1, 1, 1))
    assert e.value.status == capi.ERR_INVALID_ARGUMENT
    cube = np.ones((4, 4, 4), np.float32)
    for spc in ((0.0, 1.0, 1.0), (1.0, float("nan"), 1.0), (1.0, 1.0, 2e30)):
        with pytest.raises(capi.BirpError) as e:
            capi.Volume.upload(vol.dev, cube, (0, 0, 0), spc)
        assert e.value.status == capi.ERR_INVALID_ARGUMENT
    with pytest.raises(capi.BirpError) as e:
        capi.Volume.upload(vol.dev, cube, (0, 0, 0), (1, 1, 1), dtype=9)
    assert e.value.status == capi.ERR_INVALID_ARGUMENT
    with pytest.raises(capi.BirpError) as e:
        capi.Volume.upload(vol.dev, cube, (0, 0, 0), (1, 1, 1), interp=capi.INTERP_NEAREST, dtype=capi.DTYPE_F16)
    assert e.value.status == capi.ERR_UNSUPPORTED
    nonfinite = cube.copy()
    nonfinite[1, 1, 1] = np.inf
    with pytest.raises(capi.BirpError) as e:
        capi.Volume.upload(vol.dev, nonfinite, (0, 0, 0), (1, 1, 1), dtype=capi.DTYPE_F16)
    assert e.value.status == capi.ERR_UNSUPPORTED


@pytest.fixture(scope="module")
def cfg_small(device, oracle):
    vals = positive_cube((40, 40, 40), seed=5)
    org, spc = (-10.0, -15.0, -15.0), (30.0 / 39,) * 3
    vol = capi.Volume.upload(device, vals, org, spc)
    yield vol, oracle.OracleCube(vals, org, spc)
    vol.free()
