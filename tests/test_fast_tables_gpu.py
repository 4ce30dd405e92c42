"""The fast kernel's per-wave offset tables live in dynamic LDS sized per launch to the largest
table the launch reads (nx + ny + nz words, rounded up to 256). A long 1800 x 96 x 96 cube has
1,992 entries, near the 2,048 limit: its fast single renders, a batch of its views and a server
that mixes it with a small cube in one launch match the FP64 render of the same sample grid, and
the small cube's served images equal its own single renders, so neither table overruns the other.
"""
import numpy as np
import pytest

from birp_cpp_amd import capi
from birp_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def cubes(device):
    rng = np.random.default_rng(31)
    nz, ny, nx = 96, 96, 1800  # arrays are [z][y][x]
    z, y, x = np.meshgrid(np.linspace(-1, 1, nz), np.linspace(-1, 1, ny), np.linspace(-1, 1, nx), indexing="ij")
    long = (1e-5 * np.exp(-3.0 * (y * y + z * z)) * (1.2 + np.sin(9.0 * x)) + 1e-7).astype(np.float32)
    long *= (1.0 + 0.05 * rng.standard_normal(long.shape)).astype(np.float32)
    long = np.abs(long)
    org_l, spc_l = (-12.0, -3.0, -3.0), (30.0 / (nx - 1), 6.0 / (ny - 1), 6.0 / (nz - 1))
    vals, org, spc, _ = W.cube_128()
    vols = {
        "long32": capi.Volume.upload(device, long, org_l, spc_l),
        "long16": capi.Volume.upload(device, long, org_l, spc_l, dtype=capi.DTYPE_F16),
        "small": capi.Volume.upload(device, vals, org, spc),
    }
    yield vols
    for v in vols.values():
        v.free()


def long_views(math=capi.MATH_FAST):
    poses = [((3.0, -14.0, 4.0), (3.0, 0.0, 0.0)), ((-20.0, -8.0, 2.0), (5.0, 0.0, 0.0)),
             ((10.0, 9.0, -6.0), (0.0, 0.0, 0.0)), ((3.0, -1.0, 12.0), (2.0, 0.0, 0.0))]
    return [capi.make_view(320, 180, 40.0, 25.0, p, a, 60.0, 600, capi.SEM_EXTENDED, math) for p, a in poses]


def test_long_cube_fast_single_and_batch(cubes):
    assert cubes["long32"].nbytes() > 0
    views = long_views()
    refs = [cubes["long32"].render_f64(v) for v in views]
    for k, ref in enumerate(refs):
        assert (ref > 0).mean() > 0.05, f"view {k} misses the cube"
    for key, tol in (("long32", 1e-5), ("long16", 1e-4)):
        for k, v in enumerate(views):
            img, _ = cubes[key].render(v)
            assert np.isfinite(img).all()
            assert rel_l2(img, refs[k]) <= tol, (key, k, rel_l2(img, refs[k]))
        b = capi.Batch(cubes[key], views)
        b.launch()
        b.wait()
        for k in range(len(views)):
            assert rel_l2(b.image(k), refs[k]) <= tol, (key, "batch", k)
        b.free()


def test_server_mixes_table_sizes(device, cubes):
    views = long_views()
    small_views = [W.view(288, 168, 0.25, pos=(float(p[0]), float(p[1]), float(p[2])), math=capi.MATH_FAST)
                   for p in capi.synth_orbit(6, 13)]
    srv = capi.Server(device, memory_budget=4 << 30, max_batch=16, batch_window_ms=2.0)
    try:
        il = srv.add_volume(cubes["long32"])
        ism = srv.add_volume(cubes["small"])
        ids = [(srv.submit(il, v), ("long", k)) for k, v in enumerate(views)]
        ids += [(srv.submit(ism, v), ("small", k)) for k, v in enumerate(small_views)]
        got = {}
        for rid, tag in ids:
            st, img, _ = srv.wait(rid, 20000.0)
            assert st == capi.OK, tag
            got[tag] = img
    finally:
        srv.destroy()
    for k, v in enumerate(views):
        assert rel_l2(got[("long", k)], cubes["long32"].render_f64(v)) <= 1e-5, k
    for k, v in enumerate(small_views):
        single, _ = cubes["small"].render(v)
        assert np.array_equal(got[("small", k)].view(np.uint32), single.view(np.uint32)), k
