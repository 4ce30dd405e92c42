"""Fast, batched and served renders of cubes with negative values: every ray ends at its first
negative sample, as the reference's Camera::Render does (src/Camera.cpp:207-210).

  * The signed 40^3 noise cube of the exact sweep (15% zeros, N(2e-6, 1e-5) values): fast single
    renders, one batch and the server, under both semantics, each image within 1e-5 (relative L2)
    of the oracle's FP64 render on the CPU; served results bitwise equal to the single renders.
  * A cube whose values are negative inside a sphere (|v| bounded, so F16 keeps one volume scale):
    FP32, F16 and BF16 storage within 1e-5 of the FP64 render of the cube as each storage holds it
    (the values rounded to F16 or BF16 on the host, exactly as the upload rounds them).
The stop is not vacuous: the same views rendered with the negative values clipped to 0 differ
from the references by far more than the tolerance.
"""
import math
import os

import numpy as np
import pytest

from birp_cpp_amd import capi

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)
ORG = (-10.0, -15.0, -15.0)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oview(O, v):
    return O.make_view(v.width, v.height, v.fov_w, v.fov_h, list(v.position), list(v.aim), v.ray_length, v.samples)


def noise_cube(n=40, seed=11):
    rng = np.random.default_rng(seed)
    vals = (rng.standard_normal((n, n, n)) * 1e-5 + 2e-6).astype(np.float32)
    vals[rng.random(vals.shape) < 0.15] = 0.0
    return vals


def sphere_cube(n=48, seed=12):
    """|v| = 1e-5 exp(0.5 N(0,1)) (so the range of |v| fits F16 under one scale), negative inside
    the ball |x - (5, 0, 0)| < 6, 10% zeros."""
    rng = np.random.default_rng(seed)
    ax = ORG[0] + np.arange(n) * (30.0 / (n - 1))
    ay = ORG[1] + np.arange(n) * (30.0 / (n - 1))
    z, y, x = np.meshgrid(ay, ay, ax, indexing="ij")
    mag = 1e-5 * np.exp(0.5 * rng.standard_normal((n, n, n)))
    sign = np.where((x - 5.0) ** 2 + y ** 2 + z ** 2 < 36.0, -1.0, 1.0)
    vals = (sign * mag).astype(np.float32)
    vals[rng.random(vals.shape) < 0.10] = 0.0
    return vals


def as_f16_storage(vals):
    """The values an F16 volume holds when every nonzero |v| fits F16's normal range under one
    power-of-two scale s (vmax / s < 2^15): f16(v / s) * s."""
    a = np.abs(vals[vals != 0])
    m, e = math.frexp(float(a.max()))
    s = 2.0 ** (e - 15)
    assert float(a.min()) / s >= 2.0 ** -14  # the one-scale form, as the upload chooses it
    q = (vals / np.float32(s)).astype(np.float16).astype(np.float32) * np.float32(s)
    return q.astype(np.float32)


def as_bf16_storage(vals):
    """Round to nearest even at 16 bits, as the upload's float -> bfloat16 conversion does."""
    b = vals.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def random_views(count, seed, sizes=(16, 41), samples=(200, 901)):
    rng = np.random.default_rng(seed)
    from oracle import oracle as O
    views = []
    while len(views) < count:
        inside = rng.random() < 0.7
        pos = rng.uniform((-10.0, -15.0, -15.0), (20.0, 15.0, 15.0)) if inside else rng.uniform((-25.0, -30.0, -30.0), (35.0, 30.0, 30.0))
        aim = rng.uniform((-10.0, -15.0, -15.0), (20.0, 15.0, 15.0))
        if np.linalg.norm(pos - aim) < 0.5 or not np.isfinite(O.orient(tuple(pos), tuple(aim))):
            continue
        w, h = (int(x) for x in rng.integers(sizes[0], sizes[1], 2))
        views.append(capi.make_view(w, h, float(rng.uniform(10.0, 70.0)), float(rng.uniform(10.0, 70.0)), tuple(pos),
                                    tuple(aim), float(rng.uniform(20.0, 80.0)), int(rng.integers(*samples)), len(views) % 2,
                                    capi.MATH_FAST))
    return views


def test_signed_noise_cube_fast_batch_server(device, oracle):
    vals = noise_cube()
    spc = (30.0 / 39,) * 3
    vol = capi.Volume.upload(device, vals, ORG, spc)
    oc = oracle.OracleCube(vals, ORG, spc)
    clipped = oracle.OracleCube(np.maximum(vals, 0.0), ORG, spc)
    views = random_views(16, seed=21)
    srv = None
    b = None
    try:
        refs, singles, lit, stopped = [], [], 0, 0
        for k, v in enumerate(views):
            ref, _ = oracle.render_f64(oc, oview(oracle, v), v.semantics, nthreads=NT)
            img, _ = vol.render(v)
            refs.append(ref)
            singles.append(img)
            if not ref.any():
                assert not img.any(), k
                continue
            lit += 1
            assert rel_l2(img, ref) <= 1e-5, (k, rel_l2(img, ref))
            nostop, _ = oracle.render_f64(clipped, oview(oracle, v), v.semantics, nthreads=NT)
            stopped += rel_l2(nostop, ref) > 1e-2
        assert lit >= 10 and stopped >= 8, (lit, stopped)
        b = capi.Batch(vol, views)
        b.launch()
        b.wait()
        for k in range(len(views)):
            img = b.image(k)
            if refs[k].any():
                assert rel_l2(img, refs[k]) <= 1e-5, ("batch", k)
            else:
                assert not img.any()
        srv = capi.Server(device, memory_budget=vol.nbytes() + (64 << 20), max_batch=8, max_inflight=2)
        idx = srv.add_volume(vol)
        ids = [srv.submit(idx, v) for v in views]
        for k, rid in enumerate(ids):
            st, img, _ = srv.wait(rid, 30000.0)
            assert st == capi.OK
            assert np.array_equal(img.view(np.uint32), singles[k].view(np.uint32)), ("served", k)
    finally:
        if srv is not None:
            srv.destroy()
        if b is not None:
            b.free()
        vol.free()


@pytest.mark.parametrize("dtype", [capi.DTYPE_F32, capi.DTYPE_F16, capi.DTYPE_BF16])
def test_signed_sphere_cube_each_storage(device, oracle, dtype):
    n = 48
    vals = sphere_cube(n)
    spc = (30.0 / (n - 1),) * 3
    held = {capi.DTYPE_F32: vals, capi.DTYPE_F16: as_f16_storage(vals), capi.DTYPE_BF16: as_bf16_storage(vals)}[dtype]
    vol = capi.Volume.upload(device, vals, ORG, spc, dtype=dtype)
    oc = oracle.OracleCube(held, ORG, spc)
    clipped = oracle.OracleCube(np.maximum(held, 0.0), ORG, spc)
    try:
        lit = stopped = 0
        for k, v in enumerate(random_views(12, seed=31 + dtype)):
            ref, _ = oracle.render_f64(oc, oview(oracle, v), v.semantics, nthreads=NT)
            img, _ = vol.render(v)
            if not ref.any():
                assert not img.any(), k
                continue
            lit += 1
            err = rel_l2(img, ref)
            print(f"dtype {dtype} view {k}: rel L2 {err:.3e}")
            assert err <= 1e-5, (k, err)
            nostop, _ = oracle.render_f64(clipped, oview(oracle, v), v.semantics, nthreads=NT)
            stopped += rel_l2(nostop, ref) > 1e-2
        assert lit >= 8 and stopped >= 4, (lit, stopped)
    finally:
        vol.free()
