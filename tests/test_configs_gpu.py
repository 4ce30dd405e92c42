"""The benchmark workloads themselves, checked on the GPU against the CPU oracle:

  * the FP64 render kernel (the truth the fast-path tests compare against) against the oracle's
    FP64 render: the 576x336 x 800 view of the 256^3 cube, an in-cube orbit view under both
    semantics, a 576x336 x 1600 orbit view of the 512^3 cube, and a CMEM view -- 1e-12;
  * the batched workload: the 256 orbit poses over the 256^3 cube from F16 storage exactly as
    bench.py builds the batch -- every image within 1e-4 of the FP64 kernel's render, a seeded 8
    within 1e-4 of the oracle's FP64 render;
  * the precision workload: the 512^3 wide-range cube in F16 and BF16 storage on 8 of its 64
    poses -- F16 within 1e-4 of FP64, BF16 reported and bounded by 1e-3 (its 8-bit significand
    does not reach 1e-4), one pose also against the oracle's FP64 render;
  * the serving workload: the 8 resident cubes and a seeded 200-request slice of the Poisson
    stream through the server -- every result bitwise equal to birp_render of its view, a seeded
    subset within 1e-4 of the oracle's FP64 render of the FP32 cube.
"""
import os

import numpy as np
import pytest

from birp_cpp_amd import capi
from birp_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oview(O, v):
    return O.make_view(v.width, v.height, v.fov_w, v.fov_h, list(v.position), list(v.aim), v.ray_length, v.samples)


@pytest.fixture(scope="module")
def cube256(device, oracle):
    vals, org, spc, _ = W.cube_256()
    v32 = capi.Volume.upload(device, vals, org, spc)
    yield vals, org, spc, v32, oracle.OracleCube(vals, org, spc)
    v32.free()


def test_f64_kernel_matches_oracle_on_the_workload_views(device, oracle, cube256):
    vals, org, spc, v32, oc = cube256
    v = W.view_576x336()
    got = v32.render_f64(v)
    ref, _ = oracle.render_f64(oc, oview(oracle, v), oracle.SEM_EXTENDED, nthreads=NT)
    assert ref.any() and rel_l2(got, ref) <= 1e-12
    inside = W.view(288, 168, 0.1, pos=(12.0, 5.0, 6.0))  # a camera inside the cube: both semantics
    for sem in (capi.SEM_REFERENCE, capi.SEM_EXTENDED):
        inside.semantics = sem
        got = v32.render_f64(inside)
        ref, _ = oracle.render_f64(oc, oview(oracle, inside), sem, nthreads=NT)
        assert ref.any() and rel_l2(got, ref) <= 1e-12, sem
    cm = capi.Volume.cmem(device, capi.cmem_init())
    try:
        cv = W.view(144, 84, 0.25, pos=(1.0, 1.0, 1.0), aim=(1.0, 0.0, 0.0), length=44.0)
        got = cm.render_f64(cv)
        ref, _ = oracle.render_f64(oracle.OracleCmem(None), oview(oracle, cv), oracle.SEM_EXTENDED, nthreads=NT)
        assert ref.any() and rel_l2(got, ref) <= 1e-12
    finally:
        cm.free()


def test_batched_workload_f16_against_fp64(device, oracle, cube256):
    vals, org, spc, v32, oc = cube256
    v16 = capi.Volume.upload(device, vals, org, spc, dtype=capi.DTYPE_F16)
    tune = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "birp_cpp_amd", "tuning.jsonl")
    if os.path.exists(tune):
        capi.lib().birp_autotune_load(tune.encode())
    views = W.orbit_views_mixed(256)
    b = capi.Batch(v16, views)
    try:
        b.launch()
        b.wait()
        worst = 0.0
        for k, v in enumerate(views):
            ref = v32.render_f64(v)
            e = rel_l2(b.image(k), ref) if ref.any() else float(np.abs(b.image(k)).max())
            worst = max(worst, e)
            assert e <= 1e-4, (k, e)
        for k in np.random.default_rng(3).choice(256, 8, replace=False):
            v = views[int(k)]
            ref, _ = oracle.render_f64(oc, oview(oracle, v), v.semantics, nthreads=NT)
            if ref.any():
                assert rel_l2(b.image(int(k)), ref) <= 1e-4, int(k)
        print(f"256-view F16 batch: worst rel L2 vs FP64 {worst:.3e}")
    finally:
        b.free()
        v16.free()


def test_precision_workload_512(device, oracle):
    vals, org, spc, _ = W.cube_512_wide_range(512)
    v32 = capi.Volume.upload(device, vals, org, spc)
    v16 = capi.Volume.upload(device, vals, org, spc, dtype=capi.DTYPE_F16)
    vb = capi.Volume.upload(device, vals, org, spc, dtype=capi.DTYPE_BF16)
    try:
        views = W.orbit_views_576x336_n1600(64)
        picked = [int(k) for k in np.random.default_rng(4).choice(64, 8, replace=False)]
        lit = 0
        for k in picked:
            v = views[k]
            ref = v32.render_f64(v)
            if not ref.any():
                continue
            lit += 1
            f16, _ = v16.render(v)
            bf, _ = vb.render(v)
            e16, eb = rel_l2(f16, ref), rel_l2(bf, ref)
            print(f"pose {k}: F16 {e16:.3e}  BF16 {eb:.3e}")
            assert e16 <= 1e-4 and eb <= 1e-3, (k, e16, eb)
        assert lit >= 4
        k = next(k for k in picked if v32.render_f64(views[k]).any())
        ref, _ = oracle.render_f64(oracle.OracleCube(vals, org, spc), oview(oracle, views[k]), oracle.SEM_EXTENDED, nthreads=NT)
        f16, _ = v16.render(views[k])
        assert rel_l2(f16, ref) <= 1e-4
        assert rel_l2(v32.render_f64(views[k]), ref) <= 1e-12
    finally:
        for v in (v32, v16, vb):
            v.free()


def test_serving_workload_slice(device, oracle):
    cubes, fp32 = [], {}
    try:
        for k, (n, dt) in enumerate(W.serving_cubes()):
            vals, org, spc, _ = capi.synth_shell(n, (7.0, 12.0), 0.1, seed=500 + k)
            cubes.append(capi.Volume.upload(device, vals, org, spc, dtype=dt))
            if k in (0, 1, 2) or dt == capi.DTYPE_F32:
                fp32.setdefault(k, (vals, org, spc, dt))
            del vals
        t_arr, which, views = W.serving_requests(2000, 200.0)
        sel = np.sort(np.random.default_rng(55).choice(2000, 200, replace=False))
        srv = capi.Server(device, memory_budget=sum(c.nbytes() for c in cubes) + (1 << 30), max_batch=32, max_inflight=4,
                          batch_window_ms=0.2)
        try:
            idx = [srv.add_volume(c) for c in cubes]
            ids = [srv.submit(idx[which[k]], views[k]) for k in sel]
            served = []
            for rid in ids:
                st, img, info = srv.wait(rid, 60000.0)
                assert st == capi.OK
                served.append(img)
        finally:
            srv.destroy()
        for q, k in enumerate(sel):
            single, _ = cubes[which[k]].render(views[k])
            assert np.array_equal(single.view(np.uint32), served[q].view(np.uint32)), int(k)
        checked = 0
        for q, k in enumerate(sel):
            c = int(which[k])
            if c not in fp32 or checked >= 6:
                continue
            vals, org, spc, dt = fp32[c]
            ref, _ = oracle.render_f64(oracle.OracleCube(vals, org, spc), oview(oracle, views[k]), oracle.SEM_EXTENDED, nthreads=NT)
            if not ref.any():
                continue
            tol = 1e-5 if dt == capi.DTYPE_F32 else 1e-4
            assert rel_l2(served[q], ref) <= tol, (int(k), c, dt)
            checked += 1
        assert checked >= 3
    finally:
        for c in cubes:
            c.free()
