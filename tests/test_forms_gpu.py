"""GPU tests of the other forms: CMEM renders, the FP64 render, the fast (FP64 entry, Kahan) path
with FP32/BF16/F16 storage, batched persistent-kernel renders under a HIP graph, and the
continuous-batching server.

Tolerances (image relative L2 error):
  CMEM exact path vs the FP32 oracle ......... 1e-5 off the model's zone edges (tests/parity.py)
  FP64 render vs the oracle's FP64 render .... 1e-12
  fast FP32 storage vs FP64 .................. 1e-5
  fast F16 storage (per-brick scale) vs FP64 . 1e-4
  fast BF16 storage vs FP64 .................. reported; bounded by 1e-2
"""
import os

import numpy as np
import pytest

from birp_cpp_amd import capi
from birp_cpp_amd import workloads as W
from parity import CMEM_EDGE_BAND, cmem_render_parity

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def oview(O, v):
    return O.make_view(v.width, v.height, v.fov_w, v.fov_h, list(v.position), list(v.aim), v.ray_length, v.samples)


@pytest.fixture(scope="module")
def cube1(device, oracle):
    vals, org, spc, _ = W.cube_128()
    vols = {dt: capi.Volume.upload(device, vals, org, spc, dtype=dt) for dt in (capi.DTYPE_F32, capi.DTYPE_BF16, capi.DTYPE_F16)}
    yield vals, org, spc, vols, oracle.OracleCube(vals, org, spc)
    for v in vols.values():
        v.free()


# ---------------------------------------------------------------------------------- CMEM
@pytest.mark.parametrize("params", [None, dict(v=[400, 0, 0], b=[0, 0, 4], dipole=0.0, p0=1, p1=4, p2=6, p3=3, B=3.6, alpha=-1.0)])
def test_cmem_points_and_render(device, oracle, params):
    oc = oracle.OracleCmem(params)
    vol = capi.Volume.cmem(device, capi.cmem_state_from(oc.state()))
    try:
        # the reference's own known answers (tests/CMEM_test.cpp:14-90)
        if params is None:
            pts = [(8, 4, 4), (4, 12, 10), (3, 1.5, 1.2), (8, 10, 10), (8, -10, -10), (8, 10, -10)]
            want = [0.0, 4.46853346e-06, 0.0, 8.15947317e-07, 8.15947317e-07, 8.15947317e-07]
        else:
            pts = [(8, 4, 4), (12, 12, 10), (3, 1.5, 1.2), (12, 10, 10), (12, -10, -10), (12, 10, -10)]
            want = [0.0, 4.57952495e-07, 0.0, 5.48567925e-07, 5.48567925e-07, 5.48567925e-07]
        got = vol.sample(np.array(pts, np.float32))
        np.testing.assert_allclose(got, np.array(want, np.float32), rtol=4e-6, atol=0)
        rng = np.random.default_rng(3)
        P = rng.uniform(-20, 20, (5000, 3)).astype(np.float32)
        g = vol.sample(P)
        r = np.array([oc.sample(*map(float, p)) for p in P], np.float32)
        close = np.isclose(g, r, rtol=2e-5, atol=1e-12)
        # a point may differ by more only where a zone edge lies within the band, and then by at
        # most the model's jump there
        for k in np.nonzero(~close)[0]:
            p = tuple(map(float, P[k]))
            assert oc.edge_distance(*p) < CMEM_EDGE_BAND, (p, g[k], r[k])
            assert np.isfinite(g[k]) and abs(float(g[k]) - float(r[k])) <= oc.edge_jump(*p, CMEM_EDGE_BAND) * 1.001 + 2e-5 * abs(float(r[k])), (p, g[k], r[k])
        v = W.view(144, 84, 0.25, pos=(1.0, 1.0, 1.0), aim=(1.0, 0.0, 0.0), length=44.0)
        img, st = vol.render(v)
        ov = oview(oracle, v)
        ref, _ = oracle.render_f32(oc, ov, oracle.SEM_EXTENDED, nthreads=NT)
        ref64, _ = oracle.render_f64(oc, ov, oracle.SEM_EXTENDED, nthreads=NT)
        cmem_render_parity(img, ref, ref64, *oc.edge_bound(ov, CMEM_EDGE_BAND, nthreads=NT))
        assert st.samples_in_cube == v.width * v.height * v.samples
    finally:
        vol.free()


def test_cmem_render_576x336_800_samples(device, oracle):
    """The CMEM render bench.py times (the default model, 576x336 pixels, 800 samples per ray over
    80 RE from the workload camera): non-edge pixels within 1e-5 of the FP32 oracle, every edge
    pixel within its zone-flip bound."""
    oc = oracle.OracleCmem(None)
    vol = capi.Volume.cmem(device, capi.cmem_state_from(oc.state()))
    try:
        v = W.view_576x336()
        img, st = vol.render(v)
        ov = oview(oracle, v)
        ref, _ = oracle.render_f32(oc, ov, oracle.SEM_EXTENDED, nthreads=NT)
        ref64, _ = oracle.render_f64(oc, ov, oracle.SEM_EXTENDED, nthreads=NT)
        e32, e64, whole, frac, worst = cmem_render_parity(img, ref, ref64, *oc.edge_bound(ov, CMEM_EDGE_BAND, nthreads=NT))
        print(f"CMEM 576x336x800: non-edge {e32:.3e} vs FP32, {e64:.3e} vs FP64; whole image {whole:.3e}; "
              f"edge pixels {frac:.4f}, worst edge pixel at {worst:.3f} of its bound")
        assert st.samples_in_cube == v.width * v.height * v.samples
    finally:
        vol.free()


@pytest.mark.parametrize("params", [None, dict(v=[400, 0, 0], b=[0, 0, 4], dipole=0.0, p0=1, p1=4, p2=6, p3=3, B=3.6, alpha=-1.0)])
def test_fast_cmem_as_accurate_as_the_fp32_model(device, oracle, params):
    """The fast CMEM form (FP32 with the hardware exp2/log2, coordinate-ratio cosines and polynomial
    arc cosines; Kahan sums) against the oracle's FP64 model: the relative L2 error of every image is
    at most twice the reference FP32 model's own error against FP64 (plus 1e-6), over the whole image
    and off the zone edges; a batch of the same views gives the single renders' images to 1e-6."""
    oc = oracle.OracleCmem(params)
    vol = capi.Volume.cmem(device, capi.cmem_state_from(oc.state()))
    b = None
    try:
        views = [W.view(144, 84, 0.25, pos=(1.0, 1.0, 1.0), aim=(1.0, 0.0, 0.0), length=44.0, math=capi.MATH_FAST),
                 W.view(288, 168, 0.1, math=capi.MATH_FAST),
                 W.view(144, 84, 0.5, pos=(12.0, 9.0, 4.0), aim=(-2.0, 0.0, 0.0), length=60.0, math=capi.MATH_FAST)]
        singles = []
        for v in views:
            fast, _ = vol.render(v)
            singles.append(fast)
            ov = oview(oracle, v)
            ref32, _ = oracle.render_f32(oc, ov, oracle.SEM_EXTENDED, nthreads=NT)
            ref64, _ = oracle.render_f64(oc, ov, oracle.SEM_EXTENDED, nthreads=NT)
            mask, _ = oc.edge_bound(ov, CMEM_EDGE_BAND, nthreads=NT)
            keep = ~mask
            own, err = rel_l2(ref32, ref64), rel_l2(fast, ref64)
            own_k, err_k = rel_l2(ref32[keep], ref64[keep]), rel_l2(fast[keep], ref64[keep])
            print(f"fast CMEM {v.width}x{v.height}: vs FP64 {err:.3e} (FP32 model {own:.3e}); off edges {err_k:.3e} ({own_k:.3e})")
            assert np.isfinite(fast).all() and fast.any()
            assert err <= 2 * own + 1e-6 and err_k <= 2 * own_k + 1e-6
        b = capi.Batch(vol, views)
        b.launch()
        b.wait()
        for k in range(len(views)):
            assert rel_l2(b.image(k), singles[k]) <= 1e-6, k
    finally:
        if b is not None:
            b.free()
        vol.free()


# ---------------------------------------------------------------------------------- FP64
def test_f64_render_matches_oracle_f64(cube1, oracle):
    vals, org, spc, vols, oc = cube1
    v = W.view_144x84()
    got = vols[capi.DTYPE_F32].render_f64(v)
    ref, _ = oracle.render_f64(oc, oview(oracle, v), oracle.SEM_EXTENDED, nthreads=NT)
    assert rel_l2(got, ref) <= 1e-12
    img, _ = vols[capi.DTYPE_F32].render(v)
    e32 = rel_l2(img, ref)
    print(f"exact FP32 path vs FP64 oracle: {e32:.3e}")
    assert e32 <= 1e-5


# ---------------------------------------------------------------------------------- fast path
@pytest.mark.parametrize("dtype,tol", [(capi.DTYPE_F32, 1e-5), (capi.DTYPE_F16, 1e-4), (capi.DTYPE_BF16, 1e-2)])
def test_fast_path_vs_fp64(cube1, oracle, dtype, tol):
    vals, org, spc, vols, oc = cube1
    for pose in [W.CAM_POS, (15.0, 5.0, 8.0), (3.0, 17.0, 12.0)]:
        v = W.view(144, 84, 0.25, pos=pose, math=capi.MATH_FAST)
        ref = vols[capi.DTYPE_F32].render_f64(v)
        img, _ = vols[dtype].render(v)
        err = rel_l2(img, ref)
        mx = float(np.max(np.abs(img - ref) / np.maximum(np.abs(ref), 1e-30)))
        print(f"dtype {dtype} pose {pose}: rel L2 {err:.3e} max rel {mx:.3e}")
        assert err <= tol


def test_fast_path_random_views(cube1, oracle):
    """A seeded sweep of 32 random fast views of the 128^3 cube (cameras inside and outside it,
    16..48 px a side, 200..1200 samples, both semantics): FP32 storage within 1e-5 and F16 within
    1e-4 relative L2 of the FP64 render, and a view that sees nothing renders all zeros."""
    vals, org, spc, vols, oc = cube1
    rng = np.random.default_rng(4242)
    done = lit = 0
    while done < 32:
        inside = rng.random() < 0.7
        pos = rng.uniform((-10.0, -15.0, -15.0), (20.0, 15.0, 15.0)) if inside else rng.uniform((-30.0, -35.0, -35.0), (40.0, 35.0, 35.0))
        aim = rng.uniform((-10.0, -15.0, -15.0), (20.0, 15.0, 15.0))
        if np.linalg.norm(pos - aim) < 0.5 or not np.isfinite(oracle.orient(tuple(pos), tuple(aim))):
            continue
        w, h = (int(x) for x in rng.integers(16, 49, 2))
        v = capi.make_view(w, h, float(rng.uniform(5.0, 70.0)), float(rng.uniform(5.0, 70.0)), tuple(pos), tuple(aim),
                           float(rng.uniform(10.0, 90.0)), int(rng.integers(200, 1201)), done % 2, capi.MATH_FAST)
        ref = vols[capi.DTYPE_F32].render_f64(v)
        for dt, tol in ((capi.DTYPE_F32, 1e-5), (capi.DTYPE_F16, 1e-4)):
            img, _ = vols[dt].render(v)
            if not ref.any():
                assert not img.any(), (done, dt)
            else:
                assert rel_l2(img, ref) <= tol, (done, dt, pos, aim, w, h)
        lit += bool(ref.any())
        done += 1
    assert lit >= 16


def test_f16_scale_modes(device):
    """F16 storage keeps one power-of-two scale for the whole volume when every nonzero |v| then
    lies in F16's normal range, and a scale per 8^3 brick otherwise. A cube spanning 30 decades
    (v = 10^(-30 x/X), far beyond F16's 2^29 range) takes the per-brick path; its value range
    scaled into one brick still fits, so both paths meet the F16 tolerance against FP64."""
    n = 64
    x = np.arange(n, dtype=np.float64)
    wide = np.broadcast_to(10.0 ** (-30.0 * x / (n - 1)), (n, n, n)).astype(np.float32).copy()
    narrow = np.broadcast_to(1e-6 * (1.0 + 0.5 * np.sin(x / 5.0)), (n, n, n)).astype(np.float32).copy()
    org, spc = (-10.0, -15.0, -15.0), (30 / (n - 1),) * 3
    for vals in (wide, narrow):
        v32 = capi.Volume.upload(device, vals, org, spc, dtype=capi.DTYPE_F32)
        v16 = capi.Volume.upload(device, vals, org, spc, dtype=capi.DTYPE_F16)
        try:
            for pose in [(15.0, 5.0, 8.0), (3.0, -7.0, 2.0)]:
                view = W.view(96, 64, 0.25, pos=pose, math=capi.MATH_FAST)
                ref = v32.render_f64(view)
                img, _ = v16.render(view)
                err = rel_l2(img, ref)
                print(f"range {vals.max() / vals.min():.1e} pose {pose}: F16 rel L2 {err:.3e}")
                assert err <= 1e-4
        finally:
            v32.free()
            v16.free()


def test_fast_path_reference_semantics(cube1, oracle):
    vals, org, spc, vols, oc = cube1
    v = W.view(96, 64, 0.25, pos=(15.0, 5.0, 8.0), aim=(0.0, 0.0, 0.0), semantics=capi.SEM_REFERENCE, math=capi.MATH_FAST, length=44.0)
    ref = vols[capi.DTYPE_F32].render_f64(v)
    img, _ = vols[capi.DTYPE_F32].render(v)
    assert rel_l2(img, ref) <= 1e-5
    outside = W.view(32, 32, 0.5, semantics=capi.SEM_REFERENCE, math=capi.MATH_FAST)
    img, _ = vols[capi.DTYPE_F32].render(outside)
    assert not img.any()


# ---------------------------------------------------------------------------------- batch
def test_batch_matches_fp64_and_replays(cube1):
    vals, org, spc, vols, oc = cube1
    views = W.orbit_views_mixed(count=24)
    b = capi.Batch(vols[capi.DTYPE_F16], views)
    try:
        b.launch()  # the kernels straight onto the stream
        b.wait()
        eager = [b.image(k) for k in range(len(views))]
        for k, v in enumerate(views):
            ref = vols[capi.DTYPE_F32].render_f64(v)
            assert rel_l2(eager[k], ref) <= 1e-4, k
        # every graph replay, with the program's own eager work and syncs in between, equals
        # the direct launch bit for bit
        for rep in range(4):
            _ = vols[capi.DTYPE_F32].render(W.view_144x84())
            b.launch_graph()
            b.wait()
            for k in range(len(views)):
                assert np.array_equal(b.image(k), eager[k]), (rep, k)
            if rep == 1:
                b.launch_eager()
                b.wait()
        n, ms = b.stats()
        assert n == sum(v.width * v.height * v.samples for v in views) and ms > 0
    finally:
        b.free()


def test_batch_output_into_pinned_host_memory(cube1):
    """birp_batch_set_output: the kernel writes every image of the batch into a pinned host arena
    (graph replays and eager launches), bit for bit the images it writes to device memory; the
    output must be a birp_host_alloc buffer of the batch's pixels, and None restores device
    images."""
    vals, org, spc, vols, oc = cube1
    views = W.orbit_views_mixed(count=24, seed=9)
    b = capi.Batch(vols[capi.DTYPE_F16], views)
    try:
        b.launch()
        b.wait()
        want = [b.image(k) for k in range(len(views))]
        offs = np.cumsum([0] + [v.width * v.height for v in views])
        with pytest.raises(capi.BirpError) as e:
            b.set_output(np.zeros(offs[-1], np.float32))  # pageable memory is refused
        assert e.value.status == capi.ERR_INVALID_ARGUMENT
        arena = capi.pinned_empty((int(offs[-1]),))
        b.set_output(arena)
        for launch in (b.launch, b.launch_eager, b.launch):
            arena.fill(np.nan)
            launch()
            b.wait()
            for k, v in enumerate(views):
                img = arena[offs[k]:offs[k + 1]].reshape(v.height, v.width)
                assert np.array_equal(img.view(np.uint32), want[k].view(np.uint32)), k
                assert np.array_equal(b.image(k), want[k])
        b.set_output(None)
        arena.fill(np.nan)
        b.launch()
        b.wait()
        assert np.isnan(arena).all()  # back to device images
        assert np.array_equal(b.image(3), want[3])
    finally:
        b.free()


@pytest.mark.parametrize("tuning", [(8, 8, 1, 1), (16, 4, 1, 2), (8, 4, 2, 4), (4, 8, 2, 1), (4, 4, 4, 2), (8, 2, 4, 4)])
def test_batch_tunings_agree(cube1, tuning):
    vals, org, spc, vols, oc = cube1
    views = W.orbit_views_mixed(count=6)
    b = capi.Batch(vols[capi.DTYPE_F32], views)
    try:
        b.set_tuning(capi.Tuning(*tuning))
        b.launch()
        b.wait()
        for k, v in enumerate(views):
            ref = vols[capi.DTYPE_F32].render_f64(v)
            assert rel_l2(b.image(k), ref) <= 1e-5
    finally:
        b.free()


def test_autotune_roundtrip(cube1, tmp_path):
    vals, org, spc, vols, oc = cube1
    views = W.orbit_views_mixed(count=8)
    path = str(tmp_path / "tune.jsonl")
    capi.autotune(vols[capi.DTYPE_F16], views, path)
    assert capi.lib().birp_autotune_load(path.encode()) == capi.OK
    v = views[0]
    t = capi.autotune_lookup(capi.DTYPE_F16, v.width, v.height, v.samples)
    assert t is not None and t.tile_w * t.tile_h * t.lanes_per_ray == 64


# ---------------------------------------------------------------------------------- server
def test_server_serves_cmem_models_beside_cubes(device, cube1):
    """CMEM models are resident volumes too: their requests render in the math they name (the exact
    arithmetic through the multi-view exact kernel, or the fast CMEM kernel) in the same slot
    launches as cube requests, each result bitwise equal to birp_render of its view in that math
    (fast for the cubes), in launches that mix all of them."""
    vals, org, spc, vols, oc = cube1
    cms = [capi.Volume.cmem(device, capi.cmem_init()),
           capi.Volume.cmem(device, capi.cmem_init(capi.cmem_params(p0=1.0, p1=4, p2=6, p3=3, B=3.6, alpha=-1.0)))]
    srv = capi.Server(device, memory_budget=4 << 30, max_batch=24, max_inflight=2, batch_window_ms=0.5)
    try:
        idx = [srv.add_volume(vols[capi.DTYPE_F16]), srv.add_volume(cms[0]), srv.add_volume(cms[1]), srv.add_volume(vols[capi.DTYPE_F32])]
        spaces = [vols[capi.DTYPE_F16], cms[0], cms[1], vols[capi.DTYPE_F32]]
        views = W.orbit_views_mixed(count=40, seed=13)
        which = [k % 4 for k in range(len(views))]
        # CMEM requests: exact for odd k // 4, fast for even
        maths = [capi.MATH_EXACT if (w in (1, 2) and (k // 4) % 2) else capi.MATH_FAST for k, w in enumerate(which)]
        for v, m in zip(views, maths):
            v.math = m
        ids = [srv.submit(idx[w], v) for v, w in zip(views, which)]
        mixed = 0
        for k, rid in enumerate(ids):
            st, img, info = srv.wait(rid, 60000.0)
            assert st == capi.OK
            mixed += info.batch_size > 1
            v = capi.make_view(views[k].width, views[k].height, views[k].fov_w, views[k].fov_h, tuple(views[k].position),
                               tuple(views[k].aim), views[k].ray_length, views[k].samples, views[k].semantics, maths[k])
            want, _ = spaces[which[k]].render(v)
            assert np.array_equal(img.view(np.uint32), want.view(np.uint32)), (k, which[k])
            if which[k] in (1, 2):
                assert img.any()
        assert mixed > 0
    finally:
        srv.destroy()
        for c in cms:
            c.free()


def test_server_alone_equals_in_mix(device, cube1):
    vals, org, spc, vols, oc = cube1
    srv = capi.Server(device, memory_budget=4 << 30, max_batch=16, max_inflight=3)
    try:
        i16 = srv.add_volume(vols[capi.DTYPE_F16])
        i32 = srv.add_volume(vols[capi.DTYPE_F32])
        views = W.orbit_views_mixed(count=20)
        alone = []
        for k, v in enumerate(views):
            rid = srv.submit(i16 if k % 2 else i32, v)
            st, img, info = srv.wait(rid, 30000.0)
            assert st == capi.OK and info.batch_size == 1
            alone.append(img)
        ids = [srv.submit(i16 if k % 2 else i32, v) for k, v in enumerate(views)]
        sizes = []
        for k, rid in enumerate(ids):
            st, img, info = srv.wait(rid, 30000.0)
            assert st == capi.OK
            sizes.append(info.batch_size)
            assert np.array_equal(img, alone[k]), k
        assert max(sizes) > 1  # requests did share launches
    finally:
        srv.destroy()


def test_server_concurrent_clients(device, cube1):
    """Four client threads submit and wait concurrently (ctypes releases the GIL in the C calls);
    every result equals the same request rendered alone, and requests share launches."""
    import threading
    vals, org, spc, vols, oc = cube1
    srv = capi.Server(device, memory_budget=4 << 30, max_batch=32, max_inflight=3, batch_window_ms=0.5,
                      reserve_pixels=32 * 576 * 336)
    try:
        idx = [srv.add_volume(vols[capi.DTYPE_F16]), srv.add_volume(vols[capi.DTYPE_F32]), srv.add_volume(vols[capi.DTYPE_BF16])]
        views = W.orbit_views_mixed(count=60, seed=11)
        which = [k % 3 for k in range(len(views))]
        alone = []
        for v, w in zip(views, which):
            st, img, info = srv.wait(srv.submit(idx[w], v), 30000.0)
            assert st == capi.OK
            alone.append(img)
        results, errors, sizes = {}, [], []

        def client(part):
            try:
                ids = [(k, srv.submit(idx[which[k]], views[k])) for k in part]
                for k, rid in ids:
                    st, img, info = srv.wait(rid, 60000.0)
                    results[k] = (st, img)
                    sizes.append(info.batch_size)
            except Exception as e:  # reported below
                errors.append(e)

        threads = [threading.Thread(target=client, args=(range(c, len(views), 4),)) for c in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(120)
        assert not errors, errors
        assert len(results) == len(views)
        for k, (st, img) in results.items():
            assert st == capi.OK and np.array_equal(img, alone[k]), k
        assert max(sizes) > 1
    finally:
        srv.destroy()


def test_server_admission_and_injected_failure(device, cube1):
    vals, org, spc, vols, oc = cube1
    vb = vols[capi.DTYPE_F16].nbytes()
    # budget: the volume plus device room for ~1.5 small images
    srv = capi.Server(device, memory_budget=vb + int(1.5 * 144 * 84 * 4), max_batch=8, max_inflight=2)
    try:
        idx = srv.add_volume(vols[capi.DTYPE_F16])
        big = W.view(576, 336, 0.5, math=capi.MATH_FAST)
        st, img, info = srv.wait(srv.submit(idx, big), 30000.0)
        assert st == capi.ERR_ADMISSION and img is None
        small = [W.view(144, 84, 0.5, pos=(float(k), -20.0, 10.0), math=capi.MATH_FAST) for k in range(6)]
        ids = [srv.submit(idx, v) for v in small]
        for rid in ids:
            st, img, info = srv.wait(rid, 30000.0)
            assert st == capi.OK and info.batch_size == 1  # only one fits at a time
        device.inject_failure(1)
        ids = [srv.submit(idx, v) for v in small[:3]]
        res = [srv.wait(rid, 30000.0)[0] for rid in ids]
        assert res[0] == capi.ERR_INJECTED and res[1:] == [capi.OK, capi.OK]
        tiny = capi.Server(device, memory_budget=1024)
        try:
            with pytest.raises(capi.BirpError) as e:
                tiny.add_volume(vols[capi.DTYPE_F16])
            assert e.value.status == capi.ERR_ADMISSION
        finally:
            tiny.destroy()
    finally:
        srv.destroy()


def test_server_budget_counts_held_image_buffers(device, cube1):
    """The memory budget bounds the resident volumes plus the image buffers the slots hold between
    batches: reservations beyond it are refused, so is a volume that would exceed it, and a
    request larger than a slot's buffer gets the room by releasing an idle slot's buffer. Results
    equal those of a server with ample memory."""
    vals, org, spc, vols, oc = cube1
    vb = vols[capi.DTYPE_F16].nbytes()
    R = 144 * 84
    with pytest.raises(capi.BirpError) as e:
        capi.Server(device, memory_budget=2 * R * 4 - 1, max_batch=4, max_inflight=2, reserve_pixels=R)
    assert e.value.status == capi.ERR_ADMISSION
    srv = capi.Server(device, memory_budget=vb + 2 * R * 4, max_batch=4, max_inflight=2, reserve_pixels=R)
    ample = capi.Server(device, memory_budget=4 << 30, max_batch=4, max_inflight=2)
    try:
        idx = srv.add_volume(vols[capi.DTYPE_F16])
        with pytest.raises(capi.BirpError) as e:
            srv.add_volume(vols[capi.DTYPE_F32])
        assert e.value.status == capi.ERR_ADMISSION
        ia = ample.add_volume(vols[capi.DTYPE_F16])
        views = [W.view(144, 84, 0.5, math=capi.MATH_FAST),           # fits a slot's reservation
                 W.view(168, 108, 0.5, math=capi.MATH_FAST),          # 1.5 reservations: the other slot's is released
                 W.view(144, 84, 0.25, pos=(2.0, -20.0, 10.0), math=capi.MATH_FAST)]
        for v in views:
            st, img, info = srv.wait(srv.submit(idx, v), 30000.0)
            st2, want, _ = ample.wait(ample.submit(ia, v), 30000.0)
            assert st == st2 == capi.OK and info.batch_size == 1
            assert np.array_equal(img, want)
        # two of the larger requests at once cannot share the budget: one launch each
        ids = [srv.submit(idx, views[1]) for _ in range(2)]
        for rid in ids:
            st, img, info = srv.wait(rid, 30000.0)
            assert st == capi.OK and info.batch_size == 1
        with pytest.raises(capi.BirpError) as e:
            srv.add_volume(vols[capi.DTYPE_F32])
        assert e.value.status == capi.ERR_ADMISSION
    finally:
        srv.destroy()
        ample.destroy()


def test_served_requests_equal_single_fast_renders(device, cube1):
    """A served request is rendered with the tuning its own view selects, the one birp_render
    uses for a fast render: every request of a mixed burst (three image sizes, three sample
    counts, three storage types, so every lanes-per-ray group) equals birp_render of its view,
    bit for bit."""
    vals, org, spc, vols, oc = cube1
    srv = capi.Server(device, memory_budget=4 << 30, max_batch=32, max_inflight=3)
    try:
        dts = (capi.DTYPE_F32, capi.DTYPE_BF16, capi.DTYPE_F16)
        idx = {dt: srv.add_volume(vols[dt]) for dt in dts}
        pos = capi.synth_orbit(27, 21)
        reqs = []
        for k, p in enumerate(pos):
            w, h = W.SIZES[k % 3]
            step = (0.1, 0.25, 0.5)[(k // 3) % 3]
            reqs.append((dts[(k // 9) % 3], W.view(w, h, step, pos=tuple(float(c) for c in p), math=capi.MATH_FAST)))
        ids = [srv.submit(idx[dt], v) for dt, v in reqs]
        sizes = []
        for (dt, v), rid in zip(reqs, ids):
            st, img, info = srv.wait(rid, 30000.0)
            assert st == capi.OK
            sizes.append(info.batch_size)
            want, _ = vols[dt].render(v)
            assert np.array_equal(img.view(np.uint32), want.view(np.uint32)), (dt, v.width, v.samples)
        assert max(sizes) > 1
    finally:
        srv.destroy()


def test_server_results_into_pinned_buffers(device, cube1):
    """Requests whose result arrays come from birp_host_alloc get their images written there by the
    render kernel (zero-copy through the device mapping); one at a time and in a burst that mixes
    pinned and ordinary result buffers, every image equals the ordinary (staged) path's."""
    vals, org, spc, vols, oc = cube1
    srv = capi.Server(device, memory_budget=4 << 30, max_batch=16, max_inflight=3)
    try:
        idx = srv.add_volume(vols[capi.DTYPE_F16])
        views = W.orbit_views_mixed(count=24, seed=5)
        want = []
        for v in views:
            st, img, _ = srv.wait(srv.submit(idx, v), 30000.0)
            assert st == capi.OK
            want.append(img)
        for k, v in enumerate(views[:6]):  # alone: launches of one request
            out = capi.pinned_empty((v.height, v.width))
            out.fill(np.nan)
            st, img, info = srv.wait(srv.submit(idx, v, out=out), 30000.0)
            assert st == capi.OK and info.batch_size == 1
            assert np.array_equal(img.view(np.uint32), want[k].view(np.uint32)), k
        outs = [capi.pinned_empty((v.height, v.width)) if k % 2 == 0 else np.empty((v.height, v.width), np.float32)
                for k, v in enumerate(views)]
        for o in outs:
            o.fill(np.nan)
        ids = [srv.submit(idx, v, out=o) for v, o in zip(views, outs)]
        for k, rid in enumerate(ids):
            st, img, _ = srv.wait(rid, 30000.0)
            assert st == capi.OK and img is outs[k]
            assert np.array_equal(img.view(np.uint32), want[k].view(np.uint32)), k
        # results at offsets inside one shared pinned arena (device address = base + offset)
        offs = np.cumsum([0] + [v.width * v.height for v in views])
        arena = capi.pinned_empty((int(offs[-1]) + 7,))
        arena.fill(np.nan)
        ids = [srv.submit(idx, v, out=arena[7 + offs[k]:7 + offs[k + 1]].reshape(v.height, v.width))
               for k, v in enumerate(views)]
        for k, rid in enumerate(ids):
            st, img, _ = srv.wait(rid, 30000.0)
            assert st == capi.OK
            assert np.array_equal(img.view(np.uint32), want[k].view(np.uint32)), k
        assert np.isnan(arena[:7]).all()  # nothing written before the first result
    finally:
        srv.destroy()


def test_server_destroy_with_requests_in_flight(device, cube1):
    """birp_server_destroy with requests still queued and launched (none waited for) finishes
    their work, writes every result buffer, then frees the server; nothing hangs."""
    vals, org, spc, vols, oc = cube1
    views = W.orbit_views_mixed(count=40, seed=13)
    ref = capi.Server(device, memory_budget=4 << 30, max_batch=8, max_inflight=2)
    try:
        i = ref.add_volume(vols[capi.DTYPE_F16])
        want = []
        for v in views:
            st, img, _ = ref.wait(ref.submit(i, v), 30000.0)
            assert st == capi.OK
            want.append(img)
    finally:
        ref.destroy()
    srv = capi.Server(device, memory_budget=4 << 30, max_batch=8, max_inflight=2)
    idx = srv.add_volume(vols[capi.DTYPE_F16])
    outs = [capi.pinned_empty((v.height, v.width)) if k % 2 else np.full((v.height, v.width), np.nan, np.float32)
            for k, v in enumerate(views)]
    for o in outs:
        o.fill(np.nan)
    for v, o in zip(views, outs):
        srv.submit(idx, v, out=o)
    srv.destroy()
    for k, o in enumerate(outs):
        assert np.array_equal(o.view(np.uint32), want[k].view(np.uint32)), k
