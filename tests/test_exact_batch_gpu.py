"""Exact batches (birp_batch_create_exact): many views over several volumes -- positive, signed
and nearest-sampled cubes and a CMEM model -- in one launch per sample mode. Every image is
bit-identical to the single exact render of its view (and so to the oracle's FP32 render of the
reference algorithm), every in-cube count equals the single render's, graph replays equal the
eager launch, accumulating views add into the previous launch's images exactly as repeated
Camera::Render calls do, and images can go straight into pinned host memory."""
import os

import numpy as np
import pytest

from birp_cpp_amd import capi
from birp_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def oview(O, v):
    return O.make_view(v.width, v.height, v.fov_w, v.fov_h, list(v.position), list(v.aim), v.ray_length, v.samples)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def spaces(device, oracle):
    rng = np.random.default_rng(17)
    n = 40
    org, spc = (-10.0, -15.0, -15.0), (30.0 / (n - 1),) * 3
    signed = (rng.standard_normal((n, n, n)) * 1e-5 + 2e-6).astype(np.float32)
    signed[rng.random(signed.shape) < 0.15] = 0.0
    vals128, org128, spc128, _ = W.cube_128()
    vols = [capi.Volume.upload(device, vals128, org128, spc128),
            capi.Volume.upload(device, signed, org, spc),
            capi.Volume.upload(device, signed, org, spc, interp=capi.INTERP_NEAREST),
            capi.Volume.cmem(device, capi.cmem_init())]
    ocs = [oracle.OracleCube(vals128, org128, spc128), oracle.OracleCube(signed, org, spc),
           oracle.OracleCube(signed, org, spc, trilinear=False), oracle.OracleCmem(None)]
    yield vols, ocs
    for v in vols:
        v.free()


def mixed_views(count, seed):
    rng = np.random.default_rng(seed)
    pos = capi.synth_orbit(count, seed)
    views, which = [], []
    for k in range(count):
        w, h = (int(x) for x in rng.integers(8, 97, 2))
        p = tuple(float(c) for c in pos[k]) if k % 3 else tuple(float(c) for c in rng.uniform((-8.0, -12.0, -12.0), (18.0, 12.0, 12.0)))
        which.append(k % 4)
        n = int(rng.integers(2, 700)) if which[-1] != 3 else int(rng.integers(2, 200))
        views.append(capi.make_view(w, h, float(rng.uniform(10.0, 60.0)), float(rng.uniform(10.0, 60.0)), p, (10.0, 0.0, 0.0),
                                    float(rng.uniform(10.0, 80.0)), n, int(rng.integers(0, 2)), capi.MATH_EXACT))
    return views, which


def test_exact_batch_mixed_volumes_bitwise(spaces, oracle):
    vols, ocs = spaces
    views, which = mixed_views(32, seed=41)
    b = capi.Batch.exact(vols, views, which)
    try:
        b.launch_eager()
        b.wait()
        eager = [b.image(k) for k in range(len(views))]
        counts = b.samples()
        for k, v in enumerate(views):
            single, st = vols[which[k]].render(v)
            assert np.array_equal(bits(eager[k]), bits(single)), (k, which[k])
            assert counts[k] == st.samples_in_cube, (k, counts[k], st.samples_in_cube)
            if which[k] != 3:  # cubes: the oracle's FP32 render of the reference algorithm
                ref, cnt = oracle.render_f32(ocs[which[k]], oview(oracle, v), v.semantics, nthreads=NT)
                assert np.array_equal(bits(eager[k]), bits(ref)) and counts[k] == cnt, (k, which[k])
        assert sum(counts) > 100_000
        # graph replays, with the program's own eager work in between, equal the eager launch
        for rep in range(3):
            _ = vols[0].render(W.view_144x84())
            b.launch_graph()
            b.wait()
            for k in range(len(views)):
                assert np.array_equal(bits(b.image(k)), bits(eager[k])), (rep, k)
    finally:
        b.free()


def test_exact_batch_accumulates_and_writes_pinned_images(spaces, oracle):
    vols, ocs = spaces
    views, which = mixed_views(12, seed=43)
    for v in views[::2]:
        v.flags = capi.RENDER_ACCUMULATE
    b = capi.Batch.exact(vols, views, which)
    try:
        offs = np.cumsum([0] + [v.width * v.height for v in views])
        arena = capi.pinned_empty((int(offs[-1]),))
        arena.fill(0.0)
        b.set_output(arena)
        for rep in range(2):
            b.launch()
            b.wait()
        for k, v in enumerate(views):
            img = arena[offs[k]:offs[k + 1]].reshape(v.height, v.width)
            one, _ = vols[which[k]].render(v)  # v's accumulate flag: renders into zeros here
            if v.flags & capi.RENDER_ACCUMULATE:
                want, _ = vols[which[k]].render(v, init=one)  # a second render added on top
            else:
                want = one
            assert np.array_equal(bits(img), bits(want)), k
    finally:
        b.free()
        del arena


def test_exact_batch_of_the_256_orbit_views(device, oracle):
    """The 256 orbit poses of the batched workload (mixed image sizes and steps, 256^3 cube) as one
    exact batch: every image bitwise equal to its single exact render, a seeded subset of 8 bitwise
    equal to the oracle's FP32 render, and every in-cube count equal to the single render's."""
    vals, org, spc, _ = W.cube_256()
    vol = capi.Volume.upload(device, vals, org, spc)
    views = W.orbit_views_mixed(256, math=capi.MATH_EXACT)
    b = capi.Batch.exact([vol], views)
    try:
        b.launch()
        b.wait()
        counts = b.samples()
        for k, v in enumerate(views):
            single, st = vol.render(v)
            assert np.array_equal(bits(b.image(k)), bits(single)), k
            assert counts[k] == st.samples_in_cube, k
        oc = oracle.OracleCube(vals, org, spc)
        for k in np.random.default_rng(5).choice(256, 8, replace=False):
            v = views[int(k)]
            ref, cnt = oracle.render_f32(oc, oview(oracle, v), v.semantics, nthreads=NT)
            assert np.array_equal(bits(b.image(int(k))), bits(ref)) and counts[int(k)] == cnt, int(k)
    finally:
        b.free()
        vol.free()


def test_resident_exact_batches_bitwise(spaces, oracle):
    """Batches whose blocks the GPU holds all at once take the single renders' launch shape: 8 lanes
    per ray when every image has at most 20,000 rays and those tiles fit, else 4 when every image
    has at most 60,000 rays, 4 samples in flight, and one plan dealt in a snake over the CUs. Six
    144x84 views (4 lanes: their 8-lane tiles do not fit), one 576x336 view (2 lanes, planned) and
    four 144x84 views over positive and signed cubes (8 lanes): every image bitwise equal to its
    single render, counts too, and two images to the oracle's FP32 render."""
    vols, ocs = spaces
    pos = [tuple(float(c) for c in p) for p in capi.synth_orbit(6, 23)]
    small = [W.view(144, 84, 0.1, pos=p) for p in pos]
    head = [W.view_576x336()]
    mixed = [W.view(144, 84, 0.25, pos=p) for p in pos[:4]]
    for views, which in ((small, [0] * 6), (head, [0]), (mixed, [0, 1, 0, 1])):
        b = capi.Batch.exact(vols, views, which)
        try:
            for rep in range(2):
                b.launch()
                b.wait()
            counts = b.samples()
            for k, v in enumerate(views):
                single, st = vols[which[k]].render(v)
                assert np.array_equal(bits(b.image(k)), bits(single)), (v.width, k)
                assert counts[k] == st.samples_in_cube, (v.width, k)
            assert sum(counts) > 0
            v = views[-1]
            ref, cnt = oracle.render_f32(ocs[which[-1]], oview(oracle, v), v.semantics, nthreads=NT)
            assert np.array_equal(bits(b.image(len(views) - 1)), bits(ref)) and counts[-1] == cnt
        finally:
            b.free()


def test_exact_batch_argument_checks(spaces):
    vols, _ = spaces
    v = capi.make_view(16, 16, 30.0, 30.0, (1.0, 2.0, 3.0), (10.0, 0.0, 0.0), 40.0, 8193, 1, capi.MATH_EXACT)
    with pytest.raises(capi.BirpError) as e:
        capi.Batch.exact(vols, [v], [0])
    assert e.value.status == capi.ERR_UNSUPPORTED  # the exact path's longest ray is 8192 samples
    v.samples = 100
    with pytest.raises(capi.BirpError) as e:
        capi.Batch.exact(vols, [v], [7])
    assert e.value.status == capi.ERR_INVALID_ARGUMENT
    b = capi.Batch.exact(vols, [v], [0])
    try:
        with pytest.raises(capi.BirpError) as e:
            b.set_tuning(capi.Tuning(8, 8, 1, 4))
        assert e.value.status == capi.ERR_UNSUPPORTED
    finally:
        b.free()


def test_server_lone_exact_requests_bitwise(device, spaces):
    """Exact requests sent one at a time launch alone (the view by value to the single-view exact
    kernel, nothing uploaded): over a positive cube, a signed cube, a nearest-sampled cube and a
    CMEM model, at sizes on both sides of the lanes-per-ray thresholds, each result equals
    birp_render of its view bit for bit, also into a pinned result buffer."""
    vols, _ = spaces
    views, which = mixed_views(12, seed=93)
    views += [W.view(160, 130, 0.25), W.view(300, 210, 0.25)]  # 20,800 and 63,000 rays
    which += [0, 0]
    srv = capi.Server(device, memory_budget=8 << 30, max_batch=8, max_inflight=2)
    try:
        idx = [srv.add_volume(v) for v in vols]
        for k, (v, w) in enumerate(zip(views, which)):
            print(f"lone exact request {k}: space {w}, {v.width}x{v.height} n={v.samples}", flush=True)
            out = capi.pinned_empty((v.height, v.width)) if k % 2 else None
            st, img, info = srv.wait(srv.submit(idx[w], v, out), 60000.0)
            assert st == capi.OK, (k, st)
            assert info.batch_size == 1
            want, _ = vols[w].render(v)
            assert np.array_equal(bits(img), bits(want)), (k, w)
    finally:
        srv.destroy()


def test_server_exact_requests_bitwise(device, spaces):
    """Served requests in BIRP_MATH_EXACT over a positive cube, a signed cube (general quad mode), a
    nearest-sampled cube and a CMEM model share slot launches (one exact launch per sample mode)
    with fast requests, and every exact result equals birp_render of its view bit for bit; a 16-bit
    cube refuses exact requests at submission."""
    vols, _ = spaces
    views, which = mixed_views(40, seed=77)
    maths = [capi.MATH_FAST if (k % 5 == 0 and which[k] != 2) else capi.MATH_EXACT for k in range(len(views))]
    for v, m in zip(views, maths):
        v.math = m
    vals128, org128, spc128, _ = W.cube_128()
    v16 = capi.Volume.upload(device, vals128, org128, spc128, dtype=capi.DTYPE_F16)
    srv = capi.Server(device, memory_budget=8 << 30, max_batch=24, max_inflight=2, batch_window_ms=1.0)
    try:
        idx = [srv.add_volume(v) for v in vols]
        i16 = srv.add_volume(v16)
        ids = [srv.submit(idx[w], v) for v, w in zip(views, which)]
        batched = 0
        for k, rid in enumerate(ids):
            st, img, info = srv.wait(rid, 60000.0)
            assert st == capi.OK, (k, st)
            batched += info.batch_size > 1
            want, _ = vols[which[k]].render(views[k])
            assert np.array_equal(bits(img), bits(want)), (k, which[k], maths[k])
        assert batched > 0
        bad = capi.make_view(64, 48, 20.0, 20.0, (0.0, -20.0, 10.0), (10.0, 0.0, 0.0), 40.0, 100, capi.SEM_EXTENDED, capi.MATH_EXACT)
        with pytest.raises(capi.BirpError) as e:
            srv.submit(i16, bad)
        assert e.value.status == capi.ERR_UNSUPPORTED
    finally:
        srv.destroy()
        v16.free()
