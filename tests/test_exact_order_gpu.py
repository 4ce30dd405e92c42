"""Tile order of single exact renders. A plan (k_exact_plan) sorts a view's tiles by their
estimated trip counts, heaviest first, and the render deals them over the CUs; the natural
order (block b renders tile b) needs no plan. Checked here: every plan is a permutation of the view's tiles, ranked by the
oracle's own per-tile work (the longest in-cube run of samples of each wave, read back from the
oracle's FP32 render of a cube of ones); images and in-cube counts are bit for bit the same under
the natural order, a reused plan and a plan made on every render, through birp_render and
birp_render_async; and a reused plan follows changes of pose, image size, cube box and semantics.
"""
import os

import numpy as np
import pytest
from scipy.stats import kendalltau

from birp_cpp_amd import capi
from birp_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)
TW, TH = 4, 8  # wave tile of the exact kernel at two lanes per ray (block: 2x2 waves)


def oview(O, v):
    return O.make_view(v.width, v.height, v.fov_w, v.fov_h, list(v.position), list(v.aim), v.ray_length, v.samples)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(autouse=True)
def default_order():
    yield
    capi.exact_order(-1)


@pytest.fixture(scope="module")
def cubes(device):
    vals, org, spc, _ = W.cube_256()
    rng = np.random.default_rng(23)
    n = 48
    sorg, sspc = (-4.0, -9.0, -7.0), (0.4, 0.35, 0.3)
    signed = (rng.standard_normal((n, n, n)) * 1e-5 + 3e-6).astype(np.float32)
    vols = {
        "shell": capi.Volume.upload(device, vals, org, spc),
        "signed": capi.Volume.upload(device, signed, sorg, sspc),
        "nearest": capi.Volume.upload(device, signed, sorg, sspc, interp=capi.INTERP_NEAREST),
    }
    geo = {"shell": (vals.shape, org, spc), "signed": (signed.shape, sorg, sspc), "nearest": (signed.shape, sorg, sspc)}
    yield vols, geo
    for v in vols.values():
        v.free()


def ray_lengths(O, shape, org, spc, v, semantics):
    """In-cube samples of every ray, from the oracle's FP32 render of a cube of ones: a ray of m
    samples sums m copies of the scaled unit sample, in order, and those partial sums increase."""
    ones = np.ones(shape, np.float32)
    oc = O.OracleCube(ones, org, spc)
    img, cnt = O.render_f32(oc, oview(O, v), semantics, nthreads=NT)
    _, w = O.ray_dist(v.ray_length, v.samples)
    s1 = np.float32(O.scale_sample(1.0, w))
    partial = np.zeros(v.samples + 1, np.float32)
    acc = np.float32(0)
    for m in range(1, v.samples + 1):
        acc = np.float32(acc + s1)
        partial[m] = acc
    assert np.all(np.diff(partial) > 0)
    lens = np.searchsorted(partial, img.ravel()).reshape(img.shape)
    assert np.array_equal(partial[lens], img), "image is not a run of unit samples"
    assert lens.sum() == cnt
    return lens


def tile_work(lens):
    """Per block of 2x2 wave tiles: the sum over its waves of the wave's longest ray."""
    H, Wd = lens.shape
    wg_x, wg_y = (Wd + 2 * TW - 1) // (2 * TW), (H + 2 * TH - 1) // (2 * TH)
    pad = np.zeros((wg_y * 2 * TH, wg_x * 2 * TW), np.int64)
    pad[:H, :Wd] = lens
    waves = pad.reshape(wg_y * 2, TH, wg_x * 2, TW).max(axis=(1, 3))
    return waves.reshape(wg_y, 2, wg_x, 2).sum(axis=(1, 3)).ravel()


PLAN_VIEWS = [
    ("headline", lambda: W.view_576x336(), capi.SEM_EXTENDED),
    ("ragged", lambda: W.view(601, 333, 0.1, pos=(-3.0, -24.0, 6.0), aim=(8.0, 2.0, -1.0)), capi.SEM_EXTENDED),
    ("far", lambda: W.view(576, 336, 0.1, pos=(5.0, -55.0, 20.0), aim=(5.0, 0.0, 0.0)), capi.SEM_EXTENDED),
    ("inside", lambda: W.view(512, 288, 0.1, pos=(4.0, 1.0, -2.0), aim=(18.0, 6.0, 3.0)), capi.SEM_REFERENCE),
    ("large", lambda: W.view(1600, 900, 0.25, pos=(2.0, -26.0, 8.0), aim=(6.0, 0.0, 1.0)), capi.SEM_EXTENDED),
]


@pytest.mark.parametrize("name,mk,sem", PLAN_VIEWS, ids=[p[0] for p in PLAN_VIEWS])
def test_plan_is_a_heavy_first_permutation(cubes, oracle, name, mk, sem):
    vols, geo = cubes
    v = mk()
    v.semantics = sem
    print(f"plan view {name}: {v.width}x{v.height} n={v.samples}", flush=True)
    order = capi.exact_plan(vols["shell"], v)
    work = tile_work(ray_lengths(oracle, *geo["shell"], v, sem))
    assert order.shape == work.shape
    assert np.array_equal(np.sort(order), np.arange(work.size)), "plan is not a permutation of the tiles"
    rank = np.empty(work.size)
    rank[order] = np.arange(work.size)
    tau = kendalltau(rank, work)[0]
    print(f"  tiles {work.size}, work {work.min()}..{work.max()}, tau(rank, work) {tau:.4f}", flush=True)
    assert tau <= -0.97, tau
    # the plan's first quarter carries (nearly) the work of the heaviest quarter of the tiles
    q = work.size // 4
    assert work[order[:q]].mean() >= 0.995 * np.sort(work)[::-1][:q].mean()


def test_plan_of_dead_rays_is_still_a_permutation(cubes):
    vols, _ = cubes
    v = W.view_576x336()
    v.semantics = capi.SEM_REFERENCE  # the camera is outside: every ray's first sample is outside
    order = capi.exact_plan(vols["shell"], v)
    assert np.array_equal(np.sort(order), np.arange(order.size))


def render_all_orders(vol, v):
    """Policy 0 (natural order), 1 (a view's first render in the natural order, then planned once it
    comes back: the second render is the planned one) and 2 (planned on every render)."""
    out = {}
    for pol in (0, 1, 1, 2):
        capi.exact_order(pol)
        img, st = vol.render(v)
        out[pol] = (img, st.samples_in_cube)
    return out


ORDER_CASES = [
    ("shell-headline", "shell", lambda: W.view_576x336(), capi.SEM_EXTENDED),
    ("shell-ragged", "shell", lambda: W.view(601, 333, 0.1, pos=(-3.0, -24.0, 6.0), aim=(8.0, 2.0, -1.0)), capi.SEM_EXTENDED),
    ("shell-inside-ref", "shell", lambda: W.view(512, 288, 0.1, pos=(4.0, 1.0, -2.0), aim=(18.0, 6.0, 3.0)), capi.SEM_REFERENCE),
    ("signed", "signed", lambda: W.view(400, 300, 0.05, pos=(6.0, -14.0, 4.0), aim=(5.0, -1.0, 0.0)), capi.SEM_EXTENDED),
    ("nearest", "nearest", lambda: W.view(400, 300, 0.05, pos=(6.0, -14.0, 4.0), aim=(5.0, -1.0, 0.0)), capi.SEM_EXTENDED),
    # small images: 16 lanes per ray (72x42, 756 tiles), 8 (144x84, 378 tiles) and 4 (240x160, 600
    # tiles), all planned
    ("shell-tiny-16-lanes", "shell", lambda: W.view(72, 42, 0.1, pos=(2.0, -17.0, 8.0)), capi.SEM_EXTENDED),
    ("signed-tiny-16-lanes", "signed", lambda: W.view(90, 50, 0.05, pos=(6.0, -14.0, 4.0), aim=(5.0, -1.0, 0.0)), capi.SEM_REFERENCE),
    ("shell-small-8-lanes", "shell", lambda: W.view(144, 84, 0.1, pos=(3.0, -18.0, 9.0)), capi.SEM_EXTENDED),
    ("signed-small-4-lanes", "signed", lambda: W.view(240, 160, 0.1, pos=(6.0, -14.0, 4.0), aim=(5.0, -1.0, 0.0)), capi.SEM_EXTENDED),
    # more tiles (18,432) than a plan holds: every policy renders the natural order
    ("shell-beyond-plans", "shell", lambda: W.view(2048, 1152, 0.5, pos=(1.0, -22.0, 9.0), aim=(9.0, 0.0, 0.0)), capi.SEM_EXTENDED),
]


@pytest.mark.parametrize("name,key,mk,sem", ORDER_CASES, ids=[c[0] for c in ORDER_CASES])
def test_orders_render_the_same_bits(cubes, oracle, name, key, mk, sem):
    vols, geo = cubes
    v = mk()
    v.semantics = sem
    print(f"order case {name}: {v.width}x{v.height} n={v.samples}", flush=True)
    res = render_all_orders(vols[key], v)
    base, cnt = res[0]
    for pol in (1, 2):
        assert np.array_equal(bits(res[pol][0]), bits(base)), f"policy {pol}"
        assert res[pol][1] == cnt, f"policy {pol}: {res[pol][1]} != {cnt}"
    if name in ("shell-ragged", "shell-beyond-plans", "shell-small-8-lanes", "shell-tiny-16-lanes"):  # and against the oracle, bit for bit
        vals, org, spc, _ = W.cube_256()
        ref, rcnt = oracle.render_f32(oracle.OracleCube(vals, org, spc), oview(oracle, v), sem, nthreads=NT)
        assert np.array_equal(bits(base), bits(ref))
        assert rcnt == cnt


def test_async_renders_follow_the_plan_cache(device, cubes):
    """Device images through birp_render_async, with plans made when a view comes back, reused,
    replaced by a new pose, a new image size, another cube and other semantics, alternating
    between views held in different plan slots, and reused again; each equals its natural-order
    render, and every render starts from an image full of NaNs so a tile left out would show."""
    vols, _ = cubes
    a = W.view_576x336()
    b = W.view(576, 336, 0.1, pos=(-2.0, -22.0, 12.0), aim=(9.0, 1.0, 0.0))
    c = W.view(640, 400, 0.1)
    d = W.view(576, 336, 0.1)
    d.semantics = capi.SEM_REFERENCE
    seq = [("shell", a), ("shell", a), ("shell", b), ("shell", a), ("shell", c), ("signed", a), ("shell", d), ("shell", a),
           ("shell", b), ("shell", c), ("shell", b), ("shell", c), ("signed", a), ("signed", a), ("shell", d), ("shell", d),
           ("shell", a), ("shell", b)]
    capi.exact_order(0)
    want = {}
    for key, v in seq:
        img = capi.Image(device, v.width, v.height)
        capi.render_async(vols[key], v, img)
        want[(key, id(v))] = img.download()
        img.free()
    capi.exact_order(1)
    nan = np.full((400, 640), np.nan, np.float32)
    for k, (key, v) in enumerate(seq):
        img = capi.Image(device, v.width, v.height)
        img.upload(nan[: v.height, : v.width])
        capi.render_async(vols[key], v, img)
        got = img.download()
        img.free()
        assert np.array_equal(bits(got), bits(want[(key, id(v))])), f"render {k} ({key}, {v.width}x{v.height})"


def test_batch_plans_render_the_same_bits(device, cubes):
    """Exact batches made with per-view tile plans (the default) and in the natural order: every
    image and in-cube count is the same, over two cubes, several image sizes and a repeated view."""
    vols, _ = cubes
    pos = capi.synth_orbit(24, 11)
    views = [W.view(*W.SIZES[k % 3], 0.1 if k % 2 else 0.25, pos=tuple(float(c) for c in p)) for k, p in enumerate(pos)]
    views += [W.view_576x336()] * 3
    which = [k % 2 for k in range(len(views))]
    got = {}
    for pol in (0, 1):
        capi.exact_order(pol)
        b = capi.Batch.exact([vols["shell"], vols["signed"]], views, which)
        b.launch()
        b.wait()
        got[pol] = ([b.image(k) for k in range(len(views))], b.samples())
        b.free()
    for k in range(len(views)):
        assert np.array_equal(bits(got[0][0][k]), bits(got[1][0][k])), f"view {k}"
    assert list(got[0][1]) == list(got[1][1])
