"""Benchmark of the MI355X BIRP renderer (one JSON line on stdout).

Headline (`value`): in-cube trilinear samples per second of the plain exact FP32 path on the
256^3 shell cube at 576x336, 800 samples per ray (step 0.1 RE over 80 RE): K back-to-back renders
through the C ABI with the cube resident in HBM, between two device synchronisations.

Also measured in the same run (keys under "forms"):
  plain_renders_per_s   the drop-in Camera::Render contract: launch, wait, copy the image to host
  batched               256 orbit poses, mixed 144x84/288x168/576x336 images, steps 0.1/0.25 RE,
                        F16 storage, one persistent-kernel launch (birp_batch_launch); plain
                        renders/s of the same 256 views for the ratio
  exact_batched         the same 256 views, and 8 copies of the headline view, as exact batches
                        (bit-identical images, one launch per sample mode) against back-to-back
                        single exact launches of the same views
  cmem                  the analytic CMEM model (default Init) on the same 576x336 x 800 grid:
                        samples/s of the exact path (VALU-bound) and its error against FP64
  precision             512^3 cube stored FP32/BF16/F16 (one or per-brick power-of-two scale),
                        576x336, step 0.05 RE: relative L2 and max relative error of every image
                        against the FP64 render
  serving               Poisson requests at 200/s over 8 resident cubes: p50/p99 latency, renders/s;
                        the same requests as one burst (saturated server) and through the plain path
The roofline is the VALU issue rate: the cycles in which the kernel's SIMDs issued a VALU
instruction (rocprofv3 PMC counts committed under profiles/) over the kernel time measured here with
HIP events, against every SIMD issuing every cycle. The HBM bytes the kernel moves (FETCH_SIZE and
WRITE_SIZE, copy-calibrated) and the algorithmic 8-corner gather bytes are reported beside it: the
cube stays in L2/MALL, so HBM is not the binding resource.
The CPU baseline leg is the only part that uses the oracle: its FP32 restatement of the reference
render (a port: the reference's own Camera.cpp cannot be compiled here) times whole renders of the
headline view on the host's cores and checks the GPU image bit for bit ("error": "vs_oracle_fp32"),
and it checks the CMEM image against the oracle's FP32 render with the zone-edge rule of
tests/parity.py. `--cpu-reference` times the reference's own DataCube sampler (oracle/_ref, which
is built and run only in the development container) in the render loop, single-threaded.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "Gsamples/s, renders/s vs plain GPU at 1e-4 rel err; p99 latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def max_rel(a, b, floor_frac=1e-3):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = np.abs(b) > floor_frac * np.abs(b).max()
    return float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m]))) if m.any() else 0.0


def read_profile_value(name, key):
    """A `key: value` figure from a committed rocprofv3 summary under profiles/ (None if absent)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


PEAK_VALU_ISSUE = 256 * 4 * 2.4e9  # SIMD cycles per second: 256 CUs x 4 SIMDs x 2.4 GHz


def valu_roofline(profile, kern_ms):
    """VALU issue roofline of one kernel: the cycles in which its SIMDs issued at least one VALU
    instruction, from a committed rocprofv3 --pmc summary (SQ_ACTIVE_INST_VALU counts one quad-cycle
    per wave64 VALU instruction, SQ_ACTIVE_INST_VALU2 the quad-cycles in which a SIMD issued two, so
    their difference is the quad-cycles with any VALU issue), over this run's measured kernel time,
    against every SIMD issuing every cycle at the 2.4 GHz peak clock. This fraction cannot exceed 1.
    `frac_cycles` is the same ratio over the dispatch's own GRBM_GUI_ACTIVE cycles in the profiled
    run: it does not depend on the clock the chip ran at (2.1 - 2.3 GHz under these loads)."""
    busy = read_profile_value(profile, "valu_issue_quad_cycles_per_launch")
    insts = read_profile_value(profile, "valu_wave_insts_per_launch")
    if busy is None or insts is None:
        return None
    achieved = busy * 4 / (kern_ms * 1e-3)
    return {"bound": "valu", "achieved": round(achieved / 1e9, 1), "peak": round(PEAK_VALU_ISSUE / 1e9, 1),
            "unit": "G SIMD-cycles/s with a VALU issue", "frac": round(achieved / PEAK_VALU_ISSUE, 3),
            "lane_instructions_per_s": float(f"{insts * 64 / (kern_ms * 1e-3):.4g}"),
            "dual_issue_frac": read_profile_value(profile, "valu_dual_issue_fraction"),
            "resident_waves_per_simd": read_profile_value(profile, "resident_waves_per_simd"),
            "frac_cycles": read_profile_value(profile, "valu_issue_occupancy_cycles"),
            "source": "profiles/" + profile}


def l1_pipeline(kern_ms):
    """The vector memory pipeline of the plain kernel from the committed rocprofv3 --pmc summary:
    the TD (data return) busy cycles per launch over this run's kernel time at 2.4 GHz on 256 CUs
    (`frac`, time-based like the VALU fraction), and the per-CU-cycle rates of the profiled run."""
    prof = "plain_exact_256cube_576x336_l1_pmc.txt"
    td = read_profile_value(prof, "td_busy_cycles_per_launch")
    if td is None:
        return None
    return {"unit": "TD busy cycles per CU-cycle", "frac": round(td / (256 * kern_ms * 1e-3 * 2.4e9), 3),
            "frac_cycles": read_profile_value(prof, "td_busy_per_cu_cycle"),
            "l1_lookups_per_cu_cycle": read_profile_value(prof, "tcp_lookups_per_cu_cycle"),
            "l1_lookups_per_record_load": read_profile_value(prof, "tcp_lookups_per_record_load"),
            "ta_busy_per_cu_cycle": read_profile_value(prof, "ta_busy_per_cu_cycle"),
            "source": "profiles/" + prof}


def read_traffic():
    """HBM bytes per launch of the plain kernel from the committed rocprofv3 --pmc summary."""
    return read_profile_value("plain_exact_256cube_576x336_pmc.txt", "traffic_bytes_per_launch")


def bench_plain(capi, W, dev, steps, warmup):
    vals, org, spc, _ = W.cube_256()
    vol = capi.Volume.upload(dev, vals, org, spc)
    v = W.view_576x336()
    img = capi.Image(dev, v.width, v.height)
    host, st = vol.render(v)
    n_in = st.samples_in_cube
    for _ in range(warmup):
        capi.render_async(vol, v, img)
    dev.synchronize()
    t0 = time.perf_counter()
    dev.timer_begin()
    for _ in range(steps):
        capi.render_async(vol, v, img)
    dev_ms = dev.timer_end()  # waits for the last render
    wall = time.perf_counter() - t0
    # the same renders under the other tile orders: the natural order without a plan, and a plan
    # made on every render (what a stream of new poses pays); the timed loop above reuses its plan
    orders = {}
    for name, pol in (("natural_order_ms", 0), ("plan_every_render_ms", 2)):
        capi.exact_order(pol)
        for _ in range(warmup):
            capi.render_async(vol, v, img)
        dev.timer_begin()
        for _ in range(steps):
            capi.render_async(vol, v, img)
        orders[name] = round(dev.timer_end() / steps, 5)
    capi.exact_order(-1)
    orders["plan_reused_ms"] = round(dev_ms / steps, 5)
    # the drop-in contract: one launch + wait + host copy per image, into an image that persists
    # across renders as Camera::image does
    k2 = max(10, steps)
    for _ in range(5):
        vol.render(v, out=host)
    t1 = time.perf_counter()
    for _ in range(k2):
        vol.render(v, out=host)
    sync_rps = k2 / (time.perf_counter() - t1)
    # the same into a birp_host_alloc (pinned) image, which the render fills by DMA
    pin = capi.pinned_empty((v.height, v.width))
    for _ in range(5):
        vol.render(v, out=pin)
    t1 = time.perf_counter()
    for _ in range(k2):
        vol.render(v, out=pin)
    sync_rps_pinned = k2 / (time.perf_counter() - t1)
    del pin
    img.free()
    return dict(vol=vol, vals=vals, org=org, spc=spc, view=v, n_in=n_in, wall=wall, dev_ms=dev_ms, sync_rps=sync_rps,
                sync_rps_pinned=sync_rps_pinned, image=host, orders=orders)


def cpu_baseline_leg(W, plain, cmem_img):
    """The CPU baseline and the oracle checks of this run's images (the only part of bench.py that
    uses the oracle): the oracle's FP32 restatement of Camera::Render times whole renders of the
    headline view on the host's cores and checks the GPU image bit for bit; the CMEM image is checked
    against the oracle's FP32 render, whole-image and with the pixels near the model's zone edges
    split off and held to their zone-flip bound (tests/parity.py)."""
    from oracle import oracle as O
    nthreads = min(16, os.cpu_count() or 1)
    v, n_in = plain["view"], plain["n_in"]
    oc = O.OracleCube(plain["vals"], plain["org"], plain["spc"])
    ov = O.make_view(v.width, v.height, v.fov_w, v.fov_h, list(v.position), list(v.aim), v.ray_length, v.samples)
    # whole renders, repeated until about a second of wall time (10-20 CPU-seconds over the threads)
    reps, el = 0, 0.0
    t = time.perf_counter()
    while reps < 2 or el < 1.0:
        ref, cnt = O.render_f32(oc, ov, O.SEM_EXTENDED, nthreads=nthreads)
        assert cnt == n_in
        reps += 1
        el = time.perf_counter() - t
    base = {"value": round(reps * cnt / el / 1e9, 6), "unit": "Gsamples/s", "cores": nthreads, "kind": "port",
            "sample": f"{reps} full 576x336 renders (n=800, {cnt} in-cube samples each) of the 256^3 cube by the oracle's "
                      f"FP32 restatement of Camera::Render, {nthreads} threads, {el:.2f} s ({el * nthreads:.0f} CPU-s)"}
    same = np.array_equal(ref.view(np.uint32), plain["image"].view(np.uint32))
    checks = {"vs_oracle_fp32": "bitwise" if same else float(f"{rel_l2(plain['image'], ref):.4e}")}
    if cmem_img is not None:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from parity import CMEM_EDGE_BAND
        m = O.OracleCmem(None)
        cv = W.view_576x336()
        ocv = O.make_view(cv.width, cv.height, cv.fov_w, cv.fov_h, list(cv.position), list(cv.aim), cv.ray_length, cv.samples)
        ref32, _ = O.render_f32(m, ocv, O.SEM_EXTENDED, nthreads=nthreads)
        mask, bound = m.edge_bound(ocv, CMEM_EDGE_BAND, nthreads=nthreads)
        img = cmem_img.astype(np.float64)
        d = np.abs(img - ref32)
        allow = bound * 1.001 + 1e-5 * np.abs(ref32) + 1e-30
        checks["cmem"] = {"vs_oracle_fp32_rel_l2": float(f"{rel_l2(img, ref32):.4e}"),
                          "edge_pixel_fraction": round(float(mask.mean()), 5),
                          "vs_oracle_fp32_rel_l2_off_edges": float(f"{rel_l2(img[~mask], ref32[~mask]):.4e}"),
                          "edge_pixels_within_flip_bound": bool(np.isfinite(img).all() and (d[mask] <= allow[mask]).all()),
                          "pixels_differing": int((d > 1e-5 * np.abs(ref32)).sum())}
    return base, checks


def cpu_reference_container():
    """The reference's own DataCube::GetSample (oracle/_ref, compiled from its sources in this
    container only; it never travels) inside the oracle's Camera::Render loop, single-threaded as
    the reference runs, on the whole 576x336 image of the 256^3 cube; the image is checked against
    the oracle's FP32 render bit for bit. Prints one JSON line."""
    from birp_cpp_amd import workloads as W
    from oracle import oracle as O
    if not os.path.exists(O.REF_PATH):
        print("oracle/_ref/libbirp_ref.so is not built here (it needs the reference sources)", file=sys.stderr)
        sys.exit(2)
    vals, org, spc, _ = W.cube_256()
    v = W.view_576x336()
    ov = O.make_view(v.width, v.height, v.fov_w, v.fov_h, list(v.position), list(v.aim), v.ray_length, v.samples)
    dirs = O.ray_dirs(ov)
    rc = O.RefCube(vals, org, spc)
    try:
        t = time.perf_counter()
        img, cnt = rc.render_rows(ov, dirs, 0, v.height, O.SEM_EXTENDED)
        el = time.perf_counter() - t
    finally:
        rc.free()
    oc = O.OracleCube(vals, org, spc)
    t = time.perf_counter()
    ref, cnt1 = O.render_f32(oc, ov, O.SEM_EXTENDED, nthreads=1)
    el1 = time.perf_counter() - t
    print(json.dumps({"reference_sampler": {"gsamples_per_s": round(cnt / el / 1e9, 6), "ns_per_sample": round(el / cnt * 1e9, 1),
                                            "seconds": round(el, 2), "cores": 1},
                      "oracle_port_1_thread": {"gsamples_per_s": round(cnt1 / el1 / 1e9, 6), "ns_per_sample": round(el1 / cnt1 * 1e9, 1),
                                               "seconds": round(el1, 2), "cores": 1},
                      "in_cube_samples": int(cnt), "bitwise_equal": bool(np.array_equal(img.view(np.uint32), ref.view(np.uint32))),
                      "host": os.uname().nodename, "nproc": os.cpu_count()}))


def bench_batched(capi, W, dev, vals, org, spc, fp32_vol):
    vol16 = capi.Volume.upload(dev, vals, org, spc, dtype=capi.DTYPE_F16)
    views = W.orbit_views_mixed(256)
    tune = os.path.join(ROOT, "birp_cpp_amd", "tuning.jsonl")
    if os.path.exists(tune):
        capi.lib().birp_autotune_load(tune.encode())
    b = capi.Batch(vol16, views)
    for _ in range(2):
        b.launch()
        b.wait()
    reps = 10
    t = time.perf_counter()
    for _ in range(reps):
        b.launch()
        b.wait()
    el = (time.perf_counter() - t) / reps
    nominal, last_ms = b.stats()
    # plain path on the same 256 views: one exact render, wait and host copy each
    ex = [capi.make_view(x.width, x.height, x.fov_w, x.fov_h, list(x.position), list(x.aim), x.ray_length, x.samples,
                         capi.SEM_EXTENDED, capi.MATH_EXACT) for x in views]
    bufs = _result_buffers(ex)
    for k in range(8):
        fp32_vol.render(ex[k], out=bufs[k])
    t = time.perf_counter()
    for k, x in enumerate(ex):
        fp32_vol.render(x, out=bufs[k])
    plain_el = time.perf_counter() - t
    # accuracy of the batch against the FP64 render (a sample of images)
    errs = [rel_l2(b.image(k), fp32_vol.render_f64(views[k])) for k in range(0, 256, 16)]
    # the same batch with its images written by the kernel into pinned host memory
    # (birp_batch_set_output): a finished launch has all 256 images on the host
    pixels = sum(x.width * x.height for x in views)
    arena = capi.pinned_empty((pixels,))
    b.set_output(arena)
    for _ in range(2):
        b.launch()
        b.wait()
    t = time.perf_counter()
    for _ in range(reps):
        b.launch()
        b.wait()
    el_host = (time.perf_counter() - t) / reps
    b.set_output(None)
    # in-cube samples of the 256 views (the exact path's count; the fast path's FP64 entry can
    # differ by a sample at a face) and the batch's algorithmic bytes: 16 B per F16 sample + 4 B/px
    in_cube = sum(fp32_vol.count_samples(x) for x in ex)
    alg = 16.0 * in_cube + 4.0 * pixels
    b.free()
    del arena
    vol16.free()
    return {"poses": 256, "storage": "f16", "renders_per_s": round(256 / el, 1), "ms_per_batch": round(el * 1e3, 3),
            "renders_per_s_images_in_host_memory": round(256 / el_host, 1),
            "device_ms_per_batch": round(last_ms, 3), "nominal_gsamples_per_s": round(nominal / el / 1e9, 2),
            "plain_renders_per_s_same_views": round(256 / plain_el, 1), "speedup_vs_plain": round(plain_el / el, 2),
            "max_rel_l2_vs_fp64": float(f"{max(errs):.3e}"), "in_cube_samples": int(in_cube),
            "in_cube_gsamples_per_s": round(in_cube / el / 1e9, 2),
            "algorithmic_gb_per_s_device": round(alg / (last_ms * 1e-3) / 1e9, 1),
            "roofline": valu_roofline("batched_f16_256cube_orbit256_valu_pmc.txt", last_ms)}


def bench_exact_batched(capi, W, dev, fp32_vol, view0, reps=10):
    """Exact renders of many views in one launch per sample mode (birp_batch_create_exact), each image
    bit-identical to its single exact render: the 256 orbit views of the batched workload, and 8
    copies of the headline view; against the same views as back-to-back single exact launches on
    the device (birp_render_async into device images)."""
    out = {}
    for name, views in (("orbit256", W.orbit_views_mixed(256, math=capi.MATH_EXACT)), ("headline_x8", [view0] * 8)):
        b = capi.Batch.exact([fp32_vol], views)
        for _ in range(2):
            b.launch()
            b.wait()
        t = time.perf_counter()
        dms = []
        for _ in range(reps):
            b.launch()
            b.wait()
            dms.append(b.stats()[1])
        el = (time.perf_counter() - t) / reps
        in_cube = sum(b.samples())
        b.free()
        imgs = [capi.Image(dev, v.width, v.height) for v in views]
        for v, img in zip(views, imgs):
            capi.render_async(fp32_vol, v, img)
        dev.synchronize()
        dev.timer_begin()
        for _ in range(reps):
            for v, img in zip(views, imgs):
                capi.render_async(fp32_vol, v, img)
        single_ms = dev.timer_end() / reps
        for img in imgs:
            img.free()
        dev_ms = float(np.median(dms))
        out[name] = {"views": len(views), "device_ms_per_batch": round(dev_ms, 4), "renders_per_s": round(len(views) / el, 1),
                     "single_launches_device_ms": round(single_ms, 4), "speedup_vs_single_launches": round(single_ms / dev_ms, 3),
                     "in_cube_samples": int(in_cube), "in_cube_gsamples_per_s_device": round(in_cube / (dev_ms * 1e-3) / 1e9, 1)}
    prof = valu_roofline("exact_batched_256cube_orbit256_valu_pmc.txt", out["orbit256"]["device_ms_per_batch"])
    if prof:
        out["orbit256"]["roofline"] = prof
    return out


def bench_cmem(capi, W, dev, steps=20):
    """The analytic CMEM model (default Init) rendered by the exact path over the 576x336 image
    and ray grid: every sample is evaluated (the model has no bounds), about 25 transcendentals
    each, so this is VALU-bound."""
    vol = capi.Volume.cmem(dev, capi.cmem_init())
    v = W.view_576x336()
    img = capi.Image(dev, v.width, v.height)
    host, st = vol.render(v)
    ref64 = vol.render_f64(v)
    for _ in range(3):
        capi.render_async(vol, v, img)
    dev.synchronize()
    t0 = time.perf_counter()
    dev.timer_begin()
    for _ in range(steps):
        capi.render_async(vol, v, img)
    dev_ms = dev.timer_end() / steps
    wall = (time.perf_counter() - t0) / steps
    # the fast form of the model (hardware exp2/log2, polynomial arc cosines, Kahan sums)
    vf = W.view_576x336(math=capi.MATH_FAST)
    fast, _ = vol.render(vf)
    for _ in range(3):
        capi.render_async(vol, vf, img)
    dev.synchronize()
    t0 = time.perf_counter()
    dev.timer_begin()
    for _ in range(steps):
        capi.render_async(vol, vf, img)
    fast_ms = dev.timer_end() / steps
    fast_wall = (time.perf_counter() - t0) / steps
    img.free()
    vol.free()
    nsamp = v.width * v.height * v.samples
    return {"image": f"{v.width}x{v.height}", "samples_per_ray": v.samples, "samples": nsamp,
            "kernel_ms": round(dev_ms, 4), "gsamples_per_s": round(nsamp / wall / 1e9, 2),
            "rel_l2_vs_fp64": float(f"{rel_l2(host, ref64):.4e}"),
            "roofline": valu_roofline("cmem_exact_576x336_valu_pmc.txt", dev_ms),
            "fast": {"kernel_ms": round(fast_ms, 4), "gsamples_per_s": round(nsamp / fast_wall / 1e9, 2),
                     "rel_l2_vs_fp64": float(f"{rel_l2(fast, ref64):.4e}")}}, host


def bench_precision(capi, W, dev, n=512, count=64):
    vals, org, spc, _ = W.cube_512_wide_range(n)
    vols = {name: capi.Volume.upload(dev, vals, org, spc, dtype=dt)
            for name, dt in (("fp32", capi.DTYPE_F32), ("bf16", capi.DTYPE_BF16), ("f16", capi.DTYPE_F16))}
    del vals
    views = W.orbit_views_576x336_n1600(count)
    out = {"cube": f"{n}^3", "poses": count, "image": "576x336", "samples_per_ray": views[0].samples}
    errs = {k: [] for k in ("exact_fp32", "fast_fp32", "bf16", "f16")}
    mx = {k: [] for k in errs}
    for v in views:
        ref = vols["fp32"].render_f64(v)
        if not ref.any():
            continue
        ex = capi.make_view(v.width, v.height, v.fov_w, v.fov_h, list(v.position), list(v.aim), v.ray_length, v.samples,
                            capi.SEM_EXTENDED, capi.MATH_EXACT)
        for key, vol, view in (("exact_fp32", vols["fp32"], ex), ("fast_fp32", vols["fp32"], v), ("bf16", vols["bf16"], v),
                               ("f16", vols["f16"], v)):
            img, _ = vol.render(view)
            errs[key].append(rel_l2(img, ref))
            mx[key].append(max_rel(img, ref))
    for k in errs:
        out[k] = {"rel_l2_max": float(f"{max(errs[k]):.3e}"), "rel_l2_mean": float(f"{np.mean(errs[k]):.3e}"),
                  "max_rel_err": float(f"{max(mx[k]):.3e}")}
    # batched renders/s of the 64 poses per storage type
    for name in ("fp32", "bf16", "f16"):
        b = capi.Batch(vols[name], views)
        b.launch()
        b.wait()
        t = time.perf_counter()
        for _ in range(3):
            b.launch()
            b.wait()
        el = (time.perf_counter() - t) / 3
        out[name + "_batched_renders_per_s"] = round(count / el, 1)
        b.free()
    out["bytes"] = {k: v.nbytes() for k, v in vols.items()}
    for v in vols.values():
        v.free()
    return out


def _result_buffers(views):
    """One written-to (so already paged-in) result image per request, as a serving client that
    reuses its response buffers holds them."""
    return [np.ones((v.height, v.width), np.float32) for v in views]


def _serve(capi, srv, idx, which, views, t_arr, bufs):
    """Submit requests at their arrival offsets (all at once when t_arr is None); wait for all."""
    ids = []
    start = capi.clock_s() + 0.05
    for k in range(len(views)):
        if t_arr is not None:
            target = start + t_arr[k]
            while capi.clock_s() < target:
                pass
        ids.append(srv.submit(idx[which[k]], views[k], out=bufs[k]))
    submitted = capi.clock_s()
    lat, done_max, fails, bsz = [], 0.0, 0, []
    for rid in ids:
        st, _, info = srv.wait(rid, 60000.0)
        if st != capi.OK:
            fails += 1
            continue
        lat.append(info.done_s - info.submit_s)
        done_max = max(done_max, info.done_s)
        bsz.append(info.batch_size)
    lat = np.array(lat) * 1e3
    first = start if t_arr is not None else start - 0.05
    return {"requests": len(views), "failed": fails, "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "renders_per_s": round(len(lat) / (done_max - first), 1),
            "mean_batch": round(float(np.mean(bsz)), 2), "submit_span_ms": round((submitted - first) * 1e3, 2)}


def bench_serving(capi, W, dev, num=2000, rate=200.0):
    """Config-5 serving: 8 resident cubes, Poisson arrivals at `rate`, then the same requests as
    one burst (the server saturated; timed after one untimed burst), both again with pinned
    response buffers (birp_host_alloc), the FP32 cubes' requests in the exact arithmetic as one
    burst, then the same requests through the plain path (one birp_render and host sync per
    request) for the renders/s comparisons."""
    cubes = []
    for k, (n, dt) in enumerate(W.serving_cubes()):
        vals, org, spc, _ = capi.synth_shell(n, (7.0, 12.0), 0.1, seed=500 + k)
        cubes.append(capi.Volume.upload(dev, vals, org, spc, dtype=dt))
        del vals
    t_arr, which, views = W.serving_requests(num, rate)
    # six launches in flight; each slot reserves a full batch of the largest image at creation,
    # so no allocation happens while serving. The budget holds the cubes and those six buffers.
    max_px = max(v.width * v.height for v in views)
    budget = sum(c.nbytes() for c in cubes) + 6 * 64 * max_px * 4
    srv = capi.Server(dev, memory_budget=budget, max_batch=64, max_inflight=6, batch_window_ms=0.2, reserve_pixels=64 * max_px)
    idx = [srv.add_volume(c) for c in cubes]
    bufs = _result_buffers(views)
    out = {"rate_hz": rate, "budget_bytes": budget}
    out["poisson"] = _serve(capi, srv, idx, which, views, t_arr, bufs)
    # sustained load: one untimed burst first (as the main line's warm-up steps), since the
    # first burst after the light Poisson load still runs at ramping clocks
    _serve(capi, srv, idx, which, views, None, bufs)
    out["saturated"] = _serve(capi, srv, idx, which, views, None, bufs)
    # the same with response buffers from birp_host_alloc: the kernel writes each image straight
    # into its buffer (no device image, DMA or host copy)
    pinned = [capi.pinned_empty((v.height, v.width)) for v in views]
    for b in pinned:
        b.fill(1.0)
    out["poisson_pinned_results"] = _serve(capi, srv, idx, which, views, t_arr, pinned)
    _serve(capi, srv, idx, which, views, None, pinned)
    out["saturated_pinned_results"] = _serve(capi, srv, idx, which, views, None, pinned)
    # the requests on the FP32 cubes again in BIRP_MATH_EXACT: the reference's own images, bit for
    # bit, served through the multi-view exact kernel; against the same requests as single exact renders
    dts = [dt for _, dt in W.serving_cubes()]
    ex = [k for k in range(num) if dts[which[k]] == capi.DTYPE_F32]
    ex_views = [capi.make_view(v.width, v.height, v.fov_w, v.fov_h, list(v.position), list(v.aim), v.ray_length, v.samples,
                               v.semantics, capi.MATH_EXACT) for v in (views[k] for k in ex)]
    ex_which = [which[k] for k in ex]
    ex_bufs = [bufs[k] for k in ex]
    _serve(capi, srv, idx, ex_which, ex_views, None, ex_bufs)
    out["saturated_exact_fp32_cubes"] = _serve(capi, srv, idx, ex_which, ex_views, None, ex_bufs)
    srv.destroy()
    del pinned
    t0 = time.perf_counter()
    for w, v, b in zip(ex_which, ex_views, ex_bufs):
        cubes[w].render(v, out=b)
    ex_plain = len(ex_views) / (time.perf_counter() - t0)
    out["saturated_exact_fp32_cubes"]["plain_renders_per_s_same_requests"] = round(ex_plain, 1)
    out["saturated_exact_fp32_cubes"]["speedup_vs_plain"] = round(out["saturated_exact_fp32_cubes"]["renders_per_s"] / ex_plain, 2)
    t0 = time.perf_counter()
    for k in range(num):
        cubes[which[k]].render(views[k], out=bufs[k])
    plain_rps = num / (time.perf_counter() - t0)
    for c in cubes:
        c.free()
    out["plain_renders_per_s_same_requests"] = round(plain_rps, 1)
    out["saturated_speedup_vs_plain"] = round(out["saturated"]["renders_per_s"] / plain_rps, 2)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--forms", default="batched,exact_batched,cmem,precision,serving", help="extra forms to measure ('' for none)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle checks")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (torch) to calibrate rocprofv3 FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--cpu-reference", action="store_true",
                    help="time the reference's own sampler (oracle/_ref, built in the development container only)")
    args = ap.parse_args()
    if args.cpu_reference:
        cpu_reference_container()
        return
    if args.copy_calibration:
        import torch
        n = 1 << 28  # 1 GiB of float32
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: there is no multi-GPU path (--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    from birp_cpp_amd import capi
    from birp_cpp_amd import workloads as W

    dev = capi.Device(0)
    log("plain: 256^3 cube, 576x336 exact renders")
    p = bench_plain(capi, W, dev, args.steps, args.warmup)
    v, n_in = p["view"], p["n_in"]
    ms_step = p["wall"] * 1e3 / args.steps
    kern_ms = p["dev_ms"] / args.steps
    alg_bytes = 32.0 * n_in + 4.0 * v.width * v.height
    traffic = read_traffic()
    roof = valu_roofline("plain_exact_256cube_576x336_valu_pmc.txt", kern_ms) or {"bound": "valu", "frac": None}
    roof["traffic"] = traffic
    roof["kernel_ms"] = round(kern_ms, 5)
    roof["hbm"] = {"achieved": round(traffic / (kern_ms * 1e-3) / 1e9, 1) if traffic else None, "peak": PEAK_HBM_GBS,
                   "unit": "GB/s", "frac": round(traffic / (kern_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4) if traffic else None,
                   "source": "profiles/plain_exact_256cube_576x336_pmc.txt"}
    roof["gathers"] = {"algorithmic_bytes": int(alg_bytes), "achieved": round(alg_bytes / (kern_ms * 1e-3) / 1e9, 1),
                       "unit": "GB/s", "note": "8 FP32 corners per in-cube sample + 4 B per pixel, served mostly by L1/L2"}
    roof["l1"] = l1_pipeline(kern_ms)
    line = {
        "metric": METRIC, "value": round(n_in * args.steps / p["wall"] / 1e9, 3), "unit": "Gsamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain exact FP32 render, 256^3 shell cube (seed 2), 576x336 px, 800 samples/ray "
                               "(step 0.1 RE, L 80 RE), camera (0,-20,10) aim (10,0,0), extended semantics",
                   "in_cube_samples": n_in, "nominal_samples": v.width * v.height * v.samples},
        "roofline": roof,
        "forms": {"plain_renders_per_s": round(p["sync_rps"], 1),
                  "plain_renders_per_s_pinned_image": round(p["sync_rps_pinned"], 1),
                  "plain_renders_per_s_async": round(args.steps / p["wall"], 1),
                  "plain_tile_order": p["orders"]},
    }
    # every result carries its error against the FP64 render of the same sample grid
    ref64 = p["vol"].render_f64(v)
    line["error"] = {"vs_fp64_rel_l2": float(f"{rel_l2(p['image'], ref64):.4e}"),
                     "vs_fp64_max_rel": float(f"{max_rel(p['image'], ref64):.4e}")}
    forms = [f for f in args.forms.split(",") if f]
    if "batched" in forms:
        log("batched: 256 orbit poses, F16")
        line["forms"]["batched"] = bench_batched(capi, W, dev, p["vals"], p["org"], p["spc"], p["vol"])
    if "exact_batched" in forms:
        log("exact batched: 256 orbit poses and 8 headline views, one launch each")
        line["forms"]["exact_batched"] = bench_exact_batched(capi, W, dev, p["vol"], v)
    p["vol"].free()
    cmem_img = None
    if "cmem" in forms:
        log("cmem: analytic model, 576x336 x 800 samples")
        line["forms"]["cmem"], cmem_img = bench_cmem(capi, W, dev)
    if "precision" in forms:
        log("precision: 512^3 FP32/BF16/F16")
        line["forms"]["precision"] = bench_precision(capi, W, dev)
    if "serving" in forms:
        log("serving: Poisson 200 req/s")
        line["forms"]["serving"] = bench_serving(capi, W, dev)
    dev.close()
    if not args.no_cpu:
        log("cpu baseline: oracle render, oracle checks of the plain and CMEM images")
        line["cpu_baseline"], checks = cpu_baseline_leg(W, p, cmem_img)
        line["error"]["vs_oracle_fp32"] = checks["vs_oracle_fp32"]
        if "cmem" in checks:
            line["forms"]["cmem"]["vs_oracle"] = checks["cmem"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
