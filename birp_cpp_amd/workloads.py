"""The benchmark and test workloads: seeded cubes, camera poses and ray grids.

Geometry shared by all workloads: camera at (0,-20,10) RE aimed at (10,0,0), FOV 15.5 deg
(vertical) x 26.5 deg (horizontal), rays of length 80 RE; cubes span x in [-10,20],
y,z in [-15,15] RE. The reference's own camera (Camera.cpp:29-47) fixes L=44, n=200 and derives
W x H from a pixel size; here the size and the step are explicit (576x336 has no pixel size under
the reference's rounding rule, Camera.cpp:38-39).
"""
from __future__ import annotations

import numpy as np

from . import capi

FOV_H, FOV_W = 15.5, 26.5
CAM_POS = (0.0, -20.0, 10.0)
AIM = (10.0, 0.0, 0.0)
RAY_LENGTH = 80.0
SIZES = [(144, 84), (288, 168), (576, 336)]


def samples_for(step: float, length: float = RAY_LENGTH) -> int:
    return int(round(length / step))


def view(width, height, step, pos=CAM_POS, aim=AIM, semantics=capi.SEM_EXTENDED, math=capi.MATH_EXACT, length=RAY_LENGTH):
    return capi.make_view(width, height, FOV_W, FOV_H, pos, aim, length, samples_for(step, length), semantics, math)


def cube_128():
    """128^3 FP32 shell cube, r0 ~ U(8,11), 5% lognormal noise, seed 1."""
    return capi.synth_shell(128, (8.0, 11.0), 0.05, seed=1)


def cube_256():
    """256^3 FP32 shell cube, r0 ~ U(7,12), sigma 0.1, seed 2."""
    return capi.synth_shell(256, (7.0, 12.0), 0.1, seed=2)


def cube_512_wide_range(n=512):
    """512^3 shell cube with dynamic range ~1e-8..1e-4: background 1e-8, shell amplitude 3e-5, seed 4."""
    return capi.synth_shell(n, (7.0, 12.0), 0.1, seed=4, amp=3e-5, bg=1e-8)


def view_144x84(math=capi.MATH_EXACT):
    return view(144, 84, 0.5, math=math)


def view_576x336(math=capi.MATH_EXACT):
    return view(576, 336, 0.1, math=math)


def orbit_views_mixed(count=256, seed=3, math=capi.MATH_FAST):
    """Orbit poses (perigee 1.8, apogee 19 RE, inclination 70 deg) with image sizes drawn from
    SIZES and steps from {0.1, 0.25} RE, all from `seed`."""
    pos = capi.synth_orbit(count, seed)
    rng = np.random.default_rng(seed)
    sizes = rng.integers(0, len(SIZES), count)
    steps = rng.choice([0.1, 0.25], count)
    return [view(*SIZES[s], st, pos=tuple(float(c) for c in p), math=math) for p, s, st in zip(pos, sizes, steps)]


def orbit_views_576x336_n1600(count=64, seed=4, math=capi.MATH_FAST):
    pos = capi.synth_orbit(count, seed)
    return [view(576, 336, 0.05, pos=tuple(float(c) for c in p), math=math) for p in pos]


def serving_cubes(seed=5, count=8):
    """(n, dtype) of the serving workload's resident cubes: 8 cubes of 128^3..512^3, FP32 or F16."""
    rng = np.random.default_rng(seed)
    ns = rng.choice([128, 256, 384, 512], count)
    dts = rng.choice([capi.DTYPE_F32, capi.DTYPE_F16], count)
    return [(int(n), int(d)) for n, d in zip(ns, dts)]


def serving_requests(num, rate_hz=200.0, seed=5, ncubes=8):
    """Poisson arrivals: (arrival offsets [s], cube index, view) per request."""
    rng = np.random.default_rng(seed + 1000)
    gaps = rng.exponential(1.0 / rate_hz, num)
    t = np.cumsum(gaps) - gaps[0]
    cubes = rng.integers(0, ncubes, num)
    pos = capi.synth_orbit(num, seed + 2000)
    sizes = rng.integers(0, len(SIZES), num)
    steps = rng.choice([0.1, 0.25, 0.5], num)
    views = [view(*SIZES[s], st, pos=tuple(float(c) for c in p), math=capi.MATH_FAST) for p, s, st in zip(pos, sizes, steps)]
    return t, cubes, views
