"""ctypes bindings to libbirp_hip.so, the C ABI declared in include/birp.h.

This module is glue for Python callers (tests, bench): every render goes through the C ABI into
the HIP kernels. It never falls back to anything else: if the library is missing it raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libbirp_hip.so")

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_SINGULAR = 2
ERR_OUT_OF_MEMORY = 3
ERR_DEVICE = 4
ERR_NO_DEVICE = 5
ERR_ADMISSION = 6
ERR_NOT_FOUND = 7
ERR_TIMEOUT = 8
ERR_INJECTED = 9
ERR_UNSUPPORTED = 10

DTYPE_F32, DTYPE_BF16, DTYPE_F16 = 0, 1, 2
INTERP_NEAREST, INTERP_TRILINEAR = 0, 1
SEM_REFERENCE, SEM_EXTENDED = 0, 1
MATH_EXACT, MATH_FAST = 0, 1
RENDER_ACCUMULATE = 1

_f, _d, _i = C.c_float, C.c_double, C.c_int


class VolumeDesc(C.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i), ("origin", _f * 3), ("spacing", _f * 3), ("interp", _i), ("dtype", _i)]


class CmemState(C.Structure):
    _fields_ = [(n, _f) for n in ("p0", "p1", "p2", "p3", "bs", "A1", "A2", "B", "alpha", "beta", "ay_bs", "az_bs", "a5", "a21")] + \
               [("beta_c", _f * 4)] + [(n, _f) for n in ("charlie", "dn", "ds", "theta_n", "theta_s", "r0_lin")]


class CmemParams(C.Structure):
    _fields_ = [("has_v", _i), ("v", _i * 3), ("has_b", _i), ("b", _i * 3), ("dipole", _f), ("p0", _f), ("p1", _i), ("p2", _i),
                ("p3", _i)] + [(n, _f) for n in ("B", "alpha", "beta", "bs", "A1", "A2", "ay_bs", "az_bs", "density")]


class View(C.Structure):
    _fields_ = [("width", _i), ("height", _i), ("fov_w", _f), ("fov_h", _f), ("position", _f * 3), ("aim", _f * 3),
                ("ray_length", _f), ("samples", _i), ("semantics", _i), ("math", _i), ("flags", _i)]


class RenderStats(C.Structure):
    _fields_ = [("samples_in_cube", C.c_longlong), ("samples_nominal", C.c_longlong), ("kernel_ms", _d)]


class Tuning(C.Structure):
    _fields_ = [("tile_w", _i), ("tile_h", _i), ("lanes_per_ray", _i), ("unroll", _i)]


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_size_t), ("max_batch", _i), ("max_inflight", _i), ("batch_window_ms", _d),
                ("reserve_pixels", C.c_size_t)]


class Request(C.Structure):
    _fields_ = [("volume", _i), ("view", View)]


class ResultInfo(C.Structure):
    _fields_ = [("status", _i), ("submit_s", _d), ("start_s", _d), ("done_s", _d), ("batch_size", _i)]


class ShellParams(C.Structure):
    _fields_ = [("n", _i), ("x0", _f), ("x1", _f), ("y0", _f), ("y1", _f), ("r0_lo", _f), ("r0_hi", _f), ("amp", _f),
                ("bg", _f), ("sigma", _f), ("seed", C.c_uint64)]


# (name, restype, argtypes) of every entry point in include/birp.h
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
SIGNATURES = [
    ("birp_strerror", C.c_char_p, [_i]),
    ("birp_last_error", C.c_char_p, []),
    ("birp_version", _i, []),
    ("birp_device_open", _i, [_i, _PP]),
    ("birp_device_close", _i, [_P]),
    ("birp_device_synchronize", _i, [_P]),
    ("birp_timer_begin", _i, [_P]),
    ("birp_timer_end", _i, [_P, _DP]),
    ("birp_volume_upload", _i, [_P, C.POINTER(VolumeDesc), _FP, _PP]),
    ("birp_cmem_default_params", None, [C.POINTER(CmemParams)]),
    ("birp_cmem_init", _i, [C.POINTER(CmemParams), C.POINTER(CmemState)]),
    ("birp_cmem_create", _i, [_P, C.POINTER(CmemState), _PP]),
    ("birp_volume_free", _i, [_P]),
    ("birp_volume_bytes", _i, [_P, C.POINTER(C.c_size_t)]),
    ("birp_volume_sample", _i, [_P, _FP, C.c_longlong, _FP]),
    ("birp_render", _i, [_P, C.POINTER(View), _FP, C.POINTER(RenderStats)]),
    ("birp_space_points", _i, [_P, C.POINTER(View), _i, _i, _FP]),
    ("birp_space_integrate", _i, [_P, C.POINTER(View), _i, _i, _FP, _FP]),
    ("birp_image_create", _i, [_P, _i, _i, _PP]),
    ("birp_image_free", _i, [_P]),
    ("birp_image_download", _i, [_P, _FP]),
    ("birp_image_upload", _i, [_P, _FP]),
    ("birp_render_async", _i, [_P, C.POINTER(View), _P]),
    ("birp_render_views_noalloc", _i, [_P, C.POINTER(View), _i, C.POINTER(C.c_void_p)]),
    ("birp_count_samples", _i, [_P, C.POINTER(View), C.POINTER(C.c_longlong)]),
    ("birp_render_f64", _i, [_P, C.POINTER(View), _DP]),
    ("birp_view_basis", _i, [C.POINTER(View), _FP, _FP, _FP]),
    ("birp_batch_create", _i, [_P, C.POINTER(View), _i, _PP]),
    ("birp_batch_free", _i, [_P]),
    ("birp_batch_set_tuning", _i, [_P, C.POINTER(Tuning)]),
    ("birp_batch_launch", _i, [_P]),
    ("birp_batch_launch_eager", _i, [_P]),
    ("birp_batch_launch_graph", _i, [_P]),
    ("birp_batch_wait", _i, [_P]),
    ("birp_batch_download", _i, [_P, _i, _FP]),
    ("birp_batch_stats", _i, [_P, C.POINTER(C.c_longlong), _DP]),
    ("birp_batch_set_output", _i, [_P, _FP]),
    ("birp_batch_create_exact", _i, [C.POINTER(_P), _i, C.POINTER(View), C.POINTER(_i), _i, _PP]),
    ("birp_batch_samples", _i, [_P, C.POINTER(C.c_longlong)]),
    ("birp_autotune", _i, [_P, C.POINTER(View), _i, C.c_char_p]),
    ("birp_autotune_load", _i, [C.c_char_p]),
    ("birp_autotune_lookup", _i, [_i, _i, _i, _i, C.POINTER(Tuning)]),
    ("birp_server_create", _i, [_P, C.POINTER(ServerConfig), _PP]),
    ("birp_server_destroy", _i, [_P]),
    ("birp_server_add_volume", _i, [_P, _P, C.POINTER(_i)]),
    ("birp_server_submit", _i, [_P, C.POINTER(Request), _FP, C.POINTER(C.c_uint64)]),
    ("birp_server_wait", _i, [_P, C.c_uint64, _d, C.POINTER(ResultInfo)]),
    ("birp_clock_s", _d, []),
    ("birp_host_alloc", _i, [C.c_size_t, _PP]),
    ("birp_host_free", _i, [_P]),
    ("birp_synth_shell", _i, [C.POINTER(ShellParams), _FP, _FP, _FP, _FP]),
    ("birp_synth_orbit", _i, [_i, _f, _f, _f, C.c_uint64, _FP]),
    ("birp_debug_inject_failure", _i, [_P, _i]),
    ("birp_debug_exact_order", _i, [_i, _IP]),
    ("birp_debug_exact_plan", _i, [_P, C.POINTER(View), _IP, _i, _IP]),
]

_lib = None


class BirpError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        detail = lib().birp_last_error().decode(errors="replace")
        super().__init__(f"{where}: {lib().birp_strerror(status).decode()} ({status}): {detail}")


def lib():
    """Load libbirp_hip.so (built in-tree by birp_cpp_amd/Makefile); raise if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C birp_cpp_amd` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, where: str) -> None:
    if status != OK:
        raise BirpError(status, where)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_FP)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def make_view(width, height, fov_w, fov_h, position, aim, ray_length, samples, semantics=SEM_EXTENDED,
              math=MATH_EXACT, flags=0) -> View:
    return View(int(width), int(height), float(fov_w), float(fov_h), (_f * 3)(*position), (_f * 3)(*aim),
                float(ray_length), int(samples), int(semantics), int(math), int(flags))


# ------------------------------------------------------------------------------ objects
class Device:
    def __init__(self, ordinal: int = 0):
        self.h = C.c_void_p()
        check(lib().birp_device_open(ordinal, C.byref(self.h)), "birp_device_open")

    def close(self):
        if self.h:
            lib().birp_device_close(self.h)
            self.h = C.c_void_p()

    def synchronize(self):
        check(lib().birp_device_synchronize(self.h), "birp_device_synchronize")

    def timer_begin(self):
        check(lib().birp_timer_begin(self.h), "birp_timer_begin")

    def timer_end(self) -> float:
        ms = C.c_double()
        check(lib().birp_timer_end(self.h, C.byref(ms)), "birp_timer_end")
        return ms.value

    def inject_failure(self, count: int):
        check(lib().birp_debug_inject_failure(self.h, count), "birp_debug_inject_failure")

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Volume:
    """A cube (or CMEM model) resident on the device."""

    def __init__(self, dev: Device, h):
        self.dev = dev
        self.h = h

    @classmethod
    def upload(cls, dev: Device, values: np.ndarray, origin, spacing, interp=INTERP_TRILINEAR, dtype=DTYPE_F32):
        v = np.ascontiguousarray(values, np.float32)
        nz, ny, nx = v.shape
        desc = VolumeDesc(nx, ny, nz, (_f * 3)(*origin), (_f * 3)(*spacing), interp, dtype)
        h = C.c_void_p()
        check(lib().birp_volume_upload(dev.h, C.byref(desc), fptr(v), C.byref(h)), "birp_volume_upload")
        return cls(dev, h)

    @classmethod
    def cmem(cls, dev: Device, state: CmemState):
        h = C.c_void_p()
        check(lib().birp_cmem_create(dev.h, C.byref(state), C.byref(h)), "birp_cmem_create")
        return cls(dev, h)

    def free(self):
        if self.h:
            lib().birp_volume_free(self.h)
            self.h = C.c_void_p()

    def nbytes(self) -> int:
        b = C.c_size_t()
        check(lib().birp_volume_bytes(self.h, C.byref(b)), "birp_volume_bytes")
        return b.value

    def sample(self, xyz: np.ndarray) -> np.ndarray:
        p = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        out = np.zeros(p.shape[0], np.float32)
        check(lib().birp_volume_sample(self.h, fptr(p), p.shape[0], fptr(out)), "birp_volume_sample")
        return out

    def render(self, view: View, init: np.ndarray | None = None, out: np.ndarray | None = None):
        """Render into a new image, into `out` (H x W float32, C-contiguous; overwritten unless the
        view accumulates), or into a copy of `init` (the accumulate contract of Camera::Render)."""
        if out is not None:
            if out.dtype != np.float32 or out.shape != (view.height, view.width) or not out.flags.c_contiguous:
                raise ValueError("out must be a C-contiguous float32 array of shape (height, width)")
            img = out
            if init is not None:
                img[...] = init
        else:
            img = np.zeros((view.height, view.width), np.float32) if init is None else np.ascontiguousarray(init, np.float32).copy()
        st = RenderStats()
        check(lib().birp_render(self.h, C.byref(view), fptr(img), C.byref(st)), "birp_render")
        return img, st

    def render_f64(self, view: View) -> np.ndarray:
        img = np.zeros((view.height, view.width), np.float64)
        check(lib().birp_render_f64(self.h, C.byref(view), dptr(img)), "birp_render_f64")
        return img

    def count_samples(self, view: View) -> int:
        c = C.c_longlong()
        check(lib().birp_count_samples(self.h, C.byref(view), C.byref(c)), "birp_count_samples")
        return c.value

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


class Image:
    def __init__(self, dev: Device, width: int, height: int):
        self.dev, self.W, self.H = dev, width, height
        self.h = C.c_void_p()
        check(lib().birp_image_create(dev.h, width, height, C.byref(self.h)), "birp_image_create")

    def download(self) -> np.ndarray:
        out = np.zeros((self.H, self.W), np.float32)
        check(lib().birp_image_download(self.h, fptr(out)), "birp_image_download")
        return out

    def upload(self, host: np.ndarray) -> None:
        a = np.ascontiguousarray(host, np.float32)
        if a.shape != (self.H, self.W):
            raise ValueError(f"image is {self.H}x{self.W}, got {a.shape}")
        check(lib().birp_image_upload(self.h, fptr(a)), "birp_image_upload")

    def free(self):
        if self.h:
            lib().birp_image_free(self.h)
            self.h = C.c_void_p()


def render_async(vol: Volume, view: View, img: Image) -> None:
    check(lib().birp_render_async(vol.h, C.byref(view), img.h), "birp_render_async")


class Batch:
    def __init__(self, vol: Volume, views):
        self.vol = vol
        self.views = list(views)
        arr = (View * len(self.views))(*self.views)
        self._arr = arr
        self.h = C.c_void_p()
        check(lib().birp_batch_create(vol.h, arr, len(self.views), C.byref(self.h)), "birp_batch_create")

    @classmethod
    def exact(cls, volumes, views, volume_of=None):
        """An exact batch (birp_batch_create_exact): view k of volumes[volume_of[k]] (or of the one
        volume), each image bit-identical to Volume.render of that view with MATH_EXACT."""
        self = cls.__new__(cls)
        volumes = list(volumes) if isinstance(volumes, (list, tuple)) else [volumes]
        self.vol = volumes[0]
        self._vols = volumes
        self.views = list(views)
        self._arr = (View * len(self.views))(*self.views)
        vh = (_P * len(volumes))(*[v.h.value for v in volumes])
        self._vh = vh
        vo = None
        if volume_of is not None:
            vo = (_i * len(self.views))(*[int(x) for x in volume_of])
        self._vo = vo
        self.h = C.c_void_p()
        check(lib().birp_batch_create_exact(vh, len(volumes), self._arr, vo, len(self.views), C.byref(self.h)),
              "birp_batch_create_exact")
        return self

    def samples(self):
        """Per-view in-cube sample counts of the last finished launch (exact batches)."""
        out = (C.c_longlong * len(self.views))()
        check(lib().birp_batch_samples(self.h, out), "birp_batch_samples")
        return [int(x) for x in out]

    def set_tuning(self, tuning: Tuning | None):
        check(lib().birp_batch_set_tuning(self.h, C.byref(tuning) if tuning is not None else None), "birp_batch_set_tuning")

    def launch(self):
        check(lib().birp_batch_launch(self.h), "birp_batch_launch")

    def launch_eager(self):
        check(lib().birp_batch_launch_eager(self.h), "birp_batch_launch_eager")

    def launch_graph(self):
        check(lib().birp_batch_launch_graph(self.h), "birp_batch_launch_graph")

    def wait(self):
        check(lib().birp_batch_wait(self.h), "birp_batch_wait")

    def image(self, k: int) -> np.ndarray:
        v = self.views[k]
        out = np.zeros((v.height, v.width), np.float32)
        check(lib().birp_batch_download(self.h, k, fptr(out)), "birp_batch_download")
        return out

    def set_output(self, host: np.ndarray | None):
        """Have the kernel write every image into `host` (a pinned_empty float32 array of the
        batch's total pixels, images in view order), or back into device memory for None."""
        if host is not None:
            total = sum(v.width * v.height for v in self.views)
            if host.dtype != np.float32 or host.size < total or not host.flags.c_contiguous:
                raise ValueError("host must be a C-contiguous float32 array of the batch's total pixels")
        check(lib().birp_batch_set_output(self.h, fptr(host) if host is not None else None), "birp_batch_set_output")
        self._out = host

    def stats(self):
        n = C.c_longlong()
        ms = C.c_double()
        check(lib().birp_batch_stats(self.h, C.byref(n), C.byref(ms)), "birp_batch_stats")
        return n.value, ms.value

    def free(self):
        if self.h:
            lib().birp_batch_free(self.h)
            self.h = C.c_void_p()


class Server:
    def __init__(self, dev: Device, memory_budget: int, max_batch=64, max_inflight=3, batch_window_ms=0.0, reserve_pixels=0):
        self.dev = dev
        cfg = ServerConfig(int(memory_budget), int(max_batch), int(max_inflight), float(batch_window_ms), int(reserve_pixels))
        self.h = C.c_void_p()
        check(lib().birp_server_create(dev.h, C.byref(cfg), C.byref(self.h)), "birp_server_create")
        self._bufs = {}

    def add_volume(self, vol: Volume) -> int:
        idx = _i()
        check(lib().birp_server_add_volume(self.h, vol.h, C.byref(idx)), "birp_server_add_volume")
        return idx.value

    def submit(self, volume: int, view: View, out: np.ndarray | None = None) -> int:
        """Queue a render; the image is written into `out` (or a new array) by the time wait()
        returns for this id."""
        if out is None:
            out = np.zeros((view.height, view.width), np.float32)
        elif out.dtype != np.float32 or out.shape != (view.height, view.width) or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float32 array of shape (height, width)")
        rid = C.c_uint64()
        req = Request(volume, view)
        check(lib().birp_server_submit(self.h, C.byref(req), fptr(out), C.byref(rid)), "birp_server_submit")
        self._bufs[rid.value] = out
        return rid.value

    def wait(self, rid: int, timeout_ms: float = -1.0):
        """Returns (status, image or None, ResultInfo)."""
        info = ResultInfo()
        st = lib().birp_server_wait(self.h, rid, timeout_ms, C.byref(info))
        img = self._bufs.pop(rid, None)
        return st, (img if st == OK else None), info

    def destroy(self):
        if self.h:
            lib().birp_server_destroy(self.h)
            self.h = C.c_void_p()


def exact_order(policy: int) -> int:
    """Set the tile order of single exact renders (0 natural order, 1 planned and reused, 2 planned
    every render, -1 the default); returns the policy in force before."""
    prev = C.c_int()
    check(lib().birp_debug_exact_order(int(policy), C.byref(prev)), "birp_debug_exact_order")
    return prev.value


def exact_plan(vol: "Volume", view: View) -> np.ndarray:
    """The device's tile plan for `view` over `vol` (tile ids, heaviest estimate first)."""
    nb = C.c_int()
    probe = np.zeros(1, np.int32)
    lib().birp_debug_exact_plan(vol.h, C.byref(view), probe.ctypes.data_as(_IP), 0, C.byref(nb))  # sets nb only
    order = np.zeros(max(nb.value, 1), np.int32)
    check(lib().birp_debug_exact_plan(vol.h, C.byref(view), order.ctypes.data_as(_IP), order.size, C.byref(nb)),
          "birp_debug_exact_plan")
    return order[:nb.value].copy()


def pinned_empty(shape, dtype=np.float32) -> np.ndarray:
    """An uninitialised array in pinned host memory from birp_host_alloc (freed with the array):
    a server request whose result is such an array has its image written there by the kernel."""
    n = int(np.prod(shape)) * np.dtype(dtype).itemsize
    p = C.c_void_p()
    check(lib().birp_host_alloc(max(n, 1), C.byref(p)), "birp_host_alloc")
    buf = (C.c_char * max(n, 1)).from_address(p.value)
    buf._owner = _PinnedBlock(p.value)  # keeps the allocation alive as long as any view of it
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)


class _PinnedBlock:
    def __init__(self, addr: int):
        self.addr = addr

    def __del__(self):
        if _lib is not None and self.addr:
            _lib.birp_host_free(C.c_void_p(self.addr))
            self.addr = 0


def clock_s() -> float:
    return lib().birp_clock_s()


def autotune(vol: Volume, views, path: str | None = None) -> None:
    arr = (View * len(views))(*views)
    check(lib().birp_autotune(vol.h, arr, len(views), path.encode() if path else None), "birp_autotune")


def autotune_lookup(dtype, width, height, samples):
    t = Tuning()
    st = lib().birp_autotune_lookup(dtype, width, height, samples, C.byref(t))
    return t if st == OK else None


def view_basis(view: View):
    a = C.c_float()
    rot = np.zeros(9, np.float32)
    rinv = np.zeros(9, np.float32)
    st = lib().birp_view_basis(C.byref(view), C.byref(a), fptr(rot), fptr(rinv))
    return st, a.value, rot.reshape(3, 3), rinv.reshape(3, 3)


# ------------------------------------------------------------------------------ synthetic workloads
def synth_shell(n, r0_range, sigma, seed, amp=1e-5, bg=1e-7, x_range=(-10.0, 20.0), yz_range=(-15.0, 15.0)):
    """Seeded magnetosheath-shell cube: returns (values[n,n,n] float32 [z][y][x], origin, spacing, r0)."""
    p = ShellParams(int(n), x_range[0], x_range[1], yz_range[0], yz_range[1], r0_range[0], r0_range[1], amp, bg, sigma, seed)
    vals = np.zeros((n, n, n), np.float32)
    org = np.zeros(3, np.float32)
    spc = np.zeros(3, np.float32)
    r0 = C.c_float()
    check(lib().birp_synth_shell(C.byref(p), fptr(vals), fptr(org), fptr(spc), C.byref(r0)), "birp_synth_shell")
    return vals, tuple(float(x) for x in org), tuple(float(x) for x in spc), r0.value


def synth_orbit(count, seed, perigee=1.8, apogee=19.0, incl_deg=70.0) -> np.ndarray:
    out = np.zeros((count, 3), np.float32)
    check(lib().birp_synth_orbit(count, perigee, apogee, incl_deg, seed, fptr(out)), "birp_synth_orbit")
    return out


def cmem_params(**kw) -> CmemParams:
    """CMEM::Init's arguments: the defaults of CMEM::Init() with keyword overrides (v=[...] and
    b=[...] set the vectors; the other names are the birp_cmem_params fields)."""
    p = CmemParams()
    lib().birp_cmem_default_params(C.byref(p))
    for name, val in kw.items():
        if name in ("v", "b"):
            setattr(p, "has_" + name, 1)
            getattr(p, name)[:] = [int(x) for x in val]
        else:
            setattr(p, name, val)
    return p


def cmem_init(params: CmemParams | None = None) -> CmemState:
    """The post-Init CMEM state (CMEM::Init on the host, bit-identical to the reference)."""
    s = CmemState()
    check(lib().birp_cmem_init(C.byref(params if params is not None else cmem_params()), C.byref(s)), "birp_cmem_init")
    return s


def cmem_state_from(state: dict) -> CmemState:
    """Post-Init CMEM members (e.g. the oracle's or the C++ CMEM's) -> CmemState."""
    s = CmemState()
    for name in ("p0", "p1", "p2", "p3", "bs", "A1", "A2", "B", "alpha", "beta", "ay_bs", "az_bs", "charlie", "dn", "ds",
                 "theta_n", "theta_s", "r0_lin"):
        setattr(s, name, float(state[name]))
    s.a5 = float(state["a"][5])
    s.a21 = float(state["a"][21])
    for k in range(4):
        s.beta_c[k] = float(state["beta_c"][k])
    return s
