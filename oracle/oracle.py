"""ctypes front end to the oracle (liboracle.so) and to the reference build (oracle/_ref).

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module, and only to check results or as the timed CPU baseline; the product package
never does. See birp_oracle.h for what each function restates and what pins it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbirp_ref.so")

SEM_REFERENCE = 0
SEM_EXTENDED = 1

_f = C.c_float
_d = C.c_double
_i = C.c_int
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Cube(C.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i), ("origin", _f * 3), ("spacing", _f * 3),
                ("data", _fp), ("trilinear", _i)]


class Cmem(C.Structure):
    _fields_ = [("v", _f * 3), ("b", _f * 3), ("pressure_dynamic", _f), ("pressure_magnetic", _f),
                ("dipole", _f), ("a", _f * 22), ("beta_c", _f * 4), ("charlie", _f), ("dn", _f),
                ("ds", _f), ("theta_n", _f), ("theta_s", _f), ("r0_lin", _f), ("p0", _f), ("bs", _f),
                ("A1", _f), ("A2", _f), ("B", _f), ("alpha", _f), ("beta", _f), ("p1", _f), ("p2", _f),
                ("p3", _f), ("ay_bs", _f), ("az_bs", _f), ("density", _f)]


class Space(C.Structure):
    _fields_ = [("kind", _i), ("cube", C.POINTER(Cube)), ("cmem", C.POINTER(Cmem))]


class View(C.Structure):
    _fields_ = [("W", _i), ("H", _i), ("fov_w", _f), ("fov_h", _f), ("pos", _f * 3), ("aim", _f * 3),
                ("L", _f), ("n", _i)]


_lib = None


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_lerpf.restype = _f
        L.or_lerpf.argtypes = [_f, _f, _f]
        L.or_cube_sample.restype = _f
        L.or_cube_sample.argtypes = [C.POINTER(Cube), _f, _f, _f]
        L.or_cube_sample_f64.restype = _d
        L.or_cube_sample_f64.argtypes = [C.POINTER(Cube), _d, _d, _d]
        L.or_cube_inside.restype = _i
        L.or_cube_inside.argtypes = [C.POINTER(Cube), _f, _f, _f]
        L.or_cmem_init.restype = None
        L.or_cmem_init.argtypes = [C.POINTER(Cmem), _ip, _ip, _f, _f, _i, _i, _i, _f, _f, _f, _f, _f, _f, _f, _f, _f]
        L.or_cmem_init_default.argtypes = [C.POINTER(Cmem)]
        L.or_cmem_sample.restype = _f
        L.or_cmem_sample.argtypes = [C.POINTER(Cmem), _f, _f, _f]
        L.or_cmem_sample_f64.restype = _d
        L.or_cmem_sample_f64.argtypes = [C.POINTER(Cmem), _d, _d, _d]
        L.or_cmem_edge_distance_f64.restype = _d
        L.or_cmem_edge_distance_f64.argtypes = [C.POINTER(Cmem), _d, _d, _d]
        L.or_cmem_edge_pixels.restype = _i
        L.or_cmem_edge_pixels.argtypes = [C.POINTER(Cmem), C.POINTER(View), _d, C.POINTER(C.c_ubyte)]
        L.or_cmem_edge_jump_f64.restype = _d
        L.or_cmem_edge_jump_f64.argtypes = [C.POINTER(Cmem), _d, _d, _d, _d]
        L.or_cmem_edge_bound.restype = _i
        L.or_cmem_edge_bound.argtypes = [C.POINTER(Cmem), C.POINTER(View), _d, _i, _i, _i, C.POINTER(C.c_ubyte), _dp]
        L.or_image_dims.argtypes = [_f, _f, _f, _ip, _ip]
        L.or_ray_dist.argtypes = [_f, _i, _fp, _fp]
        L.or_orient.restype = _f
        L.or_orient.argtypes = [_fp, _fp]
        L.or_orient_f64.restype = _d
        L.or_orient_f64.argtypes = [_dp, _dp]
        L.or_matmul3.argtypes = [_fp, _fp, _fp]
        L.or_get_inverse3.restype = _i
        L.or_get_inverse3.argtypes = [_fp, _fp]
        L.or_view_basis.restype = _i
        L.or_view_basis.argtypes = [C.POINTER(View), _fp, _fp]
        L.or_view_basis_f64.restype = _i
        L.or_view_basis_f64.argtypes = [C.POINTER(View), _dp, _dp]
        L.or_ray_dir.argtypes = [C.POINTER(View), _fp, _i, _i, _fp]
        L.or_ray_dir_f64.argtypes = [C.POINTER(View), _dp, _i, _i, _dp]
        L.or_ray_dirs.argtypes = [C.POINTER(View), _fp]
        L.or_scale_sample.restype = _f
        L.or_scale_sample.argtypes = [_f, _f]
        L.or_render_f32.restype = _i
        L.or_render_f32.argtypes = [C.POINTER(Space), C.POINTER(View), _i, _i, _i, _i, _i, _fp,
                                    C.POINTER(C.c_longlong)]
        L.or_render_f64.restype = _i
        L.or_render_f64.argtypes = [C.POINTER(Space), C.POINTER(View), _i, _i, _i, _i, _dp,
                                    C.POINTER(C.c_longlong)]
        _lib = L
    return _lib


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class OracleCube:
    """A cube held for the oracle: data [z][y][x] float32 plus origin/spacing metadata."""

    def __init__(self, data: np.ndarray, origin, spacing, trilinear: bool = True):
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        nz, ny, nx = self.data.shape
        self.c = Cube(nx, ny, nz, (_f * 3)(*origin), (_f * 3)(*spacing), _fptr(self.data), int(trilinear))
        self.space = Space(0, C.pointer(self.c), None)

    def sample(self, x, y, z) -> float:
        return lib().or_cube_sample(C.byref(self.c), x, y, z)

    def sample_f64(self, x, y, z) -> float:
        return lib().or_cube_sample_f64(C.byref(self.c), x, y, z)

    def inside(self, x, y, z) -> bool:
        return bool(lib().or_cube_inside(C.byref(self.c), x, y, z))


class OracleCmem:
    def __init__(self, params=None):
        self.m = Cmem()
        if params is None:
            lib().or_cmem_init_default(C.byref(self.m))
        else:
            p = dict(params)
            v = p.pop("v", None)
            b = p.pop("b", None)
            va = (_i * 3)(*v) if v else None
            ba = (_i * 3)(*b) if b else None
            lib().or_cmem_init(C.byref(self.m), va, ba, *cmem_positional(p))
        self.space = Space(1, None, C.pointer(self.m))

    def sample(self, x, y, z) -> float:
        return lib().or_cmem_sample(C.byref(self.m), x, y, z)

    def sample_f64(self, x, y, z) -> float:
        return lib().or_cmem_sample_f64(C.byref(self.m), x, y, z)

    def edge_distance(self, x, y, z) -> float:
        """Relative distance |r - e| / r of a point from the nearest zone edge (FP64)."""
        return lib().or_cmem_edge_distance_f64(C.byref(self.m), x, y, z)

    def edge_pixels(self, view: "View", rel_band: float) -> np.ndarray:
        """Boolean HxW mask of pixels whose FP32 ray has a sample within rel_band of a zone edge."""
        mask = np.zeros((view.H, view.W), np.uint8)
        rc = lib().or_cmem_edge_pixels(C.byref(self.m), C.byref(view), rel_band,
                                       mask.ctypes.data_as(C.POINTER(C.c_ubyte)))
        if rc < 0:
            raise ValueError("singular view basis")
        return mask.astype(bool)

    def edge_jump(self, x, y, z, rel_band: float) -> float:
        """FP64 change of the model's value at a point whose r lies within rel_band of zone edges
        (sum over those edges of |zone below - zone above| at the point's r; 0 if none is near)."""
        return lib().or_cmem_edge_jump_f64(C.byref(self.m), x, y, z, rel_band)

    def edge_bound(self, view: "View", rel_band: float, rows=None, nthreads=1):
        """(mask, bound): HxW (or rows) near-edge pixel mask and, per pixel, the largest image change
        that zone flips of its near-edge samples can make (FP64, in image units)."""
        r0, r1 = (0, view.H) if rows is None else rows
        mask = np.zeros((r1 - r0, view.W), np.uint8)
        bound = np.zeros((r1 - r0, view.W), np.float64)
        rc = lib().or_cmem_edge_bound(C.byref(self.m), C.byref(view), rel_band, r0, r1, nthreads,
                                      mask.ctypes.data_as(C.POINTER(C.c_ubyte)), _dptr(bound))
        if rc < 0:
            raise ValueError("singular view basis")
        return mask.astype(bool), bound

    def state(self) -> dict:
        return {name: (list(getattr(self.m, name)) if isinstance(getattr(self.m, name), C.Array)
                       else getattr(self.m, name)) for name, _ in Cmem._fields_}


CMEM_DEFAULTS = dict(dipole=0.0, p0=0.786300004, p1=1, p2=3, p3=4, B=2.0, alpha=2.5, beta=-1.6,
                     bs=12.6400003, A1=7.2e-06, A2=3.5000000000000004e-06, ay_bs=0.0, az_bs=0.0,
                     density=0.0)
CMEM_ORDER = ["dipole", "p0", "p1", "p2", "p3", "B", "alpha", "beta", "bs", "A1", "A2", "ay_bs", "az_bs",
              "density"]


def cmem_positional(p: dict):
    q = dict(CMEM_DEFAULTS)
    q.update(p)
    return [q[k] for k in CMEM_ORDER]


def make_view(W, H, fov_w, fov_h, pos, aim, L, n) -> View:
    return View(int(W), int(H), float(fov_w), float(fov_h), (_f * 3)(*pos), (_f * 3)(*aim), float(L), int(n))


def image_dims(pixel_deg, fov_h, fov_w):
    W, H = _i(), _i()
    lib().or_image_dims(pixel_deg, fov_h, fov_w, C.byref(W), C.byref(H))
    return W.value, H.value


def ray_dist(L, n):
    t = np.zeros(n, np.float32)
    w = _f()
    lib().or_ray_dist(L, n, _fptr(t), C.byref(w))
    return t, w.value


def orient(pos, aim) -> float:
    p = np.asarray(pos, np.float32)
    a = np.asarray(aim, np.float32)
    return lib().or_orient(_fptr(p), _fptr(a))


def orient_f64(pos, aim) -> float:
    p = np.asarray(pos, np.float64)
    a = np.asarray(aim, np.float64)
    return lib().or_orient_f64(_dptr(p), _dptr(a))


def view_basis(view: View):
    rot = np.zeros(9, np.float32)
    rinv = np.zeros(9, np.float32)
    rc = lib().or_view_basis(C.byref(view), _fptr(rot), _fptr(rinv))
    return rc, rot.reshape(3, 3), rinv.reshape(3, 3)


def view_basis_f64(view: View):
    rot = np.zeros(9, np.float64)
    rinv = np.zeros(9, np.float64)
    rc = lib().or_view_basis_f64(C.byref(view), _dptr(rot), _dptr(rinv))
    return rc, rot.reshape(3, 3), rinv.reshape(3, 3)


def get_inverse3(m):
    a = np.ascontiguousarray(m, np.float32).reshape(9)
    out = np.zeros(9, np.float32)
    rc = lib().or_get_inverse3(_fptr(a), _fptr(out))
    return rc, out.reshape(3, 3)


def matmul3(a, b):
    a = np.ascontiguousarray(a, np.float32).reshape(9)
    b = np.ascontiguousarray(b, np.float32).reshape(9)
    out = np.zeros(9, np.float32)
    lib().or_matmul3(_fptr(a), _fptr(b), _fptr(out))
    return out.reshape(3, 3)


def ray_dir(view: View, rinv, i, j):
    r = np.ascontiguousarray(rinv, np.float32).reshape(9)
    d = np.zeros(3, np.float32)
    lib().or_ray_dir(C.byref(view), _fptr(r), i, j, _fptr(d))
    return d


def ray_dirs(view: View) -> np.ndarray:
    """H x W x 3 unit rays of the view, as the FP32 render traces them."""
    d = np.zeros((view.H, view.W, 3), np.float32)
    lib().or_ray_dirs(C.byref(view), _fptr(d))
    return d


def ray_dir_f64(view: View, rinv, i, j):
    r = np.ascontiguousarray(rinv, np.float64).reshape(9)
    d = np.zeros(3, np.float64)
    lib().or_ray_dir_f64(C.byref(view), _dptr(r), i, j, _dptr(d))
    return d


def scale_sample(v, w) -> float:
    return lib().or_scale_sample(v, w)


def render_f32(space, view: View, semantics=SEM_EXTENDED, rows=None, nthreads=1, init=None):
    """Render with the FP32 reference-exact oracle. Returns (image[H,W] float32, n_incube)."""
    H, W = view.H, view.W
    img = np.zeros((H, W), np.float32) if init is None else np.ascontiguousarray(init, np.float32).copy()
    r0, r1 = (0, H) if rows is None else rows
    cnt = C.c_longlong(0)
    rc = lib().or_render_f32(C.byref(space.space), C.byref(view), semantics, r0, r1, int(init is not None),
                             nthreads, _fptr(img), C.byref(cnt))
    if rc != 0:
        raise RuntimeError("Determinant is 0")
    return img, cnt.value


def render_f64(space, view: View, semantics=SEM_EXTENDED, rows=None, nthreads=1):
    H, W = view.H, view.W
    img = np.zeros((H, W), np.float64)
    r0, r1 = (0, H) if rows is None else rows
    cnt = C.c_longlong(0)
    rc = lib().or_render_f64(C.byref(space.space), C.byref(view), semantics, r0, r1, nthreads, _dptr(img),
                             C.byref(cnt))
    if rc != 0:
        raise RuntimeError("Determinant is 0")
    return img, cnt.value


# ------------------------------------------------------------------------- reference build
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own DataCube/CMEM/Helper compiled from its sources (oracle/_ref)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_cube_load.restype = C.c_void_p
        R.ref_cube_load.argtypes = [C.c_char_p, _i]
        R.ref_cube_info.argtypes = [C.c_void_p, _ip, _fp, _fp]
        R.ref_cube_values.restype = _i
        R.ref_cube_values.argtypes = [C.c_void_p, _fp, C.c_longlong]
        R.ref_cube_sample.restype = _f
        R.ref_cube_sample.argtypes = [C.c_void_p, _f, _f, _f, _ip]
        R.ref_cube_free.argtypes = [C.c_void_p]
        R.ref_cube_from_grid.restype = C.c_void_p
        R.ref_cube_from_grid.argtypes = [_i, _i, _i, _fp, _fp, _fp, _i]
        R.ref_render_rows.restype = C.c_longlong
        R.ref_render_rows.argtypes = [C.c_void_p, _i, _i, _fp, _fp, _f, _i, _i, _i, _i, _fp]
        R.ref_cmem_default.restype = C.c_void_p
        R.ref_cmem_init.restype = C.c_void_p
        R.ref_cmem_init.argtypes = [_ip, _i, _ip, _i, _f, _f, _i, _i, _i, _f, _f, _f, _f, _f, _f, _f, _f, _f]
        R.ref_cmem_sample.restype = _f
        R.ref_cmem_sample.argtypes = [C.c_void_p, _f, _f, _f]
        R.ref_cmem_free.argtypes = [C.c_void_p]
        R.ref_matmul3.argtypes = [_fp, _fp, _fp]
        R.ref_get_inverse3.restype = _i
        R.ref_get_inverse3.argtypes = [_fp, _fp]
        R.ref_gei2gse.argtypes = [_fp]
        R.ref_radec_to_gsi.argtypes = [_f, _f, _fp]
        R.ref_radec_gse.argtypes = [_f, _f, _fp, _fp]
        R.ref_two_min.argtypes = [_fp, _i, _ip, _i, _fp]
        R.ref_xy_to_radec.argtypes = [_f, _f, _fp, _fp, _fp, _fp, _i, _f, _fp]
        R.ref_interp1.argtypes = [_fp, _fp, _i, _fp, _i, _fp]
        R.ref_nearest_below.restype = _i
        R.ref_nearest_below.argtypes = [_f, _fp, _i]
        R.ref_xyz_to_spherical.argtypes = [_fp, _i, _fp]
        _ref = R
    return _ref


def ref_cmem(params=None):
    R = ref()
    if params is None:
        return R.ref_cmem_default()
    p = dict(params)
    v = p.pop("v", None) or []
    b = p.pop("b", None) or []
    va = (_i * max(len(v), 1))(*v)
    ba = (_i * max(len(b), 1))(*b)
    return R.ref_cmem_init(va, len(v), ba, len(b), *cmem_positional(p))


class RefCube:
    """The reference's own DataCube (oracle/_ref) holding an in-memory grid."""

    def __init__(self, data: np.ndarray, origin, spacing, trilinear: bool = True):
        self.data = np.ascontiguousarray(data, np.float32)
        nz, ny, nx = self.data.shape
        o = np.asarray(origin, np.float32)
        sp = np.asarray(spacing, np.float32)
        self.h = ref().ref_cube_from_grid(nx, ny, nz, _fptr(o), _fptr(sp), _fptr(self.data), int(trilinear))

    def render_rows(self, view: View, dirs: np.ndarray, row0: int, row1: int, semantics=SEM_EXTENDED):
        """Rows [row0, row1) with the reference's sampler; returns (H x W image, samples added)."""
        img = np.zeros((view.H, view.W), np.float32)
        pos = np.asarray(list(view.pos), np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        cnt = ref().ref_render_rows(self.h, view.W, view.H, _fptr(d), _fptr(pos), view.L, view.n, row0, row1,
                                    semantics, _fptr(img))
        return img, cnt

    def free(self):
        if self.h:
            ref().ref_cube_free(self.h)
            self.h = None
