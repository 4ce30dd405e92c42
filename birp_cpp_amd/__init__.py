"""birp_cpp_amd -- MI355X renderer for BIRP's hot path (ray generation, cube/CMEM march, per-ray
integration). The product is the HIP library behind the C ABI of include/birp.h
(lib/libbirp_hip.so) and the C++ API over it (lib/libbirp_api.so); this package holds their
sources (csrc/, api/), the build (Makefile) and Python bindings to the C ABI (capi.py)."""
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8) -> None:
    """Compile csrc/ and api/ for gfx950 into lib/ (always rebuilds what changed)."""
    subprocess.run(["make", "-s", "-C", PKG_DIR, f"-j{jobs}", "all"], check=True)
