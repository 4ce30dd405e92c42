"""Shared parity rules of the GPU tests.

CMEM renders (the analytic model of src/Cmem.cpp) cannot be bit-exact: the device evaluates
acosf/cosf/sinf/powf/expf with ROCm's libm, the reference with glibc's, and the two differ in
the last bit for a few percent of arguments. The model is also discontinuous: at the zone edges
r = rmp +- h and r = rbs +- h (Cmem.cpp:132-176) the ramps are normalised by rmp + h instead of
2h, so the value jumps there (from ~0.4% of the inner emissivity to all of it at rmp + h). A
sample within an ulp or so of an edge can land on either side under the two libms and moves its
pixel by that jump. The rule below therefore:
  * splits off the pixels whose ray has a sample within CMEM_EDGE_BAND (relative) of an edge, as
    found in FP64 by the oracle, and requires at most 5% of the image to be such pixels;
  * holds every other pixel to the FP32 bar, image relative L2 <= 1e-5 against the FP32 oracle;
  * holds every edge pixel to a bound: finite, and within the sum, over its near-edge samples, of
    the model's jump at those edges (FP64, in image units: or_cmem_edge_bound) plus 1e-5 of the
    pixel, so an edge pixel can only differ by the zone flips the rule allows for;
  * and requires the device's error against the FP64 oracle on the non-edge pixels to be no
    worse than twice the FP32 reference's own (plus 1e-7), so that agreement with the oracle
    does not hide a loss of accuracy.
"""
import numpy as np

CMEM_EDGE_BAND = 1e-5  # |r - edge| / r; about 100 FP32 ulps of r
CMEM_EDGE_MAX_FRACTION = 0.05


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cmem_render_parity(img, ref32, ref64, edge_mask, edge_bound, tol=1e-5):
    """Returns (rel L2 vs FP32 on non-edge pixels, same vs FP64, whole-image rel L2 vs FP32,
    edge fraction, largest |img - ref32| / bound over the edge pixels)."""
    img = np.asarray(img, np.float64)
    ref32 = np.asarray(ref32, np.float64)
    edge = np.asarray(edge_mask, bool)
    keep = ~edge
    assert edge.mean() <= CMEM_EDGE_MAX_FRACTION, f"{edge.sum()} of {edge.size} pixels touch a zone edge"
    err32 = rel_l2(img[keep], ref32[keep])
    assert err32 <= tol, f"CMEM image relative L2 {err32:.3g} on non-edge pixels"
    dev64 = rel_l2(img[keep], ref64[keep])
    ref_own = rel_l2(ref32[keep], ref64[keep])
    assert dev64 <= 2 * ref_own + 1e-7, f"device error vs FP64 {dev64:.3g}, reference FP32's own {ref_own:.3g}"
    worst = 0.0
    if edge.any():
        d = np.abs(img[edge] - ref32[edge])
        allow = np.asarray(edge_bound, np.float64)[edge] * 1.001 + tol * np.abs(ref32[edge]) + 1e-30
        assert np.isfinite(img[edge]).all(), "non-finite edge pixel"
        over = d > allow
        assert not over.any(), f"{over.sum()} edge pixels differ by more than their zone-flip bound (worst {np.max(d / allow):.3g}x)"
        worst = float(np.max(d / allow))
    return err32, dev64, rel_l2(img, ref32), float(edge.mean()), worst
