"""Comparison helpers shared by the parity tests."""
import numpy as np


def rel_err(a, b, floor_frac=1e-6):
    """Largest |a - b| / |b| over elements with |b| above floor_frac·max|b|, and the largest
    |a - b| / max|b| over all elements (the field-scaled error, which covers values near zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    diff = np.abs(a - b)
    if scale == 0.0:
        return float(diff.max(initial=0.0)), float(diff.max(initial=0.0))
    mask = np.abs(b) > floor_frac * scale
    elem = float(np.max(diff[mask] / np.abs(b[mask]), initial=0.0))
    field = float(diff.max(initial=0.0) / scale)
    return elem, field


def assert_rel_terms(a, b, scale, rtol, name=""):
    """|a − b| ≤ rtol·max(|b|, scale) elementwise, where scale is the magnitude of the terms that were
    summed into b (a sum that cancels to near zero is compared against the size of its terms)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ref = np.maximum(np.abs(b), np.asarray(scale, dtype=np.float64))
    diff = np.abs(a - b)
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = np.where(ref > 0, diff / ref, np.where(diff > 0, np.inf, 0.0))
    worst = float(rel.max(initial=0.0))
    assert worst <= rtol, f"{name}: relative error {worst:.3e} (vs max(|G|, Σ|terms|)) > {rtol:.1e}"
    return worst


def assert_rel(a, b, rtol, name=""):
    elem, field = rel_err(a, b)
    assert np.all(np.isfinite(a)) == np.all(np.isfinite(b)), f"{name}: finiteness differs"
    assert elem <= rtol and field <= rtol, f"{name}: relative error {elem:.3e} (field-scaled {field:.3e}) > {rtol:.1e}"
    return elem, field


def assert_rel_envelope(a, b, envelope, rtol, name=""):
    """|a − b| ≤ max(rtol·|b|, envelope) elementwise, and what exceeds the envelope ≤ rtol·max|b|
    over the field, where envelope is a per-element bound on the rounding-level response of b
    (see the caller)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    env = np.asarray(envelope, dtype=np.float64)
    assert np.all(np.isfinite(a)) == np.all(np.isfinite(b)), f"{name}: finiteness differs"
    diff = np.abs(a - b)
    bound = np.maximum(rtol * np.abs(b), env)
    over = diff > bound
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    excess = np.maximum(diff - env, 0.0)
    field = float(excess.max(initial=0.0) / scale) if scale > 0 else float(excess.max(initial=0.0))
    assert not over.any(), (f"{name}: {int(over.sum())} elements beyond max(rtol·|b|, envelope); worst "
                            f"|Δ| {float(diff[over].max()):.3e} against bound {float(bound[over].min()):.3e}")
    assert field <= rtol, f"{name}: field-scaled error {field:.3e} > {rtol:.1e}"
    with np.errstate(invalid="ignore", divide="ignore"):
        elem = float(np.max(np.where(np.abs(b) > 0, diff / np.abs(b), 0.0), initial=0.0))
        # how much of the envelope the elements beyond rtol·|b| use
        used = float(np.max(np.where((diff > rtol * np.abs(b)) & (env > 0), diff / env, 0.0), initial=0.0))
    return elem, field, used
