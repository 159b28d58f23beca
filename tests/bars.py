"""The precision contract's bars, shared by the GPU parity tests.

  FP64 (FP64 meshes): bit-exact to the FP64 argmin (checked with array_equal where it applies).
  FP32 + FP64 fixup: closest-triangle index identical to the FP64 argmin (lowest index on exact
    ties); distance within 1e-6 RELATIVE of the FP64 distance (the output is the FP64 distance of
    that triangle rounded to float, so about 6e-8 in practice); sign identical outside the
    1e-6 * bbox-diagonal band.
  Plain FP32: distance within 1e-6 * bbox diagonal ABSOLUTE.  Near the surface that is far above
    1e-6 relative (an FP32 distance between points of magnitude ~1 carries ~1e-7 absolute error), so
    the observed relative error is reported, with the fraction of points where it exceeds 1e-6.
"""
from __future__ import annotations

import numpy as np

REL = 1e-6      # relative distance bar of FP32 + fixup
ABS_DIAG = 1e-6  # absolute bar of plain FP32, times the bounding-box diagonal
FLOOR = 1e-30   # absolute floor where the true distance is 0


def rel_err(got, d_ref):
    """|got - d_ref| / d_ref (FP64), with FLOOR where d_ref is 0."""
    got = np.abs(np.asarray(got, np.float64))
    d_ref = np.asarray(d_ref, np.float64)
    return np.abs(got - d_ref) / np.maximum(d_ref, FLOOR)


def assert_fixup_distances(got, d_ref, what=""):
    """FP32 + fixup distances (|got| for signed outputs) within 1e-6 relative of the FP64 ones."""
    got = np.abs(np.asarray(got, np.float64))
    d_ref = np.asarray(d_ref, np.float64)
    err = np.abs(got - d_ref)
    bad = err > REL * d_ref + FLOOR
    assert not bad.any(), (f"{what}: {int(bad.sum())} distances beyond 1e-6 relative, worst rel "
                           f"{float(rel_err(got, d_ref).max()):.3g}")
    return float(rel_err(got, d_ref).max())


def plain_report(got, d_ref, diag, what=""):
    """Plain FP32: asserts the absolute bar, returns (and prints) the observed relative error."""
    got = np.abs(np.asarray(got, np.float64))
    d_ref = np.asarray(d_ref, np.float64)
    err = np.abs(got - d_ref)
    assert err.max() <= ABS_DIAG * diag, f"{what}: max |err| {err.max():.3g} > 1e-6 * diag"
    rel = rel_err(got, d_ref)
    frac = float(np.mean(rel > REL))
    print(f"[plain FP32 {what}] max abs err {err.max():.3g} (1e-6*diag = {ABS_DIAG * diag:.3g}), "
          f"max rel err {rel.max():.3g}, fraction above 1e-6 relative {frac:.3g}")
    return float(rel.max()), frac


def sign_mismatches_outside_band(got_signed, ref_signed, band):
    """Points farther than `band` from the surface whose signs differ."""
    g = np.asarray(got_signed, np.float64)
    r = np.asarray(ref_signed, np.float64)
    out = np.abs(r) > band
    return int(np.sum((np.sign(g) != np.sign(r)) & out))


def error_stats(got, d_ref, got_tri=None, ref_tri=None, got_signed=None, ref_signed=None, band=None):
    """Max and RMS distance error against the FP64 reference, and index / sign mismatch counts."""
    got = np.abs(np.asarray(got, np.float64))
    d_ref = np.asarray(d_ref, np.float64)
    e = got - d_ref
    st = {"n": int(d_ref.size), "max_err": float(np.abs(e).max()), "rms_err": float(np.sqrt(np.mean(e * e))),
          "max_rel_err": float(rel_err(got, d_ref).max())}
    if got_tri is not None:
        st["index_mismatch"] = int(np.sum(np.asarray(got_tri) != np.asarray(ref_tri)))
    if got_signed is not None:
        st["sign_mismatch_outside_band"] = sign_mismatches_outside_band(got_signed, ref_signed, band)
    return st
