"""Exact float32 arithmetic helpers for the tests (numpy has no fused multiply-add).

fma32(a, b, c) is IEEE-754 binary32 fma: a * b + c rounded once. The product of two float32 is
exact in float64; the float64 sum is made round-to-odd (if inexact, the neighbour with an odd
last bit), and rounding a round-to-odd float64 to float32 (29 extra bits) is correctly rounded.
"""
import numpy as np


def fma32(a, b, c):
    a = np.asarray(a, np.float32).astype(np.float64)
    b = np.asarray(b, np.float32).astype(np.float64)
    c = np.asarray(c, np.float32).astype(np.float64)
    p = a * b                      # exact
    s = p + c
    bb = s - p
    err = (p - (s - bb)) + (c - bb)  # TwoSum: s + err == p + c exactly
    bits = s.view(np.int64) if s.ndim else np.array(s).view(np.int64)
    odd = (bits & 1) == 1
    toward = np.nextafter(s, np.where(err > 0, np.inf, -np.inf))
    ro = np.where((err != 0) & ~odd, toward, s)
    return ro.astype(np.float32)


def dist2(p, q):
    """Squared distance fma(dz, dz, fma(dy, dy, dx * dx)) in float32, broadcasting (..., 3)."""
    d = (np.asarray(p, np.float32)[..., :3] - np.asarray(q, np.float32)[..., :3]).astype(np.float32)
    xx = (d[..., 0] * d[..., 0]).astype(np.float32)
    return fma32(d[..., 2], d[..., 2], fma32(d[..., 1], d[..., 1], xx))
