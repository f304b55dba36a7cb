"""The split form of the FP64 velocity (csrc/state64.h): an FP64 value held
as its FP32 rounding hi plus an int32 residual r.  The header's functions
are compiled for the host with g++ and checked against what the arithmetic
fixes:

* join(hi, split(u, hi)) == u bit for bit for finite u with |u| >= 2^-127
  (random 53-bit mantissas over every FP32 exponent, both signs), and the
  residual stays within |r| <= 2^29;
* the awkward points: u exactly representable in FP32 (r = 0), u that rounds
  up to a power of two (hi's exponent above u's), u one FP64 ulp either side
  of an FP32 value, the bottom of the normal FP32 range, signed zeros;
* below 2^-127 the round trip errs by at most 2^-180 in absolute value.

A planted mistake (the residual scaled by 2^(52-e) instead of 2^(53-e))
must fail the round trip, so the check is not vacuous."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "incompact3d_amd", "csrc")

PROG = r"""
#include <cstdio>
#include <cstdlib>
#include "state64.h"
#ifdef PLANT
#define SPLIT(u, h) ((int32_t)std::rint(std::ldexp((u) - (double)(h), 52 - x3d::split_exponent(h))))
#else
#define SPLIT(u, h) x3d::split_residual((u), (h))
#endif
int main() {
  // stdin: n doubles (binary); stdout: n joined doubles, then n residuals
  size_t n;
  if (fread(&n, 8, 1, stdin) != 1) return 2;
  double *u = (double *)malloc(n * 8), *o = (double *)malloc(n * 8);
  int32_t *r = (int32_t *)malloc(n * 4);
  if (fread(u, 8, n, stdin) != n) return 3;
  for (size_t i = 0; i < n; ++i) {
    const float h = (float)u[i];
    r[i] = SPLIT(u[i], h);
    o[i] = x3d::join_split(h, r[i]);
  }
  fwrite(o, 8, n, stdout);
  fwrite(r, 4, n, stdout);
  return 0;
}
"""


def _build(tmp_path, plant=False):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "rt.cpp"
    src.write_text(PROG)
    exe = tmp_path / ("rt_plant" if plant else "rt")
    cmd = [gxx, "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)]
    if plant:
        cmd.insert(1, "-DPLANT")
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def _run(exe, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    inp = np.array([u.size], dtype=np.uint64).tobytes() + u.tobytes()
    out = subprocess.run([str(exe)], input=inp, capture_output=True, check=True).stdout
    o = np.frombuffer(out[:8 * u.size], dtype=np.float64)
    r = np.frombuffer(out[8 * u.size:], dtype=np.int32)
    return o, r


def _random_normal_range(rng, n):
    # mantissa uniform in [1, 2) with all 52 fraction bits random, exponent
    # uniform over -127 .. 127, random sign
    frac = rng.integers(0, 1 << 52, size=n, dtype=np.uint64)
    m = 1.0 + frac.astype(np.float64) * 2.0 ** -52
    # (frac to float64 is exact: < 2^52)
    e = rng.integers(-127, 128, size=n)
    s = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    return s * np.ldexp(m, e)


def _edge_points():
    pts = [0.0, -0.0, 1.0, -1.0, 2.0 ** -126, -(2.0 ** -126), 2.0 ** -127, 3.0 * 2.0 ** -128]
    for e in (-126, -100, -20, -1, 0, 1, 7, 20, 60, 100, 126):
        p = 2.0 ** e
        # just below a power of two: rounds up to it (hi's exponent above u's)
        pts += [np.nextafter(p, 0.0), p - 2.0 ** (e - 25), p - 2.0 ** (e - 26)]
        # an FP32 value and its FP64 neighbours
        f = float(np.float32(1.2345678 * p))
        pts += [f, np.nextafter(f, np.inf), np.nextafter(f, -np.inf)]
        # halfway between two FP32 values, and one FP64 ulp either side
        h = f + float(np.spacing(np.float32(f))) / 2
        pts += [h, np.nextafter(h, np.inf), np.nextafter(h, -np.inf)]
    pts += [-x for x in pts]
    return np.array(pts, dtype=np.float64)


def test_round_trip_exact(tmp_path):
    exe = _build(tmp_path)
    rng = np.random.default_rng(20261203)
    u = np.concatenate([_random_normal_range(rng, 400_000), _edge_points()])
    o, r = _run(exe, u)
    assert np.array_equal(o.view(np.uint64), u.view(np.uint64)), "round trip changed a value"
    assert int(np.abs(r.astype(np.int64)).max()) <= 2 ** 29
    # exactly representable in FP32 -> zero residual
    f = u.astype(np.float32).astype(np.float64)
    assert np.all(r[f == u] == 0)


def test_round_trip_below_fp32_normal_range(tmp_path):
    exe = _build(tmp_path)
    rng = np.random.default_rng(7)
    frac = rng.integers(0, 1 << 52, size=50_000, dtype=np.uint64)
    m = 1.0 + frac.astype(np.float64) * 2.0 ** -52
    u = np.ldexp(m, rng.integers(-200, -127, size=m.size)) * np.where(rng.random(m.size) < 0.5, -1.0, 1.0)
    o, _ = _run(exe, u)
    assert np.max(np.abs(o - u)) <= 2.0 ** -180


def test_planted_scale_error_fails(tmp_path):
    exe = _build(tmp_path, plant=True)
    rng = np.random.default_rng(3)
    u = _random_normal_range(rng, 10_000)
    o, _ = _run(exe, u)
    assert not np.array_equal(o, u)
