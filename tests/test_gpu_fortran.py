"""A Fortran caller on the GPU: incompact3d_amd/fortran/check_binding (built
by build()) binds the reference's procedure pointers (DERIVATIVE_X/_Y/_YY/_Z
of module_param.f90:135-168) to the drop-in subroutines and calls them, the
staggered deryvp/derypv and poisson(rhs) as navier.f90 does, with the module
coefficient arrays, on host arrays of a 32 x 33 x 16 box with walls in y.
Every result against the oracle: FP64 1e-12 relative; the 3-D FFT against
numpy's.  The one-sided second-derivative closures at the walls have a
roundoff floor above that on a smooth O(1) field (the oracle's own FP64
evaluation is 1.5e-12 from its extended-precision one at the top wall): there
the bar is the oracle's, within twice the FP64 oracle's distance from the
same algorithm evaluated in extended precision."""
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "incompact3d_amd", "fortran", "_build", "check_binding")


def test_fortran_drop_in_matches_oracle(tmp_path, torch_cuda):
    from oracle import x3d_oracle as O
    assert os.path.exists(CHECK), "the Fortran caller is not built (python -m incompact3d_amd.build)"
    path = str(tmp_path / "fortran_out.bin")
    print("check_binding gpu", flush=True)
    r = subprocess.run([CHECK, "gpu", path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nx, ny, nz = 32, 33, 16
    nym = ny - 1
    raw = np.fromfile(path, dtype=np.float64)
    n, npr = nx * ny * nz, nx * nym * nz
    sizes = [n] * 7 + [npr, n, npr, 2 * nx * nym * (nz // 2 + 1)]
    assert raw.size == sum(sizes)
    parts, o = [], 0
    for s in sizes:
        parts.append(raw[o:o + s])
        o += s
    u = parts[0].reshape(nz, ny, nx)
    p = O.Params(nx, ny, nz, ncly1=2, nclyn=2)
    S = O.Schemes(p)
    refs = [O.der1(S, u, "x", 0), O.der1(S, u, "y", 0), O.der1(S, u, "z", 0),
            O.der2(S, u, "x", 0), O.der2(S, u, "y", 0), O.der2(S, u, "z", 0)]
    ops = [(O.der1, "x"), (O.der1, "y"), (O.der1, "z"), (O.der2, "x"), (O.der2, "y"), (O.der2, "z")]
    for k, (ref, (fn, ax)) in enumerate(zip(refs, ops)):
        got = parts[1 + k].reshape(nz, ny, nx)
        err = rel_err(got, ref)
        if err > 1e-12:
            ext = fn(S, u.astype(np.longdouble), ax, 0)
            floor = rel_err(ref, ext)
            assert rel_err(got, ext) <= 2.0 * floor, (k, err, floor)
    tp = O.dervp(S, u, "y", 0)
    assert rel_err(parts[7].reshape(nz, nym, nx), tp) <= 1e-12
    assert rel_err(parts[8].reshape(nz, ny, nx), O.derpv(S, tp, "y", 1)) <= 1e-12
    rhs = tp - tp.mean()
    assert rel_err(parts[9].reshape(nz, nym, nx), O.poisson(S, rhs)) <= 1e-12
    spec = parts[10].view(np.complex128).reshape(nz // 2 + 1, nym, nx)
    ref = np.fft.fftn(tp)[: nz // 2 + 1]
    assert np.abs(spec - ref).max() <= 1e-12 * np.abs(ref).max()
