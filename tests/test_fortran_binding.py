"""The Fortran side of the drop-in boundary, on the CPU: the generated
iso_c_binding module and the drop-in subroutines compile with amdflang, a
Fortran caller links libx3d.so, and

* x3d_schemes_host called from Fortran returns the coefficient arrays of
  schemes() bit for bit as the C ABI does from Python (and those equal the
  oracle's, tests/test_capi_host.py);
* entries called before x3d_init return X3D_ERR_STATE with a message
  instead of failing;
* the committed binding files are what make_binding.py writes.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "incompact3d_amd", "libx3d.so")


@pytest.fixture(scope="module")
def checker():
    from incompact3d_amd import build as b
    if b.flang() is None:
        pytest.skip("no Fortran compiler in this image")
    if not os.path.exists(LIB):
        pytest.fail("libx3d.so is not built (python -m incompact3d_amd.build)")
    return b.build_fortran()


def test_generated_binding_is_current():
    from incompact3d_amd.fortran import make_binding as mb
    with open(os.path.join(mb.HERE, "x3d_binding.f90")) as f:
        assert f.read() == mb.binding()
    with open(os.path.join(mb.HERE, "x3d_drop_in.F90")) as f:
        assert f.read() == mb.drop_in()


def test_fortran_caller_schemes_and_error_paths(checker):
    import incompact3d_amd as X
    out = subprocess.run([checker, "cpu"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    arrays, status, message = {}, {}, ""
    for line in out.stdout.splitlines():
        f = line.split()
        if f[0].startswith("status_"):
            status[f[0]] = int(f[1])
        elif f[0] == "message":
            message = line
        else:
            arrays.setdefault(f[0], []).append((int(f[1]), float(f[2])))
    p = X.make_params(32, 33, 16, ncly1=2, nclyn=2)
    assert len(arrays) == 10
    for name, vals in arrays.items():
        got = np.array([v for _, v in sorted(vals)])
        ref = X.schemes_host(p, name)
        assert np.array_equal(got, ref), name
    assert status == {"status_host_derx_00": 2, "status_host_poisson": 2}
    assert "x3d_init" in message
