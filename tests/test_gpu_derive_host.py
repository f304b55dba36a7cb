"""The x3d_host_* twins of every derivative and staggered entry: the same
checks as test_gpu_derive.py with the fields in host memory (numpy arrays,
staged through HBM by the library), on the small grids of each boundary
family (periodic, walls in y, free-slip, mixed faces), FP64 and FP32."""
import pytest

from tests.test_gpu_derive import CONFIGS, _check_derivative, _check_staggered, _context

pytestmark = pytest.mark.gpu

HOST_CONFIGS = [c for c in CONFIGS if c[0] in ("periodic_small", "ywalls", "freeslip", "mixed")]


@pytest.fixture(scope="module", params=HOST_CONFIGS, ids=[c[0] for c in HOST_CONFIGS])
def ctx_host(request, torch_cuda):
    c = _context(request.param)
    yield c
    c[3].finalize()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("ax", ["x", "y", "z"])
@pytest.mark.parametrize("kind", ["der1", "der2"])
@pytest.mark.parametrize("npaire", [0, 1])
def test_derivative_host_arrays(ctx_host, dtype, ax, kind, npaire):
    _check_derivative(ctx_host, dtype, ax, kind, npaire, "host")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("ax", ["x", "y", "z"])
@pytest.mark.parametrize("op", ["dervp0", "dervp1", "intervp", "derpv", "interpv"])
def test_staggered_host_arrays(ctx_host, dtype, ax, op):
    _check_staggered(ctx_host, dtype, ax, op, "host")
