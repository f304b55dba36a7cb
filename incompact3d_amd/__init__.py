"""MI355X (gfx950) implementation of the Xcompact3d per-timestep hot path.

The compute lives in ``libx3d.so`` (HIP, C-ABI in ``include/x3d.h``); this
package is the host-side mirror of the reference interface:

* :func:`init` / :func:`finalize`  -- schemes() + decomp_2d_poisson_init()
* :data:`coef`                      -- the module coefficient arrays/scalars
* :mod:`.derive`                    -- derx_00 ... interzpv (derive.f90)
* :func:`poisson`                   -- poisson_000 / poisson_010 (poisson.f90)
* :func:`transpose_x_to_y` ...      -- 2DECOMP pencil transposes
* :class:`Solver`                   -- the device-resident time step
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import Params, X3DError, check, iref, make_params
from . import derive  # noqa: F401

__all__ = ["Params", "X3DError", "make_params", "init", "finalize", "coef", "poisson",
           "transpose_x_to_y", "transpose_y_to_z", "transpose_z_to_y", "transpose_y_to_x",
           "reorder_ji", "Solver", "derive", "device_name", "schemes_host", "set_stream",
           "synchronize", "line_op_async"]


def init(p: Params) -> None:
    """x3d_init: schemes() (schemes.f90:7) and decomp_2d_poisson_init() (poisson.f90:73)."""
    check(_lib.lib().x3d_init(C.byref(p)), "x3d_init")


def finalize() -> None:
    check(_lib.lib().x3d_finalize(), "x3d_finalize")


def set_stream(handle) -> None:
    """x3d_set_stream: enqueue all library work on this HIP stream handle
    (e.g. ``torch.cuda.Stream().cuda_stream``); 0/None = default stream."""
    check(_lib.lib().x3d_set_stream(C.c_void_p(int(handle or 0))), "x3d_set_stream")


def synchronize() -> None:
    check(_lib.lib().x3d_synchronize(), "x3d_synchronize")


OPS = {"der1": 0, "der2": 1, "dervp": 2, "intervp": 3, "derpv": 4, "interpv": 5}


def line_op_async(dir_: int, op: str, npaire: int, u, t) -> None:
    """x3d_line_op_async: one line operator (library coefficient sets) on the
    library stream, no synchronisation.  u, t: device tensors (nz, ny, nx)."""
    import torch
    prec = 0 if u.dtype == torch.float64 else 1
    n3, n2, n1 = u.shape
    check(_lib.lib().x3d_line_op_async(int(dir_), OPS[op], int(npaire), prec, derive._dptr(u),
                                       derive._dptr(t), n1, n2, n3), "line_op_async")


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(_lib.lib().x3d_device_name(buf, 256), "x3d_device_name")
    return buf.value.decode()


class _Coef:
    """Attribute access to the coefficient state of the initialised context,
    by reference variable name: ``coef.ffx``, ``coef.cifxp6``, ``coef.afix``."""

    def __getattr__(self, name):
        lib = _lib.lib()
        n = C.c_int(0)
        if lib.x3d_get_coefficients(name.encode(), None, 0, C.byref(n)) == 0:
            out = np.empty(n.value, dtype=np.float64)
            check(lib.x3d_get_coefficients(name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)),
                                           n.value, C.byref(n)), name)
            return out
        v = C.c_double(0.0)
        check(lib.x3d_get_scalar(name.encode(), C.byref(v)), name)
        return v.value


coef = _Coef()


def schemes_host(p: Params, name: str):
    """Host-only evaluation of one coefficient (array or scalar) for ``p``."""
    lib = _lib.lib()
    n = C.c_int(0)
    check(lib.x3d_schemes_host(C.byref(p), name.encode(), None, 0, C.byref(n)), name)
    out = np.empty(max(n.value, 1), dtype=np.float64)
    check(lib.x3d_schemes_host(C.byref(p), name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)),
                               out.size, C.byref(n)), name)
    return out if n.value > 1 else float(out[0])


def _dims(t):
    nz, ny, nx = t.shape
    return nx, ny, nz


def poisson(rhs) -> None:
    """poisson(rhs): the procedure pointer of poisson.f90:57-65 (in place).
    rhs is a device tensor on the pressure grid, shape (nzm, nym, nxm)."""
    import torch
    n1, n2, n3 = _dims(rhs)
    fn = _lib.lib().x3d_poisson if rhs.dtype == torch.float64 else _lib.lib().x3d_poisson_f32
    check(fn(derive._dptr(rhs), iref(n1), iref(n2), iref(n3)), "poisson")


def poisson_async(rhs) -> None:
    """x3d_poisson_async: poisson(rhs) queued on the library stream, no wait."""
    import torch
    n1, n2, n3 = _dims(rhs)
    check(_lib.lib().x3d_poisson_async(0 if rhs.dtype == torch.float64 else 1, derive._dptr(rhs), n1, n2, n3),
          "poisson_async")


def fft3d_r2c(f):
    """decomp_2d_fft_3d(in, out) forward: f real (nzm, nym, nxm) device
    tensor -> complex spectrum (nzm/2+1, nym, nxm) (the reference's
    (nxm, nym, nzm/2+1) Fortran array), unnormalised, exp(-i k.x)."""
    import torch
    n1, n2, n3 = _dims(f)
    cdt = torch.complex128 if f.dtype == torch.float64 else torch.complex64
    out = torch.empty((n3 // 2 + 1, n2, n1), dtype=cdt, device=f.device)
    fn = _lib.lib().x3d_fft3d_r2c if f.dtype == torch.float64 else _lib.lib().x3d_fft3d_r2c_f32
    check(fn(derive._dptr(f), C.c_void_p(out.data_ptr()), iref(n1), iref(n2), iref(n3)), "fft3d_r2c")
    return out


def fft3d_c2r(spec, n3: int):
    """decomp_2d_fft_3d(in, out) backward: the (n3/2+1, n2, n1) complex
    spectrum of a real field -> real (n3, n2, n1), unnormalised, exp(+i k.x)."""
    import torch
    _, n2, n1 = spec.shape
    rdt = torch.float64 if spec.dtype == torch.complex128 else torch.float32
    out = torch.empty((n3, n2, n1), dtype=rdt, device=spec.device)
    fn = _lib.lib().x3d_fft3d_c2r if rdt == torch.float64 else _lib.lib().x3d_fft3d_c2r_f32
    check(fn(C.c_void_p(spec.data_ptr()), derive._dptr(out), iref(n1), iref(n2), iref(n3)), "fft3d_c2r")
    return out


def poisson_async_noalloc(rhs) -> None:
    """x3d_poisson_async_noalloc: the enqueue-only solve with the library's
    fixed device scratch (no allocation)."""
    n1, n2, n3 = _dims(rhs)
    check(_lib.lib().x3d_poisson_async_noalloc(0 if rhs.dtype == torch.float64 else 1, derive._dptr(rhs), n1,
                                                n2, n3), "x3d_poisson_async_noalloc")


def transpose_async(src, dst) -> None:
    """x3d_transpose_async: any of the four pencil transposes, queued only."""
    import torch
    n1, n2, n3 = _dims(src)
    check(_lib.lib().x3d_transpose_async(0 if src.dtype == torch.float64 else 1, derive._dptr(src),
                                          derive._dptr(dst), n1, n2, n3), "transpose_async")


def _transpose(name):
    def f(src, dst):
        import torch
        n1, n2, n3 = _dims(src)
        sfx = "" if src.dtype == torch.float64 or name == "reorder_ji" else "_f32"
        check(getattr(_lib.lib(), f"x3d_{name}{sfx}")(derive._dptr(src), derive._dptr(dst), iref(n1),
                                                       iref(n2), iref(n3)), name)
    f.__name__ = name
    return f


transpose_x_to_y = _transpose("transpose_x_to_y")
transpose_y_to_z = _transpose("transpose_y_to_z")
transpose_z_to_y = _transpose("transpose_z_to_y")
transpose_y_to_x = _transpose("transpose_y_to_x")
reorder_ji = _transpose("reorder_ji")


class Solver:
    """Device-resident TGV/periodic solver (x3d_solver_*).

    precision 0: FP64 everywhere.  precision 1: FP32 bulk with FP64 velocity
    accumulation, FP64 Poisson divide and FP64 reductions."""

    def __init__(self, p: Params, precision: int = 0):
        self._h = C.c_void_p()
        check(_lib.lib().x3d_solver_create(int(precision), C.byref(self._h)), "x3d_solver_create")
        self.precision = precision
        self._shape = (p.nz, p.ny, p.nx)
        self._dt = p.dt

    def close(self):
        if self._h:
            _lib.lib().x3d_solver_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_tgv(self):
        check(_lib.lib().x3d_solver_init_tgv(self._h), "init_tgv")

    def set_velocity(self, ux, uy, uz):
        check(_lib.lib().x3d_solver_set_velocity(self._h, derive._dptr(ux), derive._dptr(uy),
                                                 derive._dptr(uz)), "set_velocity")

    def velocity(self):
        """FP64 copies (ux, uy, uz) of the solver velocity, device tensors."""
        import torch
        out = [torch.empty(self._shape, dtype=torch.float64, device="cuda") for _ in range(3)]
        check(_lib.lib().x3d_solver_get_velocity(self._h, *[derive._dptr(t) for t in out]),
              "get_velocity")
        return out

    def step(self, nsteps: int = 1):
        check(_lib.lib().x3d_solver_step(self._h, int(nsteps)), "step")

    def tgv_stats(self):
        """(eek, eps, eps2, enst) of postprocess_tgv."""
        out = (C.c_double * 4)()
        check(_lib.lib().x3d_solver_tgv_stats(self._h, out), "tgv_stats")
        return tuple(out)

    def divergence_stats(self):
        out = (C.c_double * 2)()
        check(_lib.lib().x3d_solver_divergence_stats(self._h, out), "divergence_stats")
        return tuple(out)

    def divu_stats(self):
        """(max, mean |.|) of div u* before the last pressure correction, as
        navier.f90:348-361 computes them on every call of divergence."""
        out = (C.c_double * 2)()
        check(_lib.lib().x3d_solver_divu_stats(self._h, out), "divu_stats")
        return out[0], out[1]

    def momentum_yz(self, nrep: int = 1) -> float:
        """x3d_solver_momentum_yz: enqueue nrep times the six fused y/z
        momentum sweeps of one step on the library stream; returns the HBM
        bytes one launch moves on average."""
        b = C.c_double(0.0)
        check(_lib.lib().x3d_solver_momentum_yz(self._h, int(nrep), C.byref(b)), "momentum_yz")
        return b.value

    def init_channel(self, init_noise: float = 0.125, seed: int = 20261202 % (1 << 24)):
        """init_channel (Case-Channel.f90:96-128) with counter-based noise."""
        check(_lib.lib().x3d_solver_init_channel(self._h, float(init_noise), int(seed)), "init_channel")

    def set_forcing(self, mode: int, value: float = 2.0 / 3.0):
        """0 none, 1 constant flow rate (channel_cfr), 2 constant pressure gradient."""
        check(_lib.lib().x3d_solver_set_forcing(self._h, int(mode), float(value)), "set_forcing")

    def set_rotation(self, wrotation: float, spinup_time: int):
        """Spin-up rotation of momentum_forcing_channel (Case-Channel.f90:413-418):
        while itime < spinup_time, dux -= wrotation uy, duy += wrotation ux."""
        check(_lib.lib().x3d_solver_set_rotation(self._h, float(wrotation), int(spinup_time)), "set_rotation")

    def checkpoint(self, rundir: str, t: float = None):
        """restart(iresflg=1) of tools.f90: <rundir>/checkpoint/{ux,uy,uz,dux-2,...,pp}
        and <rundir>/restart.info; t defaults to itime * dt."""
        if t is None:
            t = self.itime * self._dt
        check(_lib.lib().x3d_solver_checkpoint(self._h, os.fsencode(rundir), float(t)), "checkpoint")

    def restart(self, rundir: str) -> float:
        """restart(iresflg=0): the state written by checkpoint; returns tfield."""
        t = C.c_double(0.0)
        check(_lib.lib().x3d_solver_restart(self._h, os.fsencode(rundir), C.byref(t)), "restart")
        return t.value

    def bulk_velocity(self) -> float:
        v = C.c_double()
        check(_lib.lib().x3d_solver_bulk_velocity(self._h, C.byref(v)), "bulk_velocity")
        return v.value

    @property
    def itime(self):
        v = C.c_int(0)
        check(_lib.lib().x3d_solver_itime(self._h, C.byref(v)), "itime")
        return v.value

    def set_les(self, model: int, cs: float = 0.15, iconserv: int = 0, beta: float = 0.0, mlp=None,
                gscale: float = 1.0, delta0: float = 1.0):
        """x3d_solver_set_les: 0 none, 1 Smagorinsky (jles=1, smagcst=cs,
        iconserv), 2 the learned closure ``mlp`` (an Mlp) with the Smagorinsky
        fallback of constant cs and bound beta."""
        self._mlp = mlp   # keep the network alive while the solver uses it
        check(_lib.lib().x3d_solver_set_les(self._h, int(model), float(cs), int(iconserv), float(beta),
                                            mlp._h if mlp is not None else None, float(gscale),
                                            float(delta0)), "set_les")

    def sgs(self):
        """SGS force (sgsx, sgsy, sgsz) of the current velocity, FP64 device tensors."""
        import torch
        out = [torch.empty(self._shape, dtype=torch.float64, device="cuda") for _ in range(3)]
        check(_lib.lib().x3d_solver_sgs(self._h, *[derive._dptr(t) for t in out]), "sgs")
        return out

    def les_fallbacks(self):
        v = C.c_longlong(0)
        check(_lib.lib().x3d_solver_les_fallbacks(self._h, C.byref(v)), "les_fallbacks")
        return v.value


def _ptr_array(ts):
    arr = (C.c_void_p * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = derive._dptr(t)
    return arr


class Mlp:
    """The closure network (x3d_mlp_*): 10-64-64-64-6, GELU, BF16 MFMA."""

    def __init__(self, params=None):
        self._h = C.c_void_p()
        check(_lib.lib().x3d_mlp_create(C.byref(self._h)), "x3d_mlp_create")
        self.nparams = _lib.lib().x3d_mlp_nparams()
        if params is not None:
            self.set_params(params)

    def close(self):
        if self._h:
            _lib.lib().x3d_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        import numpy as np
        a = np.ascontiguousarray(np.asarray(flat, dtype=np.float32).ravel())
        if a.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {a.size}")
        check(_lib.lib().x3d_mlp_set_params(self._h, a.ctypes.data), "mlp_set_params")

    def get_params(self):
        import numpy as np
        a = np.empty(self.nparams, dtype=np.float32)
        check(_lib.lib().x3d_mlp_get_params(self._h, a.ctypes.data), "mlp_get_params")
        return a

    def forward(self, x):
        """x: (n, 10) FP32 device tensor -> (n, 6)."""
        import torch
        y = torch.empty((x.shape[0], 6), dtype=torch.float32, device=x.device)
        check(_lib.lib().x3d_mlp_forward(self._h, derive._dptr(x), derive._dptr(y), int(x.shape[0])),
              "mlp_forward")
        return y

    def train_step(self, x, t, lr: float):
        loss = C.c_float(0.0)
        check(_lib.lib().x3d_mlp_train_step(self._h, derive._dptr(x), derive._dptr(t), int(x.shape[0]),
                                            float(lr), C.byref(loss)), "mlp_train_step")
        return loss.value

    def gradient(self, x, t):
        import numpy as np
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_float(0.0)
        check(_lib.lib().x3d_mlp_gradient(self._h, derive._dptr(x), derive._dptr(t), int(x.shape[0]),
                                          g.ctypes.data, C.byref(loss)), "mlp_gradient")
        return g, loss.value

    def closure(self, g, gscale, delta, delta0, cs, beta, project: int = 0):
        """T_ij (6 tensors) from the nine gradient tensors g (FP64 or FP32);
        project=1 is the eddy-viscosity projection of LES model 3."""
        import torch
        t = [torch.empty_like(g[0]) for _ in range(6)]
        nfb = C.c_longlong(0)
        prec = 0 if g[0].dtype == torch.float64 else 1
        check(_lib.lib().x3d_mlp_closure(self._h, prec, _ptr_array(g), _ptr_array(t), int(g[0].numel()),
                                         float(gscale), float(delta), float(delta0), float(cs), float(beta),
                                         int(project), C.byref(nfb)), "mlp_closure")
        return t, nfb.value


def box_filter(ux, uy, uz, f: int):
    """x3d_box_filter: filtered velocity (3) and tau_ij (6: xx yy zz xy xz yz)
    on the (nz/f, ny/f, nx/f) grid; FP64 device tensors in and out."""
    import torch
    nz, ny, nx = ux.shape
    shape = (nz // f, ny // f, nx // f)
    ub = [torch.empty(shape, dtype=torch.float64, device=ux.device) for _ in range(3)]
    tau = [torch.empty(shape, dtype=torch.float64, device=ux.device) for _ in range(6)]
    check(_lib.lib().x3d_box_filter(derive._dptr(ux), derive._dptr(uy), derive._dptr(uz), nx, ny, nz, int(f),
                                    _ptr_array(ub), _ptr_array(tau)), "box_filter")
    return ub, tau

