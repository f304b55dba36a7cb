"""Host-side mirror of derive.f90: the same subroutine names and argument
order, executed by libx3d on the GPU.

Fields (t, u) are device tensors holding Fortran-ordered arrays, i.e. torch
tensors of shape (nz, ny, nx) in C order (torch.float64 -> FP64 entry points,
torch.float32 -> the ``_f32`` entry points), or numpy arrays of the same
shape in host memory, which go to the ``x3d_host_*`` twins (staged through
HBM).  Coefficient arrays (ffx, fsx, fwx, cfx6, ...) are host arrays, as the
reference passes its module arrays; ``r`` and ``s`` are the reference's
scratch arguments and may be None.

Example, as in transeq.f90:125::

    derx(td1, ta1, None, None, c.ffxp, c.fsxp, c.fwxp, nx, ny, nz, 1)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, iref


def _dptr(t):
    if t is None:
        return None
    if not getattr(t, "is_cuda", False):
        raise TypeError("fields must be device tensors (torch, device='cuda')")
    if not t.is_contiguous():
        raise ValueError("fields must be contiguous (Fortran order = C order of (nz, ny, nx))")
    return C.c_void_p(t.data_ptr())


def _is_host(t):
    return isinstance(t, np.ndarray)


def _fptr(t, host):
    """Pointer to a field: a C-contiguous numpy array for the host twins, a
    device tensor otherwise (r and s may be None)."""
    if t is None:
        return None
    if host:
        if not isinstance(t, np.ndarray) or not t.flags.c_contiguous:
            raise TypeError("host fields must be C-contiguous numpy arrays of shape (nz, ny, nx)")
        return C.c_void_p(t.ctypes.data)
    return _dptr(t)


def _entry(name, sfx, host):
    return getattr(_lib.lib(), f"x3d_{'host_' if host else ''}{name}{sfx}")


class _HostArrays:
    """Keeps host coefficient arrays alive for the duration of one call."""

    def __init__(self, dtype):
        self.dtype = dtype
        self.keep = []

    def __call__(self, a):
        if a is None:
            return None
        arr = np.ascontiguousarray(np.asarray(a, dtype=self.dtype))
        self.keep.append(arr)
        return C.c_void_p(arr.ctypes.data)


def _suffix(t):
    if _is_host(t):
        if t.dtype == np.float64:
            return "", np.float64, C.c_double
        if t.dtype == np.float32:
            return "_f32", np.float32, C.c_float
        raise TypeError("fields must be float64 or float32")
    import torch
    if t.dtype == torch.float64:
        return "", np.float64, C.c_double
    if t.dtype == torch.float32:
        return "_f32", np.float32, C.c_float
    raise TypeError("fields must be float64 or float32")


def _scalar_ref(v, ctype):
    return C.byref(ctype(0.0 if v is None else float(v)))


def _der_x(name):
    def f(t, u, r, s, ff, fs, fw, nx, ny, nz, npaire, lind=0.0):
        sfx, dt, ct = _suffix(t)
        h = _HostArrays(dt)
        host = _is_host(t)
        fn = _entry(name, sfx, host)
        fields = (_fptr(t, host), _fptr(u, host), _fptr(r, host), _fptr(s, host))
        check(fn(*fields, h(ff), h(fs), h(fw), iref(nx), iref(ny), iref(nz), iref(npaire), _scalar_ref(lind, ct)),
              name)
    f.__name__ = name
    f.__doc__ = f"{name}(t,u,r,s,ff,fs,fw,nx,ny,nz,npaire,lind) -- derive.f90 {name}"
    return f


def _der_y(name):
    def f(t, u, r, s, ff, fs, fw, pp, nx, ny, nz, npaire, lind=0.0):
        sfx, dt, ct = _suffix(t)
        h = _HostArrays(dt)
        host = _is_host(t)
        fn = _entry(name, sfx, host)
        fields = (_fptr(t, host), _fptr(u, host), _fptr(r, host), _fptr(s, host))
        check(fn(*fields, h(ff), h(fs), h(fw), h(pp), iref(nx), iref(ny), iref(nz), iref(npaire),
                 _scalar_ref(lind, ct)), name)
    f.__name__ = name
    f.__doc__ = f"{name}(t,u,r,s,ff,fs,fw,pp,nx,ny,nz,npaire,lind) -- derive.f90 {name}"
    return f


_g = globals()
for _d in ("x", "z", "xx", "yy", "zz"):
    for _bc in ("00", "11", "12", "21", "22"):
        _g[f"der{_d}_{_bc}"] = _der_x(f"der{_d}_{_bc}")
for _bc in ("00", "11", "12", "21", "22"):
    _g[f"dery_{_bc}"] = _der_y(f"dery_{_bc}")


def _vp(name, with_pp=False):
    # (t,u,r,s,cf,cs,cw,[pp],n1,n2,n3,n4,npaire)
    def f(t, u, r, s, cf, cs, cw, *rest):
        sfx, dt, _ = _suffix(t)
        h = _HostArrays(dt)
        host = _is_host(t)
        fn = _entry(name, sfx, host)
        if with_pp:
            pp, n1, n2, n3, n4, npaire = rest
            args = [h(pp)]
        else:
            n1, n2, n3, n4, npaire = rest
            args = []
        fields = (_fptr(t, host), _fptr(u, host), _fptr(r, host), _fptr(s, host))
        check(fn(*fields, h(cf), h(cs), h(cw), *args, iref(n1), iref(n2), iref(n3), iref(n4), iref(npaire)), name)
    f.__name__ = name
    return f


def _pv(name, with_pp=False):
    # (t,u,r,s,cf1,cs1,cw1,cf2,cs2,cw2,[pp],n1,n2,n3,n4,npaire)
    def f(t, u, r, s, cf1, cs1, cw1, cf2, cs2, cw2, *rest):
        sfx, dt, _ = _suffix(t)
        h = _HostArrays(dt)
        host = _is_host(t)
        fn = _entry(name, sfx, host)
        if with_pp:
            pp, n1, n2, n3, n4, npaire = rest
            args = [h(pp)]
        else:
            n1, n2, n3, n4, npaire = rest
            args = []
        fields = (_fptr(t, host), _fptr(u, host), _fptr(r, host), _fptr(s, host))
        check(fn(*fields, h(cf1), h(cs1), h(cw1), h(cf2), h(cs2), h(cw2), *args, iref(n1), iref(n2), iref(n3),
                 iref(n4), iref(npaire)), name)
    f.__name__ = name
    return f


derxvp = _vp("derxvp")          # derive.f90:3796
interxvp = _vp("interxvp")      # derive.f90:3911
derxpv = _pv("derxpv")          # derive.f90:4041
interxpv = _pv("interxpv")      # derive.f90:4126
interyvp = _vp("interyvp")      # derive.f90:4265
deryvp = _vp("deryvp", True)    # derive.f90:4442
interypv = _pv("interypv")      # derive.f90:4587
derypv = _pv("derypv", True)    # derive.f90:4775
derzvp = _vp("derzvp")          # derive.f90:4920
interzvp = _vp("interzvp")      # derive.f90:5105
derzpv = _pv("derzpv")          # derive.f90:5287
interzpv = _pv("interzpv")      # derive.f90:5426


def bind_procedures(p) -> dict:
    """The procedure-pointer binding of schemes.f90:31-66: returns
    {'derx': derx_<bc>, 'dery': ..., 'derzz': ...} for the BCs of ``p``."""
    def pick(kind, b1, bn):
        return _g[f"{kind}_{b1}{bn}"]
    return {
        "derx": pick("derx", p.nclx1, p.nclxn), "dery": pick("dery", p.ncly1, p.nclyn),
        "derz": pick("derz", p.nclz1, p.nclzn), "derxx": pick("derxx", p.nclx1, p.nclxn),
        "deryy": pick("deryy", p.ncly1, p.nclyn), "derzz": pick("derzz", p.nclz1, p.nclzn),
    }
