"""ctypes binding of oracle/cport/tgv_step.cpp: the OpenMP CPU restatement
of one periodic time step (test infrastructure and the CPU baseline of the
benchmark, never the product).  The coefficient sets, Sherman-Morrison
vectors and the modified-wavenumber table come from the oracle's Schemes.

    python -m oracle.cport_step        # builds oracle/cport/_build/libtgvstep.so
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import x3d_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cport", "tgv_step.cpp")
BUILD = os.path.join(HERE, "cport", "_build")
LIB = os.path.join(BUILD, "libtgvstep.so")


def build() -> str:
    """g++ -O3 -fopenmp: a generic x86-64 build (no -march), so that the same
    library runs on any host it is copied to."""
    os.makedirs(BUILD, exist_ok=True)
    tmp = LIB + ".tmp"
    subprocess.run(["g++", "-O3", "-fopenmp", "-std=c++17", "-shared", "-fPIC", SRC, "-o", tmp], check=True)
    os.replace(tmp, LIB)
    return LIB


class _Dir(C.Structure):
    _fields_ = [("sc", C.c_double * 12), ("arr", (C.c_void_p * 4) * 6), ("alfa", C.c_double * 6)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.cport_create.restype = C.c_void_p
        _lib.cport_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(_Dir),
                                      C.c_void_p]
        _lib.cport_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        _lib.cport_threads.restype = C.c_int
        _lib.cport_destroy.argtypes = [C.c_void_p]
    return _lib


class CportStep:
    """AB3 steps of the periodic solver on host arrays (nz, ny, nx)."""

    def __init__(self, S: O.Schemes):
        p = S.p
        if not (p.nclx and p.ncly and p.nclz):
            raise ValueError("the CPU restatement covers triply periodic boxes")
        self.p = p
        self._keep = []
        dirs = (_Dir * 3)()
        for d, ax in enumerate("xyz"):
            d1, d2, st = S.d1[ax], S.d2[ax], S.st[ax]
            sc = [d1["afi"], d1["bfi"], d2["asi"], d2["bsi"], d2["csi"], d2["dsi"],
                  st["aci"], st["bci"], st["aici"], st["bici"], st["cici"], st["dici"]]
            for i, v in enumerate(sc):
                dirs[d].sc[i] = float(v)
            sets = [(d1["ff"], d1["fs"], d1["fw"], d1["alfai"]),
                    (d2["sf"], d2["ss"], d2["sw"], d2["alsai"]),
                    (st["cfx6"], st["csx6"], st["cwx6"], st["alcai"]),
                    (st["cifxp6"], st["cisxp6"], st["ciwxp6"], st["ailcai"]),
                    (st["cfx6"], st["csx6"], st["cwx6"], st["alcai"]),
                    (st["cifx6"], st["cisx6"], st["ciwx6"], st["ailcai"])]
            for o, (ff, fs, fw, alfa) in enumerate(sets):
                arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (ff, fs, fw)]
                arrs.append(np.ascontiguousarray(O.sm_vector(arrs[0], arrs[1], arrs[2], alfa)))
                self._keep += arrs
                for k in range(4):
                    dirs[d].arr[o][k] = arrs[k].ctypes.data
                dirs[d].alfa[o] = float(alfa)
        self.K = np.ascontiguousarray(O.waves_kxyz(S), dtype=np.float64)
        self._dirs = dirs
        self.h = lib().cport_create(p.nx, p.ny, p.nz, p.xnu, p.dt, dirs, self.K.ctypes.data)

    @staticmethod
    def threads() -> int:
        return lib().cport_threads()

    def step(self, ux, uy, uz, nsteps=1):
        for a in (ux, uy, uz):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        lib().cport_step(self.h, ux.ctypes.data, uy.ctypes.data, uz.ctypes.data, int(nsteps))

    def close(self):
        if self.h:
            lib().cport_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


if __name__ == "__main__":
    print(build())
