"""CPU oracle for the Xcompact3d hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy FP64, the reference algorithms that the HIP
library replaces (schemes.f90, derive.f90, poisson.f90, navier.f90,
transeq.f90, time_integrators.f90, Case-TGV.f90).  It exists only to check
the GPU path: it may be imported by ``tests/``, by ``__graft_entry__.smoke()``
and by ``bench.py``'s CPU-baseline leg, and by nothing else.  The product
(``incompact3d_amd``) never imports it and has no CPU fallback.

Parity status of each function is recorded in the docstring of
``oracle.x3d_oracle``.
"""
