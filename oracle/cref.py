"""ctypes loader of the reference's own C++ oracle, built by oracle/Makefile into oracle/_ref/.

TEST INFRASTRUCTURE ONLY.  Binds the two entry points of
experiments/session_a/inputs/reference.cc:12-39 the way the reference harness does
(experiments/session_a/dependencies/common.py:57-69):
  element(h, nu, ke[576])                                  PETSc Hex8Isoparametric, cube of side h
  product(nx, ny, nz, rho, u, ke, mask, emin, emax, penal, y, ce, diag)   natural layout
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_ref" / "reference.so"


def available():
    return LIB.exists()


class Reference:
    def __init__(self, library=LIB):
        self.lib = ctypes.CDLL(str(library))
        arr = np.ctypeslib.ndpointer(dtype=np.float64, ndim=1, flags="C_CONTIGUOUS")
        self.lib.product.argtypes = [ctypes.c_int] * 3 + [arr] * 4 + [ctypes.c_double] * 3 + [arr] * 3
        self.lib.product.restype = None
        self.lib.element.argtypes = [ctypes.c_double, ctypes.c_double, arr]
        self.lib.element.restype = None

    def element(self, h, nu):
        ke = np.zeros(576)
        self.lib.element(float(h), float(nu), ke)
        return ke.reshape(24, 24)

    def product(self, shape, rho, u, ke, mask, emin, emax=1.0, penal=3.0):
        nx, ny, nz = (int(s) for s in shape)
        u = np.ascontiguousarray(u, dtype=np.float64)
        y, ce, d = np.empty_like(u), np.empty(nx * ny * nz), np.empty_like(u)
        self.lib.product(nx, ny, nz, np.ascontiguousarray(rho, dtype=np.float64), u,
                         np.ascontiguousarray(ke, dtype=np.float64).ravel(),
                         np.ascontiguousarray(mask, dtype=np.float64),
                         float(emin), float(emax), float(penal), y, ce, d)
        return y, ce, d
