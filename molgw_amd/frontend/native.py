"""ctypes binding of the native integral engine (ints.cpp, built by
molgw_amd.build into libmgw_ints.so): (mu nu|P) and (P|Q) over the final
functions of BasisSet objects."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .basis import BasisSet

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libmgw_ints.so')
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f'{LIB} is not built: run molgw_amd.build.build()')
        lib = C.CDLL(LIB)
        ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
        shell_args = [C.c_int, ip, dp, ip, dp, dp, ip, dp, ip]
        lib.mgwi_eri3.argtypes = shell_args + shell_args + [dp]
        lib.mgwi_eri3.restype = C.c_int
        lib.mgwi_eri2.argtypes = shell_args + [dp]
        lib.mgwi_eri2.restype = C.c_int
        lib.mgwi_one_electron.argtypes = shell_args + [C.c_int, dp, dp, dp, dp, dp]
        lib.mgwi_one_electron.restype = C.c_int
        _lib = lib
    return _lib


def _flat(bs: BasisSet):
    """Keep-alive arrays and ctypes arguments of a BasisSet's shells."""
    sh = bs.shells
    l = np.array([s.l for s in sh], dtype=np.int32)
    center = np.ascontiguousarray(np.array([s.center for s in sh], dtype=np.float64).reshape(-1))
    poff = np.cumsum([0] + [len(s.exps) for s in sh]).astype(np.int32)
    exps = np.ascontiguousarray(np.concatenate([s.exps for s in sh]))
    coefs = np.ascontiguousarray(np.concatenate([s.coefs for s in sh]))
    tm = [np.ascontiguousarray(s.transform, dtype=np.float64).reshape(-1) for s in sh]
    toff = np.cumsum([0] + [len(t) for t in tm[:-1]]).astype(np.int32)
    tmat = np.ascontiguousarray(np.concatenate(tm))
    foff = np.asarray(bs.offsets, dtype=np.int32)
    keep = (l, center, poff, exps, coefs, toff, tmat, foff)
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    args = [len(sh), l.ctypes.data_as(ip), center.ctypes.data_as(dp), poff.ctypes.data_as(ip),
            exps.ctypes.data_as(dp), coefs.ctypes.data_as(dp), toff.ctypes.data_as(ip), tmat.ctypes.data_as(dp),
            foff.ctypes.data_as(ip)]
    return keep, args


def eri_3center(bs: BasisSet, aux: BasisSet) -> np.ndarray:
    lib = load()
    k1, a1 = _flat(bs)
    k2, a2 = _flat(aux)
    out = np.zeros((bs.nbf, bs.nbf, aux.nbf))
    rc = lib.mgwi_eri3(*a1, *a2, out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise ValueError('native integrals: angular momentum above 6')
    return out


def eri_2center(aux: BasisSet) -> np.ndarray:
    lib = load()
    k2, a2 = _flat(aux)
    out = np.zeros((aux.nbf, aux.nbf))
    rc = lib.mgwi_eri2(*a2, out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise ValueError('native integrals: angular momentum above 6')
    return out


def one_electron(bs: BasisSet, charges, coords):
    """(S, T, V) over the final functions."""
    lib = load()
    k1, a1 = _flat(bs)
    n = bs.nbf
    s, t, v = (np.zeros((n, n)) for _ in range(3))
    q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
    xyz = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).reshape(-1))
    dp = C.POINTER(C.c_double)
    rc = lib.mgwi_one_electron(*a1, len(q), q.ctypes.data_as(dp), xyz.ctypes.data_as(dp), s.ctypes.data_as(dp),
                               t.ctypes.data_as(dp), v.ctypes.data_as(dp))
    if rc != 0:
        raise ValueError('native integrals: angular momentum above 6')
    return s, t, v
