"""ctypes binding of the C++/OpenMP restatement of the VMC step (oracle/cpu_ref.cpp).

TEST INFRASTRUCTURE ONLY: the CPU baseline ("reference algorithm, CPU restatement") that bench.py
times beside the GPU, checked against the NumPy oracle by tests/test_cpu_ref.py.  ``build()``
compiles ``oracle/libcpuref.so`` with g++ and OpenMP; __graft_entry__.build() calls it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_ref.cpp")
LIB = os.path.join(HERE, "libcpuref.so")
# x86-64-v3 (AVX2 + FMA) runs on this container's and the GPU hosts' CPUs alike
CXXFLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17"]

_lib = None


def build() -> str:
    tmp = LIB + ".tmp"
    subprocess.run(["g++", *CXXFLAGS, "-o", tmp, SRC], check=True, capture_output=True, text=True)
    os.replace(tmp, LIB)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (python -c 'import oracle.cpu_ref as c; c.build()')")
        lib = ctypes.CDLL(LIB)
        lib.cpuref_vmc_step.restype = ctypes.c_int32
        P = ctypes.c_void_p
        lib.cpuref_vmc_step.argtypes = [P, ctypes.c_int64, P, P, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, P, P, P, P, P]
        _lib = lib
    return _lib


def vmc_step(desc, r_up, r_dn, n_moves, Dt, epsilon_AS, seed, move0=0, rot_step=0, n_threads=1):
    """One VMC step per walker (set-up, sweep, rotation, E_L) on the host.  ``desc`` is a
    jqmc_system_desc (jqmc_amd._pack.SystemPack(..., device=False).desc).  Returns
    (e_L (W,), r_up, r_dn, accepted (W,), seconds of the sweeps and local energies)."""
    lib = load()
    ru = np.ascontiguousarray(r_up, dtype=np.float64)
    rd = np.ascontiguousarray(r_dn, dtype=np.float64)
    W = ru.shape[0]
    e = np.zeros(W)
    ou, od = np.zeros_like(ru), np.zeros_like(rd)
    acc = np.zeros(W, dtype=np.int64)
    sec = ctypes.c_double()
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = lib.cpuref_vmc_step(ctypes.byref(desc), W, ptr(ru), ptr(rd), int(n_moves), float(Dt), float(epsilon_AS),
                             int(seed) & 0xFFFFFFFFFFFFFFFF, int(move0), int(rot_step), int(n_threads), ptr(e),
                             ptr(ou), ptr(od), ptr(acc), ctypes.byref(sec))
    if rc != 0:
        raise RuntimeError(f"cpuref_vmc_step failed ({rc})")
    return e, ou, od, acc, sec.value
