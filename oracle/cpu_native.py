"""ctypes loader of oracle/liblres_cpu.so: the C++/OpenMP CPU restatement of the
eigen-route and pole-route window passes (oracle/lres_cpu.cpp), with LAPACK
taken from scipy's cython_lapack (dsygvd, dsygvx, zgesv function pointers).

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's cpu_baseline leg); the product
never loads it.  ``build()`` compiles it with g++ -O3 -fopenmp."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "lres_cpu.cpp")
LIB = os.path.join(HERE, "liblres_cpu.so")

_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)
_vp = C.c_void_p
_lib = None


def build():
    """Compile liblres_cpu.so in place (always)."""
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-march=x86-64-v2", "-fopenmp", "-shared", "-fPIC", "-std=c++17", SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


def _lapack(name):
    from scipy.linalg import cython_lapack as cl
    cap = cl.__pyx_capi__[name]
    C.pythonapi.PyCapsule_GetName.restype = C.c_char_p
    C.pythonapi.PyCapsule_GetName.argtypes = [C.py_object]
    C.pythonapi.PyCapsule_GetPointer.restype = C.c_void_p
    C.pythonapi.PyCapsule_GetPointer.argtypes = [C.py_object, C.c_char_p]
    return C.pythonapi.PyCapsule_GetPointer(cap, C.pythonapi.PyCapsule_GetName(cap))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} not built: run oracle.cpu_native.build()")
        L = C.CDLL(LIB)
        L.lres_cpu_eig_pass.restype = C.c_int
        L.lres_cpu_eig_pass.argtypes = [_vp, _vp] + [_i64p, _i64p, _f64p] * 3 + [
            _i64p, _i64p, C.c_int64, _i64p, _i64p, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int,
            C.c_double, C.c_int, C.c_int, _f64p, _f64p, _i32p, _f64p, _f32p, _i64p]
        L.lres_cpu_pole_pass.restype = C.c_int
        L.lres_cpu_pole_pass.argtypes = [_vp] + [_i64p, _i64p, _f64p] * 3 + [
            _i64p, _i64p, C.c_int64, _i64p, _i64p, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int,
            C.c_int, _f64p, _f64p, _f64p]
        L.lres_cpu_max_threads.restype = C.c_int
        L._dsygvd, L._dsygvx, L._zgesv = _lapack("dsygvd"), _lapack("dsygvx"), _lapack("zgesv")
        _lib = L
    return _lib


def _csr(A):
    A = A.tocsr()
    A.sort_indices()
    return (np.ascontiguousarray(A.indptr, np.int64), np.ascontiguousarray(A.indices, np.int64),
            np.ascontiguousarray(A.data, np.float64))


def _p(a, t):
    return a.ctypes.data_as(t)


class Tables:
    """Flattened window tables (Wflat/Woff/Oflat/Ooff, lcao3d.py:343-357) and the
    fixed CSR data of a basis, as the native passes take them."""

    def __init__(self, basis, windows, own):
        self.Wflat = np.ascontiguousarray(np.concatenate(windows), np.int64)
        self.Woff = np.concatenate([[0], np.cumsum([len(w) for w in windows])]).astype(np.int64)
        self.Oflat = np.ascontiguousarray(np.concatenate(own), np.int64)
        self.Ooff = np.concatenate([[0], np.cumsum([len(o) for o in own])]).astype(np.int64)
        self.nwin = len(windows)
        self.nb, self.ng, self.dV = int(basis.nb), int(basis.cr.ng), float(basis.cr.dV)
        self.S, self.X = _csr(basis.S), _csr(basis.Xr)
        nW = np.diff(self.Woff)
        nown = np.diff(self.Ooff)
        self.y2off = np.concatenate([[0], np.cumsum(nW * nown)]).astype(np.int64)


def _threadpools_single():
    from threadpoolctl import threadpool_limits
    return threadpool_limits(1)      # LAPACK single-threaded inside every OpenMP thread


def eig_pass(tab, H, mu, e_lo, M, upper=None, want_density=True, nthreads=1, keep_y2=False):
    """lrft_windows' eigen route natively: (dens or None, lam, c, lamoff, y2 or None)
    with the layout of oracle.lres_oracle.eig_window_pass (y2: list of float32
    (nown, m) arrays)."""
    L = lib()
    Hc = _csr(H)
    ntot = int(tab.Woff[-1])
    lam = np.zeros(ntot)
    c = np.zeros(ntot)
    m = np.zeros(tab.nwin, np.int32)
    dens = np.zeros(tab.ng) if want_density else None
    y2 = np.zeros(int(tab.y2off[-1]), np.float32) if keep_y2 else None
    up = np.inf if upper is None else float(upper)
    with _threadpools_single():
        rc = L.lres_cpu_eig_pass(L._dsygvd, L._dsygvx, *(_p(a, t) for a, t in zip(Hc, (_i64p, _i64p, _f64p))),
                                 *(_p(a, t) for a, t in zip(tab.S, (_i64p, _i64p, _f64p))),
                                 *(_p(a, t) for a, t in zip(tab.X, (_i64p, _i64p, _f64p))),
                                 _p(tab.Wflat, _i64p), _p(tab.Woff, _i64p), tab.nwin, _p(tab.Oflat, _i64p),
                                 _p(tab.Ooff, _i64p), tab.nb, tab.dV, float(mu), float(e_lo), int(M), up,
                                 int(want_density), int(nthreads), _p(lam, _f64p), _p(c, _f64p), _p(m, _i32p),
                                 _p(dens, _f64p) if dens is not None else None,
                                 _p(y2, _f32p) if y2 is not None else None, _p(tab.y2off, _i64p))
    if rc != 0:
        raise RuntimeError(f"dsygvd/dsygvx failed (info {rc})")
    nW = np.diff(tab.Woff)
    lams, cs, ys = [], [], []
    for w in range(tab.nwin):
        l0 = tab.Woff[w] - tab.Woff[0]
        lams.append(lam[l0:l0 + m[w]])
        cs.append(c[l0:l0 + m[w]])
        if keep_y2:
            no = tab.Ooff[w + 1] - tab.Ooff[w]
            ys.append(y2[tab.y2off[w]:tab.y2off[w + 1]].reshape(no, nW[w])[:, :m[w]])
    lamoff = np.concatenate([[0], np.cumsum(m)]).astype(np.int64)
    return dens, np.concatenate(lams), np.concatenate(cs), lamoff, (ys if keep_y2 else None)


def pole_pass(tab, H, mu, e_lo, M, want_density=True, nthreads=1):
    """The pole route natively: (dens or None, N_B, E_B), no spin factor."""
    L = lib()
    Hc = _csr(H)
    dens = np.zeros(tab.ng) if want_density else None
    NB, EB = np.zeros(tab.nwin), np.zeros(tab.nwin)
    with _threadpools_single():
        rc = L.lres_cpu_pole_pass(L._zgesv, *(_p(a, t) for a, t in zip(Hc, (_i64p, _i64p, _f64p))),
                                  *(_p(a, t) for a, t in zip(tab.S, (_i64p, _i64p, _f64p))),
                                  *(_p(a, t) for a, t in zip(tab.X, (_i64p, _i64p, _f64p))),
                                  _p(tab.Wflat, _i64p), _p(tab.Woff, _i64p), tab.nwin, _p(tab.Oflat, _i64p),
                                  _p(tab.Ooff, _i64p), tab.nb, tab.dV, float(mu), float(e_lo), int(M),
                                  int(nthreads), _p(dens, _f64p) if dens is not None else None, _p(NB, _f64p),
                                  _p(EB, _f64p))
    if rc != 0:
        raise RuntimeError(f"zgesv failed (info {rc})")
    return dens, NB, EB
