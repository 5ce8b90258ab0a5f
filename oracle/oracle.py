"""ctypes front end of the CPU FP64 oracle (liblowdin_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product package never imports this module.

Each function follows the reference routine named in its docstring; the C
side (lowdin_oracle.c) cites the reference lines.  Matrices come back as
numpy arrays in the reference's row-major Eigen convention (element (r, c)
at r*n + c), which for the symmetric results equals the Fortran image.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

DBL_EPSILON = float(np.finfo(np.float64).eps)


def build():
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblowdin_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        L.orc_boys.argtypes = [C.c_int, C.c_double, dp]
        L.orc_prim_norm.argtypes = [C.c_double, C.c_int]
        L.orc_prim_norm.restype = C.c_double
        L.orc_prim_eri.argtypes = [C.c_int] * 4 + [dp] * 4 + [C.c_double] * 4 + [dp]
        L.orc_eri_shell.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_double, dp]
        L.orc_overlap_shell.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, dp]
        L.orc_norma.argtypes = [C.c_void_p, dp]
        L.orc_nbasis.argtypes = [C.c_void_p]
        L.orc_nbasis.restype = C.c_int
        L.orc_shellpair_list.argtypes = [C.c_void_p, C.c_double, ip, ip, dp]
        L.orc_shellpair_list.restype = C.c_int
        L.orc_schwarz.argtypes = [C.c_void_p, dp]
        L.orc_compute_2body_direct.argtypes = [C.c_void_p, ip, ip, C.c_int, dp, dp, C.c_double,
                                               C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, dp, lp]
        L.orc_compute_2body_direct.restype = C.c_long
        L.orc_compute_coupling_direct.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_int, C.c_double,
                                                  C.c_int, dp, lp]
        L.orc_1body.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, dp]
        L.orc_compute_2body_all.argtypes = [C.c_void_p, ip, ip, C.c_int, dp, dp, C.c_int, C.c_double,
                                            C.c_double, dp]
        L.orc_compute_2body_all.restype = C.c_long
        L.orc_compute_coupling_all.argtypes = [C.c_void_p, C.c_void_p, C.c_double, dp]
        L.orc_prim_g12.argtypes = [C.c_int] * 4 + [dp] * 4 + [C.c_double] * 4 + [C.c_int, dp, dp, dp]
        L.orc_g12_shell.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.orc_compute_g12_direct.argtypes = [C.c_void_p, ip, ip, C.c_int, dp, C.c_double, C.c_int, dp, dp,
                                             C.c_int, dp]
        L.orc_compute_g12_direct.restype = C.c_long
        L.orc_compute_g12inter_direct.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_int, C.c_double, C.c_int, dp, dp,
                                                  C.c_int, dp]
        _LIB = L
    return _LIB


class _CBasis(C.Structure):
    _fields_ = [("nshell", C.c_int), ("l", C.c_void_p), ("nprim", C.c_void_p), ("poff", C.c_void_p),
                ("exps", C.c_void_p), ("coefs", C.c_void_p), ("xyz", C.c_void_p)]


class Basis:
    """Shell arrays of one species, kept alive for the C side."""

    def __init__(self, species):
        self.species = species
        self.l, self.nprim, self.poff, self.exps, self.coefs, self.xyz = species.arrays()
        self.c = _CBasis(len(self.l), self.l.ctypes.data, self.nprim.ctypes.data, self.poff.ctypes.data,
                         self.exps.ctypes.data, self.coefs.ctypes.data, self.xyz.ctypes.data)
        self.ptr = C.cast(C.pointer(self.c), C.c_void_p)
        self.nshell = len(self.l)
        self.n = int(lib().orc_nbasis(self.ptr))
        self.sizes = (self.l + 1) * (self.l + 2) // 2
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(int)


def boys(mmax: int, T: float) -> np.ndarray:
    F = np.zeros(mmax + 1)
    lib().orc_boys(mmax, float(T), F)
    return F


def prim_norm(alpha: float, l: int) -> float:
    return lib().orc_prim_norm(alpha, l)


def prim_eri(la, lb, lc, ld, A, B, Cc, D, a, b, c, d) -> np.ndarray:
    nc = lambda l: (l + 1) * (l + 2) // 2
    out = np.zeros(nc(la) * nc(lb) * nc(lc) * nc(ld))
    f = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_prim_eri(la, lb, lc, ld, f(A), f(B), f(Cc), f(D), a, b, c, d, out)
    return out


def eri_shell(b12: Basis, s1, s2, b34: Basis, s3, s4, prim_eps=0.0) -> np.ndarray:
    """Raw contracted ERIs of (s1 s2|s3 s4), libint order (f1 slowest)."""
    n = int(b12.sizes[s1] * b12.sizes[s2] * b34.sizes[s3] * b34.sizes[s4])
    out = np.zeros(n)
    lib().orc_eri_shell(b12.ptr, s1, s2, b34.ptr, s3, s4, prim_eps, out)
    return out


def norma(b: Basis) -> np.ndarray:
    """Libint2Iface.cpp:115-127: 1/sqrt of the raw self-overlap diagonal."""
    out = np.zeros(b.n)
    lib().orc_norma(b.ptr, out)
    return out


def shellpair_list(b: Basis, threshold: float = 1e-12):
    """compute_shellpair_list (Libint2Iface.cpp:3248-3342).  Returns
    (pairs array [npair, 2] of (s1, s2 <= s1), norms [nshell, nshell])."""
    ns = b.nshell
    s1 = np.zeros(ns * (ns + 1) // 2, dtype=np.int32)
    s2 = np.zeros_like(s1)
    norms = np.zeros(ns * ns)
    np_ = lib().orc_shellpair_list(b.ptr, threshold, s1, s2, norms)
    return np.stack([s1[:np_], s2[:np_]], axis=1), norms.reshape(ns, ns)


def schwarz(b: Basis) -> np.ndarray:
    """compute_schwartz_ints (Libint2Iface.cpp:3345-3412), raw convention."""
    Q = np.zeros(b.nshell * b.nshell)
    lib().orc_schwarz(b.ptr, Q)
    return Q.reshape(b.nshell, b.nshell)


def compute_2body_direct(b: Basis, D: np.ndarray, factor: float, pairs=None, Q=None,
                         precision: float = DBL_EPSILON, prim_eps: float = 0.0,
                         nthreads: int = 0, return_counts: bool = False, s1_stride: int = 1):
    """LibintInterface::compute_2body_direct (Libint2Iface.cpp:418-618) with
    the C wrapper's per-call pair list and Schwarz matrix (3505, 3520)."""
    if pairs is None:
        pairs, _ = shellpair_list(b)
    if Q is None:
        Q = schwarz(b)
    s1 = np.ascontiguousarray(pairs[:, 0], dtype=np.int32)
    s2 = np.ascontiguousarray(pairs[:, 1], dtype=np.int32)
    G = np.zeros(b.n * b.n)
    counters = np.zeros(2, dtype=np.int64)
    screen = 1 if b.species.is_electron else 0
    nints = lib().orc_compute_2body_direct(b.ptr, s1, s2, len(s1), np.ascontiguousarray(Q, dtype=np.float64).reshape(-1),
                                           np.ascontiguousarray(D, dtype=np.float64).reshape(-1), float(factor),
                                           screen, precision, prim_eps, nthreads, s1_stride, G, counters)
    G = G.reshape(b.n, b.n)
    if return_counts:
        return G, {"eris": int(nints), "quartets": int(counters[0]), "prim_quartets": int(counters[1])}
    return G


def compute_coupling_direct(this: Basis, other: Basis, id_this: int, id_other: int,
                            D_other: np.ndarray, drop: float = 1e-10, nthreads: int = 0,
                            return_counts: bool = False):
    """LibintInterface_compute_coupling_direct (Libint2Iface.cpp:3566-3589):
    the lower species id runs the outer loop; ``permuted = id_this >
    id_other``.  Returns the n_this x n_this image (lower triangle filled in
    row-major = upper triangle of the Fortran column-major result)."""
    permuted = id_this > id_other
    outer, inner = (other, this) if permuted else (this, other)
    n = this.n
    B = np.zeros(n * n)
    counts = np.zeros(2, dtype=np.int64)
    lib().orc_compute_coupling_direct(outer.ptr, inner.ptr, np.ascontiguousarray(D_other, dtype=np.float64).reshape(-1),
                                      1 if permuted else 0, drop, nthreads, B, counts)
    B = B.reshape(n, n)
    if return_counts:
        return B, {"kept": int(counts[0]), "near_cutoff": int(counts[1])}
    return B


def one_body(b: Basis, kind: int, charges=()) -> np.ndarray:
    """S (1), T (2) or point-charge attraction (3), normalised with norma."""
    Z = np.array([z for z, _ in charges] or [0.0], dtype=np.float64)
    X = np.array([x for _, x in charges] or [[0.0, 0.0, 0.0]], dtype=np.float64).reshape(-1)
    M = np.zeros(b.n * b.n)
    lib().orc_1body(b.ptr, kind, len(charges), Z, X, M)
    return M.reshape(b.n, b.n)


def compute_2body_directAll(b: Basis, D: np.ndarray, pairs=None, Q=None, precision: float = DBL_EPSILON,
                            drop: float = 1e-10, A: np.ndarray = None) -> np.ndarray:
    """LibintInterface::compute_2body_directAll (Libint2Iface.cpp:2552-2742)
    with the C wrapper's Schwarz matrix (3736-3750).  Returns the (n, n, n, n)
    tensor, element (i j|k l) at [i, j, k, l] (memory i n^3 + j n^2 + k n + l);
    elements the store does not write keep the values of ``A`` (zeros when
    None).  ``drop`` = 0 keeps every integral of the screened quartets."""
    if pairs is None:
        pairs, _ = shellpair_list(b)
    if Q is None:
        Q = schwarz(b)
    n = b.n
    out = np.zeros(n ** 4) if A is None else np.array(A, dtype=np.float64).reshape(-1).copy()
    s1 = np.ascontiguousarray(pairs[:, 0], dtype=np.int32)
    s2 = np.ascontiguousarray(pairs[:, 1], dtype=np.int32)
    screen = 1 if b.species.is_electron else 0
    lib().orc_compute_2body_all(b.ptr, s1, s2, len(s1), np.ascontiguousarray(Q, dtype=np.float64).reshape(-1),
                                np.ascontiguousarray(D, dtype=np.float64).reshape(-1), screen, precision,
                                float(drop), out)
    return out.reshape(n, n, n, n)


def compute_2body_directIT(b: Basis, D: np.ndarray, Cmat: np.ndarray, p: int, **kw) -> np.ndarray:
    """LibintInterface::compute_2body_directIT (Libint2Iface.cpp:620-921):
    GG[j, k, l] = sum_i C(p, i) (i j|k l) over the quartets, screen and drop
    of compute_2body_directAll, with C(p, i) = Cmat[p, i] (the Eigen copy of
    the Fortran coefficient array, :3535-3537).  The reference accumulates
    each unique integral into its distinct index permutations and mirrors
    the k > l triangle (:776-875); the sum over the full store tensor is the
    same quantity."""
    A = compute_2body_directAll(b, D, **kw)
    return np.tensordot(np.asarray(Cmat, dtype=np.float64)[p, :], A, axes=(0, 0))


def _coupling_tensor(outer: Basis, inner: Basis, drop: float) -> np.ndarray:
    T = np.zeros(outer.n * outer.n * inner.n * inner.n)
    lib().orc_compute_coupling_all(outer.ptr, inner.ptr, float(drop), T)
    return T.reshape(outer.n, outer.n, inner.n, inner.n)


def compute_coupling_directAll(this: Basis, other: Basis, id_this: int, id_other: int,
                               drop: float = 1e-10) -> np.ndarray:
    """LibintInterface_compute_coupling_directAll (Libint2Iface.cpp:3752-3773,
    2744-2908): the lower species id runs the outer loops; the result is
    (n_outer, n_outer, n_inner, n_inner) with the reference's strides
    (n n_o^2, n_o^2, n_o, 1), n = this.n, n_o = other.n.  Called from the
    higher species id those strides only fit when n == n_o (ValueError
    otherwise)."""
    permuted = id_this > id_other
    if permuted and this.n != other.n:
        raise ValueError("reference layout undefined for a permuted call with different basis sizes")
    outer, inner = (other, this) if permuted else (this, other)
    return _coupling_tensor(outer, inner, drop)


def compute_coupling_directIT(this: Basis, other: Basis, id_this: int, id_other: int, Cmat: np.ndarray,
                              p: int, drop: float = 1e-10) -> np.ndarray:
    """LibintInterface::compute_coupling_directIT (Libint2Iface.cpp:1296-1622):
    GG[v, l, s] (n_this x n_other x n_other) = sum_u C(p, u) (u v|l s), u, v
    of ``this``, l, s of ``other``, over every pair of both species with the
    |raw| >= drop element sets of compute_coupling_directAll."""
    permuted = id_this > id_other
    c = np.asarray(Cmat, dtype=np.float64)[p, :]
    if permuted:
        T = _coupling_tensor(other, this, drop)          # [l, s, u, v]
        return np.einsum("u,lsuv->vls", c, T)
    T = _coupling_tensor(this, other, drop)              # [u, v, l, s]
    return np.tensordot(c, T, axes=(0, 0))


def _gem(coefficients, exponents):
    c = np.ascontiguousarray(coefficients, dtype=np.float64).reshape(-1)
    g = np.ascontiguousarray(exponents, dtype=np.float64).reshape(-1)
    if c.shape != g.shape:
        raise ValueError("geminal coefficients and exponents differ in length")
    return len(c), (c if len(c) else np.zeros(1)), (g if len(g) else np.zeros(1))


def prim_g12(la, lb, lc, ld, A, B, Cc, D, a, b, c, d, coefficients, exponents) -> np.ndarray:
    """Unnormalised primitive integrals of sum_k c_k exp(-g_k r12^2), the
    layout of prim_eri (McMurchie-Davidson with the geminal's Hermite table)."""
    nc = lambda l: (l + 1) * (l + 2) // 2
    out = np.zeros(nc(la) * nc(lb) * nc(lc) * nc(ld))
    f = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    ng, gc, gg = _gem(coefficients, exponents)
    lib().orc_prim_g12(la, lb, lc, ld, f(A), f(B), f(Cc), f(D), a, b, c, d, ng, gc, gg, out)
    return out


def g12_shell(b12: Basis, s1, s2, b34: Basis, s3, s4, coefficients, exponents) -> np.ndarray:
    """Raw contracted geminal integrals of (s1 s2|s3 s4), libint order."""
    n = int(b12.sizes[s1] * b12.sizes[s2] * b34.sizes[s3] * b34.sizes[s4])
    out = np.zeros(n)
    ng, gc, gg = _gem(coefficients, exponents)
    lib().orc_g12_shell(b12.ptr, s1, s2, b34.ptr, s3, s4, ng, gc, gg, out)
    return out


def compute_g12_direct(b: Basis, D: np.ndarray, factor: float, coefficients, exponents, pairs=None,
                       nthreads: int = 0) -> np.ndarray:
    """LibintInterface::compute_g12_direct (Libint2Iface.cpp:2193-2365) with
    the C wrapper's arguments (3687-3706): J + factor K with the contracted
    Gaussian geminal, every quartet of the pair list, no density screen."""
    if pairs is None:
        pairs, _ = shellpair_list(b)
    s1 = np.ascontiguousarray(pairs[:, 0], dtype=np.int32)
    s2 = np.ascontiguousarray(pairs[:, 1], dtype=np.int32)
    G = np.zeros(b.n * b.n)
    ng, gc, gg = _gem(coefficients, exponents)
    lib().orc_compute_g12_direct(b.ptr, s1, s2, len(s1), np.ascontiguousarray(D, dtype=np.float64).reshape(-1),
                                 float(factor), ng, gc, gg, nthreads, G)
    return G.reshape(b.n, b.n)


def compute_g12inter_direct(this: Basis, other: Basis, id_this: int, id_other: int, D_other: np.ndarray,
                            coefficients, exponents, drop: float = 1e-10, nthreads: int = 0) -> np.ndarray:
    """libintinterface_compute_g12inter_direct (Libint2Iface.cpp:3708-3734,
    2367-2550): the coupling build of compute_coupling_direct with the
    geminal operator; same image convention (lower triangle, row-major)."""
    permuted = id_this > id_other
    outer, inner = (other, this) if permuted else (this, other)
    n = this.n
    B = np.zeros(n * n)
    ng, gc, gg = _gem(coefficients, exponents)
    lib().orc_compute_g12inter_direct(outer.ptr, inner.ptr, np.ascontiguousarray(D_other, dtype=np.float64).reshape(-1),
                                      1 if permuted else 0, drop, ng, gc, gg, nthreads, B)
    return B.reshape(n, n)
