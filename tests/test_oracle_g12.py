"""Pins the oracle's Gaussian-geminal (G12) integrals and builds before the
GPU is checked against them.

The oracle evaluates sum_k c_k exp(-g_k r12^2) by McMurchie-Davidson with
the geminal's Hermite table.  It is pinned by:

* the g -> 0 limit, where the operator is 1 and every integral is the
  product of two overlap integrals (pinned through the reference energies);
* an independent evaluation written here: the integrand is a Gaussian in
  (x1, x2) per axis, so each primitive integral is a product of moments of
  a two-variable Gaussian, computed by Gaussian integration by parts over
  all four polynomial factors (no Hermite expansion, no HRR);
* the closed form of [ss|exp(-g r12^2)|ss] and its delta-function limit;
* the direct and coupling builds against brute-force contractions of the
  full G12 tensor.
"""
from functools import lru_cache

import numpy as np
import pytest

import oracle as O
from openlowdin_amd import system as S


def _comps(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def _axis_moments(mu, cen, Sig):
    """I(n0, n1, n2, n3) = <(x1-A)^n0 (x1-B)^n1 (x2-C)^n2 (x2-D)^n3> of the
    normalised Gaussian with mean mu and covariance Sig (x1: bra, x2: ket),
    by E[(x - c) f] = (mu - c) E[f] + sum_v Cov(x, x_v) E[df/dx_v]."""
    side = (0, 0, 1, 1)

    @lru_cache(maxsize=None)
    def I(n0, n1, n2, n3):
        ns = [n0, n1, n2, n3]
        if sum(ns) == 0:
            return 1.0
        k = next(i for i in range(4) if ns[i] > 0)
        ns[k] -= 1
        v = (mu[side[k]] - cen[k]) * I(*ns)
        for j in range(4):
            if ns[j] > 0:
                lower = list(ns)
                lower[j] -= 1
                v += ns[j] * Sig[side[k], side[j]] * I(*lower)
        return v
    return I


def _moment_prim(la, lb, lc, ld, A, B, C, D, a, b, c, d, cs, gs):
    """Primitive [ab|sum c exp(-g r12^2)|cd] from two-variable Gaussian moments."""
    A, B, C, D = (np.asarray(x, dtype=float) for x in (A, B, C, D))
    p, q = a + b, c + d
    P, Q = (a * A + b * B) / p, (c * C + d * D) / q
    kab = np.exp(-a * b / p * np.sum((A - B) ** 2))
    kcd = np.exp(-c * d / q * np.sum((C - D) ** 2))
    out = np.zeros(len(_comps(la)) * len(_comps(lb)) * len(_comps(lc)) * len(_comps(ld)))
    for ck, g in zip(cs, gs):
        M = np.array([[p + g, -g], [-g, q + g]])
        Dm = np.linalg.det(M)
        Sig = np.linalg.inv(2.0 * M)
        norm = (np.pi ** 2 / Dm) ** 1.5 * np.exp(-p * q * g / Dm * np.sum((P - Q) ** 2))
        tabs = [_axis_moments(np.linalg.solve(M, np.array([p * P[ax], q * Q[ax]])),
                              (A[ax], B[ax], C[ax], D[ax]), Sig) for ax in range(3)]
        m = 0
        for xa in _comps(la):
            for xb in _comps(lb):
                for xc in _comps(lc):
                    for xd in _comps(ld):
                        v = 1.0
                        for ax in range(3):
                            v *= tabs[ax](xa[ax], xb[ax], xc[ax], xd[ax])
                        out[m] += ck * norm * kab * kcd * v
                        m += 1
    return out


def test_primitive_g12_against_moments():
    rng = np.random.default_rng(3)
    cs, gs = [0.7, -1.3, 2.0], [0.0, 0.8, 5.5]
    worst = 0.0
    classes = [(la, lb, lc, ld) for la in range(3) for lb in range(3) for lc in range(3) for ld in range(3)]
    for la, lb, lc, ld in classes[::4] + [(2, 2, 2, 2), (1, 0, 0, 0), (0, 0, 0, 1)]:
        A, B, C, D = rng.normal(size=(4, 3)) * 0.8
        a, b, c, d = rng.uniform(0.3, 3.0, size=4)
        got = O.prim_g12(la, lb, lc, ld, A, B, C, D, a, b, c, d, cs, gs)
        ref = _moment_prim(la, lb, lc, ld, A, B, C, D, a, b, c, d, cs, gs)
        scale = max(1.0, np.max(np.abs(ref)))
        worst = max(worst, np.max(np.abs(got - ref)) / scale)
    print("max rel err %.2e" % worst)
    assert worst < 1e-12


def test_ssss_closed_form_and_delta_limit():
    A, B, C, D = np.array([0.0, 0.0, 0.0]), np.array([0.3, -0.2, 0.5]), np.array([1.0, 0.4, -0.3]), \
        np.array([0.8, 0.9, 0.1])
    a, b, c, d = 1.2, 0.7, 0.9, 1.6
    p, q = a + b, c + d
    P, Q = (a * A + b * B) / p, (c * C + d * D) / q
    kab = np.exp(-a * b / p * np.sum((A - B) ** 2))
    kcd = np.exp(-c * d / q * np.sum((C - D) ** 2))
    g = 2.5
    Dl = p * q + g * (p + q)
    want = kab * kcd * (np.pi ** 2 / Dl) ** 1.5 * np.exp(-p * q * g / Dl * np.sum((P - Q) ** 2))
    got = O.prim_g12(0, 0, 0, 0, A, B, C, D, a, b, c, d, [1.0], [g])[0]
    assert abs(got - want) < 1e-14 * abs(want)
    # (g/pi)^(3/2) exp(-g r12^2) -> delta(r12): the four-centre overlap
    s = p + q
    four = kab * kcd * np.exp(-p * q / s * np.sum((P - Q) ** 2)) * (np.pi / s) ** 1.5
    g = 1e6
    got = O.prim_g12(0, 0, 0, 0, A, B, C, D, a, b, c, d, [(g / np.pi) ** 1.5], [g])[0]
    assert abs(got - four) < 1e-4 * abs(four)


@pytest.fixture(scope="module")
def water():
    sp = S.water("6-31G**").species[0]
    return sp, O.Basis(sp)


def _overlap(b, s1, s2):
    out = np.zeros(int(b.sizes[s1] * b.sizes[s2]))
    O.lib().orc_overlap_shell(b.ptr, s1, b.ptr, s2, out)
    return out.reshape(b.sizes[s1], b.sizes[s2])


def test_zero_exponent_is_overlap_product(water):
    sp, b = water
    worst = 0.0
    for s1 in range(b.nshell):
        for s3 in range(b.nshell):
            s2, s4 = (s1 * 5 + 3) % b.nshell, (s3 * 7 + 1) % b.nshell
            g = O.g12_shell(b, s1, s2, b, s3, s4, [1.0], [0.0]).reshape(
                b.sizes[s1], b.sizes[s2], b.sizes[s3], b.sizes[s4])
            ref = np.einsum("ij,kl->ijkl", _overlap(b, s1, s2), _overlap(b, s3, s4))
            worst = max(worst, np.max(np.abs(g - ref)))
    assert worst < 1e-14


def _full_g12(b, cs, gs):
    n = b.n
    T = np.zeros((n, n, n, n))
    off, sz = b.offsets, b.sizes
    nrm = O.norma(b)
    for s1 in range(b.nshell):
        for s2 in range(b.nshell):
            for s3 in range(b.nshell):
                for s4 in range(b.nshell):
                    blk = O.g12_shell(b, s1, s2, b, s3, s4, cs, gs).reshape(sz[s1], sz[s2], sz[s3], sz[s4])
                    T[off[s1]:off[s1] + sz[s1], off[s2]:off[s2] + sz[s2], off[s3]:off[s3] + sz[s3],
                      off[s4]:off[s4] + sz[s4]] = blk
    return T * np.einsum("i,j,k,l->ijkl", nrm, nrm, nrm, nrm)


def test_g12_direct_build_against_full_tensor():
    sp = S.water("6-31G").species[0]
    b = O.Basis(sp)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(b.n, 3))
    D = X @ X.T
    cs, gs = [1.3, -0.4, 0.25], [0.0, 0.6, 3.0]
    T = _full_g12(b, cs, gs)
    pairs, _ = O.shellpair_list(b)
    assert len(pairs) == b.nshell * (b.nshell + 1) // 2     # no pair dropped for this molecule
    J = np.einsum("ijkl,kl->ij", T, D)
    K = np.einsum("ikjl,kl->ij", T, D)
    for f in (0.0, -1.0):
        G = O.compute_g12_direct(b, D, f, cs, gs)
        assert np.max(np.abs(G - (J + f * K))) < 1e-12


def test_g12_coupling_against_full_tensor():
    f = S.fhf_anion("6-31G")
    e, p = f.species
    be, bp = O.Basis(e), O.Basis(p)
    rng = np.random.default_rng(5)
    Dp = rng.normal(size=(bp.n, bp.n))
    Dp = Dp + Dp.T
    De = rng.normal(size=(be.n, be.n))
    De = De + De.T
    cs, gs = [0.9, -0.3], [0.05, 1.7]
    ne, npr = be.n, bp.n
    T = np.zeros((ne, ne, npr, npr))
    for s1 in range(be.nshell):
        for s2 in range(be.nshell):
            for s3 in range(bp.nshell):
                for s4 in range(bp.nshell):
                    blk = O.g12_shell(be, s1, s2, bp, s3, s4, cs, gs).reshape(
                        be.sizes[s1], be.sizes[s2], bp.sizes[s3], bp.sizes[s4])
                    T[be.offsets[s1]:be.offsets[s1] + be.sizes[s1], be.offsets[s2]:be.offsets[s2] + be.sizes[s2],
                      bp.offsets[s3]:bp.offsets[s3] + bp.sizes[s3],
                      bp.offsets[s4]:bp.offsets[s4] + bp.sizes[s4]] = blk
    ne_, np_ = O.norma(be), O.norma(bp)
    T *= np.einsum("i,j,k,l->ijkl", ne_, ne_, np_, np_)
    Be = O.compute_g12inter_direct(be, bp, 1, 2, Dp, cs, gs, drop=0.0)
    assert np.max(np.abs(np.tril(Be) - np.tril(np.einsum("ijkl,kl->ij", T, Dp)))) < 1e-12
    Bp = O.compute_g12inter_direct(bp, be, 2, 1, De, cs, gs, drop=0.0)
    assert np.max(np.abs(np.tril(Bp) - np.tril(np.einsum("ijkl,ij->kl", T, De)))) < 1e-12
