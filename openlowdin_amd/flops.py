"""Algorithmic work model of the direct Fock build (kernel independent).

Per computed shell quartet of class (la lb|lc ld) with P primitive quartets:

    F = P * (c_vrr + c_boys(L) + c_prim) + c_hrr + 12 * na nb nc nd

* c_vrr, c_hrr: floating-point operations of the textbook Obara-Saika
  vertical recursion (per primitive quartet) and Head-Gordon-Pople
  horizontal recursion (per contracted quartet), counted by
  csrc/gen_eri.py into csrc/ops_table.json;
* c_boys(L) = 16 + [L > 0] (14 + 3 L): seventh-order Taylor evaluation of
  F_L plus exp(-T) and the downward recursion to F_0;
* c_prim = 25: primitive-quartet geometry (p+q, 1/(p+q), rho, P-Q, T, the
  prefactor, W-P, W-Q, and the scaling of F_0..F_L);
* 12 per ERI: the six multiply-adds of the J/K scatter
  (openLOWDIN Libint2Iface.cpp:577-582); a coupling ERI costs 4 (two
  directions, one multiply-add each).

ERIs are counted as the reference's num_ints_computed
(Libint2Iface.cpp:539): the sum of na nb nc nd over computed quartets.
"""
import json
import os

_TABLE = None

C_PRIM = 25


def ops_table():
    global _TABLE
    if _TABLE is None:
        with open(os.path.join(os.path.dirname(__file__), "csrc", "ops_table.json")) as fh:
            _TABLE = json.load(fh)["classes"]
    return _TABLE


def ncart(l):
    return (l + 1) * (l + 2) // 2


def c_boys(L):
    return 16 + (14 + 3 * L if L > 0 else 0)


def class_flops(code: str, quartets: int, prim_quartets: int, per_eri: int = 12) -> float:
    t = ops_table()[code]
    la, lb, lc, ld = (int(c) for c in code)
    L = la + lb + lc + ld
    neri = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    return prim_quartets * (t["vrr_flops"] + c_boys(L) + C_PRIM) + quartets * (t["hrr_flops"] + per_eri * neri)


def build_flops(profile) -> float:
    """profile: [(class code, ms, quartets, prim quartets)] of one build."""
    return sum(class_flops(c, q, p) for c, _, q, p in profile)
