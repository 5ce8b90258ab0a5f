"""The wave-cooperative programs of the high-L classes (gen_eri.py, CoopGen)
executed on the CPU exactly as a wavefront runs them -- per group of
independent rounds every lane reads its LDS operands, then every lane writes -- and compared with the
oracle's primitive ERIs (McMurchie-Davidson, pinned against the reference's
own Rys code), single primitive quartets and a contracted two-primitive sum.

A slot allocated while still live, an operand offset off by one m level, a
coefficient read for the wrong direction or recursion kind, or an HRR round
ahead of its inputs gives a wrong (ab|cd) block here.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "openlowdin_amd", "csrc"))
import gen_eri as G  # noqa: E402

COOP_CLASSES = [(2, 1, 2, 1), (2, 2, 1, 1), (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2), (2, 1, 1, 1), (1, 1, 1, 1)]


def _decode(words, r):
    return words[r * 64:(r + 1) * 64]


def run_program(gen, prims, AB, CD):
    """prims: list of dicts with PA, WP, QC, WQ (3-vectors), o2p, o2q, o2pq,
    rop, roq, F (scaled Boys values F_0..F_L).  Executes the rounds in the
    kernel's groups: every lane reads the operands of all rounds of a group,
    then every lane writes."""
    vrounds, hrounds, words, out0 = gen.build()
    vg, ag, hg = gen.groups
    assert sum(n for _, n in vg) == len(vrounds) and sum(n for _, n in hg) == len(hrounds)
    info = gen.info
    R = np.full(info["region"] + 8, np.nan)
    R[G.COOP_ZERO] = R[G.COOP_ZERO + 1] = 0.0
    R[G.COOP_ACC:G.COOP_ACC + gen.nacc] = 0.0
    nv, na = len(vrounds), info["acc_rounds"]

    def vrr_op(x, y, kind, pq):
        s0, k1 = (x & 0xfff8) // 8, x & 7
        s2, k2 = ((x >> 16) & 0xfff8) // 8, (x >> 16) & 7
        s4 = (y & 0xfff8) // 8
        dst, i = ((y >> 16) & 0xfff8) // 8, (y >> 16) & 3
        c1, c2 = R[6 * kind + 2 * i], R[6 * kind + 2 * i + 1]
        if kind:
            v = c1 * R[s0] + c2 * R[s0 + 1] + k1 * pq["o2q"] * (R[s2] - pq["roq"] * R[s2 + 1]) \
                + k2 * pq["o2pq"] * R[s4]
        else:
            v = c1 * R[s0] + c2 * R[s0 + 1] + k1 * pq["o2p"] * (R[s2] - pq["rop"] * R[s2 + 1])
        return dst, v

    for pq in prims:
        R[0:6] = [pq["PA"][0], pq["WP"][0], pq["PA"][1], pq["WP"][1], pq["PA"][2], pq["WP"][2]]
        R[6:12] = [pq["QC"][0], pq["WQ"][0], pq["QC"][1], pq["WQ"][1], pq["QC"][2], pq["WQ"][2]]
        R[info["base"]:info["base"] + gen.L + 1] = pq["F"]
        r = 0
        for kind, nr in vg:
            assert all(vrounds[r + k][0] == kind for k in range(nr))
            vals = [vrr_op(x, y, kind, pq) for k in range(nr) for x, y in _decode(words, r + k)]
            for dst, v in vals:
                R[dst] = v
            r += nr
        r = nv
        for _, nr in ag:
            vals = [(y // 8, R[y // 8] + R[x // 8]) for k in range(nr) for x, y in _decode(words, r + k)]
            for dst, v in vals:
                R[dst] = v
            r += nr
    R[12:15] = AB
    R[15:18] = CD
    r = nv + na
    for kind, nr in hg:
        vals = []
        for k in range(nr):
            for x, y in _decode(words, r + k):
                xs, ys = (x & 0xffff) // 8, (x >> 16) // 8
                dst, i = (y & 0xfff8) // 8, y & 3
                vals.append((dst, R[xs] + R[(15 if kind else 12) + i] * R[ys]))
        for dst, v in vals:
            R[dst] = v
        r += nr
    return R[out0:out0 + gen.nout].copy()


def prim_params(A, B, C, D, a, b, c, d, L):
    p, q = a + b, c + d
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    rho = p * q / (p + q)
    PQ = P - Q
    T = rho * PQ @ PQ
    pref = 2.0 * np.pi ** 2.5 / (p * q * np.sqrt(p + q)) * np.exp(-a * b / p * np.sum((A - B) ** 2)) \
        * np.exp(-c * d / q * np.sum((C - D) ** 2))
    return {"PA": P - A, "WP": -(q / (p + q)) * PQ, "QC": Q - C, "WQ": (p / (p + q)) * PQ,
            "o2p": 0.5 / p, "o2q": 0.5 / q, "o2pq": 0.5 / (p + q), "rop": q / (p + q), "roq": p / (p + q),
            "F": pref * O.boys(L, T)}


@pytest.mark.parametrize("cls", COOP_CLASSES, ids=lambda c: "%d%d%d%d" % c)
def test_coop_program_matches_oracle_primitives(cls):
    la, lb, lc, ld = cls
    gen = G.CoopGen(*cls)
    rng = np.random.default_rng(sum(cls))
    for trial in range(3):
        A, B, C, D = (rng.normal(size=3) * 1.2 for _ in range(4))
        a, b, c, d = rng.uniform(0.2, 3.0, size=4)
        got = run_program(gen, [prim_params(A, B, C, D, a, b, c, d, gen.L)], A - B, C - D)
        ref = O.prim_eri(la, lb, lc, ld, A, B, C, D, a, b, c, d)
        assert np.all(np.isfinite(got))
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))), (trial, np.max(np.abs(got - ref)))


@pytest.mark.parametrize("cls", [(2, 1, 2, 1), (2, 2, 2, 2)], ids=lambda c: "%d%d%d%d" % c)
def test_coop_program_contracts_primitive_quartets(cls):
    """Two bra and two ket primitive pairs: the accumulators sum four
    primitive quartets before the (contraction-independent) HRR."""
    la, lb, lc, ld = cls
    gen = G.CoopGen(*cls)
    rng = np.random.default_rng(7)
    A, B, C, D = (rng.normal(size=3) for _ in range(4))
    ea, eb, ec, ed = (rng.uniform(0.3, 2.0, size=2) for _ in range(4))
    prims, ref = [], 0.0
    for k in range(2):
        for i in range(2):
            prims.append(prim_params(A, B, C, D, ea[i], eb[i], ec[k], ed[k], gen.L))
            ref = ref + O.prim_eri(la, lb, lc, ld, A, B, C, D, ea[i], eb[i], ec[k], ed[k])
    got = run_program(gen, prims, A - B, C - D)
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("cls", COOP_CLASSES, ids=lambda c: "%d%d%d%d" % c)
def test_coop_program_layout(cls):
    """Byte offsets fit the 16-bit descriptor fields and the region; the
    base run holds F_0..F_L."""
    gen = G.CoopGen(*cls)
    _, _, words, out0 = gen.build()
    reg = gen.info["region"]
    assert reg * 8 < 65536
    for x, y in words:
        for off in (x & 0xfff8, (x >> 16) & 0xfff8, y & 0xfff8, (y >> 16) & 0xfff8):
            assert off // 8 < reg
    assert gen.info["base"] + gen.L < reg
