"""Generate class-specialised Obara-Saika VRR and Head-Gordon-Pople HRR code.

For every canonical quartet class (la lb|lc ld) with l <= LMAX, la >= lb,
lc >= ld and pair class (la, lb) >= (lc, ld), this emits two device
functions into ``gen/eri_<class>.inc`` (plus the Gaussian-geminal VRR
``g12vrr_<class>``, see G12Gen), three kernel translation units per class
(``gen/k_<class>.hip``: Fock build, coupling and Schwarz kernels;
``gen/s_<class>.hip``: integral store and first-quarter transform;
``gen/g_<class>.hip``: G12 Fock build and coupling), so that they compile in
parallel, and the launch dispatch table ``gen/dispatch.h``:

``vrr_<la><lb><lc><ld>(F, ...)``
    Obara-Saika vertical recursion (Obara & Saika, J. Chem. Phys. 84, 3963
    (1986), eqs. 39 and its ket analogue) from the scaled Boys values
    F[m] = pref * F_m(T) to the targets [e0|f0]^(0), |e| in [la, la+lb],
    |f| in [lc, lc+ld], added into the contracted accumulator ``acc``.
    Every intermediate is a named local, emitted once in dependency order.

``hrr_<la><lb><lc><ld>(acc, AB, CD, out)``
    Head-Gordon-Pople horizontal recursion (J. Chem. Phys. 89, 5777 (1988)):
    (a, b+1_i| = (a+1_i, b| + AB_i (a b| on the bra, then the same on the ket
    with CD, producing (ab|cd) in the standard Cartesian order (xx, xy, xz,
    yy, yz, zz), index ((ia*nb + ib)*nc + ic)*nd + id.

Also emits ``vrr_flops`` / ``hrr_flops`` tables (floating-point operations
per primitive quartet and per contracted quartet) used by the roofline
accounting.

Run: python openlowdin_amd/csrc/gen_eri.py  (the package build does this).
"""
import itertools
import json
import os
import sys

LMAX = 2
XYZ = "xyz"


def comps(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def ncart(l):
    return (l + 1) * (l + 2) // 2


def cidx(v):
    return comps(sum(v)).index(tuple(v))


def add(v, i, d):
    w = list(v)
    w[i] += d
    return tuple(w)


def pick(v):
    """Direction to reduce: the largest component (ties: x, then y, then z)."""
    best = 0
    for i in range(3):
        if v[i] > v[best]:
            best = i
    return best


def canonical_classes(lmax=LMAX):
    pcs = [(a, b) for a in range(lmax + 1) for b in range(a + 1)]
    pcs.sort(key=lambda p: (p[0] + p[1], p[0]))
    out = []
    for i, X in enumerate(pcs):
        for Y in pcs[: i + 1]:
            out.append((X[0], X[1], Y[0], Y[1]))
    return pcs, out


class VRRGen:
    def __init__(self, la, lb, lc, ld):
        self.la, self.lb, self.lc, self.ld = la, lb, lc, ld
        self.lines = []
        self.memo = {}
        self.nv = 0
        self.flops = 0
        self.es = [c for l in range(la, la + lb + 1) for c in comps(l)]
        self.fs = [c for l in range(lc, lc + ld + 1) for c in comps(l)]

    def new(self, expr, flops):
        name = "v%d" % self.nv
        self.nv += 1
        self.lines.append("  const R %s = %s;" % (name, expr))
        self.flops += flops
        return name

    def need(self, e, f, m):
        if min(e) < 0 or min(f) < 0:
            return None
        key = (e, f, m)
        if key in self.memo:
            return self.memo[key]
        if sum(e) == 0 and sum(f) == 0:
            name = "F[%d]" % m
        elif sum(f) > 0:
            i = pick(f)
            f1 = add(f, i, -1)
            t = [("QC%s" % XYZ[i], self.need(e, f1, m)), ("WQ%s" % XYZ[i], self.need(e, f1, m + 1))]
            terms = ["%s*%s" % (c, v) for c, v in t if v is not None]
            fl = 2 * len(terms) - 1
            if f1[i] > 0:
                f2 = add(f1, i, -1)
                a, b = self.need(e, f2, m), self.need(e, f2, m + 1)
                terms.append("(R)%d*o2q*(%s - roq*%s)" % (f1[i], a, b))
                fl += 5
            if e[i] > 0:
                e1 = add(e, i, -1)
                c = self.need(e1, f1, m + 1)
                terms.append("(R)%d*o2pq*%s" % (e[i], c))
                fl += 3
            name = self.new(" + ".join(terms), fl)
        else:
            i = pick(e)
            e1 = add(e, i, -1)
            t = [("PA%s" % XYZ[i], self.need(e1, f, m)), ("WP%s" % XYZ[i], self.need(e1, f, m + 1))]
            terms = ["%s*%s" % (c, v) for c, v in t if v is not None]
            fl = 2 * len(terms) - 1
            if e1[i] > 0:
                e2 = add(e1, i, -1)
                a, b = self.need(e2, f, m), self.need(e2, f, m + 1)
                terms.append("(R)%d*o2p*(%s - rop*%s)" % (e1[i], a, b))
                fl += 5
            name = self.new(" + ".join(terms), fl)
        self.memo[key] = name
        return name

    def emit(self):
        body = []
        nf = len(self.fs)
        for ie, e in enumerate(self.es):
            for jf, f in enumerate(self.fs):
                v = self.need(e, f, 0)
                body.append("  acc[%d] += %s;" % (ie * nf + jf, v))
                self.flops += 1
        L = self.la + self.lb + self.lc + self.ld
        sig = ("template <typename R>\n__device__ __forceinline__ void vrr_%d%d%d%d(const R* __restrict__ F, "
               "R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, R QCx, R QCy, R QCz, R WQx, R WQy, R WQz, "
               "R o2p, R o2q, R o2pq, R rop, R roq, R* __restrict__ acc) {" % (self.la, self.lb, self.lc, self.ld))
        unused = "  (void)PAx; (void)PAy; (void)PAz; (void)WPx; (void)WPy; (void)WPz; (void)QCx; (void)QCy; (void)QCz;\n" \
                 "  (void)WQx; (void)WQy; (void)WQz; (void)o2p; (void)o2q; (void)o2pq; (void)rop; (void)roq;"
        return "\n".join([sig, unused] + self.lines + body + ["}"]), L


class G12Gen:
    """Primitive [e0|exp(-g r12^2)|f0] for one Gaussian-geminal term.

    The integrand is a Gaussian in (x1, x2) per Cartesian axis, so each
    integral is a product of three one-dimensional moments
    I(i, k) = <(x1 - A)^i (x2 - C)^k> of that Gaussian.  With mean
    (mu1, mu2) and covariance S = (2 M)^-1, M = [[p + g, -g], [-g, q + g]],
    Gaussian integration by parts gives
        I(i+1, k) = (mu1 - A) I(i, k) + i S11 I(i-1, k) + k S12 I(i, k-1)
        I(i, k+1) = (mu2 - C) I(i, k) + i S12 I(i-1, k) + k S22 I(i, k-1)
    with S11 = (q+g)/(2D), S12 = g/(2D), S22 = (p+g)/(2D), D = pq + g(p+q),
    mu1 - A = PA - (g q / D) PQ, mu2 - C = QC + (g p / D) PQ.  The caller
    supplies these and the prefactor; ``acc`` has the layout of the Coulomb
    VRR's, so the same HRR finishes the quartet."""

    def __init__(self, la, lb, lc, ld):
        self.la, self.lb, self.lc, self.ld = la, lb, lc, ld
        self.es = [c for l in range(la, la + lb + 1) for c in comps(l)]
        self.fs = [c for l in range(lc, lc + ld + 1) for c in comps(l)]

    def emit(self):
        lab, lcd = self.la + self.lb, self.lc + self.ld
        lines = []
        names = {}
        for d, ax in enumerate(XYZ):
            tab = {}
            for k in range(lcd + 1):
                for i in range(lab + 1):
                    if i == 0 and k == 0:
                        tab[(i, k)] = "pref" if d == 0 else "(R)1"
                        continue
                    if k == 0:
                        terms = ["XA%s*%s" % (ax, tab[(i - 1, 0)])]
                        if i - 1 > 0:
                            terms.append("(R)%d*s11*%s" % (i - 1, tab[(i - 2, 0)]))
                    else:
                        terms = ["XC%s*%s" % (ax, tab[(i, k - 1)])]
                        if i > 0:
                            terms.append("(R)%d*s12*%s" % (i, tab[(i - 1, k - 1)]))
                        if k - 1 > 0:
                            terms.append("(R)%d*s22*%s" % (k - 1, tab[(i, k - 2)]))
                    name = "g%s%d_%d" % (ax, i, k)
                    lines.append("  const R %s = %s;" % (name, " + ".join(terms)))
                    tab[(i, k)] = name
            names[ax] = tab
        nf = len(self.fs)
        for ie, e in enumerate(self.es):
            for jf, f in enumerate(self.fs):
                fac = [names[ax][(e[d], f[d])] for d, ax in enumerate(XYZ)]
                fac = [x for x in fac if x != "(R)1"]
                lines.append("  acc[%d] += %s;" % (ie * nf + jf, "*".join(fac)))
        sig = ("template <typename R>\n__device__ __forceinline__ void g12vrr_%d%d%d%d(R XAx, R XAy, R XAz, "
               "R XCx, R XCy, R XCz, R s11, R s12, R s22, R pref, R* __restrict__ acc) {"
               % (self.la, self.lb, self.lc, self.ld))
        unused = "  (void)XAx; (void)XAy; (void)XAz; (void)XCx; (void)XCy; (void)XCz; (void)s11; (void)s12; (void)s22;"
        return "\n".join([sig, unused] + lines + ["}"])


class HRRGen:
    def __init__(self, la, lb, lc, ld):
        self.la, self.lb, self.lc, self.ld = la, lb, lc, ld
        self.lines = []
        self.memo = {}
        self.nv = 0
        self.flops = 0
        self.es = [c for l in range(la, la + lb + 1) for c in comps(l)]
        self.fs = [c for l in range(lc, lc + ld + 1) for c in comps(l)]
        self.eidx = {e: k for k, e in enumerate(self.es)}
        self.fidx = {f: k for k, f in enumerate(self.fs)}

    def new(self, expr):
        name = "h%d" % self.nv
        self.nv += 1
        self.lines.append("  const R %s = %s;" % (name, expr))
        self.flops += 2
        return name

    def bra(self, a, b, f):
        key = ("b", a, b, f)
        if key in self.memo:
            return self.memo[key]
        if sum(b) == 0:
            name = "acc[%d]" % (self.eidx[a] * len(self.fs) + self.fidx[f])
        else:
            i = pick(b)
            b1 = add(b, i, -1)
            x = self.bra(add(a, i, 1), b1, f)
            y = self.bra(a, b1, f)
            name = self.new("%s + AB%s*%s" % (x, XYZ[i], y))
        self.memo[key] = name
        return name

    def ket(self, a, b, c, d):
        key = ("k", a, b, c, d)
        if key in self.memo:
            return self.memo[key]
        if sum(d) == 0:
            name = self.bra(a, b, c)
        else:
            i = pick(d)
            d1 = add(d, i, -1)
            x = self.ket(a, b, add(c, i, 1), d1)
            y = self.ket(a, b, c, d1)
            name = self.new("%s + CD%s*%s" % (x, XYZ[i], y))
        self.memo[key] = name
        return name

    def emit(self):
        out = []
        la, lb, lc, ld = self.la, self.lb, self.lc, self.ld
        nb, nc, nd = ncart(lb), ncart(lc), ncart(ld)
        for ia, a in enumerate(comps(la)):
            for ib, b in enumerate(comps(lb)):
                for ic, c in enumerate(comps(lc)):
                    for id_, d in enumerate(comps(ld)):
                        v = self.ket(a, b, c, d)
                        out.append("  out[%d] = %s;" % (((ia * nb + ib) * nc + ic) * nd + id_, v))
        sig = ("template <typename R>\n__device__ __forceinline__ void hrr_%d%d%d%d(const R* __restrict__ acc, "
               "R ABx, R ABy, R ABz, R CDx, R CDy, R CDz, R* __restrict__ out) {" % (la, lb, lc, ld))
        unused = "  (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;"
        return "\n".join([sig, unused] + self.lines + out + ["}"])



# ------------------------------------------------------------------------
# Wave-cooperative programs for the high-L classes.
#
# One wavefront evaluates one quartet; its intermediates live in LDS and the
# 64 lanes share every recursion level.  The VRR and HRR above become
# "rounds" of at most 64 independent operations (one per lane), grouped by
# (dependency level, recursion kind) so that the kind, and with it the
# recursion's coefficient vectors, is known at compile time; each lane reads
# its operation's LDS operands from a per-class descriptor table.
#
# Per-wave LDS region (doubles):
#   [0, 18)        coefficient vectors: (PA_i, WP_i) at 2i, (QC_i, WQ_i) at
#                  6 + 2i, AB_i at 12 + i, CD_i at 15 + i
#   18, 19         ZERO (always 0.0: operands of absent terms; a pair, since
#                  a term reads the m and m+1 slots)
#   20             TRASH (target of idle lanes)
#   [22, 22+NACC)  contracted [e0|f0] accumulators
#   [.., +NOUT)    (ab|cd) block in the HRR's output order (aliases the
#                  accumulators when the class has no HRR)
#   [.., +W)       work slots: VRR values of one primitive quartet, later
#                  the HRR intermediates (slots reused by live range)
# Descriptor words (uint2, byte offsets into the region, multiples of 8 so
# that the low 3 bits carry small fields):
#   VRR  x = s0 | k1 | (s2 | k2) << 16      y = s4 | (dst | i) << 16
#        kind 0: V = PA_i V[s0] + WP_i V[s0+1] + k1 o2p (V[s2] - rop V[s2+1])
#        kind 1: V = QC_i V[s0] + WQ_i V[s0+1] + k1 o2q (V[s2] - roq V[s2+1])
#                    + k2 o2pq V[s4]
#   ACC  x = src, y = dst                      acc[dst] += V[src]
#   HRR  x = x | (y << 16), y = dst | i        V = V[x] + C_i V[y], C = AB (kind 0) or CD (kind 1)
COOP_P, COOP_ZERO, COOP_TRASH, COOP_ACC = 0, 18, 20, 22
# Rounds of one recursion level and kind are independent: a lane loads the
# descriptors and operands of up to COOP_GROUP of them before computing.
COOP_GROUP = 4   # kCoopGroup of eri_coop.h


def coop_groups(rounds):
    """Consecutive rounds with equal (level, kind), in pieces of at most
    COOP_GROUP: [(kind, nrounds)]."""
    out = []
    prev = None
    for kind, _, key in rounds:
        if key == prev and out[-1][1] < COOP_GROUP:
            out[-1] = (kind, out[-1][1] + 1)
        else:
            out.append((kind, 1))
        prev = key
    return out


def pack_groups(groups):
    """4 bits per group (rounds - 1, kind << 3), 16 groups per word."""
    words = [0, 0]
    assert len(groups) <= 32
    for g, (kind, nr) in enumerate(groups):
        words[g // 16] |= ((nr - 1) | (kind << 3)) << (4 * (g % 16))
    return words


class CoopGen:
    def __init__(self, la, lb, lc, ld):
        self.cls = (la, lb, lc, ld)
        self.L = la + lb + lc + ld
        self.es = [c for l in range(la, la + lb + 1) for c in comps(l)]
        self.fs = [c for l in range(lc, lc + ld + 1) for c in comps(l)]
        self.nacc = len(self.es) * len(self.fs)
        self.nout = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)

    # ---- VRR dependency graph (the recursion VRRGen emits)
    def _vrr_graph(self):
        ops = {}     # key -> (kind, i, k1, k2, s0key, s2key, s4key)

        def visit(key):
            if key in ops:
                return
            e, f, m = key
            if sum(e) == 0 and sum(f) == 0:
                ops[key] = None
                return
            if sum(f) > 0:
                i = pick(f)
                f1 = add(f, i, -1)
                s0 = (e, f1, m)
                s2 = s4 = None
                k1 = k2 = 0
                if f1[i] > 0:
                    s2, k1 = (e, add(f1, i, -1), m), f1[i]
                if e[i] > 0:
                    s4, k2 = (add(e, i, -1), f1, m + 1), e[i]
                ops[key] = (1, i, k1, k2, s0, s2, s4)
            else:
                i = pick(e)
                e1 = add(e, i, -1)
                s0 = (e1, f, m)
                s2 = None
                k1 = 0
                if e1[i] > 0:
                    s2, k1 = (add(e1, i, -1), f, m), e1[i]
                ops[key] = (0, i, k1, 0, s0, s2, None)
            _, _, _, _, s0, s2, s4 = ops[key]
            for s in (s0, s2):
                if s is not None:
                    visit(s)
                    visit((s[0], s[1], s[2] + 1))
            if s4 is not None:
                visit(s4)

        for e in self.es:
            for f in self.fs:
                visit((e, f, 0))
        return ops

    @staticmethod
    def _alloc(runs, nrounds):
        """Linear scan over rounds.  runs: name -> [size, def_round, last_use];
        a run is placed before its defining round and released after its last
        use.  Returns (offsets, peak)."""
        free = [(0, 1 << 30)]
        off = {}
        peak = 0
        order = sorted(runs, key=lambda r: (runs[r][1], -runs[r][0]))
        by_def = {}
        for r in order:
            by_def.setdefault(runs[r][1], []).append(r)
        by_end = {}
        for r in runs:
            by_end.setdefault(runs[r][2], []).append(r)
        for rnd in range(-1, nrounds + 1):
            for r in by_def.get(rnd, []):
                size = runs[r][0]
                for k, (a, n) in enumerate(free):
                    if n >= size:
                        off[r] = a
                        free[k] = (a + size, n - size)
                        peak = max(peak, a + size)
                        break
            for r in by_end.get(rnd, []):
                a, size = off[r], runs[r][0]
                free.append((a, size))
                free.sort()
                merged = []
                for x in free:
                    if merged and merged[-1][0] + merged[-1][1] == x[0]:
                        merged[-1] = (merged[-1][0], merged[-1][1] + x[1])
                    else:
                        merged.append(x)
                free = merged
        return off, peak

    def build(self):
        L = self.L
        ops = self._vrr_graph()
        level = {}

        def lev(key):
            if key in level:
                return level[key]
            op = ops[key]
            if op is None:
                level[key] = 0
                return 0
            _, _, _, _, s0, s2, s4 = op
            srcs = [s0, (s0[0], s0[1], s0[2] + 1)]
            if s2 is not None:
                srcs += [s2, (s2[0], s2[1], s2[2] + 1)]
            if s4 is not None:
                srcs.append(s4)
            level[key] = 1 + max(lev(s) for s in srcs)
            return level[key]

        for k in ops:
            lev(k)
        # rounds: (level, kind) groups of <= 64 operations
        groups = {}
        for k, op in ops.items():
            if op is not None:
                groups.setdefault((level[k], op[0]), []).append(k)
        vrounds = []
        for g in sorted(groups):
            ks = sorted(groups[g], key=lambda k: (k[0], k[1], k[2]))
            for c in range(0, len(ks), 64):
                vrounds.append((g[1], ks[c:c + 64], g))
        rnd_of = {}
        for r, (_, ks, _) in enumerate(vrounds):
            for k in ks:
                rnd_of[k] = r
        # runs: one per (e, f), m contiguous from the lowest needed m
        mlo, mhi = {}, {}
        for (e, f, m) in ops:
            mlo[(e, f)] = min(mlo.get((e, f), m), m)
            mhi[(e, f)] = max(mhi.get((e, f), m), m)
        runs = {}
        for ef in mlo:
            d = -1 if ef == ((0, 0, 0), (0, 0, 0)) else min(rnd_of[(ef[0], ef[1], m)] for m in range(mlo[ef], mhi[ef] + 1)
                                                            if (ef[0], ef[1], m) in rnd_of)
            runs[ef] = [mhi[ef] - mlo[ef] + 1, d, d]
        nvr = len(vrounds)
        for k, op in ops.items():
            if op is None:
                continue
            _, _, _, _, s0, s2, s4 = op
            for s in (s0, s2, s4):
                if s is not None:
                    ef = (s[0], s[1])
                    runs[ef][2] = max(runs[ef][2], rnd_of[k])
        for e in self.es:
            for f in self.fs:
                runs[(e, f)][2] = nvr          # read by the accumulation
        assert runs[((0, 0, 0), (0, 0, 0))][0] == L + 1
        voff, vpeak = self._alloc(runs, nvr + 1)
        # ---- HRR graph (the recursion HRRGen emits)
        eidx = {e: k for k, e in enumerate(self.es)}
        fidx = {f: k for k, f in enumerate(self.fs)}
        nf = len(self.fs)
        hops = {}    # key -> (kind, i, xkey, ykey)

        def bra(a, b, f):
            key = ("b", a, b, f)
            if key in hops:
                return key
            if sum(b) == 0:
                hops[key] = ("acc", eidx[a] * nf + fidx[f])
            else:
                i = pick(b)
                b1 = add(b, i, -1)
                hops[key] = (0, i, bra(add(a, i, 1), b1, f), bra(a, b1, f))
            return key

        def ket(a, b, c, d):
            key = ("k", a, b, c, d)
            if key in hops:
                return key
            if sum(d) == 0:
                hops[key] = ("alias", bra(a, b, c))
            else:
                i = pick(d)
                d1 = add(d, i, -1)
                hops[key] = (1, i, ket(a, b, add(c, i, 1), d1), ket(a, b, c, d1))
            return key

        la, lb, lc, ld = self.cls
        nb, nc, nd = ncart(lb), ncart(lc), ncart(ld)
        outs = {}
        for ia, a in enumerate(comps(la)):
            for ib, b in enumerate(comps(lb)):
                for ic, c in enumerate(comps(lc)):
                    for id_, d in enumerate(comps(ld)):
                        outs[ket(a, b, c, d)] = ((ia * nb + ib) * nc + ic) * nd + id_

        def resolve(key):
            while hops[key][0] == "alias":
                key = hops[key][1]
            return key

        hlevel = {}

        def hlev(key):
            key = resolve(key)
            if key in hlevel:
                return hlevel[key]
            op = hops[key]
            hlevel[key] = 0 if op[0] == "acc" else 1 + max(hlev(op[2]), hlev(op[3]))
            return hlevel[key]

        real = {}     # insertion-ordered: the generated tables do not depend on string hashing
        for k in list(hops):
            rk = resolve(k)
            hlev(rk)
            if hops[rk][0] != "acc":
                real[rk] = None
        final = {}
        for k, idx in outs.items():
            final[resolve(k)] = idx
        has_hrr = len(real) > 0
        if has_hrr:
            assert all(resolve(k) in real for k in outs), "mixed HRR outputs"
        hgroups = {}
        for k in real:
            hgroups.setdefault((hlevel[k], hops[k][0]), []).append(k)
        hrounds = []
        for g in sorted(hgroups):
            ks = sorted(hgroups[g], key=str)
            for c in range(0, len(ks), 64):
                hrounds.append((g[1], ks[c:c + 64], g))
        hrnd = {}
        for r, (_, ks, _) in enumerate(hrounds):
            for k in ks:
                hrnd[k] = r
        hruns = {}
        for k in real:
            if k in final:
                continue
            hruns[k] = [1, hrnd[k], hrnd[k]]
        for k in real:
            for s in (hops[k][2], hops[k][3]):
                s = resolve(s)
                if s in hruns:
                    hruns[s][2] = max(hruns[s][2], hrnd[k])
        hoff, hpeak = self._alloc(hruns, len(hrounds) + 1)
        # ---- layout
        acc0 = COOP_ACC
        out0 = acc0 + self.nacc if has_hrr else acc0
        work0 = out0 + (self.nout if has_hrr else 0)
        W = max(vpeak, hpeak)
        region = work0 + W
        assert region * 8 < 65536

        def vslot(key):
            e, f, m = key
            return work0 + voff[(e, f)] + (m - mlo[(e, f)])

        def b(slot):
            return 8 * slot

        words = []
        idle_v = (b(COOP_ZERO), b(COOP_TRASH) << 16)
        for kind, ks, _ in vrounds:
            row = []
            for k in ks:
                _, i, k1, k2, s0, s2, s4 = ops[k]
                x = (b(vslot(s0)) | k1) | ((b(vslot(s2)) if s2 else b(COOP_ZERO)) | k2) << 16
                y = (b(vslot(s4)) if s4 else b(COOP_ZERO)) | ((b(vslot(k)) | i) << 16)
                row.append((x, y))
            row += [(b(COOP_ZERO) | (b(COOP_ZERO) << 16), b(COOP_ZERO) | (b(COOP_TRASH) << 16))] * (64 - len(row))
            words += row
        nacc_r = (self.nacc + 63) // 64
        tgt = [(e, f) for e in self.es for f in self.fs]
        for r in range(nacc_r):
            row = []
            for t in range(r * 64, min(self.nacc, r * 64 + 64)):
                e, f = tgt[t]
                row.append((b(vslot((e, f, 0))), b(acc0 + t)))
            row += [(b(COOP_ZERO), b(COOP_TRASH))] * (64 - len(row))
            words += row

        def hslot(key):
            key = resolve(key)
            op = hops[key]
            if op[0] == "acc":
                return acc0 + op[1]
            if key in final:
                return out0 + final[key]
            return work0 + hoff[key]

        for kind, ks, _ in hrounds:
            row = []
            for k in ks:
                _, i, xk, yk = hops[k]
                row.append((b(hslot(xk)) | (b(hslot(yk)) << 16), b(hslot(k)) | i))
            row += [(b(COOP_ZERO) | (b(COOP_ZERO) << 16), b(COOP_TRASH))] * (64 - len(row))
            words += row
        # idle rounds at the end: the kernels load the next group's
        # descriptors one group ahead
        words += [(b(COOP_ZERO) | (b(COOP_ZERO) << 16), b(COOP_ZERO) | (b(COOP_TRASH) << 16))] * (64 * COOP_GROUP)
        self.info = {"vrr_rounds": len(vrounds), "acc_rounds": nacc_r, "hrr_rounds": len(hrounds),
                     "region": region, "work": W, "vpeak": vpeak, "hpeak": hpeak,
                     "base": vslot(((0, 0, 0), (0, 0, 0), 0))}
        self.groups = (coop_groups(vrounds), coop_groups([(0, None, "acc")] * nacc_r), coop_groups(hrounds))
        return vrounds, hrounds, words, out0

    def emit(self):
        la, lb, lc, ld = self.cls
        k = "%d%d%d%d" % self.cls
        vrounds, hrounds, words, out0 = self.build()
        I = self.info
        lines = ["// Generated by gen_eri.py -- do not edit.",
                 "// class (%d%d|%d%d): wave-cooperative program, %d VRR + %d accumulation + %d HRR rounds, "
                 "%d doubles of LDS per wavefront" % (la, lb, lc, ld, I["vrr_rounds"], I["acc_rounds"],
                                                     I["hrr_rounds"], I["region"]),
                 "static __device__ const uint2 coop_prog_%s[%d] = {" % (k, len(words))]
        vg, ag, hg = self.groups
        for c in range(0, len(words), 4):
            lines.append("  " + ", ".join("{0x%xu, 0x%xu}" % w for w in words[c:c + 4]) + ",")
        lines += ["};",
                 "template <>", "struct CoopProgram<%d, %d, %d, %d> {" % self.cls,
                 "  __device__ static __forceinline__ const uint2* prog() { return coop_prog_%s; }" % k,
                 "  static_assert(kCoopZero == %d && kCoopAcc == %d, \"LDS layout of gen_eri.py\");" % (COOP_ZERO, COOP_ACC),
                 "  static constexpr int kVrrRounds = %d;" % I["vrr_rounds"],
                 "  static constexpr int kAccRounds = %d;" % I["acc_rounds"],
                 "  static constexpr int kHrrRounds = %d;" % I["hrr_rounds"],
                 "  static constexpr int kRegion = %d;" % I["region"],
                 "  static constexpr int kOut = %d;" % out0,
                 "  static constexpr int kBase = %d;" % I["base"],
                 "  // groups of independent rounds: 4 bits each (rounds - 1, kind << 3)",
                 "  static constexpr int kVrrGroups = %d;" % len(vg),
                 "  static constexpr unsigned long long kVrrG[2] = {0x%xull, 0x%xull};" % tuple(pack_groups(vg)),
                 "  static constexpr int kAccGroups = %d;" % len(ag),
                 "  static constexpr unsigned long long kAccG[2] = {0x%xull, 0x%xull};" % tuple(pack_groups(ag)),
                 "  static constexpr int kHrrGroups = %d;" % len(hg),
                 "  static constexpr unsigned long long kHrrG[2] = {0x%xull, 0x%xull};" % tuple(pack_groups(hg)),
                 "};"]
        return "\n".join(lines) + "\n"


KERNEL_TU = """// Generated by gen_eri.py -- do not edit.
// Kernels of quartet class (%(a)d%(b)d|%(c)d%(d)d): fused ERI + J/K, coupling%(schw)s%(coopnote)s.
#include "../eri_coop.h"

namespace lowdin {
#include "eri_%(k)s.inc"
LOWDIN_ERI_CLASS(%(a)d, %(b)d, %(c)d, %(d)d)
%(coopinc)s
void launch_jk_%(k)s(const JKArgs& args, hipStream_t s) { launch_jk_t<%(a)d, %(b)d, %(c)d, %(d)d>(args, s); }
void launch_jks_%(k)s(const JKArgs& args, hipStream_t s) { launch_jk_strip_t<%(a)d, %(b)d, %(c)d, %(d)d>(args, s); }
void launch_cp_%(k)s(const CouplingArgs& args, hipStream_t s) { launch_cp_t<%(a)d, %(b)d, %(c)d, %(d)d>(args, s); }
%(schwfn)s%(coopfn)s}  // namespace lowdin
"""

# classes with a wave-cooperative J/K kernel (total L >= 5)
COOP_MIN_L = 5


STORE_TU = """// Generated by gen_eri.py -- do not edit.
// Kernels of quartet class (%(a)d%(b)d|%(c)d%(d)d): integral store and first-quarter transform.
#include "../eri_store.h"

namespace lowdin {
#include "eri_%(k)s.inc"
LOWDIN_ERI_CLASS(%(a)d, %(b)d, %(c)d, %(d)d)

void launch_st_%(k)s(const StoreArgs& args, hipStream_t s) { launch_store_t<%(a)d, %(b)d, %(c)d, %(d)d>(args, s); }
void launch_cst_%(k)s(const CouplingStoreArgs& args, hipStream_t s) { launch_cstore_t<%(a)d, %(b)d, %(c)d, %(d)d>(args, s); }
}  // namespace lowdin
"""


G12_TU = """// Generated by gen_eri.py -- do not edit.
// Kernels of quartet class (%(a)d%(b)d|%(c)d%(d)d): G12 (Gaussian-geminal) Fock build and coupling.
#include "../eri_g12.h"

namespace lowdin {
#include "eri_%(k)s.inc"
LOWDIN_ERI_CLASS(%(a)d, %(b)d, %(c)d, %(d)d)
LOWDIN_G12_CLASS(%(a)d, %(b)d, %(c)d, %(d)d)

void launch_gjk_%(k)s(const G12Args& args, hipStream_t s) { launch_g12_t<%(a)d, %(b)d, %(c)d, %(d)d>(args, s); }
void launch_gcp_%(k)s(const G12CouplingArgs& args, hipStream_t s) { launch_g12cp_t<%(a)d, %(b)d, %(c)d, %(d)d>(args, s); }
}  // namespace lowdin
"""


def main(outdir):
    pcs, classes = canonical_classes()
    gdir = os.path.join(outdir, "gen")
    os.makedirs(gdir, exist_ok=True)
    table = {}
    for (la, lb, lc, ld) in classes:
        k = "%d%d%d%d" % (la, lb, lc, ld)
        v = VRRGen(la, lb, lc, ld)
        code, L = v.emit()
        h = HRRGen(la, lb, lc, ld)
        hcode = h.emit()
        nacc = len(v.es) * len(v.fs)
        inc = ["// Generated by gen_eri.py -- do not edit.",
               "// class (%d%d|%d%d): L = %d, %d accumulators, VRR %d flops/primitive quartet, HRR %d flops"
               % (la, lb, lc, ld, L, nacc, v.flops, h.flops),
               "#define NACC_%s %d" % (k, nacc), code, hcode, G12Gen(la, lb, lc, ld).emit(), ""]
        with open(os.path.join(gdir, "eri_%s.inc" % k), "w") as fh:
            fh.write("\n".join(inc))
        diag = (la, lb) == (lc, ld)
        schwfn = ("void launch_schwarz_%d%d(PairInfo* pairs, int npairs, const double* prim, const double* boys, "
                  "const double* norma, hipStream_t s) { launch_schwarz_t<%d, %d>(pairs, npairs, prim, boys, norma, s); }\n"
                  % (la, lb, la, lb)
                  if diag else "")
        coop = la + lb + lc + ld >= COOP_MIN_L
        coopinc = coopfn = ""
        if coop:
            with open(os.path.join(gdir, "c_%s.inc" % k), "w") as fh:
                fh.write(CoopGen(la, lb, lc, ld).emit())
            coopinc = '#include "c_%s.inc"\n' % k
            coopfn = ("bool launch_jkc_%s(const JKArgs& args, hipStream_t s) { return launch_jk_coop_t<%d, %d, %d, %d>(args, s); }\n"
                      "bool launch_cpc_%s(const CouplingArgs& args, hipStream_t s) { return launch_cp_coop_t<%d, %d, %d, %d>(args, s); }\n"
                      "void coop_attr_%s() { coop_attr_t<%d, %d, %d, %d>(); }\n"
                      % (k, la, lb, lc, ld, k, la, lb, lc, ld, k, la, lb, lc, ld))
        with open(os.path.join(gdir, "k_%s.hip" % k), "w") as fh:
            fh.write(KERNEL_TU % {"a": la, "b": lb, "c": lc, "d": ld, "k": k,
                                  "schw": ", Schwarz" if diag else "", "schwfn": schwfn,
                                  "coopnote": ", wave-cooperative J/K and coupling" if coop else "",
                                  "coopinc": coopinc, "coopfn": coopfn})
        with open(os.path.join(gdir, "g_%s.hip" % k), "w") as fh:
            fh.write(G12_TU % {"a": la, "b": lb, "c": lc, "d": ld, "k": k})
        with open(os.path.join(gdir, "s_%s.hip" % k), "w") as fh:
            fh.write(STORE_TU % {"a": la, "b": lb, "c": lc, "d": ld, "k": k})
        table[k] = {"L": L, "nacc": nacc, "vrr_flops": v.flops, "hrr_flops": h.flops}
    with open(os.path.join(outdir, "ops_table.json"), "w") as fh:
        json.dump({"note": "floating-point operations of the generated Obara-Saika VRR (per primitive "
                           "quartet, excluding Boys) and HGP HRR (per contracted quartet)",
                   "classes": table}, fh, indent=1)
    # dispatch table
    lines = ["// Generated by gen_eri.py -- do not edit.", "#pragma once", '#include "../lowdin_kernels.h"',
             "namespace lowdin {"]
    for (la, lb, lc, ld) in classes:
        k = "%d%d%d%d" % (la, lb, lc, ld)
        lines.append("void launch_jk_%s(const JKArgs&, hipStream_t);" % k)
        lines.append("void launch_jks_%s(const JKArgs&, hipStream_t);" % k)
        lines.append("void launch_cp_%s(const CouplingArgs&, hipStream_t);" % k)
        lines.append("void launch_st_%s(const StoreArgs&, hipStream_t);" % k)
        lines.append("void launch_cst_%s(const CouplingStoreArgs&, hipStream_t);" % k)
        lines.append("void launch_gjk_%s(const G12Args&, hipStream_t);" % k)
        lines.append("void launch_gcp_%s(const G12CouplingArgs&, hipStream_t);" % k)
        if la + lb + lc + ld >= COOP_MIN_L:
            lines.append("bool launch_jkc_%s(const JKArgs&, hipStream_t);" % k)
            lines.append("bool launch_cpc_%s(const CouplingArgs&, hipStream_t);" % k)
            lines.append("void coop_attr_%s();" % k)
    for (la, lb) in pcs:
        lines.append("void launch_schwarz_%d%d(PairInfo*, int, const double*, const double*, const double*, hipStream_t);"
                     % (la, lb))
    lines.append("inline bool dispatch_jk(int la, int lb, int lc, int ld, const JKArgs& a, hipStream_t s) {")
    lines.append("  switch (la * 1000 + lb * 100 + lc * 10 + ld) {")
    for (la, lb, lc, ld) in classes:
        lines.append("    case %d: launch_jk_%d%d%d%d(a, s); return true;" % (la * 1000 + lb * 100 + lc * 10 + ld, la, lb, lc, ld))
    lines.append("    default: return false;\n  }\n}")
    lines.append("inline bool dispatch_jk_strip(int la, int lb, int lc, int ld, const JKArgs& a, hipStream_t s) {")
    lines.append("  switch (la * 1000 + lb * 100 + lc * 10 + ld) {")
    for (la, lb, lc, ld) in classes:
        lines.append("    case %d: launch_jks_%d%d%d%d(a, s); return true;" % (la * 1000 + lb * 100 + lc * 10 + ld, la, lb, lc, ld))
    lines.append("    default: return false;\n  }\n}")
    lines.append("// wave-cooperative J/K kernel; false when the class has none or the launch")
    lines.append("// does not fit in LDS")
    lines.append("inline bool dispatch_jk_coop(int la, int lb, int lc, int ld, const JKArgs& a, hipStream_t s) {")
    lines.append("  switch (la * 1000 + lb * 100 + lc * 10 + ld) {")
    for (la, lb, lc, ld) in classes:
        if la + lb + lc + ld >= COOP_MIN_L:
            lines.append("    case %d: return launch_jkc_%d%d%d%d(a, s);" % (la * 1000 + lb * 100 + lc * 10 + ld, la, lb, lc, ld))
    lines.append("    default: return false;\n  }\n}")
    lines.append("inline bool has_jk_coop(int la, int lb, int lc, int ld) { return la + lb + lc + ld >= %d; }" % COOP_MIN_L)
    lines.append("// raise the dynamic LDS limit of every wave-cooperative kernel (once per process)")
    lines.append("inline void coop_attr_all() {")
    for (la, lb, lc, ld) in classes:
        if la + lb + lc + ld >= COOP_MIN_L:
            lines.append("  coop_attr_%d%d%d%d();" % (la, lb, lc, ld))
    lines.append("}")
    lines.append("// wave-cooperative coupling kernel; false when the class has none")
    lines.append("inline bool dispatch_cp_coop(int la, int lb, int lc, int ld, const CouplingArgs& a, hipStream_t s) {")
    lines.append("  switch (la * 1000 + lb * 100 + lc * 10 + ld) {")
    for (la, lb, lc, ld) in classes:
        if la + lb + lc + ld >= COOP_MIN_L:
            lines.append("    case %d: return launch_cpc_%d%d%d%d(a, s);" % (la * 1000 + lb * 100 + lc * 10 + ld, la, lb, lc, ld))
    lines.append("    default: return false;\n  }\n}")
    lines.append("inline bool dispatch_cp(int la, int lb, int lc, int ld, const CouplingArgs& a, hipStream_t s) {")
    lines.append("  switch (la * 1000 + lb * 100 + lc * 10 + ld) {")
    for (la, lb, lc, ld) in classes:
        lines.append("    case %d: launch_cp_%d%d%d%d(a, s); return true;" % (la * 1000 + lb * 100 + lc * 10 + ld, la, lb, lc, ld))
    lines.append("    default: return false;\n  }\n}")
    lines.append("inline bool dispatch_store(int la, int lb, int lc, int ld, const StoreArgs& a, hipStream_t s) {")
    lines.append("  switch (la * 1000 + lb * 100 + lc * 10 + ld) {")
    for (la, lb, lc, ld) in classes:
        lines.append("    case %d: launch_st_%d%d%d%d(a, s); return true;" % (la * 1000 + lb * 100 + lc * 10 + ld, la, lb, lc, ld))
    lines.append("    default: return false;\n  }\n}")
    lines.append("inline bool dispatch_cstore(int la, int lb, int lc, int ld, const CouplingStoreArgs& a, hipStream_t s) {")
    lines.append("  switch (la * 1000 + lb * 100 + lc * 10 + ld) {")
    for (la, lb, lc, ld) in classes:
        lines.append("    case %d: launch_cst_%d%d%d%d(a, s); return true;" % (la * 1000 + lb * 100 + lc * 10 + ld, la, lb, lc, ld))
    lines.append("    default: return false;\n  }\n}")
    lines.append("inline bool dispatch_g12(int la, int lb, int lc, int ld, const G12Args& a, hipStream_t s) {")
    lines.append("  switch (la * 1000 + lb * 100 + lc * 10 + ld) {")
    for (la, lb, lc, ld) in classes:
        lines.append("    case %d: launch_gjk_%d%d%d%d(a, s); return true;" % (la * 1000 + lb * 100 + lc * 10 + ld, la, lb, lc, ld))
    lines.append("    default: return false;\n  }\n}")
    lines.append("inline bool dispatch_g12cp(int la, int lb, int lc, int ld, const G12CouplingArgs& a, hipStream_t s) {")
    lines.append("  switch (la * 1000 + lb * 100 + lc * 10 + ld) {")
    for (la, lb, lc, ld) in classes:
        lines.append("    case %d: launch_gcp_%d%d%d%d(a, s); return true;" % (la * 1000 + lb * 100 + lc * 10 + ld, la, lb, lc, ld))
    lines.append("    default: return false;\n  }\n}")
    lines.append("inline bool dispatch_schwarz(int la, int lb, PairInfo* p, int np, const double* prim, const double* boys, "
                 "const double* norma, hipStream_t s) {")
    lines.append("  switch (la * 10 + lb) {")
    for (la, lb) in pcs:
        lines.append("    case %d: launch_schwarz_%d%d(p, np, prim, boys, norma, s); return true;" % (la * 10 + lb, la, lb))
    lines.append("    default: return false;\n  }\n}")
    lines.append("}  // namespace lowdin")
    with open(os.path.join(gdir, "dispatch.h"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return table


if __name__ == "__main__":
    d = sys.argv[1] if len(sys.argv) > 1 else os.path.dirname(os.path.abspath(__file__))
    t = main(d)
    for k, v in t.items():
        print(k, v)
