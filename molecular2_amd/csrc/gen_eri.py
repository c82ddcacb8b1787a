"""Generate straight-line HIP device code for contracted shell-quartet ERIs.

For every angular-momentum class (la lb | lc ld) with la >= lb, lc >= ld it
emits ``EriClass<la,lb,lc,ld>::run<R, PURE>(...)``:

  1. loop over bra primitive pairs x ket primitive pairs (pair data prepared
     on the device): Boys F_0..F_L, then the Obara-Saika vertical recurrence
     (HGP form: bra build [e0|00]^(m), ket build [e0|f0]^(m) with the
     electron-transfer term), accumulating the contracted [e0|f0];
  2. horizontal recurrences (ket then bra) to (ab|cd) in Cartesian form;
  3. optional Cartesian -> real-solid-harmonic transform (PURE) for l >= 2.

Only the recurrence nodes reachable from the targets are emitted, each as a
``const R`` so the compiler keeps everything in VGPRs.  The generator is
deterministic; its output is written to csrc/generated/eri_classes.inc and
compiled by hipcc (build step in molecular2_amd/build.py).

The derivative variant (gradient) emits ``EriDeriv<la,lb,lc,ld>::run`` that
accumulates the contracted, exponent-weighted [e0|f0] needed for
d/dA, d/dB, d/dC of (ab|cd) and contracts them with a 4-index density weight.
"""
from __future__ import annotations

import math
import os
import sys
from fractions import Fraction

AX = "xyz"


def comps(L):
    return [(x, y, L - x - y) for x in range(L, -1, -1) for y in range(L - x, -1, -1)]


def cidx(c):
    L = sum(c)
    x, y, _ = c
    return (L - x) * (L - x + 1) // 2 + (L - x - y)


def sub(c, i, k=1):
    t = list(c)
    t[i] -= k
    return tuple(t)


def add(c, i, k=1):
    t = list(c)
    t[i] += k
    return tuple(t)


def cname(c):
    return "%d%d%d" % c


# ---------------------------------------------------------------------------
# real solid harmonics (Helgaker, Joergensen, Olsen, eq. 6.4.47-6.4.50),
# normalised against the x^l-normalised Cartesian metric.
# ---------------------------------------------------------------------------

def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def solid_harmonics(l):
    """rows m=-l..l, columns Cartesian components (standard order)."""
    cs = comps(l)
    rows = []
    for m in range(-l, l + 1):
        am = abs(m)
        vm = Fraction(1, 2) if m < 0 else Fraction(0)
        row = [0.0] * len(cs)
        for t in range((l - am) // 2 + 1):
            for u in range(t + 1):
                vmax2 = math.floor(am / 2 - vm)
                for vi in range(vmax2 + 1):
                    v = vi + vm
                    two_v = int(2 * v)
                    sgn = -1.0 if (t + int(v - vm)) % 2 else 1.0
                    C = sgn * (0.25 ** t) * math.comb(l, t) * math.comb(l - t, am + t) * \
                        math.comb(t, u) * math.comb(am, two_v)
                    px = 2 * t + am - 2 * u - two_v
                    py = 2 * u + two_v
                    pz = l - 2 * t - am
                    if px < 0 or pz < 0:
                        continue
                    row[cidx((px, py, pz))] += C
        nrm = 0.0
        for i, ci in enumerate(cs):
            for j, cj in enumerate(cs):
                if row[i] == 0 or row[j] == 0:
                    continue
                if any((ci[k] + cj[k]) % 2 for k in range(3)):
                    continue
                nrm += row[i] * row[j] * dfact(ci[0] + cj[0] - 1) * dfact(ci[1] + cj[1] - 1) * \
                    dfact(ci[2] + cj[2] - 1) / dfact(2 * l - 1)
        rows.append([x / math.sqrt(nrm) for x in row])
    return rows


def lit(x):
    r = repr(float(x))
    if "e" not in r and "." not in r:
        r += ".0"
    return "R(%s)" % r


# ---------------------------------------------------------------------------
# VRR / HRR DAG construction
# ---------------------------------------------------------------------------

class Emitter:
    def __init__(self):
        self.lines = []
        self.done = set()

    def out(self, s):
        self.lines.append(s)


def vrr_name(e, f, m):
    return "v%s_%s_%d" % (cname(e), cname(f), m)


def build_vrr(targets, L):
    """targets: list of (e, f) needed at m=0.  Returns ordered node list with
    expressions."""
    nodes = {}

    def need(e, f, m):
        key = (e, f, m)
        if key in nodes:
            return
        if sum(f) > 0:
            best = None
            for i in range(3):
                if f[i] == 0:
                    continue
                score = (f[i] >= 2) + (e[i] > 0)
                if best is None or score < best[0]:
                    best = (score, i)
            i = best[1]
            f1 = sub(f, i)
            terms = [("QC%s" % AX[i], (e, f1, m)), ("WQ%s" % AX[i], (e, f1, m + 1))]
            deps = [(e, f1, m), (e, f1, m + 1)]
            extra = []
            if f1[i] > 0:
                f2 = sub(f1, i)
                extra.append(("ff", f1[i], (e, f2, m), (e, f2, m + 1)))
                deps += [(e, f2, m), (e, f2, m + 1)]
            if e[i] > 0:
                e1 = sub(e, i)
                extra.append(("ef", e[i], (e1, f1, m + 1)))
                deps.append((e1, f1, m + 1))
            nodes[key] = ("ket", terms, extra)
            for d in deps:
                need(*d)
        elif sum(e) > 0:
            best = None
            for i in range(3):
                if e[i] == 0:
                    continue
                score = (e[i] >= 2)
                if best is None or score < best[0]:
                    best = (score, i)
            i = best[1]
            e1 = sub(e, i)
            terms = [("PA%s" % AX[i], (e1, f, m)), ("WP%s" % AX[i], (e1, f, m + 1))]
            deps = [(e1, f, m), (e1, f, m + 1)]
            extra = []
            if e1[i] > 0:
                e2 = sub(e1, i)
                extra.append(("ee", e1[i], (e2, f, m), (e2, f, m + 1)))
                deps += [(e2, f, m), (e2, f, m + 1)]
            nodes[key] = ("bra", terms, extra)
            for d in deps:
                need(*d)
        else:
            nodes[key] = ("base", m)

    for e, f in targets:
        need(e, f, 0)
    order = sorted(nodes.keys(), key=lambda k: (sum(k[0]) + sum(k[1]), sum(k[1]), k[2]))
    return order, nodes


def emit_vrr(em, order, nodes, indent):
    for key in order:
        e, f, m = key
        kind = nodes[key]
        name = vrr_name(e, f, m)
        if kind[0] == "base":
            em.out(indent + "const R %s = Fm[%d];" % (name, m))
            continue
        _, terms, extra = kind
        expr = " + ".join("%s * %s" % (c, vrr_name(*k)) for c, k in terms)
        for x in extra:
            if x[0] == "ff":
                expr += " + %s * inv2q * (%s - rho_q * %s)" % (lit(x[1]), vrr_name(*x[2]), vrr_name(*x[3]))
            elif x[0] == "ee":
                expr += " + %s * inv2p * (%s - rho_p * %s)" % (lit(x[1]), vrr_name(*x[2]), vrr_name(*x[3]))
            elif x[0] == "ef":
                expr += " + %s * inv2pq * %s" % (lit(x[1]), vrr_name(*x[2]))
        em.out(indent + "const R %s = %s;" % (name, expr))


def hrr_ket(la, lb, lc, ld):
    """nodes for (e | c d) from acc(e, f); returns function name(e,c,d) and
    ordered definitions."""
    defs = []
    seen = {}

    def node(e, c, d):
        key = (e, c, d)
        if key in seen:
            return seen[key]
        if sum(d) == 0:
            nm = "a%s_%s" % (cname(e), cname(c))
            seen[key] = nm
            return nm
        i = next(k for k in range(3) if d[k] > 0)
        n1 = node(e, add(c, i), sub(d, i))
        n0 = node(e, c, sub(d, i))
        nm = "k%s_%s_%s" % (cname(e), cname(c), cname(d))
        defs.append("const R %s = %s + CD%s * %s;" % (nm, n1, AX[i], n0))
        seen[key] = nm
        return nm

    return node, defs


def gen_class(la, lb, lc, ld, deriv=False):
    L = la + lb + lc + ld
    em = Emitter()
    E = [c for Le in range(la, la + lb + 1) for c in comps(Le)]
    F = [c for Lf in range(lc, lc + ld + 1) for c in comps(Lf)]
    targets = [(e, f) for e in E for f in F]
    order, nodes = build_vrr(targets, L)
    na, nb, nc, nd = [(l + 1) * (l + 2) // 2 for l in (la, lb, lc, ld)]
    em.out("template <> struct EriClass<%d, %d, %d, %d> {" % (la, lb, lc, ld))
    em.out("  static constexpr int L = %d;" % L)
    em.out("  static constexpr int NCART = %d;" % (na * nb * nc * nd))
    em.out("  template <typename R, bool PURE>")
    em.out("  static __device__ __forceinline__ void run(const PrimPair* __restrict__ bra, int nbra,")
    em.out("      const PrimPair* __restrict__ ket, int nket, const R ABx, const R ABy, const R ABz,")
    em.out("      const R CDx, const R CDy, const R CDz, const BoysTables& tab, R* __restrict__ out) {")
    for e in E:
        for f in F:
            em.out("    R a%s_%s = R(0);" % (cname(e), cname(f)))
    em.out("    for (int ib = 0; ib < nbra; ++ib) {")
    em.out("      const PrimPair bp = bra[ib];")
    em.out("      const R p = R(bp.p), inv2p = R(bp.inv2p), Kab = R(bp.K);")
    em.out("      const R Px = R(bp.P[0]), Py = R(bp.P[1]), Pz = R(bp.P[2]);")
    if la + lb > 0:
        em.out("      const R PAx = R(bp.PA[0]), PAy = R(bp.PA[1]), PAz = R(bp.PA[2]);")
    em.out("      for (int ik = 0; ik < nket; ++ik) {")
    em.out("        const PrimPair kp = ket[ik];")
    em.out("        const R q = R(kp.p), Kcd = R(kp.K);")
    em.out("        const R dx = Px - R(kp.P[0]), dy = Py - R(kp.P[1]), dz = Pz - R(kp.P[2]);")
    em.out("        const R inv_pq = R(1) / (p + q);")
    em.out("        const R rho = p * q * inv_pq;")
    em.out("        const R T = rho * (dx * dx + dy * dy + dz * dz);")
    em.out("        R Fm[%d];" % (L + 1))
    em.out("        BoysEval<R, %d>::run(T, Fm, tab);" % L)
    em.out("        const R pref = Kab * Kcd * sqrt_inv<R>(p + q);")
    em.out("#pragma unroll")
    em.out("        for (int m = 0; m <= %d; ++m) Fm[m] *= pref;" % L)
    if L > 0:
        em.out("        const R rho_p = q * inv_pq, rho_q = p * inv_pq, inv2pq = R(0.5) * inv_pq;")
        em.out("        const R WPx = -rho_p * dx, WPy = -rho_p * dy, WPz = -rho_p * dz;")
        em.out("        const R WQx = rho_q * dx, WQy = rho_q * dy, WQz = rho_q * dz;")
        if lc + ld > 0:
            em.out("        const R inv2q = R(kp.inv2p);")
            em.out("        const R QCx = R(kp.PA[0]), QCy = R(kp.PA[1]), QCz = R(kp.PA[2]);")
    emit_vrr(em, order, nodes, "        ")
    for e in E:
        for f in F:
            em.out("        a%s_%s += %s;" % (cname(e), cname(f), vrr_name(e, f, 0)))
    em.out("      }")
    em.out("    }")
    # HRR: ket first for every e in E, then bra
    kn, kdefs = hrr_ket(la, lb, lc, ld)
    Cc, Dc = comps(lc), comps(ld)
    ket_names = {}
    for e in E:
        for c in Cc:
            for d in Dc:
                ket_names[(e, c, d)] = kn(e, c, d)
    for s in kdefs:
        em.out("    " + s.replace("CDx", "CDx").replace("CDy", "CDy"))
    # bra HRR over (a,b) for each (c,d)
    bseen = {}
    bdefs = []

    def bnode(a, b, c, d):
        key = (a, b, c, d)
        if key in bseen:
            return bseen[key]
        if sum(b) == 0:
            nm = ket_names[(a, c, d)]
            bseen[key] = nm
            return nm
        i = next(k for k in range(3) if b[k] > 0)
        n1 = bnode(add(a, i), sub(b, i), c, d)
        n0 = bnode(a, sub(b, i), c, d)
        nm = "h%s_%s_%s_%s" % (cname(a), cname(b), cname(c), cname(d))
        bdefs.append("const R %s = %s + AB%s * %s;" % (nm, n1, AX[i], n0))
        bseen[key] = nm
        return nm

    Ac, Bc = comps(la), comps(lb)
    cart = {}
    for a in Ac:
        for b in Bc:
            for c in Cc:
                for d in Dc:
                    cart[(a, b, c, d)] = bnode(a, b, c, d)
    for s in bdefs:
        em.out("    " + s)
    # outputs
    em.out("    if constexpr (!PURE) {")
    k = 0
    for a in Ac:
        for b in Bc:
            for c in Cc:
                for d in Dc:
                    em.out("      out[%d] = %s;" % (k, cart[(a, b, c, d)]))
                    k += 1
    em.out("    } else {")
    # pure transform: only l >= 2 dims change
    dims = [la, lb, lc, ld]
    comp_lists = [Ac, Bc, Cc, Dc]
    # represent each output as linear combination over cart tuples
    def basis_funcs(l):
        if l < 2:
            return [[(i, 1.0)] for i in range((l + 1) * (l + 2) // 2)]
        sh = solid_harmonics(l)
        return [[(i, c) for i, c in enumerate(row) if abs(c) > 1e-15] for row in sh]
    bfs = [basis_funcs(l) for l in dims]
    # stage transforms dimension by dimension, emitting intermediates
    cur = {}
    for a_i, a in enumerate(Ac):
        for b_i, b in enumerate(Bc):
            for c_i, c in enumerate(Cc):
                for d_i, d in enumerate(Dc):
                    cur[(a_i, b_i, c_i, d_i)] = cart[(a, b, c, d)]
    sizes = [len(Ac), len(Bc), len(Cc), len(Dc)]
    tcount = 0
    for dim in range(4):
        if dims[dim] < 2:
            continue
        new = {}
        nsz = list(sizes)
        nsz[dim] = len(bfs[dim])
        import itertools
        for idx in itertools.product(*[range(s) for s in nsz]):
            terms = []
            for ci, coef in bfs[dim][idx[dim]]:
                src = list(idx)
                src[dim] = ci
                terms.append("%s * %s" % (lit(coef), cur[tuple(src)]))
            nm = "s%d_%d" % (dim, tcount)
            tcount += 1
            em.out("      const R %s = %s;" % (nm, " + ".join(terms)))
            new[idx] = nm
        cur = new
        sizes = nsz
    k = 0
    import itertools
    for idx in itertools.product(*[range(s) for s in sizes]):
        em.out("      out[%d] = %s;" % (k, cur[idx]))
        k += 1
    em.out("    }")
    em.out("  }")
    em.out("};")
    return "\n".join(em.lines)


def class_list(lmax_bra, lmax_ket):
    out = []
    for la in range(lmax_bra + 1):
        for lb in range(la + 1):
            for lc in range(lmax_ket + 1):
                for ld in range(lc + 1):
                    out.append((la, lb, lc, ld))
    return out


def main(outdir):
    """Write one include file per class: generated/eri_<la><lb><lc><ld>.inc."""
    os.makedirs(outdir, exist_ok=True)
    classes = set(class_list(2, 2))
    # 3-centre DF: (P s | c d) with aux l <= 3; 2-centre (P s | Q s)
    for lp in range(4):
        for lc in range(3):
            for ld in range(lc + 1):
                classes.add((lp, 0, lc, ld))
        for lq in range(4):
            classes.add((lp, 0, lq, 0))
    written = []
    for c in sorted(classes):
        text = "// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.\n#pragma once\n" + gen_class(*c) + "\n"
        path = os.path.join(outdir, "eri_%d%d%d%d.inc" % c)
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as fh:
                fh.write(text)
        written.append(path)
    written += write_tus(outdir)
    written.append(write_sph_tables(outdir))
    return written


PAIR_CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]


def write_sph_tables(outdir, lmax=4):
    """Cartesian -> real solid harmonic coefficients for l = 0..lmax as one
    flat table: SPH_TAB[SPH_OFF[l] + m_index * ncart(l) + cart_index]."""
    vals, offs = [], []
    for l in range(lmax + 1):
        offs.append(len(vals))
        for row in solid_harmonics(l):
            vals.extend(row)
    lines = ["// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.", "#pragma once",
             "namespace m2 {",
             "constexpr int SPH_LMAX = %d;" % lmax,
             "constexpr int SPH_OFF[%d] = {%s};" % (lmax + 1, ", ".join(str(o) for o in offs)),
             "constexpr double SPH_TAB[%d] = {" % len(vals)]
    for i in range(0, len(vals), 4):
        lines.append("    " + ", ".join(repr(float(v)) for v in vals[i:i + 4]) + ",")
    lines += ["};", "}  // namespace m2"]
    text = "\n".join(lines) + "\n"
    path = os.path.join(outdir, "sph_tables.inc")
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as fh:
            fh.write(text)
    return path


def write_tus(outdir):
    """One translation unit per class combination (bra X >= ket Y) so the
    heavy d-shell classes compile in parallel, plus one per X for Schwarz."""
    paths = []

    def put(name, lines):
        text = "\n".join(lines) + "\n"
        path = os.path.join(outdir, name)
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as fh:
                fh.write(text)
        paths.append(path)

    head = ["// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.",
            '#include "../jk_kernel.hpp"', "namespace m2 {"]
    for X, (la, lb) in enumerate(PAIR_CLASSES):
        for Y in range(X + 1):
            lc, ld = PAIR_CLASSES[Y]
            lines = list(head)
            lines.append('#include "eri_%d%d%d%d.inc"' % (la, lb, lc, ld))
            lines.append("void launch_jk_%d_%d(int pure, const JKArgs& a, const BoysTables& tab, const Task* t, int nt, int mode, hipStream_t st) {" % (X, Y))
            lines.append("  if (pure) launch_jk<%d, %d, %d, %d, true>(a, tab, t, nt, mode, st);" % (la, lb, lc, ld))
            lines.append("  else launch_jk<%d, %d, %d, %d, false>(a, tab, t, nt, mode, st);" % (la, lb, lc, ld))
            lines.append("}")
            if X == Y:
                lines.append("void launch_schwarz%d(int pure, const JKArgs& a, const BoysTables& tab, int p0, int np, double* Q, hipStream_t st) {" % X)
                lines.append("  if (pure) launch_schwarz<%d, %d, true>(a, tab, p0, np, Q, st); else launch_schwarz<%d, %d, false>(a, tab, p0, np, Q, st);" % (la, lb, la, lb))
                lines.append("}")
            lines.append("}  // namespace m2")
            put("jk_%d_%d.hip" % (X, Y), lines)
    # dispatch table
    lines = ["// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.",
             '#include "../m2_common.hpp"', '#include "../boys.hpp"', "namespace m2 {"]
    for X in range(len(PAIR_CLASSES)):
        for Y in range(X + 1):
            lines.append("void launch_jk_%d_%d(int, const JKArgs&, const BoysTables&, const Task*, int, int, hipStream_t);" % (X, Y))
    lines.append("typedef void (*jk_combo_launcher)(int, const JKArgs&, const BoysTables&, const Task*, int, int, hipStream_t);")
    lines.append("const jk_combo_launcher JK_COMBO[%d][%d] = {" % (len(PAIR_CLASSES), len(PAIR_CLASSES)))
    for X in range(len(PAIR_CLASSES)):
        row = ["launch_jk_%d_%d" % (X, Y) if Y <= X else "nullptr" for Y in range(len(PAIR_CLASSES))]
        lines.append("  {" + ", ".join(row) + "},")
    lines.append("};")
    lines.append("}  // namespace m2")
    put("jk_dispatch.inc", lines)
    return paths


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)), "generated"))
