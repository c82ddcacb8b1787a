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
    post = []

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
            post.append(key)
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
            post.append(key)
        else:
            nodes[key] = ("base", m)
            post.append(key)

    for e, f in targets:
        need(e, f, 0)
    # depth-first post-order: every value is emitted right before its first use
    return post, nodes


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


def node_deps(kind):
    """the VRR nodes one node reads"""
    if kind[0] == "base":
        return []
    _, terms, extra = kind
    deps = [k for _, k in terms]
    for x in extra:
        deps += list(x[2:])
    return deps


def emit_fused_acc(em, acc_name, kind, indent, L):
    """A target no other VRR node reads, added term by term into its
    accumulator (acc += c v, one FMA each) instead of formed first and added
    after.  With L = 1 the WP / WQ factors are not formed at all: WP_x v =
    -d_x (rho_p v), the product rho_p v shared by the three components."""
    _, terms, extra = kind
    for c, k in terms:
        v = vrr_name(*k)
        if L == 1 and c[:2] == "WP":
            em.out(indent + "%s -= d%s * (rho_p * %s);" % (acc_name, c[2], v))
        elif L == 1 and c[:2] == "WQ":
            em.out(indent + "%s += d%s * (rho_q * %s);" % (acc_name, c[2], v))
        else:
            em.out(indent + "%s += %s * %s;" % (acc_name, c, v))
    for x in extra:
        if x[0] == "ff":
            em.out(indent + "%s += %s * inv2q * (%s - rho_q * %s);" % (acc_name, lit(x[1]), vrr_name(*x[2]), vrr_name(*x[3])))
        elif x[0] == "ee":
            em.out(indent + "%s += %s * inv2p * (%s - rho_p * %s);" % (acc_name, lit(x[1]), vrr_name(*x[2]), vrr_name(*x[3])))
        elif x[0] == "ef":
            em.out(indent + "%s += %s * inv2pq * %s;" % (acc_name, lit(x[1]), vrr_name(*x[2])))


def prim_body(la, lb, lc, ld, targets, interleave, pref, acc, fuse=None):
    """Lines of one bra primitive of the bra loop (record `bp`), split into
    (setup up to T, finish from the scaled F_m on); the Boys call sits in
    between.  `acc(e, f)` names the accumulator of target (e, f).  With
    `fuse` (default: with interleave) a target that no other node reads is
    added into its accumulator term by term."""
    L = la + lb + lc + ld
    if fuse is None:
        fuse = interleave
    order, nodes = build_vrr(targets, L)
    read = set()
    for key in order:
        read.update(node_deps(nodes[key]))
    ind = "        "
    setup = [ind + "const R p = R(bp.p), Kab = R(bp.K);"]
    if la + lb > 0:
        setup.append(ind + "const R inv2p = R(bp.inv2p);")
        setup.append(ind + "const R PAx = R(bp.PA[0]), PAy = R(bp.PA[1]), PAz = R(bp.PA[2]);")
    setup += [ind + "const R dx = R(bp.P[0]) - Qx, dy = R(bp.P[1]) - Qy, dz = R(bp.P[2]) - Qz;",
              ind + "const R rs_pq = sqrt_inv<R>(p + q);",
              ind + "const R inv_pq = rs_pq * rs_pq;",
              ind + "const R rho = p * q * inv_pq;",
              ind + "const R T = rho * (dx * dx + dy * dy + dz * dz);"]
    fin = Emitter()
    fin.out(ind + "const R pref = %s;" % pref)
    fin.out("#pragma unroll")
    fin.out(ind + "for (int m = 0; m <= %d; ++m) Fm[m] *= pref;" % L)
    if L > 0:
        fin.out(ind + "const R rho_p = q * inv_pq, rho_q = p * inv_pq, inv2pq = R(0.5) * inv_pq;")
        if L > 1 or not fuse:
            fin.out(ind + "const R WPx = -rho_p * dx, WPy = -rho_p * dy, WPz = -rho_p * dz;")
            fin.out(ind + "const R WQx = rho_q * dx, WQy = rho_q * dy, WQz = rho_q * dz;")
    tset = set(targets)

    def fused(key):
        e, f, m = key
        return fuse and m == 0 and (e, f) in tset and key not in read and nodes[key][0] != "base"

    if interleave or fuse:
        # each target is accumulated right after its VRR node (shorter live ranges)
        for key in order:
            e, f, m = key
            if fused(key):
                emit_fused_acc(fin, acc(e, f), nodes[key], ind, L)
                continue
            emit_vrr(fin, [key], nodes, ind)
            if m == 0 and (e, f) in tset:
                fin.out(ind + "%s += %s;" % (acc(e, f), vrr_name(e, f, 0)))
    else:
        emit_vrr(fin, order, nodes, ind)
        for e, f in targets:
            fin.out(ind + "%s += %s;" % (acc(e, f), vrr_name(e, f, 0)))
    return setup, fin.lines


def emit_bra_loop(em, L, setup, fin):
    """The bra primitive loop of one ket primitive; the wavefront stops at
    the first bra primitive below every lane's cut (bra bounds are sorted).
    (Two bra primitives per step, interleaved, measured slower on (H2O)20
    cc-pVDZ for every class tried: the extra registers cost more occupancy
    than the second chain hid latency.)"""
    em.out("      for (int ib = 0; ib < nbra; ++ib) {")
    em.out("        if (SCR && !__any(float(bq[ib]) * qk >= cut)) break;")
    em.out("        const auto& bp = bra[ib];")
    em.lines += setup
    em.out("        R Fm[%d];" % (L + 1))
    em.out("        BoysEval<R, %d>::run(T, Fm, tab);" % L)
    em.lines += fin
    em.out("      }")


def gen_class(la, lb, lc, ld, deriv=False, interleave=True, fuse=None, prefetch=False):
    L = la + lb + lc + ld
    em = Emitter()
    E = [c for Le in range(la, la + lb + 1) for c in comps(Le)]
    F = [c for Lf in range(lc, lc + ld + 1) for c in comps(Lf)]
    targets = [(e, f) for e in E for f in F]
    na, nb, nc, nd = [(l + 1) * (l + 2) // 2 for l in (la, lb, lc, ld)]
    em.out("template <> struct EriClass<%d, %d, %d, %d> {" % (la, lb, lc, ld))
    em.out("  static constexpr int L = %d;" % L)
    em.out("  static constexpr int NCART = %d;" % (na * nb * nc * nd))
    em.out("  static constexpr int NACC = %d;   // VRR accumulators per lane" % (len(E) * len(F)))
    em.out("  template <typename R, bool PURE, bool SCR = false, typename BPtr, typename KPtr,")
    em.out("            typename QBPtr = const float*>")
    em.out("  static __device__ __forceinline__ void run(BPtr bra, int nbra,")
    em.out("      KPtr ket, int nket, const R ABx, const R ABy, const R ABz,")
    em.out("      const R CDx, const R CDy, const R CDz, const BoysTables& tab, R* __restrict__ out,")
    em.out("      QBPtr bq = nullptr, const float* __restrict__ kq = nullptr, float cut = 0.f) {")
    for e in E:
        for f in F:
            em.out("    R a%s_%s = R(0);" % (cname(e), cname(f)))
    # ket primitive pairs (different per lane) outside, bra primitive pairs
    # (the same for the whole wavefront: scalar loads) inside.  Primitive
    # screening (SCR): bq / kq bound the primitive pairs, both sorted
    # descending, and a primitive quartet may be dropped when bq kq < cut
    # (the lane's own cut).  A lane stops at the first ket primitive whose
    # product with the largest bra bound is below its cut; the bra loop of
    # the wavefront stops at the first bra primitive below every lane's cut.
    # Quartets between those limits are computed, which only adds accuracy.
    em.out("    const float bq0 = SCR ? float(bq[0]) : 0.f;")
    if prefetch:
        # the next ket primitive is loaded while the current one is computed
        em.out("    auto kp_next = load_ket(ket, 0, nket);")
        em.out("    for (int ik = 0; ik < nket; ++ik) {")
        em.out("      const auto kp = kp_next;")
        em.out("      kp_next = load_ket(ket, ik + 1, nket);")
        em.out("      const float qk = SCR ? ket_q(kp, kq, ik) : 0.f;")
        em.out("      if (SCR && qk * bq0 < cut) break;")
    else:
        em.out("    for (int ik = 0; ik < nket; ++ik) {")
        em.out("      const float qk = SCR ? ket_bound(ket, kq, ik) : 0.f;")
        em.out("      if (SCR && qk * bq0 < cut) break;")
        em.out("      const auto kp = ket[ik];")
    em.out("      const R q = R(kp.p), Kcd = R(kp.K);")
    em.out("      const R Qx = R(kp.P[0]), Qy = R(kp.P[1]), Qz = R(kp.P[2]);")
    if lc + ld > 0:
        em.out("      const R inv2q = R(kp.inv2p);")
        em.out("      const R QCx = R(kp.PA[0]), QCy = R(kp.PA[1]), QCz = R(kp.PA[2]);")
    setup, fin = prim_body(la, lb, lc, ld, targets, interleave, "Kab * Kcd * rs_pq",
                           lambda e, f: "a%s_%s" % (cname(e), cname(f)), fuse=fuse)
    emit_bra_loop(em, L, setup, fin)
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


# Bra classes whose (ab|ss) quartets also get run_fam: the ket is a family
# pair of s shells (shells of one atom with identical exponents, e.g. the two
# contracted s shells of O in cc-pVDZ), whose primitive pairs are shared by up
# to FAM_MAXMEM contracted member pairs.
FAM_BRA = [(0, 0), (1, 0), (1, 1), (2, 0)]
FAM_MAXMEM = 4
# family classes that load their next family primitive ahead: (ss|ss) (4
# waves per SIMD with it: 4.8 -> 4.6 ms on (H2O)20 cc-pVDZ) and (pp|ss)
# (226 VGPRs either way, 2 waves per SIMD)
FAM_PREFETCH = {(0, 0), (1, 1)}
# register classes that load their next ket primitive ahead: those whose
# kernels stay within the same waves per SIMD with the extra record (2 waves
# at <= 256 VGPRs or 1 wave; (ps|ss), a family-bra class, is set in
# gen_class_fam)
PREFETCH = {(1, 1, 1, 0), (1, 1, 1, 1), (2, 0, 1, 1), (2, 0, 2, 0), (2, 1, 0, 0), (2, 1, 1, 0), (2, 1, 2, 0),
            (2, 2, 1, 0)}


def gen_class_fam(la, lb):
    """run_fam for (la lb|ss): per ket primitive pair the bra-contracted
    [e0|s] are formed once without any ket factor, then folded into the
    accumulators of every member with that member's ket factor
    km[FAM_MAXMEM * ik + m] (K_cd of the member's own primitive pair); the
    HRR / solid-harmonic post-processing then runs once per member, writing
    out + m * NCART."""
    base = gen_class(la, lb, 0, 0, interleave=False)
    lines = base.splitlines()
    E = [c for Le in range(la, la + lb + 1) for c in comps(Le)]
    accs = ["a%s_%s" % (cname(e), cname((0, 0, 0))) for e in E]
    i_loop = next(i for i, t in enumerate(lines) if t.strip().startswith("for (int ik = 0; ik < nket; ++ik)"))
    i_kp = next(i for i in range(i_loop, len(lines)) if lines[i].strip().startswith("const auto kp = ket[ik];"))
    i_bra = next(i for i in range(i_loop, len(lines)) if lines[i].strip().startswith("for (int ib = 0"))
    i_acc = next(i for i in range(i_bra, len(lines)) if lines[i].strip().startswith(accs[0] + " +="))
    i_end = i_acc + len(accs)
    assert lines[i_end].strip() == "}" and lines[i_end + 1].strip() == "}"
    post = lines[i_end + 2:]
    # drop the closing of run() and of the struct from the post block
    assert post[-1] == "};" and post[-2].strip() == "}"
    post = post[:-2]
    screen = lines[i_loop + 1:i_kp]
    ket_head = [t.replace("const R q = R(kp.p), Kcd = R(kp.K);", "const R q = R(kp.p);")
                for t in lines[i_kp:i_bra]]
    targets = [(e, (0, 0, 0)) for e in E]
    setup, fin = prim_body(la, lb, 0, 0, targets, False, "Kab * rs_pq",
                           lambda e, f: "t%s_%s" % (cname(e), cname(f)), fuse=True)
    bra_loop = Emitter()
    emit_bra_loop(bra_loop, la + lb, setup, fin)
    em = []
    em.append("  static constexpr int NACC_FAM = %d;   // accumulators of run_fam" % (len(accs) * (FAM_MAXMEM + 1)))
    em.append("  template <typename R, bool PURE, bool SCR, typename BPtr, typename KPtr, typename CPtr, typename QBPtr>")
    em.append("  static __device__ __forceinline__ void run_fam(BPtr bra, int nbra,")
    em.append("      KPtr ket, CPtr km, int nket, const R ABx, const R ABy, const R ABz,")
    em.append("      const BoysTables& tab, R* __restrict__ out, QBPtr bq, const float* __restrict__ kq, float cut) {")
    for m in range(FAM_MAXMEM):
        for a in accs:
            em.append("    R %s_m%d = R(0);" % (a, m))
    em.append("    const float bq0 = SCR ? float(bq[0]) : 0.f;")
    if (la, lb) in FAM_PREFETCH:
        # the next family primitive is loaded while the current one is computed
        em.append("    auto kp_next = load_ket(ket, 0, nket);")
        em.append("    for (int ik = 0; ik < nket; ++ik) {")
        em.append("      const auto kp = kp_next;")
        em.append("      kp_next = load_ket(ket, ik + 1, nket);")
        em.append("      const float qk = SCR ? ket_q(kp, kq, ik) : 0.f;")
        em.append("      if (SCR && qk * bq0 < cut) break;")
        em.extend(t for t in ket_head if t.strip() != "const auto kp = ket[ik];")
    else:
        em.append("    for (int ik = 0; ik < nket; ++ik) {")
        em.extend(screen)
        em.extend(ket_head)
    for a in accs:
        em.append("      R t%s = R(0);" % a[1:])
    em.extend(bra_loop.lines)
    for m in range(FAM_MAXMEM):
        if (la, lb) in FAM_PREFETCH:
            em.append("      const R c%d = R(fam_factor_of(kp, km, ik, %d));" % (m, m))
        else:
            em.append("      const R c%d = R(fam_factor(km, ket, ik, %d));" % (m, m))
    for m in range(FAM_MAXMEM):
        for a in accs:
            em.append("      %s_m%d += c%d * t%s;" % (a, m, m, a[1:]))
    em.append("    }")
    for m in range(FAM_MAXMEM):
        em.append("    {")
        for a in accs:
            em.append("      const R %s = %s_m%d;" % (a, a, m))
        em.append("      R* __restrict__ outm = out + %d * NCART;" % m)
        for t in post:
            em.append("  " + t.replace("out[", "outm["))
        em.append("    }")
    em.append("  }")
    # (ps|ss) loads its next ket primitive ahead (measured: (ps|ss) 8.3 -> 7.9 ms
    # on (H2O)20 cc-pVDZ; (ps|ps) and (ds|ss) lose occupancy to the registers)
    run = gen_class(la, lb, 0, 0, interleave=False, fuse=True, prefetch=(la, lb) == (1, 0)).splitlines()
    assert run[-1] == "};"
    return "\n".join(run[:-1] + em + ["};"])


# Classes whose J/K quartets run as split register code (k_jk_split): their
# VRR accumulators do not fit one lane's registers, so the quartet is
# evaluated in passes, one per Cartesian component c of shell c ("c") or per
# pair of Cartesian components (c, d) ("cd").  A pass accumulates only the
# targets [e0|f0] its ket HRR needs -- f = c + d' for the sub-monomials d'
# of its d -- and returns (a b | c d) for its d (a, b spherical for pure
# bases); the lower VRR levels are recomputed in every pass.
SPLIT = {(2, 1, 2, 1): "c", (2, 2, 1, 1): "cd", (2, 2, 2, 0): "c", (2, 2, 2, 1): "cd"}


def submonos(d):
    return [(x, y, z) for x in range(d[0] + 1) for y in range(d[1] + 1) for z in range(d[2] + 1)]


def gen_class_split(la, lb, lc, ld):
    L = la + lb + lc + ld
    E = [c for Le in range(la, la + lb + 1) for c in comps(Le)]
    Ac, Bc, Cc, Dc = comps(la), comps(lb), comps(lc), comps(ld)
    mode = SPLIT[(la, lb, lc, ld)]
    passes = [(ci, 0, len(Dc)) for ci in range(len(Cc))] if mode == "c" else \
        [(ci, di, 1) for ci in range(len(Cc)) for di in range(len(Dc))]

    def harm(l):
        if l < 2:
            return [[(i, 1.0)] for i in range((l + 1) * (l + 2) // 2)]
        return [[(i, x) for i, x in enumerate(row) if abs(x) > 1e-15] for row in solid_harmonics(l)]

    ha, hb, hc = harm(la), harm(lb), harm(lc)
    maxnd = max(p[2] for p in passes)
    em = Emitter()
    em.out("template <> struct EriSplit<%d, %d, %d, %d> {" % (la, lb, lc, ld))
    em.out("  static constexpr int L = %d;" % L)
    em.out("  static constexpr int NPASS = %d;" % len(passes))
    em.out("  // pass p: Cartesian component PC[p] of shell c, components PD0[p] .. PD0[p] + PND[p] - 1 of d")
    em.out("  static constexpr int PC[NPASS] = {%s};" % ", ".join(str(p[0]) for p in passes))
    em.out("  static constexpr int PD0[NPASS] = {%s};" % ", ".join(str(p[1]) for p in passes))
    em.out("  static constexpr int PND[NPASS] = {%s};" % ", ".join(str(p[2]) for p in passes))
    em.out("  static constexpr int NOUT = %d;   // outputs of a pass at most, (a, b, d) order" %
           (len(Ac) * len(Bc) * maxnd))
    args = ("BPtr bra, int nbra, KPtr ket, int nket, const R ABx, const R ABy, const R ABz, "
            "const R CDx, const R CDy, const R CDz, const BoysTables& tab, R* __restrict__ out, "
            "QBPtr bq, const float* __restrict__ kq, float cut")
    for pi, (ci, d0, nd) in enumerate(passes):
        c = Cc[ci]
        ds = Dc[d0:d0 + nd]
        Fs = sorted({tuple(c[i] + dd[i] for i in range(3)) for d in ds for dd in submonos(d)},
                    key=lambda f: (sum(f), f))
        targets = [(e, f) for e in E for f in Fs]
        em.out("  // pass %d: c = %s, d = %s, %d VRR targets" % (pi, cname(c), " ".join(cname(d) for d in ds),
                                                               len(targets)))
        em.out("  template <typename R, bool PURE, bool SCR, typename BPtr, typename KPtr, typename QBPtr>")
        em.out("  static __device__ __forceinline__ void pass%d(%s) {" % (pi, args))
        for e, f in targets:
            em.out("    R a%s_%s = R(0);" % (cname(e), cname(f)))
        em.out("    const float bq0 = SCR ? float(bq[0]) : 0.f;")
        em.out("    for (int ik = 0; ik < nket; ++ik) {")
        em.out("      const float qk = SCR ? kq[ik] : 0.f;")
        em.out("      if (SCR && qk * bq0 < cut) break;")
        em.out("      const auto kp = ket[ik];")
        em.out("      const R q = R(kp.p), Kcd = R(kp.K);")
        em.out("      const R Qx = R(kp.P[0]), Qy = R(kp.P[1]), Qz = R(kp.P[2]);")
        em.out("      const R inv2q = R(kp.inv2p);")
        em.out("      const R QCx = R(kp.PA[0]), QCy = R(kp.PA[1]), QCz = R(kp.PA[2]);")
        setup, fin = prim_body(la, lb, lc, ld, targets, True, "Kab * Kcd * rs_pq",
                               lambda e, f: "a%s_%s" % (cname(e), cname(f)))
        emit_bra_loop(em, L, setup, fin)
        em.out("    }")
        node, kdefs = hrr_ket(la, lb, lc, ld)
        ket_names = {(e, d): node(e, c, d) for e in E for d in ds}
        for t in kdefs:
            em.out("    " + t)
        bseen, bdefs = {}, []

        def bnode(a, b, d):
            key = (a, b, d)
            if key in bseen:
                return bseen[key]
            if sum(b) == 0:
                bseen[key] = ket_names[(a, d)]
                return bseen[key]
            i = next(k for k in range(3) if b[k] > 0)
            n1 = bnode(add(a, i), sub(b, i), d)
            n0 = bnode(a, sub(b, i), d)
            nm = "h%s_%s_%s" % (cname(a), cname(b), cname(d))
            bdefs.append("const R %s = %s + AB%s * %s;" % (nm, n1, AX[i], n0))
            bseen[key] = nm
            return nm

        cart = {(ia, ib, idd): bnode(a, b, d) for ia, a in enumerate(Ac) for ib, b in enumerate(Bc)
                for idd, d in enumerate(ds)}
        for t in bdefs:
            em.out("    " + t)
        em.out("    if constexpr (PURE) {")
        for ia, ra in enumerate(ha):
            for ib, rb in enumerate(hb):
                for idd in range(nd):
                    terms = []
                    for ca, xa in ra:
                        for cb, xb in rb:
                            x = xa * xb
                            src = cart[(ca, cb, idd)]
                            terms.append(src if x == 1.0 else "%s * %s" % (lit(x), src))
                    em.out("      out[%d] = %s;" % ((ia * len(hb) + ib) * nd + idd, " + ".join(terms)))
        em.out("    } else {")
        for (ia, ib, idd), nm in sorted(cart.items()):
            em.out("      out[%d] = %s;" % ((ia * len(Bc) + ib) * nd + idd, nm))
        em.out("    }")
        em.out("  }")
    # spherical components of shell c fed by each Cartesian component
    cm = [[(m, x) for m, row in enumerate(hc) for i, x in row if i == ci] for ci in range(len(Cc))]
    w = max(len(t) for t in cm)
    em.out("  // spherical components m of shell c fed by its Cartesian component i, and the factors")
    em.out("  static constexpr int CN[%d] = {%s};" % (len(Cc), ", ".join(str(len(t)) for t in cm)))
    em.out("  static constexpr int CM[%d][%d] = {%s};" % (len(Cc), w, ", ".join(
        "{%s}" % ", ".join(str(t[k][0]) if k < len(t) else "0" for k in range(w)) for t in cm)))
    em.out("  static constexpr double CC[%d][%d] = {%s};" % (len(Cc), w, ", ".join(
        "{%s}" % ", ".join(repr(float(t[k][1])) if k < len(t) else "0.0" for k in range(w)) for t in cm)))
    em.out("  template <int P, typename R, bool PURE, bool SCR, typename BPtr, typename KPtr, typename QBPtr>")
    em.out("  static __device__ __forceinline__ void run(%s) {" % args)
    call = "(bra, nbra, ket, nket, ABx, ABy, ABz, CDx, CDy, CDz, tab, out, bq, kq, cut)"
    for pi in range(len(passes)):
        em.out("    %sif constexpr (P == %d) pass%d<R, PURE, SCR>%s;" % ("" if pi == 0 else "else ", pi, pi, call))
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
        body = gen_class_fam(c[0], c[1]) if (c[2], c[3]) == (0, 0) and (c[0], c[1]) in FAM_BRA else \
            gen_class(*c, prefetch=c in PREFETCH)
        text = "// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.\n#pragma once\n" + body + "\n"
        path = os.path.join(outdir, "eri_%d%d%d%d.inc" % c)
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as fh:
                fh.write(text)
        written.append(path)
    for c in SPLIT:
        text = "// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.\n#pragma once\n" + \
            gen_class_split(*c) + "\n"
        path = os.path.join(outdir, "eri_split_%d%d%d%d.inc" % c)
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as fh:
                fh.write(text)
        written.append(path)
    written += write_tus(outdir)
    written += write_df_tus(outdir)
    written.append(write_sph_tables(outdir))
    written.append(write_programs(outdir))
    written.append(write_deriv_programs(outdir))
    written.append(write_deriv_register(outdir))
    # The op-count model (molecular2_amd/flop_model.json) is committed and
    # fixed; the build does not rewrite it.  `gen_eri.py --flop-model PATH`
    # regenerates it from the textbook recurrence DAG.
    return written


PAIR_CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]


def _ops(expr):
    """arithmetic operations of one emitted right-hand side (binary + - *)."""
    n = 0
    depth = 0
    prev = ""
    for ch in expr:
        if ch in "*+" or (ch == "-" and prev not in "(e"):
            n += 1
        if not ch.isspace():
            prev = ch
    return n


def flop_model(la, lb, lc, ld):
    """Algorithmic operation counts of the textbook OS-VRR + HGP-HRR scheme
    for one class, read off the generated recurrence DAG:
      f_prim: per primitive quartet (setup, Boys, VRR, accumulation),
      f_cont: per contracted quartet (HRR, solid-harmonic transform),
      n_eri:  spherical integrals per quartet."""
    L = la + lb + lc + ld
    # the textbook DAG: every VRR node formed, then added to its accumulator
    # (the emitted kernels add terminal nodes into their accumulators term
    # by term, which saves instructions, not operations of the scheme)
    text = gen_class(la, lb, lc, ld, fuse=False)
    lines = text.splitlines()
    i0 = next(i for i, t in enumerate(lines) if "for (int ib = 0" in t)
    i1 = next(i for i in range(i0, len(lines)) if lines[i].strip() == "}")
    prim = 0
    for t in lines[i0 + 1:i1]:
        t = t.strip()
        if t.startswith("const R v") or t.startswith("a") and "+=" in t:
            prim += _ops(t.split("=", 1)[1]) + (1 if "+=" in t else 0)
    # setup (distance, 1/(p+q), rho, T, prefactor, W-P, W-Q, scaling of F_m)
    # and Boys (8-term Horner + downward recursion)
    prim += 30 + (L + 1) + 21 + 3 * L
    cont_pure = cont_cart = 0
    in_pure = False
    for t in lines[i1:]:
        st = t.strip()
        if st.startswith("} else {"):
            in_pure = True
        if st.startswith("const R k") or st.startswith("const R h"):
            c = _ops(st.split("=", 1)[1])
            cont_pure += c
            cont_cart += c
        elif st.startswith("const R s") and in_pure:
            cont_pure += _ops(st.split("=", 1)[1])
    n_pure = (2 * la + 1) * (2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1)
    n_cart = ((la + 1) * (la + 2) * (lb + 1) * (lb + 2) * (lc + 1) * (lc + 2) * (ld + 1) * (ld + 2)) // 16
    return {"f_prim": prim, "f_cont_pure": cont_pure, "f_cont_cart": cont_cart,
            "n_eri_pure": n_pure, "n_eri_cart": n_cart}


def write_flop_model(path):
    import json
    model = {}
    for X, (la, lb) in enumerate(PAIR_CLASSES):
        for Y in range(X + 1):
            lc, ld = PAIR_CLASSES[Y]
            model["%d,%d" % (X, Y)] = dict(flop_model(la, lb, lc, ld), cls="(%d%d|%d%d)" % (la, lb, lc, ld))
    text = json.dumps({"digest_flops_per_eri": 12, "classes": model}, indent=1, sort_keys=True) + "\n"
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as fh:
            fh.write(text)
    return path


def write_df_tus(outdir):
    """3-centre (P s|c d) launchers per aux l (one TU per aux l) and 2-centre
    (P s|Q s) launchers (one TU)."""
    paths = []

    def put(name, lines):
        text = "\n".join(lines) + "\n"
        path = os.path.join(outdir, name)
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as fh:
                fh.write(text)
        paths.append(path)

    head = ["// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.",
            '#include "../df_kernel.hpp"', "namespace m2 {"]
    for lp in range(4):
        lines = list(head)
        for Y, (lc, ld) in enumerate(PAIR_CLASSES):
            lines.append('#include "eri_%d0%d%d.inc"' % (lp, lc, ld))
        lines.append("void launch_df3_%d(int Y, int pure, const DF3Args& a, const BoysTables& tab, const Task* t, int nt, hipStream_t st) {" % lp)
        lines.append("  switch (Y) {")
        for Y, (lc, ld) in enumerate(PAIR_CLASSES):
            lines.append("    case %d: if (pure) launch_df3<%d, %d, %d, true>(a, tab, t, nt, st); else launch_df3<%d, %d, %d, false>(a, tab, t, nt, st); break;"
                         % (Y, lp, lc, ld, lp, lc, ld))
        lines.append("    default: break;")
        lines.append("  }")
        lines.append("}")
        lines.append("}  // namespace m2")
        put("df3_%d.hip" % lp, lines)
    lines = list(head)
    for lp in range(4):
        for lq in range(lp + 1):
            lines.append('#include "eri_%d0%d0.inc"' % (lp, lq))
    lines.append("void launch_df2(int lp, int lq, int pure, const DF2Args& a, const BoysTables& tab, hipStream_t st) {")
    for lp in range(4):
        for lq in range(lp + 1):
            lines.append("  if (lp == %d && lq == %d) { if (pure) launch_df2_t<%d, %d, true>(a, tab, st); else launch_df2_t<%d, %d, false>(a, tab, st); return; }"
                         % (lp, lq, lp, lq, lp, lq))
    lines.append("}")
    lines.append("}  // namespace m2")
    put("df2.hip", lines)
    return paths


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
            if (la, lb, lc, ld) in SPLIT:
                lines = list(head)
                lines.append('#include "eri_split_%d%d%d%d.inc"' % (la, lb, lc, ld))
                lines.append("void launch_jk_%d_%d(int pure, const JKArgs& a, const BoysTables& tab, int nblocks, int mode, hipStream_t st) {" % (X, Y))
                lines.append("  if (pure) launch_jk_split<%d, %d, %d, %d, true>(a, tab, nblocks, mode, st);" % (la, lb, lc, ld))
                lines.append("  else launch_jk_split<%d, %d, %d, %d, false>(a, tab, nblocks, mode, st);" % (la, lb, lc, ld))
                lines.append("}")
                lines.append("}  // namespace m2")
                put("jk_%d_%d.hip" % (X, Y), lines)
                continue
            if (la, lb, lc, ld) in HEAVY:
                continue
            lines = list(head)
            lines.append('#include "eri_%d%d%d%d.inc"' % (la, lb, lc, ld))
            lines.append("void launch_jk_%d_%d(int pure, const JKArgs& a, const BoysTables& tab, int nblocks, int mode, hipStream_t st) {" % (X, Y))
            if (lc, ld) == (0, 0) and (la, lb) in FAM_BRA:
                lines.append("  if (a.fam_kets) {")
                lines.append("    if (pure) launch_jk_fam<%d, %d, true>(a, tab, nblocks, mode, st);" % (la, lb))
                lines.append("    else launch_jk_fam<%d, %d, false>(a, tab, nblocks, mode, st);" % (la, lb))
                lines.append("    return;")
                lines.append("  }")
            lines.append("  if (pure) launch_jk<%d, %d, %d, %d, true>(a, tab, nblocks, mode, st);" % (la, lb, lc, ld))
            lines.append("  else launch_jk<%d, %d, %d, %d, false>(a, tab, nblocks, mode, st);" % (la, lb, lc, ld))
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
    def heavy(X, Y):
        return PAIR_CLASSES[X] + PAIR_CLASSES[Y] in HEAVY

    def table_jk(X, Y):   # J/K through the table-driven heavy kernel
        return heavy(X, Y) and PAIR_CLASSES[X] + PAIR_CLASSES[Y] not in SPLIT

    for X in range(len(PAIR_CLASSES)):
        for Y in range(X + 1):
            if not table_jk(X, Y):
                lines.append("void launch_jk_%d_%d(int, const JKArgs&, const BoysTables&, int, int, hipStream_t);" % (X, Y))
    lines.append("typedef void (*jk_combo_launcher)(int, const JKArgs&, const BoysTables&, int, int, hipStream_t);")
    lines.append("const jk_combo_launcher JK_COMBO[%d][%d] = {" % (len(PAIR_CLASSES), len(PAIR_CLASSES)))
    for X in range(len(PAIR_CLASSES)):
        row = ["launch_jk_%d_%d" % (X, Y) if Y <= X and not table_jk(X, Y) else "nullptr"
               for Y in range(len(PAIR_CLASSES))]
        lines.append("  {" + ", ".join(row) + "},")
    lines.append("};")
    lines.append("typedef void (*sw_launcher)(int, const JKArgs&, const BoysTables&, int, int, double*, hipStream_t);")
    for X in range(len(PAIR_CLASSES)):
        if not heavy(X, X):
            lines.append("void launch_schwarz%d(int, const JKArgs&, const BoysTables&, int, int, double*, hipStream_t);" % X)
    lines.append("const sw_launcher SW_LAUNCH[%d] = {%s};" % (len(PAIR_CLASSES), ", ".join(
        "nullptr" if heavy(X, X) else "launch_schwarz%d" % X for X in range(len(PAIR_CLASSES)))))
    lines.append("}  // namespace m2")
    put("jk_dispatch.inc", lines)
    return paths




# ---------------------------------------------------------------------------
# Table-driven programs for heavy classes (one wavefront per quartet, LDS)
# ---------------------------------------------------------------------------
# Scalar ids available to program terms (filled per primitive quartet in LDS)
SID = {"one": 0, "PAx": 1, "PAy": 2, "PAz": 3, "WPx": 4, "WPy": 5, "WPz": 6,
       "QCx": 7, "QCy": 8, "QCz": 9, "WQx": 10, "WQy": 11, "WQz": 12,
       "inv2p": 13, "inv2q": 14, "inv2pq": 15, "inv2p_rho_p": 16, "inv2q_rho_q": 17,
       "ABx": 18, "ABy": 19, "ABz": 20, "CDx": 21, "CDy": 22, "CDz": 23}
NSCALAR = 24
MAXTERMS = 6


def gen_program(la, lb, lc, ld, pure):
    """Recurrence DAG of one class as level-ordered instructions
    dst = sum_t coef_t * S[sidx_t] * X[src_t].  Returns a dict:
      nslots, base (slots base..base+L hold [00|00]^(m)), acc (list of
      (acc_slot, vrr_slot)), vrr_levels / post_levels (lists of instruction
      lists), out (output slots in (a,b,c,d) order), sizes."""
    L = la + lb + lc + ld
    E = [c for Le in range(la, la + lb + 1) for c in comps(Le)]
    F = [c for Lf in range(lc, lc + ld + 1) for c in comps(Lf)]
    order, nodes = build_vrr([(e, f) for e in E for f in F], L)
    nacc = len(E) * len(F)
    acc_slot = {}
    k = 0
    for e in E:
        for f in F:
            acc_slot[(e, f)] = k
            k += 1
    nxt = [nacc]
    slot = {}

    def alloc():
        s = nxt[0]
        nxt[0] += 1
        return s

    base = nxt[0]
    for m in range(L + 1):
        slot[((0, 0, 0), (0, 0, 0), m)] = alloc()
    vrr_levels = [[] for _ in range(L + 1)]
    for key in order:
        e, f, m = key
        kind = nodes[key]
        if kind[0] == "base":
            continue
        _, terms, extra = kind
        t = []
        for c, src in terms:
            t.append((slot[src], SID[c], 1.0))
        for x in extra:
            if x[0] == "ff":
                t.append((slot[x[2]], SID["inv2q"], float(x[1])))
                t.append((slot[x[3]], SID["inv2q_rho_q"], -float(x[1])))
            elif x[0] == "ee":
                t.append((slot[x[2]], SID["inv2p"], float(x[1])))
                t.append((slot[x[3]], SID["inv2p_rho_p"], -float(x[1])))
            elif x[0] == "ef":
                t.append((slot[x[2]], SID["inv2pq"], float(x[1])))
        assert len(t) <= MAXTERMS
        d = alloc()
        slot[key] = d
        vrr_levels[sum(e) + sum(f)].append((d, t))
    vrr_levels = [lv for lv in vrr_levels if lv]
    vrr_end = nxt[0]
    acc = [(acc_slot[(e, f)], slot[(e, f, 0)]) for e in E for f in F]
    # post phase built on symbolic values, then slot-allocated with reuse:
    # ("acc", e, f) are the accumulators, other keys are HRR / transform nodes
    Cc, Dc, Ac, Bc = comps(lc), comps(ld), comps(la), comps(lb)
    sym_levels = []

    def kget(e, c, d):
        return ("acc", e, c) if sum(d) == 0 else ("k", e, c, d)

    for lev in range(1, ld + 1):
        ins = []
        for e in E:
            for Lc in range(lc, lc + ld - lev + 1):
                for c in comps(Lc):
                    for d in comps(lev):
                        i = next(q for q in range(3) if d[q] > 0)
                        ins.append((("k", e, c, d), [(kget(e, add(c, i), sub(d, i)), SID["one"], 1.0),
                                                     (kget(e, c, sub(d, i)), SID["CD" + AX[i]], 1.0)]))
        sym_levels.append(ins)

    def bget(a, b, c, d):
        return kget(a, c, d) if sum(b) == 0 else ("b", a, b, c, d)

    for lev in range(1, lb + 1):
        ins = []
        for La in range(la, la + lb - lev + 1):
            for a in comps(La):
                for b in comps(lev):
                    for c in Cc:
                        for d in Dc:
                            i = next(q for q in range(3) if b[q] > 0)
                            ins.append((("b", a, b, c, d), [(bget(add(a, i), sub(b, i), c, d), SID["one"], 1.0),
                                                            (bget(a, sub(b, i), c, d), SID["AB" + AX[i]], 1.0)]))
        sym_levels.append(ins)
    import itertools
    cur = {}
    for ia, a in enumerate(Ac):
        for ib, b in enumerate(Bc):
            for ic, c in enumerate(Cc):
                for idd, d in enumerate(Dc):
                    cur[(ia, ib, ic, idd)] = bget(a, b, c, d)
    sizes = [len(Ac), len(Bc), len(Cc), len(Dc)]
    dims = [la, lb, lc, ld]
    if pure:
        for dim in range(4):
            if dims[dim] < 2:
                continue
            rows = solid_harmonics(dims[dim])
            nsz = list(sizes)
            nsz[dim] = len(rows)
            new = {}
            ins = []
            for idx in itertools.product(*[range(s_) for s_ in nsz]):
                t = []
                for ci, coef in enumerate(rows[idx[dim]]):
                    if abs(coef) < 1e-15:
                        continue
                    src = list(idx)
                    src[dim] = ci
                    t.append((cur[tuple(src)], SID["one"], coef))
                assert len(t) <= MAXTERMS
                key = ("t", dim, idx)
                new[idx] = key
                ins.append((key, t))
            sym_levels.append(ins)
            cur = new
            sizes = nsz
    out_sym = [cur[idx] for idx in itertools.product(*[range(s_) for s_ in sizes])]
    # allocation with reuse: a slot is freed after the last level reading it
    last = {}
    for li, ins in enumerate(sym_levels):
        for _, t in ins:
            for src, _, _ in t:
                last[src] = li
    keep = set(out_sym)
    where = {("acc",) + ef: acc_slot[ef] for ef in acc_slot}
    free = list(range(nacc, vrr_end))
    top = vrr_end
    post = []
    import heapq
    heapq.heapify(free)
    for li, ins in enumerate(sym_levels):
        rows = []
        for key, t in ins:
            if free:
                d = heapq.heappop(free)
            else:
                d = top
                top += 1
            where[key] = d
            rows.append((d, [(where[src], sid, c) for src, sid, c in t]))
        post.append(rows)
        for v, lu in last.items():
            if lu == li and v not in keep and v in where:
                heapq.heappush(free, where[v])
    out = [where[k] for k in out_sym]
    nxt[0] = top
    return {"L": L, "nslots": max(vrr_end, nxt[0]), "base": base, "acc": acc, "nacc": nacc,
            "vrr_levels": vrr_levels, "post_levels": post, "out": out, "sizes": sizes}


HEAVY = [(2, 1, 2, 1), (2, 2, 1, 1), (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2)]


def write_programs(outdir, classes=None):
    """Emit all heavy-class programs (pure and Cartesian) as flat tables:
    PROG_INS[i] = {dst, nterm, src0..5, sid0..5} (14 ints), PROG_COEF[i][6],
    PROG_DESC[k] describes class k (see ClassProgram in heavy_kernel.hpp)."""
    classes = classes or HEAVY
    ins_i, ins_c, lev, accs, outs, desc = [], [], [], [], [], []
    for (la, lb, lc, ld) in classes:
        for pure in (0, 1):
            P = gen_program(la, lb, lc, ld, pure)
            d = {"la": la, "lb": lb, "lc": lc, "ld": ld, "pure": pure, "L": P["L"], "nslots": P["nslots"],
                 "base": P["base"], "nacc": P["nacc"], "acc_off": len(accs) // 2}
            for a_, v_ in P["acc"]:
                accs += [a_, v_]
            d["vrr_lev_off"] = len(lev)
            d["n_vrr_lev"] = len(P["vrr_levels"])
            for lvl in P["vrr_levels"]:
                lev.append(len(ins_c) // MAXTERMS)
                for dst, t in lvl:
                    row = [dst, len(t)] + [x[0] for x in t] + [0] * (MAXTERMS - len(t)) + \
                        [x[1] for x in t] + [0] * (MAXTERMS - len(t))
                    ins_i += row
                    ins_c += [x[2] for x in t] + [0.0] * (MAXTERMS - len(t))
            lev.append(len(ins_c) // MAXTERMS)
            d["post_lev_off"] = len(lev)
            d["n_post_lev"] = len(P["post_levels"])
            for lvl in P["post_levels"]:
                lev.append(len(ins_c) // MAXTERMS)
                for dst, t in lvl:
                    row = [dst, len(t)] + [x[0] for x in t] + [0] * (MAXTERMS - len(t)) + \
                        [x[1] for x in t] + [0] * (MAXTERMS - len(t))
                    ins_i += row
                    ins_c += [x[2] for x in t] + [0.0] * (MAXTERMS - len(t))
            lev.append(len(ins_c) // MAXTERMS)
            d["out_off"] = len(outs)
            d["nout"] = len(P["out"])
            outs += P["out"]
            d["na"], d["nb"], d["nc"], d["nd"] = P["sizes"]
            desc.append(d)
    keys = ["la", "lb", "lc", "ld", "pure", "L", "nslots", "base", "nacc", "acc_off", "vrr_lev_off",
            "n_vrr_lev", "post_lev_off", "n_post_lev", "out_off", "nout", "na", "nb", "nc", "nd"]
    lines = ["// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.", "#pragma once", "namespace m2 {",
             "constexpr int PROG_MAXTERMS = %d;" % MAXTERMS,
             "constexpr int PROG_NCLASS = %d;" % len(desc),
             "constexpr int PROG_DESC_FIELDS = %d;" % len(keys),
             "// fields: " + " ".join(keys),
             "constexpr int PROG_DESC[%d] = {" % (len(desc) * len(keys))]
    for d in desc:
        lines.append("  " + ", ".join(str(d[k]) for k in keys) + ",")
    lines.append("};")

    def arr(name, typ, vals, per=16):
        lines.append("constexpr %s %s[%d] = {" % (typ, name, max(1, len(vals))))
        for i in range(0, len(vals), per):
            lines.append("  " + ", ".join(repr(v) if typ == "double" else str(v) for v in vals[i:i + per]) + ",")
        if not vals:
            lines.append("  0")
        lines.append("};")

    # VRR instruction words of the heavy J/K kernel in pass order: for every
    # VRR level ceil(n / 64) passes; in a pass lane j executes instruction
    # 64 q + j of the level.  Per pass and lane 7 ints, stored [pass][word]
    # [lane] so one pass's words load coalesced: header dst | nterm << 16 (-1:
    # idle lane), then per term src | sid << 13 | (coef & 63) << 18 (coef a
    # small signed integer).  The VRR part of a class's program does not
    # depend on pure / Cartesian output.
    vdesc, vwords, vend = [], [], []
    for (la, lb, lc, ld) in classes:
        P0 = gen_program(la, lb, lc, ld, 0)
        P1 = gen_program(la, lb, lc, ld, 1)
        assert P0["vrr_levels"] == P1["vrr_levels"] and P0["base"] == P1["base"]
        passes, ends = [], []
        for lvl in P0["vrr_levels"]:
            npass = (len(lvl) + 63) // 64
            for q in range(npass):
                w = [[-1] * 64] + [[0] * 64 for _ in range(MAXTERMS)]
                for j in range(64):
                    k = 64 * q + j
                    if k >= len(lvl):
                        continue
                    dst, t = lvl[k]
                    assert dst < 65536 and len(t) <= MAXTERMS
                    w[0][j] = dst | (len(t) << 16)
                    for ti, (src, sid, c) in enumerate(t):
                        ci = int(round(c))
                        assert abs(ci - c) < 1e-12 and -32 <= ci < 32 and src < 8192 and sid < 32
                        w[1 + ti][j] = src | (sid << 13) | ((ci & 63) << 18)
                passes.append(w)
                ends.append(q == npass - 1)
        assert len(passes) <= 64
        vdesc.append((la, lb, lc, ld, len(passes), len(vwords)))
        for w in passes:
            for row in w:
                vwords += row
        vend.append(sum(1 << i for i, e in enumerate(ends) if e))
    # Post (HRR, solid-harmonic) instruction words in pass order, per class
    # and output kind, laid out as the VRR words.  A post term is either
    # coef * X[src] (solid-harmonic transform, scalar "one") or S[AB/CD] *
    # X[src] with coefficient 1 (HRR), so a term word is src | sid << 13 with
    # the program's non-unit coefficients as extra scalars S[24 + i]
    # (POSTW_COEF, <= 8 per program); the passes that end a level are listed
    # in POSTW_END.
    pdesc, pwords, pcoef, pend = [], [], [], []
    for (la, lb, lc, ld) in classes:
        for pure in (0, 1):
            P = gen_program(la, lb, lc, ld, pure)
            coefs = []
            word_off, coef_off, end_off = len(pwords), len(pcoef), len(pend)
            npass = 0
            for lvl in P["post_levels"]:
                nq = (len(lvl) + 63) // 64
                for q in range(nq):
                    w = [[-1] * 64] + [[0] * 64 for _ in range(MAXTERMS)]
                    for j in range(64):
                        k = 64 * q + j
                        if k >= len(lvl):
                            continue
                        dst, t = lvl[k]
                        w[0][j] = dst | (len(t) << 16)
                        for ti, (src, sid, c) in enumerate(t):
                            if c != 1.0:
                                assert sid == SID["one"]
                                if c not in coefs:
                                    coefs.append(c)
                                sid = NSCALAR + coefs.index(c)
                            assert src < 8192 and sid < NSCALAR + 8
                            w[1 + ti][j] = src | (sid << 13)
                    for row in w:
                        pwords += row
                    pend.append(1 if q == nq - 1 else 0)
                    npass += 1
            assert len(coefs) <= 8
            pcoef += coefs
            pdesc.append((la, lb, lc, ld, pure, npass, word_off, coef_off, len(coefs), end_off))
    lines.append("constexpr int POSTW_NPROG = %d;" % len(pdesc))
    lines.append("// per program: la lb lc ld pure npass word_offset coef_offset ncoef end_offset")
    lines.append("constexpr int POSTW_DESC[%d] = {%s};" % (10 * len(pdesc), ", ".join(str(x) for d in pdesc for x in d)))
    arr("POSTW_WORDS", "int", pwords)
    arr("POSTW_COEF", "double", [float(c) for c in pcoef], 8)
    arr("POSTW_END", "int", pend)
    lines.append("constexpr int VRRW_NCLASS = %d;" % len(vdesc))
    lines.append("// per class: la lb lc ld npass word_offset")
    lines.append("constexpr int VRRW_DESC[%d] = {%s};" % (6 * len(vdesc), ", ".join(str(x) for d in vdesc for x in d)))
    lines.append("// per class: bit p set when VRR pass p ends a level")
    lines.append("constexpr unsigned long long VRRW_LEVEND[%d] = {%s};" % (
        len(vend), ", ".join("%dull" % e for e in vend)))
    arr("VRRW_WORDS", "int", vwords)
    arr("PROG_INS", "int", ins_i)
    arr("PROG_COEF", "double", [float(x) for x in ins_c], 8)
    arr("PROG_LEV", "int", lev)
    arr("PROG_ACC", "int", accs)
    arr("PROG_OUT", "int", outs)
    lines.append("}  // namespace m2")
    text = "\n".join(lines) + "\n"
    path = os.path.join(outdir, "programs.inc")
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as fh:
            fh.write(text)
    return path


# ---------------------------------------------------------------------------
# Derivative (gradient adjoint) programs
# ---------------------------------------------------------------------------
DSID = dict(SID, wA=24, wB=25, wC=26)
NSCALAR_D = 27


def _alloc_levels(levels, reads_after, pinned, first_free, preset=None):
    """Slot allocation with reuse for instruction levels.
    levels: list of lists of (key, [(src_key, sid, coef)])
    reads_after: extra {key: last level index} uses outside the instruction
    lists (e.g. accumulation right after the level that defines the key)
    pinned: keys whose slots are never freed
    preset: {key: slot} already allocated (inputs)
    Returns (slot map, allocated instruction levels, top slot)."""
    import heapq
    last = dict(reads_after)
    for li, ins in enumerate(levels):
        for _, t in ins:
            for src, _, _ in t:
                last[src] = max(last.get(src, -1), li)
    where = dict(preset or {})
    free = []
    top = first_free
    out = []
    for li, ins in enumerate(levels):
        rows = []
        for key, t in ins:
            if free:
                d = heapq.heappop(free)
            else:
                d = top
                top += 1
            where[key] = d
            rows.append((d, [(where[src], sid, c) for src, sid, c in t]))
        out.append(rows)
        for v, lu in list(last.items()):
            if lu == li and v not in pinned and v in where:
                heapq.heappush(free, where[v])
                del last[v]
    return where, out, top


def gen_deriv_program(la, lb, lc, ld):
    """Cartesian derivative-adjoint program of class (la lb|lc ld).

    VRR per primitive quartet to the nodes needed by
      S0 = sum w [e|f]       e in [la-1, la+lb],   f in [lc-1, lc+ld]
      SA = sum 2a w [e|f]    e in [la+1, la+lb+1], f in [lc, lc+ld]
      SB = sum 2b w [e|f]    e in [la, la+lb+1],   f in [lc, lc+ld]
      SC = sum 2c w [e|f]    e in [la, la+lb],     f in [lc+1, lc+ld+1]
    (accumulated right after the VRR level that produces each node), then
    HRR to the shifted Cartesian blocks (la+-1, ...), (la, lb+-1, ...),
    (..., lc+-1, ld), and a contraction list
      g[3*X + i] += coef * Gamma[abcd] * X[slot]
    with dX_i (ab|cd) = 2x (..x+1_i..) - x_i (..x-1_i..)."""
    Lp = la + lb + lc + ld + 1

    def rng(lo, hi):
        return [c for L_ in range(max(lo, 0), hi + 1) for c in comps(L_)]

    sets = {
        "S0": (rng(la - 1, la + lb), rng(lc - 1, lc + ld), "one"),
        "SA": (rng(la + 1, la + lb + 1), rng(lc, lc + ld), "wA"),
        "SB": (rng(la, la + lb + 1), rng(lc, lc + ld), "wB"),
        "SC": (rng(la, la + lb), rng(lc + 1, lc + ld + 1), "wC"),
    }
    targets = set()
    for E, F, _ in sets.values():
        for e in E:
            for f in F:
                targets.add((e, f))
    order, nodes = build_vrr(sorted(targets), Lp)
    # accumulator slots first (pinned)
    acc_slot = {}
    k = 0
    for name, (E, F, _) in sets.items():
        for e in E:
            for f in F:
                acc_slot[(name, e, f)] = k
                k += 1
    nacc = k
    base = nacc
    preset = {((0, 0, 0), (0, 0, 0), m): base + m for m in range(Lp + 1)}
    levels = [[] for _ in range(Lp + 1)]
    for key in order:
        e, f, m = key
        kind = nodes[key]
        if kind[0] == "base":
            continue
        _, terms, extra = kind
        t = [(src, DSID[c], 1.0) for c, src in terms]
        for x in extra:
            if x[0] == "ff":
                t += [(x[2], DSID["inv2q"], float(x[1])), (x[3], DSID["inv2q_rho_q"], -float(x[1]))]
            elif x[0] == "ee":
                t += [(x[2], DSID["inv2p"], float(x[1])), (x[3], DSID["inv2p_rho_p"], -float(x[1]))]
            elif x[0] == "ef":
                t.append((x[2], DSID["inv2pq"], float(x[1])))
        levels[sum(e) + sum(f)].append((key, t))
    # accumulation after the level of each target node
    acc_lists = [[] for _ in range(Lp + 1)]
    reads_after = {}
    for name, (E, F, w) in sets.items():
        for e in E:
            for f in F:
                lv = sum(e) + sum(f)
                acc_lists[lv].append((acc_slot[(name, e, f)], (e, f, 0), DSID[w]))
                reads_after[(e, f, 0)] = max(reads_after.get((e, f, 0), -1), lv)
    where, vl, top = _alloc_levels(levels, reads_after, set(), base + Lp + 1, preset)
    acc_ins = [[(a, where[key], sid) for a, key, sid in lst] for lst in acc_lists]
    # drop empty leading levels but keep level/acc pairing
    vrr_levels, acc_levels = [], []
    for li in range(Lp + 1):
        if vl[li] or acc_ins[li]:
            vrr_levels.append(vl[li])
            acc_levels.append(acc_ins[li])
    vrr_top = top
    # post phase: six shifted blocks
    post_sym = []
    blocks = {}

    def hrr_block(name, sa, sb, sc, sd):
        """symbolic HRR from accumulator set `name` to Cartesian (sa sb|sc sd)"""
        E = rng(sa, sa + sb)
        F = rng(sc, sc + sd)
        lv_k = []

        def kget(e, c, d):
            return ("acc", name, e, c) if sum(d) == 0 else ("k", name, sa, sb, sc, sd, e, c, d)
        for lev in range(1, sd + 1):
            ins = []
            for e in E:
                for Lc in range(sc, sc + sd - lev + 1):
                    for c in comps(Lc):
                        for d in comps(lev):
                            i = next(q for q in range(3) if d[q] > 0)
                            ins.append((kget(e, c, d), [(kget(e, add(c, i), sub(d, i)), DSID["one"], 1.0),
                                                        (kget(e, c, sub(d, i)), DSID["CD" + AX[i]], 1.0)]))
            lv_k.append(ins)

        def bget(a, b, c, d):
            return kget(a, c, d) if sum(b) == 0 else ("b", name, sa, sb, sc, sd, a, b, c, d)
        lv_b = []
        for lev in range(1, sb + 1):
            ins = []
            for La in range(sa, sa + sb - lev + 1):
                for a in comps(La):
                    for b in comps(lev):
                        for c in comps(sc):
                            for d in comps(sd):
                                i = next(q for q in range(3) if b[q] > 0)
                                ins.append((bget(a, b, c, d), [(bget(add(a, i), sub(b, i), c, d), DSID["one"], 1.0),
                                                               (bget(a, sub(b, i), c, d), DSID["AB" + AX[i]], 1.0)]))
            lv_b.append(ins)
        vals = {}
        for a in comps(sa):
            for b in comps(sb):
                for c in comps(sc):
                    for d in comps(sd):
                        vals[(a, b, c, d)] = bget(a, b, c, d)
        return lv_k, lv_b, vals

    terms = [("A+", "SA", (la + 1, lb, lc, ld), 0, +1), ("B+", "SB", (la, lb + 1, lc, ld), 1, +1),
             ("C+", "SC", (la, lb, lc + 1, ld), 2, +1)]
    if la > 0:
        terms.append(("A-", "S0", (la - 1, lb, lc, ld), 0, -1))
    if lb > 0:
        terms.append(("B-", "S0", (la, lb - 1, lc, ld), 1, -1))
    if lc > 0:
        terms.append(("C-", "S0", (la, lb, lc - 1, ld), 2, -1))
    # terms one after another (each term's HRR levels, then the next term):
    # only one term's intermediates are live at a time
    seq = []
    for nm, accn, cls, ctr, sgn in terms:
        lv_k, lv_b, vals = hrr_block(accn, *cls)
        blocks[nm] = (cls, ctr, sgn, vals)
        seq.extend(lv_k)
        seq.extend(lv_b)
    # dedupe identical symbolic nodes (S0 blocks share intermediates)
    post_sym = []
    seen = set()
    for lv in seq:
        ins = []
        for key, t in lv:
            if key in seen:
                continue
            seen.add(key)
            ins.append((key, t))
        post_sym.append(ins)
    # contraction list (reads block values after the post phase)
    gam = {}
    ca, cb, cc, cd = comps(la), comps(lb), comps(lc), comps(ld)
    for ia, a in enumerate(ca):
        for ib, b in enumerate(cb):
            for ic, c in enumerate(cc):
                for idd, d in enumerate(cd):
                    gam[(a, b, c, d)] = ((ia * len(cb) + ib) * len(cc) + ic) * len(cd) + idd
    contr = []
    for nm, (cls, ctr, sgn, vals) in blocks.items():
        for (a, b, c, d), key in vals.items():
            quad = [a, b, c, d]
            for i in range(3):
                q = list(quad[ctr])
                if sgn > 0:
                    if q[i] == 0:
                        continue
                    q[i] -= 1
                    coef = 1.0
                else:
                    q[i] += 1
                    coef = -float(q[i])
                src = list(quad)
                src[ctr] = tuple(q)
                contr.append((key, gam[tuple(src)], coef, 3 * ctr + i))
    # contract every block value right after the post level that produces
    # it (accumulator-sourced values before the first level), so its slot
    # can be reused by later levels
    defined = {}
    for li, lv in enumerate(post_sym):
        for key, _ in lv:
            defined[key] = li
    preset_post = {("acc",) + tuple(k): v for k, v in acc_slot.items()}
    reads = {}
    for key, g, c, x in contr:
        if key in defined:
            reads[key] = max(reads.get(key, -1), defined[key])
    where2, pl, top2 = _alloc_levels(post_sym, reads, set(preset_post.keys()), nacc, preset_post)
    contr_levels = [[] for _ in range(len(post_sym) + 1)]   # [0]: before level 0
    for key, g, c, x in contr:
        li = defined.get(key, -1)
        contr_levels[li + 1].append((where2[key], g, c, x))
    return {"L": Lp, "nslots": max(vrr_top, top2), "base": base, "nacc": nacc,
            "vrr_levels": vrr_levels, "acc_levels": acc_levels, "post_levels": pl,
            "contr_levels": contr_levels, "contr": [e for lv in contr_levels for e in lv], "ngam": len(gam),
            "sizes": [len(ca), len(cb), len(cc), len(cd)]}


def write_deriv_programs(outdir):
    """Orbital class combinations (bra class >= ket class) not in DERIV_REGISTER."""
    ins_i, ins_c, lev, accs, contr_i, contr_c, desc = [], [], [], [], [], [], []
    for X, (la, lb) in enumerate(PAIR_CLASSES):
        for Y in range(X + 1):
            lc, ld = PAIR_CLASSES[Y]
            if (la, lb, lc, ld) in DERIV_REGISTER:
                continue   # register code (deriv_register.inc)
            P = gen_deriv_program(la, lb, lc, ld)
            d = {"la": la, "lb": lb, "lc": lc, "ld": ld, "L": P["L"], "nslots": P["nslots"], "base": P["base"],
                 "nacc": P["nacc"], "lev_off": len(lev), "nvlev": len(P["vrr_levels"])}
            # per VRR level: instruction range, then accumulation range
            for lv, al in zip(P["vrr_levels"], P["acc_levels"]):
                lev.append(len(ins_c) // MAXTERMS)
                for dst, t in lv:
                    ins_i += [dst, len(t)] + [x[0] for x in t] + [0] * (MAXTERMS - len(t)) + \
                        [x[1] for x in t] + [0] * (MAXTERMS - len(t))
                    ins_c += [x[2] for x in t] + [0.0] * (MAXTERMS - len(t))
                lev.append(len(ins_c) // MAXTERMS)
                lev.append(len(accs) // 3)
                for a_, v_, sid in al:
                    accs += [a_, v_, sid]
                lev.append(len(accs) // 3)
            # post phase: contraction list before level 0, then for each
            # level its instruction range and its contraction range
            d["post_off"] = len(lev)
            d["nplev"] = len(P["post_levels"])
            cl = P["contr_levels"]
            lev.append(len(contr_c))
            for slot, g, c, x in cl[0]:
                contr_i += [slot, g, x]
                contr_c.append(c)
            lev.append(len(contr_c))
            for li, lv in enumerate(P["post_levels"]):
                lev.append(len(ins_c) // MAXTERMS)
                for dst, t in lv:
                    ins_i += [dst, len(t)] + [x[0] for x in t] + [0] * (MAXTERMS - len(t)) + \
                        [x[1] for x in t] + [0] * (MAXTERMS - len(t))
                    ins_c += [x[2] for x in t] + [0.0] * (MAXTERMS - len(t))
                lev.append(len(ins_c) // MAXTERMS)
                lev.append(len(contr_c))
                for slot, g, c, x in cl[li + 1]:
                    contr_i += [slot, g, x]
                    contr_c.append(c)
                lev.append(len(contr_c))
            d["contr_off"] = 0
            d["ncontr"] = len(P["contr"])
            d["na"], d["nb"], d["nc"], d["nd"] = P["sizes"]
            desc.append(d)
    keys = ["la", "lb", "lc", "ld", "L", "nslots", "base", "nacc", "lev_off", "nvlev", "post_off", "nplev",
            "contr_off", "ncontr", "na", "nb", "nc", "nd"]
    lines = ["// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.", "#pragma once", "namespace m2 {",
             "constexpr int DPROG_NCLASS = %d;" % len(desc),
             "constexpr int DPROG_DESC_FIELDS = %d;" % len(keys),
             "// fields: " + " ".join(keys),
             "constexpr int DPROG_DESC[%d] = {" % (len(desc) * len(keys))]
    for d in desc:
        lines.append("  " + ", ".join(str(d[k]) for k in keys) + ",")
    lines.append("};")

    def arr(name, typ, vals, per=16):
        lines.append("constexpr %s %s[%d] = {" % (typ, name, max(1, len(vals))))
        for i in range(0, len(vals), per):
            lines.append("  " + ", ".join(repr(v) if typ == "double" else str(v) for v in vals[i:i + per]) + ",")
        if not vals:
            lines.append("  0")
        lines.append("};")

    # The VRR phase (run once per primitive quartet) also in pass order with
    # prefetch, as the J/K heavy VRR: per pass and lane 7 ints.  Instruction
    # passes: header dst | nterm << 16 (-1: idle), terms src | sid << 14 |
    # (coef & 63) << 19.  Accumulation passes (X[acc] += S[sid] X[vrr]):
    # word 0 = acc | sid << 16 (-1: idle), word 1 = vrr slot.  DVRR_KIND per
    # pass: 0 instructions, 1 accumulation, + 4 when a barrier follows.
    dwords, dkind, ddesc = [], [], []
    for X, (la, lb) in enumerate(PAIR_CLASSES):
        for Y in range(X + 1):
            lc, ld = PAIR_CLASSES[Y]
            if (la, lb, lc, ld) in DERIV_REGISTER:
                continue
            P = gen_deriv_program(la, lb, lc, ld)
            first = len(dkind)
            for lv, al in zip(P["vrr_levels"], P["acc_levels"]):
                for kind, items in ((0, lv), (1, al)):
                    nq = (len(items) + 63) // 64
                    for q in range(nq):
                        w = [[-1] * 64] + [[0] * 64 for _ in range(MAXTERMS)]
                        for j in range(64):
                            k = 64 * q + j
                            if k >= len(items):
                                continue
                            if kind == 0:
                                dst, t = items[k]
                                assert dst < 65536 and len(t) <= MAXTERMS
                                w[0][j] = dst | (len(t) << 16)
                                for ti, (src, sid, c) in enumerate(t):
                                    ci = int(round(c))
                                    assert abs(ci - c) < 1e-12 and -32 <= ci < 32 and src < 16384 and sid < 32
                                    w[1 + ti][j] = src | (sid << 14) | ((ci & 63) << 19)
                            else:
                                acc_, vrr_, sid = items[k]
                                assert acc_ < 65536 and sid < 32
                                w[0][j] = acc_ | (sid << 16)
                                w[1][j] = vrr_
                        for row in w:
                            dwords += row
                        dkind.append(kind + (4 if q == nq - 1 else 0))
            ddesc.append((la, lb, lc, ld, first, len(dkind) - first))
    lines.append("constexpr int DVRR_NCLASS = %d;" % len(ddesc))
    lines.append("// per class: la lb lc ld first_pass npass")
    lines.append("constexpr int DVRR_DESC[%d] = {%s};" % (6 * len(ddesc), ", ".join(str(x) for d in ddesc for x in d)))
    arr("DVRR_KIND", "int", dkind)
    arr("DVRR_WORDS", "int", dwords)
    arr("DPROG_INS", "int", ins_i)
    arr("DPROG_COEF", "double", [float(x) for x in ins_c], 8)
    arr("DPROG_LEV", "int", lev)
    arr("DPROG_ACC", "int", accs)
    arr("DPROG_CONTR", "int", contr_i)
    arr("DPROG_CCOEF", "double", [float(x) for x in contr_c], 8)
    lines.append("}  // namespace m2")
    text = "\n".join(lines) + "\n"
    path = os.path.join(outdir, "deriv_programs.inc")
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as fh:
            fh.write(text)
    return path


DERIV_REGISTER = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (2, 0, 0, 0), (1, 1, 1, 0), (2, 0, 1, 0),
                  (2, 1, 0, 0), (1, 1, 1, 1), (2, 0, 1, 1), (2, 0, 2, 0), (2, 1, 1, 0), (2, 1, 2, 0), (2, 2, 0, 0),
                  (2, 2, 1, 0), (2, 1, 1, 1), (2, 1, 2, 1), (2, 2, 2, 0), (2, 2, 1, 1), (2, 2, 2, 1)]
# classes whose three derivative centres run as three lighter passes
DERIV_SPLIT = {(1, 1, 1, 0), (2, 0, 1, 0), (1, 1, 1, 1), (2, 0, 1, 1), (2, 0, 2, 0), (2, 1, 1, 0), (2, 1, 2, 0),
               (2, 2, 0, 0), (2, 2, 1, 0), (2, 1, 1, 1), (2, 1, 2, 1), (2, 2, 2, 0), (2, 2, 1, 1), (2, 2, 2, 1)}
# split classes whose every centre pass is split again over the c components
# of Gamma (fewer accumulators per pass, each pass with its own VRR).  The
# FP64 passes of (dd|pp) and (dd|dp) still spill (1.3-4.3 K VGPRs each, to
# scratch) and run faster than the LDS interpreter all the same: (H2O)20
# cc-pVDZ gradient 0.176 -> 0.165 s.  (dd|dd) stays in the interpreter: its
# register passes take over half an hour to compile.
DERIV_CSPLIT = {(2, 1, 1, 1), (2, 1, 2, 1), (2, 2, 2, 0), (2, 2, 1, 1), (2, 2, 2, 1)}


def gen_deriv_class(la, lb, lc, ld, mask=7):
    """Register-resident derivative adjoint of one small class:
    EriDeriv<la,lb,lc,ld,mask>::run<R>(..., gam_fill, g) adds to g[0..8] the
    contraction of d/dA (mask & 1), d/dB (mask & 2), d/dC (mask & 4) of
    (ab|cd) (Cartesian) with Gamma; the other entries of g are left alone.
    mask >> 3 = k > 0 keeps only the Gamma entries whose c component is
    comps(lc)[k - 1] (one of the passes of a class split over its c
    components): the HRR is generated on demand for those outputs, and only
    the accumulators it reads are kept."""
    Lp = la + lb + lc + ld + 1
    part = mask >> 3
    cmask = mask & 7

    def rng(lo, hi):
        return [c for L_ in range(max(lo, 0), hi + 1) for c in comps(L_)]

    def rect(E, F):
        return [(e, f) for e in E for f in F]

    ca, cb, cc, cd = comps(la), comps(lb), comps(lc), comps(ld)
    gidx = {}
    for ia, a in enumerate(ca):
        for ib, b in enumerate(cb):
            for ic, c in enumerate(cc):
                for idd, d in enumerate(cd):
                    gidx[(a, b, c, d)] = ((ia * len(cb) + ib) * len(cc) + ic) * len(cd) + idd
    keep_c = cc[part - 1] if part else None

    # post phase first (into its own emitter): HRR per term, on demand, then
    # the contraction; records the accumulators it reads
    post = Emitter()
    used = set()
    defined = set()

    def hrr_get(acc, sa, sb, sc, sd, quad):
        memo = {}

        def kget(e, c, d):
            if sum(d) == 0:
                used.add((acc, e, c))
                return "%s_%s_%s" % (acc, cname(e), cname(c))
            key = ("k", e, c, d)
            if key in memo:
                return memo[key]
            i = next(q for q in range(3) if d[q] > 0)
            n1 = kget(e, add(c, i), sub(d, i))
            n0 = kget(e, c, sub(d, i))
            nm = "k%s_%d%d%d%d_%s_%s_%s" % (acc, sa, sb, sc, sd, cname(e), cname(c), cname(d))
            if nm not in defined:
                post.out("    const R %s = %s + CD%s * %s;" % (nm, n1, AX[i], n0))
                defined.add(nm)
            memo[key] = nm
            return nm

        def bget(a, b, c, d):
            if sum(b) == 0:
                return kget(a, c, d)
            key = ("b", a, b, c, d)
            if key in memo:
                return memo[key]
            i = next(q for q in range(3) if b[q] > 0)
            n1 = bget(add(a, i), sub(b, i), c, d)
            n0 = bget(a, sub(b, i), c, d)
            nm = "h%s_%d%d%d%d_%s_%s_%s_%s" % (acc, sa, sb, sc, sd, cname(a), cname(b), cname(c), cname(d))
            if nm not in defined:
                post.out("    const R %s = %s + AB%s * %s;" % (nm, n1, AX[i], n0))
                defined.add(nm)
            memo[key] = nm
            return nm

        return bget(*quad)

    terms = []
    if cmask & 1:
        terms.append(("sa", (la + 1, lb, lc, ld), 0, +1))
    if cmask & 2:
        terms.append(("sb", (la, lb + 1, lc, ld), 1, +1))
    if cmask & 4:
        terms.append(("sc", (la, lb, lc + 1, ld), 2, +1))
    if la > 0 and cmask & 1:
        terms.append(("s0", (la - 1, lb, lc, ld), 0, -1))
    if lb > 0 and cmask & 2:
        terms.append(("s0", (la, lb - 1, lc, ld), 1, -1))
    if lc > 0 and cmask & 4:
        terms.append(("s0", (la, lb, lc - 1, ld), 2, -1))
    sums = {k: [] for k in range(9)}
    for acc, cls, ctr, sgn in terms:
        for a in comps(cls[0]):
            for b in comps(cls[1]):
                for c in comps(cls[2]):
                    for d in comps(cls[3]):
                        quad = [a, b, c, d]
                        for i in range(3):
                            q = list(quad[ctr])
                            if sgn > 0:
                                if q[i] == 0:
                                    continue
                                q[i] -= 1
                                coef = None
                            else:
                                q[i] += 1
                                coef = float(q[i])
                            src = list(quad)
                            src[ctr] = tuple(q)
                            if keep_c is not None and src[2] != keep_c:
                                continue
                            nm = hrr_get(acc, *cls, (a, b, c, d))
                            sums[3 * ctr + i].append(("+" if coef is None else "-", coef, gidx[tuple(src)], nm))
    for k in range(9):
        if not sums[k]:
            continue
        parts = []
        for sg, coef, gi, nm in sums[k]:
            if coef is None:
                parts.append("+ Gam[%d] * %s" % (gi, nm))
            else:
                parts.append("- %s * Gam[%d] * %s" % (lit(coef), gi, nm))
        post.out("    g[%d] += R(0) %s;" % (k, " ".join(parts)))

    # accumulator sets, pruned to what the post phase reads
    weight = {"s0": None, "sa": "wA", "sb": "wB", "sc": "wC"}
    sets = {}
    for nm in ("s0", "sa", "sb", "sc"):
        T_ = sorted((e, f) for a_, e, f in used if a_ == nm)
        if T_:
            sets[nm] = (T_, weight[nm])
    targets = sorted({t for T_, _ in sets.values() for t in T_})
    order, nodes = build_vrr(targets, Lp)
    em = Emitter()
    em.out("template <> struct EriDeriv<%d, %d, %d, %d, %d> {" % (la, lb, lc, ld, mask))
    ngam = len(ca) * len(cb) * len(cc) * len(cd)
    em.out("  static constexpr int NACC = %d;" % sum(len(T_) for T_, _ in sets.values()))
    em.out("  template <typename R, typename GamFill>")
    em.out("  static __device__ __forceinline__ void run(const PrimPair* __restrict__ bra, int nbra,")
    em.out("      const PrimPair* __restrict__ ket, int nket, const R ABx, const R ABy, const R ABz,")
    em.out("      const R CDx, const R CDy, const R CDz, const BoysTables& tab, const GamFill& gam_fill, R* __restrict__ g,")
    em.out("      const float* __restrict__ bq, const float* __restrict__ kq, float cut) {")
    for nm, (T_, _) in sets.items():
        for e, f in T_:
            em.out("    R %s_%s_%s = R(0);" % (nm, cname(e), cname(f)))
    # primitive screening as in the J/K classes: primitive pair bounds sorted
    # descending, a lane stops at its first ket primitive below its cut, the
    # lanes of the bra loop at the first bra primitive below every lane's cut
    em.out("    const float bq0 = bq[0];")
    em.out("    for (int ik = 0; ik < nket; ++ik) {")
    em.out("      const float qk = kq[ik];")
    em.out("      if (qk * bq0 < cut) break;")
    em.out("      const PrimPair kp = ket[ik];")
    em.out("      const R q = R(kp.p), Kcd = R(kp.K), wC = R(2) * R(kp.a);")
    em.out("      const R Qx = R(kp.P[0]), Qy = R(kp.P[1]), Qz = R(kp.P[2]);")
    em.out("      const R inv2q = R(kp.inv2p);")
    em.out("      const R QCx = R(kp.PA[0]), QCy = R(kp.PA[1]), QCz = R(kp.PA[2]);")
    em.out("      for (int ib = 0; ib < nbra; ++ib) {")
    em.out("        if (!__any(bq[ib] * qk >= cut)) break;")
    em.out("        const PrimPair bp = bra[ib];")
    em.out("        const R p = R(bp.p), Kab = R(bp.K), wA = R(2) * R(bp.a), wB = R(2) * (R(bp.p) - R(bp.a));")
    em.out("        const R inv2p = R(bp.inv2p);")
    em.out("        const R PAx = R(bp.PA[0]), PAy = R(bp.PA[1]), PAz = R(bp.PA[2]);")
    em.out("        const R dx = R(bp.P[0]) - Qx, dy = R(bp.P[1]) - Qy, dz = R(bp.P[2]) - Qz;")
    em.out("        const R rs_pq = sqrt_inv<R>(p + q);")
    em.out("        const R inv_pq = rs_pq * rs_pq;")
    em.out("        const R rho = p * q * inv_pq;")
    em.out("        const R T = rho * (dx * dx + dy * dy + dz * dz);")
    em.out("        R Fm[%d];" % (Lp + 1))
    em.out("        BoysEval<R, %d>::run(T, Fm, tab);" % Lp)
    em.out("        const R pref = Kab * Kcd * rs_pq;")
    em.out("#pragma unroll")
    em.out("        for (int m = 0; m <= %d; ++m) Fm[m] *= pref;" % Lp)
    em.out("        const R rho_p = q * inv_pq, rho_q = p * inv_pq, inv2pq = R(0.5) * inv_pq;")
    em.out("        const R WPx = -rho_p * dx, WPy = -rho_p * dy, WPz = -rho_p * dz;")
    em.out("        const R WQx = rho_q * dx, WQy = rho_q * dy, WQz = rho_q * dz;")
    # each target is accumulated right after its VRR node, so that it does
    # not stay live to the end of the primitive iteration
    acc_of = {}
    for nm, (T_, w) in sets.items():
        for e, f in T_:
            acc_of.setdefault((e, f), []).append((nm, w))
    for key in order:
        emit_vrr(em, [key], nodes, "        ")
        e, f, m = key
        if m == 0 and (e, f) in acc_of:
            v = vrr_name(e, f, 0)
            for nm, w in acc_of[(e, f)]:
                em.out("        %s_%s_%s += %s;" % (nm, cname(e), cname(f), v if w is None else "%s * %s" % (w, v)))
    em.out("      }")
    em.out("    }")
    # Gamma is formed only now, so that it does not hold registers in the loop
    em.out("    R Gam[%d];" % ngam)
    em.out("    gam_fill(Gam);")
    em.lines.extend(post.lines)
    em.out("  }")
    em.out("};")
    return "\n".join(em.lines)


def write_deriv_register(outdir):
    """One translation unit per register derivative class (grad_reg_ABCD.hip,
    compiled in parallel), the host-side class table (deriv_classes.inc) and
    the dispatcher (deriv_dispatch.inc)."""
    def put(path, text):
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as fh:
                fh.write(text)

    head = "// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit."
    names = []
    for c in DERIV_REGISTER:
        tag = "%d%d%d%d" % c
        masks = (1, 2, 4) if c in DERIV_SPLIT else (7,)
        if c in DERIV_CSPLIT:
            masks = tuple(m | (k << 3) for m in masks for k in range(1, len(comps(c[2])) + 1))
        # classes split over the c components get one translation unit per
        # derivative centre (grad_reg_ABCD_k.hip), so they compile in parallel
        groups = [masks] if c not in DERIV_CSPLIT else [tuple(m for m in masks if m & 7 == ctr) for ctr in (1, 2, 4)]
        sig = "(const GradArgs& a, const BoysTables& tab, const Task* t, int nt, int natom, int fp32, hipStream_t st)"
        for gi, grp in enumerate(groups):
            body = [head, '#include "../grad_reg.hpp"', "namespace m2 {"]
            for m in grp:
                body.append(gen_deriv_class(*c, mask=m))
            fn = "launch_grad_reg_%s" % tag if len(groups) == 1 else "launch_grad_reg_%s_%d" % (tag, gi)
            body.append("void %s%s {" % (fn, sig))
            for m in grp:
                body.append("  if (fp32) launch_reg<%d, %d, %d, %d, %d, float>(a, tab, t, nt, natom, st);" % (c + (m,)))
                body.append("  else launch_reg<%d, %d, %d, %d, %d, double>(a, tab, t, nt, natom, st);" % (c + (m,)))
            body.append("}")
            if len(groups) > 1 and gi == len(groups) - 1:
                for k in range(len(groups) - 1):
                    body.append("void launch_grad_reg_%s_%d%s;" % (tag, k, sig))
                body.append("void launch_grad_reg_%s%s {" % (tag, sig))
                for k in range(len(groups)):
                    body.append("  launch_grad_reg_%s_%d(a, tab, t, nt, natom, fp32, st);" % (tag, k))
                body.append("}")
            body += ["}  // namespace m2", ""]
            fname = "grad_reg_%s.hip" % tag if len(groups) == 1 else "grad_reg_%s_%d.hip" % (tag, gi)
            put(os.path.join(outdir, fname), "\n".join(body))
        if len(groups) > 1 and os.path.exists(os.path.join(outdir, "grad_reg_%s.hip" % tag)):
            os.remove(os.path.join(outdir, "grad_reg_%s.hip" % tag))
        names.append((c, tag))
    disp = [head, "#pragma once"]
    for c, tag in names:
        disp.append("void launch_grad_reg_%s(const GradArgs&, const BoysTables&, const Task*, int, int, int, hipStream_t);"
                    % tag)
    disp.append("// true when (la, lb, lc, ld) has register code (and was launched, in FP32 arithmetic when fp32)")
    disp.append("inline bool launch_grad_reg(int la, int lb, int lc, int ld, const GradArgs& a, const BoysTables& tab,"
                " const Task* t, int nt, int natom, int fp32, hipStream_t st) {")
    for c, tag in names:
        disp.append("  if (la == %d && lb == %d && lc == %d && ld == %d) {"
                    " launch_grad_reg_%s(a, tab, t, nt, natom, fp32, st); return true; }" % (c + (tag,)))
    disp += ["  return false;", "}", ""]
    put(os.path.join(outdir, "deriv_dispatch.inc"), "\n".join(disp))
    cls = [head, "#pragma once",
           "constexpr int DERIV_REG_N = %d;" % len(DERIV_REGISTER),
           "constexpr int DERIV_REG_L[%d][4] = {%s};" % (len(DERIV_REGISTER), ", ".join(
               "{%d, %d, %d, %d}" % c for c in DERIV_REGISTER)), ""]
    put(os.path.join(outdir, "deriv_classes.inc"), "\n".join(cls))


if __name__ == "__main__":
    if len(sys.argv) > 2 and sys.argv[1] == "--flop-model":
        write_flop_model(sys.argv[2])
    else:
        main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                "generated"))
