"""Generate galerkin_gen.hpp: the wall-normal Galerkin products for the z-kernel.

The reference tabulates the products of its wall-normal bases by hand
(program/modes.M4.f90:14-81, program/modes.M6.f90:14-79).  Here they are
derived from the product-to-sum identities with exact rational arithmetic and
emitted in the streaming form the fused z-kernel needs: the 3*K0-1 base values
v of a point are held in registers, then each x-derivative (vx) and
z-derivative (vz) component is folded into the accumulators as it arrives,
and finally the terms that involve only v (the d/dy products) are added.

Bases (beta = pi/2):
  u/w family h_k, k = 0..K0-1 : 1, sin(b y), cos(2 b y), sin(3 b y), ...
  v family   g_k, k = 1..K0-1 : cos(b y), sin(2 b y), cos(3 b y), ...
Projection onto h_k or g_k is the coefficient of that function (the mean for
h_0), truncated at K0-1.

  ans_u = P_u(u ux) + P_u(v du/dy) + P_u(w uz)
  ans_v = P_v(u vx) + P_v(v dv/dy) + P_v(w vz)
  ans_w = P_u(u wx) + P_u(v dw/dy) + P_u(w wz)

Flux (conservative) form, flux_gen.hpp.  The Galerkin velocity field is
divergence-free as a trigonometric polynomial (var_mpt2spec builds u, v, w
from psi and phi, program/variables.f90:390-434, so d/dx u + dv/dy + d/dz w
= 0 term by term; the laminar profile depends on y alone), hence
(u.grad)u = div(u u).  With the coefficient-of-basis projections above being
orthogonal projections on one period in y, d/dy moves onto the basis exactly:

  P_h,i(dF/dy) =  ak1(i) P_g,i(F),   P_g,i(dF/dy) = -ak1(i) P_h,i(F),
  ak1(i) = (-1)^i i beta  (h_i' = -ak1(i) g_i, g_i' = ak1(i) h_i),

so the three projected momentum components need only six projected
product fields (rows) of the base values, with d/dx and d/dz applied in
spectral space:

  N_u,i = ddx uu_i + ak1(i) uv_i + ddz uw_i         (uv_0 term absent)
  N_v,i = ddx uv_i - ak1(i) vv_i + ddz vw_i
  N_w,i = ddx uw_i + ak1(i) vw_i + ddz ww_i         (vw_0 term absent)

  uu_i = P_h,i(u u), uw_i = P_h,i(u w), ww_i = P_h,i(w w)   (i < K0)
  uv_i = P_g,i(u v), vw_i = P_g,i(v w), vv_i = P_h,i(v v)   (1 <= i < K0)

The z kernel of this form transforms the rows forward and emits, per spec
slot s, a line B_s to be differentiated in x (uu, uv, uw) and a line A_s
(the d/dy and d/dz terms); the x kernel forms N_s = i alpha m B_s + A_s.

Run:  python mwf_amd/csrc/gen_galerkin.py   (rewrites galerkin_gen.hpp and flux_gen.hpp)
"""
import os
from fractions import Fraction as F

C, S = "C", "S"


def utype(k):
    return C if k % 2 == 0 else S


def vtype(k):
    return C if k % 2 == 1 else S


def prod(ta, fa, tb, fb):
    """(type, freq, coef) terms of trig(ta, fa) * trig(tb, fb)."""
    h = F(1, 2)
    if ta == C and tb == C:
        return [(C, fa + fb, h), (C, abs(fa - fb), h)]
    if ta == S and tb == S:
        return [(C, abs(fa - fb), h), (C, fa + fb, -h)]
    fs, fc = (fa, fb) if ta == S else (fb, fa)
    out = [(S, fs + fc, h)]
    d = fs - fc
    if d != 0:
        out.append((S, abs(d), h if d > 0 else -h))
    return out


def proj(fam, i, ta, fa, tb, fb):
    ti = utype(i) if fam == "u" else vtype(i)
    return sum((c for t, f, c in prod(ta, fa, tb, fb) if f == i and t == ti), F(0))


def deriv(t, f):
    """d/dy of trig(t, f) in units of beta: (type, freq, factor)."""
    return (S, f, F(-f)) if t == C else (C, f, F(f))


def tables(k0):
    U = range(k0)
    V = range(1, k0)
    uw = {(i, j, l): proj("u", i, utype(j), j, utype(l), l) for i in U for j in U for l in U}
    uv = {(i, j, l): proj("v", i, utype(j), j, vtype(l), l) for i in V for j in U for l in V}
    uyv = {}
    for i in U:
        for j in range(1, k0):
            td, fd, fac = deriv(utype(j), j)
            for l in V:
                uyv[(i, j, l)] = fac * proj("u", i, vtype(l), l, td, fd)
    vvy = {}
    for i in V:
        for j in V:
            for l in V:
                td, fd, fac = deriv(vtype(l), l)
                vvy[(i, j, l)] = fac * proj("v", i, vtype(j), j, td, fd)
    return uw, uv, uyv, vvy


def slots(k0):
    """spec slot (0-based) of u_k, v_k, w_k."""
    return (lambda k: k), (lambda k: k0 - 1 + k), (lambda k: 2 * k0 - 1 + k)


def lit(c):
    """C++ literal for a rational coefficient."""
    if c.denominator == 1:
        return f"{c.numerator}.0"
    return f"({c.numerator}.0/{c.denominator}.0)"


def stream_terms(k0):
    """For derivative family 'x' (partner u) and 'z' (partner w): per derivative
    slot s, list of (out_slot, base_slot, coef)."""
    uw, uv, uyv, vvy = tables(k0)
    su, sv, sw = slots(k0)
    out = {"x": {}, "z": {}}
    for fam, partner in (("x", su), ("z", sw)):
        for s in range(3 * k0 - 1):
            terms = []
            # which velocity component the derivative belongs to
            if s < k0:              # d u_l
                l, eq, tab, outslot = s, "u", uw, su
            elif s < 2 * k0 - 1:    # d v_l
                l, eq, tab, outslot = s - (k0 - 1), "v", uv, sv
            else:                   # d w_l
                l, eq, tab, outslot = s - (2 * k0 - 1), "w", uw, sw
            rng = range(1, k0) if eq == "v" else range(k0)
            for i in rng:
                for j in range(k0):
                    c = tab.get((i, j, l), F(0))
                    if c != 0:
                        terms.append((outslot(i), partner(j), c))
            out[fam][s] = terms
    # terms with only base values: P_u(v du/dy), P_v(v dv/dy), P_u(v dw/dy)
    y = []
    for i in range(k0):
        for j in range(1, k0):
            for l in range(1, k0):
                c = uyv[(i, j, l)]
                if c != 0:
                    y.append((su(i), su(j), sv(l), c))
                    y.append((sw(i), sw(j), sv(l), c))
    for i in range(1, k0):
        for j in range(1, k0):
            for l in range(1, k0):
                c = vvy[(i, j, l)]
                if c != 0:
                    y.append((sv(i), sv(j), sv(l), c))
    y.sort()
    return out, y


def emit(k0s=(4, 6)):
    lines = [
        "// GENERATED by gen_galerkin.py -- do not edit.",
        "// Wall-normal Galerkin products (the udotgradu closures of the M4/M6",
        "// bases, program/modes.M4.f90:14-81, program/modes.M6.f90:14-79) in the",
        "// streaming form used by the fused z-kernel.",
        "#pragma once",
        "",
        "template <int K0> struct Gal;",
        "",
    ]
    for k0 in k0s:
        st, y = stream_terms(k0)
        kk = 3 * k0 - 1
        lines.append(f"template <> struct Gal<{k0}> {{")
        lines.append(f"  static constexpr int KK = {kk};")
        for fam in ("x", "z"):
            lines.append(f"  // fold derivative component d{fam} of spec slot S into a[]")
            lines.append("  template <int S, typename R>")
            lines.append(f"  __host__ __device__ static __forceinline__ void d{fam}(const R* v, R d, R* a) {{")
            first = True
            for s in range(kk):
                kw = "if constexpr" if first else "else if constexpr"
                first = False
                body = " ".join(
                    f"a[{o}] += " + ("" if c == 1 else ("-" if c == -1 else f"(R){lit(c)} * ")) + f"(v[{b}] * d);"
                    for o, b, c in st[fam][s])
                lines.append(f"    {kw} (S == {s}) {{ {body} }}")
            lines.append("  }")
        lines.append("  // terms of v only: P_u(v du/dy), P_v(v dv/dy), P_u(v dw/dy)")
        lines.append("  template <typename R>")
        lines.append("  __host__ __device__ static __forceinline__ void yterms(const R* v, R* a, R beta) {")
        for o, j, l, c in y:
            lines.append(f"    a[{o}] += ((R){lit(c)} * beta) * (v[{j}] * v[{l}]);")
        lines.append("  }")
        lines.append("};")
        lines.append("")
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------- flux form

def flux_rows(k0):
    """Product rows of the flux form: list of (name, i, terms) with terms
    [(slot a, slot b, coef)] of the row's value sum coef * v[a] * v[b]
    (slot a <= slot b: symmetric products combined)."""
    su, sv, sw = slots(k0)
    rows = []

    def add(name, i, fam, ta, sa, ra, tb, sb, rb):
        acc = {}
        for j in ra:
            for l in rb:
                c = proj(fam, i, ta(j), j, tb(l), l)
                if c != 0:
                    a, b = sorted((sa(j), sb(l)))
                    acc[(a, b)] = acc.get((a, b), F(0)) + c
        rows.append((name, i, sorted((a, b, c) for (a, b), c in acc.items() if c != 0)))

    U, V = range(k0), range(1, k0)
    for i in range(k0):
        add("uu", i, "u", utype, su, U, utype, su, U)
        add("uw", i, "u", utype, su, U, utype, sw, U)
        add("ww", i, "u", utype, sw, U, utype, sw, U)
    for i in range(1, k0):
        add("uv", i, "v", utype, su, U, vtype, sv, V)
        add("vw", i, "v", utype, sw, U, vtype, sv, V)
        add("vv", i, "u", vtype, sv, V, vtype, sv, V)
    return rows


def flux_order(k0):
    """Forward order of the rows (indices into flux_rows): i = 0 first, then per
    i the rows of one wall-normal index together, so that partial lines live
    briefly."""
    rows = flux_rows(k0)
    idx = {(n, i): r for r, (n, i, _) in enumerate(rows)}
    order = [idx[("uu", 0)], idx[("uw", 0)], idx[("ww", 0)]]
    for i in range(1, k0):
        order += [idx[(n, i)] for n in ("uv", "uw", "vw", "ww", "uu", "vv")]
    return order


def ak1_sign(i):
    """ak1(i) / beta = (-1)^i i (program/variables.f90:528-529)."""
    return (-1) ** i * i


def flux_lines(k0):
    """For each output line L < 2 KK (B lines at L = slot, A lines at KK + slot):
    [(row, kind, s)]: kind 0 adds the row, kind 1 adds s * beta * row, kind 2
    adds i gamma n * row."""
    rows = flux_rows(k0)
    idx = {(n, i): r for r, (n, i, _) in enumerate(rows)}
    su, sv, sw = slots(k0)
    kk = 3 * k0 - 1
    lines = [[] for _ in range(2 * kk)]
    for i in range(k0):
        lines[su(i)].append((idx[("uu", i)], 0, 1))
        lines[sw(i)].append((idx[("uw", i)], 0, 1))
        if i:
            lines[kk + su(i)].append((idx[("uv", i)], 1, ak1_sign(i)))
        lines[kk + su(i)].append((idx[("uw", i)], 2, 1))
        if i:
            lines[kk + sw(i)].append((idx[("vw", i)], 1, ak1_sign(i)))
        lines[kk + sw(i)].append((idx[("ww", i)], 2, 1))
    for i in range(1, k0):
        lines[sv(i)].append((idx[("uv", i)], 0, 1))
        lines[kk + sv(i)].append((idx[("vv", i)], 1, -ak1_sign(i)))
        lines[kk + sv(i)].append((idx[("vw", i)], 2, 1))
    return lines


def emit_flux(k0s=(4, 6)):
    out = [
        "// GENERATED by gen_galerkin.py -- do not edit.",
        "// Flux (conservative) form of the wall-normal Galerkin products: the",
        "// projected product rows uu, uw, ww, uv, vw, vv of the base values and",
        "// how the z kernel assembles them into the B (d/dx) and A (d/dy + d/dz)",
        "// lines; see gen_galerkin.py for the derivation.",
        "#pragma once",
        "",
        "template <int K0> struct Flux;",
        "",
    ]
    for k0 in k0s:
        rows, order, lines = flux_rows(k0), flux_order(k0), flux_lines(k0)
        kk, nr = 3 * k0 - 1, len(rows)
        pos = {r: q for q, r in enumerate(order)}
        uses = [[] for _ in range(nr)]
        for L, terms in enumerate(lines):
            for r, kind, sg in terms:
                uses[r].append((L, kind, sg))
        last = [max(pos[r] for r, _, _ in lines[L]) for L in range(2 * kk)]
        first = [min(pos[r] for r, _, _ in lines[L]) for L in range(2 * kk)]
        arr = lambda v: "{" + ", ".join(str(x) for x in v) + "}"
        out.append(f"template <> struct Flux<{k0}> {{")
        out.append(f"  static constexpr int KK = {kk};")
        out.append(f"  static constexpr int NR = {nr};  // product rows")
        out.append("  // row at forward position q (two rows per complex sequence); -1 = padding")
        out.append(f"  __host__ __device__ static constexpr int order(int q) {{ constexpr int o[] = {arr(order)}; "
                   f"return q < NR ? o[q] : -1; }}")
        out.append("  // uses of row r: line, kind (0: +row, 1: + s beta row, 2: + i gamma n row), s")
        out.append(f"  __host__ __device__ static constexpr int nuse(int r) {{ constexpr int o[] = {arr(len(u) for u in uses)}; return o[r]; }}")
        for nm, j in (("use_line", 0), ("use_kind", 1), ("use_sign", 2)):
            flat = [u[k][j] if k < len(u) else 0 for u in uses for k in range(2)]
            out.append(f"  __host__ __device__ static constexpr int {nm}(int r, int u) {{ constexpr int o[] = {arr(flat)}; "
                       f"return o[2 * r + u]; }}")
        out.append("  // forward position of the last row of line L (the line is complete after it)")
        out.append(f"  __host__ __device__ static constexpr int last(int L) {{ constexpr int o[] = {arr(last)}; return o[L]; }}")
        out.append("  // forward position of the first row of line L (the line starts there)")
        out.append(f"  __host__ __device__ static constexpr int first(int L) {{ constexpr int o[] = {arr(first)}; return o[L]; }}")
        out.append("  // value of product row ROW from the base values v[KK] of one point")
        out.append("  template <int ROW, typename R>")
        out.append("  __host__ __device__ static __forceinline__ R row(const R* v) {")
        for r, (name, i, terms) in enumerate(rows):
            kw = "if constexpr" if r == 0 else "else if constexpr"
            expr = " + ".join(("" if c == 1 else f"(R){lit(c)} * ") + f"(v[{a}] * v[{b}])" for a, b, c in terms)
            out.append(f"    {kw} (ROW == {r}) {{ return {expr or '(R)0'}; }}  // {name}_{i}")
        out.append("    else { return (R)0; }")
        out.append("  }")
        out.append("};")
        out.append("")
    return "\n".join(out) + "\n"


def _write(path, text):
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "galerkin_gen.hpp")
    _write(path, emit())
    _write(os.path.join(here, "flux_gen.hpp"), emit_flux())
    return path


if __name__ == "__main__":
    print(main())
