#!/usr/bin/env python3
"""Generate the straight-line HIP code that evaluates the MB-pol polynomials.

Input: the monomial tables in ../data (one line per monomial, "basis v1 [v2 [v3 [v4]]]";
every basis polynomial of poly_2b_v6x / poly_3b_v2x is a sum of distinct monomials with
unit coefficient, so E(x) = sum over monomials of a[basis] * prod x).

Evaluation scheme (ours; the reference uses Maple-generated Horner code):
  * monomials are stored in a trie whose path runs from the most frequently used
    variable to the least; the value of a node is V_N = c_N + sum_children x_j V_{Nj}
    (Horner over the trie) and E = V_root;
  * the gradient is the reverse sweep of the same trie: along an edge N -> Nj,
    dE/dx_j += P_N * V_{Nj}, where P_N is the product of the variables on the path to N.
    A depth-first walk carries P down and V up, so only O(depth) temporaries are live
    beyond the variables and the gradient accumulators: the code runs from registers.
  * the top-level subtrees are packed into ``nslices`` slices of balanced operation count.
    Each slice is one device function producing a partial (E, dE/dx); the
    slices of one cluster run in different wavefronts and their partials are summed
    afterwards.  This spreads one cluster's polynomial over several wavefronts so that a
    few thousand clusters fill the 256 CUs.

Output: pip{2,3}b_gen.inc with
    template <typename T> __device__ void pip{N}b_slice_<s>(
        const T* __restrict__ xv, T* __restrict__ part, int c, int nc,
        const T* __restrict__ C);
  xv:   variables, SoA [nvars][nc]   (cluster c reads xv[v*nc + c])
  part: partials, SoA [nslices][nvars+1][nc] (slot 0 = E, slot 1+v = dE/dx_v)
  C:    the coefficients as the code addresses them (class Stream: by default the fitted
        coefficient vector a[] itself); pip{N}b_stream_basis[] maps every slot of C to
        its basis index in a[].
and a dispatcher pip{N}b_slice(s, ...).
"""
import collections
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(HERE), "data")


def load(which):
    mono = []
    with open(os.path.join(DATA, f"poly{which}b_monomials.txt")) as f:
        for line in f:
            t = [int(s) for s in line.split()]
            if t:
                mono.append((t[0], tuple(t[1:])))
    return mono


class Node:
    __slots__ = ("coef", "ch")

    def __init__(self):
        self.coef = None
        self.ch = {}


def build_trie(mono):
    freq = collections.Counter(v for _, vs in mono for v in vs)
    order = [v for v, _ in sorted(freq.items(), key=lambda kv: (-kv[1], kv[0]))]
    rank = {v: i for i, v in enumerate(order)}      # 0 = top of the path
    root = Node()
    for b, vs in mono:
        n = root
        for v in sorted(vs, key=lambda v: rank[v]):
            n = n.ch.setdefault(v, Node())
        if n.coef is not None:
            raise SystemExit("duplicate monomial")
        n.coef = b
    return root


def cost(n):
    c = 0
    for ch in n.ch.values():
        c += (3 + cost(ch)) if ch.ch else 2
    return c


class Stream:
    """Coefficient addressing of the generated code.

    "basis" (default): the code reads C[basis index] from the fitted coefficient vector
    itself (1163 doubles for the 3B PIP, 9 KB), which the scalar data cache of a CU holds
    whole, however many wavefronts run the slice at different points of its code.
    "stream": every read gets its own slot in consumption order (29 000 slots, 230 KB for
    the 3B PIP), read once and in sequence."""

    def __init__(self, mode, nbasis, grouped=False):
        self.mode = mode
        self.grouped = grouped and mode == "basis"
        self.basis = list(range(nbasis)) if mode == "basis" else []

    def take(self, b):
        if self.mode == "basis":
            return b
        self.basis.append(b)
        return len(self.basis) - 1

    def ref(self, k):
        # in "basis" mode a coefficient is read at every use through an opaque copy of the
        # base pointer (pip_coef), so that the compiler neither merges the reads of one
        # coefficient nor keeps it live between its uses
        return f"pip_coef(C, {k})" if self.mode == "basis" else f"C[{k}]"


def emit_edge(out, st, node, var, P, V, depth, ind, K=None):
    """Edge from a parent with prefix product P and value accumulator V to `node` via x[var].
    K: name of the enclosing block's coefficient pointer (grouped reads), or None."""
    if not node.ch:
        k = st.take(node.coef)
        a = f"{K}[{k}]" if K else st.ref(k)
        out.append(f"{ind}{{ const T a = {a}; {V} = fma(a, x[{var}], {V}); "
                   f"g[{var}] = fma(a, {P}, g[{var}]); }}")
        return
    p, v = f"p{depth}", f"v{depth}"
    out.append(ind + "{")
    i2 = ind + "  "
    Kn = None
    if st.grouped:
        Kn = f"K{depth}"
        out.append(f"{i2}const pip_kptr<T> {Kn} = pip_base(C);")
    out.append(f"{i2}const T {p} = {P} * x[{var}];")
    if node.coef is not None:
        k = st.take(node.coef)
        out.append(f"{i2}T {v} = " + (f"{Kn}[{k}];" if Kn else f"{st.ref(k)};"))
    else:
        out.append(f"{i2}T {v} = T(0);")
    for cv in sorted(node.ch):
        emit_edge(out, st, node.ch[cv], cv, p, v, depth + 1, i2, Kn)
    out.append(f"{i2}{V} = fma(x[{var}], {v}, {V});")
    out.append(f"{i2}g[{var}] = fma({P}, {v}, g[{var}]);")
    out.append(ind + "}")


def make_units(root):
    """Work units: (top variable k, child variable j or None for k's own coefficient)."""
    units = []
    for k in sorted(root.ch):
        nk = root.ch[k]
        if nk.coef is not None:
            units.append((k, None, 2))
        for j in sorted(nk.ch):
            nj = nk.ch[j]
            units.append((k, j, (3 + cost(nj)) if nj.ch else 2))
    return units


def partition(units, nslices):
    """Greedy longest-first packing into nslices bins of balanced cost."""
    bins = [[] for _ in range(nslices)]
    load = [0] * nslices
    for u in sorted(units, key=lambda u: -u[2]):
        s = min(range(nslices), key=lambda i: load[i])
        bins[s].append(u)
        load[s] += u[2]
    return bins, load


def generate(which, nslices):
    mono = load(which)
    nv = 31 if which == 2 else 36
    root = build_trie(mono)
    units = make_units(root)
    bins, loads = partition(units, nslices)
    with open(os.path.join(DATA, f"poly{which}b_coeffs.txt")) as f:
        coef = [float(t) for t in f.read().split()]
    st = Stream(os.environ.get("MBGPU_PIP_COEF", "basis"), len(coef), os.environ.get("MBGPU_PIP_GROUP", "1") == "1")
    lines = [f"// generated by gen_pip.py from data/poly{which}b_monomials.txt -- do not edit",
             "#ifndef PIP_COEF_DEFINED",
             "#define PIP_COEF_DEFINED",
             "template <typename T>",
             "using pip_kptr = const __attribute__((address_space(4))) T*;",
             "template <typename T>",
             "__device__ __forceinline__ pip_kptr<T> pip_base(const T* C)",
             "{",
             "  pip_kptr<T> K = (pip_kptr<T>)C;",
             "  asm volatile(\"\" : \"+s\"(K));",
             "  return K;",
             "}",
             "template <typename T>",
             "__device__ __forceinline__ T pip_coef(const T* C, int i)",
             "{",
             "  // constant address space: a uniform read of read-only data (scalar load)",
             "  const __attribute__((address_space(4))) T* K = (const __attribute__((address_space(4))) T*)C;",
             "  asm volatile(\"\" : \"+s\"(K));",
             "  return K[i];",
             "}",
             "#endif",
             f"// {len(mono)} monomials, trie cost {cost(root)} fma/mul, {nslices} slices",
             f"#define PIP{which}B_NVARS {nv}",
             f"#define PIP{which}B_NSLICES {nslices}",
             ""]
    for s, units_s in enumerate(bins):
        lines.append(f"// slice {s}: {loads[s]} operations")
        lines.append("template <typename T>")
        lines.append(f"__device__ __forceinline__ void pip{which}b_slice_{s}(const T* __restrict__ xv, "
                     "T* __restrict__ part, int c, int nc, const T* __restrict__ C)")
        lines.append("{")
        lines.append(f"  T x[{nv}], g[{nv}];")
        lines.append(f"  for (int v = 0; v < {nv}; ++v) {{ x[v] = xv[v * nc + c]; g[v] = T(0); }}")
        lines.append("  T e = T(0);")
        byk = collections.OrderedDict()
        for k, j, _ in sorted(units_s, key=lambda u: (u[0], -1 if u[1] is None else u[1])):
            byk.setdefault(k, []).append(j)
        for k, js in byk.items():
            nk = root.ch[k]
            lines.append("  {")
            own = None in js
            K1 = None
            if st.grouped:
                K1 = "K1"
                lines.append("    const pip_kptr<T> K1 = pip_base(C);")
            if own:
                kk = st.take(nk.coef)
                lines.append(f"    T v1 = " + (f"K1[{kk}];" if K1 else f"{st.ref(kk)};"))
            else:
                lines.append("    T v1 = T(0);")
            for j in js:
                if j is None:
                    continue
                emit_edge(lines, st, nk.ch[j], j, f"x[{k}]", "v1", 2, "    ", K1)
            lines.append(f"    e = fma(x[{k}], v1, e);")
            lines.append(f"    g[{k}] += v1;")
            lines.append("  }")
        lines.append(f"  T* __restrict__ o = part + (size_t){s} * {nv + 1} * nc + c;")
        lines.append("  o[0] = e;")
        lines.append(f"  for (int v = 0; v < {nv}; ++v) o[(size_t)(v + 1) * nc] = g[v];")
        lines.append("}")
        lines.append("")
    lines.append("template <typename T>")
    lines.append(f"__device__ __forceinline__ void pip{which}b_slice(int s, const T* __restrict__ xv, "
                 "T* __restrict__ part, int c, int nc, const T* __restrict__ C)")
    lines.append("{")
    lines.append("  switch (s) {")
    for s in range(nslices):
        lines.append(f"  case {s}: pip{which}b_slice_{s}<T>(xv, part, c, nc, C); break;")
    lines.append("  default: break;")
    lines.append("  }")
    lines.append("}")
    lines.append("")
    lines.append(f"#define PIP{which}B_STREAM_LEN {len(st.basis)}")
    lines.append(f"// basis index of every coefficient-stream slot, in consumption order")
    lines.append(f"static const short pip{which}b_stream_basis[{len(st.basis)}] = {{")
    for i in range(0, len(st.basis), 16):
        lines.append("  " + ", ".join(str(b) for b in st.basis[i:i + 16]) + ",")
    lines.append("};")
    lines.append(f"// the coefficient stream itself: a[pip{which}b_stream_basis[i]]")
    lines.append(f"static const double pip{which}b_stream_coef[{len(st.basis)}] = {{")
    for i in range(0, len(st.basis), 4):
        lines.append("  " + ", ".join(repr(coef[b]) for b in st.basis[i:i + 4]) + ",")
    lines.append("};")
    return "\n".join(lines) + "\n", cost(root)


def main():
    outdir = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "gen")
    ns2 = int(os.environ.get("MBGPU_SLICES_2B", "8"))
    ns3 = int(os.environ.get("MBGPU_SLICES_3B", "8"))
    os.makedirs(outdir, exist_ok=True)
    for which, ns in ((2, ns2), (3, ns3)):
        text, c = generate(which, ns)
        path = os.path.join(outdir, f"pip{which}b_gen.inc")
        old = open(path).read() if os.path.exists(path) else None
        if old != text:
            with open(path, "w") as f:
                f.write(text)
        print(f"pip{which}b: trie cost {c} ops in {ns} slices -> {path}")


if __name__ == "__main__":
    main()
