#!/usr/bin/env python3
"""CPU study: can the MB-pol polynomials run in FP32 within RMS dF <= 1e-5 kcal/mol/A?

Test infrastructure, run by hand (about a minute): ``python oracle/precision_study.py``.
It takes every dimer and trimer of the seeded 512-water box (8 168 and 34 927 clusters),
their polynomial variables from the oracle (ora_vars_2b / ora_vars_3b, i.e.
mbpol-2b-poly.cpp:415-445 and mbpol-3b-poly.cpp:170-206), and evaluates the polynomials
(monomial tables extracted from poly-2b-v6x.cpp / poly-3b-v2x.cpp) with their gradients
in three ways:

  fp64      every product and sum in FP64 (the reference's arithmetic; the error baseline);
  fp32      variables, coefficients, products and sums in FP32 (what mbgpu_set_precision
            does on the GPU, summation order aside);
  shifted   the polynomial re-centred at x0 = the mean variable vector of the box's
            clusters: P(x0 + d) = sum_b c'_b d^b with c' computed exactly in FP64; the
            constant and linear terms in FP64, every term of degree >= 2 in FP32.

The variable-gradient errors are mapped to atom forces through the chain rule (Jacobians of
the variables by central differences of the oracle's variable routines, and the switching
functions of mbpol-2b-poly.cpp:340-354 / mbpol-3b-poly.cpp:130-144), summed per atom over
all clusters, and reported as RMS and max force errors (kcal/mol/A) with the energy error
per molecule.  The per-atom heat-flux tensors W_a (the force-matrix contraction the heat
current needs) are accumulated the same way, and the relative heat-current error of the
MB-pol part is reported for the configuration's velocities.
"""
import ctypes
import importlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))
import oracle_lib as o  # noqa: E402

wb = importlib.import_module("dlpoly2-mbpol-heat_amd.waterbox")
DATA = os.path.join(REPO, "dlpoly2-mbpol-heat_amd", "data")


def load_poly(which):
    mons = [list(map(int, line.split())) for line in open(os.path.join(DATA, f"poly{which}_monomials.txt"))]
    a = np.loadtxt(os.path.join(DATA, f"poly{which}_coeffs.txt"))
    return [(a[m[0]], tuple(sorted(m[1:]))) for m in mons]


def merge(terms):
    out = {}
    for c, v in terms:
        out[v] = out.get(v, 0.0) + c
    return [(c, v) for v, c in out.items() if c != 0.0]


def shift(terms, x0):
    """Exact re-centring: sum_m c prod_k (x0_vk + d_vk) as monomials in d (FP64)."""
    out = []
    for c, v in terms:
        n = len(v)
        for mask in range(1 << n):
            sel = tuple(v[k] for k in range(n) if mask >> k & 1)
            rest = c
            for k in range(n):
                if not mask >> k & 1:
                    rest *= x0[v[k]]
            out.append((rest, sel))
    return merge(out)


class Poly:
    """Dense evaluation of a list of (coef, variable tuple) monomials for many clusters."""

    def __init__(self, terms, nvar):
        self.nvar = nvar
        self.groups = {}
        for c, v in terms:
            self.groups.setdefault(len(v), []).append((c, v))
        self.groups = {d: (np.array([c for c, _ in g]), np.array([v for _, v in g], dtype=np.int64).reshape(len(g), d))
                       for d, g in self.groups.items()}

    def eval(self, X, dt, min_degree=0, max_degree=9):
        """E and dE/dx for X (nc, nvar) in precision dt (products and sums)."""
        X = X.astype(dt)
        nc = X.shape[0]
        E = np.zeros(nc, dt)
        G = np.zeros((nc, self.nvar), dt)
        for d, (c, V) in self.groups.items():
            if d < min_degree or d > max_degree:
                continue
            c = c.astype(dt)
            if d == 0:
                E += dt(c.sum()) if dt is not np.float32 else np.float32(np.sum(c, dtype=np.float32))
                continue
            cols = [X[:, V[:, k]] for k in range(d)]              # (nc, nmon) each
            prod = np.ones_like(cols[0])
            for k in range(d):
                prod = prod * cols[k]
            E += prod @ c
            for k in range(d):
                others = np.ones_like(cols[0])
                for q in range(d):
                    if q != k:
                        others = others * cols[q]
                contrib = others * c                               # (nc, nmon)
                ind = np.zeros((V.shape[0], self.nvar), dt)
                ind[np.arange(V.shape[0]), V[:, k]] = 1
                G += contrib @ ind
        return E.astype(np.float64), G.astype(np.float64)


def switch(r, ri, rf):
    if r > rf:
        return 0.0
    if r > ri:
        return 0.5 * (1.0 + np.cos(np.pi * (r - ri) / (rf - ri)))
    return 1.0


def s2(w):
    return switch(np.linalg.norm(w[0:3] - w[9:12]), 4.5, 6.5)


def s3(w):
    d = [np.linalg.norm(w[9 * a:9 * a + 3] - w[9 * b:9 * b + 3]) for a, b in ((0, 1), (0, 2), (1, 2))]
    sw = [switch(x, 0.0, 4.5) for x in d]
    return sw[0] * sw[1] + sw[0] * sw[2] + sw[1] * sw[2]


def jacobian(fun, w, nvar, h=1e-6):
    J = np.zeros((nvar, w.size))
    for i in range(w.size):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        J[:, i] = (fun(wp) - fun(wm)) / (2 * h)
    return J


def grad_s(fun, w, h=1e-6):
    g = np.zeros(w.size)
    for i in range(w.size):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        g[i] = (fun(wp) - fun(wm)) / (2 * h)
    return g


def clusters(box):
    r = o.mbpol_forces(box, lists=True)
    L = box.L
    mon = box.xyz[:box.nreal].reshape(box.nw, 3, 3).copy()
    for w in range(box.nw):           # image the hydrogens next to their oxygen
        d = mon[w, 1:] - mon[w, 0]
        mon[w, 1:] = mon[w, 0] + d - L * np.round(d / L)

    def near(wi, ref):
        d = mon[wi, 0] - mon[ref, 0]
        return (mon[wi] - L * np.round(d / L)).reshape(-1)

    dim = [np.concatenate([near(i, j), mon[j].reshape(-1)]) for i, j in r["dimers"]]
    tri = [np.concatenate([near(i, j), mon[j].reshape(-1), near(k, j)]) for i, j, k in r["trimers"]]
    return r["dimers"], np.array(dim), r["trimers"], np.array(tri)


def study():
    t0 = time.time()
    box = wb.box_512()
    di, dw, ti, tw = clusters(box)
    print(f"{len(dw)} dimers, {len(tw)} trimers ({time.time() - t0:.0f} s)", flush=True)
    f2 = lambda w: o.vars_2b(w[:9], w[9:])
    f3 = lambda w: o.vars_3b(w[:9], w[9:18], w[18:])
    X2 = np.array([f2(w) for w in dw])
    X3 = np.array([f3(w) for w in tw])
    P2 = merge(load_poly("2b"))
    P3 = merge(load_poly("3b"))
    res = {}
    for name in ("fp64", "fp32", "shifted"):
        res[name] = {}
    for tag, P, X in (("2b", P2, X2), ("3b", P3, X3)):
        nv = X.shape[1]
        full = Poly(P, nv)
        E64, G64 = [], []
        E32, G32 = [], []
        x0 = X.mean(0)
        Psh = shift(P, x0)
        sh = Poly(Psh, nv)
        ES, GS = [], []
        for b in range(0, len(X), 2048):
            xb = X[b:b + 2048]
            e, g = full.eval(xb, np.float64)
            E64.append(e), G64.append(g)
            e, g = full.eval(xb, np.float32)
            E32.append(e), G32.append(g)
            d = xb - x0
            el, gl = sh.eval(d, np.float64, max_degree=1)
            eh, gh = sh.eval(d, np.float32, min_degree=2)
            ES.append(el + eh), GS.append(gl + gh)
        res["fp64"][tag] = (np.concatenate(E64), np.concatenate(G64))
        res["fp32"][tag] = (np.concatenate(E32), np.concatenate(G32))
        res["shifted"][tag] = (np.concatenate(ES), np.concatenate(GS))
        print(f"{tag}: {len(P)} monomials, {len(Psh)} after re-centring ({time.time() - t0:.0f} s)", flush=True)
    # chain rule to atoms
    print("Jacobians ...", flush=True)
    J2 = [jacobian(f2, w, 31) for w in dw]
    J3 = [jacobian(f3, w, 36) for w in tw]
    S2 = np.array([s2(w) for w in dw])
    S3 = np.array([s3(w) for w in tw])
    gS2 = [grad_s(s2, w) for w in dw]
    gS3 = [grad_s(s3, w) for w in tw]
    print(f"done ({time.time() - t0:.0f} s)", flush=True)
    nr = box.nreal
    vel = box.vel[:nr]
    out = {}
    for name in ("fp32", "shifted"):
        dF = np.zeros((nr, 3))
        dW = np.zeros((nr, 3, 3))
        dE = 0.0
        for tag, ids, W, J, S, gS, k in (("2b", di, dw, J2, S2, gS2, 6.0), ("3b", ti, tw, J3, S3, gS3, 9.0)):
            e64, g64 = res["fp64"][tag]
            e, g = res[name][tag]
            de, dg = e - e64, g - g64
            dE += float((S * de).sum())
            for c in range(len(W)):
                # dE/dr of the cluster: s J^T g + P grad s; its error from de, dg
                gr = S[c] * (J[c].T @ dg[c]) + de[c] * gS[c]
                nm = len(ids[c])
                Sc = W[c].reshape(nm * 3, 3).sum(0)
                for m in range(nm):
                    wat = ids[c][m]
                    for a in range(3):
                        f = -gr[9 * m + 3 * a:9 * m + 3 * a + 3]
                        dF[3 * wat + a] += f
                        dW[3 * wat + a] += np.outer(f / k, k * W[c][9 * m + 3 * a:9 * m + 3 * a + 3] - Sc)
        dJ = np.einsum("ia,iab->b", vel, dW)
        out[name] = dict(rms_dF=float(np.sqrt((dF ** 2).mean())), max_dF=float(np.abs(dF).max()),
                         dE_per_molecule=dE / box.nw, dJpot=np.abs(dJ).max())
    # the MB-pol part of J_pot in FP64, for scale
    r = o.mbpol_forces(box, lists=False)
    Jm = np.einsum("ia,iab->b", vel, r["heat_w"][:nr].reshape(nr, 3, 3) / 418.4)
    for name, v in out.items():
        v["rel_dJ_mbpol"] = float(v["dJpot"] / np.abs(Jm).max())
        print(f"{name:8s} RMS dF {v['rms_dF']:.3e}  max dF {v['max_dF']:.3e} kcal/mol/A   "
              f"dE/molecule {v['dE_per_molecule']:+.3e} kcal/mol   |dJ_pot(MB-pol)|/|J_pot(MB-pol)| "
              f"{v['rel_dJ_mbpol']:.2e}", flush=True)
    return out


if __name__ == "__main__":
    study()
