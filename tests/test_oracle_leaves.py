"""The oracle's MB-pol leaves against the reference's own outputs and against invariants.

Pins for oracle/pip_oracle.c:
  * golden vectors written by the reference's compiled C++ leaves (tests/golden);
  * the raw polynomials with random coefficients (pins the monomial tables alone);
  * permutation symmetry, translation and rotation invariance, finite differences,
    cutoff behaviour.
Tolerances are scaled by S = sum over monomials |a_b * monomial| at the cluster: the
polynomials cancel by a factor S/|E| ~ 1e6 on liquid-water clusters, so FP64 evaluations
in different orders agree to ~1e-17 * S, not to 1e-16 * |E|.  A dropped or misplaced
monomial shifts E by ~1e-6 * S or more.
"""
import numpy as np
import pytest

import oracle_lib as o

TOL = 1e-15


def test_raw_polynomials_match_reference(golden, pkg):
    lib = o.oracle()
    for tag, which in (("q", 2), ("p", 3)):
        a = golden[f"{tag}_a"]
        mono = [list(map(int, line.split())) for line in
                open(f"{o.DATA}/poly{which}b_monomials.txt")]
        basis = np.array([m[0] for m in mono])
        for x, E, g in zip(golden[f"{tag}_x"], golden[f"{tag}_E"], golden[f"{tag}_g"]):
            gg = np.zeros(x.size)
            e = lib.ora_poly_eval(which, o.P(a), o.P(x.copy()), o.P(gg))
            S = sum(abs(a[m[0]] * np.prod(x[m[1:]])) for m in mono)
            assert abs(e - E) <= TOL * S
            assert np.abs(gg - g).max() <= TOL * S
        assert basis.max() == a.size - 1


def test_leaf_2b_golden(golden):
    for w, E, g, S in zip(golden["d_w"], golden["d_E"], golden["d_g"], golden["d_S"]):
        w1, w2 = w[:9], w[9:]
        if np.linalg.norm(w1[:3] - w2[:3]) > 6.5:
            e, gg = o.leaf_2b(w1, w2)
            assert e == 0.0 == E
            continue
        e, gg = o.leaf_2b(w1, w2)
        assert abs(e - E) <= TOL * S, (e, E, S)
        assert np.abs(gg - g).max() <= TOL * S


def test_leaf_3b_golden(golden):
    for w, E, g, S in zip(golden["t_w"], golden["t_E"], golden["t_g"], golden["t_S"]):
        e, gg = o.leaf_3b(w[:9], w[9:18], w[18:])
        assert abs(e - E) <= TOL * S, (e, E, S)
        assert np.abs(gg - g).max() <= TOL * S


def _swap_h(w):
    w = w.copy()
    w[3:6], w[6:9] = w[6:9].copy(), w[3:6].copy()
    return w


def test_2b_symmetries(golden):
    for w, S in zip(golden["d_w"][:40], golden["d_S"][:40]):
        w1, w2 = w[:9], w[9:]
        e, g = o.leaf_2b(w1, w2)
        # monomer exchange
        e2, g2 = o.leaf_2b(w2, w1)
        assert abs(e2 - e) <= 1e-14 * S
        assert np.abs(np.concatenate([g2[9:], g2[:9]]) - g).max() <= 1e-13 * S
        # H1 <-> H2 inside either monomer
        e3, _ = o.leaf_2b(_swap_h(w1), w2)
        e4, _ = o.leaf_2b(w1, _swap_h(w2))
        assert abs(e3 - e) <= 1e-14 * S and abs(e4 - e) <= 1e-14 * S


def test_3b_symmetries(golden):
    import itertools
    for w, S in zip(golden["t_w"][:20], golden["t_S"][:20]):
        ws = [w[:9], w[9:18], w[18:]]
        e, g = o.leaf_3b(*ws)
        for perm in itertools.permutations(range(3)):
            ep, gp = o.leaf_3b(*[ws[p] for p in perm])
            assert abs(ep - e) <= 1e-14 * S
            back = np.zeros(27)
            for slot, p in enumerate(perm):
                back[9 * p:9 * p + 9] = gp[9 * slot:9 * slot + 9]
            assert np.abs(back - g).max() <= 1e-13 * S
        e5, _ = o.leaf_3b(_swap_h(ws[0]), ws[1], _swap_h(ws[2]))
        assert abs(e5 - e) <= 1e-14 * S


def _rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


@pytest.mark.parametrize("which", [2, 3])
def test_translation_rotation_invariance(golden, which):
    rng = np.random.default_rng(5)
    key = "d" if which == 2 else "t"
    leaf = o.leaf_2b if which == 2 else o.leaf_3b
    nat = 6 if which == 2 else 9
    for w, S in zip(golden[f"{key}_w"][:20], golden[f"{key}_S"][:20]):
        ws = [w[9 * m:9 * m + 9] for m in range(nat // 3)]
        e, g = leaf(*ws)
        gs = g.reshape(nat, 3)
        x = w.reshape(nat, 3)
        # zero net force and torque
        assert np.abs(gs.sum(0)).max() <= 1e-13 * S
        assert np.abs(np.cross(x - x.mean(0), gs).sum(0)).max() <= 1e-12 * S
        R = _rot(rng)
        shift = rng.uniform(-5, 5, 3)
        xr = (x - x[0]) @ R.T + x[0] + shift
        er, gr = leaf(*[xr[3 * m:3 * m + 3].ravel() for m in range(nat // 3)])
        assert abs(er - e) <= 1e-14 * S
        assert np.abs(gr.reshape(nat, 3) - gs @ R.T).max() <= 1e-13 * S


@pytest.mark.parametrize("which", [2, 3])
def test_finite_differences(golden, which):
    key = "d" if which == 2 else "t"
    leaf = o.leaf_2b if which == 2 else o.leaf_3b
    nat = 6 if which == 2 else 9
    h = 1e-5
    for w in golden[f"{key}_w"][:6]:
        ws = [w[9 * m:9 * m + 9].copy() for m in range(nat // 3)]
        if which == 2 and np.linalg.norm(ws[0][:3] - ws[1][:3]) > 6.4:
            continue
        _, g = leaf(*ws)
        for k in range(3 * nat):
            wp = [v.copy() for v in ws]
            wm = [v.copy() for v in ws]
            wp[k // 9][k % 9] += h
            wm[k // 9][k % 9] -= h
            num = (leaf(*wp)[0] - leaf(*wm)[0]) / (2 * h)
            assert abs(num - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, num, g[k])


def test_cutoff_behaviour(golden):
    w = golden["d_w"][0]
    w1, w2 = w[:9].copy(), w[9:].copy()
    w1 = w1 - np.tile(w1[:3], 3) + np.tile(w2[:3] + np.array([6.5001, 0, 0]), 3)
    g1, g2 = np.full(9, 7.0), np.full(9, 7.0)
    e = o.oracle().ora_mbpol_2b(o.P(w1), o.P(w2), o.P(g1), o.P(g2))
    assert e == 0.0 and np.all(g1 == 7.0) and np.all(g2 == 7.0)
    # trimer with every O-O edge beyond 4.5 A: zero energy and gradient
    t = golden["t_w"][0]
    ws = [t[:9].copy(), t[9:18].copy(), t[18:].copy()]
    for m, off in ((0, np.array([0.0, 0, 0])), (1, np.array([5.0, 0, 0])), (2, np.array([0, 5.0, 0]))):
        ws[m] = ws[m] - np.tile(ws[m][:3], 3) + np.tile(off, 3)
    e, g = o.leaf_3b(*ws)
    assert e == 0.0 and not np.any(g)
