"""The particle-mesh Ewald oracle (charges and dipoles) and the Ewald reciprocal sum it
approximates, pinned against each other and against what the mathematics fixes.

* B-splines: partition of unity, derivatives summing to zero, M_4 against its closed form,
  and the derivatives against finite differences of the splines.
* Ewald reciprocal sum: U = 1/2 sum (q phi - mu . E) and, by central differences of U,
  phi = dU/dq, E = -dU/dmu, F = -dU/dr (a triclinic cell).
* SPME against the Ewald sum: the error falls with the order and the grid and reaches
  1e-13 at order 12 on a 64^3 grid.  The two computations share nothing but the cell
  (splines, FFT-free DFT and influence function against a plain sum over k).
"""
import numpy as np
import pytest

import oracle_lib as o


def system(n=24, seed=1):
    rng = np.random.default_rng(seed)
    L = 12.0
    cell = np.diag([L, 1.1 * L, 0.95 * L])
    cell[1, 0] = 1.0
    xyz = rng.uniform(0, 1, (n, 3)) @ cell
    q = rng.normal(size=n)
    q -= q.mean()
    mu = rng.normal(size=(n, 3)) * 0.3
    return cell, xyz, q, mu


@pytest.mark.parametrize("p", [4, 6, 8, 10, 12])
def test_bspline_identities(p):
    for w in (0.0, 0.137, 0.5, 0.93):
        a, d, d2 = o.bspline(p, w)
        assert abs(a.sum() - 1.0) < 1e-14
        assert abs(d.sum()) < 1e-14 and abs(d2.sum()) < 1e-13
        assert (a >= 0).all()
        h = 1e-6
        if h <= w <= 1.0 - h:
            ap, dp, _ = o.bspline(p, w + h)
            am, dm, _ = o.bspline(p, w - h)
            assert np.abs((ap - am) / (2 * h) - d).max() < 1e-8
            assert np.abs((dp - dm) / (2 * h) - d2).max() < 1e-7


def test_bspline_order4_closed_form():
    # M_4(x) = x^3/6 on [0,1], (-3x^3 + 12x^2 - 12x + 4)/6 on [1,2], symmetric about 2
    def m4(x):
        if x < 0 or x > 4:
            return 0.0
        if x > 2:
            x = 4 - x
        return x ** 3 / 6 if x <= 1 else (-3 * x ** 3 + 12 * x ** 2 - 12 * x + 4) / 6
    w = 0.3
    a, _, _ = o.bspline(4, w)
    assert np.allclose(a, [m4(w + t) for t in range(4)], rtol=0, atol=1e-15)


def test_ewald_recip_identities_and_derivatives():
    cell, xyz, q, mu = system()
    km, al = (10, 10, 10), 0.35
    ew = o.ewald_recip(cell, km, al, xyz, q, mu)
    assert abs(ew["U"] - 0.5 * (q @ ew["phi"] - (mu * ew["E"]).sum())) < 1e-12 * abs(ew["U"])
    h, i = 1e-5, 3
    U = lambda **kw: o.ewald_recip(cell, km, al, kw.get("xyz", xyz), kw.get("q", q), kw.get("mu", mu))["U"]
    for d in range(3):
        xp, xm = xyz.copy(), xyz.copy()
        xp[i, d] += h
        xm[i, d] -= h
        assert abs(-(U(xyz=xp) - U(xyz=xm)) / (2 * h) - ew["F"][i, d]) < 1e-7
        mp, mm = mu.copy(), mu.copy()
        mp[i, d] += h
        mm[i, d] -= h
        assert abs(-(U(mu=mp) - U(mu=mm)) / (2 * h) - ew["E"][i, d]) < 1e-7
    qp, qm = q.copy(), q.copy()
    qp[i] += h
    qm[i] -= h
    assert abs((U(q=qp) - U(q=qm)) / (2 * h) - ew["phi"][i]) < 1e-7


def test_spme_converges_to_ewald_sum():
    cell, xyz, q, mu = system()
    al = 0.35
    ew = o.ewald_recip(cell, (14, 14, 14), al, xyz, q, mu)
    errs = []
    for K, p in ((24, 4), (32, 6), (48, 8), (64, 12)):
        sp = o.spme_recip(cell, p, (K, K, K), al, xyz, q, mu)
        e = max(abs(sp["U"] - ew["U"]) / abs(ew["U"]),
                np.abs(sp["E"] - ew["E"]).max() / np.abs(ew["E"]).max(),
                np.abs(sp["F"] - ew["F"]).max() / np.abs(ew["F"]).max())
        errs.append(e)
    assert all(b < a for a, b in zip(errs, errs[1:]))
    assert errs[-1] < 1e-12


def test_spme_charges_only_and_dipoles_only_add_up():
    # U is quadratic in (q, mu): U(q, mu) = U(q, 0) + U(0, mu) + cross term, and the cross
    # term is sum q phi_mu (phi of the dipoles alone)
    cell, xyz, q, mu = system(n=12, seed=3)
    args = (cell, 8, (32, 32, 32), 0.35, xyz)
    both = o.spme_recip(*args, q, mu)
    qo = o.spme_recip(*args, q, None)
    mo = o.spme_recip(*args, np.zeros_like(q), mu)
    cross = q @ mo["phi"]
    assert abs(both["U"] - (qo["U"] + mo["U"] + cross)) < 1e-12 * abs(both["U"])
    assert np.abs(both["phi"] - (qo["phi"] + mo["phi"])).max() < 1e-12
