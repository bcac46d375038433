"""The ring-polymer oracle (normal modes and springs, centroid_matrix.f and
centroid_spring_force.f) against what the mathematics of a ring fixes.

* Default ordering: nmode_to_cart . cart_to_nmode = I, and the normal modes diagonalise the
  ring's Laplacian: cart_to_nmode . K . nmode_to_cart = diag(lambda / nbead) with
  K = circulant(2, -1, ..., -1), i.e. lambda_n / nbead = 2 (1 - cos(2 pi n / nbead)).
* Langevin ordering: the matrix is orthonormal.
* The spring energy equals its normal-mode form 1/2 coeff sum_i m_i sum_n lambda_n q_n,i^2,
  the spring forces are minus its gradient (finite differences), and Epot_deriv is the
  centroid-virial sum computed directly.
* Transforms: positions through cart_to_nmode and back through nmode_to_cart are unchanged.
"""
import numpy as np
import pytest

import oracle_lib as o


@pytest.mark.parametrize("nb", [2, 4, 8, 32])
def test_default_matrices(nb):
    n2c, c2n, tc2n, lam = o.ring_matrices(nb, 0)
    assert np.abs(n2c @ c2n - np.eye(nb)).max() < 1e-13
    assert np.array_equal(tc2n, n2c.T)
    K = 2 * np.eye(nb) - np.roll(np.eye(nb), 1, 1) - np.roll(np.eye(nb), -1, 1)
    if nb == 2:
        K = 2 * np.eye(2) - 2 * np.roll(np.eye(2), 1, 1)    # both neighbours are the other bead
    D = c2n @ K @ n2c
    assert np.abs(D - np.diag(np.diag(D))).max() < 1e-12
    assert np.abs(np.diag(D) - lam / nb).max() < 1e-12
    assert lam[0] == 0.0 and lam[-1] == 4.0 * nb


@pytest.mark.parametrize("nb", [4, 16])
def test_langevin_matrices_orthonormal(nb):
    n2c, c2n, tc2n, lam = o.ring_matrices(nb, 1)
    assert np.abs(n2c.T @ n2c - np.eye(nb)).max() < 1e-13
    assert np.array_equal(c2n, n2c.T) and np.all(lam == 1.0)


def test_spring_energy_force_and_virial():
    rng = np.random.default_rng(3)
    nb, n = 8, 12
    mass = rng.uniform(1.0, 16.0, n)
    x = rng.normal(size=(3, n))[None] + 0.05 * rng.normal(size=(nb, 3, n))
    f = rng.normal(size=(nb, 3, n))
    coeff = 37.5
    n2c, c2n, tc2n, lam = o.ring_matrices(nb, 0)
    q = o.ring_transform(c2n, x.reshape(nb, -1)).reshape(nb, 3, n)
    assert np.abs(o.ring_transform(n2c, q.reshape(nb, -1)).reshape(nb, 3, n) - x).max() < 1e-13
    xc = q[0]
    assert np.abs(xc - x.mean(0)).max() < 1e-14
    f1, es, ed = o.ring_spring(x, xc, f, mass, coeff)
    # normal-mode form of the spring energy
    e_nm = 0.5 * coeff * np.sum(mass[None, None, :] * lam[:, None, None] * q ** 2)
    assert abs(es - e_nm) <= 1e-12 * abs(es)
    # centroid virial term
    assert abs(ed - 0.5 * np.sum(-(x - xc[None]) * f)) <= 1e-12 * abs(ed)

    def espring(xx):
        d = xx - np.roll(xx, 1, axis=0)
        return 0.5 * coeff * np.sum(mass[None, None, :] * d ** 2)

    h = 1e-6
    g = np.zeros_like(x)
    for idx in [(0, 0, 0), (3, 1, 5), (7, 2, 11), (4, 0, 7)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        g[idx] = (espring(xp) - espring(xm)) / (2 * h)
        assert abs((f1 - f)[idx] + g[idx]) <= 1e-6 * max(1.0, abs(g[idx]))
