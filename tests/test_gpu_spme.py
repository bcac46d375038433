"""Particle-mesh Ewald reciprocal space on the device (mbgpu_spme_recip_dev) against the
oracle's SPME and the Ewald reciprocal sum it approximates.

* Same order and grid as the oracle: energy, potential, field and force within 1e-10
  (the device sums the spread density in 64-bit fixed point, 2^-44 e per contribution),
  on grids of powers of two and of mixed radices 2, 3, 5 (48^3, 24 x 30 x 40, 45 x 20 x 36,
  250 x 12 x 10); the oracle transforms by direct DFT sums.
* Order 12 on a 64^3 grid against the Ewald sum over k: within 1e-9 (a triclinic cell).
* The 512-water geometry with DMS-sized charges and induced-dipole-sized dipoles, order 8
  on 64^3, against the Ewald sum at kmax 16: the SPME discretisation error, 1e-5.
* Repeat calls bit-identical; positions shifted by lattice vectors give the same result
  (1e-11); U = 1/2 sum (q phi - mu . E) at 8192 sites; charges alone; no sites and one
  site; argument and input errors reported.
"""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


def system(n=24, seed=1, L=12.0):
    rng = np.random.default_rng(seed)
    cell = np.diag([L, 1.1 * L, 0.95 * L])
    cell[1, 0] = 1.0
    xyz = rng.uniform(0, 1, (n, 3)) @ cell
    q = rng.normal(size=n)
    q -= q.mean()
    mu = rng.normal(size=(n, 3)) * 0.3
    return cell, xyz, q, mu


class Dev:
    def __init__(self, pkg):
        import torch
        self.t = torch
        self.ctx = pkg.Context(8, 32)

    def run(self, cell, order, grid, alpha, xyz, q, mu=None, outputs=True):
        t = self.t
        f64 = dict(dtype=t.float64, device="cuda")
        n = len(q)
        xyz = np.asarray(xyz, float)
        x, y, z = (t.as_tensor(np.ascontiguousarray(xyz[:, a]), **f64) for a in range(3))
        qd = t.as_tensor(np.asarray(q, float), **f64)
        md = None if mu is None else t.as_tensor(np.ascontiguousarray(mu, float).reshape(-1), **f64)
        phi, E, F = t.zeros(n, **f64), t.zeros(3 * n, **f64), t.zeros(3 * n, **f64)
        t.cuda.synchronize()
        ptr = lambda a: None if a is None else a.data_ptr()
        outs = (phi, E, F) if outputs is True else (phi, E, None) if outputs == "field" else (None,) * 3
        U = self.ctx.spme_recip_dev(cell, order, grid, alpha, n, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                                    qd.data_ptr(), ptr(md), *map(ptr, outs))
        return dict(U=U, phi=phi.cpu().numpy(), E=E.cpu().numpy().reshape(n, 3), F=F.cpu().numpy().reshape(n, 3))


@pytest.fixture(scope="module")
def dev(pkg):
    return Dev(pkg)


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


@pytest.mark.parametrize("order,grid", [(4, (16, 16, 16)), (8, (32, 16, 64)), (12, (32, 32, 32)), (6, (48, 48, 48)),
                                        (8, (24, 30, 40)), (10, (45, 20, 36)), (6, (250, 12, 10))])
def test_matches_oracle_spme(dev, order, grid):
    cell, xyz, q, mu = system()
    print(f"spme order {order} grid {grid}", flush=True)
    g = dev.run(cell, order, grid, 0.35, xyz, q, mu)
    r = o.spme_recip(cell, order, grid, 0.35, xyz, q, mu)
    assert abs(g["U"] - r["U"]) <= 1e-10 * abs(r["U"])
    for k in ("phi", "E", "F"):
        assert rel(g[k], r[k]) <= 1e-10, k


def test_converges_to_ewald_sum(dev):
    cell, xyz, q, mu = system()
    g = dev.run(cell, 12, (64, 64, 64), 0.35, xyz, q, mu)
    ew = o.ewald_recip(cell, (14, 14, 14), 0.35, xyz, q, mu)
    assert abs(g["U"] - ew["U"]) <= 1e-9 * abs(ew["U"])
    for k in ("phi", "E", "F"):
        assert rel(g[k], ew[k]) <= 1e-9, k


def test_water_box_against_ewald_sum(dev, waterbox):
    box = waterbox.make_waterbox(512, 24.8342, 8, seed=11)
    rng = np.random.default_rng(2)
    xyz = np.concatenate([box.xyz[:box.nreal], waterbox.m_sites(box.xyz[:box.nreal], box.nw, box.L)])
    n = len(xyz)
    q = np.concatenate([np.tile([0.0, 0.6, 0.6], box.nw), -1.2 * np.ones(box.nw)]) + rng.normal(size=n) * 0.02
    q -= q.mean()
    mu = rng.normal(size=(n, 3)) * 0.05
    cell = np.diag([box.L] * 3)
    print("512-water geometry, order 8, 64^3", flush=True)
    g = dev.run(cell, 8, (64, 64, 64), 0.3, xyz, q, mu)
    ew = o.ewald_recip(cell, (16, 16, 16), 0.3, xyz, q, mu)
    assert abs(g["U"] - ew["U"]) <= 1e-5 * abs(ew["U"])
    assert rel(g["E"], ew["E"]) <= 1e-4 and rel(g["F"], ew["F"]) <= 1e-4


def test_repeat_shift_identity_and_charges_only(dev):
    cell, xyz, q, mu = system(n=8192, seed=4, L=40.0)
    a = dev.run(cell, 8, (64, 64, 64), 0.3, xyz, q, mu)
    b = dev.run(cell, 8, (64, 64, 64), 0.3, xyz, q, mu)
    assert a["U"] == b["U"] and all(np.array_equal(a[k], b[k]) for k in ("phi", "E", "F"))
    shift = xyz + np.array([1, -2, 3]) @ cell
    s = dev.run(cell, 8, (64, 64, 64), 0.3, shift, q, mu)
    assert abs(s["U"] - a["U"]) <= 1e-11 * abs(a["U"]) and rel(s["F"], a["F"]) <= 1e-9
    u2 = 0.5 * (q @ a["phi"] - (mu * a["E"]).sum())
    assert abs(u2 - a["U"]) <= 1e-10 * abs(a["U"])
    c = dev.run(cell, 8, (64, 64, 64), 0.3, xyz, q, None)
    assert abs(c["U"] - 0.5 * q @ c["phi"]) <= 1e-10 * abs(c["U"])
    e = dev.run(cell, 8, (64, 64, 64), 0.3, xyz, q, mu, outputs=False)
    assert e["U"] == a["U"]
    # potential and field without forces (no second derivatives gathered): the same values
    fo = dev.run(cell, 8, (64, 64, 64), 0.3, xyz, q, mu, outputs="field")
    assert fo["U"] == a["U"] and np.array_equal(fo["phi"], a["phi"]) and np.array_equal(fo["E"], a["E"])
    assert not fo["F"].any()


def test_argument_and_input_errors(pkg, dev):
    cell, xyz, q, mu = system()
    for order, grid, what in ((5, (16, 16, 16), "order"), (14, (32, 32, 32), "order"),
                              (8, (14, 16, 16), "3\\^b 5\\^c"), (8, (16, 16, 4), "3\\^b 5\\^c"),
                              (8, (512, 16, 16), "3\\^b 5\\^c"), (8, (16, 21, 16), "3\\^b 5\\^c"),
                              (12, (16, 10, 16), "3\\^b 5\\^c")):
        with pytest.raises(pkg.MbgpuError, match=what):
            dev.run(cell, order, grid, 0.35, xyz, q, mu)
    bad = xyz.copy()
    bad[3, 1] = np.nan
    with pytest.raises(pkg.MbgpuError, match="not finite"):
        dev.run(cell, 8, (16, 16, 16), 0.35, bad, q, mu)
    big = mu.copy()
    big[5] = [1e5, 0, 0]
    with pytest.raises(pkg.MbgpuError, match="too large"):
        dev.run(cell, 8, (16, 16, 16), 0.35, xyz, q, big)
    ok = dev.run(cell, 8, (16, 16, 16), 0.35, xyz, q, mu)
    r = o.spme_recip(cell, 8, (16, 16, 16), 0.35, xyz, q, mu)
    assert abs(ok["U"] - r["U"]) <= 1e-10 * abs(r["U"])
    # no sites: zero energy; one site: the oracle's value
    assert dev.ctx.spme_recip_dev(cell, 8, (16, 16, 16), 0.35, 0, None, None, None, None) == 0.0
    one = dev.run(cell, 6, (16, 32, 16), 0.35, xyz[:1], [0.7], mu[:1])
    r1 = o.spme_recip(cell, 6, (16, 32, 16), 0.35, xyz[:1], [0.7], mu[:1])
    assert abs(one["U"] - r1["U"]) <= 1e-10 * abs(r1["U"])
    # a lone site's reciprocal force is the mesh's small self-force: compare absolutely
    assert np.abs(one["F"] - r1["F"]).max() <= 1e-10 * max(1.0, np.abs(r1["E"]).max())
