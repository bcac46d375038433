"""The oracle's mbpol_forces restatement: list predicates, decomposition, heat flux.

Pins for oracle/mbpol_oracle.c and oracle/heat_oracle.c:
  * dimer / trimer sets against an independent numpy brute force of the predicates
    (mbpol.f90:496-503, 843-869);
  * the reference's in-place re-imaging (mbpol.f90:505-509, 875-882) changes nothing but
    rounding;
  * the per-atom decomposition sums back to the totals (heat_check.f90:24-45) and the
    force-matrix rows sum to the forces (heatcurrent.f90:841-852);
  * the per-atom tensor form of the heat flux equals the dense force-matrix form of
    compute_heat_flux2 (heatcurrent.f90:703-728), and J_pot is linear in v;
  * Newton's third law, virial = -trace(stress), energy finite differences.
"""
import itertools

import numpy as np
import pytest

import oracle_lib as o


@pytest.fixture(scope="module")
def box64(waterbox):
    return waterbox.box_64()


@pytest.fixture(scope="module")
def res64(box64):
    return o.mbpol_forces(box64, compat_inplace=1, dense=True)


def _min_image_dist(box):
    O = box.xyz[:box.nreal:3]
    d = O[:, None, :] - O[None, :, :]
    d -= box.L * np.round(d / box.L)
    return np.sqrt((d ** 2).sum(-1))


def test_lists_match_brute_force(box64, res64):
    r = _min_image_dist(box64)
    nw = box64.nw
    dim = {(i, j) for i in range(nw) for j in range(i + 1, nw) if r[i, j] <= 6.5}
    assert set(map(tuple, res64["dimers"].tolist())) == dim
    tri = set()
    for i, j, k in itertools.combinations(range(nw), 3):
        short = int(r[i, j] <= 4.5) + int(r[i, k] <= 4.5) + int(r[j, k] <= 4.5)
        if short >= 2:
            tri.add((i, j, k))
    assert set(map(tuple, res64["trimers"].tolist())) == tri
    assert len(res64["trimers"]) == len(tri)


def test_inplace_reimaging_is_rounding_only(box64, res64):
    clean = o.mbpol_forces(box64, compat_inplace=0)
    assert np.array_equal(clean["dimers"], res64["dimers"])
    assert np.array_equal(clean["trimers"], res64["trimers"])
    assert abs(clean["eng"] - res64["eng"]) <= 1e-12 * abs(res64["eng"])
    assert np.abs(clean["f"] - res64["f"]).max() <= 1e-12 * np.abs(res64["f"]).max()


def test_decomposition_sums(box64, res64):
    ae = res64["atomic_energy"]
    assert abs(ae.sum() - (res64["eng2"] + res64["eng3"])) <= 1e-12 * abs(res64["eng"])
    assert np.all(ae[box64.nreal:] == 0.0)
    fm = res64["force_matrix"].reshape(3, box64.natms, box64.natms)   # [k][j][i]
    rows = fm.sum(axis=1).T                                             # sum over j
    assert np.abs(rows - res64["f"]).max() <= 1e-10 * np.abs(res64["f"]).max()


def test_newton_virial_stress(box64, res64):
    f = res64["f"]
    assert np.abs(f.sum(0)).max() <= 1e-9 * np.abs(f).max()
    s = res64["stress"]
    assert abs((s[0] + s[4] + s[8]) + res64["vir"]) <= 1e-12 * abs(res64["vir"])
    assert s[1] == s[3] and s[2] == s[6] and s[5] == s[7]


def test_heat_flux_tensor_equals_dense(box64, res64):
    rng = np.random.default_rng(3)
    vel = box64.vel
    ek = 0.5 * box64.mass * (vel ** 2).sum(1)
    ep = res64["atomic_energy"]
    sew = rng.normal(size=(box64.natms, 9))
    dense = o.heat_flux_dense(box64, res64["force_matrix"], vel, ek, ep, sew.ravel().copy())
    tens = o.heat_flux_tensor(box64, res64["heat_w"], vel, ek, ep, sew.ravel().copy())
    for a, b in zip(dense, tens):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()
    # J_pot is linear in the velocities
    j2 = o.heat_flux_tensor(box64, res64["heat_w"], 2 * vel, ek, ep, None)
    j1 = o.heat_flux_tensor(box64, res64["heat_w"], vel, ek, ep, None)
    assert np.allclose(j2[1], 2 * j1[1], rtol=1e-12, atol=0)


def test_energy_finite_difference(waterbox):
    small = waterbox.make_waterbox(8, 6.21, 2, seed=7)
    base = o.mbpol_forces(small, lists=False)
    h = 1e-5
    for atom in (0, 1, 2, 13):
        for ax in range(3):
            bp, bm = _shift(small, atom, ax, h), _shift(small, atom, ax, -h)
            ep = o.mbpol_forces(bp, lists=False)["eng"]
            em = o.mbpol_forces(bm, lists=False)["eng"]
            num = -(ep - em) / (2 * h)
            assert abs(num - base["f"][atom, ax]) <= 1e-5 * max(1.0, abs(base["f"][atom, ax]))


def _shift(box, atom, ax, h):
    import copy
    b = copy.deepcopy(box)
    b.xyz = box.xyz.copy()
    b.xyz[atom, ax] += h
    return b


def test_unsupported_imcon_is_an_error(box64):
    """imcon 4-7 (truncated octahedron, ...) are not images this path handles: the reference
    stops in fixme (mbpol.f90:1357); imcon 0 is the isolated cluster, without images."""
    import copy
    b = copy.deepcopy(box64)
    b.imcon = 4
    with pytest.raises(RuntimeError):
        o.mbpol_forces(b)
