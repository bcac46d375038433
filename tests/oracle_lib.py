"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference's own
compiled leaves (oracle/_ref/libmbpol_ref.so, present only where the reference tree was
available to build them).  Test infrastructure: the product never imports this."""
import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
DATA = os.path.join(REPO, "dlpoly2-mbpol-heat_amd", "data")
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def P(a):
    return None if a is None else a.ctypes.data_as(_dp)


class MbpolArgs(ctypes.Structure):
    _fields_ = [("imcon", ctypes.c_int), ("cell", ctypes.c_double * 9), ("nw", ctypes.c_int),
                ("listttm2", _ip), ("xxx", _dp), ("yyy", _dp), ("zzz", _dp), ("natms", ctypes.c_int),
                ("compat_inplace", ctypes.c_int), ("fxx", _dp), ("fyy", _dp), ("fzz", _dp),
                ("engsrp", ctypes.c_double), ("virsrp", ctypes.c_double), ("stress", ctypes.c_double * 9),
                ("eng2", ctypes.c_double), ("eng3", ctypes.c_double), ("atomic_energy", _dp), ("heat_w", _dp),
                ("force_matrix", _dp), ("dimers", _ip), ("ndimers", ctypes.c_long), ("cap_dimers", ctypes.c_long),
                ("trimers", _ip), ("ntrimers", ctypes.c_long), ("cap_trimers", ctypes.c_long),
                ("skip_3b", ctypes.c_int)]


_ora = None
_ora_exact = None


def oracle_exact():
    """The oracle built with the C library's erfc in place of the reference's approximation
    (liboracle_exact.so): for checks of the mathematics that the approximation blurs."""
    global _ora_exact
    if _ora_exact is None:
        so = os.path.join(ORACLE_DIR, "liboracle_exact.so")
        if not os.path.exists(so):
            subprocess.check_call(["make", "-C", ORACLE_DIR, "oracle"], stdout=subprocess.DEVNULL)
        lib = ctypes.CDLL(so)
        lib.ora_erfcc.restype = ctypes.c_double
        lib.ora_erfcc.argtypes = [ctypes.c_double]
        lib.ora_msite_distribute.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        if lib.ora_dms_load(os.path.join(DATA, "dms_table.txt").encode()) != 0:
            raise RuntimeError("oracle could not load the DMS table")
        _ora_exact = lib
    return _ora_exact


def oracle():
    global _ora
    if _ora is not None:
        return _ora
    so = os.path.join(ORACLE_DIR, "liboracle.so")
    if not os.path.exists(so):
        subprocess.check_call(["make", "-C", ORACLE_DIR, "oracle"], stdout=subprocess.DEVNULL)
    lib = ctypes.CDLL(so)
    lib.ora_poly_eval.restype = ctypes.c_double
    lib.ora_poly_scale.restype = ctypes.c_double
    lib.ora_mbpol_2b.restype = ctypes.c_double
    lib.ora_mbpol_3b.restype = ctypes.c_double
    files = [os.path.join(DATA, f).encode() for f in
             ("poly2b_monomials.txt", "poly2b_coeffs.txt", "poly3b_monomials.txt", "poly3b_coeffs.txt")]
    if lib.ora_pip_load(*files) != 0:
        raise RuntimeError("oracle could not load the polynomial tables")
    _ora = lib
    return lib


def reference_leaves():
    """The reference's compiled C++ leaves, or None when they were not built here."""
    so = os.path.join(ORACLE_DIR, "_ref", "libmbpol_ref.so")
    if not os.path.exists(so):
        return None
    lib = ctypes.CDLL(so)
    return lib


def leaf_2b(w1, w2):
    g1, g2 = np.zeros(9), np.zeros(9)
    e = oracle().ora_mbpol_2b(P(np.ascontiguousarray(w1, float)), P(np.ascontiguousarray(w2, float)), P(g1), P(g2))
    return e, np.concatenate([g1, g2])


def leaf_3b(w1, w2, w3):
    g = [np.zeros(9) for _ in range(3)]
    ws = [np.ascontiguousarray(w, float) for w in (w1, w2, w3)]
    e = oracle().ora_mbpol_3b(P(ws[0]), P(ws[1]), P(ws[2]), P(g[0]), P(g[1]), P(g[2]))
    return e, np.concatenate(g)


def vars_2b(w1, w2):
    v = np.zeros(31)
    oracle().ora_vars_2b(P(np.ascontiguousarray(w1, float)), P(np.ascontiguousarray(w2, float)), P(v))
    return v


def vars_3b(w1, w2, w3):
    v = np.zeros(36)
    ws = [np.ascontiguousarray(w, float) for w in (w1, w2, w3)]
    oracle().ora_vars_3b(P(ws[0]), P(ws[1]), P(ws[2]), P(v))
    return v


def poly_scale(which, x):
    return oracle().ora_poly_scale(which, P(np.ascontiguousarray(x, float)))


def mbpol_forces(box, compat_inplace=1, dense=False, lists=True, skip_3b=False):
    """Run the oracle's mbpol_forces on a WaterBox; returns a dict of results."""
    lib = oracle()
    a = MbpolArgs()
    a.imcon = box.imcon
    for q in range(9):
        a.cell[q] = box.cell[q]
    a.nw = box.nw
    lst = np.ascontiguousarray(box.listttm2, dtype=np.int32)
    a.listttm2 = lst.ctypes.data_as(_ip)
    x, y, z = box.x, box.y, box.z
    a.xxx, a.yyy, a.zzz = P(x), P(y), P(z)
    a.natms = box.natms
    a.compat_inplace = compat_inplace
    f = np.zeros((3, box.natms))
    a.fxx, a.fyy, a.fzz = P(f[0]), P(f[1]), P(f[2])
    ae = np.zeros(box.natms)
    hw = np.zeros(9 * box.natms)
    a.atomic_energy, a.heat_w = P(ae), P(hw)
    fm = np.zeros(3 * box.natms * box.natms) if dense else None
    a.force_matrix = P(fm)
    cap_d, cap_t = 64 * box.nw, 512 * box.nw
    d = np.zeros(2 * cap_d, dtype=np.int32) if lists else None
    t = np.zeros(3 * cap_t, dtype=np.int32) if lists else None
    if lists:
        a.dimers, a.cap_dimers = d.ctypes.data_as(_ip), cap_d
        a.trimers, a.cap_trimers = t.ctypes.data_as(_ip), cap_t
    a.skip_3b = int(skip_3b)
    rc = lib.ora_mbpol_forces(ctypes.byref(a))
    if rc != 0:
        raise RuntimeError(f"oracle mbpol_forces returned {rc}")
    out = dict(f=f.T.copy(), eng=a.engsrp, vir=a.virsrp, stress=np.array(a.stress[:]), eng2=a.eng2,
               eng3=a.eng3, atomic_energy=ae, heat_w=hw.reshape(box.natms, 3, 3), nd=a.ndimers, nt=a.ntrimers)
    if lists:
        out["dimers"] = d[:2 * a.ndimers].reshape(-1, 2)
        out["trimers"] = t[:3 * a.ntrimers].reshape(-1, 3)
    if dense:
        out["force_matrix"] = fm
    return out


def heat_flux_dense(box, fm, vel, e_kin, e_pot, stress_ew1p=None):
    jk, jp, jt = np.zeros(3), np.zeros(3), np.zeros(3)
    v = np.ascontiguousarray(vel.T)
    oracle().ora_heat_flux_dense(box.natms, box.nreal, P(fm), P(box.x), P(box.y), P(box.z), P(v[0]), P(v[1]),
                                 P(v[2]), P(np.ascontiguousarray(e_kin)), P(np.ascontiguousarray(e_pot)),
                                 P(stress_ew1p), P(jk), P(jp), P(jt))
    return jk, jp, jt


def heat_flux_tensor(box, heat_w, vel, e_kin, e_pot, stress_ew1p=None):
    jk, jp, jt = np.zeros(3), np.zeros(3), np.zeros(3)
    v = np.ascontiguousarray(vel.T)
    hw = np.ascontiguousarray(heat_w, dtype=float).reshape(-1)
    oracle().ora_heat_flux_tensor(box.nreal, P(hw), P(v[0]), P(v[1]), P(v[2]), P(np.ascontiguousarray(e_kin)),
                                  P(np.ascontiguousarray(e_pot)), P(stress_ew1p), P(jk), P(jp), P(jt))
    return jk, jp, jt


# ---------------------------------------------------------------- electrostatics ----

class ElecArgs(ctypes.Structure):
    _fields_ = [("imcon", ctypes.c_int), ("cell", ctypes.c_double * 9), ("nw", ctypes.c_int),
                ("x", _dp), ("y", _dp), ("z", _dp), ("rcut", ctypes.c_double), ("alpha", ctypes.c_double),
                ("kmax", ctypes.c_int * 3), ("toler", ctypes.c_double), ("omega", ctypes.c_double),
                ("max_iter", ctypes.c_int), ("gamma", ctypes.c_double), ("q_o", ctypes.c_double),
                ("polr_site", ctypes.c_double * 3), ("polr2_site", ctypes.c_double * 3),
                ("athole", ctypes.c_double), ("athole12", ctypes.c_double), ("athole13", ctypes.c_double),
                ("ascc", ctypes.c_double), ("ascd", ctypes.c_double), ("compat_qd_column", ctypes.c_int),
                ("dip", _dp), ("f", _dp), ("chge", _dp), ("vesp", _dp), ("efield", _dp), ("f_ew1p", _dp),
                ("eng_ew1p", ctypes.c_double), ("eng_ew2p", ctypes.c_double), ("eng_ew3p", ctypes.c_double),
                ("eng_polar", ctypes.c_double), ("eng_cpe", ctypes.c_double), ("vir_cpe", ctypes.c_double),
                ("vir_ew1p", ctypes.c_double), ("vir_qd", ctypes.c_double), ("stress", ctypes.c_double * 9),
                ("stress_ew1p", ctypes.c_double * 9),
                ("e_ew1p", _dp), ("e_ew2p", _dp), ("e_ew3p", _dp), ("e_polar", _dp), ("s_ew1p", _dp),
                ("heat_w", _dp), ("force_matrix", _dp), ("npairs", ctypes.c_long), ("niter", ctypes.c_int),
                ("last_delta", ctypes.c_double), ("volm", ctypes.c_double),
                ("stress_qd_atoms", ctypes.c_double * 9)]


_dms_loaded = False


def elec_lib():
    global _dms_loaded
    lib = oracle()
    if not _dms_loaded:
        for name in ("ora_erfcc", "ora_gammln", "ora_gammq"):
            getattr(lib, name).restype = ctypes.c_double
        lib.ora_erfcc.argtypes = [ctypes.c_double]
        lib.ora_gammln.argtypes = [ctypes.c_double]
        lib.ora_gammq.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.ora_erfc_grid.argtypes = [ctypes.c_double, _dp]
        lib.ora_erfc_table.argtypes = [ctypes.c_double, ctypes.c_double, _dp]
        lib.ora_msite_distribute.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        if lib.ora_dms_load(os.path.join(DATA, "dms_table.txt").encode()) != 0:
            raise RuntimeError("oracle could not load the DMS table")
        _dms_loaded = True
    return lib


def erfc_table(alpha, rcut):
    lib = elec_lib()
    dr = ctypes.c_double()
    n = lib.ora_erfc_grid(rcut, ctypes.byref(dr))
    t = np.zeros((n + 1) * 4)
    lib.ora_erfc_table(alpha, rcut, P(t))
    return t.reshape(n + 1, 4), dr.value


def dms_charges(w9, gamma=0.426706882, imcon=1, cell=None):
    lib = elec_lib()
    cell = np.asarray(cell if cell is not None else [100.0, 0, 0, 0, 100.0, 0, 0, 0, 100.0], float)
    q = np.zeros(3)
    rc = lib.ora_dms_charges(imcon, P(cell), P(np.ascontiguousarray(w9, float)), ctypes.c_double(gamma), P(q))
    assert rc == 0
    return q


def dms_p(w9, imcon=1, cell=None):
    """p1, p2 and (dp1/dr1, dp1/dr2, dp1/dcos), (dp2/dr1, dp2/dr2, dp2/dcos) of one water."""
    lib = elec_lib()
    cell = np.asarray(cell if cell is not None else [100.0, 0, 0, 0, 100.0, 0, 0, 0, 100.0], float)
    p = np.zeros(8)
    assert lib.ora_dms_p(imcon, P(cell), P(np.ascontiguousarray(w9, float)), P(p)) == 0
    return p[0], p[1], p[2:5].copy(), p[5:8].copy()


def qdforce(xyz, vesp, gamma, imcon, cell, compat_qd_column=1):
    """The oracle's qdforce alone for given sites (4*nw, M sites as given) and potentials:
    (forces (n, 3), stress[9], virial, sum of the angular per-atom stresses [9], force
    matrix [3][n][n] with the M rows moved onto O, H1, H2)."""
    lib = elec_lib()
    xyz = np.asarray(xyz, float)
    n = xyz.shape[0]
    a = ElecArgs()
    a.imcon = imcon
    for q in range(9):
        a.cell[q] = cell[q]
    a.nw = n // 4
    x, y, z = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    a.x, a.y, a.z = P(x), P(y), P(z)
    a.gamma, a.compat_qd_column = gamma, compat_qd_column
    v = np.ascontiguousarray(vesp, float)
    f, fm = np.zeros(3 * n), np.zeros(3 * n * n)
    a.vesp, a.f, a.force_matrix = P(v), P(f), P(fm)
    assert lib.ora_qdforce(ctypes.byref(a)) == 0
    return (f.reshape(n, 3), np.array(a.stress[:]), a.vir_qd, np.array(a.stress_qd_atoms[:]), fm.reshape(3, n, n))


def site_xyz(box):
    """4*nw site positions (O, H1, H2 per water, then M sites), interleaved (n, 3)."""
    return np.ascontiguousarray(box.xyz[:4 * box.nw], float).copy()


def elec_forces(box, p, dip0=None, dense=False, xyz=None, cell=None, imcon=None, distribute=True, exact_erfc=False):
    """Run the oracle's polarizable electrostatics on a WaterBox with ElecParams p.
    Returns a dict; per-atom arrays are redistributed from the M sites when
    distribute=True (forces of real atoms, energies, ew1p stresses).  exact_erfc: the
    oracle build with the exact erfc (oracle_exact)."""
    lib = oracle_exact() if exact_erfc else elec_lib()
    nw = box.nw
    n = 4 * nw
    a = ElecArgs()
    a.imcon = box.imcon if imcon is None else imcon
    c = box.cell if cell is None else cell
    for q in range(9):
        a.cell[q] = c[q]
    a.nw = nw
    r = site_xyz(box) if xyz is None else np.array(xyz, float)
    x, y, z = (np.ascontiguousarray(r[:, k]) for k in range(3))
    a.x, a.y, a.z = P(x), P(y), P(z)
    a.rcut, a.alpha = p.rcut, p.alpha
    for q in range(3):
        a.kmax[q] = p.kmax[q]
    a.toler, a.omega, a.max_iter = p.toler, p.omega, p.max_iter
    a.gamma, a.q_o = p.gamma, p.q_o
    for q in range(3):
        a.polr_site[q] = p.polr[q]
        a.polr2_site[q] = p.polr2[q]
    a.athole, a.athole12, a.athole13, a.ascc, a.ascd = p.athole, p.athole12, p.athole13, p.ascc, p.ascd
    a.compat_qd_column = p.compat_qd_column
    dip = np.zeros(3 * n) if dip0 is None else np.array(dip0, float).reshape(-1).copy()
    f = np.zeros(3 * n)
    chge, vesp, efield, fk = np.zeros(n), np.zeros(n), np.zeros(3 * n), np.zeros(3 * n)
    e1, e2, e3, ep = (np.zeros(n) for _ in range(4))
    s1 = np.zeros(9 * n)
    hw = np.zeros(9 * 3 * nw)
    fm = np.zeros(3 * n * n) if dense else None
    a.dip, a.f, a.chge, a.vesp, a.efield, a.f_ew1p = P(dip), P(f), P(chge), P(vesp), P(efield), P(fk)
    a.e_ew1p, a.e_ew2p, a.e_ew3p, a.e_polar, a.s_ew1p = P(e1), P(e2), P(e3), P(ep), P(s1)
    a.heat_w, a.force_matrix = P(hw), P(fm)
    rc = lib.ora_elec_forces(ctypes.byref(a))
    if rc != 0:
        raise RuntimeError(f"oracle elec_forces returned {rc} (niter {a.niter})")
    out = dict(f_sites=f.reshape(n, 3).copy(), chge=chge, vesp=vesp, dip=dip.reshape(n, 3), efield=efield.reshape(n, 3),
               eng_ew1p=a.eng_ew1p, eng_ew2p=a.eng_ew2p, eng_ew3p=a.eng_ew3p, eng_polar=a.eng_polar,
               eng=a.eng_cpe, vir=a.vir_cpe, vir_ew1p=a.vir_ew1p, vir_qd=a.vir_qd, stress=np.array(a.stress[:]),
               stress_ew1p=np.array(a.stress_ew1p[:]),
               e_sites=np.stack([e1, e2, e3, ep]), s_ew1p_sites=s1.reshape(n, 9).copy(), heat_w=hw.reshape(3 * nw, 3, 3),
               npairs=a.npairs, niter=a.niter, last_delta=a.last_delta, xyz=np.stack([x, y, z], 1),
               stress_qd_atoms=np.array(a.stress_qd_atoms[:]))
    out["f_ew1p_sites"] = fk.reshape(n, 3).copy()
    if distribute:
        e = (e1 + e2 + e3 + ep).copy()
        lib.ora_msite_distribute(nw, p.gamma, P(f), P(e), P(s1))
        lib.ora_msite_distribute(nw, p.gamma, P(fk), None, None)
        out["f_ew1p"] = fk.reshape(n, 3)[:3 * nw].copy()
        out["f"] = f.reshape(n, 3)[:3 * nw].copy()
        out["atomic_energy"] = e[:3 * nw].copy()
        out["s_ew1p"] = s1.reshape(n, 9)[:3 * nw].copy()
    if dense:
        out["force_matrix"] = fm
    return out


class WaterArgs(ctypes.Structure):
    _fields_ = [("imcon", ctypes.c_int), ("cell", ctypes.c_double * 9), ("nw", ctypes.c_int),
                ("x", _dp), ("y", _dp), ("z", _dp), ("rcut", ctypes.c_double), ("rvdw", ctypes.c_double),
                ("c6", ctypes.c_double * 3), ("a6", ctypes.c_double * 3), ("c8", ctypes.c_double * 3),
                ("a8", ctypes.c_double * 3), ("bond_key", ctypes.c_int * 2), ("bond_prm", (ctypes.c_double * 3) * 2),
                ("ps_monomer", ctypes.c_int), ("dispersion", ctypes.c_int), ("long_range", ctypes.c_int),
                ("f", _dp), ("atomic_energy", _dp), ("heat_w", _dp),
                ("force_matrix", _dp), ("eng_bnd", ctypes.c_double), ("vir_bnd", ctypes.c_double),
                ("eng_ang", ctypes.c_double), ("vir_ang", ctypes.c_double),
                ("eng_srp", ctypes.c_double), ("vir_srp", ctypes.c_double), ("elrc", ctypes.c_double),
                ("virlrc", ctypes.c_double), ("stress", ctypes.c_double * 9)]


_ps_loaded = False


def water_lib():
    global _ps_loaded
    lib = oracle()
    if not _ps_loaded:
        lib.ora_tang_toennies.restype = ctypes.c_double
        lib.ora_tang_toennies.argtypes = [ctypes.c_int, ctypes.c_double]
        lib.ora_ps_couple.argtypes = [ctypes.c_double, _dp, ctypes.c_double, _dp, ctypes.c_double, _dp, _dp, _dp]
        lib.ora_ps_coefficients.argtypes = [_dp, _dp]
        if lib.ora_ps_load(os.path.join(DATA, "ps_table.txt").encode(),
                           os.path.join(DATA, "ps_mass.txt").encode()) != 0:
            raise RuntimeError("oracle could not load the PS tables")
        _ps_loaded = True
    return lib


def tang_toennies(n, x):
    return water_lib().ora_tang_toennies(n, x)


def ps_coefficients():
    lib = water_lib()
    c5z, cm = np.zeros(245), np.zeros(9)
    lib.ora_ps_coefficients(P(c5z), P(cm))
    return c5z, cm


def ps_couple(roh1, u1, roh2, u2, theta):
    lib = water_lib()
    pt, vt, fd = ctypes.c_double(), ctypes.c_double(), np.zeros(6)
    lib.ora_ps_couple(roh1, P(np.ascontiguousarray(u1, float)), roh2, P(np.ascontiguousarray(u2, float)), theta,
                      ctypes.byref(pt), P(fd), ctypes.byref(vt))
    return pt.value, fd.reshape(2, 3), vt.value


def water_terms(box, p, xyz=None, cell=None, imcon=None, dense=False):
    """PS monomer + TT dispersion + long-range correction of the oracle on a WaterBox with
    WaterParams p (terms switched off in p are dropped by zeroing their inputs)."""
    lib = water_lib()
    nw = box.nw
    n = 4 * nw
    a = WaterArgs()
    a.imcon = box.imcon if imcon is None else imcon
    c = box.cell if cell is None else cell
    for q in range(9):
        a.cell[q] = c[q]
    a.nw = nw
    r = site_xyz(box) if xyz is None else np.array(xyz, float)
    x, y, z = (np.ascontiguousarray(r[:, k]) for k in range(3))
    a.x, a.y, a.z = P(x), P(y), P(z)
    a.rcut, a.rvdw = p.rcut, p.rvdw
    c6, a6, c8, a8 = p.internal()
    for q in range(3):
        a.c6[q], a.a6[q], a.c8[q], a.a8[q] = c6[q], a6[q], c8[q], a8[q]
    for t, (key, prm) in enumerate(p.bonds_internal()):
        a.bond_key[t] = key
        for q in range(3):
            a.bond_prm[t][q] = prm[q]
    a.ps_monomer, a.dispersion, a.long_range = p.ps_monomer, p.dispersion, p.long_range
    f = np.zeros(3 * n)
    ae = np.zeros(3 * nw)
    hw = np.zeros(9 * 3 * nw)
    fm = np.zeros(3 * n * n) if dense else None
    a.f, a.atomic_energy, a.heat_w, a.force_matrix = P(f), P(ae), P(hw), P(fm)
    rc = lib.ora_water_terms(ctypes.byref(a))
    if rc != 0:
        raise RuntimeError(f"oracle water_terms returned {rc}")
    out = dict(f=f.reshape(n, 3)[:3 * nw].copy(), atomic_energy=ae, heat_w=hw.reshape(3 * nw, 3, 3),
               eng_bnd=a.eng_bnd, vir_bnd=a.vir_bnd, eng_ang=a.eng_ang, vir_ang=a.vir_ang, eng_srp=a.eng_srp, vir_srp=a.vir_srp, elrc=a.elrc,
               virlrc=a.virlrc, stress=np.array(a.stress[:]), xyz=np.stack([x, y, z], 1))
    if dense:
        out["force_matrix"] = fm
    return out


# ---------------------------------------------------------------- ring polymer ----

def ring_lib():
    lib = oracle()
    lib.ora_ring_matrices.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
    lib.ora_ring_transform.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
    lib.ora_ring_spring.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]
    return lib


def ring_matrices(nb, langevin=0):
    lib = ring_lib()
    a, b, c, lam = np.zeros(nb * nb), np.zeros(nb * nb), np.zeros(nb * nb), np.zeros(nb)
    lib.ora_ring_matrices(nb, langevin, P(a), P(b), P(c), P(lam))
    return a.reshape(nb, nb), b.reshape(nb, nb), c.reshape(nb, nb), lam


def ring_transform(M, x):
    """x: (P, ncoord) -> (P, ncoord)"""
    lib = ring_lib()
    x = np.ascontiguousarray(x, float)
    out = np.zeros_like(x)
    lib.ora_ring_transform(x.shape[0], x.shape[1], P(np.ascontiguousarray(M, float)), P(x), P(out))
    return out


def ring_spring(x, xc, f, mass, coeff, fac3=1.0, force_scale=1.0):
    """x, f: (P, 3, natom); xc: (3, natom).  Returns (f_new, Epot_spring, Epot_deriv)."""
    lib = ring_lib()
    x = np.ascontiguousarray(x, float)
    f = np.ascontiguousarray(f, float).copy()
    nb, _, n = x.shape
    es, ed = ctypes.c_double(), ctypes.c_double()
    lib.ora_ring_spring(nb, n, coeff, fac3, force_scale, P(np.ascontiguousarray(mass, float)), P(x),
                        P(np.ascontiguousarray(xc, float)), P(f), ctypes.byref(es), ctypes.byref(ed))
    return f, es.value, ed.value


# ---------------------------------------------------------------- particle-mesh Ewald ----

def spme_lib():
    lib = oracle()
    _ip3 = ctypes.POINTER(ctypes.c_int)
    lib.ora_bspline.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
    lib.ora_spme_bmod.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
    lib.ora_spme_recip.restype = ctypes.c_double
    lib.ora_spme_recip.argtypes = [_dp, ctypes.c_int, _ip3, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                   _dp, _dp, _dp]
    lib.ora_ewald_recip.restype = ctypes.c_double
    lib.ora_ewald_recip.argtypes = [_dp, _ip3, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                    _dp]
    return lib


def bspline(p, w):
    """(M_p(w + t), M_p'(w + t), M_p''(w + t)) for t = 0..p-1."""
    a, d, d2 = np.zeros(p), np.zeros(p), np.zeros(p)
    spme_lib().ora_bspline(p, w, P(a), P(d), P(d2))
    return a, d, d2


def _recip_args(cell, xyz, q, mu):
    cell = np.ascontiguousarray(cell, float).reshape(9)
    xyz = np.asarray(xyz, float)
    x, y, z = (np.ascontiguousarray(xyz[:, a]) for a in range(3))
    q = np.ascontiguousarray(q, float)
    mu = None if mu is None else np.ascontiguousarray(mu, float).reshape(-1)
    n = len(q)
    return cell, x, y, z, q, mu, n, np.zeros(n), np.zeros(3 * n), np.zeros(3 * n)


def spme_recip(cell, order, grid, alpha, xyz, q, mu=None):
    """SPME reciprocal space (Coulomb constant 1): dict(U, phi (n), E (n, 3), F (n, 3))."""
    cell, x, y, z, q, mu, n, phi, E, F = _recip_args(cell, xyz, q, mu)
    K = (ctypes.c_int * 3)(*grid)
    U = spme_lib().ora_spme_recip(P(cell), order, K, alpha, n, P(x), P(y), P(z), P(q), P(mu), P(phi), P(E), P(F))
    return dict(U=U, phi=phi, E=E.reshape(n, 3), F=F.reshape(n, 3))


def ewald_recip(cell, kmax, alpha, xyz, q, mu=None):
    """Ewald reciprocal sum over |m_a| <= kmax_a: dict(U, phi, E, F) as spme_recip."""
    cell, x, y, z, q, mu, n, phi, E, F = _recip_args(cell, xyz, q, mu)
    K = (ctypes.c_int * 3)(*kmax)
    U = spme_lib().ora_ewald_recip(P(cell), K, alpha, n, P(x), P(y), P(z), P(q), P(mu), P(phi), P(E), P(F))
    return dict(U=U, phi=phi, E=E.reshape(n, 3), F=F.reshape(n, 3))
