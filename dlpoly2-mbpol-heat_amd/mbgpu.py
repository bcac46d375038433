"""ctypes binding of libmbgpu.so, the gfx950 MB-pol force / heat-current library.

Host-side mirror of the reference's interface for this path:

  * ``mbpol_2b_poly`` / ``mbpol_3b_poly`` / ``mbpol_2b_cutoff`` / ``mbpol_3b_cutoff``: the
    C++ leaves (mbpol-2b-poly.cpp:1878-1908, mbpol-3b-poly.cpp:1518-1546);
  * ``Context.mbpol_forces``: the Fortran ``mbpol_forces`` subroutine (mbpol.f90:98-449)
    with the same argument meaning and accumulate-into semantics;
  * ``Context.heat_flux``: ``compute_heat_flux2`` (heatcurrent.f90:703-728).

There is no CPU fallback: loading fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmbgpu.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None

# name -> (restype, argtypes); every function declared in include/mbgpu.h
_SIGS = {
    "mbpol_2b_poly_": (None, [_dp, _dp, _dp, _dp, _dp, _dp]),
    "mbpol_3b_poly_": (None, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mbpol_2b_cutoff_": (None, [_dp]),
    "mbpol_3b_cutoff_": (None, [_dp]),
    "mbpol_2b_poly": (None, [_dp, _dp, _dp, _dp, _dp, _dp]),
    "mbpol_3b_poly": (None, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mbpol_2b_cutoff": (None, [_dp]),
    "mbpol_3b_cutoff": (None, [_dp]),
    "mbgpu_create": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int]),
    "mbgpu_destroy": (None, [ctypes.c_void_p]),
    "mbgpu_create_queues": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "mbgpu_last_error": (ctypes.c_char_p, []),
    "mbgpu_check_listttm2": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int]),
    "mbgpu_mbpol_forces": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, _dp, _dp,
                                          _dp, _dp, _dp, _dp, _dp, _dp]),
    "mbgpu_mbpol_forces_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, _dp, _dp, _dp]),
    "mbgpu_get_decomposition": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "mbgpu_set_precision": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mbgpu_set_elec_reciprocal": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "mbgpu_set_graphs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mbgpu_get_graph_counts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "mbgpu_get_evaluate_graph_counts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "mbgpu_evaluate_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "mbgpu_evaluate_forces_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
                                  + [ctypes.c_void_p] * 16),
    "mbgpu_get_forces": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "mbgpu_get_lists": (ctypes.c_int, [ctypes.c_void_p, _ip, ctypes.c_long, ctypes.POINTER(ctypes.c_long), _ip,
                                       ctypes.c_long, ctypes.POINTER(ctypes.c_long)]),
    "mbgpu_heat_flux": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mbgpu_heat_flux_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _dp,
                                           _dp, _dp]),
    "mbgpu_time_forces": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp]),
    "mbgpu_time_elec": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp]),
    "mbgpu_polar_electrostatics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_void_p, _dp, _dp,
                                                  _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]),
    "mbgpu_polar_electrostatics_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_void_p,
                                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                      ctypes.c_void_p, ctypes.c_void_p, _dp, _dp, _dp, _ip]),
    "mbgpu_get_elec_details": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mbgpu_get_dipoles": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "mbgpu_heat_acf": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_long, ctypes.c_int, _dp]),
    "mbgpu_ring_matrices": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]),
    "mbgpu_ring_transform_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_long, _dp, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "mbgpu_ring_spring_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double)]),
    "mbgpu_ring_spring_dev_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                     ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.POINTER(ctypes.c_double),
                                                     ctypes.POINTER(ctypes.c_double)]),
    "mbgpu_spme_recip_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_double, ctypes.c_int] + [ctypes.c_void_p] * 8
                             + [ctypes.POINTER(ctypes.c_double)]),
    "mbgpu_heat_acf_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                                          ctypes.c_void_p]),
    "mbgpu_verlet_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double] + [ctypes.c_void_p] * 12
                           + [_dp]),
    "mbgpu_water_terms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_void_p, _dp, _dp, _dp, _dp,
                                         _dp, _dp, _dp, _dp, _dp]),
    "mbgpu_water_terms_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, _dp, _dp, _dp]),
    "mbgpu_get_water_details": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "mbgpu_heat_check": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _dp]),
}


class ElecParams(ctypes.Structure):
    """mbgpu_elec_params (include/mbgpu.h)."""
    _fields_ = [("rcut", ctypes.c_double), ("alpha", ctypes.c_double), ("kmax", ctypes.c_int * 3),
                ("toler", ctypes.c_double), ("omega", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("gamma", ctypes.c_double), ("q_o", ctypes.c_double), ("polr", ctypes.c_double * 3),
                ("polr2", ctypes.c_double * 3), ("athole", ctypes.c_double), ("athole12", ctypes.c_double),
                ("athole13", ctypes.c_double), ("ascc", ctypes.c_double), ("ascd", ctypes.c_double),
                ("compat_qd_column", ctypes.c_int), ("distribute_msite", ctypes.c_int), ("solver", ctypes.c_int)]

    @classmethod
    def from_params(cls, p, distribute_msite=1, solver=None):
        """From an elec_params.ElecParams (or any object with the same attributes)."""
        e = cls()
        e.rcut, e.alpha = p.rcut, p.alpha
        for i in range(3):
            e.kmax[i] = p.kmax[i]
            e.polr[i] = p.polr[i]
            e.polr2[i] = p.polr2[i]
        e.toler, e.omega, e.max_iter = p.toler, p.omega, p.max_iter
        e.gamma, e.q_o = p.gamma, p.q_o
        e.athole, e.athole12, e.athole13, e.ascc, e.ascd = p.athole, p.athole12, p.athole13, p.ascc, p.ascd
        e.compat_qd_column = p.compat_qd_column
        e.distribute_msite = distribute_msite
        e.solver = getattr(p, "solver", 0) if solver is None else solver
        return e


class WaterParams(ctypes.Structure):
    """mbgpu_water_params (include/mbgpu.h)."""
    _fields_ = [("rcut", ctypes.c_double), ("rvdw", ctypes.c_double), ("c6", ctypes.c_double * 3),
                ("a6", ctypes.c_double * 3), ("c8", ctypes.c_double * 3), ("a8", ctypes.c_double * 3),
                ("bond_key", ctypes.c_int * 2), ("bond_prm", (ctypes.c_double * 3) * 2),
                ("ps_monomer", ctypes.c_int), ("dispersion", ctypes.c_int), ("long_range", ctypes.c_int)]

    @classmethod
    def from_params(cls, p):
        """From a water_params.WaterParams (energies converted to the internal unit)."""
        w = cls()
        w.rcut, w.rvdw = p.rcut, p.rvdw
        c6, a6, c8, a8 = p.internal()
        for i in range(3):
            w.c6[i], w.a6[i], w.c8[i], w.a8[i] = c6[i], a6[i], c8[i], a8[i]
        for t, (key, prm) in enumerate(p.bonds_internal()):
            w.bond_key[t] = key
            for q in range(3):
                w.bond_prm[t][q] = prm[q]
        w.ps_monomer, w.dispersion, w.long_range = p.ps_monomer, p.dispersion, p.long_range
        return w


class MbgpuError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Load libmbgpu.so (built by __graft_entry__.build()); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MbgpuError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    try:
        # PyTorch ships its own HIP runtime under the same soname (libamdhip64.so.7).  Loaded
        # first, it is the one libmbgpu.so binds to, so the process holds a single runtime;
        # loaded second, it would bring a second runtime whose initialisation then fails.
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _check(rc, what):
    if rc != 0:
        raise MbgpuError(f"{what} failed ({rc}): {load_library().mbgpu_last_error().decode()}")


# ------------------------------------------------------------------------ leaves ----
def mbpol_2b_cutoff():
    r = ctypes.c_double()
    load_library().mbpol_2b_cutoff_(ctypes.byref(r))
    return r.value


def mbpol_3b_cutoff():
    r = ctypes.c_double()
    load_library().mbpol_3b_cutoff_(ctypes.byref(r))
    return r.value


def mbpol_2b_poly(w1, w2, g1=None, g2=None):
    """E (kcal/mol), g1, g2 (overwritten unless r_OO > 6.5, as the reference leaf)."""
    w1, w2 = _f64(w1, 9), _f64(w2, 9)
    g1 = np.zeros(9) if g1 is None else g1
    g2 = np.zeros(9) if g2 is None else g2
    e = ctypes.c_double()
    fm = np.zeros(108)
    load_library().mbpol_2b_poly_(_p(w1), _p(w2), ctypes.byref(e), _p(g1), _p(g2), _p(fm))
    return e.value, g1, g2


def mbpol_3b_poly(w1, w2, w3, g1=None, g2=None, g3=None):
    """E (kcal/mol); g1..g3 are accumulated into (zeros if not given), as the reference."""
    w1, w2, w3 = _f64(w1, 9), _f64(w2, 9), _f64(w3, 9)
    g1 = np.zeros(9) if g1 is None else g1
    g2 = np.zeros(9) if g2 is None else g2
    g3 = np.zeros(9) if g3 is None else g3
    e = ctypes.c_double()
    fm = np.zeros(243)
    load_library().mbpol_3b_poly_(_p(w1), _p(w2), _p(w3), ctypes.byref(e), _p(g1), _p(g2), _p(g3), _p(fm))
    return e.value, g1, g2, g3


# ----------------------------------------------------------------------- context ----
def check_listttm2(listttm2, natms):
    """mbgpu_check_listttm2: 0 when every water's entries are consecutive atom ids a, a+1,
    a+2 within 1..natms, else the 1-based index of the first bad water (-1: size not a
    multiple of 3).  Runs on the host; the device entries apply the same predicate."""
    a = np.ascontiguousarray(listttm2, dtype=np.int32)
    return load_library().mbgpu_check_listttm2(a.size, a.ctypes.data_as(_ip), natms)


def ring_matrices(nbead, langevin=0):
    """Normal-mode matrices of transformation_matrix (mbgpu_ring_matrices): (nmode_to_cart,
    cart_to_nmode, tcart_to_tnmode, lambda), each matrix [k][b] = M(k+1, b+1)."""
    P = int(nbead)
    a, b, c, lam = np.zeros(P * P), np.zeros(P * P), np.zeros(P * P), np.zeros(P)
    _check(load_library().mbgpu_ring_matrices(P, int(langevin), _p(a), _p(b), _p(c), _p(lam)),
           "mbgpu_ring_matrices")
    return a.reshape(P, P), b.reshape(P, P), c.reshape(P, P), lam


class Context:
    """One device context for a system of nw waters and natms sites.  dedicated_queues: its
    two streams get hardware queues of their own (mbgpu_create_queues), for one
    latency-bound context; the default shares the process's queues (many contexts)."""

    def __init__(self, nw, natms, dedicated_queues=False):
        lib = load_library()
        self.nw, self.natms = nw, natms
        self.dedicated_queues = bool(dedicated_queues)
        self._h = lib.mbgpu_create_queues(nw, natms, int(self.dedicated_queues))
        if not self._h:
            raise MbgpuError("mbgpu_create failed: " + lib.mbgpu_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            load_library().mbgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mbpol_forces(self, imcon, cell, nttm2, listttm2, xxx, yyy, zzz, fxx, fyy, fzz, engsrp=0.0, virsrp=0.0,
                     stress=None):
        """mbpol_forces (mbpol.f90:98-449): fxx/fyy/fzz and stress accumulate in place;
        returns (engsrp, virsrp) with this call's energy and virial added."""
        cell = _f64(cell, 9)
        lst = np.ascontiguousarray(listttm2, dtype=np.int32)
        x, y, z = _f64(xxx, self.natms), _f64(yyy, self.natms), _f64(zzz, self.natms)
        for f in (fxx, fyy, fzz):
            if f.dtype != np.float64 or not f.flags.c_contiguous or f.size != self.natms:
                raise ValueError("force arrays must be contiguous float64 of length natms")
        stress = np.zeros(9) if stress is None else stress
        e, v = ctypes.c_double(engsrp), ctypes.c_double(virsrp)
        rc = load_library().mbgpu_mbpol_forces(self._h, imcon, _p(cell), nttm2, lst.ctypes.data_as(_ip), _p(x),
                                               _p(y), _p(z), _p(fxx), _p(fyy), _p(fzz), ctypes.byref(e),
                                               ctypes.byref(v), _p(stress))
        _check(rc, "mbgpu_mbpol_forces")
        return e.value, v.value

    def mbpol_forces_dev(self, imcon, cell, list_ptr, x_ptr, y_ptr, z_ptr, fx_ptr=None, fy_ptr=None, fz_ptr=None):
        """Device-pointer variant (ints are raw device addresses, e.g. torch data_ptr())."""
        cell = _f64(cell, 9)
        e, v = ctypes.c_double(0.0), ctypes.c_double(0.0)
        stress = np.zeros(9)
        rc = load_library().mbgpu_mbpol_forces_dev(self._h, imcon, _p(cell), 3 * self.nw, list_ptr, x_ptr, y_ptr,
                                                   z_ptr, fx_ptr, fy_ptr, fz_ptr, ctypes.byref(e), ctypes.byref(v),
                                                   _p(stress))
        _check(rc, "mbgpu_mbpol_forces_dev")
        return e.value, v.value, stress

    def set_precision(self, mode):
        """Bit mask: 0 = FP64; bit 0 / bit 1 = the 2B / 3B polynomial in FP32 (FP64 accumulation)."""
        _check(load_library().mbgpu_set_precision(self._h, int(mode)), "mbgpu_set_precision")

    def set_elec_reciprocal(self, mode, order=6, grid=(32, 32, 32)):
        """Reciprocal space of the electrostatics: mode 0 / "ewald" = the Ewald sum over k
        (ewald1p.f); mode 1 / "spme" = particle-mesh Ewald of charges and dipoles with
        B-splines of the given even order on the given grid (sizes 2^a 3^b 5^c)."""
        m = {"ewald": 0, "spme": 1}.get(mode, mode)
        g = (ctypes.c_int * 3)(*[int(k) for k in grid])
        _check(load_library().mbgpu_set_elec_reciprocal(self._h, int(m), int(order), g), "mbgpu_set_elec_reciprocal")

    def set_graphs(self, on):
        """Graph replay of the electrostatics: True (default) = capture one evaluation's kernels
        once the iteration count is known and replay them; False = launch kernel by kernel."""
        _check(load_library().mbgpu_set_graphs(self._h, 1 if on else 0), "mbgpu_set_graphs")

    def evaluate_dev(self, imcon, cell, nttm2, list_ptr, eparams, wparams, x_ptr, y_ptr, z_ptr, dip_ptr, vx_ptr,
                     vy_ptr, vz_ptr, ek_ptr):
        """mbgpu_evaluate_dev: one complete evaluation (MB-pol, electrostatics, water terms, heat
        flux) of device arrays; dip_ptr may be None (dipoles from zero).  Returns the 64-value
        output vector (see the header for its layout)."""
        out = np.zeros(64)
        cell = _f64(cell, 9)
        rc = load_library().mbgpu_evaluate_dev(self._h, int(imcon), _p(cell), int(nttm2), list_ptr,
                                               ctypes.byref(eparams), ctypes.byref(wparams), x_ptr, y_ptr, z_ptr,
                                               dip_ptr, vx_ptr, vy_ptr, vz_ptr, ek_ptr, _p(out))
        _check(rc, "mbgpu_evaluate_dev")
        return out

    def evaluate_forces_dev(self, imcon, cell, nttm2, list_ptr, eparams, wparams, x_ptr, y_ptr, z_ptr, dip_ptr,
                            fx_ptr, fy_ptr, fz_ptr, fsite_ptr, vx_ptr=None, vy_ptr=None, vz_ptr=None, ek_ptr=None):
        """mbgpu_evaluate_forces_dev: evaluate_dev that also WRITES the forces (MB-pol and water
        terms into fx/fy/fz, natms each; electrostatics into fsite, 3 per site); without
        velocities no heat flux (out[40:49] zero).  Returns the 64-value output vector."""
        out = np.zeros(64)
        cell = _f64(cell, 9)
        rc = load_library().mbgpu_evaluate_forces_dev(self._h, int(imcon), _p(cell), int(nttm2), list_ptr,
                                                      ctypes.byref(eparams), ctypes.byref(wparams), x_ptr, y_ptr,
                                                      z_ptr, dip_ptr, fx_ptr, fy_ptr, fz_ptr, fsite_ptr, vx_ptr,
                                                      vy_ptr, vz_ptr, ek_ptr, _p(out))
        _check(rc, "mbgpu_evaluate_forces_dev")
        return out

    def total_forces(self):
        """Total forces of the last complete evaluation on the real atoms, (3*nw, 3):
        (MB-pol + electrostatics) + water terms, summed on the device (mbgpu_get_forces)."""
        f = np.zeros((3 * self.nw, 3))
        _check(load_library().mbgpu_get_forces(self._h, _p(f)), "mbgpu_get_forces")
        return f

    def evaluate_graph_counts(self):
        """(complete evaluations captured as graphs, replays) in this context."""
        c = (ctypes.c_long * 2)()
        _check(load_library().mbgpu_get_evaluate_graph_counts(self._h, c), "mbgpu_get_evaluate_graph_counts")
        return int(c[0]), int(c[1])

    def graph_counts(self):
        """(graphs captured, replays launched) by this context's electrostatics."""
        c = (ctypes.c_long * 2)()
        _check(load_library().mbgpu_get_graph_counts(self._h, c), "mbgpu_get_graph_counts")
        return int(c[0]), int(c[1])

    def decomposition(self):
        ae = np.zeros(self.natms)
        hw = np.zeros(9 * self.natms)
        f = np.zeros(3 * self.natms)
        _check(load_library().mbgpu_get_decomposition(self._h, _p(ae), _p(hw), _p(f)), "mbgpu_get_decomposition")
        return ae, hw.reshape(self.natms, 3, 3), f.reshape(self.natms, 3)

    def lists(self):
        nd, nt = ctypes.c_long(), ctypes.c_long()
        lib = load_library()
        _check(lib.mbgpu_get_lists(self._h, None, 0, ctypes.byref(nd), None, 0, ctypes.byref(nt)), "lists")
        d = np.zeros((max(nd.value, 1), 2), dtype=np.int32)
        t = np.zeros((max(nt.value, 1), 3), dtype=np.int32)
        _check(lib.mbgpu_get_lists(self._h, d.ctypes.data_as(_ip), nd.value, ctypes.byref(nd),
                                   t.ctypes.data_as(_ip), nt.value, ctypes.byref(nt)), "lists")
        return d[:nd.value], t[:nt.value]

    def heat_flux(self, vxx, vyy, vzz, e_kin, e_pot_extra=None, w_extra=None, stress_ew1p=None):
        """compute_heat_flux2 (heatcurrent.f90:703-728) over the real atoms: (jkin, jpot, jtot)."""
        nr = 3 * self.nw
        vx, vy, vz = _f64(vxx)[:nr].copy(), _f64(vyy)[:nr].copy(), _f64(vzz)[:nr].copy()
        ek = _f64(e_kin)[:nr].copy()
        ep = None if e_pot_extra is None else _f64(e_pot_extra)[:nr].copy()
        we = None if w_extra is None else _f64(w_extra).reshape(-1)[:9 * nr].copy()
        se = None if stress_ew1p is None else _f64(stress_ew1p).reshape(-1)[:9 * nr].copy()
        jk, jp, jt = np.zeros(3), np.zeros(3), np.zeros(3)
        rc = load_library().mbgpu_heat_flux(self._h, _p(vx), _p(vy), _p(vz), _p(ek), _p(ep), _p(we), _p(se),
                                            _p(jk), _p(jp), _p(jt))
        _check(rc, "mbgpu_heat_flux")
        return jk, jp, jt

    def heat_flux_dev(self, vx_ptr, vy_ptr, vz_ptr, ekin_ptr, epot_ptr=None, w_ptr=None, sew_ptr=None):
        """Device-pointer variant of heat_flux (raw device addresses)."""
        jk, jp, jt = np.zeros(3), np.zeros(3), np.zeros(3)
        rc = load_library().mbgpu_heat_flux_dev(self._h, vx_ptr, vy_ptr, vz_ptr, ekin_ptr, epot_ptr, w_ptr, sew_ptr,
                                                _p(jk), _p(jp), _p(jt))
        _check(rc, "mbgpu_heat_flux_dev")
        return jk, jp, jt

    def time_forces(self, reps):
        ms, k3 = ctypes.c_double(), ctypes.c_double()
        _check(load_library().mbgpu_time_forces(self._h, reps, ctypes.byref(ms), ctypes.byref(k3)), "time_forces")
        return ms.value, k3.value

    def time_elec(self, reps):
        """(ms per launch, FLOP per launch) of the dipole structure-factor kernel."""
        ms, fl = ctypes.c_double(), ctypes.c_double()
        _check(load_library().mbgpu_time_elec(self._h, reps, ctypes.byref(ms), ctypes.byref(fl)), "time_elec")
        return ms.value, fl.value

    # ------------------------------------------------------------ electrostatics ----
    def polar_electrostatics(self, imcon, cell, params, xxx, yyy, zzz, dipx, dipy, dipz, fxx=None, fyy=None,
                             fzz=None, distribute_msite=1):
        """Polarizable Ewald electrostatics (forces.f:268-404 and the routines it calls):
        xxx..zzz (natms >= 4*nw; M entries overwritten), dipx..dipz (in/out) and fxx..fzz
        (+=) are modified in place.  Returns dict(eng, vir, stress, niter, chge)."""
        n = 4 * self.nw
        arrs = [xxx, yyy, zzz, dipx, dipy, dipz] + [a for a in (fxx, fyy, fzz) if a is not None]
        for a in arrs:
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.size < n:
                raise ValueError("position, dipole and force arrays must be contiguous float64 of >= 4*nw values")
        cell = _f64(cell, 9)
        p = params if isinstance(params, ElecParams) else ElecParams.from_params(params, distribute_msite)
        chge = np.zeros(n)
        e, v, it = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_int(0)
        stress = np.zeros(9)
        rc = load_library().mbgpu_polar_electrostatics(self._h, imcon, _p(cell), ctypes.byref(p), _p(xxx), _p(yyy),
                                                       _p(zzz), _p(dipx), _p(dipy), _p(dipz), _p(chge), _p(fxx),
                                                       _p(fyy), _p(fzz), ctypes.byref(e), ctypes.byref(v),
                                                       _p(stress), ctypes.byref(it))
        _check(rc, "mbgpu_polar_electrostatics")
        return dict(eng=e.value, vir=v.value, stress=stress, niter=it.value, chge=chge)

    def polar_electrostatics_dev(self, imcon, cell, params, x_ptr, y_ptr, z_ptr, dip_ptr, f_ptr=None,
                                 distribute_msite=1):
        """Device-pointer variant: dip_ptr holds 3 interleaved values per site (in/out)."""
        cell = _f64(cell, 9)
        p = params if isinstance(params, ElecParams) else ElecParams.from_params(params, distribute_msite)
        e, v, it = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_int(0)
        stress = np.zeros(9)
        rc = load_library().mbgpu_polar_electrostatics_dev(self._h, imcon, _p(cell), ctypes.byref(p), x_ptr, y_ptr,
                                                           z_ptr, dip_ptr, f_ptr, ctypes.byref(e), ctypes.byref(v),
                                                           _p(stress), ctypes.byref(it))
        _check(rc, "mbgpu_polar_electrostatics_dev")
        return e.value, v.value, stress, it.value

    # --------------------------------------------- monomer, dispersion, long range ----
    def water_terms(self, imcon, cell, params, xxx, yyy, zzz, fxx=None, fyy=None, fzz=None):
        """Monomer bonds, PS coupling terms, dispersion and long-range correction
        (bndfrc, angfrc, srfrce2, lrcorrect): fxx..fzz (+=) modified in place.
        Returns dict(eng_bnd, eng_ang, eng_srp, elrc, vir_*, virlrc, stress)."""
        nr = 3 * self.nw
        for a in [xxx, yyy, zzz] + [a for a in (fxx, fyy, fzz) if a is not None]:
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.size < nr:
                raise ValueError("position and force arrays must be contiguous float64 of >= 3*nw values")
        cell = _f64(cell, 9)
        p = params if isinstance(params, WaterParams) else WaterParams.from_params(params)
        eng, vir, stress = np.zeros(4), np.zeros(4), np.zeros(9)
        rc = load_library().mbgpu_water_terms(self._h, imcon, _p(cell), ctypes.byref(p), _p(xxx), _p(yyy), _p(zzz),
                                              _p(fxx), _p(fyy), _p(fzz), _p(eng), _p(vir), _p(stress))
        _check(rc, "mbgpu_water_terms")
        return dict(eng_bnd=eng[0], eng_ang=eng[1], eng_srp=eng[2], elrc=eng[3], vir_bnd=vir[0], vir_ang=vir[1],
                    vir_srp=vir[2], virlrc=vir[3], stress=stress)

    def water_terms_dev(self, imcon, cell, params, x_ptr, y_ptr, z_ptr, fx_ptr=None, fy_ptr=None, fz_ptr=None):
        """Device-pointer variant; returns (eng[4], vir[4], stress[9])."""
        cell = _f64(cell, 9)
        p = params if isinstance(params, WaterParams) else WaterParams.from_params(params)
        eng, vir, stress = np.zeros(4), np.zeros(4), np.zeros(9)
        rc = load_library().mbgpu_water_terms_dev(self._h, imcon, _p(cell), ctypes.byref(p), x_ptr, y_ptr, z_ptr,
                                                  fx_ptr, fy_ptr, fz_ptr, _p(eng), _p(vir), _p(stress))
        _check(rc, "mbgpu_water_terms_dev")
        return eng, vir, stress

    def water_details(self):
        """Per-atom energy, W and forces of the last water-terms call."""
        nr = 3 * self.nw
        e, w, f = np.zeros(nr), np.zeros(9 * nr), np.zeros(3 * nr)
        _check(load_library().mbgpu_get_water_details(self._h, _p(e), _p(w), _p(f)), "mbgpu_get_water_details")
        return dict(atomic_energy=e, heat_w=w.reshape(nr, 3, 3), f=f.reshape(nr, 3))

    def heat_check(self, ekin_ptr=None):
        """write_check's invariants (heat_check.f90:13-49): (sum e_kin, sum e_pot, sum e_ew1p,
        sum of the per-atom stresses minus the reciprocal ones as a 3x3), internal units."""
        out = np.zeros(12)
        _check(load_library().mbgpu_heat_check(self._h, ekin_ptr, _p(out)), "mbgpu_heat_check")
        return out[0], out[1], out[2], out[3:].reshape(3, 3)

    def verlet_stage(self, stage, dt, mass, x, y, z, vx, vy, vz, fx=None, fy=None, fz=None, fsite=None, ekin=None):
        """mbgpu_verlet_stage with device pointers; returns the kinetic energy (stage 2)."""
        ke = ctypes.c_double(0.0)
        _check(load_library().mbgpu_verlet_stage(self._h, stage, dt, mass, x, y, z, vx, vy, vz, fx, fy, fz, fsite,
                                                 ekin, ctypes.byref(ke)), "mbgpu_verlet_stage")
        return ke.value

    def heat_acf(self, j, ntau):
        """Autocorrelation of a current series j (nt, 3) for lags 0..ntau-1 (mbgpu_heat_acf)."""
        jj = np.ascontiguousarray(j, dtype=np.float64).reshape(-1, 3)
        acf = np.zeros(int(ntau))
        _check(load_library().mbgpu_heat_acf(self._h, _p(jj), len(jj), int(ntau), _p(acf)), "mbgpu_heat_acf")
        return acf

    def heat_acf_dev(self, j_ptr, nt, ntau, acf_ptr):
        """mbgpu_heat_acf_dev with device pointers ([nt][3] in, [ntau] out)."""
        _check(load_library().mbgpu_heat_acf_dev(self._h, j_ptr, int(nt), int(ntau), acf_ptr), "mbgpu_heat_acf_dev")

    def ring_transform_dev(self, nbead, ncoord, M, in_ptr, out_ptr):
        """mbgpu_ring_transform_dev: out[k] = sum_b M[k][b] in[b] over bead-major device arrays."""
        m = np.ascontiguousarray(M, dtype=np.float64).reshape(-1)
        if m.size != nbead * nbead:
            raise MbgpuError("the bead matrix must be nbead x nbead")
        _check(load_library().mbgpu_ring_transform_dev(self._h, int(nbead), int(ncoord), _p(m), in_ptr, out_ptr),
               "mbgpu_ring_transform_dev")

    def ring_spring_dev(self, nbead, natom, coeff, fac3, force_scale, mass_ptr, x_ptr, xc_ptr, f_ptr):
        """mbgpu_ring_spring_dev; returns (Epot_spring, Epot_deriv)."""
        es, ed = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(load_library().mbgpu_ring_spring_dev(self._h, int(nbead), int(natom), float(coeff), float(fac3),
                                                    float(force_scale), mass_ptr, x_ptr, xc_ptr, f_ptr,
                                                    ctypes.byref(es), ctypes.byref(ed)), "mbgpu_ring_spring_dev")
        return es.value, ed.value

    def ring_spring_dev_noalloc(self, nbead, natom, coeff, fac3, force_scale, mass_ptr, x_ptr, xc_ptr, f_ptr):
        """mbgpu_ring_spring_dev_noalloc (no allocation, fixed per-GPU scratch); returns
        (Epot_spring, Epot_deriv)."""
        es, ed = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(load_library().mbgpu_ring_spring_dev_noalloc(self._h, int(nbead), int(natom), float(coeff), float(fac3),
                                                            float(force_scale), mass_ptr, x_ptr, xc_ptr, f_ptr,
                                                            ctypes.byref(es), ctypes.byref(ed)),
               "mbgpu_ring_spring_dev_noalloc")
        return es.value, ed.value

    def spme_recip_dev(self, cell, order, grid, alpha, n, x_ptr, y_ptr, z_ptr, q_ptr, mu_ptr=None, phi_ptr=None,
                       efield_ptr=None, f_ptr=None):
        """mbgpu_spme_recip_dev with device pointers; returns the reciprocal energy (Coulomb
        constant 1).  mu, phi, E and F are [n][3] / [n] double arrays or None."""
        c = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
        g = (ctypes.c_int * 3)(*[int(k) for k in grid])
        u = ctypes.c_double(0.0)
        _check(load_library().mbgpu_spme_recip_dev(self._h, _p(c), int(order), g, float(alpha), int(n), x_ptr, y_ptr,
                                                   z_ptr, q_ptr, mu_ptr, phi_ptr, efield_ptr, f_ptr, ctypes.byref(u)),
               "mbgpu_spme_recip_dev")
        return u.value

    def dipoles(self):
        """Converged induced dipoles of the last electrostatics call, (4*nw, 3)."""
        d = np.zeros(12 * self.nw)
        _check(load_library().mbgpu_get_dipoles(self._h, _p(d)), "mbgpu_get_dipoles")
        return d.reshape(4 * self.nw, 3)

    def elec_details(self):
        """Results of the last electrostatics call (see mbgpu_get_elec_details)."""
        n, nr = 4 * self.nw, 3 * self.nw
        scal = np.zeros(20)
        chge, vesp = np.zeros(n), np.zeros(n)
        e, w, s, f = np.zeros(nr), np.zeros(9 * nr), np.zeros(9 * nr), np.zeros(3 * n)
        _check(load_library().mbgpu_get_elec_details(self._h, _p(scal), _p(chge), _p(vesp), _p(e), _p(w), _p(s),
                                                     _p(f)), "mbgpu_get_elec_details")
        names = ["eng_ew1p", "eng_ew2p", "eng_ew3p", "eng_polar", "eng", "vir", "vir_ew1p", "vir_qd", "niter",
                 "last_delta"]
        out = {k: scal[i] for i, k in enumerate(names)}
        out.update(stress=scal[10:19].copy(), k_vectors=scal[19], chge=chge, vesp=vesp, atomic_energy=e, heat_w=w.reshape(nr, 3, 3),
                   s_ew1p=s.reshape(nr, 9), f_sites=f.reshape(n, 3))
        return out
