"""NVE molecular dynamics on the device with the heat current every step.

`NVE(slot, dt_fs)` advances one configuration held by an engine.Slot: velocity Verlet on
the real atoms (mbgpu_verlet_stage), a complete force evaluation per step (MB-pol 2B/3B,
polarizable electrostatics with the induced dipoles started from the previous step's, as
the reference keeps them, the monomer / dispersion / long-range terms), per-atom kinetic
energies, and J = J_kin + J_pot from the library's heat flux.  Positions, velocities,
forces and dipoles stay in HBM; each step returns the scalars (potential and kinetic
energy, J).

`Heatflux2Writer` appends J records in the layout of write_heat_flux2
(heatcurrent.f90:730-740): the time as es12.5, then J_kin, J_pot and J_tot, each as
3es15.7, in internal units times junit = 1e3.

`green_kubo` turns a J series into the heat-current autocorrelation (on the device,
mbgpu_heat_acf) and the running Green-Kubo conductivity
kappa(t) = 1 / (3 V k_B T^2) * int_0^t <J(0) . J(s)> ds (trapezoid rule), the analysis the
reference leaves to post-processing of heatflux2.  The helper integrates whatever current it
is given: in water (several atomic species) J must first be freed of its correlation with
the non-conserved mass currents (multi-component Green-Kubo / cepstral analysis), and the
series must be long (tens of ps) for the integral to converge.
"""
import numpy as np

JUNIT = 1.0e3     # heatcurrent.f90:19
# SI value of the internal units: energy 10 J/mol per particle, velocity Angstrom/ps
KB_SI = 1.380649e-23
E_INT_SI = 10.0 / 6.02214076e23          # J
V_INT_SI = 1.0e-10 / 1.0e-12             # m/s


def fortran_es(v, w, d):
    """Fortran ESw.d edit descriptor for one value (1 digit before the point, a two-digit
    exponent with sign, or three digits without the letter E when |exponent| > 99)."""
    if not np.isfinite(v):
        return "*" * w
    s = "%.*E" % (d, v)
    mant, exp = s.split("E")
    e = int(exp)
    t = mant + ("E%+03d" % e if abs(e) <= 99 else "%+04d" % e)
    return t.rjust(w) if len(t) <= w else "*" * w


def heatflux2_record(time, jkin, jpot, jtot):
    """The four lines write_heat_flux2 writes for one time (J in internal units)."""
    lines = [fortran_es(time, 12, 5)]
    for j in (jkin, jpot, jtot):
        lines.append("".join(fortran_es(float(v) * JUNIT, 15, 7) for v in j))
    return "\n".join(lines) + "\n"


class Heatflux2Writer:
    def __init__(self, path):
        self.path = path

    def write(self, time, jkin, jpot, jtot):
        with open(self.path, "a") as f:
            f.write(heatflux2_record(time, jkin, jpot, jtot))


def check_record(time, tot_kin, tot_pot, stress9, engunit=418.4, strsunit=1.0):
    """One record of write_check's CHECK_ENERGY_STRESS file (heat_check.f90:33-35): the time
    (es12.5), total kinetic and potential energy / engunit (2es15.7) and the 9 components of
    total minus reciprocal-space per-atom stress * strsunit (9es15.7).  The inputs are those
    mbgpu_heat_check returns (internal units)."""
    s9 = [float(v) * strsunit for v in np.asarray(stress9).reshape(-1)]
    return (fortran_es(time, 12, 5) + "\n" + fortran_es(tot_kin / engunit, 15, 7) + fortran_es(tot_pot / engunit, 15, 7)
            + "\n" + "".join(fortran_es(v, 15, 7) for v in s9) + "\n")


class CheckWriter:
    """Appends check_record lines to a CHECK_ENERGY_STRESS file."""

    def __init__(self, path):
        self.path = path

    def write(self, time, tot_kin, tot_pot, stress9, engunit=418.4, strsunit=1.0):
        with open(self.path, "a") as f:
            f.write(check_record(time, tot_kin, tot_pot, stress9, engunit, strsunit))


class NVE:
    def __init__(self, slot, dt_fs):
        import torch
        self.t = torch
        self.slot = slot
        self.dt = dt_fs * 1.0e-3                   # ps
        box = slot.box
        nr, n = box.nreal, box.natms
        f64 = dict(dtype=torch.float64, device=slot.dev)
        self.mass = torch.as_tensor(np.ascontiguousarray(box.mass[:nr]), **f64)
        self.f = torch.zeros(3, n, **f64)          # MB-pol + water forces (SoA)
        self.fsite = torch.zeros(3 * n, **f64)     # electrostatics forces (3 per site)
        self.dip = torch.zeros(3 * n, **f64)       # induced dipoles, kept between steps
        torch.cuda.synchronize(slot.dev)
        self.time = 0.0
        self.nstep = 0
        self.pot = self._forces()
        self.kin = self._kinetic()
        torch.cuda.synchronize(slot.dev)          # ek is read by the library's stream

    def _ptrs(self):
        s = self.slot
        return s.x.data_ptr(), s.y.data_ptr(), s.z.data_ptr()

    def _forces(self):
        s, b, c = self.slot, self.slot.box, self.slot.ctx
        x, y, z = self._ptrs()
        f = self.f
        fp = (f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr())
        # MB-pol and the water terms add into f; the electrostatics write fsite (own stream)
        (e_el, _, _, self.iterations), ((e_mb, _, _), (e_w, _, _)) = s.split(
            lambda: c.polar_electrostatics_dev(b.imcon, b.cell, s.eparams, x, y, z, self.dip.data_ptr(),
                                               self.fsite.data_ptr()),
            lambda: (c.mbpol_forces_dev(b.imcon, b.cell, s.lst.data_ptr(), x, y, z, *fp),
                     c.water_terms_dev(b.imcon, b.cell, s.wparams, x, y, z, *fp)))
        return e_mb + e_el + float(np.sum(e_w))

    def _kinetic(self):
        v = self.slot.v
        ke = 0.5 * (self.mass * (v * v).sum(0))
        self.slot.ek.copy_(ke)
        return float(ke.sum())

    def step(self):
        """One time step; returns dict(time, pot, kin, tot, jkin, jpot, jtot)."""
        s, c = self.slot, self.slot.ctx
        x, y, z = self._ptrs()
        v, f = s.v, self.f
        vp = (v[0].data_ptr(), v[1].data_ptr(), v[2].data_ptr())
        fp = (f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr())
        c.verlet_stage(1, self.dt, self.mass.data_ptr(), x, y, z, *vp, *fp, self.fsite.data_ptr())
        self.pot = self._forces()
        self.kin = c.verlet_stage(2, self.dt, self.mass.data_ptr(), x, y, z, *vp, *fp, self.fsite.data_ptr(),
                                  s.ek.data_ptr())
        jk, jp, jt = c.heat_flux_dev(*vp, s.ek.data_ptr())
        self.nstep += 1
        self.time += self.dt
        return dict(time=self.time, pot=self.pot, kin=self.kin, tot=self.pot + self.kin, jkin=np.array(jk),
                    jpot=np.array(jp), jtot=np.array(jt), iterations=self.iterations)

    def positions(self):
        s = self.slot
        self.t.cuda.synchronize(s.dev)
        nr = s.box.nreal
        return np.stack([s.x[:nr].cpu().numpy(), s.y[:nr].cpu().numpy(), s.z[:nr].cpu().numpy()], 1)

    def velocities(self):
        self.t.cuda.synchronize(self.slot.dev)
        return self.slot.v.cpu().numpy().T.copy()


def kappa_from_acf(acf, dt_ps, volume_A3, temperature_K):
    """Running Green-Kubo integral in W/(m K) of an autocorrelation of J in internal units
    ((10 J/mol) Angstrom/ps per particle), sampled every dt_ps: kappa[k] = integral over
    [0, k dt] by the trapezoid rule."""
    c = np.asarray(acf, dtype=np.float64) * (E_INT_SI * V_INT_SI) ** 2
    dt = dt_ps * 1.0e-12
    integral = np.concatenate([[0.0], np.cumsum(0.5 * (c[1:] + c[:-1]) * dt)])
    return integral / (3.0 * volume_A3 * 1.0e-30 * KB_SI * temperature_K ** 2)


def green_kubo(ctx, jseries, dt_ps, volume_A3, temperature_K, ntau):
    """Autocorrelation of J (nt, 3) on the device for lags 0..ntau-1 and the running
    conductivity; returns dict(time_ps, acf, kappa_W_mK)."""
    acf = ctx.heat_acf(jseries, ntau)
    return dict(time_ps=dt_ps * np.arange(ntau), acf=acf,
                kappa_W_mK=kappa_from_acf(acf, dt_ps, volume_A3, temperature_K))
