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

`large_box_run` is the large-box workflow: a thermostatted FP64 relaxation, then an NVE
run with the polynomials in FP32 and the heat current every step, its autocorrelation on
the device, and every `shadow_every` steps an FP64 evaluation of the same configuration
(the shadow) against which the run's forces, energy and J are compared.

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
    """extrapolate: the solver's start dipoles each step.  0: the previous step's converged
    dipoles, as the reference keeps them; 1: the line through the last two (2 mu1 - mu2); 2:
    the parabola through the last three (3 mu1 - 3 mu2 + mu3).  The solve runs to the same
    stopping test from any start, so the extrapolation only saves iterations; the dipoles
    differ from the warm-started ones within that test."""

    def __init__(self, slot, dt_fs, extrapolate=0):
        import torch
        self.t = torch
        self.slot = slot
        self.dt = dt_fs * 1.0e-3                   # ps
        if extrapolate not in (0, 1, 2):
            raise ValueError("extrapolate is 0, 1 or 2")
        self.extrapolate = extrapolate
        self._hist = []                            # converged dipoles of the last steps, newest first
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

    def _start_dipoles(self):
        h, t = self._hist, self.t
        k = min(self.extrapolate, len(h) - 1)
        if k == 1:
            t.sub(2.0 * h[0], h[1], out=self.dip)
        elif k == 2:
            t.add(3.0 * (h[0] - h[1]), h[2], out=self.dip)
        if k > 0:
            t.cuda.current_stream(self.slot.dev).synchronize()   # read by the library's stream

    def _remember(self):
        if self.extrapolate > 0:
            self._hist.insert(0, self.dip.clone())
            del self._hist[self.extrapolate + 1:]

    def _forces(self):
        s, b, c = self.slot, self.slot.box, self.slot.ctx
        x, y, z = self._ptrs()
        f = self.f
        self._start_dipoles()
        fp = (f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr())
        if s.fused:
            # one library call (graph replays): MB-pol and water forces written into f, the
            # electrostatics' into fsite; no heat flux here (J follows the second kick)
            o = c.evaluate_forces_dev(b.imcon, b.cell, 3 * b.nw, s.lst.data_ptr(), s.eparams, s.wparams, x, y, z,
                                      self.dip.data_ptr(), *fp, self.fsite.data_ptr())
            self.iterations = int(o[22])
            self._remember()
            return float(o[0]) + float(o[11]) + float(np.sum(o[23:27]))
        # MB-pol and the water terms add into f; the electrostatics write fsite (own stream)
        (e_el, _, _, self.iterations), ((e_mb, _, _), (e_w, _, _)) = s.split(
            lambda: c.polar_electrostatics_dev(b.imcon, b.cell, s.eparams, x, y, z, self.dip.data_ptr(),
                                               self.fsite.data_ptr()),
            lambda: (c.mbpol_forces_dev(b.imcon, b.cell, s.lst.data_ptr(), x, y, z, *fp),
                     c.water_terms_dev(b.imcon, b.cell, s.wparams, x, y, z, *fp)))
        self._remember()
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


def _copy_state(src, dst):
    """Positions, velocities and per-atom kinetic energies of slot src into slot dst (same
    size), device to device; both slots' streams are idle when this is called."""
    t = src.torch
    t.cuda.synchronize(src.dev)
    for a, b in ((src.x, dst.x), (src.y, dst.y), (src.z, dst.z), (src.v, dst.v), (src.ek, dst.ek)):
        b.copy_(a)
    t.cuda.synchronize(src.dev)


def large_box_run(box, relax_steps=500, nve_steps=10000, shadow_every=100, dt_fs=0.2, temperature=300.0,
                  tau_fs=20.0, recip=("spme", 6, (48, 48, 48)), ntau=None, progress=None, extrapolate=2):
    """The large-box workflow on one configuration (box: a waterbox.WaterBox).

    1. relax_steps steps of FP64 MD at dt_fs with Berendsen velocity scaling towards
       `temperature` (time constant tau_fs): lambda^2 = 1 + dt / tau (T0 / T - 1), applied
       to the velocities after each step, T from the kinetic energy over 3 N_atoms - 3
       degrees of freedom;
    2. nve_steps steps of NVE from the relaxed state with the 2B and 3B polynomials in FP32
       (the rest in FP64), the heat current every step;
    3. every shadow_every steps (and at step 0) the same positions and velocities evaluated
       in FP64 by a second slot: RMS and max force difference, energy difference per
       molecule and relative |J - J_FP64|;
    4. the autocorrelation of the run's J_tot for ntau lags on the device.
    The reciprocal space is `recip` (engine.Slot's argument; None: the Ewald sum over k);
    `extrapolate` is NVE's choice of start dipoles (2: the parabola through the last three).
    progress(stage, step, dict) is called every shadow_every steps.  Returns a dict of the
    series and summaries (internal units: energy 10 J/mol, J in (10 J/mol) A/ps)."""
    import time as _time
    from . import engine
    dt_ps = dt_fs * 1.0e-3
    nr = box.nreal
    ndf = 3 * nr - 3
    kb = 0.831446261815324          # Boltzmann's constant in 10 J/mol per K
    out = dict(waters=box.nw, relax_steps=relax_steps, nve_steps=nve_steps, shadow_every=shadow_every,
               dt_fs=dt_fs, recip=None if recip is None else list(recip))
    # 1. relaxation (FP64)
    fp64 = engine.Slot(box, precision=0, recip=recip, dedicated_queues=True)
    md = NVE(fp64, dt_fs, extrapolate=extrapolate)
    temps = []
    t0 = _time.perf_counter()
    for k in range(relax_steps):
        r = md.step()
        temp = 2.0 * r["kin"] / (ndf * kb)
        lam = np.sqrt(max(0.0, 1.0 + dt_fs / tau_fs * (temperature / temp - 1.0)))
        fp64.torch.cuda.synchronize(fp64.dev)
        fp64.v.mul_(lam)
        fp64.ek.mul_(lam * lam)
        fp64.torch.cuda.synchronize(fp64.dev)
        temps.append(temp)
        if progress and (k + 1) % shadow_every == 0:
            progress("relax", k + 1, dict(temperature=temp, pot=r["pot"]))
    out["relax_ms_per_step"] = 1e3 * (_time.perf_counter() - t0) / max(1, relax_steps)
    out["relax_temperature_last"] = temps[-1] if temps else None
    # 2.-3. mixed-precision NVE with the FP64 shadow
    mixed = engine.Slot(box, precision=3, recip=recip, dedicated_queues=True)
    _copy_state(fp64, mixed)
    shadow = fp64
    nve = NVE(mixed, dt_fs, extrapolate=extrapolate)
    e0, k0 = nve.pot + nve.kin, nve.kin
    etot, jser, checks, iters = [], [], [], []

    def shadow_check(step, jmix):
        _copy_state(mixed, shadow)
        rs = shadow.evaluate()
        fm, fs = mixed.forces(), shadow.forces()
        d = fm - fs
        checks.append(dict(step=step, rms_dF_kcal_mol_A=float(np.sqrt((d ** 2).mean())) / 418.4,
                           max_dF_kcal_mol_A=float(np.abs(d).max()) / 418.4,
                           dE_kcal_mol_per_molecule=(nve.pot - rs["eng"]) / 418.4 / box.nw,
                           rel_dJ=float(np.linalg.norm(jmix - rs["jtot"]) / np.linalg.norm(rs["jtot"]))))

    shadow_check(0, _current_j(nve))
    t0 = _time.perf_counter()
    for k in range(nve_steps):
        r = nve.step()
        etot.append(r["tot"])
        jser.append(r["jtot"])
        iters.append(r["iterations"])
        if (k + 1) % shadow_every == 0:
            t1 = _time.perf_counter()
            shadow_check(k + 1, r["jtot"])
            t0 += _time.perf_counter() - t1          # the shadow is not part of the run's time
            if progress:
                progress("nve", k + 1, checks[-1])
    out["nve_ms_per_step"] = 1e3 * (_time.perf_counter() - t0) / max(1, nve_steps)
    etot = np.array(etot)
    tt = dt_ps * np.arange(1, len(etot) + 1)
    slope = float(np.polyfit(tt, etot, 1)[0]) if len(etot) > 1 else 0.0
    out.update(drift_per_ps_over_KE0=slope / k0, max_abs_dE_over_KE0=float(np.abs(etot - e0).max() / k0),
               dipole_iterations_median=float(np.median(iters)) if iters else None, shadow=checks,
               max_rms_dF_kcal_mol_A=max(c["rms_dF_kcal_mol_A"] for c in checks),
               max_rel_dJ=max(c["rel_dJ"] for c in checks))
    jser = np.array(jser)
    if ntau is None:
        ntau = max(1, min(1000, len(jser) // 2))
    t1 = _time.perf_counter()
    acf = mixed.ctx.heat_acf(jser, ntau) if len(jser) else np.zeros(0)
    out["acf_ms"] = 1e3 * (_time.perf_counter() - t1)
    out["acf"] = acf
    out["jtot"] = jser
    out["etot"] = etot
    mixed.close()
    fp64.close()
    return out


def _current_j(nve):
    """J of the NVE run's current state (the library's heat flux of the last evaluation)."""
    s = nve.slot
    v = s.v
    _, _, jt = s.ctx.heat_flux_dev(v[0].data_ptr(), v[1].data_ptr(), v[2].data_ptr(), s.ek.data_ptr())
    return np.array(jt)
