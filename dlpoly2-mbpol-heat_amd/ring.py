"""In-process ring polymer: beads, normal modes and springs on one GPU.

The reference runs path-integral MD with one group of MPI ranks per bead, normal-mode
matrices from transformation_matrix (centroid_matrix.f:3-135) and spring forces from
centroid_spring_force.f:3-157, the bead coordinates travelling around the ring by MPI.
Here every bead is an engine.Slot (one library context with its own stream, so the beads'
force evaluations are independent configurations) whose positions, velocities and forces
are views of three bead-major device arrays [nbead][3][natms]; the transforms and springs run
on those arrays through mbgpu_ring_transform_dev and mbgpu_ring_spring_dev.

`RingPolymer.forces()` gives, per bead, the physical forces of the complete MB-pol +
electrostatics + water-term evaluation plus the spring forces, and returns the bead
potentials, the spring energy (Epot_spring) and the centroid-virial term (Epot_deriv) as
the reference defines them.  `RingPolymer.step()` advances the ring by velocity Verlet with
the physical masses (the microcanonical ring-polymer Hamiltonian
H = sum_k [K_k + V(x_k)] + Epot_spring, without the reference's thermostats, fictitious
normal-mode masses or barostat, which stay out of scope); `energy()` returns H.

Units are the reference's internal ones; hbar and k_B as centroid_module.f90:19-21.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import elec_params as _ep
from . import water_params as _wp
from .engine import Slot
from .mbgpu import ring_matrices

HBAR = 6.350780668      # centroid_module.f90:19, 10 J/mol ps
KB = 0.831451115        # centroid_module.f90:21, 10 J/mol / K


def spring_coefficient(nbead, temperature):
    """coeff = nbead / (hbar beta)^2 (centroid_spring_force.f:43)."""
    beta = 1.0 / (KB * temperature)
    return nbead / (HBAR * beta) ** 2


class RingPolymer:
    def __init__(self, boxes, temperature, elec=None, water=None, langevin=False, solver=1, force_scale=1.0):
        import torch
        self.t = torch
        if len(boxes) < 2 or len(boxes) % 2:
            raise ValueError("a ring polymer needs an even number of beads, at least 2")
        b0 = boxes[0]
        if any(b.nw != b0.nw for b in boxes):
            raise ValueError("every bead holds the same waters")
        self.P, self.temperature = len(boxes), float(temperature)
        self.nw, self.natms, self.nr = b0.nw, b0.natms, b0.nreal
        self.langevin, self.force_scale = bool(langevin), float(force_scale)
        self.fac1 = 1.0 / np.sqrt(self.P) if langevin else 1.0          # centroid_spring_force.f:56-63
        self.fac3 = self.force_scale
        self.coeff = spring_coefficient(self.P, self.temperature)
        self.n2c, self.c2n, self.tc2n, self.lam = ring_matrices(self.P, int(langevin))
        elec = elec or _ep.for_box(b0.L)
        water = water or _wp.for_box(b0.L)
        self.slots = [Slot(b, solver=solver, elec=elec, water=water) for b in boxes]
        self.dev = self.slots[0].dev
        f64 = dict(dtype=torch.float64, device=self.dev)
        P, n = self.P, self.natms
        self.X = torch.zeros(P, 3, n, **f64)          # bead positions (M sites rebuilt per evaluation)
        self.V = torch.zeros(P, 3, n, **f64)          # velocities (real atoms)
        self.F = torch.zeros(P, 3, n, **f64)          # physical + spring forces
        self.Q = torch.zeros(P, 3, n, **f64)          # normal-mode positions
        self.fsite = torch.zeros(P, 3 * n, **f64)     # electrostatics forces per site
        self.dip = torch.zeros(P, 3 * n, **f64)       # induced dipoles, kept between steps
        self.ek = torch.zeros(P, self.nr, **f64)
        m = np.zeros(n)
        m[:self.nr] = b0.mass[:self.nr]
        self.mass_all = torch.as_tensor(m, **f64)     # M sites: no mass, no spring
        self.mass = self.mass_all[:self.nr].contiguous()
        for k, (s, b) in enumerate(zip(self.slots, boxes)):
            self.X[k, 0].copy_(s.x)
            self.X[k, 1].copy_(s.y)
            self.X[k, 2].copy_(s.z)
            s.x, s.y, s.z = self.X[k, 0], self.X[k, 1], self.X[k, 2]
            self.V[k, :, :self.nr].copy_(torch.as_tensor(np.ascontiguousarray(b.vel[:self.nr].T), **f64))
        torch.cuda.synchronize(self.dev)
        self.pot = None
        self._pool = None
        self.forces()

    # ---------------------------------------------------------------- forces ----
    def _bead_forces(self, k):
        s, b, c = self.slots[k], self.slots[k].box, self.slots[k].ctx
        x, y, z = s.x.data_ptr(), s.y.data_ptr(), s.z.data_ptr()
        f = self.F[k]
        fp = (f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr())
        if s.fused:
            # one library call per bead (graph replays), forces written into F[k] and fsite[k]
            o = c.evaluate_forces_dev(b.imcon, b.cell, 3 * b.nw, s.lst.data_ptr(), s.eparams, s.wparams, x, y, z,
                                      self.dip[k].data_ptr(), *fp, self.fsite[k].data_ptr())
            return float(o[0]) + float(o[11]) + float(np.sum(o[23:27])), int(o[22])
        (e_el, _, _, its), ((e_mb, _, _), (e_w, _, _)) = s.split(
            lambda: c.polar_electrostatics_dev(b.imcon, b.cell, s.eparams, x, y, z, self.dip[k].data_ptr(),
                                               self.fsite[k].data_ptr()),
            lambda: (c.mbpol_forces_dev(b.imcon, b.cell, s.lst.data_ptr(), x, y, z, *fp),
                     c.water_terms_dev(b.imcon, b.cell, s.wparams, x, y, z, *fp)))
        return e_mb + e_el + float(np.sum(e_w)), its

    def forces(self):
        """Physical forces of every bead plus the springs into F; returns dict(pot (per bead),
        spring, deriv, iterations)."""
        t = self.t
        self.F.zero_()
        self.fsite.zero_()
        t.cuda.synchronize(self.dev)
        # the beads are independent configurations: one host thread each, so that their
        # kernels overlap on the GPU (the library calls release the interpreter lock)
        if self._pool is None:
            self._pool = ThreadPoolExecutor(max_workers=self.P)
        res = list(self._pool.map(self._bead_forces, range(self.P)))
        pots, its = [r[0] for r in res], [r[1] for r in res]
        # electrostatics forces (3 per site, already moved off the M sites) onto the atoms
        self.F[:, :, :self.nr] += self.fsite.view(self.P, self.natms, 3)[:, :self.nr, :].transpose(1, 2)
        t.cuda.synchronize(self.dev)
        c = self.slots[0].ctx
        c.ring_transform_dev(self.P, 3 * self.natms, self.c2n, self.X.data_ptr(), self.Q.data_ptr())
        xc = (self.Q[0] * self.fac1).contiguous()
        t.cuda.synchronize(self.dev)
        es, ed = c.ring_spring_dev(self.P, self.natms, self.coeff, self.fac3, self.force_scale,
                                   self.mass_all.data_ptr(), self.X.data_ptr(), xc.data_ptr(), self.F.data_ptr())
        self.pot = np.array(pots)
        self.spring, self.deriv, self.iterations = es, ed, its
        return dict(pot=self.pot.copy(), spring=es, deriv=ed, iterations=its)

    # ---------------------------------------------------------------- dynamics ----
    def kinetic(self):
        v = self.V[:, :, :self.nr]
        return float((0.5 * self.mass * (v * v).sum(1)).sum())

    def energy(self):
        """Ring-polymer Hamiltonian sum_k [K_k + V(x_k)] + Epot_spring."""
        return self.kinetic() + float(self.pot.sum()) + self.spring

    def step(self, dt_fs):
        """One velocity-Verlet step of every bead (physical masses); returns energy()."""
        dt = dt_fs * 1.0e-3
        t = self.t
        t.cuda.synchronize(self.dev)
        for k, s in enumerate(self.slots):
            v, f = self.V[k], self.F[k]
            s.ctx.verlet_stage(1, dt, self.mass.data_ptr(), s.x.data_ptr(), s.y.data_ptr(), s.z.data_ptr(),
                               v[0].data_ptr(), v[1].data_ptr(), v[2].data_ptr(), f[0].data_ptr(),
                               f[1].data_ptr(), f[2].data_ptr(), None)
        self.forces()
        for k, s in enumerate(self.slots):
            v, f = self.V[k], self.F[k]
            s.ctx.verlet_stage(2, dt, self.mass.data_ptr(), s.x.data_ptr(), s.y.data_ptr(), s.z.data_ptr(),
                               v[0].data_ptr(), v[1].data_ptr(), v[2].data_ptr(), f[0].data_ptr(),
                               f[1].data_ptr(), f[2].data_ptr(), None, self.ek[k].data_ptr())
        return self.energy()

    def close(self):
        if self._pool is not None:
            self._pool.shutdown()
            self._pool = None
        for s in self.slots:
            s.close()
            s.ctx.close()
