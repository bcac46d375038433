"""Device-resident force evaluations of MB-pol water boxes, alone or several at once.

A `Slot` owns one library context (its own HIP stream and device buffers) and the
device-resident inputs of one configuration: positions (with the M sites), the water list,
velocities and per-atom kinetic energies.  `Slot.evaluate()` runs one complete force
evaluation in the order of the reference's force routine (dlpoly_cmd.f:2287-2504 with
forces.f): MB-pol 2B/3B, the polarizable Ewald electrostatics with induced dipoles solved
from zero, the monomer, dispersion and long-range terms, and the heat flux over all of
them.  Forces stay on the device; the scalars come back to the host.

`run_concurrently(slots, steps)` drives several slots from one host thread each, so that
their kernels overlap on the GPU (each context launches on its own stream).  A context's
results do not depend on what runs beside it: every kernel sums in a fixed order.
"""
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import elec_params as _ep
from . import water_params as _wp
from .mbgpu import Context, ElecParams, WaterParams

ENGUNIT = 418.4


class Slot:
    def __init__(self, box, precision=0, solver=1, elec=None, water=None, device="cuda:0", overlap=True,
                 recip=None, fused=True, dedicated_queues=False):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.box = box
        # dedicated_queues: the context's streams get hardware queues of their own (one
        # latency-bound slot, e.g. an MD run); shared queues suit many concurrent slots
        self.ctx = Context(box.nw, box.natms, dedicated_queues=dedicated_queues)
        self.ctx.set_precision(precision)
        # recip: None (the Ewald sum over k) or ("spme", order, grid)
        if recip is not None:
            self.ctx.set_elec_reciprocal(recip[0], recip[1], recip[2])
        self.precision = precision
        self.solver = solver
        self.overlap = bool(overlap)
        # fused: one library call per evaluation (mbgpu_evaluate_dev: every part, the two
        # streams and, once the buffers are sized, one captured graph); otherwise the parts
        # are called in turn (the electrostatics from a helper thread when overlap is set)
        self.fused = bool(fused)
        self._helper = None
        self.set_params(elec or _ep.for_box(box.L), water or _wp.for_box(box.L))
        self.stream = torch.cuda.Stream(device=self.dev)
        self._alloc(box)
        self.load(box)

    def set_params(self, elec, water):
        """Electrostatics (elec_params.ElecParams) and water-term (water_params.WaterParams)
        parameters of the following evaluations."""
        self.eparams = ElecParams.from_params(elec, distribute_msite=1, solver=self.solver)
        self.wparams = WaterParams.from_params(water)

    def _alloc(self, box):
        # allocated and cleared on the slot's stream, where load() then writes them: a
        # clear on another stream would not be ordered before the upload (torch's streams
        # do not wait for the default stream)
        t = self.torch
        n, nr = box.natms, box.nreal
        f64 = dict(dtype=t.float64, device=self.dev)
        m = 3 * n + 4 * nr
        with t.cuda.stream(self.stream):
            # positions, velocities and kinetic energies are views of one buffer, so that a
            # configuration is uploaded in one copy
            self._dbuf = t.zeros(m, **f64)
            self.x, self.y, self.z = self._dbuf[:n], self._dbuf[n:2 * n], self._dbuf[2 * n:3 * n]
            self.v = self._dbuf[3 * n:3 * n + 3 * nr].view(3, nr)
            self.ek = self._dbuf[3 * n + 3 * nr:]
            self.lst = t.zeros(nr, dtype=t.int32, device=self.dev)
        self._hbuf = t.empty(m, dtype=t.float64, pin_memory=True)   # pinned staging of load()
        self._hlst = None                                            # the list last uploaded

    def load(self, box):
        """Upload a configuration of the same size (positions, list, velocities)."""
        t = self.torch
        if box.nw != self.box.nw:
            raise ValueError("a slot holds configurations of one size")
        self.box = box
        n, nr = box.natms, box.nreal
        h = self._hbuf.numpy()        # free: the previous load() synchronized its copy
        h[:n], h[n:2 * n], h[2 * n:3 * n] = box.x, box.y, box.z
        h[3 * n:3 * n + 3 * nr] = box.vel[:nr].T.ravel()
        h[3 * n + 3 * nr:] = 0.5 * box.mass[:nr] * (box.vel[:nr] ** 2).sum(1)
        lst = np.ascontiguousarray(box.listttm2, dtype=np.int32)
        with t.cuda.stream(self.stream):
            self._dbuf.copy_(self._hbuf, non_blocking=True)
            if self._hlst is None or not np.array_equal(lst, self._hlst):
                self.lst.copy_(t.as_tensor(lst))
                self._hlst = lst.copy()
        self.stream.synchronize()

    def evaluate(self):
        """One complete force evaluation; returns a dict of scalars (internal units).
        The induced dipoles start from zero inside the library and the forces stay in the
        context (see forces()).  With `overlap` set, the electrostatics (the context's second
        stream) run from a helper thread while this thread runs MB-pol and the water terms;
        the parts share no buffers and the heat flux starts after both have returned."""
        b, c = self.box, self.ctx
        x, y, z = self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr()
        if self.fused:
            o = c.evaluate_dev(b.imcon, b.cell, 3 * b.nw, self.lst.data_ptr(), self.eparams, self.wparams, x, y, z,
                               None, self.v[0].data_ptr(), self.v[1].data_ptr(), self.v[2].data_ptr(),
                               self.ek.data_ptr())
            e_mb, v_mb, s_mb = float(o[0]), float(o[1]), o[2:11].copy()
            e_el, v_el, s_el, it = float(o[11]), float(o[12]), o[13:22].copy(), int(o[22])
            e_w, v_w, s_w = [float(t) for t in o[23:27]], [float(t) for t in o[27:31]], o[31:40].copy()
            jk, jp, jt = o[40:43].copy(), o[43:46].copy(), o[46:49].copy()
            return dict(eng=e_mb + e_el + float(np.sum(e_w)), eng_mbpol=e_mb, eng_elec=e_el, eng_water=np.array(e_w),
                        vir_water=np.array(v_w), vir=v_mb + v_el + float(np.sum(v_w)), stress=s_mb + s_el + s_w,
                        dipole_iterations=it, jkin=jk, jpot=jp, jtot=jt)
        (e_el, v_el, s_el, it), ((e_mb, v_mb, s_mb), (e_w, v_w, s_w)) = self.split(
            lambda: c.polar_electrostatics_dev(b.imcon, b.cell, self.eparams, x, y, z, None),
            lambda: (c.mbpol_forces_dev(b.imcon, b.cell, self.lst.data_ptr(), x, y, z),
                     c.water_terms_dev(b.imcon, b.cell, self.wparams, x, y, z)))
        jk, jp, jt = c.heat_flux_dev(self.v[0].data_ptr(), self.v[1].data_ptr(), self.v[2].data_ptr(),
                                     self.ek.data_ptr())
        return dict(eng=e_mb + e_el + float(np.sum(e_w)), eng_mbpol=e_mb, eng_elec=e_el, eng_water=np.array(e_w),
                    vir_water=np.array(v_w),
                    vir=v_mb + v_el + float(np.sum(v_w)), stress=np.array(s_mb) + s_el + s_w, dipole_iterations=it,
                    jkin=np.array(jk), jpot=np.array(jp), jtot=np.array(jt))

    def split(self, elec, rest):
        """Returns (elec(), rest()).  With `overlap` set, elec() (the electrostatics, which
        the library runs on the context's second stream) runs from a helper thread while
        rest() (MB-pol and water terms on the first stream) runs from this one; every
        library entry waits for its stream before it returns, so both parts are complete
        here.  Without it, rest() then elec()."""
        if not self.overlap:
            r = rest()
            return elec(), r
        if self._helper is None:
            self._helper = ThreadPoolExecutor(max_workers=1)
        fut = self._helper.submit(elec)
        try:
            r = rest()
        finally:
            e = fut.result()
        return e, r

    def close(self):
        if self._helper is not None:
            self._helper.shutdown()
            self._helper = None

    def forces(self):
        """Total forces of the last evaluation on the real atoms, (3*nw, 3), internal units."""
        # (MB-pol + electrostatics) + water terms, summed on the device in that order
        return self.ctx.total_forces()


    def water_dipoles(self):
        """Total dipole of every water of the last evaluation (e*Angstrom): the site charges
        times the site positions relative to the oxygen (minimum image) plus the induced
        dipoles of the four sites."""
        nw = self.box.nw
        chge = self.ctx.elec_details()["chge"]
        dip = self.ctx.dipoles()
        self.torch.cuda.synchronize(self.dev)
        xyz = np.stack([self.x.cpu().numpy(), self.y.cpu().numpy(), self.z.cpu().numpy()], 1)
        cell = np.asarray(self.box.cell, dtype=float).reshape(3, 3)     # rows: cell vectors
        inv = np.linalg.inv(cell)
        out = np.zeros((nw, 3))
        for w in range(nw):
            sites = [3 * w, 3 * w + 1, 3 * w + 2, 3 * nw + w]
            o = xyz[3 * w]
            for s in sites:
                d = xyz[s] - o
                d = d - np.rint(d @ inv) @ cell
                out[w] += chge[s] * d + dip[s]
        return out


def run_concurrently(slots, steps, warmup=0):
    """Evaluate every slot `steps` times, one host thread per slot.  Returns (seconds,
    per-slot lists of results); the timed region starts after `warmup` evaluations."""
    import torch
    for s in slots:
        for _ in range(warmup):
            s.evaluate()
    torch.cuda.synchronize()
    out = [[] for _ in slots]
    errors = []

    def loop(i, s):
        try:
            for _ in range(steps):
                out[i].append(s.evaluate())
        except Exception as e:          # surfaced after the join
            errors.append(e)

    threads = [threading.Thread(target=loop, args=(i, s)) for i, s in enumerate(slots)]
    t0 = time.perf_counter()
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return dt, out
