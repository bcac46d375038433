"""i-PI driver: serves force evaluations to an i-PI server over its socket protocol.

Stands in for the reference's i-PI client (ipi.f90:65-198 over sockets.c:47-160, opened and
stepped from dlpoly_cmd.f:1737-1785): the program connects to the server and answers its
12-byte, space-padded headers.

  STATUS    -> "HAVEDATA" when forces are ready, "READY" after INIT, else "NEEDINIT"
  INIT      <- int32 replica id, int32 length, parameter string (ignored, as the reference)
  POSDATA   <- cell h (9 float64), its inverse (9 float64), int32 natoms, 3*natoms float64
               positions; the driver evaluates the configuration
  GETFORCE  -> "FORCEREADY", float64 energy, int32 natoms, 3*natoms float64 forces, 9 float64
               virial, int32 length and the extra string (the water dipoles, Debye)
  EXIT      ends the loop (the reference stops on any unknown header)

Units follow the reference, including its conversion constants declared as default REAL,
i.e. rounded to float32 before use: bohr -> Angstrom 0.52917721067, internal energy ->
Hartree 3.8087989e-6, and the factor 418.6799938/418.4 (ipi.f90:88-90).  Positions and the
cell arrive in bohr; the cell matrix is transposed as the reference does
(cell = reshape(transpose(h)) * conv, ipi.f90:184).  Energy, forces and virial leave in
Hartree units: force * conv_e * conv_l * ratio, virial = transpose(stress) * conv_e * ratio,
energy * conv_e * ratio (ipi.f90:119-121), in that order of multiplication.

The atoms are waters in O, H, H order (natoms = 3 * nwat, ipi.f90:173).  The dipole string
holds each water's total dipole (permanent + induced, Debye) as numbers separated by
spaces; its layout is this package's (the reference uses Fortran list-directed output).
"""
import socket
import struct

import numpy as np

from . import waterbox as _wb

HDR = 12
CONV_LENGTH = float(np.float32(0.52917721067))            # Angstrom per bohr
CONV_ENERGY = float(np.float32(3.8087989e-06))            # Hartree per 10 J/mol
CONS_RATIO = float(np.float32(418.6799938 / 418.4))
EATD = 4.803239                                           # Debye per e*Angstrom (module.f90:38)


def _hdr(s):
    return s.ljust(HDR).encode()


class Connection:
    def __init__(self, sock):
        self.sock = sock

    def recv(self, n):
        buf = bytearray()
        while len(buf) < n:
            chunk = self.sock.recv(n - len(buf))
            if not chunk:
                raise ConnectionError("i-PI server closed the connection")
            buf += chunk
        return bytes(buf)

    def send(self, data):
        self.sock.sendall(data)

    def recv_int(self):
        return struct.unpack("<i", self.recv(4))[0]

    def recv_f64(self, n):
        return np.frombuffer(self.recv(8 * n), dtype="<f8").copy()


def connect(address):
    """address as the reference's serveraddr: "host:port" (INET) or "unix:name" (the UNIX
    socket /tmp/ipi_<name>, sockets.c:98-101)."""
    head, _, tail = address.rpartition(":")
    if head == "unix" or address.startswith("unix:"):
        name = address.split(":", 1)[1].rsplit(":", 1)[0] if address.count(":") > 1 else tail
        s = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
        s.connect("/tmp/ipi_" + name)
    else:
        s = socket.create_connection((head, int(tail)))
    return Connection(s)


def cell_from_h(h9):
    """DL_POLY cell(9) in Angstrom from the 9 doubles of i-PI's h (ipi.f90:159-160,184)."""
    h = np.asarray(h9, dtype=float).reshape(3, 3, order="F")
    return (h.T.reshape(-1, order="F") * CONV_LENGTH).copy()


def imcon_of(cell):
    c = np.asarray(cell).reshape(3, 3)
    off = np.abs(c - np.diag(np.diag(c))).max()
    if off == 0.0:
        return 1 if (c[0, 0] == c[1, 1] == c[2, 2]) else 2
    return 3


def box_from_posdata(cell, pos):
    """A WaterBox (no velocities) from a cell (Angstrom) and 3*nwat positions (Angstrom)."""
    nat = pos.shape[0]
    if nat % 3:
        raise ValueError("i-PI atoms must be waters in O, H, H order")
    nw = nat // 3
    c = np.asarray(cell, dtype=float)
    L = float(c[0])
    xyz = np.zeros((4 * nw, 3))
    xyz[:nat] = pos
    if imcon_of(c) == 1:
        xyz[nat:] = _wb.m_sites(pos, nw, L)
    mass = np.zeros(4 * nw)
    mass[:nat] = np.tile([_wb.MASS_O, _wb.MASS_H, _wb.MASS_H], nw)
    box = _wb.WaterBox(nw, L, xyz, np.zeros((4 * nw, 3)), mass, np.arange(1, nat + 1, dtype=np.int32))
    box.cell = c
    box.imcon = imcon_of(c)
    return box


def dipole_string(dip_debye):
    return " ".join("%.16e" % v for v in np.asarray(dip_debye).reshape(-1))


class Driver:
    """Answers one i-PI server.  `evaluate(box)` returns (energy, forces (nat, 3),
    stress (9), dipoles (nwat, 3) in Debye), energies in the internal unit."""

    def __init__(self, conn, evaluate):
        self.conn, self.evaluate = conn, evaluate
        self.hasdata = False
        self.imginit = False
        self.replica = None
        self.steps = 0

    def run(self, max_steps=None):
        c = self.conn
        while True:
            header = c.recv(HDR).decode().strip()
            if header == "STATUS":
                c.send(_hdr("HAVEDATA" if self.hasdata else ("READY" if self.imginit else "NEEDINIT")))
            elif header == "INIT":
                self.replica = c.recv_int()
                n = c.recv_int()
                c.recv(n)
                self.imginit = True
            elif header == "POSDATA":
                h = c.recv_f64(9)
                c.recv_f64(9)                       # the inverse cell (unused, as the reference)
                nat = c.recv_int()
                pos = c.recv_f64(3 * nat).reshape(nat, 3) * CONV_LENGTH
                cell = cell_from_h(h)
                self.pot, self.forces, self.stress, self.dip = self.evaluate(box_from_posdata(cell, pos))
                self.hasdata = True
            elif header == "GETFORCE":
                if self.hasdata:
                    f = np.asarray(self.forces, dtype=float).reshape(-1) * CONV_ENERGY * CONV_LENGTH * CONS_RATIO
                    vir = np.asarray(self.stress, dtype=float).reshape(3, 3, order="F").T * CONV_ENERGY * CONS_RATIO
                    pot = self.pot * CONV_ENERGY * CONS_RATIO
                    extra = dipole_string(self.dip).encode()
                    c.send(_hdr("FORCEREADY") + struct.pack("<d", pot) + struct.pack("<i", f.size // 3)
                           + f.astype("<f8").tobytes() + vir.reshape(-1, order="F").astype("<f8").tobytes()
                           + struct.pack("<i", len(extra)) + extra)
                    self.hasdata = False
                    self.imginit = False
                    self.steps += 1
                    if max_steps is not None and self.steps >= max_steps:
                        return
            elif header == "EXIT":
                return
            else:
                raise RuntimeError(f"Unexpected message from wrapper {header!r}")


def slot_evaluator(precision=0, solver=1):
    """evaluate(box) through engine.Slot (one slot per water count, reused)."""
    from . import elec_params as _ep
    from . import water_params as _wp
    from .engine import Slot
    slots = {}

    def evaluate(box):
        elec = _ep.for_box(box.L)
        water = _wp.for_box(box.L)
        s = slots.get(box.nw)
        if s is None:
            s = slots[box.nw] = Slot(box, precision=precision, solver=solver, elec=elec, water=water)
        else:
            s.set_params(elec, water)
            s.load(box)
        r = s.evaluate()
        f = s.forces()
        stress = np.array(r["stress"], dtype=float)
        for k in (0, 4, 8):                  # long-range dispersion (dlpoly_cmd.f:2572-2574)
            stress[k] -= r["vir_water"][3] / 3.0
        return r["eng"], f, stress, s.water_dipoles() * EATD
    return evaluate
