"""Seeded synthetic water configurations in the reference's DL_POLY site layout.

Site layout follows the reference's FIELD convention for MB-pol runs
(ttm2list.f:1-77, exclude_ttm.f:15-109): waters first as O, H, H triples, then the
block of M sites (one per water, the last molecule type).  ``listttm2`` holds the
1-based ids of the water atoms in the order O, H, H per molecule, as mbpol_forces
expects (mbpol.f90:184-189).

Generator (per configuration, numpy PCG64 seeded):
  * monomer: r_OH = 0.9572 A and angle HOH = 104.52 deg, each O-H length perturbed by
    N(0, 0.02 A) and the angle by N(0, 3 deg);
  * a uniform random rotation (unit quaternion from three uniforms, Shoemake);
  * O on a cubic lattice site plus U(-0.3, 0.3) A per axis; when fewer molecules than
    sites are requested, the occupied sites are a random subset;
  * Maxwell-Boltzmann velocities at T in A/ps (DL_POLY internal units, boltz =
    0.831451115 in 10 J/mol/K, dl_params.inc) with the centre-of-mass momentum removed;
  * M = O + gamma/2 (img(H1-O) + img(H2-O)) with gamma = 0.426706882
    (dlpoly_cmd.f:2041-2057, simdef_cmd.f:1054).
"""
import numpy as np

BOLTZ = 0.831451115        # dl_params.inc
MASS_O = 15.9994
MASS_H = 1.008
GAMMA_M = 0.426706882      # simdef_cmd.f:1054 (vmbpol preset)
R_OH = 0.9572
THETA_HOH = 104.52


class WaterBox:
    """Positions, velocities and bookkeeping of one synthetic water configuration."""

    def __init__(self, nw, L, xyz, vel, mass, listttm2):
        self.nw = nw
        self.L = L
        self.cell = np.array([L, 0, 0, 0, L, 0, 0, 0, L], dtype=np.float64)
        self.imcon = 1
        self.xyz = xyz            # (4*nw, 3) sites: O,H,H per water then M block
        self.vel = vel            # (4*nw, 3), zero on M sites
        self.mass = mass          # (4*nw,)
        self.listttm2 = listttm2  # (3*nw,) int32, 1-based
        self.natms = 4 * nw
        self.nreal = 3 * nw

    @property
    def x(self):
        return np.ascontiguousarray(self.xyz[:, 0])

    @property
    def y(self):
        return np.ascontiguousarray(self.xyz[:, 1])

    @property
    def z(self):
        return np.ascontiguousarray(self.xyz[:, 2])


def _rotation(rng):
    u1, u2, u3 = rng.random(3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def m_sites(xyz_real, nw, L, gamma=GAMMA_M):
    """M-site positions, dlpoly_cmd.f:2041-2057 (imcon 1 minimum image of O-H)."""
    o = xyz_real[0::3][:nw]
    h1 = xyz_real[1::3][:nw]
    h2 = xyz_real[2::3][:nw]
    d1 = h1 - o
    d2 = h2 - o
    d1 -= L * np.round(d1 / L)
    d2 -= L * np.round(d2 / L)
    return o + 0.5 * gamma * (d1 + d2)


def make_waterbox(nw, L, nlat, seed, temperature=300.0, jitter=0.3, wrap=True):
    """nw waters on an nlat^3 lattice in a cubic box of edge L (Angstrom)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nsites = nlat ** 3
    if nw > nsites:
        raise ValueError("more molecules than lattice sites")
    sites = np.arange(nsites) if nw == nsites else np.sort(rng.choice(nsites, nw, replace=False))
    a = L / nlat
    real = np.zeros((3 * nw, 3))
    for m, s in enumerate(sites):
        ix, iy, iz = s // (nlat * nlat), (s // nlat) % nlat, s % nlat
        o = (np.array([ix, iy, iz]) + 0.5) * a + rng.uniform(-jitter, jitter, 3)
        r1 = R_OH + rng.normal(0.0, 0.02)
        r2 = R_OH + rng.normal(0.0, 0.02)
        th = np.deg2rad(THETA_HOH + rng.normal(0.0, 3.0))
        h1 = np.array([r1 * np.sin(th / 2), 0.0, r1 * np.cos(th / 2)])
        h2 = np.array([-r2 * np.sin(th / 2), 0.0, r2 * np.cos(th / 2)])
        R = _rotation(rng)
        real[3 * m] = o
        real[3 * m + 1] = o + R @ h1
        real[3 * m + 2] = o + R @ h2
    if wrap:
        real -= L * np.floor(real / L)
    mass_real = np.tile([MASS_O, MASS_H, MASS_H], nw)
    v = rng.normal(size=(3 * nw, 3)) * np.sqrt(BOLTZ * temperature / mass_real)[:, None]
    p = (mass_real[:, None] * v).sum(0)
    v -= p / mass_real.sum()
    xyz = np.zeros((4 * nw, 3))
    xyz[:3 * nw] = real
    xyz[3 * nw:] = m_sites(real, nw, L)
    vel = np.zeros((4 * nw, 3))
    vel[:3 * nw] = v
    mass = np.zeros(4 * nw)
    mass[:3 * nw] = mass_real
    listttm2 = np.arange(1, 3 * nw + 1, dtype=np.int32)
    return WaterBox(nw, L, xyz, vel, mass, listttm2)


# The workloads, by size (seeds fixed here so every caller sees the same inputs).
def box_64(seed=1):
    """64 waters, L = 12.42 A, 4^3 lattice."""
    return make_waterbox(64, 12.42, 4, seed)


def box_512(seed=2):
    """512 waters, L = 24.84 A, 8^3 lattice."""
    return make_waterbox(512, 24.84, 8, seed)


def box_256(seed=3):
    """256 waters, L = 19.72 A (one PIMD bead of the serving mix)."""
    return make_waterbox(256, 19.72, 7, seed)


def box_2048(seed=4):
    """2048 of 2197 sites of a 13^3 lattice, L = 39.43 A."""
    return make_waterbox(2048, 39.43, 13, seed)


RHO_WATER = 0.997 * 6.02214076e23 / 18.01528 * 1e-24   # molecules per A^3 at 0.997 g/cm^3


def make_cluster(n, seed, margin=12.0, dmin=2.6, temperature=300.0):
    """An isolated cluster of n waters built by random sequential addition: oxygens placed
    uniformly in a sphere holding n molecules at 0.997 g/cm^3, rejected below an O-O distance
    of dmin, each molecule a perturbed monomer with a random orientation as in
    make_waterbox.  The cluster is isolated (imcon 0: no images, direct sums); it is
    centred in a cube of side L = 2R + margin, which bounds every site pair's distance and
    is kept as the box's cell for bookkeeping only."""
    rng = np.random.Generator(np.random.PCG64(seed))
    R = (3.0 * n / (4.0 * np.pi * RHO_WATER)) ** (1.0 / 3.0)
    L = 2.0 * R + margin
    centre = np.full(3, 0.5 * L)
    oxy = []
    tries = 0
    while len(oxy) < n:
        tries += 1
        if tries > 200000:
            raise RuntimeError("random sequential addition did not place every molecule")
        u = rng.uniform(-1.0, 1.0, 3)
        if u @ u > 1.0:
            continue
        o = centre + R * u
        if oxy and np.min(np.linalg.norm(np.asarray(oxy) - o, axis=1)) < dmin:
            continue
        oxy.append(o)
    real = np.zeros((3 * n, 3))
    for m, o in enumerate(oxy):
        r1 = R_OH + rng.normal(0.0, 0.02)
        r2 = R_OH + rng.normal(0.0, 0.02)
        th = np.deg2rad(THETA_HOH + rng.normal(0.0, 3.0))
        h1 = np.array([r1 * np.sin(th / 2), 0.0, r1 * np.cos(th / 2)])
        h2 = np.array([-r2 * np.sin(th / 2), 0.0, r2 * np.cos(th / 2)])
        Rm = _rotation(rng)
        real[3 * m] = o
        real[3 * m + 1] = o + Rm @ h1
        real[3 * m + 2] = o + Rm @ h2
    mass_real = np.tile([MASS_O, MASS_H, MASS_H], n)
    v = rng.normal(size=(3 * n, 3)) * np.sqrt(BOLTZ * temperature / mass_real)[:, None]
    v -= (mass_real[:, None] * v).sum(0) / mass_real.sum()
    xyz = np.zeros((4 * n, 3))
    xyz[:3 * n] = real
    xyz[3 * n:] = m_sites(real, n, L)
    vel = np.zeros((4 * n, 3))
    vel[:3 * n] = v
    mass = np.zeros(4 * n)
    mass[:3 * n] = mass_real
    box = WaterBox(n, L, xyz, vel, mass, np.arange(1, 3 * n + 1, dtype=np.int32))
    box.imcon = 0
    return box


def displaced(box, sigma, seed):
    """A copy of box with every atom displaced by N(0, sigma) per axis (a PIMD bead of the
    serving mix); the M sites are rebuilt."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nw = box.nw
    real = box.xyz[:3 * nw] + rng.normal(0.0, sigma, (3 * nw, 3))
    xyz = np.zeros_like(box.xyz)
    xyz[:3 * nw] = real
    xyz[3 * nw:] = m_sites(real, nw, box.L)
    return WaterBox(nw, box.L, xyz, box.vel.copy(), box.mass.copy(), box.listttm2.copy())
