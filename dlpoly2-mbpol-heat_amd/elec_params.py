"""Parameters of the MB-pol polarizable electrostatics.

Model constants come from the reference's `vmbpol` preset (simdef_cmd.f:1036-1058):
nthole = 4 with smeared charge-charge and charge-dipole interactions, gamma = 0.426706882,
athole = 0.055 (non-bonded dipole-dipole), athole12 = 0.626, athole13 = 0.055,
ascc = ascd = 0.4, new DMS charge assignment.  The site polarizabilities and the oxygen
charge are FIELD-file inputs in the reference (none ships with it); the values here are
MB-pol's: alpha_O = 1.310 A^3, alpha_H = 0.294 A^3, the M site unpolarizable
(polr2 = 0) but smeared like the oxygen (polr = 1.310), q_O = 0 (the DMS puts the
negative charge on M; qdetermine.f90:301-305 requires O + H1 + H2 + M = 0).

Ewald parameters per workload follow the benchmark configurations: rcut 9 A, alpha 0.3
1/A and kmax 16 at 512 waters (L = 24.84 A); the 64-water box (L = 12.42 A) cannot hold a
9 A cutoff under the minimum image, so it uses rcut 6 A, alpha 0.55 1/A, kmax 8.
The dipole tolerance is 1e-12 Debye expressed in e*A (eatd = 4.803239 Debye per e*A,
module.f90:38); sor_omega defaults to 0.3 (simdef_cmd.f:239).
"""
from dataclasses import dataclass, field

EATD = 4.803239


@dataclass
class ElecParams:
    rcut: float = 9.0
    alpha: float = 0.3
    kmax: tuple = (16, 16, 16)
    toler: float = 1e-12 / EATD
    omega: float = 0.3
    max_iter: int = 22581
    gamma: float = 0.426706882
    q_o: float = 0.0
    polr: tuple = (1.310, 0.294, 1.310)     # O, H, M (smearing)
    polr2: tuple = (1.310, 0.294, 0.0)      # O, H, M (induction)
    athole: float = 0.055
    athole12: float = 0.626
    athole13: float = 0.055
    ascc: float = 0.4
    ascd: float = 0.4
    compat_qd_column: int = 1
    solver: int = 0                          # 0 SOR (iter.f), 1 conjugate gradients (library only)
    extra: dict = field(default_factory=dict)


def for_box(L):
    """Ewald parameters for a cubic box of side L (Angstrom)."""
    if L >= 18.0:
        k = 16 if L < 30 else 24
        return ElecParams(rcut=9.0, alpha=0.3, kmax=(k, k, k))
    return ElecParams(rcut=6.0, alpha=0.55, kmax=(8, 8, 8))


def for_cluster(L):
    """Parameters for an isolated cluster (imcon 0) inside a cube of side L
    (waterbox.make_cluster): direct sums, no Ewald split (alpha = 0), every site pair
    interacting (rcut = L bounds the cluster's extent); kmax is not used."""
    return ElecParams(rcut=float(L), alpha=0.0, kmax=(1, 1, 1))
