"""Parameters of the water-monomer and dispersion terms that accompany MB-pol.

The reference reads the dispersion from FIELD `vdw` entries of type 12 ("tt68": C6, a6,
C8, a8 with -C6 f6(a6 r)/r^6 - C8 f8(a8 r)/r^8, sysdef_cmd.f:1505-1510, 1598-1602), with
C6 and C8 in FIELD energy units times A^6 / A^8 and converted to the internal unit
(10 J/mol) by engunit (418.4 for kcal/mol, sysdef_cmd.f:219).  No FIELD ships with the
reference; the values here are MB-pol's published dispersion (Babin, Leforestier and
Paesani, J. Chem. Theory Comput. 9, 5395 (2013)): C6 in kcal/mol A^6 and d6 in 1/A for
O-O, O-H and H-H, no C8 term.  C8 = 0 still needs a positive a8: the long-range
correction divides by powers of a8*rvdw (lrcorrect.f el12), so a8 = a6 is used.

The monomer is Partridge-Schwenke: the coupling expansion is compiled into the library
(ps_type_couple.f90, module.f90 ps_type_pes), while its two-body parts are FIELD bonds in
the reference, handled by bndfrc: a Morse O-H bond ("mors": D, r0, a;
bndfrc.f:73-76) and an exponential H-H repulsion ("expo": A, B; bndfrc.f:107-108).  The
defaults are the Partridge-Schwenke values (J. Chem. Phys. 106, 4618 (1997)):
D = 42290.92019288289 cm^-1, r0 = 0.9519607159623009 A, a = 2.587949757553683 1/A,
A = 16.94879431193463 exp(12.66426998162947) cm^-1, B = 12.66426998162947 1/A, with
cm^-1 converted by engcnv = 1.19627 (module.f90:486).  With them the monomer minimum lies
at r_OH = 0.95783 A, angle 104.508 deg, the published PS equilibrium.

rvdw = 9 A as in the benchmark configurations; `rcut` is the largest cutoff of the run,
which sets the number of table points (parset.f:210-215).
"""
import math
from dataclasses import dataclass

ENGUNIT_KCAL = 418.4

MBPOL_C6 = (237.3212214147944, 83.49556669872743, 20.09358600184719)   # OO, OH, HH
MBPOL_D6 = (9.295485815062264, 9.775202425217957, 9.406475169954112)

ENGCNV_CM = 1.19627   # cm^-1 -> 10 J/mol
PS_BOND_OH = (2, (42290.92019288289 * ENGCNV_CM, 0.9519607159623009, 2.587949757553683))
PS_BOND_HH = (7, (16.94879431193463 * math.exp(12.66426998162947) * ENGCNV_CM, 12.66426998162947, 0.0))


@dataclass
class WaterParams:
    rvdw: float = 9.0
    rcut: float = 9.0
    c6: tuple = MBPOL_C6          # kcal/mol A^6
    a6: tuple = MBPOL_D6          # 1/A
    c8: tuple = (0.0, 0.0, 0.0)   # kcal/mol A^8
    a8: tuple = MBPOL_D6          # 1/A
    engunit: float = ENGUNIT_KCAL
    bond_oh: tuple = PS_BOND_OH   # (bndfrc key, prmbnd(1..3)) in internal units
    bond_hh: tuple = PS_BOND_HH
    ps_monomer: int = 1
    dispersion: int = 1
    long_range: int = 1

    def internal(self):
        """(c6, a6, c8, a8) with energies in the internal unit."""
        return (tuple(c * self.engunit for c in self.c6), tuple(self.a6),
                tuple(c * self.engunit for c in self.c8), tuple(self.a8))

    def bonds_internal(self):
        """[(key, prm3)] for the O-H and H-H bonds, energies in the internal unit."""
        return [(int(b[0]), tuple(float(v) for v in b[1])) for b in (self.bond_oh, self.bond_hh)]


def for_box(L):
    """Cutoffs for a cubic box of side L (Angstrom): 9 A, or 6 A below 18 A."""
    r = 9.0 if L >= 18.0 else 6.0
    return WaterParams(rvdw=r, rcut=r)


def for_cluster(L):
    """Cutoffs for an isolated cluster (imcon 0): the model's 9 A dispersion cutoff (no
    images to keep clear of); L is not used."""
    return WaterParams(rvdw=9.0, rcut=9.0)
