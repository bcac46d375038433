/*
 * oracle.h -- CPU restatement of the reference's MB-pol force / heat-current path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP path is tested against this library: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.  The
 * product library (libmbgpu.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Units follow the
 * reference: Angstrom, kcal/mol inside the polynomial leaves, DL_POLY internal units
 * (10 J/mol, engunit = 418.4 per kcal/mol) at the mbpol_forces level.
 */
#ifndef MBPOL_ORACLE_H
#define MBPOL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- polynomial tables (pip_oracle.c) -------------------------------------- */

/* Load the monomial and coefficient tables; returns 0 on success. */
int ora_pip_load(const char *mono2b, const char *coef2b,
                 const char *mono3b, const char *coef3b);

/* E(a, x) = sum_i a_i p_i(x) and dE/dx for an arbitrary coefficient vector
 * (poly_2b_v6x::eval, poly-2b-v6x.h:15-17; poly_3b_v2x::eval, poly-3b-v2x.h:14-16). */
double ora_poly_eval(int which, const double *a, const double *x, double *g);

/* Sum over monomials of |a_b * monomial(x)| with the fitted coefficients: the magnitude
 * against which E(x) cancels (used to scale parity tolerances). */
double ora_poly_scale(int which, const double *x);

/* Leaf restatements: x2b_v6x_p::operator() (mbpol-2b-poly.cpp:358-682) and
 * x3b_v2x_p::operator() (mbpol-3b-poly.cpp:148-312).  Same conventions as the
 * reference C ABI: 2B overwrites g1,g2 (unless r_OO > 6.5, then untouched);
 * 3B accumulates into g1..g3.  Returns E in kcal/mol. */
double ora_mbpol_2b(const double *w1, const double *w2, double *g1, double *g2);
double ora_mbpol_3b(const double *w1, const double *w2, const double *w3,
                    double *g1, double *g2, double *g3);

/* The 31 / 36 polynomial variables of one cluster (for kernel-stage parity). */
void ora_vars_2b(const double *w1, const double *w2, double *v31);
void ora_vars_3b(const double *w1, const double *w2, const double *w3, double *v36);

/* ---- mbpol_forces (mbpol_oracle.c) ------------------------------------------ */

typedef struct {
    /* inputs */
    int imcon;               /* 1, 2 or 3 (mbpol.f90:1319-1357) */
    double cell[9];
    int nw;                  /* number of waters */
    const int *listttm2;     /* 3*nw, 1-based atom ids O,H,H per water */
    const double *xxx, *yyy, *zzz;
    int natms;               /* length of per-atom arrays below */
    int compat_inplace;      /* 1: keep the reference's in-place re-imaging of the
                                shared monomer array (mbpol.f90:505-509,875-882) */
    /* outputs (accumulated, reference semantics) */
    double *fxx, *fyy, *fzz; /* internal units */
    double engsrp, virsrp, stress[9];
    double eng2, eng3;       /* 2B and 3B energies, internal units */
    double *atomic_energy;   /* natms: += E/6, E/9 (heatcurrent.f90 update_energy_mbpol) */
    double *heat_w;          /* natms*9 row-major [a][alpha][beta]: per-atom
                                sum_j F_ij,alpha (x_i - x_j)_beta of the MB-pol force-matrix
                                entries (the contraction compute_heat_flux2 makes) */
    double *force_matrix;    /* optional dense natms*natms*3, layout [k][j][i] as
                                Fortran force_matrix(i,j,k); NULL to skip */
    /* lists (optional outputs, 0-based water indices) */
    int *dimers;  long ndimers, cap_dimers;    /* pairs i<j */
    int *trimers; long ntrimers, cap_trimers;  /* triples i<j<k */
    int skip_3b;             /* debugging aid: 1 = 2B only */
} ora_mbpol_args;

/* mbpol_forces (mbpol.f90:98-449) with do_dimer (:453-790) and do_trimer (:795-1271),
 * single rank (bead_size = 1).  Returns 0, or -1 for an unsupported imcon (the
 * reference stops in fixme, mbpol.f90:1357). */
int ora_mbpol_forces(ora_mbpol_args *a);

/* Route the oracle's mbpol_forces through other leaf implementations with the reference
 * C ABI (e.g. the reference's own compiled mbpol_2b_poly_ / mbpol_3b_poly_); NULL
 * restores the oracle's leaves. */
typedef void (*ora_leaf2_fn)(const double *, const double *, double *, double *, double *, double *);
typedef void (*ora_leaf3_fn)(const double *, const double *, const double *, double *, double *, double *,
                             double *, double *);
void ora_set_leaves(ora_leaf2_fn f2, ora_leaf3_fn f3);

/* image_vectors (mbpol.f90:1305-1361) */
int ora_image_vectors(int imcon, const double *cell, int n, double *dx);

/* ---- polarizable electrostatics (elec_oracle.c) ------------------------------ */

/* erfcc.f, gammq.f (gammln, gammq), erfcgenp.f with the grid of parset.f:210-217
 * (ora_erfc_grid returns mxegrd; the table holds (mxegrd+1)*4 doubles, row i at r=i*drewd). */
double ora_erfcc(double x);
double ora_gammln(double x);
double ora_gammq(double a, double x);
int ora_erfc_grid(double rcut, double *drewd);
void ora_erfc_table(double alpha, double rcut, double *table);

/* DMS table (dms_table.txt: 84 rows "i1 i2 i3 coef"); returns 0 on success. */
int ora_dms_load(const char *path);
/* DMS charges of one water (w9: O, H1, H2 xyz) -> q(H1), q(H2), q(M) (qdetermine.f90). */
int ora_dms_charges(int imcon, const double *cell, const double *w9, double gamma, double *q3);
/* PS dipole charges of one water (qdetermine.f90:130-236), bonds by minimum image in cell:
 * p8 = p1, p2, dp1/dr1, dp1/dr2,
 * dp1/dcos, dp2/dr1, dp2/dr2, dp2/dcos (r in Angstrom, cos of the HOH angle). */
int ora_dms_p(int imcon, const double *cell, const double *w9, double *p8);

typedef struct {
    /* inputs */
    int imcon;                 /* 1, 2, 3 */
    double cell[9];
    int nw;                    /* waters; sites O,H1,H2 of water w at 3w..3w+2, M at 3nw+w */
    double *x, *y, *z;         /* 4*nw positions; the M-site entries are (re)computed */
    double rcut, alpha;        /* real-space cutoff, Ewald alpha (1/A) */
    int kmax[3];               /* Ewald kmax1..3 */
    double toler, omega;       /* dipole tolerance (e A), SOR mixing (old-dipole weight) */
    int max_iter;              /* stop after this many iterations (reference: 22581) */
    double gamma, q_o;         /* M-site gamma; oxygen charge from the FIELD */
    double polr_site[3];       /* smearing polarizabilities O, H, M (A^3) */
    double polr2_site[3];      /* induction polarizabilities O, H, M (A^3) */
    double athole, athole12, athole13, ascc, ascd;
    int compat_qd_column;      /* 1: the M column of qdforce's force-matrix entries is
                                  n_water+imol as in qdforce.f90:332-337 */
    /* in/out */
    double *dip;               /* 3*4nw dipoles: start values in, converged out */
    /* outputs (site level, before the M-site redistribution; f and the per-atom arrays
     * accumulate, the scalars are set) */
    double *f;                 /* 3*4nw interleaved forces, internal units */
    double *chge;              /* 4nw charges */
    double *vesp;              /* 4nw electrostatic potential at the sites */
    double *efield;            /* optional 3*4nw total field at convergence */
    double *f_ew1p;            /* optional 3*4nw reciprocal-space forces (force_ew1p) */
    double eng_ew1p, eng_ew2p, eng_ew3p, eng_polar, eng_cpe;
    double vir_cpe, vir_ew1p, vir_qd, stress[9];
    double stress_ew1p[9];     /* the reciprocal-space part of stress */
    double *e_ew1p, *e_ew2p, *e_ew3p, *e_polar;  /* optional 4nw per-atom energies */
    double *s_ew1p;            /* optional 4nw*9 per-atom reciprocal stress [i][a][b] */
    double *heat_w;            /* optional 3nw*9 per-atom W_i over the electrostatic
                                  force-matrix entries (M rows redistributed) */
    double *force_matrix;      /* optional dense [k][j][i] over 4nw sites (M rows
                                  redistributed) */
    long npairs;               /* site pairs within rcut */
    int niter;                 /* solver iterations */
    double last_delta;         /* RMS dipole change of the last iteration */
    double volm;
    double stress_qd_atoms[9]; /* sum over atoms of the per-atom qd stresses the reference
                                  records when HEAT_STRESS is not defined: the angular terms
                                  of update_stress_qd (qdforce.f90:504-531), set */
} ora_elec_args;

/* The polarizable-Ewald part of forces (forces.f:268-404,585-610,700-712,874-892) with
 * qdetermine and the M sites.  Returns 0, -1 (imcon), -2 (no convergence within
 * max_iter: the reference stops), -3 (DMS table not loaded). */
int ora_elec_forces(ora_elec_args *a);

/* qdforce alone (qdforce.f90:13-569) for the site potentials in a->vesp (a test aid):
 * accumulates a->f, a->stress, a->force_matrix, a->heat_w; sets a->vir_qd and
 * a->stress_qd_atoms.  Positions are used as given (M sites not rebuilt). */
int ora_qdforce(ora_elec_args *a);

/* M-site redistribution of forces (3 per site), energies and 9-component stresses. */
void ora_msite_distribute(int nw, double gamma, double *f, double *e, double *s9);

/* ---- monomer, dispersion, long-range tail (water_oracle.c) -------------------- */

/* PS potential tables (ps_table.txt, ps_mass.txt); init_ps_type_pes processing. */
int ora_ps_load(const char *table, const char *mass);
/* the processed c5z[245] and cmass[9] (internal units) */
void ora_ps_coefficients(double *c5z, double *cmass);
/* ps_type_couple for one water: bond lengths, unit O->H vectors, angle -> energy,
 * forces on H1 (fdum[0..2]) and H2 (fdum[3..5]), virial */
void ora_ps_couple(double roh1, const double *u1, double roh2, const double *u2, double theta,
                   double *pterm, double *fdum, double *vterm);
/* damping function f_n(x) = 1 - exp(-x) sum_{k<=n} x^k/k! (tang-toennies.f90) */
double ora_tang_toennies(int n, double x);

typedef struct {
    /* inputs */
    int imcon;
    double cell[9];
    int nw;                    /* waters; O,H1,H2 of water w at 3w..3w+2, M at 3nw+w */
    const double *x, *y, *z;   /* 4*nw positions */
    double rcut;               /* largest cutoff: sets the table grid (parset.f:210-215) */
    double rvdw;               /* dispersion cutoff */
    double c6[3], a6[3], c8[3], a8[3];   /* TT dispersion per pair type OO, OH, HH
                                            (internal energy units, 1/A) */
    int bond_key[2];           /* bndfrc key of the O-H bonds [0] and the H-H bond [1]:
                                  0 none, 1 harm, 2 mors, 7 expo */
    double bond_prm[2][3];     /* their prmbnd(1..3), energies in internal units */
    int ps_monomer, dispersion, long_range;   /* switches of the three terms */
    /* outputs: arrays accumulate, scalars are set */
    double *f;                 /* 3*4nw interleaved forces */
    double *atomic_energy;     /* optional 3nw */
    double *heat_w;            /* optional 3nw*9 W_i[a][b] */
    double *force_matrix;      /* optional dense [k][j][i] over 4nw sites */
    double eng_bnd, vir_bnd;   /* monomer bonds */
    double eng_ang, vir_ang;   /* PS monomer coupling terms */
    double eng_srp, vir_srp;   /* dispersion pairs */
    double elrc, virlrc;       /* long-range correction */
    double stress[9];          /* bonds + PS + dispersion (accumulates) */
} ora_water_args;

/* Monomer bonds (bndfrc.f), PS monomer coupling (angfrc.f keya = 9), dispersion
 * (srfrce2.f) and lrcorrect for a box of MB-pol waters.  Returns 0, -1 (imcon), -3 (PS tables not loaded). */
int ora_water_terms(ora_water_args *a);

/* ---- heat current (heat_oracle.c) ------------------------------------------ */

/* compute_heat_flux2 (heatcurrent.f90:703-728) from the dense force matrix, layout
 * [k][j][i] over natms_f = 4*nw sites (3*nw real atoms first).  stress_ew1p is
 * natms*9 row-major [i][a][b] (NULL = zero). */
void ora_heat_flux_dense(int natms, int nreal, const double *force_matrix,
                         const double *x, const double *y, const double *z,
                         const double *vx, const double *vy, const double *vz,
                         const double *e_kin, const double *e_pot,
                         const double *stress_ew1p,
                         double jkin[3], double jpot[3], double jtot[3]);

/* Same flux from per-atom contracted tensors W_i[a][b] = sum_j F_ij,a (x_i-x_j)_b. */
void ora_heat_flux_tensor(int nreal, const double *heat_w,
                          const double *vx, const double *vy, const double *vz,
                          const double *e_kin, const double *e_pot,
                          const double *stress_ew1p,
                          double jkin[3], double jpot[3], double jtot[3]);

/* ---- ring polymer (ring_oracle.c) ------------------------------------------------- */

/* transformation_matrix (centroid_matrix.f:3-135); P x P row-major, [k][b] = M(k+1, b+1) */
void ora_ring_matrices(int P, int langevin, double *n2c, double *c2n, double *tc2n, double *lambda);
/* out[k] = sum_b M[k][b] in[b] per coordinate (centroid_matrix.f:140-476) */
void ora_ring_transform(int P, int ncoord, const double *M, const double *in, double *out);
/* centroid_spring_force.f:41-157 over all beads; x, f bead-major [P][3][natom] */
void ora_ring_spring(int P, int natom, double coeff, double fac3, double force_scale, const double *mass,
                     const double *x, const double *xc, double *f, double *epot_spring, double *epot_deriv);

/* ---- smooth particle-mesh Ewald, charges and dipoles (spme_oracle.c) --------------- */
/* M_p(w + t), M_p', M_p'' for t = 0..p-1 (bspgen.f:27-100); p even, 4..12 */
void ora_bspline(int p, double w, double *a, double *d, double *d2);
/* |b(m)|^2, m = 0..K-1 (bspcoe.f) */
void ora_spme_bmod(int p, int K, double *bmod);
/* SPME reciprocal energy, potential, field and force per site (Coulomb constant 1) */
double ora_spme_recip(const double *cell, int p, const int *K, double alpha, int n, const double *x,
                      const double *y, const double *z, const double *q, const double *mu, double *phi, double *E,
                      double *F);
/* the Ewald reciprocal sum over |m_a| <= kmax_a that SPME approximates */
double ora_ewald_recip(const double *cell, const int *kmax, double alpha, int n, const double *x, const double *y,
                       const double *z, const double *q, const double *mu, double *phi, double *E, double *F);

#ifdef __cplusplus
}
#endif

#endif
