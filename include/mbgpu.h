/*
 * mbgpu.h -- C ABI of libmbgpu.so, the MI355X (gfx950) MB-pol force / heat-current path.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success and a nonzero
 * status otherwise; mbgpu_last_error() gives a message for the calling thread's last
 * failure.  Units follow the reference: Angstrom in, DL_POLY internal energy units
 * (10 J/mol; engunit = 418.4 per kcal/mol) out, except the leaves, which work in kcal/mol
 * like the reference's C++ leaves.
 *
 * Ownership: a context owns its device buffers (created by mbgpu_create, released by
 * mbgpu_destroy); caller arrays stay the caller's.  One context per host thread.
 */
#ifndef MBGPU_H
#define MBGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mbgpu_ctx mbgpu_ctx;

/* ------------------------------------------------------------------------------------
 * Leaf entry points: identical signatures and conventions to the reference's C ABI.
 * Each call evaluates one cluster on the GPU (device 0, the library's own stream).
 * ---------------------------------------------------------------------------------- */

/* Replaces mbpol_2b_poly_ (mbpol-2b-poly.cpp:1878-1893), called from do_dimer
 * (mbpol.f90:517).  w1, w2: 9 doubles each (O, H1, H2; x y z; Angstrom), already imaged
 * into one cluster.  E in kcal/mol; g1, g2 (9 each) are overwritten with dE/dx, except
 * when r_OO > 6.5 A, where E = 0 and g1, g2 are left untouched (as the reference).
 * force_matrix (108 doubles) is not written: the reference fills it only partially and
 * nothing on the active path reads it (mbpol.f90:533-556 uses g instead). */
void mbpol_2b_poly_(const double* w1, const double* w2, double* E, double* g1, double* g2,
                    double* force_matrix);

/* Replaces mbpol_3b_poly_ (mbpol-3b-poly.cpp:1518-1531), called from do_trimer
 * (mbpol.f90:894).  g1..g3 are ACCUMULATED (the caller zeroes them, mbpol.f90:890-892).
 * force_matrix (243 doubles) is not written (see above). */
void mbpol_3b_poly_(const double* w1, const double* w2, const double* w3, double* E, double* g1,
                    double* g2, double* g3, double* force_matrix);

/* Replace mbpol_2b_cutoff_ / mbpol_3b_cutoff_ (mbpol-2b-poly.cpp:1900-1904,
 * mbpol-3b-poly.cpp:1538-1542), called from mbpol_init (mbpol.f90:64-65): 6.5 and 4.5. */
void mbpol_2b_cutoff_(double* r);
void mbpol_3b_cutoff_(double* r);

/* Spellings without the trailing underscore (the reference's -DBGQ build,
 * mbpol-2b-poly.cpp:1884-1901, mbpol-3b-poly.cpp:1522-1539). */
void mbpol_2b_poly(const double* w1, const double* w2, double* E, double* g1, double* g2,
                   double* force_matrix);
void mbpol_3b_poly(const double* w1, const double* w2, const double* w3, double* E, double* g1,
                   double* g2, double* g3, double* force_matrix);
void mbpol_2b_cutoff(double* r);
void mbpol_3b_cutoff(double* r);

/* ------------------------------------------------------------------------------------
 * Context (stands in for mbpol_init / mbpol_fini, mbpol.f90:52-94, and init_heat,
 * heatcurrent.f90:47-110, for the MB-pol part of the decomposition).
 * ---------------------------------------------------------------------------------- */

/* nw waters, natms sites in the per-atom arrays (4*nw with M sites).  Returns NULL on
 * failure (no GPU, out of memory, tables missing); see mbgpu_last_error(). */
mbgpu_ctx* mbgpu_create(int nw, int natms);
/* mbgpu_create with a choice of hardware queues: dedicated = 1 creates the context's two
 * streams (MB-pol and water terms; electrostatics) with hardware queues of their own, so
 * that they always run side by side, for one latency-bound context (an MD run); dedicated =
 * 0 is mbgpu_create (queues shared among all streams of the process, the better choice
 * for many concurrent contexts).  Results are identical either way.  No counterpart in the
 * reference. */
mbgpu_ctx* mbgpu_create_queues(int nw, int natms, int dedicated);
void mbgpu_destroy(mbgpu_ctx* ctx);
const char* mbgpu_last_error(void);

/* Batched MB-pol forces: stands in for mbpol_forces (mbpol.f90:98-449), called from
 * dlpoly_cmd.f:2504, single rank.  Host arrays, Fortran semantics:
 *   listttm2: 3*nttm2/3 1-based atom ids (O, H, H per water; consecutive ids per water);
 *   x/y/zzz: natms positions (Angstrom);
 *   f{x,y,z}: += MB-pol forces (internal units);
 *   engsrp, virsrp: += 2B+3B energy and virial; stress[9]: += stress.
 * imcon 1, 2, 3 as image_vectors (mbpol.f90:1319-1357); imcon 0 is an isolated cluster
 * (no images; the cell is not used; the reference stops in fixme there, mbpol.f90:1357,
 * so these semantics are this library's); other values return an error.  The per-atom energies (E/6 per dimer atom, E/9 per
 * trimer atom) and the per-atom heat-flux tensors stay on the device for
 * mbgpu_heat_flux and can be read with mbgpu_get_decomposition. */
int mbgpu_mbpol_forces(mbgpu_ctx* ctx, int imcon, const double cell[9], int nttm2, const int* listttm2,
                       const double* xxx, const double* yyy, const double* zzz, double* fxx, double* fyy,
                       double* fzz, double* engsrp, double* virsrp, double stress[9]);

/* Same with device pointers (inputs resident in HBM; f{x,y,z} device arrays or NULL).
 * Every _dev entry reads and writes its device arrays on the context's own streams and
 * waits for them before it returns; the caller finishes writing the inputs first (its
 * stream synchronized), since those streams are not ordered after other streams' work.
 * The device list is checked on the device before any atom array is read with it; a bad
 * entry returns -5 and no forces are added. */
int mbgpu_mbpol_forces_dev(mbgpu_ctx* ctx, int imcon, const double cell[9], int nttm2,
                           const int* listttm2_dev, const double* x_dev, const double* y_dev,
                           const double* z_dev, double* fx_dev, double* fy_dev, double* fz_dev,
                           double* engsrp, double* virsrp, double stress[9]);

/* The check both mbgpu_mbpol_forces entries apply to listttm2 (host array here, no GPU
 * needed): every water's three entries must be consecutive atom ids a, a+1, a+2 with
 * 1 <= a and a+2 <= natms, as DL_POLY's FIELD layout gives them (the reference reads
 * listttm2 at mbpol.f90:184-188 and addresses a water's atoms as o1, o1+1, o1+2 with
 * o1 = listttm2(3i-2), mbpol.f90:229,526-544, without checking either).  Returns 0 when the
 * list is valid, the 1-based index of the first bad water otherwise, -1 for a bad nttm2
 * (negative or not a multiple of 3) or a NULL list.  No counterpart in the reference. */
int mbgpu_check_listttm2(int nttm2, const int* listttm2, int natms);

/* Per-atom results of the last mbgpu_mbpol_forces call (host arrays of natms, natms and
 * 9*natms, any may be NULL):
 *   atomic_energy_mbpol  (heatcurrent.f90 update_energy_mbpol sums),
 *   heat_w[i][a][b] = sum_j F_ij,a (x_i - x_j)_b over the MB-pol force-matrix entries
 *     (the contraction compute_heat_flux2 makes, heatcurrent.f90:716-719),
 *   forces: the MB-pol forces alone, 3*natms interleaved. */
int mbgpu_get_decomposition(mbgpu_ctx* ctx, double* atomic_energy_mbpol, double* heat_w, double* forces);

/* Precision of the MB-pol polynomials in this context, a bit mask: 0 = FP64 (default,
 * the reference's arithmetic); bit 0 = the 2B polynomial, bit 1 = the 3B polynomial in
 * FP32 (values and gradients; variables rounded from FP64), with the slice sums, switches,
 * chain rule, per-atom accumulation, virial and stress in FP64.  No counterpart in the reference; the electrostatics, water terms and
 * heat flux are FP64 in both modes. */
int mbgpu_set_precision(mbgpu_ctx* ctx, int mode);

/* Dimer / trimer sets of the last call (0-based water indices; dimers i<j, trimers
 * i<j<k, in the order evaluated).  *nd, *nt receive the counts; arrays may be NULL. */
int mbgpu_get_lists(mbgpu_ctx* ctx, int* dimers, long dcap, long* nd, int* trimers, long tcap, long* nt);

/* Heat flux: stands in for compute_heat_flux2 (heatcurrent.f90:703-728) restricted to the
 * terms this library computes, plus caller-supplied per-atom terms:
 *   vxx..vzz  velocities (natms, only the first nreal = 3*nw are used),
 *   e_kin     per-atom kinetic energy (update_kinetic_energy sums), nreal,
 *   e_pot_extra  other per-atom potential energies to add (NULL = none), nreal,
 *   w_extra      other per-atom tensors W_i (9 per atom, NULL = none),
 *   stress_ew1p  per-atom reciprocal-space stress (9 per atom, NULL = none).
 * jkin, jpot, jtot receive the three components (internal units, not scaled by junit). */
int mbgpu_heat_flux(mbgpu_ctx* ctx, const double* vxx, const double* vyy, const double* vzz,
                    const double* e_kin, const double* e_pot_extra, const double* w_extra,
                    const double* stress_ew1p, double jkin[3], double jpot[3], double jtot[3]);

/* Same with device pointers (arrays of nreal = 3*nw; extras may be NULL). */
int mbgpu_heat_flux_dev(mbgpu_ctx* ctx, const double* vx_dev, const double* vy_dev, const double* vz_dev,
                        const double* e_kin_dev, const double* e_pot_extra_dev, const double* w_extra_dev,
                        const double* stress_ew1p_dev, double jkin[3], double jpot[3], double jtot[3]);

/* ------------------------------------------------------------------------------------
 * Polarizable electrostatics (MB-pol / TTM4-F induced dipoles with Ewald sums).
 * ---------------------------------------------------------------------------------- */

/* Parameters of the polarizable Ewald electrostatics.  The reference reads them from
 * CONTROL (ewald sum, cut, dipole tolerance, sor_omega), the vmbpol preset
 * (simdef_cmd.f:1036-1058: gamma, athole, athole12, athole13, ascc, ascd, nthole = 4) and
 * FIELD (site polarizabilities, oxygen charge). */
typedef struct {
    double rcut;              /* real-space cutoff (A) */
    double alpha;             /* Ewald convergence parameter (1/A) */
    int kmax[3];              /* Ewald kmax1..3 */
    double toler;             /* dipole tolerance, e*A (the reference converts from Debye) */
    double omega;             /* sor_omega: weight of the old dipole in the SOR update */
    int max_iter;             /* error after this many iterations (the reference: 22581) */
    double gamma;             /* M-site position parameter (0.426706882) */
    double q_o;               /* oxygen charge (0) */
    double polr[3];           /* smearing polarizabilities of O, H, M (A^3) */
    double polr2[3];          /* induction polarizabilities of O, H, M (A^3; M must be 0) */
    double athole, athole12, athole13, ascc, ascd;
    int compat_qd_column;     /* 1: the heat-current force-matrix column of the M-site
                                 charge-derivative forces is n_water+imol as in
                                 qdforce.f90:332-337; 0: the M site's own column */
    int distribute_msite;     /* 1: move the M-site forces onto O, H1, H2 and zero them
                                 (dlpoly_cmd.f:2583-2597); 0: return them on the M sites */
    int solver;               /* 0: SOR exactly as iter.f (omega, toler as above);
                                 1: preconditioned conjugate gradients on the same linear
                                 system, stopped by the same test (the RMS of the step an
                                 SOR update would take is <= toler); same fixed point,
                                 fewer iterations.  No counterpart in the reference. */
} mbgpu_elec_params;

/* Stands in for the polarizable-Ewald part of forces (forces.f:268-404 with rfield,
 * ewald1p, iter, coulomb_polar; forces.f:585-610 ewald2p; forces.f:700-712 ewald3p;
 * forces.f:874-892 with qdforce) together with the M-site construction and DMS charges
 * that precede it (dlpoly_cmd.f:2041-2065, qdetermine.f90), single rank, Verlet lists.
 * Site layout as ttm2list.f builds it: O, H1, H2 of water w at 3w, 3w+1, 3w+2 (0-based),
 * M of water w at 3*nw + w; the context must have natms >= 4*nw.
 *   xxx..zzz: positions; the M-site entries are overwritten with the constructed sites.
 *   dipx..dipz: induced dipoles (4*nw), start values in (the reference keeps the previous
 *     step's dipoles), converged dipoles out.
 *   chge: receives the 4*nw site charges (may be NULL).
 *   fxx..fzz: += forces (internal units) at the sites, M-site forces moved to O, H1, H2
 *     when distribute_msite is set.
 *   engcpe, vircpe, stress: += electrostatic energy, virial and stress.
 *   niter: receives the number of solver iterations.
 * Per-atom energies, reciprocal-space stresses and force-matrix contractions stay on the
 * device and enter mbgpu_heat_flux automatically.  Errors: no convergence within
 * max_iter (the reference stops, forces.f:351-359), unsupported imcon, capacity.
 * imcon 0 (an isolated cluster; the reference rejects it, forces.f:271-273): no images and
 * no Ewald split -- alpha must be 0, so that the real-space kernels are the bare ones (the
 * -erf parts vanish, as ewald3p.f sets them below alpha 1e-6), there is no reciprocal sum
 * and no self term; every site pair within rcut interacts, so rcut should cover the
 * cluster; kmax is not used. */
int mbgpu_polar_electrostatics(mbgpu_ctx* ctx, int imcon, const double cell[9], const mbgpu_elec_params* p,
                               double* xxx, double* yyy, double* zzz, double* dipx, double* dipy, double* dipz,
                               double* chge, double* fxx, double* fyy, double* fzz, double* engcpe,
                               double* vircpe, double stress[9], int* niter);

/* Same with device pointers: x_dev..z_dev (4*nw, M entries written), dip_dev (3 per site,
 * interleaved, in/out; NULL = start from zero dipoles and keep the converged ones inside
 * the context), f_dev (3 per site, interleaved, +=; may be NULL). */
int mbgpu_polar_electrostatics_dev(mbgpu_ctx* ctx, int imcon, const double cell[9], const mbgpu_elec_params* p,
                                   double* x_dev, double* y_dev, double* z_dev, double* dip_dev, double* f_dev,
                                   double* engcpe, double* vircpe, double stress[9], int* niter);

/* Reciprocal space of the electrostatics in this context: mode 0 = the Ewald sum over k
 * (ewald1p.f, the reference's method for the polarizable model; default); mode 1 =
 * smooth particle-mesh Ewald of charges and induced dipoles with cardinal B-splines of
 * even order 4..12 on a grid[0] x grid[1] x grid[2] mesh (each size 2^a 3^b 5^c in
 * [max(8, order), 256]), used for the solver's dipole fields, the reciprocal energy,
 * forces, potentials, per-atom energies and per-atom stresses (the terms the heat current
 * takes from ewald1p.f:550-693, 771-825), with the same alpha.  The reference's own SPME
 * (ewald_spme.f, spme_for.f with bspgen.f, bspcoe.f, dlpfft3.f) handles point charges
 * only; this carries the dipoles through the spread density.  Takes effect at the next
 * electrostatics call.  No counterpart in the reference for the polarizable path. */
int mbgpu_set_elec_reciprocal(mbgpu_ctx* ctx, int mode, int order, const int grid[3]);

/* Graph replay of the electrostatics in this context: 1 (default) = once an evaluation has
 * set the solver's iteration count, later evaluations run as one captured HIP graph (the
 * start, that count + 2 solver iterations, the final pass) with one read-back, captured
 * again whenever the parameters, the cell or a buffer changes; results are bit-identical
 * to launching kernel by kernel.  0 = launch kernel by kernel.  No counterpart in the
 * reference (callers that evaluate configurations of one size many times: serving, MD). */
int mbgpu_set_graphs(mbgpu_ctx* ctx, int on);
/* counts[0]: graphs the electrostatics captured in this context, counts[1]: replays. */
int mbgpu_get_graph_counts(mbgpu_ctx* ctx, long counts[2]);

/* Results of the last electrostatics call (host arrays, any may be NULL):
 *   scal[20]: eng_ew1p, eng_ew2p, eng_ew3p, eng_polar, eng_cpe, vir_cpe, vir_ew1p, vir_qd,
 *             iterations, last rms dipole change, stress[9] (in scal[10..18]), k vectors;
 *   chge, vesp: 4*nw site charges and electrostatic potentials;
 *   e_real: per-real-atom energy (ew1p + ew2p + ew3p + polar, M share included);
 *   w_real: 9 per real atom, force-matrix contraction (M rows included);
 *   s_real: 9 per real atom, per-atom reciprocal-space stress (M share included);
 *   f_site: 3 per site, forces of the last call. */
int mbgpu_get_elec_details(mbgpu_ctx* ctx, double* scal, double* chge, double* vesp, double* e_real,
                           double* w_real, double* s_real, double* f_site);
/* The invariants write_check prints to CHECK_ENERGY_STRESS (heat_check.f90:13-49, called
 * under HEAT_CHECK), from the per-atom arrays of the last evaluations in this context
 * (MB-pol, electrostatics and water terms, each included when its last call completed):
 *   out[0]     sum of e_kin_dev over the 3*nw real atoms (total_kinetic_energy; e_kin_dev
 *              may be NULL, then 0);
 *   out[1]     sum of the per-atom potential energies (total_potential_energy,
 *              potential_energy_per_atom, heatcurrent.f90:282-290);
 *   out[2]     sum of the per-atom reciprocal-space energies (total_energy_ew1p);
 *   out[3..11] sum of the per-atom stresses minus those of the reciprocal space
 *              (total_stress - total_stress_ew1p), row-major 3x3.  With the reference's
 *              build flags (HEAT_CURRENT and HEAT_CHECK without HEAT_STRESS) the only other
 *              per-atom stresses recorded are the angular terms of the DMS charge
 *              derivatives (qdforce.f90:504-531), so this is their sum.
 * Internal units; the caller divides by engunit and multiplies by strsunit as write_check
 * does.  No kernels: device-to-host copies of the per-atom arrays, summed on the host. */
int mbgpu_heat_check(mbgpu_ctx* ctx, const double* e_kin_dev, double out[12]);

/* One velocity-Verlet stage on the 3*nw real atoms (device arrays, internal units: amu,
 * Angstrom, ps, 10 J/mol).  The reference integrates with the Verlet leapfrog of nve_0
 * (nve_0.f:50-110), which gives the same positions; this form has the full-step velocities
 * the heat flux needs.  F = (fx, fy, fz) [three arrays, or all NULL] + fsite [3 per site,
 * interleaved, or NULL].
 *   stage 1: v += dt/2 F/m; x += dt v; then F is cleared (the next force evaluation
 *            accumulates into it);
 *   stage 2: v += dt/2 F/m; ekin_dev[i] = m v^2/2 (may be NULL); *ekin_total = the sum.
 * The M sites are not moved: the next electrostatics call rebuilds them. */
int mbgpu_verlet_stage(mbgpu_ctx* ctx, int stage, double dt, const double* mass_dev, double* x_dev,
                       double* y_dev, double* z_dev, double* vx_dev, double* vy_dev, double* vz_dev,
                       double* fx_dev, double* fy_dev, double* fz_dev, double* fsite_dev, double* ekin_dev,
                       double* ekin_total);

/* Heat-current autocorrelation for Green-Kubo conductivities: acf[tau] = (1/(nt - tau)) *
 * sum_{t < nt - tau} J(t) . J(t + tau) for tau = 0 .. ntau-1, J stored [nt][3] (for example
 * the J_tot of every step of a run, in the units the caller keeps them).  The reference
 * leaves this to post-processing of its heatflux2 file (heatcurrent.f90:730-740); here it
 * runs on the device.  The _dev form takes device arrays; the host form copies through
 * temporary device buffers.  1 <= ntau <= nt. */
int mbgpu_heat_acf(mbgpu_ctx* ctx, const double* j, long nt, int ntau, double* acf);
int mbgpu_heat_acf_dev(mbgpu_ctx* ctx, const double* j_dev, long nt, int ntau, double* acf_dev);

/* ---- ring polymer (in-process beads) ----------------------------------------------
 * Beads are stored bead-major, a bead's coordinates as [3][natom] (its xxx, yyy, zzz
 * blocks).  Matrices are nbead x nbead, row-major, element [k][b] = M(k+1, b+1) of the
 * reference's Fortran arrays. */
#define MBGPU_RING_MAX_BEADS 256

/* Normal-mode matrices and eigenvalues of transformation_matrix (centroid_matrix.f:3-135):
 * langevin = 0 gives the reference's default ordering (cart_to_nmode = C^T / nbead),
 * langevin = 1 the orthonormal ordering of its Langevin branch.  Host computation; nbead
 * even, 2 .. MBGPU_RING_MAX_BEADS. */
int mbgpu_ring_matrices(int nbead, int langevin, double* nmode_to_cart, double* cart_to_nmode,
                        double* tcart_to_tnmode, double* lambda);
/* out[k][c] = sum_b M[k][b] in[b][c] over the ncoord coordinates of every bead: the
 * position / velocity / force transforms of centroid_matrix.f:140-476 (M = cart_to_nmode,
 * tcart_to_tnmode or nmode_to_cart; M on the host, bead arrays on the device). */
int mbgpu_ring_transform_dev(mbgpu_ctx* ctx, int nbead, long ncoord, const double* M, const double* in_dev,
                             double* out_dev);
/* Spring forces of centroid_spring_force.f:41-157 for all beads at once: with
 * coeff = nbead / (hbar beta)^2, the centroid xc ([3][natom], mode-1 coordinates times
 * fac1), f_k -= coeff m (2 x_k - x_{k-1} - x_{k+1}) fac3; *epot_spring = coeff/2 sum m
 * |x_k - x_{k-1}|^2 and *epot_deriv = 1/2 sum -(x_k - xc) . f_k / force_scale with f the
 * physical forces on entry.  Device arrays: mass [natom], x and f [nbead][3][natom]. */
int mbgpu_ring_spring_dev(mbgpu_ctx* ctx, int nbead, int natom, double coeff, double fac3, double force_scale,
                          const double* mass_dev, const double* x_dev, const double* xc_dev, double* f_dev,
                          double* epot_spring, double* epot_deriv);

/* mbgpu_ring_spring_dev without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and results as
 * mbgpu_ring_spring_dev; its scratch (the kernels' block partials) is a fixed device array
 * of MBGPU_NOALLOC_SCRATCH_DOUBLES doubles per GPU, held until the process ends, instead
 * of context buffers grown on demand.  It needs 2 * nbead * ceil(natom / 256) + 2 doubles;
 * larger calls return -6.  All calls on one GPU share that GPU's scratch: calls that
 * overlap on one GPU (from several threads or contexts) race. */
#define MBGPU_NOALLOC_SCRATCH_DOUBLES 32768
int mbgpu_ring_spring_dev_noalloc(mbgpu_ctx* ctx, int nbead, int natom, double coeff, double fac3,
                                  double force_scale, const double* mass_dev, const double* x_dev,
                                  const double* xc_dev, double* f_dev, double* epot_spring, double* epot_deriv);

/* ---- smooth particle-mesh Ewald reciprocal space ---------------------------------- */
#define MBGPU_SPME_MAX_ORDER 12
#define MBGPU_SPME_MAX_GRID 256
/* Reciprocal-space Ewald energy, potential, field and force of point charges q[n] and point
 * dipoles mu[3n] (NULL: none) by smooth particle-mesh Ewald: cardinal B-splines of even
 * order 4..12 (bspgen.f:27-100), |b(m)|^2 moduli (bspcoe.f), spreading, 3-D FFT, influence
 * function exp(-pi^2 m^2/alpha^2)/(pi V m^2) and gather as ewald_spme.f:1-362 and
 * spme_for.f do for charges, here with the dipoles carried in the spread density, so that
 * it converges to the reciprocal part of the induced-dipole Ewald sum (ewald1p.f).  grid:
 * powers of two in [max(8, order), 256].  Coulomb constant 1 (the caller multiplies by
 * r4pie0/epsq).  Outputs per site (device, each may be NULL): phi[n], the field E[3n] =
 * -grad phi and the force F[3n] = q E + (mu . grad) E, each including the site's own image
 * sum, as a reciprocal sum does; *energy = 1/2 sum (q phi - mu . E).  Deterministic: the
 * spread density is summed in 64-bit fixed point (2^-44 e). */
int mbgpu_spme_recip_dev(mbgpu_ctx* ctx, const double cell[9], int order, const int grid[3], double alpha, int n,
                         const double* x_dev, const double* y_dev, const double* z_dev, const double* q_dev,
                         const double* mu_dev, double* phi_dev, double* efield_dev, double* f_dev, double* energy);

/* Converged induced dipoles of the last electrostatics call: 3 per site, interleaved, 4*nw
 * sites (the reference's dipx/y/z after iter; e*Angstrom). */
int mbgpu_get_dipoles(mbgpu_ctx* ctx, double* dip);

/* Parameters of the monomer, dispersion and long-range terms (FIELD inputs in the
 * reference; energies in the internal unit, 10 J/mol). */
typedef struct {
    double rcut;          /* largest cutoff of the run: sets the dispersion table grid,
                             max(1000, int(rcut/1e-4 + 0.5) + 4) points (parset.f:210-215) */
    double rvdw;          /* dispersion cutoff, table spacing rvdw/(grid - 4) (forgen.f:97) */
    double c6[3], a6[3], c8[3], a8[3];   /* "tt68" vdw parameters (sysdef_cmd.f:1505-1510,
                                            1598-1602) of the O-O, O-H and H-H pairs */
    int bond_key[2];      /* bndfrc key of the two O-H bonds [0] and the H-H bond [1]:
                             0 none, 1 harm, 2 mors, 7 expo (sysdef_cmd.f:615-644) */
    double bond_prm[2][3];   /* their prmbnd(1..3) */
    int ps_monomer;       /* 1: Partridge-Schwenke coupling terms (angle key 9) */
    int dispersion;       /* 1: pair dispersion within rvdw */
    int long_range;       /* 1: long-range dispersion correction */
} mbgpu_water_params;

/* Stands in for the intramolecular and dispersion terms of a force evaluation of MB-pol
 * water: bndfrc (dlpoly_cmd.f:2422; bndfrc.f:136-330) for the monomer bonds, angfrc with
 * the Partridge-Schwenke branch (dlpoly_cmd.f:2430; angfrc.f:160-545 -> ps_type_couple,
 * ps_type_couple.f90:15-274), the dispersion part of the pair loop (forces.f:549-556 ->
 * srfrce2, srfrce2.f:60-281, with the forgen tables, forgen.f:90-100, 262-269, 338-357)
 * and lrcorrect (dlpoly_cmd.f:2388; lrcorrect.f:100-209).  Same site layout as
 * mbgpu_polar_electrostatics; only the 3*nw real atoms take part (the M sites have no
 * dispersion).  Dispersion pairs are all pairs of atoms of different waters within rvdw.
 *   fxx..fzz: += forces (internal units) on the real atoms.
 *   eng[4]  : set to engbnd, engang, engsrp (dispersion pairs) and elrc;
 *   vir[4]  : set to virbnd, virang, virsrp and virlrc;
 *   stress  : += bond, PS and dispersion stress (the caller adds -virlrc/3 to the diagonal,
 *             as dlpoly_cmd.f:2572-2574 does).
 * Per-atom energies (bond halves, pterm/3, dispersion halves; the long-range correction
 * has no per-atom share, its per-atom update is commented out in lrcorrect.f:226-232) and the
 * force-matrix contractions stay on the device and enter mbgpu_heat_flux automatically.
 * imcon 0 (an isolated cluster): no images and no long-range correction (lrcorrect.f:114). */
int mbgpu_water_terms(mbgpu_ctx* ctx, int imcon, const double cell[9], const mbgpu_water_params* p,
                      const double* xxx, const double* yyy, const double* zzz, double* fxx, double* fyy,
                      double* fzz, double eng[4], double vir[4], double stress[9]);

/* Same with device pointers (x_dev..z_dev at least 3*nw long; f*_dev +=, may be NULL). */
int mbgpu_water_terms_dev(mbgpu_ctx* ctx, int imcon, const double cell[9], const mbgpu_water_params* p,
                          const double* x_dev, const double* y_dev, const double* z_dev, double* fx_dev,
                          double* fy_dev, double* fz_dev, double eng[4], double vir[4], double stress[9]);

/* Results of the last water-terms call (host arrays of 3*nw, 9*3*nw and 3*3*nw entries,
 * any may be NULL): per-atom energy, force-matrix contraction W_i[a][b], forces. */
int mbgpu_get_water_details(mbgpu_ctx* ctx, double* e_real, double* w_real, double* f_real);

/* One complete evaluation with device arrays: MB-pol (mbgpu_mbpol_forces_dev), the
 * polarizable electrostatics (mbgpu_polar_electrostatics_dev, dipoles from dip_dev or from
 * zero when it is NULL, written back when given), the water terms (mbgpu_water_terms_dev) and
 * the heat flux over all of them (mbgpu_heat_flux_dev with no extras), in one call: the
 * electrostatics run on the context's second stream beside MB-pol and the water terms, and
 * the results come back in one read.  Forces stay in the context (mbgpu_get_decomposition,
 * mbgpu_get_elec_details, mbgpu_get_water_details).  Once an evaluation through the separate
 * entries has sized the buffers (the first call makes one), the device work is three captured
 * HIP graphs (the electrostatics on the second stream; MB-pol and the water terms on the
 * first; the per-atom sums, the heat flux and the packed scalars after both) replayed while
 * the parameters, the cell and the buffers stay the same (unless mbgpu_set_graphs(ctx, 0));
 * a list past its capacity, overflowed neighbour rows or a solve
 * longer than the graph holds repeats the evaluation through the separate entries.  Results
 * are bit-identical to the separate entries.  out[MBGPU_EVAL_OUT]:
 *   [0] MB-pol energy, [1] its virial, [2..10] its stress;
 *   [11] electrostatic energy (engcpe), [12] its virial, [13..21] its stress, [22] solver
 *        iterations;
 *   [23..26] water-term energies (bonds, angles/PS, dispersion, long-range tail),
 *   [27..30] their virials, [31..39] their stress;
 *   [40..42] J_kin, [43..45] J_pot, [46..48] J_tot; [49] 1 when one enqueue or replay ran,
 *   0 when the separate entries did.
 * No counterpart in the reference (its force routine calls the parts in turn, forces.f). */
#define MBGPU_EVAL_OUT 64
int mbgpu_evaluate_dev(mbgpu_ctx* ctx, int imcon, const double cell[9], int nttm2, const int* listttm2_dev,
                       const mbgpu_elec_params* ep, const mbgpu_water_params* wp, double* x_dev, double* y_dev,
                       double* z_dev, double* dip_dev, const double* vx_dev, const double* vy_dev,
                       const double* vz_dev, const double* e_kin_dev, double out[MBGPU_EVAL_OUT]);
/* mbgpu_evaluate_dev with forces, for integrators: fx/fy/fz_dev (natms entries each) receive
 * the MB-pol and water-term forces and fsite_dev (3 interleaved values per site) the
 * electrostatic forces (with distribute_msite, moved onto the real atoms).  The forces are
 * WRITTEN, not added: a call that repeats its evaluation through the separate entries
 * clears them first.  vx/vy/vz_dev and e_kin_dev may all be NULL: no heat flux then
 * (out[40..48] = 0).  Results are bit-identical to mbgpu_mbpol_forces_dev,
 * mbgpu_polar_electrostatics_dev and mbgpu_water_terms_dev called in turn on cleared force
 * arrays.  No counterpart in the reference (its force routine calls the parts in turn,
 * forces.f). */
int mbgpu_evaluate_forces_dev(mbgpu_ctx* ctx, int imcon, const double cell[9], int nttm2, const int* listttm2_dev,
                              const mbgpu_elec_params* ep, const mbgpu_water_params* wp, double* x_dev,
                              double* y_dev, double* z_dev, double* dip_dev, double* fx_dev, double* fy_dev,
                              double* fz_dev, double* fsite_dev, const double* vx_dev, const double* vy_dev,
                              const double* vz_dev, const double* e_kin_dev, double out[MBGPU_EVAL_OUT]);
/* counts[0]: complete evaluations captured as graphs in this context, counts[1]: replays. */
int mbgpu_get_evaluate_graph_counts(mbgpu_ctx* ctx, long counts[2]);
/* Total forces of the last complete evaluation on the 3*nw real atoms, f[3*i + q] (host
 * array): (MB-pol + electrostatics) + water terms per element, summed on the device in that
 * order, one copy.  Needs an evaluation of all three parts since the context was created or
 * resized.  The reference's counterpart is the summed fxx/fyy/fzz after forces
 * (dlpoly_cmd.f:2287-2603); no separate routine. */
int mbgpu_get_forces(mbgpu_ctx* ctx, double* f);


/* Timing helper: runs the MB-pol force evaluation `reps` times on device-resident inputs
 * already uploaded by a previous mbgpu_mbpol_forces(_dev) call and returns the mean
 * milliseconds per evaluation, measured with HIP events on the library's stream. */
int mbgpu_time_forces(mbgpu_ctx* ctx, int reps, double* ms_per_eval, double* ms_pip3b_kernel);

/* Timing helper: `reps` launches of the reciprocal-space dipole structure-factor kernel on
 * the dipoles of the last mbgpu_polar_electrostatics(_dev) call, timed with HIP events on
 * the library's stream.  Returns the mean milliseconds per launch and the algorithmic
 * FLOP of one launch: 18 per (k vector, polarizable site) -- six multiply-adds of the
 * dipole components into Re/Im S(k) and the complex phase recurrence along n. */
int mbgpu_time_elec(mbgpu_ctx* ctx, int reps, double* ms_sfac, double* flop_sfac);

#ifdef __cplusplus
}
#endif
#endif
