// mbgpu_internal.h -- host-side declarations shared by the translation units of libmbgpu.
#pragma once
#include <hip/hip_runtime.h>

#include <stdlib.h>

#include <string>
#include <vector>
#include <type_traits>

#include "../../include/mbgpu.h"

#include "mbgpu_dev.h"

// Device allocation of every buffer of the library.  With MBGPU_POISON=1 in the environment
// each new floating-point buffer is filled with 0xff bytes (a NaN in every element), so that a
// kernel that reads an element it did not write during the same evaluation yields NaN
// instead of silently using whatever the memory held.  MBGPU_POISON=a:b poisons only the
// allocations numbered a <= serial < b (mbgpu_debug_alloc_serial counts them) and
// MBGPU_POISON=!a:b every allocation but those, which are zeroed: ways to find the buffer.
long mbgpu_alloc_serial();     // allocations made so far (every buffer, in order)
int mbgpu_poison_this(long serial);   // 0: leave, 1: fill with 0xff, 2: fill with zeros

template <typename T>
inline hipError_t mbgpu_alloc(T** p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    const long serial = mbgpu_alloc_serial();
    if constexpr (std::is_floating_point<T>::value) {
        const int how = e == hipSuccess ? mbgpu_poison_this(serial) : 0;
        if (how) e = hipMemset(*p, how == 1 ? 0xff : 0, n * sizeof(T));
    }
    return e;
}

// Launch wrappers for the generated polynomial kernels (pip2b.hip / pip3b.hip): the first
// *ncnt clusters (a device count) of buffers laid out with stride ncap, ncap lanes launched.
// xv: variables SoA [nvars][ncap]; part: SoA [nslices][nvars+1][ncap]; cs: coefficient stream.
void launch_pip2b_f64(const double* xv, double* part, const int* ncnt, int ncap, const double* cs, hipStream_t s);
void launch_pip3b_f64(const double* xv, double* part, const int* ncnt, int ncap, const double* cs, hipStream_t s);
// FP32 instantiations of the same code (mixed-precision mode)
void launch_pip2b_f32(const float* xv, float* part, const int* ncnt, int ncap, const float* cs, hipStream_t s);
void launch_pip3b_f32(const float* xv, float* part, const int* ncnt, int ncap, const float* cs, hipStream_t s);
int pip2b_nslices();
int pip3b_nslices();
int pip2b_stream_len();
int pip3b_stream_len();
const short* pip2b_stream_table();
const short* pip3b_stream_table();
const double* pip2b_stream_values();
const double* pip3b_stream_values();

// Device buffers of one MB-pol evaluation context.
// ---- link cells (cells.hip) ----
// particles from which a search takes the cells (below, scanning them all is faster: at
// 2 048 waters the scans of the MB-pol O-O lists and of the dispersion rows beat the cell
// build, the electrostatics rows gain a little)
#define CELLS_MIN_SITES 8192       // electrostatics neighbour rows
#define CELLS_MIN_ATOMS 16384      // dispersion rows
#define CELLS_MIN_WATERS 8192      // MB-pol O-O lists
#define CELLS_MAX_PER_AXIS 16
#define CELLS_MAX_CELLS (CELLS_MAX_PER_AXIS * CELLS_MAX_PER_AXIS * CELLS_MAX_PER_AXIS)
struct CellList {
    CellView V{};                  // what the kernels see (V.on = 0: no cells)
    int ncell = 0;
    int *key = nullptr, *sidx = nullptr, *start = nullptr;   // cell per particle, particles by cell
    int n_cap = 0, ncell_cap = 0;
};
// cells at least `cutoff` wide for n particles in the cell (rows: cell vectors); leaves
// V.on = 0 when an axis holds fewer than three or n < min_particles; -1 with err
int cells_plan(CellList& C, const double cell[9], double cutoff, int n, int min_particles, std::string& err);
// the cells of the first n particles (after cells_plan), coordinates of particle i at
// x[i * stride], y[i * stride], z[i * stride], on stream s; 0 or a HIP error
int cells_build(CellList& C, int n, const double* x, const double* y, const double* z, int stride, hipStream_t s);
void cells_free(CellList& C);

struct MbBuffers {
    int nw = 0, natms = 0;
    int cap2 = 0, cap3 = 0;        // neighbour capacities per water (6.5 A and 4.5 A lists)
    CellList cells;                // link cells of the oxygens (O-O lists)
    long dcap = 0, tcap = 0;       // cluster capacities
    // geometry
    double* mon = nullptr;         // [nw][9] imaged monomers
    int* list = nullptr;           // [3*nw] listttm2 (1-based atom ids)
    int* wbase = nullptr;          // [nw] checked 0-based atom offset of each water's O (water_base)
    double *x = nullptr, *y = nullptr, *z = nullptr;   // [natms] raw positions
    // neighbour lists
    int *nbr2 = nullptr, *cnt2 = nullptr, *off2 = nullptr;   // partners j > i within 6.5 A
    int *nbr2l = nullptr, *cnt2l = nullptr;                  // partners j < i within 6.5 A
    int *nbr3 = nullptr, *cnt3 = nullptr;                    // partners within 4.5 A (64 per water)
    int *tcnt = nullptr, *toff = nullptr;
    unsigned long long* kmask = nullptr;  // [nw][64] kept-pair masks of the trimer enumeration
    int* rowoff = nullptr;                // [nw][64] kept pairs before each row
    int* counts = nullptr;         // [4]: ndimers, ntrimers, error flags, spare
    int* dimers = nullptr;         // [dcap][2]
    int* trimers = nullptr;        // [tcap][3]
    // cluster data
    double *clus2 = nullptr, *xv2 = nullptr, *part2 = nullptr, *e2 = nullptr, *g2 = nullptr, *s2 = nullptr;
    double *clus3 = nullptr, *xv3 = nullptr, *part3 = nullptr, *e3 = nullptr, *g3 = nullptr, *s3 = nullptr;
    // accumulators
    long long* acc = nullptr;      // fixed-point sums: [2][13*natms] (high halves, low halves) of
                                   // forces (3n), energies (n), heat-flux tensors (9n)
    unsigned char* is_water = nullptr;  // natms (rounded up to 4)
    double* blk = nullptr;         // per-block partial scalars
    int nblk_cap = 0;
    double* scal = nullptr;        // [16] reduced scalars
    // outputs kept on the device for the heat flux
    double* atomic_energy = nullptr;   // [natms]
    double* heat_w = nullptr;          // [9*natms]
    double* fout = nullptr;            // [3*natms] MB-pol forces of the last call
    // coefficient streams
    double *cs2 = nullptr, *cs3 = nullptr;
    // mixed precision: FP32 variables, slice partials and coefficient streams for the
    // polynomial kernels (the slice sums, chain rule and accumulation stay FP64)
    int mixed = 0;
    float *xv2f = nullptr, *part2f = nullptr, *cs2f = nullptr;
    float *xv3f = nullptr, *part3f = nullptr, *cs3f = nullptr;
    long fcap2 = 0, fcap3 = 0;     // cluster capacities of the FP32 buffers
};

// ------------------------------------------------------------------ electrostatics ----

// Kernel-argument block of the polarizable electrostatics (passed by value).
struct ElecDev {
    int nw, n, nreal, nrow;        // waters, sites (4nw), real atoms (3nw), rows carrying dipoles
    CellGeom g;
    CellView cells;                // link cells of the sites (cells.on = 0: full scans)
    double rcut, alpha, gamma, q_o;
    double polr[3], polr2[3];      // O, H, M
    double athole, athole12, athole13, ascc, ascd;
    double toler, omega;
    double drewd;                  // B-function table spacing
    int compat_qd;                 // qdforce M column as the reference (n_water + imol)
    int capA, capP;                // neighbour-row capacities (all partners / polarizable partners)
    int K, nrun, kchunks;          // k vectors, runs of consecutive n, k chunks of the site-parallel passes
    int rpc, kmaxch;               // runs per k chunk, most k vectors in one chunk (final pass)
    int fchunks, frpc, fkmax;      // the same for the field passes
    int km[3], kstride;            // kmax per axis, phase-table stride (max kmax + 1)
    double scale, scal1, rvolm, self_dd;   // 4 rvolm r4pie0, 2 rvolm r4pie0, 2 pi/V, 4 r4pie0 a^3/(3 sqrt(pi))
};

// Device buffers of the electrostatics (owned by the context).
struct ElecBuffers {
    int n = 0, capA = 0, capP = 0, K = 0, nrun = 0, kchunks = 0, schunks = 0, ntab = 0, kstride = 0;
    int spc = 0, dchunks = 0, dspc = 0, rpc = 0, kmaxch = 0;   // site chunks of the structure-factor passes
    int fchunks = 0, frpc = 0, fkmax = 0;                      // k chunks of the field passes
    double *x = nullptr, *y = nullptr, *z = nullptr;          // [n] site positions
    double *chge = nullptr, *polr = nullptr, *polr2 = nullptr; // [n]
    double* dg = nullptr;                                      // [nw][6] DMS derivatives
    double* tab = nullptr;                                     // [(ntab+1)*4] B-function table
    int *nbrA = nullptr, *cntA = nullptr;                      // [n][capA] partners within rcut
    int *nbrP = nullptr, *cntP = nullptr;                      // [n][capP] polarizable partners
    double* tens = nullptr;                                    // [6][n*capP] dipole tensors
    double* tin = nullptr;                                     // [nw][6 pairs][6] intramolecular tensors
    double *dip = nullptr, *efk = nullptr, *emu = nullptr, *efdc = nullptr, *efdd = nullptr;  // [3n]
    double* dip0 = nullptr;                                    // [3n] start dipoles of the evaluation
    double *cg_p = nullptr, *cg_q = nullptr, *cg_r = nullptr, *cg_z = nullptr;                // [3n] solver vectors
    double *cg_u = nullptr, *cg_w = nullptr, *cg_m = nullptr, *cg_s = nullptr;                // [3n]
    double* cg_minv = nullptr;                                 // [nw][45] inverse preconditioner blocks
    double *cg_tmu = nullptr, *cg_smu = nullptr;               // [3n] T mu, [K][6] S(mu) carried by the solver
    double *cg_sp = nullptr, *cg_su = nullptr, *cg_sq = nullptr;   // [K][6] S(p), S(u), S(q) carried by the solver
    double* cgblk = nullptr;                                   // block partials of the solver kernels
    double* ksum = nullptr;                                    // [16][n] chunk sums of the final pass
    double* state = nullptr;                                   // [16] solver scalars
    double* kv = nullptr;                                      // [K][8] kx ky kz akk akv akv2 . .
    int* runs = nullptr;                                       // [nrun][4] l, m, n0, count; k0 in runk
    int* runk = nullptr;                                       // [nrun] index of the run's first k
    int* klmn = nullptr;                                       // [K] l << 16 | (m+128) << 8 | (n+128)
    double *phc = nullptr, *phs = nullptr;                     // [3][kstride][n] cos / sin of l*2pi*s_a
    double *S = nullptr, *Spart = nullptr;                     // [K][8], [schunks][K][8]
    double* Sq = nullptr;                                      // [K][2] charge structure factors
    double *kc = nullptr, *kf = nullptr;                       // [K][16] per-k constants of the passes
    double* zero_state = nullptr;                              // [8] zeros (solver state of final passes)
    double* Fpart = nullptr;                                   // [kchunks][16][n] site partials
    double *vk = nullptr, *vr = nullptr, *vs = nullptr, *vc = nullptr, *vdck = nullptr, *vdcr = nullptr,
           *vesp = nullptr;                                    // [n]
    double *fk = nullptr, *f2 = nullptr, *f3 = nullptr, *fsite = nullptr;  // [3n]
    double *e1 = nullptr, *e2 = nullptr, *e3 = nullptr, *ep = nullptr;     // [n]
    double *s1 = nullptr, *w2 = nullptr, *w3 = nullptr;        // [9n]
    double *fmsum = nullptr, *fmx = nullptr;                   // [3n], [9n] M-row sums of ewald2p
    double *eout = nullptr, *wout = nullptr, *sout = nullptr;  // [nreal], [9 nreal] x2: per real atom
    double* blk = nullptr;                                     // [nblk][16] block partials
    int nblk = 0;
    double* scal = nullptr;                                    // [64] reduced scalars (41 used)
    int* err = nullptr;                                        // [4] device error flags
    double* pk = nullptr;                                      // [64] packed read-back (EL_PK_N used)
};

struct SpmeBuffers;
struct SpmeGeom;
// the particle-mesh reciprocal space of the electrostatics (null: the Ewald sum over k)
struct ElecSpme {
    SpmeBuffers* B;
    const SpmeGeom* G;
    double kap;          // V / (8 pi): from the full-space mesh sums to the half-space k sums
};
int el_launch_setup(ElecBuffers& E, const ElecDev& P, CellList& cells, hipStream_t s);
int el_launch_prep(ElecBuffers& E, const ElecDev& P, const double* rcell_dev, const ElecSpme* sp, hipStream_t s);
int el_launch_iteration(ElecBuffers& E, const ElecDev& P, int first, const ElecSpme* sp, hipStream_t s);
int el_launch_final(ElecBuffers& E, const ElecDev& P, int distribute, double g34, double gln34, int refresh,
                    const ElecSpme* sp, hipStream_t s);
int el_launch_cg_init(ElecBuffers& E, const ElecDev& P, const ElecSpme* sp, hipStream_t s);
int el_launch_cg_iteration(ElecBuffers& E, const ElecDev& P, int it, const ElecSpme* sp, hipStream_t s);
int el_time_sfac(ElecBuffers& E, const ElecDev& P, int reps, hipEvent_t a, hipEvent_t b, hipStream_t s);
void el_launch_zero(double* p, long n, hipStream_t s);
void el_launch_add(double* dst, const double* src, long n, hipStream_t s);
void el_launch_copy(double* dst, const double* src, long n, hipStream_t s);
// nseg <= EL_NSEG copies (a null source clears its destination) in one launch.
constexpr int EL_NSEG = 6;
void el_launch_copy_segs(int nseg, double* const* dst, const double* const* src, const long* n, hipStream_t s);
// nseg <= EL_NSEG sums dst_k = a_k + b_k in one launch.
void el_launch_sum_segs(int nseg, double* const* dst, const double* const* a, const double* const* b, const long* n,
                        hipStream_t s);
// Read-back of one electrostatics evaluation packed in doubles: the solver state (8), the
// error flag, the 41 reduced scalars, the particle mesh's two input flags (0 without one).
constexpr int EL_PK_ST = 0, EL_PK_ERR = 8, EL_PK_SCAL = 9, EL_PK_SPF = 50, EL_PK_N = 52;
void el_launch_pack(const double* state, const int* err, const double* scal, const int* spflag, double* out,
                    hipStream_t s);

// Host-side state of the electrostatics of one context.
struct ElecState {
    ElecBuffers E;
    ElecDev P{};
    int recip = 0;                      // 0 Ewald sum over k, 1 particle mesh (order, grid)
    int sp_order = 0, sp_grid[3] = {0, 0, 0};
    SpmeBuffers* sp = nullptr;          // particle-mesh buffers and geometry (recip = 1)
    SpmeGeom* spg = nullptr;
    CellList cells;                     // link cells of the sites (neighbour rows)
    bool ready = false;                 // the last evaluation completed
    int last_iters = 0;                 // solver iterations of the last evaluation
    int nw = 0;
    double tab_alpha = -1.0, tab_rcut = -1.0;           // key of the B-function table
    double k_cell[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};     // key of the k-vector list
    int k_kmax[3] = {0, 0, 0};
    double k_alpha = -1.0, k_rcut = -1.0;
    double* rcell = nullptr;            // [9] device copy of the inverse cell
    double scal[48];                    // host copy of the reduced scalars
    double* hpk = nullptr;              // pinned host copy of E.pk (EL_PK_N doubles)
    double out[20];                     // eng_ew1p, eng_ew2p, eng_ew3p, eng_polar, eng_cpe, vir_cpe,
                                        // vir_ew1p, vir_qd, iterations, rms change, stress[9], pairs
    // One evaluation's device work captured as a graph and replayed while its key (the
    // launch parameters and every buffer address) is unchanged: elec_evaluate.
    int graphs = 1;                     // 0: always launch kernel by kernel
    hipGraphExec_t gexec = nullptr;
    std::vector<unsigned char> gkey;
    int gbatch = 0;                     // solver iterations in the graph
    long gcaptures = 0, greplays = 0;   // graphs captured, replays launched
};

int elec_configure(ElecState& S, int nw, const CellGeom& g, const double* cell, const mbgpu_elec_params* p,
                   std::string& err);
int elec_evaluate(ElecState& S, int max_iter, int distribute, int solver, hipStream_t s, std::string& err);
void elec_release(ElecState& S);
void elec_drop_graph(ElecState& S);
// pieces of an evaluation for a caller that enqueues it with other work (one graph of a
// complete evaluation): the solver iterations a capture holds, the device work, the key
// bytes of that work, and the host side of the end
int elec_batch(const ElecState& S);
int elec_enqueue_full(ElecState& S, int distribute, int solver, int gb, hipStream_t s);
void elec_key_append(const ElecState& S, int distribute, int solver, int gb, std::vector<unsigned char>& k);
int elec_finish(ElecState& S, const double st[8], const int spflag[2], std::string& err);
// A graph capture holds the library's entry gate alone (mbgpu_api.hip): call around it.
void capture_gate_begin();
void capture_gate_end();

// ------------------------------------------------------------- monomer and dispersion ----

// Kernel-argument block of the water terms (passed by value).
struct WaterDev {
    int nw, nr;                    // waters, real atoms (3nw)
    CellGeom g;
    CellView cells;                // link cells of the real atoms (cells.on = 0: full scans)
    int bond_key[2];               // O-H, H-H bond keys (0 none, 1 harm, 2 mors, 7 expo)
    double bond_prm[2][3];
    int ps_monomer, dispersion;    // switches
    int tmask;                     // bit t: dispersion table of pair type t is non-zero
    int ngrid;                     // table row stride (mxgrid + 3)
    double rc2, rdr;               // rvdw^2, 1/dlrpot
};

// Device buffers of the water terms (owned by the context).
struct WaterBuffers {
    double* mono = nullptr;        // [nr][13] per-atom monomer terms: f3, e, W9
    double* wsc = nullptr;         // [nw][10] per-water scalars
    double* asc = nullptr;         // [nr][8] per-atom dispersion scalars
    double* fw = nullptr;          // [3][nr] forces of the water terms
    double *eout = nullptr, *wout = nullptr;   // [nr], [nr][9] per-atom energy and W
    double *vvv = nullptr, *ggg = nullptr;     // [3][ngrid] dispersion tables
    double *psc = nullptr, *pcm = nullptr;     // processed PS coefficients [245], [9]
    int *pidx = nullptr, *pidxm = nullptr;     // packed exponents i1 | i2 << 8 | i3 << 16
    double* scal = nullptr;        // [24] reduced scalars
    int ngrid = 0;
};

struct WaterState {
    WaterBuffers B;
    WaterDev P{};
    CellList cells;                // link cells of the real atoms (dispersion rows)
    bool ready = false;            // the last evaluation completed
    int nw = 0;
    double tab_key[14];            // rcut, rvdw, c6, a6, c8, a8 of the tables
    bool have_tab = false;
    double out[8];                 // eng_bnd, vir_bnd, eng_ang, vir_ang, eng_srp, vir_srp, elrc, virlrc
    double stress[9];              // bonds + PS + dispersion
};

int water_configure(WaterState& S, int nw, const CellGeom& g, const double* cell, const mbgpu_water_params* p,
                    hipStream_t s, std::string& err);
int water_evaluate(WaterState& S, const double* x, const double* y, const double* z, double* fx, double* fy,
                   double* fz, hipStream_t s, std::string& err);
int water_enqueue(WaterState& S, const double* x, const double* y, const double* z, double* fx, double* fy,
                  double* fz, hipStream_t s);
void water_finish(WaterState& S, const double r[18]);
void water_release(WaterState& S);

// ---- smooth particle-mesh Ewald, charges and dipoles (spme.hip) ----
#define SPME_MAX_ORDER 12
#define SPME_MAX_GRID 256
#define SP_FFT_THREADS 128     // threads of a transform workgroup
#define SP_FFT_ELEMS 256       // L lines x K elements staged per transform workgroup
struct SpmeGeom {
    double rv[3][3];   // reciprocal vectors a*_a (rows), a*_a . a_b = delta_ab
    double vol;
    double alpha;
    int K[3];
    int p;
    int rows;          // grid rows per spreading tile (a divisor of K[1], at most 16)
};
// one axis of the grid as lines for the transform passes (see k_sp_fft): K elements per
// line (K = 2^a 3^b 5^c), element stride se, ni lines of stride si in each of no outer
// blocks of stride so, L lines per workgroup (L divides ni); the radices of the passes
#define SPME_MAX_FACTORS 8
struct SpmeAxis {
    int K, se, ni, si, no, L;
    long so;
    int nf, fac[SPME_MAX_FACTORS];
};
struct SpmeBuffers {
    int4* gi = nullptr;                    // per site: grid index of t = 0 per axis, ok
    double* th = nullptr;                  // per site and axis: M_p, M_p', M_p'' (SPME_MAX_ORDER each)
    int* bin_off = nullptr;                // [K1 + 1] offsets of the grid planes' site lists
    int* bin_site = nullptr;               // sites by grid plane
    long n_cap = 0;
    double2* grid = nullptr;               // complex grids, ng_cap points apart
    double* bmod = nullptr;                // |b(m)|^2 of the three axes, concatenated
    double2* tw = nullptr;                 // exp(-2 pi i t / K_a), t < K_a, per axis
    double* part = nullptr;                // energy block partials, [grid][block]
    double* out = nullptr;                 // energies, one per grid
    int* flag = nullptr;                   // input errors found on the device (2 ints)
    long ng_cap = 0;                       // points per grid the allocation holds
    int ngr_cap = 0;                       // grids the allocation holds
    int cfg[4] = {0, 0, 0, 0};             // order and grid of bmod and tw
    int twoff[3] = {0, 0, 0};
    SpmeAxis axis[3] = {};
};
// The density spread onto grid g: q[i] (charge) when q_on[g], the dipole mu_i when mu_on[g],
// and the component comp[g] of mu_i as a charge when comp[g] >= 0.
#define SP_MAX_GRIDS 11
struct SpSpec {
    int ngr;
    signed char q_on[SP_MAX_GRIDS], mu_on[SP_MAX_GRIDS], comp[SP_MAX_GRIDS];
    // > 0: the dipoles are spread times state[scale_state] (a power of two, so exactly; 0 or
    // less there means 1), which the field pass divides out again -- for dipole-only spreads
    // of small vectors, whose contributions would otherwise fall below the fixed point's
    // resolution
    int scale_state;
};
// Geometry, tables, transform plans and buffers for n sites and ngr grids; 0 or -1 with err.
int spme_plan(SpmeBuffers& B, SpmeGeom& G, const double cell[9], int order, const int grid[3], double alpha, int n,
              int ngr, std::string& err);
void spme_free(SpmeBuffers& B);
inline long spme_points(const SpmeGeom& G) { return (long)G.K[0] * G.K[1] * G.K[2]; }
// Device steps on stream s; `state` (may be null) is the electrostatics solver state, whose
// done flag (state[2] != 0) turns the steps into no-ops.
void sp_launch_prep(SpmeBuffers& B, const SpmeGeom& G, int n, const double* x, const double* y, const double* z,
                    hipStream_t s);
// fwdz: the grids leave the spreading with their axis-2 forward pass done
void sp_launch_spread(SpmeBuffers& B, const SpmeGeom& G, int n, const double* q, const double* mu, const SpSpec& spec,
                      double2* grids, bool fwdz, const double* state, hipStream_t s);
// zdone: skip the axis-2 forward pass (the spreading did it)
void sp_launch_fft(SpmeBuffers& B, const SpmeGeom& G, double2* grids, int ngr, bool inverse, const double* state,
                   hipStream_t s, bool zdone);
// the whole convolution of ngr grids in place: forward transform, Q^ *= B C, backward
// transform (no energy; the axis-0 passes and the influence function in one step)
void sp_launch_convolve(SpmeBuffers& B, const SpmeGeom& G, double2* grids, int ngr, bool zdone, const double* state,
                        hipStream_t s);
// in place on ngr grids: Q^ *= B C; with part, the energy partials of each grid
void sp_launch_influence(SpmeBuffers& B, const SpmeGeom& G, double2* grids, int ngr, bool energy,
                         const double* state, hipStream_t s);
// energies of the first ngr grids from their partials into B.out[0..ngr)
void sp_launch_usum(SpmeBuffers& B, const SpmeGeom& G, int ngr, hipStream_t s);
// from the transformed grid src: src *= B C in place, dst[t] = B C akv2 k_a k_b src (t: xx
// xy xz yy yz zz, akv2 = 2 (1 / k^2 + 1 / (4 alpha^2)); dst grids ng_cap points apart)
void sp_launch_influence_stress(SpmeBuffers& B, const SpmeGeom& G, double2* src, double2* dst,
                                hipStream_t s);
// E = -grad phi of grid y at the first n sites into out[(3 y + c) * ostride + j] * fac
void sp_launch_field(SpmeBuffers& B, const SpmeGeom& G, int n, const double2* grids, int ngr, double fac,
                     double* out, long ostride, const double* state, hipStream_t s, int scale_state = 0);
// every per-site reciprocal quantity of the polarizable electrostatics from the eleven
// transformed grids of the final pass, as the k-sum site pass's partial sums (times kap):
// ks[t * n + j] (t < 15) and per-block partials blk[8 * b + v] (see k_sp_elec_final)
void sp_launch_elec_final(SpmeBuffers& B, const SpmeGeom& G, int n, const double* q, const double* mu,
                          const double2* grids, double kap, double* ks, double* blk, hipStream_t s);
int spme_launch(SpmeBuffers& B, const SpmeGeom& G, int n, const double* x, const double* y, const double* z,
                const double* q, const double* mu, double* phi, double* ef, double* f, hipStream_t s);
