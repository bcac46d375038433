// elec_api.hip -- host-side set-up and driving of the electrostatics kernels
// (elec_kernels.hip): buffers, the B-function table, the k-vector list, the solver loop
// and the assembly of the scalar results.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mbgpu.h"
#include "mbgpu_internal.h"
#include <cstdlib>

#define R4PIE0 138935.4835
#define SQRPI 1.7724538509055159
#ifndef EL_DCHUNK_MAX
#define EL_DCHUNK_MAX 32      // site chunks of the dipole structure-factor pass (partials re-read by the sum)
#endif
#ifndef EL_SPLIT
#define EL_SPLIT 2  // threads per k run in the structure-factor pass (elec_kernels.hip)
#endif
#ifndef EL_KCHUNK_WAVES
#define EL_KCHUNK_WAVES 4096  // site-block x k-chunk product the reciprocal field passes aim at
#endif
#ifndef EL_KCHUNK_RUNS
#define EL_KCHUNK_RUNS 32     // at most this many k runs per chunk of the site-parallel passes
#endif
#ifndef EL_FCHUNK_WAVES
#define EL_FCHUNK_WAVES 3072  // wavefronts the reciprocal field passes aim at
#endif
#ifndef EL_FCHUNK_RUNS
#define EL_FCHUNK_RUNS 48     // at most this many k runs per chunk of the field passes
#endif
#ifndef EL_KG
#define EL_KG 4               // run groups (wavefronts) per field block (elec_kernels.hip)
#endif
#ifndef EL_DCHUNK_WAVES
#define EL_DCHUNK_WAVES 3072  // wavefronts the structure-factor passes aim at
#endif
#ifndef EL_SG
#define EL_SG 4               // site groups (wavefronts) per structure-factor block (elec_kernels.hip)
#endif
#define PI_D 3.141592653589793

namespace {

template <typename T>
hipError_t dalloc(T** p, size_t n)
{
    return mbgpu_alloc(p, n);
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

#define ECHK(x)                                                                            \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            err = std::string(#x) + ": " + hipGetErrorString(e_);                          \
            return (int)e_;                                                                \
        }                                                                                  \
    } while (0)

// erfcc.f with its default-REAL literals (float values), as the table builder uses it
double erfcc_host(double x)
{
    const double z = fabs(x);
    const double t = 1.0 / (1.0 + 0.5 * z);
    double r = t * exp(-z * z - (double)1.26551223f +
                       t * ((double)1.00002368f +
                            t * ((double).37409196f +
                                 t * ((double).09678418f +
                                      t * ((double)-.18628806f +
                                           t * ((double).27886807f +
                                                t * ((double)-1.13520398f +
                                                     t * ((double)1.48851587f +
                                                          t * ((double)-0.82215223f + t * (double).17087277f)))))))));
    return x < 0.0 ? 2.0 - r : r;
}

// gammln of gammq.f (Lanczos)
double gammln_host(double xx)
{
    static const double cof[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                  -1.231739572450155, .1208650973866179e-2, -.5395239384953e-5};
    double x = xx, y = x;
    double tmp = x + 5.5;
    tmp = (x + 0.5) * log(tmp) - tmp;
    double ser = 1.000000000190015;
    for (int j = 0; j < 6; ++j) {
        y = y + 1.0;
        ser = ser + cof[j] / y;
    }
    return tmp + log(2.5066282746310005 * ser / x);
}

void free_buffers(ElecBuffers& E)
{
    dfree(E.x); dfree(E.y); dfree(E.z); dfree(E.chge); dfree(E.polr); dfree(E.polr2); dfree(E.dg); dfree(E.tab);
    dfree(E.nbrA); dfree(E.cntA); dfree(E.nbrP); dfree(E.cntP); dfree(E.tens); dfree(E.tin);
    dfree(E.dip); dfree(E.dip0); dfree(E.efk); dfree(E.emu); dfree(E.efdc); dfree(E.efdd);
    dfree(E.cg_p); dfree(E.cg_q); dfree(E.cg_r); dfree(E.cg_z); dfree(E.cg_minv); dfree(E.cg_tmu); dfree(E.cg_smu); dfree(E.cgblk);
    dfree(E.cg_u); dfree(E.cg_w); dfree(E.cg_m); dfree(E.cg_s); dfree(E.cg_sp); dfree(E.cg_su); dfree(E.cg_sq); dfree(E.ksum); dfree(E.state); dfree(E.zero_state);
    dfree(E.kv); dfree(E.klmn); dfree(E.runs); dfree(E.runk); dfree(E.phc); dfree(E.phs);
    dfree(E.S); dfree(E.Spart); dfree(E.Sq); dfree(E.kc); dfree(E.kf); dfree(E.Fpart);
    dfree(E.vk); dfree(E.vr); dfree(E.vs); dfree(E.vc); dfree(E.vdck); dfree(E.vdcr); dfree(E.vesp);
    dfree(E.fk); dfree(E.f2); dfree(E.f3); dfree(E.fsite); dfree(E.e1); dfree(E.e2); dfree(E.e3); dfree(E.ep);
    dfree(E.s1); dfree(E.w2); dfree(E.w3); dfree(E.fmsum); dfree(E.fmx); dfree(E.eout); dfree(E.wout);
    dfree(E.sout); dfree(E.blk); dfree(E.scal); dfree(E.err); dfree(E.pk);
    E.n = E.capA = E.capP = E.K = E.nrun = 0;
}

// site-sized buffers (allocated once per context size)
int alloc_sites(ElecState& S, int nw, std::string& err)
{
    ElecBuffers& E = S.E;
    const int n = 4 * nw, nreal = 3 * nw;
    E.n = n;
    ECHK(dalloc(&E.x, n)); ECHK(dalloc(&E.y, n)); ECHK(dalloc(&E.z, n));
    ECHK(dalloc(&E.chge, n)); ECHK(dalloc(&E.dg, (size_t)6 * nw)); ECHK(dalloc(&E.tin, (size_t)36 * nw));
    ECHK(dalloc(&E.cntA, n)); ECHK(dalloc(&E.cntP, n));
    ECHK(dalloc(&E.dip, (size_t)3 * n)); ECHK(dalloc(&E.efk, (size_t)3 * n)); ECHK(dalloc(&E.emu, (size_t)3 * n));
    ECHK(dalloc(&E.dip0, (size_t)3 * n));
    ECHK(dalloc(&E.efdc, (size_t)3 * n)); ECHK(dalloc(&E.efdd, (size_t)3 * n));
    ECHK(dalloc(&E.cg_p, (size_t)3 * n)); ECHK(dalloc(&E.cg_q, (size_t)3 * n));
    ECHK(dalloc(&E.cg_r, (size_t)3 * n)); ECHK(dalloc(&E.cg_z, (size_t)3 * n));
    ECHK(dalloc(&E.cg_u, (size_t)3 * n)); ECHK(dalloc(&E.cg_w, (size_t)3 * n));
    ECHK(dalloc(&E.cg_m, (size_t)3 * n)); ECHK(dalloc(&E.cg_s, (size_t)3 * n));
    ECHK(dalloc(&E.cg_minv, (size_t)45 * nw));
    ECHK(dalloc(&E.cg_tmu, (size_t)3 * n));
    // the solver's vector update writes 5 partials per block of 7 waters
    ECHK(dalloc(&E.cgblk, (size_t)5 * ((nw + 6) / 7) + 64));
    ECHK(dalloc(&E.ksum, (size_t)16 * n));
    ECHK(dalloc(&E.state, 16)); ECHK(dalloc(&E.zero_state, 8));
    ECHK(dalloc(&E.vk, n)); ECHK(dalloc(&E.vr, n)); ECHK(dalloc(&E.vs, n)); ECHK(dalloc(&E.vc, n));
    ECHK(dalloc(&E.vdck, n)); ECHK(dalloc(&E.vdcr, n)); ECHK(dalloc(&E.vesp, n));
    ECHK(dalloc(&E.fk, (size_t)3 * n)); ECHK(dalloc(&E.f2, (size_t)3 * n)); ECHK(dalloc(&E.f3, (size_t)3 * n));
    ECHK(dalloc(&E.fsite, (size_t)3 * n));
    ECHK(dalloc(&E.e1, n)); ECHK(dalloc(&E.e2, n)); ECHK(dalloc(&E.e3, n)); ECHK(dalloc(&E.ep, n));
    ECHK(dalloc(&E.s1, (size_t)9 * n)); ECHK(dalloc(&E.w2, (size_t)9 * n)); ECHK(dalloc(&E.w3, (size_t)9 * n));
    ECHK(dalloc(&E.fmsum, (size_t)3 * n)); ECHK(dalloc(&E.fmx, (size_t)9 * n));
    ECHK(dalloc(&E.eout, nreal)); ECHK(dalloc(&E.wout, (size_t)9 * nreal)); ECHK(dalloc(&E.sout, (size_t)9 * nreal));
    ECHK(dalloc(&E.scal, 64)); ECHK(dalloc(&E.err, 4)); ECHK(dalloc(&S.rcell, 9));
    ECHK(dalloc(&E.pk, 64));
    if (!S.hpk) ECHK(hipHostMalloc((void**)&S.hpk, sizeof(double) * 64, hipHostMallocDefault));
    ECHK(hipMemset(E.zero_state, 0, 8 * sizeof(double)));
    ECHK(hipMemset(E.w2, 0, sizeof(double) * 9 * n));
    ECHK(hipMemset(E.fmsum, 0, sizeof(double) * 3 * n));
    ECHK(hipMemset(E.fmx, 0, sizeof(double) * 9 * n));
    S.nw = nw;
    return 0;
}

int alloc_rows(ElecState& S, int capA, int capP, std::string& err)
{
    ElecBuffers& E = S.E;
    dfree(E.nbrA);
    dfree(E.nbrP);
    dfree(E.tens);
    E.capA = capA;
    E.capP = capP;
    ECHK(dalloc(&E.nbrA, (size_t)E.n * capA));
    ECHK(dalloc(&E.nbrP, (size_t)E.n * capP));
    ECHK(dalloc(&E.tens, (size_t)6 * E.n * capP));
    return 0;
}

// erfcgenp.f (table of B0..B3 at r_i = i * drewd, i = 1..mxegrd), grid of parset.f:210-217
int build_table(ElecState& S, double alpha, double rcut, std::string& err)
{
    ElecBuffers& E = S.E;
    int mxegrd = (int)(rcut / 0.0001 + 0.5) + 4;
    if (mxegrd < 1000) mxegrd = 1000;
    const double drewd = rcut / (double)(mxegrd - 4);
    std::vector<double> t((size_t)4 * (mxegrd + 1), 0.0);
    for (int i = 1; i <= mxegrd && alpha >= 1.0e-6; ++i) {
        // (no Ewald split, alpha = 0: the -erf kernels vanish, as ewald3p.f sets them)
        const double rrr = (double)i * drewd;
        const double rsqi = 1.0 / rrr / rrr;
        const double a = alpha * rrr;
        double* e = &t[(size_t)4 * i];
        e[0] = (erfcc_host(a) - 1.0) / rrr;
        const double exp2a = exp(-a * a);
        for (int n = 1; n <= 3; ++n) {
            const double fn = (double)n;
            e[n] = rsqi * ((2.0 * fn - 1.0) * e[n - 1] + pow(2.0 * alpha * alpha, fn) * exp2a / alpha / SQRPI);
        }
    }
    dfree(E.tab);
    ECHK(dalloc(&E.tab, t.size()));
    ECHK(hipMemcpy(E.tab, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    E.ntab = mxegrd;
    S.P.drewd = drewd;
    S.tab_alpha = alpha;
    S.tab_rcut = rcut;
    return 0;
}

// k vectors of ewald1p.f:151-159, 248-345 in the reference's order (l = 0..kmax1, then
// m, then n, half space), grouped into runs of consecutive n at fixed (l, m).
int build_klist(ElecState& S, const double* cell, const int* kmax, double alpha, std::string& err)
{
    ElecBuffers& E = S.E;
    const double* a = cell;
    double rc[9];
    rc[0] = a[4] * a[8] - a[5] * a[7];
    rc[1] = a[2] * a[7] - a[1] * a[8];
    rc[2] = a[1] * a[5] - a[2] * a[4];
    rc[3] = a[5] * a[6] - a[3] * a[8];
    rc[4] = a[0] * a[8] - a[2] * a[6];
    rc[5] = a[2] * a[3] - a[0] * a[5];
    rc[6] = a[3] * a[7] - a[4] * a[6];
    rc[7] = a[1] * a[6] - a[0] * a[7];
    rc[8] = a[0] * a[4] - a[1] * a[3];
    const double det = a[0] * rc[0] + a[3] * rc[1] + a[6] * rc[2];
    if (fabs(det) < 1e-6) {
        err = "cell is singular";
        return -1;
    }
    for (int i = 0; i < 9; ++i) rc[i] /= det;
    const double volm = fabs(det);
    // dcell(rcell): widths of the reciprocal cell
    const double ax[3] = {rc[0], rc[1], rc[2]}, bx[3] = {rc[3], rc[4], rc[5]}, cx[3] = {rc[6], rc[7], rc[8]};
    const double bc[3] = {bx[1] * cx[2] - bx[2] * cx[1], bx[2] * cx[0] - bx[0] * cx[2], bx[0] * cx[1] - bx[1] * cx[0]};
    const double ca[3] = {cx[1] * ax[2] - cx[2] * ax[1], cx[2] * ax[0] - cx[0] * ax[2], cx[0] * ax[1] - cx[1] * ax[0]};
    const double ab[3] = {ax[1] * bx[2] - ax[2] * bx[1], ax[2] * bx[0] - ax[0] * bx[2], ax[0] * bx[1] - ax[1] * bx[0]};
    const double vol = fabs(ax[0] * bc[0] + ax[1] * bc[1] + ax[2] * bc[2]);
    const double w0 = vol / sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
    const double w1 = vol / sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]);
    const double w2 = vol / sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
    const double twopi = 2.0 * PI_D;
    double rcpcut = std::min(std::min(kmax[0] * w0, kmax[1] * w1), kmax[2] * w2);
    rcpcut = rcpcut * 1.05 * twopi;
    const double rcpct2 = rcpcut * rcpcut;
    const double ralph = -0.25 / (alpha * alpha);

    std::vector<double> kv;
    std::vector<int> runs, runk, klmn;
    if (kmax[0] + kmax[1] + kmax[2] + 3 > 80) {
        err = "kmax1 + kmax2 + kmax3 above 77 is not supported";
        return -1;
    }
    int K = 0;
    int mmin = 0, nmin = 1;
    for (int ll = 0; ll <= kmax[0]; ++ll) {
        const double tl = twopi * (double)ll;
        const double rkx1 = tl * rc[0], rky1 = tl * rc[3], rkz1 = tl * rc[6];
        for (int mm = mmin; mm <= kmax[1]; ++mm) {
            const double tm = twopi * (double)mm;
            const double rkx2 = rkx1 + tm * rc[1], rky2 = rky1 + tm * rc[4], rkz2 = rkz1 + tm * rc[7];
            int run_open = -1;
            for (int nn = nmin; nn <= kmax[2]; ++nn) {
                const double tn = twopi * (double)nn;
                const double k0 = rkx2 + tn * rc[2], k1 = rky2 + tn * rc[5], k2 = rkz2 + tn * rc[8];
                const double rksq = k0 * k0 + k1 * k1 + k2 * k2;
                if (!(rksq <= rcpct2)) {
                    run_open = -1;
                    continue;
                }
                const double rrksq = 1.0 / rksq;
                const double akk = exp(ralph * rksq) * rrksq;
                const double akv = 2.0 * akk * (rrksq - ralph);
                const double akv2 = 2.0 * (rrksq - ralph);
                const double rec[8] = {k0, k1, k2, akk, akv, akv2, 0.0, 0.0};
                kv.insert(kv.end(), rec, rec + 8);
                klmn.push_back((ll << 16) | ((mm + 128) << 8) | (nn + 128));
                if (run_open < 0 || runs[4 * run_open + 3] == 8) {
                    run_open = (int)runk.size();
                    const int r4[4] = {ll, mm, nn, 0};
                    runs.insert(runs.end(), r4, r4 + 4);
                    runk.push_back(K);
                }
                runs[4 * run_open + 3] += 1;
                ++K;
            }
            nmin = -kmax[2];
        }
        mmin = -kmax[1];
    }
    dfree(E.kv);
    dfree(E.klmn);
    dfree(E.runs);
    dfree(E.runk);
    dfree(E.S);
    dfree(E.Sq);
    dfree(E.kc);
    dfree(E.kf);
    dfree(E.Spart);
    E.K = K;
    E.nrun = (int)runk.size();
    ECHK(dalloc(&E.kv, kv.size()));
    ECHK(dalloc(&E.runs, runs.size()));
    ECHK(dalloc(&E.runk, runk.size()));
    ECHK(hipMemcpy(E.kv, kv.data(), sizeof(double) * kv.size(), hipMemcpyHostToDevice));
    ECHK(dalloc(&E.klmn, klmn.size()));
    ECHK(hipMemcpy(E.klmn, klmn.data(), sizeof(int) * klmn.size(), hipMemcpyHostToDevice));
    ECHK(hipMemcpy(E.runs, runs.data(), sizeof(int) * runs.size(), hipMemcpyHostToDevice));
    ECHK(hipMemcpy(E.runk, runk.data(), sizeof(int) * runk.size(), hipMemcpyHostToDevice));
    // site chunks of the structure-factor passes (their partials are re-read by the chunk
    // sum): about EL_DCHUNK_WAVES wavefronts, EL_SG per block of 64 half-runs, and at least
    // 64 sites per chunk
    const int rblocks = (EL_SPLIT * E.nrun + 63) / 64;
    const int nrow = 3 * S.nw;
    E.schunks = std::max(1, std::min(16, std::min((E.n + 63) / 64, EL_DCHUNK_WAVES / (EL_SG * rblocks))));
    E.dchunks = std::max(1, std::min(EL_DCHUNK_MAX, std::min((nrow + 63) / 64, EL_DCHUNK_WAVES / (EL_SG * rblocks))));
    // MBGPU_ELEC_WAVE_CHUNKS=1 (any nonzero integer) keeps the wavefront rule alone for the k
    // chunks of the final pass below (a test aid that compares chunkings); unset, empty or 0
    // leaves the bound on
    const char* wave_env = getenv("MBGPU_ELEC_WAVE_CHUNKS");
    const bool wave_rules_only = wave_env && atoi(wave_env) != 0;
    E.spc = (E.n + E.schunks - 1) / E.schunks;
    E.dspc = (nrow + E.dchunks - 1) / E.dchunks;
    ECHK(dalloc(&E.S, (size_t)8 * K));
    dfree(E.cg_smu);
    dfree(E.cg_sp);
    dfree(E.cg_su);
    dfree(E.cg_sq);
    ECHK(dalloc(&E.cg_smu, (size_t)6 * K));
    ECHK(dalloc(&E.cg_sp, (size_t)6 * K));
    ECHK(dalloc(&E.cg_su, (size_t)6 * K));
    ECHK(dalloc(&E.cg_sq, (size_t)6 * K));
    ECHK(dalloc(&E.Sq, (size_t)2 * K));
    ECHK(dalloc(&E.kc, (size_t)16 * K));
    ECHK(dalloc(&E.kf, (size_t)16 * K));
    ECHK(dalloc(&E.Spart, (size_t)std::max(8 * E.schunks, 6 * E.dchunks) * K));
    ECHK(hipMemset(E.S, 0, sizeof(double) * 8 * K));
    ECHK(hipMemset(E.kc, 0, sizeof(double) * 16 * K));
    // phase tables
    const int ks = std::max(std::max(kmax[0], kmax[1]), kmax[2]) + 1;
    if (ks != E.kstride) {
        dfree(E.phc);
        dfree(E.phs);
        E.kstride = ks;
        ECHK(dalloc(&E.phc, (size_t)3 * ks * E.n));
        ECHK(dalloc(&E.phs, (size_t)3 * ks * E.n));
    }
    // k split of the site-parallel passes: about 2048 wavefronts over the sites
    E.kchunks = std::max(1, std::min(E.nrun, EL_KCHUNK_WAVES / std::max(1, (E.n + 63) / 64)));
    // and at most EL_KCHUNK_RUNS runs per chunk: in large boxes the rule above leaves a few
    // hundred workgroups with long k loops each
    if (!wave_rules_only)
        E.kchunks = std::max(E.kchunks, std::min(E.nrun, (E.nrun + EL_KCHUNK_RUNS - 1) / EL_KCHUNK_RUNS));
    E.rpc = (E.nrun + E.kchunks - 1) / E.kchunks;
    E.kchunks = (E.nrun + E.rpc - 1) / E.rpc;
    E.kmaxch = 0;
    for (int c = 0; c < E.kchunks; ++c) {
        const int r0 = c * E.rpc, r1 = std::min(E.nrun, r0 + E.rpc);
        const int k0 = runk[r0], k1 = r1 < E.nrun ? runk[r1] : K;
        E.kmaxch = std::max(E.kmaxch, k1 - k0);
    }
    // k chunks of the field passes: about EL_FCHUNK_WAVES wavefronts (EL_KG per block of
    // 128 sites), at most EL_FCHUNK_RUNS runs per chunk (their constants are staged in LDS)
    {
        const int sblocks = (nrow + 127) / 128;
        E.fchunks = std::max(1, std::min(E.nrun, EL_FCHUNK_WAVES / (EL_KG * sblocks)));
        E.fchunks = std::max(E.fchunks, std::min(E.nrun, (E.nrun + EL_FCHUNK_RUNS - 1) / EL_FCHUNK_RUNS));
        E.frpc = (E.nrun + E.fchunks - 1) / E.fchunks;
        E.fchunks = (E.nrun + E.frpc - 1) / E.frpc;
        E.fkmax = 0;
        for (int c = 0; c < E.fchunks; ++c) {
            const int r0 = c * E.frpc, r1 = std::min(E.nrun, r0 + E.frpc);
            const int k0 = runk[r0], k1 = r1 < E.nrun ? runk[r1] : K;
            E.fkmax = std::max(E.fkmax, k1 - k0);
        }
    }
    dfree(E.Fpart);
    ECHK(dalloc(&E.Fpart, std::max((size_t)E.kchunks * 16, (size_t)E.fchunks * 6) * E.n));
    // block partials of the final passes
    // (the particle-mesh final pass writes a partial per 4 sites where the k sum writes one
    // per 256 k)
    const int nsb = (E.n + 255) / 256, nkb = std::max((K + 255) / 256, (E.n + 3) / 4), nwb = (S.nw + 63) / 64;
    dfree(E.blk);
    E.nblk = nsb + 8 * nkb + 2 * nsb + 8 * E.n + 24 * nwb + 2 * nsb + 64;
    ECHK(dalloc(&E.blk, (size_t)E.nblk));
    ECHK(hipMemcpy(S.rcell, rc, sizeof(double) * 9, hipMemcpyHostToDevice));
    S.P.rvolm = twopi / volm;
    for (int i = 0; i < 9; ++i) S.k_cell[i] = cell[i];
    for (int i = 0; i < 3; ++i) S.k_kmax[i] = kmax[i];
    S.k_alpha = alpha;
    return 0;
}

}  // namespace

namespace {

// An isolated cluster (imcon 0): no reciprocal space.  The buffers the k-sum set-up would
// size are sized for the site count alone: the field partials (one chunk of 6 components)
// and the block partials of the final passes.
int alloc_direct(ElecState& S, std::string& err)
{
    ElecBuffers& E = S.E;
    dfree(E.Fpart);
    ECHK(dalloc(&E.Fpart, (size_t)6 * E.n));
    const int nsb = (E.n + 255) / 256, nkb = (E.n + 3) / 4, nwb = (S.nw + 63) / 64;
    dfree(E.blk);
    E.nblk = nsb + 8 * nkb + 2 * nsb + 8 * E.n + 24 * nwb + 2 * nsb + 64;
    ECHK(dalloc(&E.blk, (size_t)E.nblk));
    E.K = E.nrun = 0;
    E.kchunks = E.fchunks = 1;
    E.rpc = E.kmaxch = E.frpc = E.fkmax = 0;
    S.k_alpha = -1.0;   // a later periodic call builds its k list
    S.P.rvolm = 0.0;
    return 0;
}

}  // namespace

void elec_release(ElecState& S)
{
    elec_drop_graph(S);
    free_buffers(S.E);
    dfree(S.rcell);
    if (S.hpk) (void)hipHostFree(S.hpk);
    S.hpk = nullptr;
    if (S.sp) {
        spme_free(*S.sp);
        delete S.sp;
        S.sp = nullptr;
    }
    delete S.spg;
    S.spg = nullptr;
    cells_free(S.cells);
    S.ready = false;
}

// Prepare buffers, tables and the k list for a call; positions and dipoles are uploaded by
// the caller into S.E.x/y/z (real atoms) and S.E.dip afterwards.
int elec_configure(ElecState& S, int nw, const CellGeom& g, const double* cell, const mbgpu_elec_params* p,
                   std::string& err)
{
    if (!p) {
        err = "null parameters";
        return -1;
    }
    if (g.imcon == 0) {
        // an isolated cluster: direct sums over all site pairs within rcut, no Ewald split
        if (!(p->rcut > 0.0) || p->alpha != 0.0) {
            err = "electrostatics parameters: an isolated cluster (imcon 0) needs rcut > 0 and alpha = 0 "
                  "(direct sums; rcut at least the cluster's extent)";
            return -1;
        }
        if (S.recip) {
            err = "electrostatics parameters: the particle mesh needs a periodic cell (imcon 1-3)";
            return -1;
        }
    } else if (!(p->rcut > 0.0) || !(p->alpha > 1e-6) || p->kmax[0] < 1 || p->kmax[1] < 1 || p->kmax[2] < 1) {
        err = "electrostatics parameters: need rcut > 0, alpha > 1e-6 and kmax >= 1";
        return -1;
    }
    if (p->polr2[2] > 1e-6) {
        err = "electrostatics parameters: the M site must be unpolarizable (polr2[2] = 0)";
        return -1;
    }
    if (!(p->polr[0] > 0.0 && p->polr[1] > 0.0 && p->polr[2] > 0.0)) {
        err = "electrostatics parameters: smearing polarizabilities must be positive";
        return -1;
    }
    if (S.nw != nw) {
        elec_release(S);
        if (alloc_sites(S, nw, err)) return -1;
    }
    ElecDev& P = S.P;
    P.nw = nw;
    P.n = 4 * nw;
    P.nreal = 3 * nw;
    P.nrow = 3 * nw;
    P.g = g;
    P.rcut = p->rcut;
    P.alpha = p->alpha;
    P.gamma = p->gamma;
    P.q_o = p->q_o;
    for (int i = 0; i < 3; ++i) {
        P.polr[i] = p->polr[i];
        P.polr2[i] = p->polr2[i];
        P.km[i] = p->kmax[i];
    }
    P.athole = p->athole;
    P.athole12 = p->athole12;
    P.athole13 = p->athole13;
    P.ascc = p->ascc;
    P.ascd = p->ascd;
    P.toler = p->toler;
    P.omega = p->omega;
    P.compat_qd = p->compat_qd_column;
    if (S.tab_alpha != p->alpha || S.tab_rcut != p->rcut)
        if (build_table(S, p->alpha, p->rcut, err)) return -1;
    if (g.imcon == 0) {
        if (alloc_direct(S, err)) return -1;
    } else {
        bool same_k = S.k_alpha == p->alpha && S.E.K > 0;
        for (int i = 0; i < 9; ++i) same_k = same_k && S.k_cell[i] == cell[i];
        for (int i = 0; i < 3; ++i) same_k = same_k && S.k_kmax[i] == p->kmax[i];
        if (!same_k && build_klist(S, cell, p->kmax, p->alpha, err)) return -1;
    }
    S.cells.V.on = 0;   // (an isolated cluster has no cells to wrap)
    if (g.imcon != 0 && cells_plan(S.cells, cell, p->rcut, 4 * nw, CELLS_MIN_SITES, err)) return -1;
    S.P.cells = S.cells.V;
    if (S.recip) {
        // particle mesh: eleven grids (the final pass's), splines of every site
        if (!S.sp) S.sp = new SpmeBuffers();
        if (!S.spg) S.spg = new SpmeGeom();
        if (spme_plan(*S.sp, *S.spg, cell, S.sp_order, S.sp_grid, p->alpha, 4 * nw, SP_MAX_GRIDS, err)) {
            err = "particle mesh: " + err;
            return -1;
        }
    }
    if (S.E.capA == 0) {
        // rows: sites within rcut at a density of up to 0.15 sites per A^3, rounded to 64
        const double vol = 4.0 / 3.0 * PI_D * p->rcut * p->rcut * p->rcut;
        int cap = (int)(0.15 * vol) + 64;
        // MBGPU_ELEC_ROWS=c starts from rows of c entries instead (a test aid: it makes the
        // row-growth path run on ordinary configurations)
        if (const char* ev = getenv("MBGPU_ELEC_ROWS")) cap = std::max(1, atoi(ev));
        cap = std::min(((cap + 63) / 64) * 64, ((P.n + 63) / 64) * 64);
        if (alloc_rows(S, cap, cap, err)) return -1;
    }
    P.capA = S.E.capA;
    P.capP = S.E.capP;
    P.K = S.E.K;
    P.nrun = S.E.nrun;
    P.kchunks = S.E.kchunks;
    P.kstride = S.E.kstride;
    P.rpc = S.E.rpc;
    P.kmaxch = S.E.kmaxch;
    P.fchunks = S.E.fchunks;
    P.frpc = S.E.frpc;
    P.fkmax = S.E.fkmax;
    P.scale = 4.0 * P.rvolm * R4PIE0;
    P.scal1 = 2.0 * P.rvolm * R4PIE0;
    P.self_dd = 4.0 * R4PIE0 * P.alpha * P.alpha * P.alpha / SQRPI / 3.0;
    return 0;
}

namespace {

// Enqueue one attempt of an evaluation up to `nbatch` solver iterations: the start dipoles
// kept in dip0 (first attempt) or restored from it (a repeat after the rows grew), the
// rows, the preparation and the solver's start.  Kernels only: a graph may capture it.
int enqueue_attempt(ElecState& S, const ElecSpme* sp, int solver, bool first, int nbatch, hipStream_t s)
{
    ElecBuffers& E = S.E;
    const ElecDev& P = S.P;
    {
        // the start dipoles kept (or restored), and the four int flags cleared
        double* d[2] = {first ? E.dip0 : E.dip, reinterpret_cast<double*>(E.err)};
        const double* sr[2] = {first ? E.dip : E.dip0, nullptr};
        const long len[2] = {3L * P.n, 2};
        el_launch_copy_segs(2, d, sr, len, s);
    }
    int rc = el_launch_setup(E, P, S.cells, s);
    if (rc) return rc;
    {
        // solver state and T mu cleared; M sites carry no induced dipole (polr2 = 0): any
        // start value cleared
        double* d[3] = {E.state, E.emu, E.dip + 3 * P.nreal};
        const double* sr[3] = {nullptr, nullptr, nullptr};
        const long len[3] = {16, 3L * P.n, 3L * (P.n - P.nreal)};
        el_launch_copy_segs(3, d, sr, len, s);
    }
    rc = el_launch_prep(E, P, S.rcell, sp, s);
    if (!rc && solver == 1) rc = el_launch_cg_init(E, P, sp, s);
    for (int b = 0; !rc && b < nbatch; ++b)
        rc = solver == 1 ? el_launch_cg_iteration(E, P, b, sp, s) : el_launch_iteration(E, P, b == 0, sp, s);
    return rc;
}

// The key of a captured evaluation: the bytes of everything its launches were recorded
// with (kernel-argument blocks, buffer addresses and sizes, the mesh and link-cell plans)
// and the run's own choices.
std::vector<unsigned char> graph_key(const ElecState& S, int distribute, int solver, int nbatch)
{
    std::vector<unsigned char> k;
    auto add = [&](const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        k.insert(k.end(), c, c + n);
    };
    add(&S.P, sizeof S.P);
    add(&S.E, sizeof S.E);
    add(&S.cells, sizeof S.cells);
    add(&S.rcell, sizeof S.rcell);
    const int v[5] = {distribute, solver, nbatch, S.recip, S.P.g.imcon};
    add(v, sizeof v);
    if (S.recip && S.sp && S.spg) {
        add(S.sp, sizeof(SpmeBuffers));
        add(S.spg, sizeof(SpmeGeom));
    }
    return k;
}

void drop_graph(ElecState& S)
{
    if (S.gexec) (void)hipGraphExecDestroy(S.gexec);
    S.gexec = nullptr;
    S.gkey.clear();
}

}  // namespace

void elec_drop_graph(ElecState& S) { drop_graph(S); }

namespace {
// the reciprocal space of an evaluation: the k-sum (null), the particle mesh, or none (an
// isolated cluster: a description with no mesh buffers)
const ElecSpme* recip_of(const ElecState& S, ElecSpme& spx, ElecSpme& none)
{
    spx = ElecSpme{S.sp, S.spg, S.spg ? S.spg->vol / (8.0 * PI_D) : 0.0};
    none = ElecSpme{nullptr, nullptr, 0.0};
    return S.P.g.imcon == 0 ? &none : S.recip ? &spx : nullptr;
}
}  // namespace

// Solver iterations a captured evaluation holds (0 before the first evaluation set a count):
// the last count + 2, kept while it covers the solve and is not far too long (iterations past
// convergence are no-ops of about a microsecond per kernel; a new capture costs more), so
// counts that drift by one or two do not capture again.  MBGPU_GRAPH_BATCH=k caps it (a test
// aid for the continuation after a graph).
int elec_batch(const ElecState& S)
{
    int gb = 0;
    if (S.last_iters > 0) {
        gb = std::min(S.last_iters + 2, 64);
        if (S.gbatch >= S.last_iters && S.gbatch <= S.last_iters + 8) gb = S.gbatch;
    }
    if (const char* ev = getenv("MBGPU_GRAPH_BATCH")) gb = gb ? std::min(gb, std::max(1, atoi(ev))) : 0;
    return gb;
}

// One evaluation's device work with gb solver iterations and the final pass (what a graph
// holds), kernels only.
int elec_enqueue_full(ElecState& S, int distribute, int solver, int gb, hipStream_t s)
{
    ElecSpme spx, none;
    const ElecSpme* sp = recip_of(S, spx, none);
    const double g34 = exp(gammln_host(0.75)), gln34 = gammln_host(0.75);
    int rc = enqueue_attempt(S, sp, solver, true, gb, s);
    if (!rc) rc = el_launch_final(S.E, S.P, distribute, g34, gln34, solver == 1, sp, s);
    return rc;
}

void elec_key_append(const ElecState& S, int distribute, int solver, int gb, std::vector<unsigned char>& k)
{
    const std::vector<unsigned char> e = graph_key(S, distribute, solver, gb);
    k.insert(k.end(), e.begin(), e.end());
}

// The host side of an evaluation's end: the particle mesh's flags, then the scalars of
// S.out from the reduced sums in S.scal (read back by the caller) and the solver state st.
int elec_finish(ElecState& S, const double st[8], const int spflag[2], std::string& err)
{
    const ElecDev& P = S.P;
    const ElecBuffers& E = S.E;
    if (spflag[0] == 1) {
        err = "particle mesh: a site position is not finite or too far from the cell to index the grid";
        return -1;
    }
    if (spflag[0] == 2) {
        err = "particle mesh: a charge or dipole is too large for the fixed-point grid";
        return -1;
    }
    const double* q = S.scal;
    const double engsic = -R4PIE0 * P.alpha * q[9] / SQRPI;
    const double eng1 = P.scal1 * q[1] + engsic;
    const double engdum = q[2];
    const double omg_xx = q[3], omg_yy = q[4], omg_zz = q[5], omg_xy = q[6], omg_xz = q[7], omg_yz = q[8];
    double st9[9];
    st9[0] = P.scal1 * (omg_xx + engdum);
    st9[1] = P.scal1 * omg_xy;
    st9[2] = P.scal1 * omg_xz;
    st9[3] = st9[1];
    st9[4] = P.scal1 * (omg_yy + engdum);
    st9[5] = P.scal1 * omg_yz;
    st9[6] = st9[2];
    st9[7] = st9[5];
    st9[8] = P.scal1 * (omg_zz + engdum);
    const double vir1 = -P.scal1 * (omg_xx + omg_yy + omg_zz + 3.0 * engdum);
    // pair, excluded-pair and qd sums: [energy, virial, xx, xy, xz, yy, yz, zz]
    const double* pp[3] = {q + 11, q + 19, q + 27};
    double vir = vir1;
    for (int t = 0; t < 3; ++t) {
        const double* r = pp[t];
        vir += r[1];
        st9[0] += r[2];
        st9[1] += r[3];
        st9[2] += r[4];
        st9[3] += r[3];
        st9[4] += r[5];
        st9[5] += r[6];
        st9[6] += r[4];
        st9[7] += r[6];
        st9[8] += r[7];
    }
    double* o = S.out;
    o[0] = eng1;
    o[1] = q[11];
    o[2] = q[19];
    o[3] = q[0];
    o[4] = q[0] + eng1 + q[11] + q[19];
    o[5] = vir;
    o[6] = vir1;
    o[7] = q[28];
    o[8] = st[0];
    o[9] = st[1];
    for (int i = 0; i < 9; ++i) o[10 + i] = st9[i];
    o[19] = (double)E.K;
    S.ready = true;
    return 0;
}

// One evaluation with positions (real atoms) in E.x/y/z and start dipoles in E.dip.
//
// Once an evaluation has set the iteration count, the next ones run as one captured graph:
// the start, last count + 2 solver iterations (a converged state turns the rest into no-ops)
// and the final pass, then one read-back of the solver state, the row-overflow flag and the
// scalars.  A graph that ends unconverged continues with iterations launched one by one and
// runs the final pass again (it rewrites every output); a row overflow falls back to the
// repeat below.  The graph is captured again whenever its key changes.
int elec_evaluate(ElecState& S, int max_iter, int distribute, int solver, hipStream_t s, std::string& err)
{
    ElecBuffers& E = S.E;
    ElecDev& P = S.P;
    S.ready = false;
    // the reciprocal space: the k-sum (sp null), the particle mesh, or none (an isolated
    // cluster: sp with no mesh buffers)
    ElecSpme spx{S.sp, S.spg, S.spg ? S.spg->vol / (8.0 * PI_D) : 0.0};
    ElecSpme none{nullptr, nullptr, 0.0};
    const ElecSpme* sp = P.g.imcon == 0 ? &none : S.recip ? &spx : nullptr;
    const double g34 = exp(gammln_host(0.75)), gln34 = gammln_host(0.75);
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int herr = 0;
    bool need_final = true;
    // every read-back is one copy of the packed scalars into pinned memory (the solver state,
    // the row-overflow flag, the reduced scalars, the mesh's input flags); `pack` launches
    // the pack kernel first (a graph holds its own)
    const int* spf = sp && sp->B ? S.sp->flag : nullptr;
    auto readback = [&](bool pack) -> hipError_t {
        if (pack) el_launch_pack(E.state, E.err, E.scal, spf, E.pk, s);
        hipError_t e = hipMemcpyAsync(S.hpk, E.pk, sizeof(double) * EL_PK_N, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        for (int i = 0; i < 8; ++i) st[i] = S.hpk[EL_PK_ST + i];
        herr = (int)S.hpk[EL_PK_ERR];
        return hipSuccess;
    };
    const int gb = elec_batch(S);
    bool from_graph = false;
    // MBGPU_GRAPHS=0 launches kernel by kernel whatever the context's setting (a test aid)
    const char* gv = getenv("MBGPU_GRAPHS");
    const bool graphs = S.graphs && !(gv && atoi(gv) == 0);
    if (graphs && gb > 0) {
        std::vector<unsigned char> key = graph_key(S, distribute, solver, gb);
        if (!S.gexec || key != S.gkey) {
            drop_graph(S);
            hipGraph_t graph = nullptr;
            // no other entry runs while this thread captures (capture_gate_begin)
            capture_gate_begin();
            hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            int rc = 0;
            if (e == hipSuccess) {
                rc = elec_enqueue_full(S, distribute, solver, gb, s);
                if (!rc) el_launch_pack(E.state, E.err, E.scal, spf, E.pk, s);
                e = hipStreamEndCapture(s, &graph);
            }
            capture_gate_end();
            if (e == hipSuccess && !rc && graph) e = hipGraphInstantiate(&S.gexec, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (e != hipSuccess || rc || !S.gexec) {
                // no graph for this state: launch kernel by kernel from now on
                (void)hipGetLastError();
                drop_graph(S);
                S.graphs = 0;
            } else {
                S.gkey.swap(key);
                S.gbatch = gb;
                ++S.gcaptures;
            }
        }
        if (S.gexec) {
            ECHK(hipGraphLaunch(S.gexec, s));
            ++S.greplays;
            ECHK(readback(false));
            from_graph = !herr;
            need_final = !(st[2] != 0.0);
        }
    }
    // The neighbour rows' capacity check is read back with the solver's first convergence
    // check instead of after the row build (one host round trip less): the rows are built
    // with clamped counts, so an overflowed attempt runs in bounds and is simply repeated
    // with larger rows, from the same start dipoles (kept in dip0).
    for (int attempt = from_graph ? -1 : (S.gexec && herr ? 1 : 0); attempt >= 0 || from_graph;) {
        int launched = 0, batch = S.last_iters > 0 ? std::min(S.last_iters + 2, 64) : (solver == 1 ? 8 : 12);
        int rc = 0;
        if (from_graph) {
            // continue after the graph's iterations
            launched = S.gbatch;
            batch = solver == 1 ? 4 : 8;
            from_graph = false;
            attempt = -1;
            if (st[2] != 0.0) break;
        } else {
            if (attempt > 0) {
                // the rows overflowed: grow them and repeat from the start dipoles
                const int full = ((P.n + 63) / 64) * 64;
                if (attempt > 2 || E.capA >= full) {
                    err = "electrostatics neighbour rows exceed their capacity";
                    return -2;
                }
                if (alloc_rows(S, std::min(full, 2 * E.capA), std::min(full, 2 * E.capP), err)) return -1;
                P.capA = E.capA;
                P.capP = E.capP;
            }
            rc = enqueue_attempt(S, sp, solver, attempt == 0, batch, s);
            if (rc) {
                err = std::string("electrostatics kernels failed: ") + hipGetErrorString((hipError_t)rc);
                return rc;
            }
            launched = batch;
            batch = solver == 1 ? 4 : 8;
        }
        bool overflow = false;
        for (;;) {
            ECHK(readback(true));
            if (herr) {
                overflow = true;
                break;
            }
            if (st[2] != 0.0) break;
            if (!(st[1] == st[1])) {
                err = "induced dipoles became NaN";
                return -3;
            }
            if (st[0] >= (double)max_iter) {
                char buf[160];
                snprintf(buf, sizeof buf, "induced dipoles did not converge in %d iterations (rms change %.3e)",
                         (int)st[0], st[1]);
                err = buf;
                return -3;
            }
            for (int b = 0; b < batch; ++b) {
                rc = solver == 1 ? el_launch_cg_iteration(E, P, launched, sp, s)
                                 : el_launch_iteration(E, P, launched == 0, sp, s);
                ++launched;
                if (rc) {
                    err = std::string("electrostatics iteration kernels failed: ") +
                          hipGetErrorString((hipError_t)rc);
                    return rc;
                }
            }
        }
        need_final = true;
        if (!overflow) break;
        attempt = attempt < 0 ? 1 : attempt + 1;
    }
    S.last_iters = (int)st[0];
    if (need_final) {
        const int rc = el_launch_final(E, P, distribute, g34, gln34, solver == 1, sp, s);
        if (rc) {
            err = std::string("electrostatics final kernels failed: ") + hipGetErrorString((hipError_t)rc);
            return rc;
        }
    }
    if (need_final) {
        // the scalars of the final pass (a graph that converged read them with its state)
        const double keep[8] = {st[0], st[1], st[2], st[3], st[4], st[5], st[6], st[7]};
        ECHK(readback(true));
        for (int i = 0; i < 8; ++i) st[i] = keep[i];
    }
    for (int i = 0; i < 41; ++i) S.scal[i] = S.hpk[EL_PK_SCAL + i];
    const int spflag[2] = {(int)S.hpk[EL_PK_SPF], (int)S.hpk[EL_PK_SPF + 1]};
    return elec_finish(S, st, spflag, err);
}
