// mbgpu_api.hip -- C ABI of libmbgpu.so: contexts, device buffers and the host-side
// orchestration of the MB-pol kernels.  Declarations and reference citations live in
// include/mbgpu.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mbgpu.h"
#include "mbgpu_internal.h"

int mb_launch_lists(MbBuffers& B, const CellGeom& g, hipStream_t s);
int mb_launch_dimers(MbBuffers& B, const int* ncnt, int ncap, const CellGeom& g, bool gather, hipStream_t s);
int mb_launch_trimers(MbBuffers& B, const int* ncnt, int ncap, int blk_off, const CellGeom& g, bool gather,
                      hipStream_t s);
int mb_launch_accumulate(MbBuffers& B, double* fx, double* fy, double* fz, hipStream_t s);
int md_launch_stage(int stage, int nr, double dt, const double* mass, double* x, double* y, double* z, double* vx,
                    double* vy, double* vz, double* fx, double* fy, double* fz, double* fsite, double* ekin,
                    double* blk, double* out, hipStream_t s);
int md_launch_hcacf(long nt, int ntau, const double* j, double* acf, hipStream_t s);
int ring_launch_transform(int P, long ncoord, const double* M, const double* in, double* out, hipStream_t s);
int ring_launch_spring(int P, int natom, double coeff, double fac3, const double* mass, const double* x,
                       const double* xc, double* f, double* blk, double* out, hipStream_t s);
int mb_launch_heat_flux(int nreal, const double* vx, const double* vy, const double* vz, const double* ekin,
                        const double* epot, const double* epot2, const double* w, const double* w2,
                        const double* sew, double* blk, int nblk_cap, double* out, hipStream_t s);

static std::atomic<long> g_alloc_serial{0};

long mbgpu_alloc_serial() { return g_alloc_serial.fetch_add(1); }

// Gate between the library's entries and a graph capture.  A capture must not overlap
// another thread's null-stream work (a synchronous copy or set during a set-up, a free):
// HIP rejects that work while any stream captures.  Every entry holds the gate shared while
// it runs (the entries wait for their own streams before returning, so no library work is
// in flight outside an entry); a capture holds it alone.  New entries wait while a capture
// is pending, so a stream of entries cannot starve it.  An entry called from inside another
// on the same thread does not take the gate again.
namespace {
class EntryGate {
    std::mutex m_;
    std::condition_variable cv_;
    int readers_ = 0, waiting_ = 0;
    bool writer_ = false;

public:
    void lock_shared()
    {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return !writer_ && waiting_ == 0; });
        ++readers_;
    }
    void unlock_shared()
    {
        std::lock_guard<std::mutex> l(m_);
        if (--readers_ == 0) cv_.notify_all();
    }
    void lock()
    {
        std::unique_lock<std::mutex> l(m_);
        ++waiting_;
        cv_.wait(l, [&] { return !writer_ && readers_ == 0; });
        --waiting_;
        writer_ = true;
    }
    void unlock()
    {
        std::lock_guard<std::mutex> l(m_);
        writer_ = false;
        cv_.notify_all();
    }
};
EntryGate g_gate;
thread_local int g_gate_depth = 0;

struct EntryGuard {
    EntryGuard()
    {
        if (g_gate_depth++ == 0) g_gate.lock_shared();
    }
    ~EntryGuard()
    {
        if (--g_gate_depth == 0) g_gate.unlock_shared();
    }
};
}  // namespace

void capture_gate_begin()
{
    if (g_gate_depth > 0) g_gate.unlock_shared();
    g_gate.lock();
}

void capture_gate_end()
{
    g_gate.unlock();
    if (g_gate_depth > 0) g_gate.lock_shared();
}



int mbgpu_poison_this(long serial)
{
    const char* v = getenv("MBGPU_POISON");
    if (!v || !v[0] || v[0] == '0') return 0;
    long a = 0, b = 0;
    if (v[0] == '!' && sscanf(v + 1, "%ld:%ld", &a, &b) == 2) return (serial >= a && serial < b) ? 2 : 1;
    if (sscanf(v, "%ld:%ld", &a, &b) == 2) return (serial >= a && serial < b) ? 1 : 0;
    return 1;
}

extern "C" long mbgpu_debug_alloc_serial(void) { return g_alloc_serial.load(); }

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code ? code : -1;
}

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) return fail((int)e_, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

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

// Packed scalars of one complete evaluation (mbgpu_evaluate_dev), in doubles: MB-pol's
// dimer/trimer counts and flag (4 ints), MB-pol's energy/virial/stress sums (16), the solver
// state (8), the electrostatics' error flag, their scalar sums (41), the particle mesh's input
// flags (2 ints), the water terms' sums (18), the heat-flux sums (8).  Integers are exact.
constexpr int PK_CNT = 0, PK_MB = 4, PK_ST = 20, PK_ERR = 28, PK_EL = 29, PK_SPF = 70, PK_WT = 72, PK_J = 90,
              PK_N = 98;

__global__ void k_pack_eval(const int* __restrict__ cnt, const double* __restrict__ mb,
                            const double* __restrict__ st, const int* __restrict__ err,
                            const double* __restrict__ els, const int* __restrict__ spf,
                            const double* __restrict__ ws, const double* __restrict__ jd, double* __restrict__ out)
{
    const int t = threadIdx.x;
    if (t >= PK_N) return;
    double v;
    if (t < PK_MB) v = (double)cnt[t];
    else if (t < PK_ST) v = mb[t - PK_MB];
    else if (t < PK_ERR) v = st[t - PK_ST];
    else if (t < PK_EL) v = (double)err[0];
    else if (t < PK_SPF) v = els[t - PK_EL];
    else if (t < PK_WT) v = spf ? (double)spf[t - PK_SPF] : 0.0;
    else if (t < PK_J) v = ws[t - PK_WT];
    else v = jd[t - PK_J];
    out[t] = v;
}

// Total force on real atom i, component q: (MB-pol + electrostatics) + water terms, in that
// order (the order of the host sum over the three parts' details).
__global__ void k_sum_forces(int nr, const double* __restrict__ fmb, const double* __restrict__ fel,
                             const double* __restrict__ fwt, double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nr) return;
    const int i = t / 3, q = t - 3 * i;
    out[t] = (fmb[t] + fel[t]) + fwt[(long)q * nr + i];
}

bool make_geom(int imcon, const double* cell, CellGeom& g)
{
    if (imcon < 0 || imcon > 3) return false;
    g.imcon = imcon;
    if (imcon == 0) {
        // an isolated cluster: no images; the cell (if any) is kept but not used
        for (int i = 0; i < 9; ++i) g.c[i] = cell ? cell[i] : 0.0;
        for (int i = 0; i < 3; ++i) g.r[i] = 0.0;
        for (int i = 0; i < 9; ++i) g.rc[i] = 0.0;
        return true;
    }
    for (int i = 0; i < 9; ++i) g.c[i] = cell[i];
    g.r[0] = 1.0 / cell[0];
    g.r[1] = 1.0 / cell[4];
    g.r[2] = 1.0 / cell[8];
    const double* a = cell;
    double* b = g.rc;
    b[0] = a[4] * a[8] - a[5] * a[7];
    b[1] = a[2] * a[7] - a[1] * a[8];
    b[2] = a[1] * a[5] - a[2] * a[4];
    b[3] = a[5] * a[6] - a[3] * a[8];
    b[4] = a[0] * a[8] - a[2] * a[6];
    b[5] = a[2] * a[3] - a[0] * a[5];
    b[6] = a[3] * a[7] - a[4] * a[6];
    b[7] = a[1] * a[6] - a[0] * a[7];
    b[8] = a[0] * a[4] - a[1] * a[3];
    const double det = a[0] * b[0] + a[3] * b[1] + a[6] * b[2];
    const double r = det != 0.0 ? 1.0 / det : 0.0;
    for (int i = 0; i < 9; ++i) b[i] *= r;
    return true;
}

}  // namespace

struct mbgpu_ctx {
    MbBuffers B;
    hipStream_t stream = nullptr;
    // the electrostatics run on a stream of their own, so that a caller may evaluate them
    // from one host thread while the MB-pol and water terms run from another (the parts
    // share no buffers; the heat flux, called after both, reads their per-atom results)
    hipStream_t el_stream = nullptr;
    int device = 0;
    // host-side staging for the host-pointer entry points
    double *hx = nullptr, *hy = nullptr, *hz = nullptr;   // device copies of positions
    int* hlist = nullptr;
    long nd = 0, nt = 0;
    CellGeom geom{};
    bool have_inputs = false;
    bool mb_ready = false;            // the last MB-pol evaluation completed
    // heat-flux scratch
    double *dv = nullptr, *dek = nullptr, *dep = nullptr, *dw = nullptr, *dsew = nullptr, *dout = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    ElecState el;                     // polarizable electrostatics
    WaterState wt;                    // monomer, dispersion and long-range terms
    // ring polymer: bead-index matrix of the last transform, spring block partials
    double *ring_m = nullptr, *ring_blk = nullptr, *ring_out = nullptr;
    long ring_m_cap = 0, ring_blk_cap = 0;
    SpmeBuffers sp;                   // particle-mesh Ewald reciprocal space
    // a complete evaluation (mbgpu_evaluate_dev) captured as three graphs (the electrostatics
    // on their stream; MB-pol and the water terms; the per-atom sums, heat flux and packed
    // scalars after the join), replayed while their key is unchanged.  One graph with both
    // branches runs them one after the other on this runtime; two launches on two streams
    // overlap.
    hipGraphExec_t fexec[3] = {nullptr, nullptr, nullptr};
    std::vector<unsigned char> fkey;
    long fcaptures = 0, freplays = 0;
    // every scalar of a complete evaluation, packed on the device and read back in one copy
    // into pinned host memory
    double* dpk = nullptr;
    double* hpk = nullptr;
    double* dfsum = nullptr;          // total forces of the last evaluation (mbgpu_get_forces)
};

namespace {

int alloc_fixed(mbgpu_ctx* c)
{
    MbBuffers& B = c->B;
    const int nw = B.nw, na = B.natms;
    B.cap2 = 128;
    B.cap3 = 64;
    HIPCHK(dalloc(&B.mon, (size_t)9 * nw));
    HIPCHK(dalloc(&B.wbase, nw));
    HIPCHK(dalloc(&B.nbr2, (size_t)nw * B.cap2));
    HIPCHK(dalloc(&B.cnt2, nw));
    HIPCHK(dalloc(&B.off2, nw));
    HIPCHK(dalloc(&B.nbr3, (size_t)nw * B.cap3));
    HIPCHK(dalloc(&B.nbr2l, (size_t)nw * B.cap2));
    HIPCHK(dalloc(&B.cnt2l, nw));
    HIPCHK(dalloc(&B.kmask, (size_t)nw * 64));
    HIPCHK(dalloc(&B.rowoff, (size_t)nw * 64));
    HIPCHK(dalloc(&B.cnt3, nw));
    HIPCHK(dalloc(&B.tcnt, nw));
    HIPCHK(dalloc(&B.toff, nw));
    HIPCHK(dalloc(&B.counts, 4));
    HIPCHK(dalloc(&B.is_water, (size_t)4 * ((na + 3) / 4)));
    HIPCHK(dalloc(&B.scal, 16));
    HIPCHK(dalloc(&B.atomic_energy, na));
    HIPCHK(dalloc(&B.heat_w, (size_t)9 * na));
    HIPCHK(dalloc(&B.fout, (size_t)3 * na));
    HIPCHK(dalloc(&c->hx, na));
    HIPCHK(dalloc(&c->hy, na));
    HIPCHK(dalloc(&c->hz, na));
    HIPCHK(dalloc(&c->hlist, (size_t)3 * nw));
    HIPCHK(dalloc(&c->dv, (size_t)3 * na));
    HIPCHK(dalloc(&c->dek, na));
    HIPCHK(dalloc(&c->dep, na));
    HIPCHK(dalloc(&c->dw, (size_t)9 * na));
    HIPCHK(dalloc(&c->dsew, (size_t)9 * na));
    HIPCHK(dalloc(&c->dout, 16));
    HIPCHK(dalloc(&c->dpk, PK_N));
    HIPCHK(dalloc(&c->dfsum, (size_t)3 * na));
    HIPCHK(hipHostMalloc((void**)&c->hpk, sizeof(double) * PK_N, hipHostMallocDefault));
    // coefficient streams
    const int l2 = pip2b_stream_len(), l3 = pip3b_stream_len();
    HIPCHK(dalloc(&B.cs2, l2));
    HIPCHK(dalloc(&B.cs3, l3));
    HIPCHK(hipMemcpy(B.cs2, pip2b_stream_values(), sizeof(double) * l2, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(B.cs3, pip3b_stream_values(), sizeof(double) * l3, hipMemcpyHostToDevice));
    return 0;
}

int ensure_clusters(mbgpu_ctx* c, long nd, long nt)
{
    MbBuffers& B = c->B;
    if (nd > B.dcap) {
        dfree(B.dimers);
        dfree(B.clus2);
        dfree(B.xv2);
        dfree(B.part2);
        dfree(B.e2);
        dfree(B.g2);
        dfree(B.s2);
        B.dcap = std::max<long>(nd + nd / 2, 1024);
        HIPCHK(dalloc(&B.dimers, (size_t)2 * B.dcap));
        HIPCHK(dalloc(&B.clus2, (size_t)18 * B.dcap));
        HIPCHK(dalloc(&B.xv2, (size_t)31 * B.dcap));
        HIPCHK(dalloc(&B.part2, (size_t)pip2b_nslices() * 32 * B.dcap));
        HIPCHK(dalloc(&B.e2, B.dcap));
        HIPCHK(dalloc(&B.g2, (size_t)18 * B.dcap));
        HIPCHK(dalloc(&B.s2, (size_t)3 * B.dcap));
    }
    if (nt > B.tcap) {
        dfree(B.trimers);
        dfree(B.clus3);
        dfree(B.xv3);
        dfree(B.part3);
        dfree(B.e3);
        dfree(B.g3);
        dfree(B.s3);
        B.tcap = std::max<long>(nt + nt / 2, 1024);
        HIPCHK(dalloc(&B.trimers, (size_t)3 * B.tcap));
        HIPCHK(dalloc(&B.clus3, (size_t)27 * B.tcap));
        HIPCHK(dalloc(&B.xv3, (size_t)36 * B.tcap));
        HIPCHK(dalloc(&B.part3, (size_t)pip3b_nslices() * 37 * B.tcap));
        HIPCHK(dalloc(&B.e3, B.tcap));
        HIPCHK(dalloc(&B.g3, (size_t)27 * B.tcap));
        HIPCHK(dalloc(&B.s3, (size_t)3 * B.tcap));
    }
    if (B.mixed && (B.fcap2 != B.dcap || B.fcap3 != B.tcap)) {
        dfree(B.xv2f);
        dfree(B.part2f);
        dfree(B.xv3f);
        dfree(B.part3f);
        HIPCHK(dalloc(&B.xv2f, (size_t)31 * B.dcap));
        HIPCHK(dalloc(&B.part2f, (size_t)pip2b_nslices() * 32 * B.dcap));
        HIPCHK(dalloc(&B.xv3f, (size_t)36 * B.tcap));
        HIPCHK(dalloc(&B.part3f, (size_t)pip3b_nslices() * 37 * B.tcap));
        B.fcap2 = B.dcap;
        B.fcap3 = B.tcap;
    }
    const int nblk = (int)((B.dcap + 127) / 128 + (B.tcap + 127) / 128 + (B.natms + 255) / 256 + 8);
    if (nblk > B.nblk_cap) {
        dfree(B.blk);
        B.nblk_cap = nblk;
        HIPCHK(dalloc(&B.blk, (size_t)8 * nblk));
    }
    return 0;
}

// Build the lists and size the cluster buffers; returns nd, nt through the context.
int build_lists(mbgpu_ctx* c)
{
    MbBuffers& B = c->B;
    {
        // link cells of the oxygens, at least the 2B cut-off (the larger one) wide
        std::string err;
        B.cells.V.on = 0;
        if (c->geom.imcon != 0 && cells_plan(B.cells, c->geom.c, MBGPU_R2F, B.nw, CELLS_MIN_WATERS, err))
            return fail(-1, "%s", err.c_str());
    }
    for (int attempt = 0; attempt < 3; ++attempt) {
        if (B.dcap == 0 && ensure_clusters(c, 1024, 1024)) return -1;
        int rc = mb_launch_lists(B, c->geom, c->stream);
        if (rc) return fail(rc, "list kernels failed: %s", hipGetErrorString((hipError_t)rc));
        int cnt[4];
        HIPCHK(hipMemcpyAsync(cnt, B.counts, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (cnt[2] & 2)
            return fail(-5, "listttm2: some water's entries are not three consecutive atom ids a, a+1, a+2 "
                            "within 1..%d", B.natms);
        if (cnt[2]) return fail(-2, "neighbour list capacity exceeded (more than %d/%d O-O neighbours)", B.cap2,
                                B.cap3);
        c->nd = cnt[0];
        c->nt = cnt[1];
        if (c->nd <= B.dcap && c->nt <= B.tcap) return 0;
        if (ensure_clusters(c, c->nd, c->nt)) return -1;  // lists were truncated: rebuild
    }
    return fail(-3, "cluster list sizing did not converge");
}

int run_clusters(mbgpu_ctx* c, double* fx, double* fy, double* fz)
{
    MbBuffers& B = c->B;
    int rc = mb_launch_dimers(B, B.counts, (int)B.dcap, c->geom, true, c->stream);
    if (!rc) rc = mb_launch_trimers(B, B.counts + 1, (int)B.tcap, (int)((B.dcap + 127) / 128), c->geom, true, c->stream);
    if (!rc) rc = mb_launch_accumulate(B, fx, fy, fz, c->stream);
    if (rc) return fail(rc, "cluster kernels failed (%d)", rc);
    return 0;
}

// MB-pol's energy, virial and stress (+=) from the 16 reduced sums s (2B then 3B)
void add_mbpol_scalars(const double s[16], double* engsrp, double* virsrp, double* stress)
{
    if (engsrp) *engsrp += s[0] + s[8];
    if (virsrp) *virsrp += s[1] + s[9];
    if (stress) {
        const double t[6] = {s[2] + s[10], s[3] + s[11], s[4] + s[12], s[5] + s[13], s[6] + s[14], s[7] + s[15]};
        stress[0] += t[0];
        stress[1] += t[1];
        stress[2] += t[2];
        stress[3] += t[1];
        stress[4] += t[3];
        stress[5] += t[4];
        stress[6] += t[2];
        stress[7] += t[4];
        stress[8] += t[5];
    }
}

int finish_scalars(mbgpu_ctx* c, double* engsrp, double* virsrp, double* stress)
{
    double s[16];
    HIPCHK(hipMemcpyAsync(s, c->B.scal, sizeof s, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    add_mbpol_scalars(s, engsrp, virsrp, stress);
    return 0;
}

// ---------------------------------------------------------------- leaf context ----
std::mutex g_leaf_mu;
mbgpu_ctx* g_leaf = nullptr;

mbgpu_ctx* leaf_ctx()
{
    if (!g_leaf) {
        g_leaf = mbgpu_create(3, 12);
        if (g_leaf && ensure_clusters(g_leaf, 1, 1)) {
            mbgpu_destroy(g_leaf);
            g_leaf = nullptr;
        }
    }
    return g_leaf;
}

int leaf_eval(int nmon, const double* const* w, double* E, double* gout)
{
    std::lock_guard<std::mutex> lk(g_leaf_mu);
    mbgpu_ctx* c = leaf_ctx();
    if (!c) return -1;
    MbBuffers& B = c->B;
    // SoA with one cluster: element k of cluster 0 sits at k*nc with nc = 1
    std::vector<double> soa(9 * nmon);
    for (int m = 0; m < nmon; ++m)
        for (int k = 0; k < 9; ++k) soa[9 * m + k] = w[m][k];
    double* dclus = nmon == 2 ? B.clus2 : B.clus3;
    HIPCHK(hipMemcpyAsync(dclus, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice, c->stream));
    // one cluster (the count 1 in counts[0] and counts[1]), stride 1
    const int one[2] = {1, 1};
    HIPCHK(hipMemcpyAsync(B.counts, one, sizeof one, hipMemcpyHostToDevice, c->stream));
    int rc = nmon == 2 ? mb_launch_dimers(B, B.counts, 1, c->geom, false, c->stream)
                       : mb_launch_trimers(B, B.counts + 1, 1, 0, c->geom, false, c->stream);
    if (rc) return fail(rc, "leaf kernels failed");
    HIPCHK(hipMemcpyAsync(E, nmon == 2 ? B.e2 : B.e3, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(gout, nmon == 2 ? B.g2 : B.g3, sizeof(double) * 9 * nmon, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// Sum the per-atom terms of the model parts whose last evaluation in this context completed
// (the electrostatics' energies, contractions W and reciprocal stresses, the water terms'
// energies and W) into the heat-flux extras.  `x` is either NULL or already a context
// scratch buffer (then the terms are added in place); a NULL extra becomes the single
// term itself, or the scratch buffer holding the sum of several.
void add_terms(mbgpu_ctx* c, const double*& x, const double* const* terms, int nterms, double* scratch, long n,
               hipStream_t s)
{
    int k = 0;
    if (!x) {
        if (nterms == 0) return;
        if (nterms == 1) {
            x = terms[0];
            return;
        }
        el_launch_copy(scratch, terms[0], n, s);
        x = scratch;
        k = 1;
    }
    for (; k < nterms; ++k) el_launch_add(const_cast<double*>(x), terms[k], n, s);
}

bool have_model_terms(const mbgpu_ctx* c) { return c->el.ready || c->wt.ready; }

void with_model_terms(mbgpu_ctx* c, const double*& ep, const double*& we, const double*& se)
{
    const int nr = 3 * c->B.nw;
    const double *es[2], *ws[2], *ss[1];
    int ne = 0, ns = 0;
    if (c->el.ready) {
        es[ne] = c->el.E.eout;
        ws[ne++] = c->el.E.wout;
        ss[ns++] = c->el.E.sout;
    }
    if (c->wt.ready) {
        es[ne] = c->wt.B.eout;
        ws[ne++] = c->wt.B.wout;
    }
    add_terms(c, ep, es, ne, c->dep, nr, c->stream);
    add_terms(c, we, ws, ne, c->dw, 9L * nr, c->stream);
    add_terms(c, se, ss, ns, c->dsew, 9L * nr, c->stream);
}

}  // namespace

// ------------------------------------------------------- no-allocation scratch ----
// Scratch of mbgpu_ring_spring_dev_noalloc: the block partials of the spring kernels and
// their two sums.  A file-scope device array, so each GPU the library runs on holds its own
// copy (256 KiB) until the process ends.  Host code never uses the array's name as a device
// address: its device address is looked up once per device (hipGetSymbolAddress, with that
// device current) and kept.  Every call on one GPU shares that GPU's copy, so calls that
// overlap on one GPU race on it.
static __device__ __align__(256) double scratch[MBGPU_NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// device address of this GPU's scratch (the context's device must be current)
double* noalloc_scratch(mbgpu_ctx* c)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        fail(-1, "no current HIP device");
        return nullptr;
    }
    if (dev != c->device) {
        fail(-1, "the context belongs to device %d but device %d is current", c->device, dev);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fail((int)e, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

// ====================================================================== C ABI =======
namespace {

// The device work of one complete evaluation (mbgpu_evaluate_dev), kernels only, in three
// parts.  The electrostatics, on their stream b: positions in, start dipoles, gb solver
// iterations, the final pass, M sites and dipoles out.
int full_enqueue_elec(mbgpu_ctx* c, const mbgpu_elec_params* ep, int gb, double* x, double* y, double* z,
                      double* dip, double* fsite, hipStream_t b)
{
    ElecBuffers& E = c->el.E;
    const int nw = c->B.nw, n = 4 * nw, nr = 3 * nw;
    {
        // positions in; start dipoles in, or zero
        double* d[4] = {E.x, E.y, E.z, E.dip};
        const double* sr[4] = {x, y, z, dip};
        const long len[4] = {n, n, n, 3L * n};
        el_launch_copy_segs(4, d, sr, len, b);
    }
    int rc = elec_enqueue_full(c->el, ep->distribute_msite, ep->solver, gb, b);
    if (rc) return fail(rc, "electrostatics kernels failed: %s", hipGetErrorString((hipError_t)rc));
    {
        // M-site positions out; dipoles and per-site forces out when asked for
        double* d[5] = {x + nr, y + nr, z + nr, nullptr, nullptr};
        const double* sr[5] = {E.x + nr, E.y + nr, E.z + nr, nullptr, nullptr};
        long len[5] = {nw, nw, nw, 0, 0};
        int k = 3;
        if (dip) {
            d[k] = dip;
            sr[k] = E.dip;
            len[k++] = 3L * n;
        }
        if (fsite) {
            d[k] = fsite;
            sr[k] = E.fsite;
            len[k++] = 3L * n;
        }
        el_launch_copy_segs(k, d, sr, len, b);
    }
    return 0;
}

// MB-pol's lists and clusters and the water terms, on the context's stream a.
int full_enqueue_mbpol_water(mbgpu_ctx* c, double* x, double* y, double* z, double* fx, double* fy, double* fz,
                             hipStream_t a)
{
    MbBuffers& B = c->B;
    if (fx) {
        el_launch_zero(fx, B.natms, a);
        el_launch_zero(fy, B.natms, a);
        el_launch_zero(fz, B.natms, a);
    }
    int rc = mb_launch_lists(B, c->geom, a);
    if (!rc) rc = mb_launch_dimers(B, B.counts, (int)B.dcap, c->geom, true, a);
    if (!rc) rc = mb_launch_trimers(B, B.counts + 1, (int)B.tcap, (int)((B.dcap + 127) / 128), c->geom, true, a);
    if (!rc) rc = mb_launch_accumulate(B, fx, fy, fz, a);
    if (!rc) rc = water_enqueue(c->wt, x, y, z, fx, fy, fz, a);
    if (rc) return fail(rc, "MB-pol or water-term kernels failed: %s", hipGetErrorString((hipError_t)rc));
    return 0;
}

// After both, on stream a: the per-atom terms of the electrostatics, then of the water terms
// (with_model_terms' order), the heat flux over every term (skipped without velocities), and
// every scalar packed into dpk.
int full_enqueue_tail(mbgpu_ctx* c, const double* vx, const double* vy, const double* vz, const double* ek,
                      const int* spflag, hipStream_t a)
{
    MbBuffers& B = c->B;
    ElecBuffers& E = c->el.E;
    const long nr = 3L * B.nw;
    // per-atom energies and W: electrostatics + water terms (with_model_terms' order, one
    // launch); the reciprocal stresses are the electrostatics' own
    const double* ep_ = c->dep;
    const double* we_ = c->dw;
    const double* se_ = E.sout;
    if (vx) {
        double* d[2] = {c->dep, c->dw};
        const double* ta[2] = {E.eout, E.wout};
        const double* tb[2] = {c->wt.B.eout, c->wt.B.wout};
        const long len[2] = {nr, 9L * nr};
        el_launch_sum_segs(2, d, ta, tb, len, a);
    }
    if (vx) {
        const int rc = mb_launch_heat_flux((int)nr, vx, vy, vz, ek, B.atomic_energy, ep_, B.heat_w, we_, se_,
                                           B.blk, B.nblk_cap, c->dout, a);
        if (rc) return fail(rc, "heat-flux kernels failed");
    }
    hipLaunchKernelGGL(k_pack_eval, dim3(1), dim3(128), 0, a, B.counts, B.scal, E.state, E.err, E.scal, spflag,
                       c->wt.B.scal, c->dout, c->dpk);
    HIPCHK(hipGetLastError());
    return 0;
}

// One part captured on stream s into ex (kernel nodes only); 0 or an error code, ex null then.
template <typename F>
int capture_part(hipStream_t s, F&& enqueue, hipGraphExec_t& ex)
{
    hipGraph_t graph = nullptr;
    int rc = 0;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = enqueue();
        const hipError_t e2 = hipStreamEndCapture(s, &graph);
        if (e2 != hipSuccess) e = e2;
    }
    if (e == hipSuccess && !rc && graph) e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess || rc || !ex) {
        (void)hipGetLastError();
        if (ex) (void)hipGraphExecDestroy(ex);
        ex = nullptr;
        return rc ? rc : (e != hipSuccess ? (int)e : -1);
    }
    return 0;
}

// The key of a captured complete evaluation: everything its launches were recorded with.
std::vector<unsigned char> full_key(const mbgpu_ctx* c, const mbgpu_elec_params* ep, int gb, const void* const* ptrs,
                                    int nptr)
{
    std::vector<unsigned char> k;
    auto add = [&](const void* p, size_t n) {
        const unsigned char* q = static_cast<const unsigned char*>(p);
        k.insert(k.end(), q, q + n);
    };
    add(&c->B, sizeof c->B);
    add(&c->geom, sizeof c->geom);
    add(&c->wt.B, sizeof c->wt.B);
    add(&c->wt.P, sizeof c->wt.P);
    add(&c->wt.cells, sizeof c->wt.cells);
    add(ptrs, sizeof(void*) * nptr);
    const int v[2] = {ep->distribute_msite, ep->solver};
    add(v, sizeof v);
    elec_key_append(c->el, ep->distribute_msite, ep->solver, gb, k);
    return k;
}

}  // namespace

extern "C" {

const char* mbgpu_last_error(void) { return g_err.c_str(); }

int mbgpu_check_listttm2(int nttm2, const int* listttm2, int natms)
{
    if (nttm2 < 0 || nttm2 % 3 != 0 || (nttm2 > 0 && !listttm2)) return -1;
    for (int w = 0; w < nttm2 / 3; ++w)
        if (water_base(listttm2, w, natms) < 0) return w + 1;
    return 0;
}

mbgpu_ctx* mbgpu_create(int nw, int natms) { return mbgpu_create_queues(nw, natms, 0); }

mbgpu_ctx* mbgpu_create_queues(int nw, int natms, int dedicated)
{
    EntryGuard gate_;
    if (dedicated != 0 && dedicated != 1) {
        fail(-1, "mbgpu_create_queues: dedicated is 0 or 1 (got %d)", dedicated);
        return nullptr;
    }
    if (nw < 1 || natms < 3 * nw) {
        fail(-1, "mbgpu_create: need nw >= 1 and natms >= 3*nw (got %d, %d)", nw, natms);
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        fail(-1, "mbgpu_create: no HIP device");
        return nullptr;
    }
    mbgpu_ctx* c = new mbgpu_ctx();
    c->B.nw = nw;
    c->B.natms = natms;
    (void)hipGetDevice(&c->device);
    // the electrostatics' stream does not synchronize with the null stream: it is captured
    // into graphs (mbgpu_set_graphs), and another thread's null-stream work during a capture
    // would otherwise invalidate it; nothing relies on that implicit ordering, since every
    // entry waits for its streams before it returns and takes inputs the caller has finished
    // With `dedicated`, both streams are created with a full CU mask, for which the runtime
    // makes hardware queues of their own: the electrostatics then always overlap MB-pol and
    // the water terms (streams of shared queues can land on one queue).  Such streams are
    // blocking ones.
    hipError_t se = hipSuccess;
    if (dedicated) {
        hipDeviceProp_t prop;
        se = hipGetDeviceProperties(&prop, c->device);
        const int ncu = se == hipSuccess ? prop.multiProcessorCount : 0, nwd = (ncu + 31) / 32;
        std::vector<uint32_t> mask(nwd > 0 ? nwd : 1, 0xffffffffu);
        if (ncu % 32) mask[nwd - 1] = (1u << (ncu % 32)) - 1u;
        if (se == hipSuccess && ncu < 1) se = hipErrorInvalidDevice;
        if (se == hipSuccess) se = hipExtStreamCreateWithCUMask(&c->stream, nwd, mask.data());
        if (se == hipSuccess) se = hipExtStreamCreateWithCUMask(&c->el_stream, nwd, mask.data());
    } else {
        se = hipStreamCreate(&c->stream);
        if (se == hipSuccess) se = hipStreamCreateWithFlags(&c->el_stream, hipStreamNonBlocking);
    }
    if (se != hipSuccess) fail((int)se, "streams: %s", hipGetErrorString(se));
    if (se != hipSuccess || alloc_fixed(c) ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess) {
        std::string msg = g_err;
        mbgpu_destroy(c);
        g_err = "mbgpu_create: " + msg;
        return nullptr;
    }
    double cell[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    make_geom(1, cell, c->geom);
    return c;
}

void mbgpu_destroy(mbgpu_ctx* c)
{
    EntryGuard gate_;
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->el_stream) (void)hipStreamSynchronize(c->el_stream);
    for (hipGraphExec_t& g : c->fexec)
        if (g) (void)hipGraphExecDestroy(g);
    MbBuffers& B = c->B;
    dfree(B.mon); dfree(B.wbase); dfree(B.nbr2); dfree(B.cnt2); dfree(B.off2); dfree(B.nbr3); dfree(B.cnt3);
    dfree(B.tcnt); dfree(B.toff); dfree(B.counts); dfree(B.dimers); dfree(B.trimers);
    dfree(B.clus2); dfree(B.xv2); dfree(B.part2); dfree(B.e2); dfree(B.g2); dfree(B.s2);
    dfree(B.clus3); dfree(B.xv3); dfree(B.part3); dfree(B.e3); dfree(B.g3); dfree(B.s3);
    dfree(B.is_water); dfree(B.nbr2l); dfree(B.cnt2l); dfree(B.kmask); dfree(B.rowoff); dfree(B.blk); dfree(B.scal); dfree(B.atomic_energy); dfree(B.heat_w); dfree(B.fout);
    dfree(B.cs2); dfree(B.cs3);
    cells_free(B.cells);
    dfree(B.xv2f); dfree(B.part2f); dfree(B.cs2f); dfree(B.xv3f); dfree(B.part3f); dfree(B.cs3f);
    dfree(c->hx); dfree(c->hy); dfree(c->hz); dfree(c->hlist);
    dfree(c->dv); dfree(c->dek); dfree(c->dep); dfree(c->dw); dfree(c->dsew); dfree(c->dout); dfree(c->dpk); dfree(c->dfsum);
    if (c->hpk) (void)hipHostFree(c->hpk);
    dfree(c->ring_m); dfree(c->ring_blk); dfree(c->ring_out);
    spme_free(c->sp);
    elec_release(c->el);
    water_release(c->wt);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->ev3) (void)hipEventDestroy(c->ev3);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->el_stream) (void)hipStreamDestroy(c->el_stream);
    delete c;
}

int mbgpu_set_elec_reciprocal(mbgpu_ctx* c, int mode, int order, const int grid[3])
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (mode == 0) {
        c->el.recip = 0;
        c->el.ready = false;
        return 0;
    }
    if (mode != 1) return fail(-1, "reciprocal mode %d: 0 = Ewald sum over k, 1 = particle mesh", mode);
    if (!grid) return fail(-1, "null grid");
    if (order < 4 || order > SPME_MAX_ORDER || (order & 1))
        return fail(-1, "order must be even, 4..%d (got %d)", SPME_MAX_ORDER, order);
    for (int a = 0; a < 3; ++a) {
        int K = grid[a];
        const int K0 = K;
        for (int r : {2, 3, 5})
            while (K > 0 && K % r == 0) K /= r;
        if (K0 < 8 || K0 > SPME_MAX_GRID || K != 1 || K0 < order)
            return fail(-1, "grid[%d] = %d: each grid size must be 2^a 3^b 5^c in [max(8, order), %d]", a, K0,
                        SPME_MAX_GRID);
    }
    c->el.recip = 1;
    c->el.sp_order = order;
    for (int a = 0; a < 3; ++a) c->el.sp_grid[a] = grid[a];
    c->el.ready = false;
    return 0;
}

int mbgpu_set_graphs(mbgpu_ctx* c, int on)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (on != 0 && on != 1) return fail(-1, "graphs: 0 = launch kernel by kernel, 1 = capture and replay (got %d)", on);
    c->el.graphs = on;
    if (!on) elec_drop_graph(c->el);
    return 0;
}

int mbgpu_get_graph_counts(mbgpu_ctx* c, long counts[2])
{
    EntryGuard gate_;
    if (!c || !counts) return fail(-1, "null context or output");
    counts[0] = c->el.gcaptures;
    counts[1] = c->el.greplays;
    return 0;
}

int mbgpu_set_precision(mbgpu_ctx* c, int mode)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (mode < 0 || mode > 3)
        return fail(-1, "precision mode %d: bit 0 = FP32 2B polynomial, bit 1 = FP32 3B polynomial", mode);
    MbBuffers& B = c->B;
    if (mode != 0 && !B.cs2f) {
        const int l2 = pip2b_stream_len(), l3 = pip3b_stream_len();
        std::vector<float> f2(l2), f3(l3);
        for (int i = 0; i < l2; ++i) f2[i] = (float)pip2b_stream_values()[i];
        for (int i = 0; i < l3; ++i) f3[i] = (float)pip3b_stream_values()[i];
        HIPCHK(dalloc(&B.cs2f, l2));
        HIPCHK(dalloc(&B.cs3f, l3));
        HIPCHK(hipMemcpy(B.cs2f, f2.data(), sizeof(float) * l2, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(B.cs3f, f3.data(), sizeof(float) * l3, hipMemcpyHostToDevice));
    }
    B.mixed = mode;
    if (mode != 0 && B.dcap > 0 && ensure_clusters(c, B.dcap, B.tcap)) return -1;
    return 0;
}

int mbgpu_mbpol_forces_dev(mbgpu_ctx* c, int imcon, const double cell[9], int nttm2, const int* listttm2_dev,
                           const double* x_dev, const double* y_dev, const double* z_dev, double* fx_dev,
                           double* fy_dev, double* fz_dev, double* engsrp, double* virsrp, double stress[9])
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (nttm2 != 3 * c->B.nw) return fail(-1, "nttm2 = %d does not match the context (3*%d)", nttm2, c->B.nw);
    if (!make_geom(imcon, cell, c->geom))
        return fail(-4, "imcon = %d is not supported by image_vectors (reference stops in fixme)", imcon);
    MbBuffers& B = c->B;
    B.list = const_cast<int*>(listttm2_dev);
    B.x = const_cast<double*>(x_dev);
    B.y = const_cast<double*>(y_dev);
    B.z = const_cast<double*>(z_dev);
    c->have_inputs = true;
    c->mb_ready = false;
    if (build_lists(c)) return -1;
    if (run_clusters(c, fx_dev, fy_dev, fz_dev)) return -1;
    if (finish_scalars(c, engsrp, virsrp, stress)) return -1;
    c->mb_ready = true;
    return 0;
}

int mbgpu_mbpol_forces(mbgpu_ctx* c, int imcon, const double cell[9], int nttm2, const int* listttm2,
                       const double* xxx, const double* yyy, const double* zzz, double* fxx, double* fyy,
                       double* fzz, double* engsrp, double* virsrp, double stress[9])
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    const int na = c->B.natms;
    if (nttm2 != 3 * c->B.nw) return fail(-1, "nttm2 = %d does not match the context (3*%d)", nttm2, c->B.nw);
    if (!listttm2) return fail(-1, "null listttm2");
    if (const int bad = mbgpu_check_listttm2(nttm2, listttm2, na))
        return fail(-5, "listttm2(%d..%d) = %d %d %d: not three consecutive atom ids within 1..%d", 3 * bad - 2,
                    3 * bad, listttm2[3 * bad - 3], listttm2[3 * bad - 2], listttm2[3 * bad - 1], na);
    HIPCHK(hipMemcpyAsync(c->hlist, listttm2, sizeof(int) * nttm2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->hx, xxx, sizeof(double) * na, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->hy, yyy, sizeof(double) * na, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->hz, zzz, sizeof(double) * na, hipMemcpyHostToDevice, c->stream));
    int rc = mbgpu_mbpol_forces_dev(c, imcon, cell, nttm2, c->hlist, c->hx, c->hy, c->hz, nullptr, nullptr,
                                    nullptr, engsrp, virsrp, stress);
    if (rc) return rc;
    if (fxx || fyy || fzz) {
        std::vector<double> f((size_t)3 * na);
        HIPCHK(hipMemcpyAsync(f.data(), c->B.fout, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < na; ++i) {
            if (fxx) fxx[i] += f[3 * i];
            if (fyy) fyy[i] += f[3 * i + 1];
            if (fzz) fzz[i] += f[3 * i + 2];
        }
    }
    return 0;
}

int mbgpu_get_decomposition(mbgpu_ctx* c, double* ae, double* hw, double* forces)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    const int na = c->B.natms;
    if (ae) HIPCHK(hipMemcpyAsync(ae, c->B.atomic_energy, sizeof(double) * na, hipMemcpyDeviceToHost, c->stream));
    if (hw) HIPCHK(hipMemcpyAsync(hw, c->B.heat_w, sizeof(double) * 9 * na, hipMemcpyDeviceToHost, c->stream));
    if (forces)
        HIPCHK(hipMemcpyAsync(forces, c->B.fout, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int mbgpu_get_lists(mbgpu_ctx* c, int* dimers, long dcap, long* nd, int* trimers, long tcap, long* nt)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (nd) *nd = c->nd;
    if (nt) *nt = c->nt;
    if (dimers && c->nd)
        HIPCHK(hipMemcpyAsync(dimers, c->B.dimers, sizeof(int) * 2 * std::min(dcap, c->nd), hipMemcpyDeviceToHost,
                              c->stream));
    if (trimers && c->nt)
        HIPCHK(hipMemcpyAsync(trimers, c->B.trimers, sizeof(int) * 3 * std::min(tcap, c->nt),
                              hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int mbgpu_heat_flux(mbgpu_ctx* c, const double* vxx, const double* vyy, const double* vzz, const double* e_kin,
                    const double* e_pot_extra, const double* w_extra, const double* stress_ew1p, double jkin[3],
                    double jpot[3], double jtot[3])
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    const int nr = 3 * c->B.nw;
    double* v = c->dv;
    HIPCHK(hipMemcpyAsync(v, vxx, sizeof(double) * nr, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(v + nr, vyy, sizeof(double) * nr, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(v + 2 * nr, vzz, sizeof(double) * nr, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dek, e_kin, sizeof(double) * nr, hipMemcpyHostToDevice, c->stream));
    if (e_pot_extra) HIPCHK(hipMemcpyAsync(c->dep, e_pot_extra, sizeof(double) * nr, hipMemcpyHostToDevice, c->stream));
    if (w_extra) HIPCHK(hipMemcpyAsync(c->dw, w_extra, sizeof(double) * 9 * nr, hipMemcpyHostToDevice, c->stream));
    if (stress_ew1p)
        HIPCHK(hipMemcpyAsync(c->dsew, stress_ew1p, sizeof(double) * 9 * nr, hipMemcpyHostToDevice, c->stream));
    if (c->B.nblk_cap == 0 && ensure_clusters(c, 1, 1)) return -1;
    const double *ep = e_pot_extra ? c->dep : nullptr, *we = w_extra ? c->dw : nullptr,
                 *se = stress_ew1p ? c->dsew : nullptr;
    with_model_terms(c, ep, we, se);
    int rc = mb_launch_heat_flux(nr, v, v + nr, v + 2 * nr, c->dek, c->B.atomic_energy, ep, c->B.heat_w, we, se,
                                 c->B.blk, c->B.nblk_cap, c->dout, c->stream);
    if (rc) return fail(rc, "heat-flux kernels failed");
    double out[8];
    HIPCHK(hipMemcpyAsync(out, c->dout, sizeof out, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; ++a) {
        jkin[a] = out[a];
        jpot[a] = out[3 + a];
        jtot[a] = out[a] + out[3 + a];
    }
    return 0;
}

int mbgpu_heat_flux_dev(mbgpu_ctx* c, const double* vx, const double* vy, const double* vz, const double* e_kin,
                        const double* e_pot_extra, const double* w_extra, const double* stress_ew1p, double jkin[3],
                        double jpot[3], double jtot[3])
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    const int nr = 3 * c->B.nw;
    if (c->B.nblk_cap == 0 && ensure_clusters(c, 1, 1)) return -1;
    const double *ep = e_pot_extra, *we = w_extra, *se = stress_ew1p;
    if (have_model_terms(c)) {
        // the model terms are added to copies of the caller's extras
        if (ep) HIPCHK(hipMemcpyAsync(c->dep, ep, sizeof(double) * nr, hipMemcpyDeviceToDevice, c->stream));
        if (we) HIPCHK(hipMemcpyAsync(c->dw, we, sizeof(double) * 9 * nr, hipMemcpyDeviceToDevice, c->stream));
        if (se) HIPCHK(hipMemcpyAsync(c->dsew, se, sizeof(double) * 9 * nr, hipMemcpyDeviceToDevice, c->stream));
        ep = ep ? c->dep : nullptr;
        we = we ? c->dw : nullptr;
        se = se ? c->dsew : nullptr;
    }
    with_model_terms(c, ep, we, se);
    int rc = mb_launch_heat_flux(nr, vx, vy, vz, e_kin, c->B.atomic_energy, ep, c->B.heat_w, we, se, c->B.blk,
                                 c->B.nblk_cap, c->dout, c->stream);
    if (rc) return fail(rc, "heat-flux kernels failed");
    double out[8];
    HIPCHK(hipMemcpyAsync(out, c->dout, sizeof out, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; ++a) {
        jkin[a] = out[a];
        jpot[a] = out[3 + a];
        jtot[a] = out[a] + out[3 + a];
    }
    return 0;
}

}  // extern "C"

namespace {

// mbgpu_evaluate_dev and mbgpu_evaluate_forces_dev: forces written into fx/fy/fz (MB-pol and
// water terms, natms each) and fsite (electrostatics, 3 per site) when given; the heat flux
// when the velocities are given.
int evaluate_impl(mbgpu_ctx* c, int imcon, const double cell[9], int nttm2, const int* listttm2_dev,
                  const mbgpu_elec_params* ep, const mbgpu_water_params* wp, double* x_dev, double* y_dev,
                  double* z_dev, double* dip_dev, double* fx_dev, double* fy_dev, double* fz_dev,
                  double* fsite_dev, const double* vx_dev, const double* vy_dev, const double* vz_dev,
                  const double* e_kin_dev, double out[MBGPU_EVAL_OUT])
{
    const int nw = c->B.nw;
    const bool heat = vx_dev != nullptr;
    if (nttm2 != 3 * nw) return fail(-1, "nttm2 = %d does not match the context (3*%d)", nttm2, nw);
    if (c->B.natms < 4 * nw)
        return fail(-1, "the context holds %d sites; the electrostatics need 4*nw = %d", c->B.natms, 4 * nw);
    for (int i = 0; i < MBGPU_EVAL_OUT; ++i) out[i] = 0.0;
    CellGeom g;
    if (!make_geom(imcon, cell, g) || !make_geom(imcon, cell, c->geom))
        return fail(-4, "imcon = %d is not supported by images (reference stops in fixme)", imcon);
    // host-side set-up of the three parts: tables, k vectors, plans, buffers
    MbBuffers& B = c->B;
    B.list = const_cast<int*>(listttm2_dev);
    B.x = x_dev;
    B.y = y_dev;
    B.z = z_dev;
    c->have_inputs = true;
    c->mb_ready = false;
    std::string err;
    B.cells.V.on = 0;
    if (imcon != 0 && cells_plan(B.cells, c->geom.c, MBGPU_R2F, nw, CELLS_MIN_WATERS, err))
        return fail(-1, "%s", err.c_str());
    c->el.ready = false;
    if (elec_configure(c->el, nw, g, cell, ep, err)) return fail(-1, "%s", err.c_str());
    c->wt.ready = false;
    if (water_configure(c->wt, nw, g, cell, wp, c->stream, err)) return fail(-1, "%s", err.c_str());
    const int gb = elec_batch(c->el);
    bool fused_ran = false;
    // one enqueue (a graph replay once captured) when every buffer has been sized by an
    // evaluation through the separate entries
    if (B.dcap > 0 && B.nblk_cap > 0 && gb > 0) {
        const int* spf = c->el.recip && c->el.sp && imcon != 0 ? c->el.sp->flag : nullptr;
        const void* ptrs[18] = {listttm2_dev, x_dev, y_dev, z_dev, dip_dev, vx_dev, vy_dev, vz_dev, e_kin_dev,
                                c->dep, c->dw, c->dsew, spf, c->dpk, fx_dev, fy_dev, fz_dev, fsite_dev};
        const char* gv = getenv("MBGPU_GRAPHS");
        const bool graphs = c->el.graphs && !(gv && atoi(gv) == 0);
        hipStream_t a = c->stream, b = c->el_stream;
        auto part_el = [&] { return full_enqueue_elec(c, ep, gb, x_dev, y_dev, z_dev, dip_dev, fsite_dev, b); };
        auto part_mb = [&] { return full_enqueue_mbpol_water(c, x_dev, y_dev, z_dev, fx_dev, fy_dev, fz_dev, a); };
        auto part_tail = [&] { return full_enqueue_tail(c, vx_dev, vy_dev, vz_dev, e_kin_dev, spf, b); };
        bool replay = false;
        if (graphs) {
            std::vector<unsigned char> key = full_key(c, ep, gb, ptrs, 18);
            if (!c->fexec[0] || key != c->fkey) {
                for (hipGraphExec_t& g : c->fexec)
                    if (g) (void)hipGraphExecDestroy(g);
                c->fexec[0] = c->fexec[1] = c->fexec[2] = nullptr;
                c->fkey.clear();
                capture_gate_begin();
                int cr = capture_part(b, part_el, c->fexec[0]);
                if (!cr) cr = capture_part(a, part_mb, c->fexec[1]);
                if (!cr) cr = capture_part(b, part_tail, c->fexec[2]);
                capture_gate_end();
                if (cr) {
                    // the runtime refused a capture: this context launches kernel by kernel from now on
                    for (hipGraphExec_t& g : c->fexec)
                        if (g) (void)hipGraphExecDestroy(g);
                    c->fexec[0] = c->fexec[1] = c->fexec[2] = nullptr;
                    c->el.graphs = 0;
                } else {
                    c->fkey.swap(key);
                    c->el.gbatch = gb;
                    ++c->fcaptures;
                }
            }
            replay = c->fexec[0] != nullptr;
        }
        // The electrostatics on b, MB-pol and the water terms on a, joined on the host: the
        // tail follows the electrostatics on b once a has finished.  No stream waits on an
        // event of the other: with more streams than hardware queues (contexts evaluated
        // concurrently) such a wait holds up whatever other stream shares its queue.  The
        // inputs are complete before the call, and the set-up above enqueues on a only.
        int rc = 0;
        if (replay) {
            HIPCHK(hipGraphLaunch(c->fexec[0], b));
            HIPCHK(hipGraphLaunch(c->fexec[1], a));
            ++c->freplays;
        } else {
            rc = part_el();
            if (!rc) rc = part_mb();
        }
        if (!rc) {
            HIPCHK(hipStreamSynchronize(a));
            if (replay) HIPCHK(hipGraphLaunch(c->fexec[2], b));
            else rc = part_tail();
        }
        if (rc) {
            (void)hipStreamSynchronize(b);
            (void)hipStreamSynchronize(a);
            return rc;
        }
        fused_ran = true;
        const double* pk = c->hpk;
        HIPCHK(hipMemcpyAsync(c->hpk, c->dpk, sizeof(double) * PK_N, hipMemcpyDeviceToHost, b));
        HIPCHK(hipStreamSynchronize(b));
        int cnt[4], spflag[2];
        for (int i = 0; i < 4; ++i) cnt[i] = (int)pk[PK_CNT + i];
        const double* mb = pk + PK_MB;
        const double* st = pk + PK_ST;
        const int herr = (int)pk[PK_ERR];
        for (int i = 0; i < 41; ++i) c->el.scal[i] = pk[PK_EL + i];
        spflag[0] = (int)pk[PK_SPF];
        spflag[1] = (int)pk[PK_SPF + 1];
        const double* wr = pk + PK_WT;
        const double* jd = pk + PK_J;
        // anything out of the ordinary (a list past its capacity or with a flag, neighbour rows
        // overflowed, a solve longer than the graph holds) goes through the separate entries
        const bool ok = cnt[2] == 0 && cnt[0] <= B.dcap && cnt[1] <= B.tcap && !herr && st[2] != 0.0;
        if (ok) {
            c->nd = cnt[0];
            c->nt = cnt[1];
            c->mb_ready = true;
            add_mbpol_scalars(mb, &out[0], &out[1], &out[2]);
            c->el.last_iters = (int)st[0];
            if (elec_finish(c->el, st, spflag, err)) return fail(-1, "%s", err.c_str());
            const double* o = c->el.out;
            out[11] = o[4];
            out[12] = o[5];
            for (int i = 0; i < 9; ++i) out[13 + i] = o[10 + i];
            out[22] = o[8];
            water_finish(c->wt, wr);
            const double* w = c->wt.out;
            for (int i = 0; i < 4; ++i) {
                out[23 + i] = w[2 * i];
                out[27 + i] = w[2 * i + 1];
            }
            for (int i = 0; i < 9; ++i) out[31 + i] = c->wt.stress[i];
            if (heat)
                for (int a2 = 0; a2 < 3; ++a2) {
                    out[40 + a2] = jd[a2];
                    out[43 + a2] = jd[3 + a2];
                    out[46 + a2] = jd[a2] + jd[3 + a2];
                }
            out[49] = 1.0;
            return 0;
        }
    }
    // the separate entries: they size every buffer, repeat overflowed builds, continue
    // unconverged solves and report errors.  After a one-call attempt they start again from
    // the caller's start dipoles (kept in the electrostatics' dip0; the attempt wrote its
    // dipoles back into dip_dev).
    if (fused_ran && dip_dev) {
        el_launch_copy(dip_dev, c->el.E.dip0, 3L * 4 * nw, c->el_stream);
        HIPCHK(hipStreamSynchronize(c->el_stream));
    }
    // forces are written: cleared here, then added by the separate entries
    if (fx_dev) {
        el_launch_zero(fx_dev, c->B.natms, c->stream);
        el_launch_zero(fy_dev, c->B.natms, c->stream);
        el_launch_zero(fz_dev, c->B.natms, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (fsite_dev) {
        el_launch_zero(fsite_dev, 3L * 4 * nw, c->el_stream);
        HIPCHK(hipStreamSynchronize(c->el_stream));
    }
    int niter = 0;
    int rc = mbgpu_mbpol_forces_dev(c, imcon, cell, nttm2, listttm2_dev, x_dev, y_dev, z_dev, fx_dev, fy_dev,
                                    fz_dev, &out[0], &out[1], &out[2]);
    if (rc) return rc;
    rc = mbgpu_polar_electrostatics_dev(c, imcon, cell, ep, x_dev, y_dev, z_dev, dip_dev, fsite_dev, &out[11],
                                        &out[12], &out[13], &niter);
    if (rc) return rc;
    out[22] = niter;
    double ew[4], vw[4];
    rc = mbgpu_water_terms_dev(c, imcon, cell, wp, x_dev, y_dev, z_dev, fx_dev, fy_dev, fz_dev, ew, vw,
                               &out[31]);
    if (rc) return rc;
    for (int i = 0; i < 4; ++i) {
        out[23 + i] = ew[i];
        out[27 + i] = vw[i];
    }
    if (!heat) return 0;
    return mbgpu_heat_flux_dev(c, vx_dev, vy_dev, vz_dev, e_kin_dev, nullptr, nullptr, nullptr, &out[40], &out[43],
                               &out[46]);
}

}  // namespace

extern "C" {

int mbgpu_evaluate_dev(mbgpu_ctx* c, int imcon, const double cell[9], int nttm2, const int* listttm2_dev,
                       const mbgpu_elec_params* ep, const mbgpu_water_params* wp, double* x_dev, double* y_dev,
                       double* z_dev, double* dip_dev, const double* vx_dev, const double* vy_dev,
                       const double* vz_dev, const double* e_kin_dev, double out[MBGPU_EVAL_OUT])
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (!ep || !wp) return fail(-1, "null electrostatics or water-term parameters");
    if (!out || !listttm2_dev || !x_dev || !y_dev || !z_dev || !vx_dev || !vy_dev || !vz_dev || !e_kin_dev)
        return fail(-1, "null argument");
    return evaluate_impl(c, imcon, cell, nttm2, listttm2_dev, ep, wp, x_dev, y_dev, z_dev, dip_dev, nullptr,
                         nullptr, nullptr, nullptr, vx_dev, vy_dev, vz_dev, e_kin_dev, out);
}

int mbgpu_evaluate_forces_dev(mbgpu_ctx* c, int imcon, const double cell[9], int nttm2, const int* listttm2_dev,
                              const mbgpu_elec_params* ep, const mbgpu_water_params* wp, double* x_dev,
                              double* y_dev, double* z_dev, double* dip_dev, double* fx_dev, double* fy_dev,
                              double* fz_dev, double* fsite_dev, const double* vx_dev, const double* vy_dev,
                              const double* vz_dev, const double* e_kin_dev, double out[MBGPU_EVAL_OUT])
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (!ep || !wp) return fail(-1, "null electrostatics or water-term parameters");
    if (!out || !listttm2_dev || !x_dev || !y_dev || !z_dev || !fx_dev || !fy_dev || !fz_dev || !fsite_dev)
        return fail(-1, "null argument");
    const bool v0 = !vx_dev, v1 = !vy_dev, v2 = !vz_dev, v3 = !e_kin_dev;
    if (!(v0 == v1 && v1 == v2 && v2 == v3))
        return fail(-1, "velocities and kinetic energies: give all four arrays or none");
    return evaluate_impl(c, imcon, cell, nttm2, listttm2_dev, ep, wp, x_dev, y_dev, z_dev, dip_dev, fx_dev, fy_dev,
                         fz_dev, fsite_dev, vx_dev, vy_dev, vz_dev, e_kin_dev, out);
}

int mbgpu_get_forces(mbgpu_ctx* c, double* f)
{
    EntryGuard gate_;
    if (!c || !f) return fail(-1, "null context or output");
    if (!c->mb_ready || !c->el.ready || !c->wt.ready)
        return fail(-1, "no complete evaluation (MB-pol, electrostatics and water terms) in this context");
    const int nr = 3 * c->B.nw;
    hipLaunchKernelGGL(k_sum_forces, dim3((3 * nr + 255) / 256), dim3(256), 0, c->stream, nr, c->B.fout,
                       c->el.E.fsite, c->wt.B.fw, c->dfsum);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(f, c->dfsum, sizeof(double) * 3 * nr, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int mbgpu_get_evaluate_graph_counts(mbgpu_ctx* c, long counts[2])
{
    EntryGuard gate_;
    if (!c || !counts) return fail(-1, "null context or output");
    counts[0] = c->fcaptures;
    counts[1] = c->freplays;
    return 0;
}

int mbgpu_polar_electrostatics_dev(mbgpu_ctx* c, int imcon, const double cell[9], const mbgpu_elec_params* p,
                                   double* x_dev, double* y_dev, double* z_dev, double* dip_dev, double* f_dev,
                                   double* engcpe, double* vircpe, double stress[9], int* niter)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    const int nw = c->B.nw, n = 4 * nw;
    if (c->B.natms < n) return fail(-1, "the context holds %d sites; the electrostatics need 4*nw = %d", c->B.natms, n);
    CellGeom g;
    if (!make_geom(imcon, cell, g))
        return fail(-4, "imcon = %d is not supported by images (reference stops in fixme)", imcon);
    if (!p) return fail(-1, "null electrostatics parameters");
    std::string err;
    c->el.ready = false;
    if (elec_configure(c->el, nw, g, cell, p, err)) return fail(-1, "%s", err.c_str());
    ElecBuffers& E = c->el.E;
    HIPCHK(hipMemcpyAsync(E.x, x_dev, sizeof(double) * n, hipMemcpyDeviceToDevice, c->el_stream));
    HIPCHK(hipMemcpyAsync(E.y, y_dev, sizeof(double) * n, hipMemcpyDeviceToDevice, c->el_stream));
    HIPCHK(hipMemcpyAsync(E.z, z_dev, sizeof(double) * n, hipMemcpyDeviceToDevice, c->el_stream));
    if (dip_dev)
        HIPCHK(hipMemcpyAsync(E.dip, dip_dev, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->el_stream));
    else
        el_launch_zero(E.dip, 3L * n, c->el_stream);   // start from zero dipoles
    int rc = elec_evaluate(c->el, p->max_iter, p->distribute_msite, p->solver, c->el_stream, err);
    if (rc) return fail(rc, "%s", err.c_str());
    const int nr = 3 * nw;
    HIPCHK(hipMemcpyAsync(x_dev + nr, E.x + nr, sizeof(double) * nw, hipMemcpyDeviceToDevice, c->el_stream));
    HIPCHK(hipMemcpyAsync(y_dev + nr, E.y + nr, sizeof(double) * nw, hipMemcpyDeviceToDevice, c->el_stream));
    HIPCHK(hipMemcpyAsync(z_dev + nr, E.z + nr, sizeof(double) * nw, hipMemcpyDeviceToDevice, c->el_stream));
    if (dip_dev)
        HIPCHK(hipMemcpyAsync(dip_dev, E.dip, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->el_stream));
    if (f_dev) el_launch_add(f_dev, E.fsite, 3L * n, c->el_stream);
    HIPCHK(hipStreamSynchronize(c->el_stream));
    const double* o = c->el.out;
    if (engcpe) *engcpe += o[4];
    if (vircpe) *vircpe += o[5];
    if (stress)
        for (int i = 0; i < 9; ++i) stress[i] += o[10 + i];
    if (niter) *niter = (int)o[8];
    return 0;
}

int mbgpu_polar_electrostatics(mbgpu_ctx* c, int imcon, const double cell[9], const mbgpu_elec_params* p,
                               double* xxx, double* yyy, double* zzz, double* dipx, double* dipy, double* dipz,
                               double* chge, double* fxx, double* fyy, double* fzz, double* engcpe, double* vircpe,
                               double stress[9], int* niter)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    const int nw = c->B.nw, n = 4 * nw;
    if (c->B.natms < n) return fail(-1, "the context holds %d sites; the electrostatics need 4*nw = %d", c->B.natms, n);
    if (!xxx || !yyy || !zzz || !dipx || !dipy || !dipz) return fail(-1, "null position or dipole array");
    CellGeom g;
    if (!make_geom(imcon, cell, g))
        return fail(-4, "imcon = %d is not supported by images (reference stops in fixme)", imcon);
    if (!p) return fail(-1, "null electrostatics parameters");
    std::string err;
    c->el.ready = false;
    if (elec_configure(c->el, nw, g, cell, p, err)) return fail(-1, "%s", err.c_str());
    ElecBuffers& E = c->el.E;
    std::vector<double> dip((size_t)3 * n);
    for (int i = 0; i < n; ++i) {
        dip[3 * i] = dipx[i];
        dip[3 * i + 1] = dipy[i];
        dip[3 * i + 2] = dipz[i];
    }
    HIPCHK(hipMemcpyAsync(E.x, xxx, sizeof(double) * n, hipMemcpyHostToDevice, c->el_stream));
    HIPCHK(hipMemcpyAsync(E.y, yyy, sizeof(double) * n, hipMemcpyHostToDevice, c->el_stream));
    HIPCHK(hipMemcpyAsync(E.z, zzz, sizeof(double) * n, hipMemcpyHostToDevice, c->el_stream));
    HIPCHK(hipMemcpyAsync(E.dip, dip.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->el_stream));
    int rc = elec_evaluate(c->el, p->max_iter, p->distribute_msite, p->solver, c->el_stream, err);
    if (rc) return fail(rc, "%s", err.c_str());
    const int nr = 3 * nw;
    std::vector<double> f((size_t)3 * n);
    HIPCHK(hipMemcpyAsync(xxx + nr, E.x + nr, sizeof(double) * nw, hipMemcpyDeviceToHost, c->el_stream));
    HIPCHK(hipMemcpyAsync(yyy + nr, E.y + nr, sizeof(double) * nw, hipMemcpyDeviceToHost, c->el_stream));
    HIPCHK(hipMemcpyAsync(zzz + nr, E.z + nr, sizeof(double) * nw, hipMemcpyDeviceToHost, c->el_stream));
    HIPCHK(hipMemcpyAsync(dip.data(), E.dip, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->el_stream));
    HIPCHK(hipMemcpyAsync(f.data(), E.fsite, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->el_stream));
    if (chge) HIPCHK(hipMemcpyAsync(chge, E.chge, sizeof(double) * n, hipMemcpyDeviceToHost, c->el_stream));
    HIPCHK(hipStreamSynchronize(c->el_stream));
    for (int i = 0; i < n; ++i) {
        dipx[i] = dip[3 * i];
        dipy[i] = dip[3 * i + 1];
        dipz[i] = dip[3 * i + 2];
        if (fxx) fxx[i] += f[3 * i];
        if (fyy) fyy[i] += f[3 * i + 1];
        if (fzz) fzz[i] += f[3 * i + 2];
    }
    const double* o = c->el.out;
    if (engcpe) *engcpe += o[4];
    if (vircpe) *vircpe += o[5];
    if (stress)
        for (int i = 0; i < 9; ++i) stress[i] += o[10 + i];
    if (niter) *niter = (int)o[8];
    return 0;
}

int mbgpu_get_elec_details(mbgpu_ctx* c, double* scal, double* chge, double* vesp, double* e_real, double* w_real,
                           double* s_real, double* f_site)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (!c->el.ready) return fail(-1, "no completed electrostatics evaluation in this context");
    const ElecBuffers& E = c->el.E;
    const int nw = c->B.nw, n = 4 * nw, nr = 3 * nw;
    if (scal)
        for (int i = 0; i < 20; ++i) scal[i] = c->el.out[i];
    if (chge) HIPCHK(hipMemcpyAsync(chge, E.chge, sizeof(double) * n, hipMemcpyDeviceToHost, c->el_stream));
    if (vesp) HIPCHK(hipMemcpyAsync(vesp, E.vesp, sizeof(double) * n, hipMemcpyDeviceToHost, c->el_stream));
    if (e_real) HIPCHK(hipMemcpyAsync(e_real, E.eout, sizeof(double) * nr, hipMemcpyDeviceToHost, c->el_stream));
    if (w_real) HIPCHK(hipMemcpyAsync(w_real, E.wout, sizeof(double) * 9 * nr, hipMemcpyDeviceToHost, c->el_stream));
    if (s_real) HIPCHK(hipMemcpyAsync(s_real, E.sout, sizeof(double) * 9 * nr, hipMemcpyDeviceToHost, c->el_stream));
    if (f_site) HIPCHK(hipMemcpyAsync(f_site, E.fsite, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->el_stream));
    HIPCHK(hipStreamSynchronize(c->el_stream));
    return 0;
}

int mbgpu_water_terms_dev(mbgpu_ctx* c, int imcon, const double cell[9], const mbgpu_water_params* p,
                          const double* x_dev, const double* y_dev, const double* z_dev, double* fx_dev,
                          double* fy_dev, double* fz_dev, double eng[4], double vir[4], double stress[9])
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (!p) return fail(-1, "null water-term parameters");
    if (!x_dev || !y_dev || !z_dev) return fail(-1, "null position array");
    CellGeom g;
    if (!make_geom(imcon, cell, g))
        return fail(-4, "imcon = %d is not supported by images (reference stops in fixme)", imcon);
    std::string err;
    c->wt.ready = false;
    if (water_configure(c->wt, c->B.nw, g, cell, p, c->stream, err)) return fail(-1, "%s", err.c_str());
    if (water_evaluate(c->wt, x_dev, y_dev, z_dev, fx_dev, fy_dev, fz_dev, c->stream, err))
        return fail(-1, "%s", err.c_str());
    const double* o = c->wt.out;
    if (eng) {
        eng[0] = o[0];
        eng[1] = o[2];
        eng[2] = o[4];
        eng[3] = o[6];
    }
    if (vir) {
        vir[0] = o[1];
        vir[1] = o[3];
        vir[2] = o[5];
        vir[3] = o[7];
    }
    if (stress)
        for (int i = 0; i < 9; ++i) stress[i] += c->wt.stress[i];
    return 0;
}

int mbgpu_water_terms(mbgpu_ctx* c, int imcon, const double cell[9], const mbgpu_water_params* p,
                      const double* xxx, const double* yyy, const double* zzz, double* fxx, double* fyy,
                      double* fzz, double eng[4], double vir[4], double stress[9])
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (!xxx || !yyy || !zzz) return fail(-1, "null position array");
    const int nr = 3 * c->B.nw;
    HIPCHK(hipMemcpyAsync(c->hx, xxx, sizeof(double) * nr, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->hy, yyy, sizeof(double) * nr, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->hz, zzz, sizeof(double) * nr, hipMemcpyHostToDevice, c->stream));
    int rc = mbgpu_water_terms_dev(c, imcon, cell, p, c->hx, c->hy, c->hz, nullptr, nullptr, nullptr, eng, vir,
                                   stress);
    if (rc) return rc;
    if (fxx || fyy || fzz) {
        std::vector<double> f((size_t)3 * nr);
        HIPCHK(hipMemcpyAsync(f.data(), c->wt.B.fw, sizeof(double) * 3 * nr, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < nr; ++i) {
            if (fxx) fxx[i] += f[i];
            if (fyy) fyy[i] += f[nr + i];
            if (fzz) fzz[i] += f[2 * nr + i];
        }
    }
    return 0;
}

int mbgpu_get_water_details(mbgpu_ctx* c, double* e_real, double* w_real, double* f_real)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (!c->wt.ready) return fail(-1, "no completed water-terms evaluation in this context");
    const WaterBuffers& W = c->wt.B;
    const int nr = 3 * c->B.nw;
    if (e_real) HIPCHK(hipMemcpyAsync(e_real, W.eout, sizeof(double) * nr, hipMemcpyDeviceToHost, c->stream));
    if (w_real) HIPCHK(hipMemcpyAsync(w_real, W.wout, sizeof(double) * 9 * nr, hipMemcpyDeviceToHost, c->stream));
    if (f_real) {
        std::vector<double> f((size_t)3 * nr);
        HIPCHK(hipMemcpyAsync(f.data(), W.fw, sizeof(double) * 3 * nr, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < nr; ++i)
            for (int k = 0; k < 3; ++k) f_real[3 * i + k] = f[(size_t)k * nr + i];
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int mbgpu_heat_check(mbgpu_ctx* c, const double* e_kin_dev, double out[12])
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (!out) return fail(-1, "null output");
    const int na = c->B.natms, nr = 3 * c->B.nw;
    for (int k = 0; k < 12; ++k) out[k] = 0.0;
    std::vector<double> buf((size_t)std::max(na, 4 * c->B.nw));
    auto sum_dev = [&](const double* p, int n, double& acc) -> int {
        HIPCHK(hipMemcpyAsync(buf.data(), p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < n; ++i) acc += buf[i];
        return 0;
    };
    if (e_kin_dev && sum_dev(e_kin_dev, nr, out[0])) return -1;
    if (c->mb_ready && sum_dev(c->B.atomic_energy, na, out[1])) return -1;
    if (c->el.ready) {
        if (sum_dev(c->el.E.eout, nr, out[1])) return -1;
        if (sum_dev(c->el.E.e1, 4 * c->B.nw, out[2])) return -1;
        const double* q = c->el.scal + 35;   // xx, xy, xz, yy, yz, zz
        const double s9[9] = {q[0], q[1], q[2], q[1], q[3], q[4], q[2], q[4], q[5]};
        for (int k = 0; k < 9; ++k) out[3 + k] = s9[k];
    }
    if (c->wt.ready && sum_dev(c->wt.B.eout, nr, out[1])) return -1;
    return 0;
}

int mbgpu_verlet_stage(mbgpu_ctx* c, int stage, double dt, const double* mass_dev, double* x_dev, double* y_dev,
                       double* z_dev, double* vx_dev, double* vy_dev, double* vz_dev, double* fx_dev,
                       double* fy_dev, double* fz_dev, double* fsite_dev, double* ekin_dev, double* ekin_total)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (stage != 1 && stage != 2) return fail(-1, "stage must be 1 or 2 (got %d)", stage);
    if (!mass_dev || !x_dev || !y_dev || !z_dev || !vx_dev || !vy_dev || !vz_dev)
        return fail(-1, "null mass, position or velocity array");
    if ((fx_dev != nullptr) != (fy_dev != nullptr) || (fx_dev != nullptr) != (fz_dev != nullptr))
        return fail(-1, "fx, fy, fz must all be given or all be NULL");
    if (!(dt > 0.0)) return fail(-1, "dt must be positive");
    if (c->B.nblk_cap == 0 && ensure_clusters(c, 1, 1)) return -1;
    const int nr = 3 * c->B.nw;
    if ((nr + 255) / 256 > c->B.nblk_cap) return fail(-1, "block-partial capacity");
    int rc = md_launch_stage(stage, nr, dt, mass_dev, x_dev, y_dev, z_dev, vx_dev, vy_dev, vz_dev, fx_dev, fy_dev,
                             fz_dev, fsite_dev, ekin_dev, c->B.blk, c->dout, c->stream);
    if (rc) return fail(rc, "verlet kernels failed");
    double ke = 0.0;
    HIPCHK(hipMemcpyAsync(&ke, c->dout, sizeof ke, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (ekin_total) *ekin_total = ke;
    return 0;
}

int mbgpu_heat_acf_dev(mbgpu_ctx* c, const double* j_dev, long nt, int ntau, double* acf_dev)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (!j_dev || !acf_dev) return fail(-1, "null current or autocorrelation array");
    if (nt < 1 || ntau < 1 || ntau > nt) return fail(-1, "need 1 <= ntau <= nt (got nt %ld, ntau %d)", nt, ntau);
    int rc = md_launch_hcacf(nt, ntau, j_dev, acf_dev, c->stream);
    if (rc) return fail(rc, "autocorrelation kernel failed");
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int mbgpu_heat_acf(mbgpu_ctx* c, const double* j, long nt, int ntau, double* acf)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (!j || !acf) return fail(-1, "null current or autocorrelation array");
    if (nt < 1 || ntau < 1 || ntau > nt) return fail(-1, "need 1 <= ntau <= nt (got nt %ld, ntau %d)", nt, ntau);
    double *jd = nullptr, *ad = nullptr;
    if (mbgpu_alloc(&jd, (size_t)3 * nt) != hipSuccess || mbgpu_alloc(&ad, (size_t)ntau) != hipSuccess) {
        if (jd) (void)hipFree(jd);
        return fail(-1, "cannot allocate %ld bytes for the current series", (long)(sizeof(double) * (3 * nt + ntau)));
    }
    int rc = 0;
    if (hipMemcpyAsync(jd, j, sizeof(double) * 3 * nt, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = fail(-1, "copy of the current series failed");
    if (!rc) rc = mbgpu_heat_acf_dev(c, jd, nt, ntau, ad);
    if (!rc && (hipMemcpyAsync(acf, ad, sizeof(double) * ntau, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
        rc = fail(-1, "copy of the autocorrelation failed");
    (void)hipFree(jd);
    (void)hipFree(ad);
    return rc;
}

int mbgpu_ring_matrices(int nbead, int langevin, double* nmode_to_cart, double* cart_to_nmode,
                        double* tcart_to_tnmode, double* lambda)
{
    EntryGuard gate_;
    if (nbead < 2 || nbead % 2 || nbead > MBGPU_RING_MAX_BEADS)
        return fail(-1, "nbead must be even, 2 <= nbead <= %d (got %d)", MBGPU_RING_MAX_BEADS, nbead);
    if (!nmode_to_cart || !cart_to_nmode || !tcart_to_tnmode || !lambda) return fail(-1, "null output array");
    const int P = nbead;
    const double pi = 3.141592653589793;
    std::vector<double> C((size_t)P * P, 0.0);
    auto at = [&](int n1, int n2) -> double& { return C[(size_t)(n1 - 1) * P + (n2 - 1)]; };
    for (int n1 = 1; n1 <= P; ++n1) at(n1, 1) = 1.0;
    for (int n1 = 1; n1 <= P / 2; ++n1) {
        at(2 * n1, P) = -1.0;
        at(2 * n1 - 1, P) = 1.0;
    }
    for (int n1 = 1; n1 <= P; ++n1) {
        const double t = 2.0 * pi * (double)(n1 - 1) / (double)P;
        for (int n2 = 1; n2 <= P / 2 - 1; ++n2) {
            at(n1, 2 * n2) = std::sqrt(2.0) * std::cos(t * (double)n2);
            at(n1, 2 * n2 + 1) = -std::sqrt(2.0) * std::sin(t * (double)n2);
        }
    }
    double scale = 1.0 / (double)P;
    if (langevin) {
        for (int n1 = 1; n1 <= P; ++n1) {
            const double f1 = std::sqrt(1.0 / P), f2 = std::sqrt(2.0 / P);
            at(n1, 1) = f1;
            at(n1, P / 2 + 1) = (n1 % 2) ? -f1 : f1;
            const double t = 2.0 * pi * (double)n1 / (double)P;
            for (int n2 = 2; n2 <= P / 2; ++n2) {
                at(n1, n2) = f2 * std::cos(t * (double)(n2 - 1));
                at(n1, n2 + P / 2) = f2 * std::sin(t * ((double)n2 + (double)P / 2.0 - 1.0));
            }
        }
        scale = 1.0;
    }
    for (int n1 = 1; n1 <= P; ++n1)
        for (int n2 = 1; n2 <= P; ++n2) {
            nmode_to_cart[(n1 - 1) * P + (n2 - 1)] = at(n1, n2);
            cart_to_nmode[(n2 - 1) * P + (n1 - 1)] = at(n1, n2) * scale;
            tcart_to_tnmode[(n2 - 1) * P + (n1 - 1)] = at(n1, n2);
        }
    if (langevin) {
        for (int n = 0; n < P; ++n) lambda[n] = 1.0;
    } else {
        lambda[0] = 0.0;
        for (int n = 1; n <= P / 2 - 1; ++n) {
            const double t = 2.0 * (1.0 - std::cos(2.0 * pi * (double)n / (double)P));
            lambda[2 * n - 1] = t * (double)P;
            lambda[2 * n] = t * (double)P;
        }
        lambda[P - 1] = 4.0 * (double)P;
    }
    return 0;
}

int mbgpu_ring_transform_dev(mbgpu_ctx* c, int nbead, long ncoord, const double* M, const double* in_dev,
                             double* out_dev)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (nbead < 1 || nbead > MBGPU_RING_MAX_BEADS) return fail(-1, "nbead out of range (got %d)", nbead);
    if (ncoord < 1) return fail(-1, "ncoord must be positive");
    if (!M || !in_dev || !out_dev) return fail(-1, "null matrix or bead array");
    if (in_dev == out_dev) return fail(-1, "the transform cannot work in place");
    const long need = (long)nbead * nbead;
    if (c->ring_m_cap < need) {
        dfree(c->ring_m);
        HIPCHK(dalloc(&c->ring_m, (size_t)need));
        c->ring_m_cap = need;
    }
    HIPCHK(hipMemcpyAsync(c->ring_m, M, sizeof(double) * need, hipMemcpyHostToDevice, c->stream));
    int rc = ring_launch_transform(nbead, ncoord, c->ring_m, in_dev, out_dev, c->stream);
    if (rc) return fail(rc, "ring transform kernel failed");
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int mbgpu_ring_spring_dev(mbgpu_ctx* c, int nbead, int natom, double coeff, double fac3, double force_scale,
                          const double* mass_dev, const double* x_dev, const double* xc_dev, double* f_dev,
                          double* epot_spring, double* epot_deriv)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (nbead < 1 || nbead > MBGPU_RING_MAX_BEADS) return fail(-1, "nbead out of range (got %d)", nbead);
    if (natom < 1) return fail(-1, "natom must be positive");
    if (!mass_dev || !x_dev || !xc_dev || !f_dev) return fail(-1, "null mass, position, centroid or force array");
    if (!(force_scale != 0.0)) return fail(-1, "force_scale must be nonzero");
    const long need = 2L * nbead * ((natom + 255) / 256);
    if (c->ring_blk_cap < need) {
        dfree(c->ring_blk);
        HIPCHK(dalloc(&c->ring_blk, (size_t)need));
        c->ring_blk_cap = need;
    }
    if (!c->ring_out) HIPCHK(dalloc(&c->ring_out, 2));
    int rc = ring_launch_spring(nbead, natom, coeff, fac3, mass_dev, x_dev, xc_dev, f_dev, c->ring_blk, c->ring_out,
                                c->stream);
    if (rc) return fail(rc, "ring spring kernels failed");
    double o[2] = {0.0, 0.0};
    HIPCHK(hipMemcpyAsync(o, c->ring_out, sizeof o, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (epot_spring) *epot_spring = 0.5 * coeff * o[0];
    if (epot_deriv) *epot_deriv = 0.5 * o[1] / force_scale;
    return 0;
}

int mbgpu_ring_spring_dev_noalloc(mbgpu_ctx* c, int nbead, int natom, double coeff, double fac3, double force_scale,
                                  const double* mass_dev, const double* x_dev, const double* xc_dev, double* f_dev,
                                  double* epot_spring, double* epot_deriv)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (nbead < 1 || nbead > MBGPU_RING_MAX_BEADS) return fail(-1, "nbead out of range (got %d)", nbead);
    if (natom < 1) return fail(-1, "natom must be positive");
    if (!mass_dev || !x_dev || !xc_dev || !f_dev) return fail(-1, "null mass, position, centroid or force array");
    if (!(force_scale != 0.0)) return fail(-1, "force_scale must be nonzero");
    const long need = 2L * nbead * ((natom + 255) / 256) + 2;
    if (need > MBGPU_NOALLOC_SCRATCH_DOUBLES)
        return fail(-6, "nbead %d x natom %d needs %ld doubles of scratch; the no-allocation entry holds %d "
                        "(2 per bead per 256 atoms, plus 2)", nbead, natom, need, MBGPU_NOALLOC_SCRATCH_DOUBLES);
    double* blk = noalloc_scratch(c);
    if (!blk) return -1;
    // every block partial and both sums are written by the kernels of this call: nothing in
    // the scratch needs clearing
    double* out = blk + (need - 2);
    int rc = ring_launch_spring(nbead, natom, coeff, fac3, mass_dev, x_dev, xc_dev, f_dev, blk, out, c->stream);
    if (rc) return fail(rc, "ring spring kernels failed");
    double o[2] = {0.0, 0.0};
    HIPCHK(hipMemcpyAsync(o, out, sizeof o, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (epot_spring) *epot_spring = 0.5 * coeff * o[0];
    if (epot_deriv) *epot_deriv = 0.5 * o[1] / force_scale;
    return 0;
}

int mbgpu_spme_recip_dev(mbgpu_ctx* c, const double cell[9], int order, const int grid[3], double alpha, int n,
                         const double* x_dev, const double* y_dev, const double* z_dev, const double* q_dev,
                         const double* mu_dev, double* phi_dev, double* efield_dev, double* f_dev, double* energy)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (n < 0) return fail(-1, "n must not be negative");
    if (n > 0 && (!x_dev || !y_dev || !z_dev || !q_dev)) return fail(-1, "null position or charge array");
    SpmeGeom G{};
    SpmeBuffers& B = c->sp;
    std::string err;
    if (spme_plan(B, G, cell, order, grid, alpha, n, 1, err)) return fail(-1, "%s", err.c_str());
    int rc = spme_launch(B, G, n, x_dev, y_dev, z_dev, q_dev, mu_dev, phi_dev, efield_dev, f_dev, c->el_stream);
    if (rc) return fail(rc, "particle-mesh kernels failed");
    double u = 0.0;
    int flag[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&u, B.out, sizeof u, hipMemcpyDeviceToHost, c->el_stream));
    HIPCHK(hipMemcpyAsync(flag, B.flag, sizeof flag, hipMemcpyDeviceToHost, c->el_stream));
    HIPCHK(hipStreamSynchronize(c->el_stream));
    if (flag[0] == 1) return fail(-1, "a site position is not finite or too far from the cell to index the grid");
    if (flag[0] == 2) return fail(-1, "a charge or dipole is too large for the fixed-point grid (|contribution| >= 1024 e)");
    if (energy) *energy = u;
    return 0;
}

int mbgpu_get_dipoles(mbgpu_ctx* c, double* dip)
{
    EntryGuard gate_;
    if (!c) return fail(-1, "null context");
    if (!c->el.ready) return fail(-1, "no completed electrostatics evaluation in this context");
    if (dip)
        HIPCHK(hipMemcpyAsync(dip, c->el.E.dip, sizeof(double) * 3 * 4 * c->B.nw, hipMemcpyDeviceToHost, c->el_stream));
    HIPCHK(hipStreamSynchronize(c->el_stream));
    return 0;
}

int mbgpu_time_forces(mbgpu_ctx* c, int reps, double* ms_per_eval, double* ms_pip3b)
{
    EntryGuard gate_;
    if (!c || !c->have_inputs) return fail(-1, "mbgpu_time_forces: run mbgpu_mbpol_forces first");
    MbBuffers& B = c->B;
    float total = 0.f, k3 = 0.f;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; ++r) {
        if (build_lists(c)) return -1;
        int rc = mb_launch_dimers(B, B.counts, (int)B.dcap, c->geom, true, c->stream);
        if (!rc)
            rc = mb_launch_trimers(B, B.counts + 1, (int)B.tcap, (int)((B.dcap + 127) / 128), c->geom, true,
                                   c->stream);
        if (!rc) rc = mb_launch_accumulate(B, nullptr, nullptr, nullptr, c->stream);
        if (rc) return fail(rc, "timing run failed");
        double s[16];
        HIPCHK(hipMemcpyAsync(s, B.scal, sizeof s, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&total, c->ev0, c->ev1));
    // the 3-body polynomial kernel alone, same stream
    HIPCHK(hipEventRecord(c->ev2, c->stream));
    for (int r = 0; r < reps; ++r) launch_pip3b_f64(B.xv3, B.part3, B.counts + 1, (int)B.tcap, B.cs3, c->stream);
    HIPCHK(hipEventRecord(c->ev3, c->stream));
    HIPCHK(hipEventSynchronize(c->ev3));
    HIPCHK(hipEventElapsedTime(&k3, c->ev2, c->ev3));
    if (ms_per_eval) *ms_per_eval = total / reps;
    if (ms_pip3b) *ms_pip3b = k3 / reps;
    return 0;
}

int mbgpu_time_elec(mbgpu_ctx* c, int reps, double* ms_sfac, double* flop_sfac)
{
    EntryGuard gate_;
    if (!c || !c->el.ready) return fail(-1, "mbgpu_time_elec: run mbgpu_polar_electrostatics first");
    if (reps < 1) return fail(-1, "reps must be positive");
    ElecState& S = c->el;
    if (S.P.K < 1 || S.recip) return fail(-1, "mbgpu_time_elec: the last electrostatics call had no k-sum");
    int rc = el_time_sfac(S.E, S.P, reps, c->ev0, c->ev1, c->el_stream);
    if (rc) return fail(rc, "timing run failed");
    float ms = 0.f;
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    // polarizable sites: the O and H of every water, plus the M sites if they carry polr2
    const int npol = 2 * S.nw * (S.P.polr2[1] > 1.e-6) + S.nw * (S.P.polr2[0] > 1.e-6) +
                     S.nw * (S.P.polr2[2] > 1.e-6);
    if (ms_sfac) *ms_sfac = ms / reps;
    if (flop_sfac) *flop_sfac = 18.0 * S.P.K * npol;
    return 0;
}

// ------------------------------------------------------------------------- leaves ----
void mbpol_2b_cutoff_(double* r) { *r = MBGPU_R2F; }
void mbpol_3b_cutoff_(double* r) { *r = MBGPU_R3F; }
void mbpol_2b_cutoff(double* r) { *r = MBGPU_R2F; }
void mbpol_3b_cutoff(double* r) { *r = MBGPU_R3F; }

void mbpol_2b_poly_(const double* w1, const double* w2, double* E, double* g1, double* g2, double* force_matrix)
{
    EntryGuard gate_;
    (void)force_matrix;
    const double d[3] = {w1[0] - w2[0], w1[1] - w2[1], w1[2] - w2[2]};
    if (sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) > MBGPU_R2F) {  // leaf early return
        *E = 0.0;
        return;
    }
    const double* w[2] = {w1, w2};
    double g[18];
    if (leaf_eval(2, w, E, g)) {
        fprintf(stderr, "mbpol_2b_poly_: %s\n", g_err.c_str());
        *E = NAN;
        return;
    }
    for (int k = 0; k < 9; ++k) {
        g1[k] = g[k];
        g2[k] = g[9 + k];
    }
}

void mbpol_3b_poly_(const double* w1, const double* w2, const double* w3, double* E, double* g1, double* g2,
                    double* g3, double* force_matrix)
{
    EntryGuard gate_;
    (void)force_matrix;
    const double* w[3] = {w1, w2, w3};
    double g[27];
    if (leaf_eval(3, w, E, g)) {
        fprintf(stderr, "mbpol_3b_poly_: %s\n", g_err.c_str());
        *E = NAN;
        return;
    }
    for (int k = 0; k < 9; ++k) {
        g1[k] += g[k];
        g2[k] += g[9 + k];
        g3[k] += g[18 + k];
    }
}

void mbpol_2b_poly(const double* w1, const double* w2, double* E, double* g1, double* g2, double* fm)
{
    EntryGuard gate_;
    mbpol_2b_poly_(w1, w2, E, g1, g2, fm);
}

void mbpol_3b_poly(const double* w1, const double* w2, const double* w3, double* E, double* g1, double* g2,
                   double* g3, double* fm)
{
    EntryGuard gate_;
    mbpol_3b_poly_(w1, w2, w3, E, g1, g2, g3, fm);
}

}  // extern "C"
