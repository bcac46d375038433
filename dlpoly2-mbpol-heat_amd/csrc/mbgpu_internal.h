// mbgpu_internal.h -- host-side declarations shared by the translation units of libmbgpu.
#pragma once
#include <hip/hip_runtime.h>

#include "mbgpu_dev.h"

// Launch wrappers for the generated polynomial kernels (pip2b.hip / pip3b.hip).
// xv: variables SoA [nvars][nc]; part: SoA [nslices][nvars+1][nc]; cs: coefficient stream.
void launch_pip2b_f64(const double* xv, double* part, int nc, const double* cs, hipStream_t s);
void launch_pip3b_f64(const double* xv, double* part, int nc, const double* cs, hipStream_t s);
int pip2b_nslices();
int pip3b_nslices();
int pip2b_stream_len();
int pip3b_stream_len();
const short* pip2b_stream_table();
const short* pip3b_stream_table();
const double* pip2b_stream_values();
const double* pip3b_stream_values();

// Device buffers of one MB-pol evaluation context.
struct MbBuffers {
    int nw = 0, natms = 0;
    int cap2 = 0, cap3 = 0;        // neighbour capacities per water (6.5 A and 4.5 A lists)
    long dcap = 0, tcap = 0;       // cluster capacities
    // geometry
    double* mon = nullptr;         // [nw][9] imaged monomers
    int* list = nullptr;           // [3*nw] listttm2 (1-based atom ids)
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
};
