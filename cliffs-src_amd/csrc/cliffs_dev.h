// Device-side data layout and kernel argument blocks.
//
// HBM layout (one allocation per field, FP32):
//   node (i, j) of field F is F[j*pitch + i]; pitch is a multiple of 64
//   floats, so every row starts on a 256-byte boundary and a wave's 64
//   consecutive columns are whole 128-byte lines.  x is contiguous, exactly
//   the reference's Fortran (nx, ny) order.
// Fields: S (background celerity, read-only), CP (celerity perturbation),
// U (x-velocity), V (y-velocity), HM (max CP), VM (max u^2+v^2).
//
// CP, U and V exist twice (double buffer, see DevGrid).
// Line flags ("line has a wet node", cliffs.f:113) are byte arrays, one
// byte per row or column, produced by the previous kernel in the chain:
// the x-sweep produces column flags for the y-sweep, the y-sweep produces
// row flags for the next x-sweep.  Three buffers rotate: kernel m reads
// F[m%3], writes F[(m+1)%3] (only ever storing 1s) and clears F[(m+2)%3].
#pragma once

#include "vtcs_node.h"

namespace cliffs {

// Records msg as cliffs_last_error() and returns code (capi.cpp).
int set_error(int code, const char* msg);

constexpr int kXStrip = 248;    // output nodes per x-sweep wave (window 256)
constexpr int kXRows = 8;       // rows per x-sweep wave
constexpr int kYRows = 128;     // output rows per y-sweep block

// The state CP, U, V is double-buffered: a sweep reads the current copy and
// writes the other one (neighbouring work items read each other's halo
// lines, so updating in place would race); lines a sweep leaves untouched
// are copied.  CPn/Un/Vn point at the copy being written.
struct DevGrid {
    int nx, ny, pitch;
    float* S;
    float* CP;
    float* U;
    float* V;
    float* CPn;
    float* Un;
    float* Vn;
    float* HM;
    float* VM;
};

// Result of health_kernel (cliffs_health).
struct HealthOut {
    unsigned long long nonfinite;   // cells with a non-finite cp, u or v
    unsigned cmax_bits;             // largest Courant number (float bits)
    unsigned pad;
};

struct XArgs {
    DevGrid g;
    const float4* xm;     // per node i: (1/b(i-1), 1/b(i), 1/(b(i-1)+b(i)), b(i)-b(i-1))
    const float2* xr;     // per row j: (1/a(j), a(j));  dx(i,j) = a(j) b(i)
    const float* edges;   // [6][ny]: west ub, vb, g*eta; east ub, vb, g*eta (or null)
    const unsigned char* fin;   // row flags
    unsigned char* fout;        // produced flags (columns, or rows in 1-D)
    unsigned* fclr;             // words to clear
    int nclr;
    Params p;
    int nstrips;
    int row_lo, row_hi;   // rows visited (0-based, inclusive)
    int act_lo, act_hi;   // rows swept
    int flag_rows;        // produce row flags (1-D) instead of column flags
    int fuse_max;         // fuse the max-wave update (1-D step)
};

struct YArgs {
    DevGrid g;
    const float4* ym;     // per row j: (1/s2(j-1), 1/s2(j), 1/(s2(j-1)+s2(j)), s2(j)-s2(j-1))
    const float* zeta;    // per row j
    const float* edges;   // [6][nx]: south ub, vb, g*eta; north ub, vb, g*eta (or null)
    const unsigned char* fin;   // column flags
    unsigned char* fout;        // produced flags (rows, or columns in 1-D)
    unsigned* fclr;
    int nclr;
    Params p;
    int col_base;         // first column of the launch grid
    int col_span;         // columns covered by the launch grid
    int col_lo, col_hi;   // columns swept (0-based, inclusive)
    int flag_cols;        // produce column flags (1-D) instead of row flags
    int fuse_max;         // fuse the max-wave update (2-D step)
};

// Fused 2-D step (step_fused.hip): tiles of kFStrip columns x `rows` rows.
constexpr int kFStrip = 248;    // output columns per tile (x window of 256)
constexpr int kFStripNarrow = 120;  // narrow tiles (x window of 128, 2 nodes per lane)
// Strip width of a tile whose x window holds N nodes per lane (64 lanes)
constexpr int strip_of(int n) { return 64 * n - 8; }
constexpr int kFMaxRows = 256;  // rows per tile, at most (the tile's row metrics sit in LDS)
constexpr int kXCDs = 8;        // MI355X: 8 accelerator complex dies, one L2 each
constexpr int kFRoleThreads = 256;  // threads of the x role of a fused-step block
// threads of a fused-step block with N nodes per lane: the x role's four
// waves plus one y thread per window node
constexpr int fused_threads(int n) { return kFRoleThreads + 64 * n; }

// Line flags of the fused step are per-tile contributions (one byte per tile
// and line, written 0 or 1 by the tile that owns it), so a tile recomputed by
// the fixup kernel overwrites its own contribution; a line's flag is the OR
// over the tiles that cross it.
struct FArgs {
    DevGrid g;                    // g.S is the base of all nine fields (one allocation)
    unsigned fs;                  // floats from one field to the next (9 * fs * 4 < 2^32)
    int cur;                      // current state copy (0 or 1)
    const float4* xm;     // x-sweep metric (XArgs::xm), xr, edges [6][ny] or null
    const float2* xr;
    const float* xedges;
    const float4* ym;     // y-sweep metric (YArgs::ym), zeta, edges [6][nx] or null
    const float* zeta;
    const float* yedges;
    Params p;
    int strip_w;                  // tile width: kFStrip or kFStripNarrow
    int nstrips, nchunks, rows, ntiles;
    int cstride;                  // bytes per column of cpr: roundup(nchunks, 16)
    const unsigned char* rfl_in;  // [nstrips][ny]: row j wet within strip s, state entering the step
    unsigned char* rfl_out;       // the same for the state leaving the step
    const unsigned char* cpr_in;  // [nx][cstride]: column i wet within chunk c, entering the step
    unsigned char* cpr_out;       // the same, leaving the step
    unsigned char* ctr;           // [nx]: column i wet after the x-sweep (only 1s stored)
    unsigned char* ctr_next;      // [nx]: the next step's ctr, cleared by the fixup kernel
    unsigned char* mpend;         // [ntiles]: a deferred max-wave update of the tile changes a maximum
    int row_lo, row_hi;           // rows the x-sweep covers (0-based, inclusive)
    int col_lo, col_hi;           // columns the y-sweep covers
    int maxwave;                  // update hmax/velmax (0 only for timing trials)
};

// One-way nest feed (nest.hip)
struct NestArgs {
    DevGrid g;
    Params p;
    double grav;
    int nxc, nyc;
    const int* we_ny;      // 1-based parent rows, per child row
    const double* we_cy;
    const int* sn_nx;      // 1-based parent columns, per child column
    const double* sn_cx;
    int w_nx, e_nx, s_ny, n_ny;
    double w_cx, e_cx, s_cy, n_cy;
    int ii1, ii2, kk1, kk2;  // footprint (1-based, inclusive)
    float gthr;              // g * 0.1f * cuke
    float cuke;
    unsigned* flag;
    double* uvq;             // west (nyc,3), east (nyc,3), south (nxc,3), north (nxc,3)
};

}  // namespace cliffs
