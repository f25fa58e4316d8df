// C-ABI of the cliffs_gpu library (include/cliffs_gpu.h): device state,
// host <-> device conversion and the launch sequence of a time step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cliffs_gpu.h"
#include "cliffs_dev.h"
#include "convert.h"

namespace cliffs {
hipError_t launch_sweep_x(const XArgs& a, hipStream_t st);
hipError_t launch_sweep_y(const YArgs& a, hipStream_t st);
hipError_t launch_line_flags(const DevGrid& g, const Params& p, unsigned char* fout, int dir,
                             unsigned* fclr, int nclr, hipStream_t st);
hipError_t launch_maxwave(const DevGrid& g, hipStream_t st);
hipError_t launch_copy_lines(const DevGrid& g, int dir, int first, int last, hipStream_t st);
hipError_t gauge_scratch(int dev, void** ptr);
size_t gauge_scratch_words();
hipError_t launch_gauge_sample(const DevGrid& g, const Params& p, const int* I, const int* J, int n,
                               float* series, int slot, hipStream_t st);
hipError_t launch_reset_max(const DevGrid& g, hipStream_t st);
hipError_t launch_health(const DevGrid& g, const Params& p, const float4* xm, const float2* xr,
                         const float4* ym, int xrun, int yrun, HealthOut* out, hipStream_t st);
hipError_t launch_nest_arrival(const NestArgs& a, hipStream_t st);
hipError_t launch_nest_feed(const NestArgs& a, hipStream_t st);
hipError_t launch_child_edges(const double* r0, const double* r1, double cc, double omc, double grav,
                              int nxc, int nyc, int zero, float* xe, float* ye, hipStream_t st);
hipError_t launch_record_wave(const double* rec, int nxc, int nyc, double* wave, hipStream_t st);
hipError_t launch_step_fused(const FArgs& a, hipStream_t st);
hipError_t launch_step_fixup(const FArgs& a, hipStream_t st);
hipError_t launch_fused_flags(const FArgs& a, unsigned char* ctr0, unsigned char* ctr1,
                              hipStream_t st);
hipError_t fused_blocks_per_cu(int strip_w, int* n);
hipError_t launch_step_batch(const FArgs* args, const int2* bmap, int nblocks, int strip_w, hipStream_t st);
hipError_t launch_fixup_batch(const FArgs* args, const int2* bmap, int nblocks, int strip_w, hipStream_t st);
hipError_t launch_ensemble_gauges(const FArgs* args, int nm, const int* ij, int n, int nsamples, int slot,
                                  float* series, hipStream_t st);
void ensemble_block_map(const int* ntiles, int n, std::vector<int2>& bmap);
void fused_geometry(int nx, int ny, int strip_w, int cus, int blocks_per_cu, int* nstrips, int* nchunks,
                    int* rows);
}  // namespace cliffs

using namespace cliffs;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
int hipfail(hipError_t e, const char* where) {
    return fail(CLIFFS_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}
#define HIPCHK(call)                                  \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hipfail(e_, #call); \
    } while (0)

constexpr int kEdgeSlots = 16;

// Left-next lattice index by bisection and its interpolation weight
// (NestedGrid.f:129-148): 1-based ip with xx(ip) < xp <= xx(ip+1), clamped
// to [1, nx-1]; cp = (xx(ip+1)-xp)/(xx(ip+1)-xx(ip)), 0.5 for nx = 1.  The
// bisection order is kept so that maps equal the reference's bit for bit.
void pinpoint(const double* xx, int nx, double xp, int* ip_out, double* cp) {
    int ip = 1, ipp = nx;
    while (ip + 2 <= ipp) {
        const int med = (ip + ipp) / 2;
        if (xp > xx[med - 1]) ip = med;
        else ipp = med;
    }
    *cp = nx > 1 ? (xx[ip] - xp) / (xx[ip] - xx[ip - 1]) : 0.5;
    *ip_out = ip;
}

}  // namespace

// The other modules of the library report through cliffs_last_error too.
int cliffs::set_error(int code, const char* msg) { return fail(code, msg); }

struct cliffs_handle {
    int dev = 0;
    hipStream_t st = nullptr;
    int nx = 0, ny = 0, pitch = 0;
    Consts64 k{};
    Params p{};
    std::vector<double> dep;   // FP64 depth (host copy, for conversions)
    std::vector<float> s;      // device background celerity (host copy)
    // line ranges (cliffs_main.f:140-164), 0-based inclusive
    bool xrun = false, yrun = false;
    int ii1 = 0, ii2 = 0, jj1 = 0, jj2 = 0, mx = 1;
    // device memory
    float* fields = nullptr;   // S, CP0, U0, V0, CP1, U1, V1, HM, VM
    int cur = 0;               // which of the two state copies is current
    size_t fstride = 0;        // floats per field
    float4* xm = nullptr;
    float2* xr = nullptr;
    float4* ym = nullptr;
    float* zeta = nullptr;
    HealthOut* health = nullptr;  // cliffs_health result (allocated on first use)
    float* edges = nullptr;    // kEdgeSlots x (6*nx + 6*ny) floats
    float* hedges = nullptr;   // pinned staging, same size
    hipEvent_t edge_ev[kEdgeSlots] = {};
    int edge_slot = 0;
    size_t edge_slot_floats = 0;
    unsigned char* flags = nullptr;  // 3 buffers of fbytes
    size_t fbytes = 0;
    int m = 0;                 // kernel sequence number
    bool flags_valid = false;
    // fused 2-D step (step_fused.hip)
    bool fused = true;         // 2-D steps run as one fused kernel (+ fixup)
    bool fused_ok = false;     // the nine fields span < 4 GiB (32-bit offsets)
    int f_width = kFStrip;     // tile width (kFStrip or kFStripNarrow)
    int f_nstrips = 0, f_nchunks = 0, f_rows = 0;
    int f_cus = 0, f_bpc = 0, f_bpc_narrow = 0;  // CUs of the device, fused-kernel blocks per CU
    int f_cstride = 0;         // bytes per column of cpr: roundup(nchunks, 16)
    unsigned char* fflags = nullptr;  // rfl[2][nstrips][ny], cpr[2][nx][cstride], ctr[2][nx], mpend[ntiles]
    int fn = 0;                // fused step parity
    bool fused_valid = false;
    bool state = false;
    std::vector<float> hbuf;   // pitched host staging
    // nests (as parent or coupled child) and ensembles that use this grid's
    // device memory and stream; the grid cannot be destroyed while any is live
    int users = 0;
    int ens_users = 0;         // ensembles among them (they fix the tile geometry)

    DevGrid grid() const {
        DevGrid g;
        g.nx = nx;
        g.ny = ny;
        g.pitch = pitch;
        g.S = fields;
        float* a = fields + (1 + 3 * (size_t)cur) * fstride;      // current copy
        float* b = fields + (1 + 3 * (size_t)(1 - cur)) * fstride;  // copy to write
        g.CP = a;
        g.U = a + fstride;
        g.V = a + 2 * fstride;
        g.CPn = b;
        g.Un = b + fstride;
        g.Vn = b + 2 * fstride;
        g.HM = fields + 7 * fstride;
        g.VM = fields + 8 * fstride;
        return g;
    }
    unsigned char* fbuf(int i) { return flags + (size_t)(((i % 3) + 3) % 3) * fbytes; }
    unsigned char* rfl(int b) { return fflags + (size_t)b * f_nstrips * ny; }
    size_t rfl_bytes() const { return ((size_t)f_nstrips * ny + 255) / 256 * 256; }
    size_t cpr_bytes() const { return (size_t)f_cstride * nx; }
    unsigned char* cpr(int b) { return fflags + 2 * rfl_bytes() + (size_t)b * cpr_bytes(); }
    unsigned char* ctr(int b) {
        return fflags + 2 * rfl_bytes() + 2 * cpr_bytes() + (size_t)b * nx;
    }
    unsigned char* mpend() { return ctr(2); }
};

struct cliffs_nest {
    cliffs_handle* parent = nullptr;
    int nxc = 0, nyc = 0;
    int* we_ny = nullptr;   // device, 1-based parent rows
    int* sn_nx = nullptr;   // device, 1-based parent columns
    double* we_cy = nullptr;
    double* sn_cx = nullptr;
    int w_nx = 0, e_nx = 0, s_ny = 0, n_ny = 0;
    double w_cx = 0, e_cx = 0, s_cy = 0, n_cy = 0;
    int ii1 = 0, ii2 = 0, kk1 = 0, kk2 = 0;
    unsigned* flag = nullptr;
    double* uvq = nullptr;
    size_t nuvq = 0;
    std::vector<double> huvq;
    // device coupling to a child grid (cliffs_nest_couple)
    cliffs_handle* child = nullptr;
    int cap = 0;                       // feed records kept on the device (ring)
    double* ring = nullptr;            // cap records of nuvq doubles
    std::vector<hipEvent_t> rec_ev;    // record written (parent stream)
    std::vector<hipEvent_t> use_ev;    // record last read (child stream)
    std::vector<double> tinput;        // record times, as get_boundary_feed reads them
    int irec = 1;                      // child's next record (1-based, cliffs_main.f:305-315)
    bool bndr = true;                  // bndr_input: feed still running for the child
    bool ended = false;                // no more records will come
    float* cedges = nullptr;           // child edges: x [6][nyc], y [6][nxc]
    // child start rule (get_boundary_feed.f:144-167): the first record whose
    // largest edge |eta| reaches the child's threshold
    double* dwave = nullptr;           // per ring slot: largest |eta| of its record
    double start_cuke = 0.0;           // child's threshold (0: the first record)
    bool start_set = false;            // threshold given by cliffs_nest_child_start
    int start = 0;                     // 1-based start record, 0 while unknown
    int scanned = 0;                   // records found below the threshold
    int waited = 0;                    // record the child stream last waited for
    long run_k = 0;                    // cliffs_nest_run: record intervals run so far
    long run_kc = 0;                   // and child steps taken
    bool counted = false;              // the parent's use count includes this nest
    NestArgs args(double cuke) const;
};

namespace {

void free_handle(cliffs_handle* h) {
    if (!h) return;
    if (h->dev >= 0) (void)hipSetDevice(h->dev);
    if (h->st) (void)hipStreamSynchronize(h->st);
    for (auto& e : h->edge_ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(h->fields);
    (void)hipFree(h->xm);
    (void)hipFree(h->xr);
    (void)hipFree(h->ym);
    (void)hipFree(h->zeta);
    (void)hipFree(h->edges);
    (void)hipFree(h->flags);
    (void)hipFree(h->fflags);
    (void)hipFree(h->health);
    if (h->hedges) (void)hipHostFree(h->hedges);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}

// pitched FP32 host buffer <-> device field: host and device rows have the
// same pitch, so the field is one contiguous block (a 2-D copy from pageable
// memory would be staged row by row)
int put_field(cliffs_handle* h, float* dst, const std::vector<float>& src) {
    HIPCHK(hipMemcpyAsync(dst, src.data(), (size_t)h->pitch * h->ny * sizeof(float),
                          hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return CLIFFS_OK;
}
int get_field(cliffs_handle* h, const float* src, std::vector<float>& dst) {
    dst.resize((size_t)h->pitch * h->ny);
    HIPCHK(hipMemcpyAsync(dst.data(), src, (size_t)h->pitch * h->ny * sizeof(float),
                          hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return CLIFFS_OK;
}
inline size_t hidx(const cliffs_handle* h, int i, int j) {
    return (size_t)j * h->pitch + i;
}

// Host threads for the per-cell conversions of uploads and downloads: up to
// 16 (fewer if OMP_NUM_THREADS says so).
int host_threads() {
    static const int n = [] {
        int t = (int)std::thread::hardware_concurrency();
        if (t < 1) t = 1;
        if (const char* e = std::getenv("OMP_NUM_THREADS")) {
            const int v = std::atoi(e);
            if (v >= 1) t = std::min(t, v);
        }
        return std::min(t, 16);
    }();
    return n;
}

// f(j0, j1) over the rows [0, ny) of a grid of `cells` cells.  Cells convert
// independently of each other, so the rows are split over host threads;
// grids under 2^20 cells stay on the calling thread.
template <class F>
void for_rows(int ny, size_t cells, const F& f) {
    const int nt = std::min(host_threads(), ny);
    if (cells < ((size_t)1 << 20) || nt <= 1) {
        f(0, ny);
        return;
    }
    std::vector<std::thread> th;
    th.reserve((size_t)nt - 1);
    for (int t = 1; t < nt; t++) {
        const int j0 = (int)((long)ny * t / nt), j1 = (int)((long)ny * (t + 1) / nt);
        try {
            th.emplace_back([&f, j0, j1] { f(j0, j1); });
        } catch (...) {
            f(j0, j1);  // no thread available: convert these rows here
        }
    }
    f(0, (int)((long)ny / nt));
    for (auto& x : th) x.join();
}

// Edge arrays (mx,3) FP64 -> device [6][n] (ub, vb, g*eta for edge1, then
// edge2) for a sweep of direction iu over lines of count n.
int stage_edges(cliffs_handle* h, int iu, const double* e1, const double* e2, int mx,
                const float** dev_out) {
    *dev_out = nullptr;
    if (!e1 && !e2) return CLIFFS_OK;
    const int n = iu == 1 ? h->ny : h->nx;
    if (mx < n) return fail(CLIFFS_ERR_ARG, "edge leading dimension mx smaller than line count");
    const int slot = h->edge_slot;
    h->edge_slot = (h->edge_slot + 1) % kEdgeSlots;
    HIPCHK(hipEventSynchronize(h->edge_ev[slot]));
    float* hs = h->hedges + (size_t)slot * h->edge_slot_floats;
    const int cn = iu == 1 ? 0 : 1;  // column of the normal velocity in edge(:,1:2)
    const int ct = 1 - cn;
    for (int side = 0; side < 2; side++) {
        const double* e = side == 0 ? e1 : e2;
        float* o = hs + (size_t)side * 3 * n;
        for (int l = 0; l < n; l++) {
            o[l] = e ? (float)e[l + (size_t)mx * cn] : 0.0f;
            o[n + l] = e ? (float)e[l + (size_t)mx * ct] : 0.0f;
            o[2 * n + l] = e ? (float)(h->k.grav * e[l + (size_t)mx * 2]) : 0.0f;
        }
    }
    float* dv = h->edges + (size_t)slot * h->edge_slot_floats;
    HIPCHK(hipMemcpyAsync(dv, hs, sizeof(float) * 6 * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipEventRecord(h->edge_ev[slot], h->st));
    *dev_out = dv;
    return CLIFFS_OK;
}

int clear_flag_words(cliffs_handle* h) { return (int)(h->fbytes / 4); }

// Line flags of direction dir (0 rows, 1 columns) from the current state.
int launch_flags(cliffs_handle* h, int dir) {
    const hipError_t e = launch_line_flags(h->grid(), h->p, h->fbuf(h->m + 1), dir,
                                           reinterpret_cast<unsigned*>(h->fbuf(h->m + 2)),
                                           clear_flag_words(h), h->st);
    if (e != hipSuccess) return hipfail(e, "line_flags");
    h->m++;
    return CLIFFS_OK;
}

int ensure_flags(cliffs_handle* h, int dir) {
    if (h->flags_valid) return CLIFFS_OK;
    HIPCHK(hipMemsetAsync(h->flags, 0, 3 * h->fbytes, h->st));
    h->m = 0;
    const int rc = launch_flags(h, dir);
    if (rc) return rc;
    h->flags_valid = true;
    return CLIFFS_OK;
}

XArgs xargs(cliffs_handle* h) {
    XArgs a{};
    a.g = h->grid();
    a.xm = h->xm;
    a.xr = h->xr;
    a.p = h->p;
    a.nstrips = (h->nx + kXStrip - 1) / kXStrip;
    return a;
}
YArgs yargs(cliffs_handle* h) {
    YArgs a{};
    a.g = h->grid();
    a.ym = h->ym;
    a.zeta = h->zeta;
    a.p = h->p;
    return a;
}

// one chained x-sweep kernel: reads F[m], writes F[m+1], clears F[m+2]
int chained_x(cliffs_handle* h, const float* edges, bool one_d) {
    XArgs a = xargs(h);
    a.edges = edges;
    a.fin = h->fbuf(h->m);
    a.fout = h->fbuf(h->m + 1);
    a.fclr = reinterpret_cast<unsigned*>(h->fbuf(h->m + 2));
    a.nclr = clear_flag_words(h);
    if (one_d) {
        a.row_lo = 0;
        a.row_hi = h->ny - 1;
        a.act_lo = a.act_hi = 0;
        a.flag_rows = 1;
        a.fuse_max = 1;
    } else {
        a.row_lo = 0;
        a.row_hi = h->ny - 1;
        a.act_lo = h->jj1;
        a.act_hi = h->jj2;
        a.flag_rows = 0;
        a.fuse_max = 0;
    }
    const hipError_t e = launch_sweep_x(a, h->st);
    if (e != hipSuccess) return hipfail(e, "sweep_x");
    h->m++;
    h->cur ^= 1;
    return CLIFFS_OK;
}

int chained_y(cliffs_handle* h, const float* edges, bool one_d) {
    YArgs a = yargs(h);
    a.edges = edges;
    a.fin = h->fbuf(h->m);
    a.fout = h->fbuf(h->m + 1);
    a.fclr = reinterpret_cast<unsigned*>(h->fbuf(h->m + 2));
    a.nclr = clear_flag_words(h);
    a.col_base = 0;
    if (one_d) {
        a.col_lo = a.col_hi = 0;
        a.flag_cols = 1;
    } else {
        a.col_lo = h->ii1;
        a.col_hi = h->ii2;
        a.flag_cols = 0;
    }
    a.fuse_max = 1;
    a.col_span = h->nx;  // every column: fused max-wave and flags
    const hipError_t e = launch_sweep_y(a, h->st);
    if (e != hipSuccess) return hipfail(e, "sweep_y");
    h->m++;
    h->cur ^= 1;
    return CLIFFS_OK;
}

FArgs fargs(cliffs_handle* h, const float* xe, const float* ye) {
    FArgs a{};
    a.g = h->grid();
    a.fs = (unsigned)h->fstride;
    a.cur = h->cur;
    a.xm = h->xm;
    a.xr = h->xr;
    a.xedges = xe;
    a.ym = h->ym;
    a.zeta = h->zeta;
    a.yedges = ye;
    a.p = h->p;
    a.strip_w = h->f_width;
    a.nstrips = h->f_nstrips;
    a.nchunks = h->f_nchunks;
    a.rows = h->f_rows;
    a.ntiles = h->f_nstrips * h->f_nchunks;
    a.cstride = h->f_cstride;
    const int b = h->fn & 1;
    a.rfl_in = h->rfl(b);
    a.rfl_out = h->rfl(b ^ 1);
    a.cpr_in = h->cpr(b);
    a.cpr_out = h->cpr(b ^ 1);
    a.ctr = h->ctr(b);
    a.ctr_next = h->ctr(b ^ 1);
    a.mpend = h->mpend();
    a.row_lo = h->jj1;
    a.row_hi = h->jj2;
    a.col_lo = h->ii1;
    a.col_hi = h->ii2;
    a.maxwave = 1;
    return a;
}

// Line flags of the current state for the next fused step.
int ensure_fused(cliffs_handle* h) {
    if (!h->fflags) return fail(CLIFFS_ERR_STATE, "no fused-step flag buffers (a tile-height change failed)");
    if (h->fused_valid) return CLIFFS_OK;
    FArgs a = fargs(h, nullptr, nullptr);
    a.rfl_out = h->rfl(h->fn & 1);
    a.cpr_out = h->cpr(h->fn & 1);
    const hipError_t e = launch_fused_flags(a, h->ctr(0), h->ctr(1), h->st);
    if (e != hipSuccess) return hipfail(e, "fused_flags");
    h->fused_valid = true;
    return CLIFFS_OK;
}

// Tile geometry of the fused step: tile width (kFStrip or kFStripNarrow),
// rows per tile (0: enough tiles to fill the device once, for a grid stepped
// alone) and the per-tile flag buffers.
int set_fused_shape(cliffs_handle* h, int width, int rows) {
    int ns, nc, r;
    if (rows <= 0) {
        fused_geometry(h->nx, h->ny, width, h->f_cus, width == kFStrip ? h->f_bpc : h->f_bpc_narrow,
                       &ns, &nc, &r);
    } else {
        ns = (h->nx + width - 1) / width;
        r = std::min(std::min(rows, kFMaxRows), h->ny);
        nc = (h->ny + r - 1) / r;
        // the same number of chunks, of equal height: no short last chunk
        const int rb = (h->ny + nc - 1) / nc;
        if (rb >= 8) r = rb;
    }
    if (h->fflags && width == h->f_width && ns == h->f_nstrips && nc == h->f_nchunks && r == h->f_rows)
        return CLIFFS_OK;
    if (h->fflags) {
        HIPCHK(hipStreamSynchronize(h->st));
        HIPCHK(hipFree(h->fflags));
        h->fflags = nullptr;
    }
    h->f_width = width;
    h->f_nstrips = ns;
    h->f_nchunks = nc;
    h->f_rows = r;
    h->f_cstride = (nc + 15) / 16 * 16;
    const size_t fb = 2 * (h->rfl_bytes() + h->cpr_bytes() + (size_t)h->nx) + (size_t)ns * nc;
    // cpr bytes beyond nchunks are read as zeros and never written
    h->fused_valid = false;
    if (hipMalloc(&h->fflags, fb) != hipSuccess) {
        h->fflags = nullptr;  // ensure_fused refuses to step until a geometry is set again
        return fail(CLIFFS_ERR_NOMEM, "flag allocation");
    }
    // on the grid's own (non-blocking) stream: a null-stream clear would not
    // be ordered before the flag kernels that follow it there
    HIPCHK(hipMemsetAsync(h->fflags, 0, fb, h->st));
    return CLIFFS_OK;
}

int set_fused_rows(cliffs_handle* h, int rows) { return set_fused_shape(h, h->f_width, rows); }

// Optional events bracketing the kernels of a step: ev[0], ev[1] around the
// x-sweep (or the fused step kernel), ev[2], ev[3] around the y-sweep (or
// the fixup kernel).
int record(hipEvent_t* ev, int k, hipStream_t st) {
    if (ev) HIPCHK(hipEventRecord(ev[k], st));
    return CLIFFS_OK;
}

int do_step(cliffs_handle* h, const float* xe, const float* ye, hipEvent_t* ev = nullptr) {
    int rc;
    if (h->xrun && h->yrun && h->fused && h->fused_ok) {
        if ((rc = ensure_fused(h))) return rc;
        const FArgs a = fargs(h, xe, ye);
        if ((rc = record(ev, 0, h->st))) return rc;
        hipError_t e = launch_step_fused(a, h->st);
        if (e != hipSuccess) return hipfail(e, "step_fused");
        if ((rc = record(ev, 1, h->st)) || (rc = record(ev, 2, h->st))) return rc;
        e = launch_step_fixup(a, h->st);
        if (e != hipSuccess) return hipfail(e, "step_fixup");
        if ((rc = record(ev, 3, h->st))) return rc;
        h->cur ^= 1;
        h->fn++;
        h->flags_valid = false;
        return CLIFFS_OK;
    }
    h->fused_valid = false;
    if (h->xrun && h->yrun) {
        if ((rc = ensure_flags(h, 0))) return rc;
        if ((rc = record(ev, 0, h->st)) || (rc = chained_x(h, xe, false)) ||
            (rc = record(ev, 1, h->st)))
            return rc;
        if ((rc = record(ev, 2, h->st)) || (rc = chained_y(h, ye, false)) ||
            (rc = record(ev, 3, h->st)))
            return rc;
    } else if (h->xrun) {
        if ((rc = ensure_flags(h, 0))) return rc;
        if ((rc = record(ev, 0, h->st)) || (rc = chained_x(h, xe, true)) ||
            (rc = record(ev, 1, h->st)))
            return rc;
    } else if (h->yrun) {
        if ((rc = ensure_flags(h, 1))) return rc;
        if ((rc = record(ev, 2, h->st)) || (rc = chained_y(h, ye, true)) ||
            (rc = record(ev, 3, h->st)))
            return rc;
    } else {
        const hipError_t e = launch_maxwave(h->grid(), h->st);
        if (e != hipSuccess) return hipfail(e, "maxwave");
    }
    return CLIFFS_OK;
}

}  // namespace

extern "C" {

const char* cliffs_last_error(void) { return g_err.c_str(); }

int cliffs_create(const cliffs_grid_desc* d, cliffs_handle** out) {
    if (!d || !out) return fail(CLIFFS_ERR_ARG, "NULL descriptor or output pointer");
    *out = nullptr;
    if (d->nx < 1 || d->ny < 1 || !d->dep)
        return fail(CLIFFS_ERR_ARG, "grid must be at least 1x1 with a depth array");
    if (!(d->ground > 0.0)) return fail(CLIFFS_ERR_ARG, "ground must be > 0 (c_min = sqrt(g ground))");
    if (!(d->dt > 0.0)) return fail(CLIFFS_ERR_ARG, "dt must be > 0");
    if (d->nx > 2 && !d->s1) return fail(CLIFFS_ERR_ARG, "s1 required when nx > 2");
    if (d->ny > 2 && (!d->s2 || !d->zeta)) return fail(CLIFFS_ERR_ARG, "s2 and zeta required when ny > 2");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(CLIFFS_ERR_NODEVICE, "no HIP device");
    if (d->device < 0 || d->device >= ndev) return fail(CLIFFS_ERR_NODEVICE, "device index out of range");

    cliffs_handle* h = new (std::nothrow) cliffs_handle;
    if (!h) return fail(CLIFFS_ERR_NOMEM, "host allocation");
    const int nx = d->nx, ny = d->ny;
    h->dev = d->device;
    h->nx = nx;
    h->ny = ny;
    // rows are whole 256-byte segments (the x-sweep's float4 accesses stay
    // inside the row); strides that are multiples of 2 KiB get one more
    // segment so consecutive rows do not fall on the same HBM channels
    h->pitch = (nx + 63) / 64 * 64;
    if (h->pitch % 512 == 0) h->pitch += 64;
    h->k = make_consts(d->ground);
    h->p.dt = (float)d->dt;
    h->p.grav = 9.8f;
    h->p.celmin = h->k.celmin32;
    h->p.gground = (float)(h->k.grav * d->ground);
    h->p.gn2 = (float)(h->k.grav * d->crough);
    h->p.sqrt2 = 1.41421356f;
    h->p.friction = d->crough > 0.0 ? 1 : 0;
    // line ranges (cliffs_main.f:140-164), 0-based
    h->xrun = nx > 2;
    h->yrun = ny > 2;
    h->jj1 = h->yrun ? 1 : 0;
    h->jj2 = h->yrun ? ny - 2 : 0;
    h->ii1 = h->xrun ? 1 : 0;
    h->ii2 = h->xrun ? nx - 2 : 0;
    h->mx = std::max(h->xrun ? ny : 1, h->yrun ? nx : 1);

    auto bail = [&](int rc) {
        free_handle(h);
        return rc;
    };
    if (hipSetDevice(h->dev) != hipSuccess) return bail(fail(CLIFFS_ERR_HIP, "hipSetDevice"));
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(CLIFFS_ERR_HIP, "hipStreamCreate"));

    const size_t m = (size_t)nx * ny;
    h->dep.resize(m);
    h->s.resize(m);
    for_rows(ny, m, [&](int j0, int j1) {
        for (size_t c = (size_t)j0 * nx; c < (size_t)j1 * nx; c++) {
            h->dep[c] = d->dep[c];
            h->s[c] = to_s(d->dep[c], h->k);
        }
    });

    // ---- metrics.  x: s1(i,j) = a(j) b(i)  (MasterGrid.f:86-99)
    std::vector<float4> xm(nx, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<float2> xr(ny, make_float2(1.f, 1.f));
    if (nx > 2) {
        const double* s1 = d->s1;
        const int l = nx - 1;
        std::vector<double> b(l), aj(ny);
        for (int i = 0; i < l; i++) b[i] = s1[i];
        for (int j = 0; j < ny; j++) aj[j] = s1[(size_t)l * j] / b[0];
        std::atomic<bool> separable{true};
        for_rows(ny, m, [&](int j0, int j1) {
            for (int j = j0; j < j1; j++)
                for (int i = 0; i < l; i++) {
                    const double v = s1[i + (size_t)l * j];
                    if (!(v > 0.0) || std::fabs(aj[j] * b[i] - v) > 1e-9 * v) separable = false;
                }
        });
        if (!separable)
            return bail(fail(CLIFFS_ERR_METRIC, "s1(i,j) must be positive and separable a(j)*b(i)"));
        for (int i = 0; i < nx; i++) {
            const double bm = i >= 1 ? b[i - 1] : b[0];
            const double bp = i < l ? b[i] : b[l - 1];
            xm[i] = make_float4((float)(1.0 / bm), (float)(1.0 / bp), (float)(1.0 / (bm + bp)),
                                (float)(bp - bm));
        }
        for (int j = 0; j < ny; j++) xr[j] = make_float2((float)(1.0 / aj[j]), (float)aj[j]);
    }
    // ---- metrics.  y: s2(j), zeta(j)  (MasterGrid.f:103-116)
    std::vector<float4> ym(ny, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<float> zt(ny, 0.0f);
    if (ny > 2) {
        const double* s2 = d->s2;
        for (int j = 0; j < ny - 1; j++)
            if (!(s2[j] > 0.0)) return bail(fail(CLIFFS_ERR_METRIC, "s2 must be positive"));
        for (int j = 0; j < ny; j++) {
            const double bm = j >= 1 ? s2[j - 1] : s2[0];
            const double bp = j < ny - 1 ? s2[j] : s2[ny - 2];
            ym[j] = make_float4((float)(1.0 / bm), (float)(1.0 / bp), (float)(1.0 / (bm + bp)),
                                (float)(bp - bm));
            zt[j] = (float)d->zeta[j];
        }
    }

    // ---- device allocations
    h->fstride = (size_t)h->pitch * ny + 64;
    // the fused step addresses all fields by 32-bit byte offsets from one base;
    // larger grids (> ~119M cells) step as two sweep kernels instead
    h->fused_ok = 9 * h->fstride * sizeof(float) < ((size_t)1 << 32);
    if (hipMalloc(&h->fields, 9 * h->fstride * sizeof(float)) != hipSuccess)
        return bail(fail(CLIFFS_ERR_NOMEM, "device state allocation"));
    if (hipMalloc(&h->xm, sizeof(float4) * nx) != hipSuccess ||
        hipMalloc(&h->xr, sizeof(float2) * ny) != hipSuccess ||
        hipMalloc(&h->ym, sizeof(float4) * ny) != hipSuccess ||
        hipMalloc(&h->zeta, sizeof(float) * ny) != hipSuccess)
        return bail(fail(CLIFFS_ERR_NOMEM, "device metric allocation"));
    h->edge_slot_floats = ((size_t)6 * nx + 6 * ny + 63) / 64 * 64;
    if (hipMalloc(&h->edges, sizeof(float) * h->edge_slot_floats * kEdgeSlots) != hipSuccess ||
        hipHostMalloc(&h->hedges, sizeof(float) * h->edge_slot_floats * kEdgeSlots) != hipSuccess)
        return bail(fail(CLIFFS_ERR_NOMEM, "edge staging allocation"));
    for (auto& e : h->edge_ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return bail(fail(CLIFFS_ERR_HIP, "hipEventCreate"));
    h->fbytes = ((size_t)std::max(nx, ny) + 8 + 255) / 256 * 256;
    if (hipMalloc(&h->flags, 3 * h->fbytes) != hipSuccess)
        return bail(fail(CLIFFS_ERR_NOMEM, "flag allocation"));
    {  // fused step: tile geometry for this device, per-tile line flags
        if (hipDeviceGetAttribute(&h->f_cus, hipDeviceAttributeMultiprocessorCount, h->dev) !=
                hipSuccess ||
            fused_blocks_per_cu(kFStrip, &h->f_bpc) != hipSuccess ||
            fused_blocks_per_cu(kFStripNarrow, &h->f_bpc_narrow) != hipSuccess || h->f_cus < 1 ||
            h->f_bpc < 1 || h->f_bpc_narrow < 1)
            return bail(fail(CLIFFS_ERR_HIP, "device occupancy query"));
        const int rc = set_fused_rows(h, 0);
        if (rc) return bail(rc);
    }

    // every copy and clear on the grid's stream, which is non-blocking: work
    // on the null stream would not be ordered with the grid's kernels (with
    // several host threads creating and stepping grids, a late clear of the
    // fields or flags has wiped state already uploaded)
    if (hipMemcpyAsync(h->xm, xm.data(), sizeof(float4) * nx, hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(h->xr, xr.data(), sizeof(float2) * ny, hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(h->ym, ym.data(), sizeof(float4) * ny, hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(h->zeta, zt.data(), sizeof(float) * ny, hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipStreamSynchronize(h->st) != hipSuccess)
        return bail(fail(CLIFFS_ERR_HIP, "metric upload"));
    if (hipMemsetAsync(h->fields, 0, 9 * h->fstride * sizeof(float), h->st) != hipSuccess)
        return bail(fail(CLIFFS_ERR_HIP, "state clear"));

    // background celerity
    h->hbuf.assign((size_t)h->pitch * ny, 0.0f);
    for_rows(ny, m, [&](int j0, int j1) {
        for (int j = j0; j < j1; j++)
            for (int i = 0; i < nx; i++) h->hbuf[hidx(h, i, j)] = h->s[(size_t)i + (size_t)nx * j];
    });
    int rc = put_field(h, h->grid().S, h->hbuf);
    if (rc) return bail(rc);
    rc = cliffs_reset_max(h);
    if (rc) return bail(rc);
    *out = h;
    return CLIFFS_OK;
}

int cliffs_destroy(cliffs_handle* h) {
    if (h && h->users > 0)
        return fail(CLIFFS_ERR_STATE, "grid still used by " + std::to_string(h->users) +
                                          " nest(s) or ensemble(s): destroy those first");
    free_handle(h);
    return CLIFFS_OK;
}

int cliffs_upload_state(cliffs_handle* h, const double* cel, const double* xvel,
                        const double* yvel) {
    if (!h || !cel || !xvel || !yvel) return fail(CLIFFS_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(h->dev));
    const int nx = h->nx, ny = h->ny;
    std::vector<float>& b = h->hbuf;
    const size_t m = (size_t)nx * ny;
    b.assign((size_t)h->pitch * ny, 0.0f);
    for_rows(ny, m, [&](int j0, int j1) {
        for (int j = j0; j < j1; j++)
            for (int i = 0; i < nx; i++) {
                const size_t c = (size_t)i + (size_t)nx * j;
                b[hidx(h, i, j)] = to_cp(cel[c], h->dep[c], h->s[c], h->k);
            }
    });
    int rc = put_field(h, h->grid().CP, b);
    if (rc) return rc;
    for (const double* f : {xvel, yvel}) {
        for_rows(ny, m, [&](int j0, int j1) {
            for (int j = j0; j < j1; j++)
                for (int i = 0; i < nx; i++) b[hidx(h, i, j)] = (float)f[(size_t)i + (size_t)nx * j];
        });
        if ((rc = put_field(h, f == xvel ? h->grid().U : h->grid().V, b))) return rc;
    }
    h->state = true;
    h->flags_valid = false;
    h->fused_valid = false;
    return CLIFFS_OK;
}

int cliffs_download_state(cliffs_handle* h, double* cel, double* xvel, double* yvel) {
    if (!h || !cel || !xvel || !yvel) return fail(CLIFFS_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(h->dev));
    const int nx = h->nx, ny = h->ny;
    std::vector<float> b;
    const size_t m = (size_t)nx * ny;
    int rc = get_field(h, h->grid().CP, b);
    if (rc) return rc;
    for_rows(ny, m, [&](int j0, int j1) {
        for (int j = j0; j < j1; j++)
            for (int i = 0; i < nx; i++) {
                const size_t c = (size_t)i + (size_t)nx * j;
                cel[c] = to_cel(h->s[c], b[hidx(h, i, j)], h->dep[c], h->k);
            }
    });
    for (double* f : {xvel, yvel}) {
        if ((rc = get_field(h, f == xvel ? h->grid().U : h->grid().V, b))) return rc;
        for_rows(ny, m, [&](int j0, int j1) {
            for (int j = j0; j < j1; j++)
                for (int i = 0; i < nx; i++) f[(size_t)i + (size_t)nx * j] = b[hidx(h, i, j)];
        });
    }
    return CLIFFS_OK;
}

int cliffs_download_max(cliffs_handle* h, double* hmax, double* velmax) {
    if (!h || !hmax || !velmax) return fail(CLIFFS_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(h->dev));
    const int nx = h->nx, ny = h->ny;
    std::vector<float> b;
    const size_t m = (size_t)nx * ny;
    int rc = get_field(h, h->grid().HM, b);
    if (rc) return rc;
    for_rows(ny, m, [&](int j0, int j1) {
        for (int j = j0; j < j1; j++)
            for (int i = 0; i < nx; i++) {
                const size_t c = (size_t)i + (size_t)nx * j;
                hmax[c] = to_cel(h->s[c], b[hidx(h, i, j)], h->dep[c], h->k);
            }
    });
    if ((rc = get_field(h, h->grid().VM, b))) return rc;
    for_rows(ny, m, [&](int j0, int j1) {
        for (int j = j0; j < j1; j++)
            for (int i = 0; i < nx; i++) velmax[(size_t)i + (size_t)nx * j] = b[hidx(h, i, j)];
    });
    return CLIFFS_OK;
}

int cliffs_reset_max(cliffs_handle* h) {
    if (!h) return fail(CLIFFS_ERR_ARG, "NULL handle");
    HIPCHK(hipSetDevice(h->dev));
    // hmax = 0 means c = 0, i.e. cp = -max(s,0), from the device's S
    const hipError_t e = launch_reset_max(h->grid(), h->st);
    if (e != hipSuccess) return hipfail(e, "reset_max");
    HIPCHK(hipStreamSynchronize(h->st));
    return CLIFFS_OK;
}

int cliffs_sweep(cliffs_handle* h, int iu, int first, int last, const double* edge1,
                 const double* edge2, int mx) {
    if (!h) return fail(CLIFFS_ERR_ARG, "NULL handle");
    if (!h->state) return fail(CLIFFS_ERR_STATE, "no state uploaded");
    if (iu != 1 && iu != 2) return fail(CLIFFS_ERR_ARG, "iu must be 1 or 2");
    const int nl = iu == 1 ? h->ny : h->nx;
    const int n = iu == 1 ? h->nx : h->ny;
    if (first < 1 || last > nl || first > last) return fail(CLIFFS_ERR_ARG, "line range out of grid");
    if (n < 3) return fail(CLIFFS_ERR_ARG, "line length must be >= 3");
    HIPCHK(hipSetDevice(h->dev));
    const float* ed = nullptr;
    int rc = stage_edges(h, iu, edge1, edge2, mx, &ed);
    if (rc) return rc;
    // flags of this direction from the current state, then the sweep
    HIPCHK(hipMemsetAsync(h->flags, 0, 3 * h->fbytes, h->st));
    h->m = 0;
    if ((rc = launch_flags(h, iu == 1 ? 0 : 1))) return rc;
    if (iu == 1) {
        XArgs a = xargs(h);
        a.edges = ed;
        a.fin = h->fbuf(h->m);
        a.fout = h->fbuf(h->m + 1);
        a.fclr = reinterpret_cast<unsigned*>(h->fbuf(h->m + 2));
        a.nclr = clear_flag_words(h);
        a.row_lo = a.act_lo = first - 1;
        a.row_hi = a.act_hi = last - 1;
        a.flag_rows = 1;
        a.fuse_max = 0;
        const hipError_t e = launch_sweep_x(a, h->st);
        if (e != hipSuccess) return hipfail(e, "sweep_x");
    } else {
        YArgs a = yargs(h);
        a.edges = ed;
        a.fin = h->fbuf(h->m);
        a.fout = h->fbuf(h->m + 1);
        a.fclr = reinterpret_cast<unsigned*>(h->fbuf(h->m + 2));
        a.nclr = clear_flag_words(h);
        a.col_base = (first - 1) & ~63;
        a.col_span = last - a.col_base;
        a.col_lo = first - 1;
        a.col_hi = last - 1;
        a.flag_cols = 1;
        a.fuse_max = 0;
        const hipError_t e = launch_sweep_y(a, h->st);
        if (e != hipSuccess) return hipfail(e, "sweep_y");
    }
    h->m++;
    // only lines first..last were written: copy them back into the current copy
    const hipError_t e = launch_copy_lines(h->grid(), iu == 1 ? 0 : 1, first - 1, last - 1, h->st);
    if (e != hipSuccess) return hipfail(e, "copy_lines");
    h->flags_valid = false;
    h->fused_valid = false;
    return CLIFFS_OK;
}

int cliffs_maxwave(cliffs_handle* h) {
    if (!h) return fail(CLIFFS_ERR_ARG, "NULL handle");
    HIPCHK(hipSetDevice(h->dev));
    const hipError_t e = launch_maxwave(h->grid(), h->st);
    if (e != hipSuccess) return hipfail(e, "maxwave");
    return CLIFFS_OK;
}

int cliffs_step(cliffs_handle* h, const double* xedge1, const double* xedge2,
                const double* yedge1, const double* yedge2, int mx) {
    if (!h) return fail(CLIFFS_ERR_ARG, "NULL handle");
    if (!h->state) return fail(CLIFFS_ERR_STATE, "no state uploaded");
    HIPCHK(hipSetDevice(h->dev));
    const float *xe = nullptr, *ye = nullptr;
    int rc;
    if (h->xrun && (rc = stage_edges(h, 1, xedge1, xedge2, mx, &xe))) return rc;
    if (h->yrun && (rc = stage_edges(h, 2, yedge1, yedge2, mx, &ye))) return rc;
    return do_step(h, xe, ye);
}

int cliffs_run(cliffs_handle* h, int nsteps) {
    if (!h) return fail(CLIFFS_ERR_ARG, "NULL handle");
    if (!h->state) return fail(CLIFFS_ERR_STATE, "no state uploaded");
    if (nsteps < 0) return fail(CLIFFS_ERR_ARG, "nsteps < 0");
    HIPCHK(hipSetDevice(h->dev));
    for (int s = 0; s < nsteps; s++) {
        const int rc = do_step(h, nullptr, nullptr);
        if (rc) return rc;
    }
    return CLIFFS_OK;
}

// Events bracket the sweeps of every kTimedStride-th step.  They live in a
// ring of kTimedSlots bracketed steps: before a slot is reused its (long
// finished) events are read and accumulated, so the host never waits on
// recent work and the event count stays bounded.
constexpr int kTimedSlots = 256;   // event ring: slots of four events (x/y launch brackets)
constexpr int kTimedStride = 8;   // one bracketed step in kTimedStride

namespace {
struct EventSet {
    std::vector<hipEvent_t> ev;
    ~EventSet() {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};
}  // namespace

int cliffs_run_timed(cliffs_handle* h, int nsteps, double* ms_total, double* ms_x, double* ms_y) {
    if (!h || !ms_total || !ms_x || !ms_y) return fail(CLIFFS_ERR_ARG, "NULL argument");
    if (!h->state) return fail(CLIFFS_ERR_STATE, "no state uploaded");
    if (nsteps < 1) return fail(CLIFFS_ERR_ARG, "nsteps < 1");
    HIPCHK(hipSetDevice(h->dev));
    int rc;
    const int slots = std::min((nsteps + kTimedStride - 1) / kTimedStride, kTimedSlots);
    EventSet es;
    es.ev.assign(4 * (size_t)slots + 2, nullptr);
    for (auto& e : es.ev) HIPCHK(hipEventCreate(&e));
    hipEvent_t* ev = es.ev.data();
    double sx = 0, sy = 0;
    float t;
    auto harvest = [&](int q) -> int {  // accumulate the sweep times of slot q
        hipEvent_t* e = ev + 2 + 4 * (size_t)q;
        if (h->xrun) {
            HIPCHK(hipEventSynchronize(e[1]));
            HIPCHK(hipEventElapsedTime(&t, e[0], e[1]));
            sx += t;
        }
        if (h->yrun) {
            HIPCHK(hipEventSynchronize(e[3]));
            HIPCHK(hipEventElapsedTime(&t, e[2], e[3]));
            sy += t;
        }
        return CLIFFS_OK;
    };
    // every kTimedStride-th step is bracketed (event records between launches
    // cost the stream a few microseconds each); its launches are a sample of
    // the region's, in the same sequence
    const int nsampled = (nsteps + kTimedStride - 1) / kTimedStride;
    HIPCHK(hipEventRecord(ev[0], h->st));
    for (int s = 0, k = 0; s < nsteps; s++) {
        if (s % kTimedStride) {
            if ((rc = do_step(h, nullptr, nullptr))) return rc;
            continue;
        }
        const int q = k % slots;
        if (k >= slots && (rc = harvest(q))) return rc;
        if ((rc = do_step(h, nullptr, nullptr, ev + 2 + 4 * (size_t)q))) return rc;
        k++;
    }
    HIPCHK(hipEventRecord(ev[1], h->st));
    HIPCHK(hipEventSynchronize(ev[1]));
    HIPCHK(hipEventElapsedTime(&t, ev[0], ev[1]));
    *ms_total = t;
    for (int k = std::max(nsampled - slots, 0); k < nsampled; k++)
        if ((rc = harvest(k % slots))) return rc;
    *ms_x = h->xrun ? sx / nsampled : 0.0;
    *ms_y = h->yrun ? sy / nsampled : 0.0;
    return CLIFFS_OK;
}

}  // extern "C"

namespace {

int gauge_args(cliffs_handle* h, int every, int nsamples, int n, const int* I, const int* J,
               float* eta) {
    if (!h || every < 1 || nsamples < 1 || n < 1 || !I || !J || !eta)
        return fail(CLIFFS_ERR_ARG, "bad gauge-series arguments");
    if (!h->state) return fail(CLIFFS_ERR_STATE, "no state uploaded");
    for (int q = 0; q < n; q++)
        if (I[q] < 1 || I[q] > h->nx || J[q] < 1 || J[q] > h->ny)
            return fail(CLIFFS_ERR_ARG, "gauge index out of grid");
    return CLIFFS_OK;
}

// The steps and samples of a gauge series into device buffers dij (2n ints)
// and ds (n * nsamples floats), the series copied to eta; waits for the end.
int run_gauges_into(cliffs_handle* h, int every, int nsamples, int n, const int* I, const int* J,
                    float* eta, int* dij, float* ds) {
    int rc = CLIFFS_OK;
    hipError_t e = hipMemcpyAsync(dij, I, sizeof(int) * n, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dij + n, J, sizeof(int) * n, hipMemcpyHostToDevice, h->st);
    for (int k = 0; k < nsamples && e == hipSuccess && rc == CLIFFS_OK; k++) {
        for (int s = 0; s < every && rc == CLIFFS_OK; s++) rc = do_step(h, nullptr, nullptr);
        if (rc == CLIFFS_OK) e = launch_gauge_sample(h->grid(), h->p, dij, dij + n, n, ds, k, h->st);
    }
    if (e == hipSuccess && rc == CLIFFS_OK)
        e = hipMemcpyAsync(eta, ds, sizeof(float) * (size_t)n * nsamples, hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess && rc == CLIFFS_OK) e = hipStreamSynchronize(h->st);
    if (rc) return rc;
    if (e != hipSuccess) return hipfail(e, "gauge series");
    return CLIFFS_OK;
}

}  // namespace

extern "C" {

int cliffs_run_gauges(cliffs_handle* h, int every, int nsamples, int n, const int* I, const int* J,
                      float* eta) {
    int rc = gauge_args(h, every, nsamples, n, I, J, eta);
    if (rc) return rc;
    HIPCHK(hipSetDevice(h->dev));
    int* dij = nullptr;
    float* ds = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(dij);
        (void)hipFree(ds);
    };
    if (hipMalloc(&dij, sizeof(int) * 2 * n) != hipSuccess ||
        hipMalloc(&ds, sizeof(float) * (size_t)n * nsamples) != hipSuccess) {
        cleanup();
        return fail(CLIFFS_ERR_NOMEM, "gauge series allocation");
    }
    rc = run_gauges_into(h, every, nsamples, n, I, J, eta, dij, ds);
    cleanup();
    return rc;
}

int cliffs_run_gauges_noalloc(cliffs_handle* h, int every, int nsamples, int n, const int* I,
                              const int* J, float* eta) {
    int rc = gauge_args(h, every, nsamples, n, I, J, eta);
    if (rc) return rc;
    const size_t words = gauge_scratch_words();
    if ((size_t)n * ((size_t)nsamples + 2) > words)
        return fail(CLIFFS_ERR_ARG, "gauge series too large for the device scratch: n*(nsamples+2) must be <= " +
                                        std::to_string(words));
    HIPCHK(hipSetDevice(h->dev));
    void* base = nullptr;
    const hipError_t e = gauge_scratch(h->dev, &base);
    if (e != hipSuccess) return hipfail(e, "hipGetSymbolAddress(scratch)");
    int* dij = static_cast<int*>(base);
    float* ds = reinterpret_cast<float*>(dij + 2 * (size_t)n);
    // every slot of the series is written by its sample; nothing to clear
    return run_gauges_into(h, every, nsamples, n, I, J, eta, dij, ds);
}

int cliffs_health(cliffs_handle* h, double* courant, long long* nonfinite) {
    if (!h || !courant || !nonfinite) return fail(CLIFFS_ERR_ARG, "NULL argument");
    if (!h->state) return fail(CLIFFS_ERR_STATE, "no state uploaded");
    HIPCHK(hipSetDevice(h->dev));
    if (!h->health && hipMalloc(&h->health, sizeof(HealthOut)) != hipSuccess) {
        h->health = nullptr;
        return fail(CLIFFS_ERR_NOMEM, "health result allocation");
    }
    HIPCHK(hipMemsetAsync(h->health, 0, sizeof(HealthOut), h->st));
    const hipError_t e = launch_health(h->grid(), h->p, h->xm, h->xr, h->ym, h->xrun ? 1 : 0,
                                       h->yrun ? 1 : 0, h->health, h->st);
    if (e != hipSuccess) return hipfail(e, "health");
    HealthOut r;
    HIPCHK(hipMemcpyAsync(&r, h->health, sizeof r, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    float c;
    std::memcpy(&c, &r.cmax_bits, sizeof c);
    *courant = c;
    *nonfinite = (long long)r.nonfinite;
    return CLIFFS_OK;
}

int cliffs_set_fused(cliffs_handle* h, int on) {
    if (!h) return fail(CLIFFS_ERR_ARG, "NULL handle");
    h->fused = on != 0;
    return CLIFFS_OK;
}

int cliffs_set_tile_width(cliffs_handle* h, int width) {
    if (!h) return fail(CLIFFS_ERR_ARG, "NULL handle");
    if (width != kFStrip && width != kFStripNarrow)
        return fail(CLIFFS_ERR_ARG, "tile width must be 248 or 120");
    if (!h->fused_ok) return fail(CLIFFS_ERR_ARG, "grid has no fused step (1-D or too large)");
    if (h->ens_users > 0)
        return fail(CLIFFS_ERR_STATE, "grid belongs to an ensemble, whose tuner sets its tile shape");
    HIPCHK(hipSetDevice(h->dev));
    return set_fused_shape(h, width, 0);
}

int cliffs_set_tile_rows(cliffs_handle* h, int rows) {
    if (!h) return fail(CLIFFS_ERR_ARG, "NULL handle");
    if (rows != 0 && (rows < 8 || rows > kFMaxRows)) return fail(CLIFFS_ERR_ARG, "tile rows must be 0 or 8..256");
    if (!h->fused_ok) return fail(CLIFFS_ERR_ARG, "grid has no fused step (1-D or too large)");
    if (h->ens_users > 0)
        return fail(CLIFFS_ERR_STATE, "grid belongs to an ensemble, whose tuner sets its tile height");
    HIPCHK(hipSetDevice(h->dev));
    return set_fused_rows(h, rows);
}

// ---- ensembles --------------------------------------------------------------
}  // extern "C"

struct cliffs_ensemble {
    int dev = 0;
    hipStream_t st = nullptr;
    std::vector<cliffs_handle*> m;
    int width = kFStrip;          // tile width of every member (one kernel per launch)
    int rows = 0;                 // common tile height (0: each member's own)
    FArgs* hargs = nullptr;       // pinned staging, [2][n]: this step's and the next step's
    FArgs* dargs = nullptr;       // device copy
    int2* dmap = nullptr;         // block -> (member, tile)
    int nblocks = 0;
    hipEvent_t ev = nullptr;      // stream hand-over between the members and the ensemble
    hipEvent_t argev = nullptr;   // the argument copy has executed (staging reusable)
    hipGraph_t graph = nullptr;   // two steps (both parities)
    hipGraphExec_t gexec = nullptr;
    std::vector<int> tiles;       // members' tile counts the block map and graph were built for
    bool counted = false;         // members' use counts include this ensemble
};

namespace {

void free_ensemble(cliffs_ensemble* e) {
    if (!e) return;
    (void)hipSetDevice(e->dev);
    if (e->st) (void)hipStreamSynchronize(e->st);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->ev) (void)hipEventDestroy(e->ev);
    if (e->argev) (void)hipEventDestroy(e->argev);
    (void)hipFree(e->dargs);
    (void)hipFree(e->dmap);
    if (e->hargs) (void)hipHostFree(e->hargs);
    if (e->st) (void)hipStreamDestroy(e->st);
    if (e->counted)
        for (cliffs_handle* h : e->m) {
            h->users--;
            h->ens_users--;
        }
    delete e;
}

// Block map and the captured two-step graph for the members' current geometry.
int ensemble_rebuild(cliffs_ensemble* e) {
    const int n = (int)e->m.size();
    std::vector<int> nt(n);
    for (int i = 0; i < n; i++) nt[i] = e->m[i]->f_nstrips * e->m[i]->f_nchunks;
    e->tiles.clear();
    std::vector<int2> bmap;
    ensemble_block_map(nt.data(), n, bmap);
    HIPCHK(hipStreamSynchronize(e->st));
    if (e->gexec) HIPCHK(hipGraphExecDestroy(e->gexec));
    if (e->graph) HIPCHK(hipGraphDestroy(e->graph));
    e->gexec = nullptr;
    e->graph = nullptr;
    HIPCHK(hipFree(e->dmap));
    e->dmap = nullptr;
    e->nblocks = (int)bmap.size();
    if (hipMalloc(&e->dmap, sizeof(int2) * bmap.size()) != hipSuccess)
        return fail(CLIFFS_ERR_NOMEM, "ensemble block map");
    HIPCHK(hipMemcpyAsync(e->dmap, bmap.data(), sizeof(int2) * bmap.size(), hipMemcpyHostToDevice, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    // two steps: the argument blocks of this step, then of the next
    HIPCHK(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    for (int k = 0; k < 2 && err == hipSuccess; k++) {
        err = launch_step_batch(e->dargs + (size_t)k * n, e->dmap, e->nblocks, e->width, e->st);
        if (err == hipSuccess)
            err = launch_fixup_batch(e->dargs + (size_t)k * n, e->dmap, e->nblocks, e->width, e->st);
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(e->st, &g);
    if (err != hipSuccess) return hipfail(err, "ensemble capture");
    if (ec != hipSuccess) return hipfail(ec, "hipStreamEndCapture");
    e->graph = g;
    HIPCHK(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
    e->tiles = nt;
    return CLIFFS_OK;
}

// Members' work so far happens before the ensemble's next work; flags of the
// fused step are made current on each member's stream first.
int ensemble_take(cliffs_ensemble* e) {
    if (!e->gexec) return fail(CLIFFS_ERR_STATE, "ensemble has no step graph (a rebuild failed)");
    // a member shared with another ensemble may have been re-tiled by that
    // ensemble's tuner: give it this ensemble's tile width back, and make the
    // block map and graph match its tiles
    bool rebuild = false;
    for (size_t i = 0; i < e->m.size(); i++) {
        cliffs_handle* h = e->m[i];
        if (h->f_width != e->width) {
            HIPCHK(hipStreamSynchronize(e->st));
            const int rc = set_fused_shape(h, e->width, e->rows);
            if (rc) return rc;
        }
        if (h->f_nstrips * h->f_nchunks != e->tiles[i]) rebuild = true;
    }
    if (rebuild) {
        const int rc = ensemble_rebuild(e);
        if (rc) return rc;
    }
    for (cliffs_handle* h : e->m) {
        int rc;
        if ((rc = ensure_fused(h))) return rc;
        HIPCHK(hipEventRecord(e->ev, h->st));
        HIPCHK(hipStreamWaitEvent(e->st, e->ev, 0));
    }
    return CLIFFS_OK;
}

// The ensemble's work so far happens before each member's next work.
int ensemble_give(cliffs_ensemble* e) {
    HIPCHK(hipEventRecord(e->ev, e->st));
    for (cliffs_handle* h : e->m) HIPCHK(hipStreamWaitEvent(h->st, e->ev, 0));
    return CLIFFS_OK;
}

// Argument blocks of every member for its current step (slot 0) and the
// step after it (slot 1), copied to the device on the ensemble's stream.
// maxwave = 0 (timing trials) leaves the running maxima untouched.
int ensemble_args(cliffs_ensemble* e, int maxwave = 1) {
    const int n = (int)e->m.size();
    HIPCHK(hipEventSynchronize(e->argev));  // the previous copy has left the staging buffer
    for (int i = 0; i < n; i++) {
        cliffs_handle* h = e->m[i];
        e->hargs[i] = fargs(h, nullptr, nullptr);
        h->cur ^= 1;
        h->fn++;
        e->hargs[n + i] = fargs(h, nullptr, nullptr);
        h->cur ^= 1;
        h->fn--;
        e->hargs[i].maxwave = e->hargs[n + i].maxwave = maxwave;
    }
    HIPCHK(hipMemcpyAsync(e->dargs, e->hargs, sizeof(FArgs) * 2 * n, hipMemcpyHostToDevice, e->st));
    HIPCHK(hipEventRecord(e->argev, e->st));
    return CLIFFS_OK;
}

int ensemble_single(cliffs_ensemble* e, int slot) {
    const FArgs* a = e->dargs + (size_t)slot * e->m.size();
    hipError_t err = launch_step_batch(a, e->dmap, e->nblocks, e->width, e->st);
    if (err != hipSuccess) return hipfail(err, "ensemble step");
    err = launch_fixup_batch(a, e->dmap, e->nblocks, e->width, e->st);
    if (err != hipSuccess) return hipfail(err, "ensemble fixup");
    return CLIFFS_OK;
}

int ensemble_set_shape(cliffs_ensemble* e, int width, int rows) {
    HIPCHK(hipStreamSynchronize(e->st));
    for (cliffs_handle* h : e->m) {
        const int rc = set_fused_shape(h, width, rows);
        if (rc) {
            // the captured graph and block map no longer match the members'
            // tiles: refuse to run until a rebuild succeeds
            if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
            if (e->graph) (void)hipGraphDestroy(e->graph);
            e->gexec = nullptr;
            e->graph = nullptr;
            return rc;
        }
    }
    e->width = width;
    e->rows = rows;
    return ensemble_rebuild(e);
}

}  // namespace

extern "C" {

int cliffs_ensemble_create(cliffs_handle* const* members, int n, cliffs_ensemble** out) {
    if (!members || n < 1 || !out) return fail(CLIFFS_ERR_ARG, "bad ensemble arguments");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        cliffs_handle* h = members[i];
        if (!h) return fail(CLIFFS_ERR_ARG, "NULL member");
        if (h->dev != members[0]->dev) return fail(CLIFFS_ERR_ARG, "members on different devices");
        if (!(h->xrun && h->yrun) || !h->fused_ok)
            return fail(CLIFFS_ERR_ARG, "ensemble members must be 2-D grids (nx, ny > 2) of < ~119M cells");
        if (!h->state) return fail(CLIFFS_ERR_STATE, "member without uploaded state");
        for (int k = 0; k < i; k++)
            if (members[k] == h) return fail(CLIFFS_ERR_ARG, "member listed twice");
        if (h->f_width != members[0]->f_width)
            return fail(CLIFFS_ERR_ARG, "members with different tile widths");
    }
    cliffs_ensemble* e = new (std::nothrow) cliffs_ensemble;
    if (!e) return fail(CLIFFS_ERR_NOMEM, "host allocation");
    e->dev = members[0]->dev;
    e->width = members[0]->f_width;
    e->m.assign(members, members + n);
    auto bail = [&](int rc) {
        free_ensemble(e);
        return rc;
    };
    if (hipSetDevice(e->dev) != hipSuccess) return bail(fail(CLIFFS_ERR_HIP, "hipSetDevice"));
    if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->argev, hipEventDisableTiming) != hipSuccess)
        return bail(fail(CLIFFS_ERR_HIP, "ensemble stream/events"));
    if (hipHostMalloc(&e->hargs, sizeof(FArgs) * 2 * n) != hipSuccess ||
        hipMalloc(&e->dargs, sizeof(FArgs) * 2 * n) != hipSuccess)
        return bail(fail(CLIFFS_ERR_NOMEM, "ensemble argument blocks"));
    if (hipEventRecord(e->argev, e->st) != hipSuccess) return bail(fail(CLIFFS_ERR_HIP, "event"));
    const int rc = ensemble_rebuild(e);
    if (rc) return bail(rc);
    for (cliffs_handle* h : e->m) {
        h->users++;
        h->ens_users++;
    }
    e->counted = true;
    *out = e;
    return CLIFFS_OK;
}

int cliffs_ensemble_destroy(cliffs_ensemble* e) {
    free_ensemble(e);
    return CLIFFS_OK;
}

int cliffs_ensemble_run(cliffs_ensemble* e, int nsteps) {
    if (!e || nsteps < 0) return fail(CLIFFS_ERR_ARG, "bad ensemble run arguments");
    if (nsteps == 0) return CLIFFS_OK;
    HIPCHK(hipSetDevice(e->dev));
    int rc;
    if ((rc = ensemble_take(e)) || (rc = ensemble_args(e))) return rc;
    for (int k = 0; k < nsteps / 2; k++) HIPCHK(hipGraphLaunch(e->gexec, e->st));
    if ((nsteps & 1) && (rc = ensemble_single(e, 0))) return rc;
    for (cliffs_handle* h : e->m) {
        h->cur ^= (nsteps & 1);
        h->fn += nsteps;
        h->flags_valid = false;
    }
    return ensemble_give(e);
}

int cliffs_ensemble_run_timed(cliffs_ensemble* e, int nsteps, double* ms_total) {
    if (!e || nsteps < 1 || !ms_total) return fail(CLIFFS_ERR_ARG, "bad ensemble run arguments");
    HIPCHK(hipSetDevice(e->dev));
    EventSet es;
    es.ev.assign(2, nullptr);
    for (auto& x : es.ev) HIPCHK(hipEventCreate(&x));
    int rc;
    if ((rc = ensemble_take(e)) || (rc = ensemble_args(e))) return rc;
    HIPCHK(hipEventRecord(es.ev[0], e->st));
    for (int k = 0; k < nsteps / 2; k++) HIPCHK(hipGraphLaunch(e->gexec, e->st));
    if ((nsteps & 1) && (rc = ensemble_single(e, 0))) return rc;
    HIPCHK(hipEventRecord(es.ev[1], e->st));
    for (cliffs_handle* h : e->m) {
        h->cur ^= (nsteps & 1);
        h->fn += nsteps;
        h->flags_valid = false;
    }
    if ((rc = ensemble_give(e))) return rc;
    HIPCHK(hipEventSynchronize(es.ev[1]));
    float t;
    HIPCHK(hipEventElapsedTime(&t, es.ev[0], es.ev[1]));
    *ms_total = t;
    return CLIFFS_OK;
}

int cliffs_ensemble_tune(cliffs_ensemble* e, int trial_steps, int* rows_out) {
    if (!e || trial_steps < 1) return fail(CLIFFS_ERR_ARG, "bad tune arguments");
    HIPCHK(hipSetDevice(e->dev));
    static const int kCand[] = {16, 24, 32, 48, 64, 96, 128, 160, 192, 256};
    static const int kWidths[] = {kFStrip, kFStripNarrow};
    EventSet es;
    es.ev.assign(2, nullptr);
    for (auto& x : es.ev) HIPCHK(hipEventCreate(&x));
    int best = 0, best_w = kFStrip, rc;
    float best_ms = 0.f;
    for (int wd : kWidths)
    for (int c : kCand) {
        if ((rc = ensemble_set_shape(e, wd, c))) return rc;
        if ((rc = ensemble_take(e)) || (rc = ensemble_args(e, 0))) return rc;
        // the same step, repeated: it reads the current state and writes the
        // other copy, so the state does not advance; the running maxima are
        // not touched (maxwave = 0) and the line flags are rebuilt afterwards
        if ((rc = ensemble_single(e, 0))) return rc;
        HIPCHK(hipEventRecord(es.ev[0], e->st));
        for (int k = 0; k < trial_steps; k++)
            if ((rc = ensemble_single(e, 0))) return rc;
        HIPCHK(hipEventRecord(es.ev[1], e->st));
        HIPCHK(hipEventSynchronize(es.ev[1]));
        float t;
        HIPCHK(hipEventElapsedTime(&t, es.ev[0], es.ev[1]));
        for (cliffs_handle* h : e->m) h->fused_valid = false;
        if (best == 0 || t < best_ms) {
            best = c;
            best_w = wd;
            best_ms = t;
        }
    }
    if ((rc = ensemble_set_shape(e, best_w, best))) return rc;
    for (cliffs_handle* h : e->m) h->fused_valid = false;
    if ((rc = ensemble_give(e))) return rc;
    if (rows_out) *rows_out = best;
    return CLIFFS_OK;
}

int cliffs_ensemble_shape(cliffs_ensemble* e, int* width, int* rows) {
    if (!e || !width || !rows) return fail(CLIFFS_ERR_ARG, "NULL argument");
    *width = e->width;
    *rows = e->rows;
    return CLIFFS_OK;
}

int cliffs_ensemble_run_gauges(cliffs_ensemble* e, int every, int nsamples, int n, const int* I,
                               const int* J, float* eta) {
    if (!e || every < 2 || (every & 1) || nsamples < 1 || n < 1 || !I || !J || !eta)
        return fail(CLIFFS_ERR_ARG, "bad ensemble gauge-series arguments (every must be even)");
    const int nm = (int)e->m.size();
    for (int m = 0; m < nm; m++)
        for (int q = 0; q < n; q++) {
            const int i = I[(size_t)m * n + q], j = J[(size_t)m * n + q];
            if (i < 1 || i > e->m[m]->nx || j < 1 || j > e->m[m]->ny)
                return fail(CLIFFS_ERR_ARG, "gauge index out of a member's grid");
        }
    HIPCHK(hipSetDevice(e->dev));
    int* dij = nullptr;
    float* ds = nullptr;
    const size_t nij = (size_t)nm * n, nser = (size_t)nm * nsamples * n;
    if (hipMalloc(&dij, sizeof(int) * 2 * nij) != hipSuccess ||
        hipMalloc(&ds, sizeof(float) * nser) != hipSuccess) {
        (void)hipFree(dij);
        return fail(CLIFFS_ERR_NOMEM, "ensemble gauge series allocation");
    }
    int rc = CLIFFS_OK;
    hipError_t err = hipSuccess;
    if ((rc = ensemble_take(e)) == CLIFFS_OK && (rc = ensemble_args(e)) == CLIFFS_OK) {
        err = hipMemcpyAsync(dij, I, sizeof(int) * nij, hipMemcpyHostToDevice, e->st);
        if (err == hipSuccess) err = hipMemcpyAsync(dij + nij, J, sizeof(int) * nij, hipMemcpyHostToDevice, e->st);
        // `every` is even: after each group of steps the state is again the
        // copy the slot-0 argument blocks read
        for (int k = 0; k < nsamples && err == hipSuccess; k++) {
            for (int s = 0; s < every / 2 && err == hipSuccess; s++) err = hipGraphLaunch(e->gexec, e->st);
            if (err == hipSuccess) err = launch_ensemble_gauges(e->dargs, nm, dij, n, nsamples, k, ds, e->st);
        }
        for (cliffs_handle* h : e->m) {
            h->fn += every * nsamples;
            h->flags_valid = false;
        }
        if (err == hipSuccess) err = hipMemcpyAsync(eta, ds, sizeof(float) * nser, hipMemcpyDeviceToHost, e->st);
        if (err == hipSuccess) err = hipStreamSynchronize(e->st);
        if (err == hipSuccess) rc = ensemble_give(e);
    }
    (void)hipStreamSynchronize(e->st);
    (void)hipFree(dij);
    (void)hipFree(ds);
    if (rc) return rc;
    if (err != hipSuccess) return hipfail(err, "ensemble gauge series");
    return CLIFFS_OK;
}

int cliffs_ensemble_sync(cliffs_ensemble* e) {
    if (!e) return fail(CLIFFS_ERR_ARG, "NULL ensemble");
    HIPCHK(hipSetDevice(e->dev));
    HIPCHK(hipStreamSynchronize(e->st));
    HIPCHK(hipGetLastError());
    return CLIFFS_OK;
}

int cliffs_gauges(cliffs_handle* h, int n, const int* I, const int* J, double* eta, double* u,
                  double* v) {
    if (!h || n < 0 || (n > 0 && (!I || !J || !eta || !u || !v)))
        return fail(CLIFFS_ERR_ARG, "bad gauge arguments");
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipStreamSynchronize(h->st));
    const DevGrid g = h->grid();
    for (int q = 0; q < n; q++) {
        if (I[q] < 1 || I[q] > h->nx || J[q] < 1 || J[q] > h->ny)
            return fail(CLIFFS_ERR_ARG, "gauge index out of grid");
        const int i = I[q] - 1, j = J[q] - 1;
        const size_t o = hidx(h, i, j);
        float cp, uu, vv;
        HIPCHK(hipMemcpy(&cp, g.CP + o, sizeof(float), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&uu, g.U + o, sizeof(float), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&vv, g.V + o, sizeof(float), hipMemcpyDeviceToHost));
        const size_t c = (size_t)i + (size_t)h->nx * j;
        const double cel = to_cel(h->s[c], cp, h->dep[c], h->k);
        eta[q] = cel > h->k.celmin ? cel * cel / h->k.grav - h->dep[c] : (double)(float)-1.e32;
        u[q] = uu;
        v[q] = vv;
    }
    return CLIFFS_OK;
}

// ---- one-way nesting -------------------------------------------------------

int cliffs_nest_maps(int nxp, const double* xp, int nyp, const double* yp, int nxc,
                     const double* xc, int nyc, const double* yc, int* we_ny, double* we_cy,
                     int* sn_nx, double* sn_cx, int* w_nx, double* w_cx, int* e_nx, double* e_cx,
                     int* s_ny, double* s_cy, int* n_ny, double* n_cy) {
    if (nxp < 2 || nyp < 2 || nxc < 1 || nyc < 1 || !xp || !yp || !xc || !yc || !we_ny ||
        !we_cy || !sn_nx || !sn_cx || !w_nx || !w_cx || !e_nx || !e_cx || !s_ny || !s_cy ||
        !n_ny || !n_cy)
        return fail(CLIFFS_ERR_ARG, "bad nest map arguments (nests need a 2-D parent)");
    // containment check of NestedGrid.f:84-96
    if (xc[0] < xp[0] || xc[nxc - 1] > xp[nxp - 1] || yc[0] < yp[0] || yc[nyc - 1] > yp[nyp - 1])
        return fail(CLIFFS_ERR_ARG, "child grid not within the parent grid limits");
    pinpoint(yp, nyp, yc[0], s_ny, s_cy);
    pinpoint(yp, nyp, yc[nyc - 1], n_ny, n_cy);
    for (int i = 0; i < nxc; i++) pinpoint(xp, nxp, xc[i], &sn_nx[i], &sn_cx[i]);
    pinpoint(xp, nxp, xc[0], w_nx, w_cx);
    pinpoint(xp, nxp, xc[nxc - 1], e_nx, e_cx);
    for (int i = 0; i < nyc; i++) pinpoint(yp, nyp, yc[i], &we_ny[i], &we_cy[i]);
    return CLIFFS_OK;
}

int cliffs_nest_create(cliffs_handle* h, int nxc, int nyc, const int* we_ny, const double* we_cy,
                       const int* sn_nx, const double* sn_cx, int w_nx, double w_cx, int e_nx,
                       double e_cx, int s_ny, double s_cy, int n_ny, double n_cy,
                       cliffs_nest** out) {
    if (!h || !out || nxc < 1 || nyc < 1 || !we_ny || !we_cy || !sn_nx || !sn_cx)
        return fail(CLIFFS_ERR_ARG, "bad nest arguments");
    *out = nullptr;
    auto okx = [&](int i) { return i >= 1 && i <= h->nx - 1; };
    auto oky = [&](int j) { return j >= 1 && j <= h->ny - 1; };
    for (int q = 0; q < nyc; q++)
        if (!oky(we_ny[q])) return fail(CLIFFS_ERR_ARG, "we_ny out of the parent grid");
    for (int q = 0; q < nxc; q++)
        if (!okx(sn_nx[q])) return fail(CLIFFS_ERR_ARG, "sn_nx out of the parent grid");
    if (!okx(w_nx) || !okx(e_nx) || !oky(s_ny) || !oky(n_ny))
        return fail(CLIFFS_ERR_ARG, "edge line index out of the parent grid");
    HIPCHK(hipSetDevice(h->dev));
    cliffs_nest* n = new (std::nothrow) cliffs_nest;
    if (!n) return fail(CLIFFS_ERR_NOMEM, "host allocation");
    n->parent = h;
    n->nxc = nxc;
    n->nyc = nyc;
    n->w_nx = w_nx; n->e_nx = e_nx; n->s_ny = s_ny; n->n_ny = n_ny;
    n->w_cx = w_cx; n->e_cx = e_cx; n->s_cy = s_cy; n->n_cy = n_cy;
    n->ii1 = sn_nx[0];
    n->ii2 = sn_nx[nxc - 1] + 1;
    n->kk1 = we_ny[0];
    n->kk2 = we_ny[nyc - 1] + 1;
    n->nuvq = (size_t)6 * nyc + (size_t)6 * nxc;
    if (hipMalloc(&n->we_ny, sizeof(int) * nyc) != hipSuccess ||
        hipMalloc(&n->sn_nx, sizeof(int) * nxc) != hipSuccess ||
        hipMalloc(&n->we_cy, sizeof(double) * nyc) != hipSuccess ||
        hipMalloc(&n->sn_cx, sizeof(double) * nxc) != hipSuccess ||
        hipMalloc(&n->flag, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&n->uvq, sizeof(double) * n->nuvq) != hipSuccess) {
        cliffs_nest_destroy(n);
        return fail(CLIFFS_ERR_NOMEM, "nest device allocation");
    }
    // on the parent's (non-blocking) stream, which runs the nest's kernels
    if (hipMemcpyAsync(n->we_ny, we_ny, sizeof(int) * nyc, hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(n->sn_nx, sn_nx, sizeof(int) * nxc, hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(n->we_cy, we_cy, sizeof(double) * nyc, hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipMemcpyAsync(n->sn_cx, sn_cx, sizeof(double) * nxc, hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipStreamSynchronize(h->st) != hipSuccess) {
        cliffs_nest_destroy(n);
        return fail(CLIFFS_ERR_HIP, "nest map upload");
    }
    n->huvq.resize(n->nuvq);
    h->users++;
    n->counted = true;
    *out = n;
    return CLIFFS_OK;
}

int cliffs_nest_destroy(cliffs_nest* n) {
    if (!n) return CLIFFS_OK;
    if (n->parent) {
        (void)hipSetDevice(n->parent->dev);
        if (n->parent->st) (void)hipStreamSynchronize(n->parent->st);
        if (n->counted) n->parent->users--;
    }
    if (n->child) {
        if (n->child->st) (void)hipStreamSynchronize(n->child->st);
        n->child->users--;
    }
    for (auto e : n->rec_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto e : n->use_ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(n->ring);
    (void)hipFree(n->cedges);
    (void)hipFree(n->dwave);
    (void)hipFree(n->we_ny);
    (void)hipFree(n->sn_nx);
    (void)hipFree(n->we_cy);
    (void)hipFree(n->sn_cx);
    (void)hipFree(n->flag);
    (void)hipFree(n->uvq);
    delete n;
    return CLIFFS_OK;
}

}  // extern "C"

NestArgs cliffs_nest::args(double cuke) const {
    const cliffs_handle* h = parent;
    NestArgs a{};
    a.g = h->grid();
    a.p = h->p;
    a.grav = h->k.grav;
    a.nxc = nxc;
    a.nyc = nyc;
    a.we_ny = we_ny;
    a.we_cy = we_cy;
    a.sn_nx = sn_nx;
    a.sn_cx = sn_cx;
    a.w_nx = w_nx; a.e_nx = e_nx; a.s_ny = s_ny; a.n_ny = n_ny;
    a.w_cx = w_cx; a.e_cx = e_cx; a.s_cy = s_cy; a.n_cy = n_cy;
    a.ii1 = ii1; a.ii2 = ii2; a.kk1 = kk1; a.kk2 = kk2;
    a.gthr = (float)(h->k.grav * ((double)0.1f * cuke));  // feed_children_grid.f:30
    a.cuke = (float)cuke;
    a.flag = flag;
    a.uvq = uvq;
    return a;
}

namespace {

// Arrival test of feed_children_grid.f:26-46 (until the latch is set; the
// host reads the flag), on the parent's stream.
int nest_arrival(cliffs_nest* n, const NestArgs& a, int* iarv) {
    cliffs_handle* h = n->parent;
    if (*iarv) return CLIFFS_OK;
    HIPCHK(hipMemsetAsync(n->flag, 0, sizeof(unsigned), h->st));
    hipError_t e = launch_nest_arrival(a, h->st);
    if (e != hipSuccess) return hipfail(e, "nest_arrival");
    unsigned f = 0;
    HIPCHK(hipMemcpyAsync(&f, n->flag, sizeof(unsigned), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    *iarv = f ? 1 : 0;
    return CLIFFS_OK;
}

}  // namespace

extern "C" {

int cliffs_nest_feed(cliffs_nest* n, double cuke, double hmin, int* iarv, int* fed,
                     double* uvq_w, double* uvq_e, double* uvq_s, double* uvq_n) {
    if (!n || !iarv || !fed || !uvq_w || !uvq_e || !uvq_s || !uvq_n)
        return fail(CLIFFS_ERR_ARG, "NULL argument");
    cliffs_handle* h = n->parent;
    if (!h->state) return fail(CLIFFS_ERR_STATE, "no state uploaded");
    // the driver passes `ground` (cliffs_main.f:297, 362); the device decides
    // dep > ground exactly, any other threshold is not supported
    if (hmin != h->k.ground) return fail(CLIFFS_ERR_ARG, "hmin must equal the grid's ground");
    HIPCHK(hipSetDevice(h->dev));
    const NestArgs a = n->args(cuke);
    *fed = 0;
    int rc;
    if ((rc = nest_arrival(n, a, iarv))) return rc;
    if (!*iarv) return CLIFFS_OK;
    hipError_t e = launch_nest_feed(a, h->st);
    if (e != hipSuccess) return hipfail(e, "nest_feed");
    HIPCHK(hipMemcpyAsync(n->huvq.data(), n->uvq, sizeof(double) * n->nuvq, hipMemcpyDeviceToHost,
                          h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    const size_t ly = (size_t)3 * n->nyc, lx = (size_t)3 * n->nxc;
    std::memcpy(uvq_w, n->huvq.data(), sizeof(double) * ly);
    std::memcpy(uvq_e, n->huvq.data() + ly, sizeof(double) * ly);
    std::memcpy(uvq_s, n->huvq.data() + 2 * ly, sizeof(double) * lx);
    std::memcpy(uvq_n, n->huvq.data() + 2 * ly + lx, sizeof(double) * lx);
    *fed = 1;
    return CLIFFS_OK;
}

int cliffs_nest_couple(cliffs_nest* n, cliffs_handle* child, int capacity) {
    if (!n || !child || capacity < 2) return fail(CLIFFS_ERR_ARG, "bad coupling arguments");
    if (n->child) return fail(CLIFFS_ERR_STATE, "nest already coupled");
    if (child->nx != n->nxc || child->ny != n->nyc)
        return fail(CLIFFS_ERR_ARG, "child grid size differs from the nest maps");
    if (child->dev != n->parent->dev) return fail(CLIFFS_ERR_ARG, "child on another device");
    if (!child->xrun || !child->yrun) return fail(CLIFFS_ERR_ARG, "child must be a 2-D grid");
    HIPCHK(hipSetDevice(child->dev));
    n->cap = capacity;
    n->rec_ev.assign(capacity, nullptr);
    n->use_ev.assign(capacity, nullptr);
    for (int k = 0; k < capacity; k++) {
        HIPCHK(hipEventCreateWithFlags(&n->rec_ev[k], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&n->use_ev[k], hipEventDisableTiming));
        HIPCHK(hipEventRecord(n->use_ev[k], child->st));
    }
    if (hipMalloc(&n->ring, sizeof(double) * n->nuvq * capacity) != hipSuccess ||
        hipMalloc(&n->dwave, sizeof(double) * capacity) != hipSuccess ||
        hipMalloc(&n->cedges, sizeof(float) * (6 * (size_t)n->nyc + 6 * (size_t)n->nxc)) != hipSuccess)
        return fail(CLIFFS_ERR_NOMEM, "nest ring allocation");
    n->child = child;
    child->users++;
    n->tinput.clear();
    n->irec = 1;
    n->bndr = true;
    n->ended = false;
    n->start_cuke = 0.0;
    n->start_set = false;
    n->start = 0;
    n->scanned = 0;
    n->run_k = 0;
    n->run_kc = 0;
    n->waited = 0;
    return CLIFFS_OK;
}

}  // extern "C"

namespace {

// Start rule of get_boundary_feed.f:144-167 over the records written so
// far: the first record whose largest edge |eta| reaches start_cuke, and an
// error when the feed ended without one before its last record (it >= nTn).
int nest_find_start(cliffs_nest* n) {
    if (n->start > 0) return CLIFFS_OK;
    const int nrec = (int)n->tinput.size();
    while (n->scanned < nrec) {
        const int k = n->scanned + 1, slot = (k - 1) % n->cap;
        double w = 0.0;
        HIPCHK(hipEventSynchronize(n->rec_ev[slot]));
        HIPCHK(hipMemcpy(&w, n->dwave + slot, sizeof(double), hipMemcpyDeviceToHost));
        if (w >= n->start_cuke) {
            if (n->ended && k >= nrec)
                return fail(CLIFFS_ERR_STATE, "only the feed's last record reaches the child's threshold");
            n->start = k;
            n->irec = std::max(n->irec, k);
            return CLIFFS_OK;
        }
        n->scanned = k;
    }
    return fail(CLIFFS_ERR_STATE, n->ended ? "feed ended below the child's threshold"
                                           : "no feed record reaches the child's threshold yet");
}

}  // namespace

extern "C" {

int cliffs_nest_record(cliffs_nest* n, double cuke, double hmin, double t, int* iarv, int* nrec) {
    if (!n || !iarv || !nrec) return fail(CLIFFS_ERR_ARG, "NULL argument");
    if (!n->child) return fail(CLIFFS_ERR_STATE, "nest not coupled to a child");
    cliffs_handle* h = n->parent;
    if (!h->state) return fail(CLIFFS_ERR_STATE, "no state uploaded");
    if (hmin != h->k.ground) return fail(CLIFFS_ERR_ARG, "hmin must equal the grid's ground");
    if (n->ended) return fail(CLIFFS_ERR_STATE, "feed already ended");
    HIPCHK(hipSetDevice(h->dev));
    NestArgs a = n->args(cuke);
    int rc;
    if ((rc = nest_arrival(n, a, iarv))) return rc;
    *nrec = (int)n->tinput.size();
    if (!*iarv) return CLIFFS_OK;
    const int k = (int)n->tinput.size() + 1;  // 1-based record being written
    // the slot's previous record (k - cap) must be past the child's need:
    // the child reads records irec-1 and irec from its start record on;
    // before the start is known, records found below the threshold are free
    const int need_lo = n->start > 0 ? std::max(n->irec - 1, n->start) : n->scanned + 1;
    if (k > n->cap && k - n->cap >= need_lo)
        return fail(CLIFFS_ERR_STATE, "parent feed too far ahead of the child (raise the capacity)");
    const int slot = (k - 1) % n->cap;
    a.uvq = n->ring + (size_t)slot * n->nuvq;
    HIPCHK(hipStreamWaitEvent(h->st, n->use_ev[slot], 0));
    hipError_t e = launch_nest_feed(a, h->st);
    if (e != hipSuccess) return hipfail(e, "nest_feed");
    e = launch_record_wave(a.uvq, n->nxc, n->nyc, n->dwave + slot, h->st);
    if (e != hipSuccess) return hipfail(e, "record_wave");
    HIPCHK(hipEventRecord(n->rec_ev[slot], h->st));
    n->tinput.push_back(t);
    *nrec = k;
    // with the child's threshold known and its start not yet found, check
    // this record now, so that records below the threshold leave the ring
    if (n->start_set && n->start == 0) {
        const int rc2 = nest_find_start(n);
        if (rc2 && rc2 != CLIFFS_ERR_STATE) return rc2;
    }
    return CLIFFS_OK;
}

int cliffs_nest_end_feed(cliffs_nest* n) {
    if (!n || !n->child) return fail(CLIFFS_ERR_STATE, "nest not coupled to a child");
    n->ended = true;
    return CLIFFS_OK;
}

}  // extern "C"


extern "C" {

int cliffs_nest_child_start(cliffs_nest* n, double cuke, double* t0) {
    if (!n || !t0 || !(cuke >= 0.0)) return fail(CLIFFS_ERR_ARG, "bad child start arguments");
    if (!n->child) return fail(CLIFFS_ERR_STATE, "nest not coupled to a child");
    if (n->start_set && cuke != n->start_cuke)
        return fail(CLIFFS_ERR_STATE, "the child's threshold is already set");
    if (!n->start_set && n->start > 0)
        return fail(CLIFFS_ERR_STATE, "the child already started from the first record");
    HIPCHK(hipSetDevice(n->child->dev));
    n->start_cuke = cuke;
    n->start_set = true;
    const int rc = nest_find_start(n);
    if (rc) return rc;
    *t0 = n->tinput[n->start - 1];
    return CLIFFS_OK;
}

int cliffs_nest_child_step(cliffs_nest* n, double t) {
    if (!n || !n->child) return fail(CLIFFS_ERR_STATE, "nest not coupled to a child");
    cliffs_handle* c = n->child;
    if (!c->state) return fail(CLIFFS_ERR_STATE, "child state not uploaded");
    const int nrec = (int)n->tinput.size();
    if (nrec < 1) return fail(CLIFFS_ERR_STATE, "no feed record yet");
    HIPCHK(hipSetDevice(c->dev));
    int rc0;
    if ((rc0 = nest_find_start(n))) return rc0;  // threshold 0 unless set: the first record
    if (t < n->tinput[n->start - 1]) return fail(CLIFFS_ERR_ARG, "child time before its start record");
    // cliffs_main.f:305-315: tinput(irec-1) <= t < tinput(irec)
    while (n->bndr && t >= n->tinput[n->irec - 1]) {
        if (n->irec == nrec) {
            if (!n->ended) return fail(CLIFFS_ERR_STATE, "child ahead of the parent feed");
            n->bndr = false;  // feed ended: edges 0 from now on
        } else {
            // record irec-1 leaves the child's use: its slot is free for the
            // parent once the child work enqueued so far has run
            if (n->irec >= 2) HIPCHK(hipEventRecord(n->use_ev[(n->irec - 2) % n->cap], c->st));
            n->irec++;
        }
    }
    const int zero = n->bndr ? 0 : 1;
    double cc = 0.0;
    const double *r0 = n->ring, *r1 = n->ring;
    if (!zero) {
        if (n->irec < 2) return fail(CLIFFS_ERR_STATE, "child time before the second feed record");
        cc = (n->tinput[n->irec - 1] - t) / (n->tinput[n->irec - 1] - n->tinput[n->irec - 2]);
        const int s0 = (n->irec - 2) % n->cap, s1 = (n->irec - 1) % n->cap;
        r0 = n->ring + (size_t)s0 * n->nuvq;
        r1 = n->ring + (size_t)s1 * n->nuvq;
        // the newest record needed is written (once per record: the parent
        // writes records in order, so the older one is written too)
        if (n->waited != n->irec) {
            HIPCHK(hipStreamWaitEvent(c->st, n->rec_ev[s1], 0));
            n->waited = n->irec;
        }
    }
    hipError_t e = launch_child_edges(r0, r1, cc, 1.0 - cc, c->k.grav, n->nxc, n->nyc, zero,
                                      n->cedges, n->cedges + 6 * (size_t)n->nyc, c->st);
    if (e != hipSuccess) return hipfail(e, "child_edges");
    return do_step(c, n->cedges, n->cedges + 6 * (size_t)n->nyc);
}

int cliffs_nest_run(cliffs_nest* n, int nrec, int every, double t0, double dt, int per,
                    double dt_child, double cuke, double hmin, int* iarv) {
    if (!n || !iarv || nrec < 0 || every < 1 || per < 1 || !(dt > 0.0) || !(dt_child > 0.0))
        return fail(CLIFFS_ERR_ARG, "bad coupled-run arguments");
    if (!n->child) return fail(CLIFFS_ERR_STATE, "nest not coupled to a child");
    for (int q = 0; q < nrec; q++) {
        int rc = cliffs_run(n->parent, every);
        if (rc) return rc;
        n->run_k++;
        const double t = t0 + (double)(n->run_k * every) * dt;
        int nr = 0;
        if ((rc = cliffs_nest_record(n, cuke, hmin, t, iarv, &nr))) return rc;
        if (nr < 2) continue;
        // the child's clock starts at its start record (the first one
        // without a threshold); until that is known the child waits
        if (n->start == 0) {
            rc = nest_find_start(n);
            if (rc == CLIFFS_ERR_STATE) continue;
            if (rc) return rc;
        }
        const long s = n->start - 1, have = (long)n->tinput.size() - s;
        // step kc at tinput(s + j) + (kc - j per) dt_child, j = kc / per,
        // while the records bracketing the next step's time exist
        while (n->run_kc + 1 < (long)per * (have - 1)) {
            const long kc = ++n->run_kc, j = kc / per;
            if ((rc = cliffs_nest_child_step(n, n->tinput[s + j] + (double)(kc - j * per) * dt_child)))
                return rc;
        }
    }
    return CLIFFS_OK;
}

int cliffs_sync(cliffs_handle* h) {
    if (!h) return fail(CLIFFS_ERR_ARG, "NULL handle");
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipGetLastError());
    return CLIFFS_OK;
}

int cliffs_stream(cliffs_handle* h, void** stream) {
    if (!h || !stream) return fail(CLIFFS_ERR_ARG, "NULL argument");
    *stream = (void*)h->st;
    return CLIFFS_OK;
}

int cliffs_info(cliffs_handle* h, double* grav, double* celmin, int* ii1, int* ii2, int* jj1,
                int* jj2) {
    if (!h) return fail(CLIFFS_ERR_ARG, "NULL handle");
    if (grav) *grav = h->k.grav;
    if (celmin) *celmin = h->k.celmin;
    if (ii1) *ii1 = h->ii1 + 1;
    if (ii2) *ii2 = h->ii2 + 1;
    if (jj1) *jj1 = h->jj1 + 1;
    if (jj2) *jj2 = h->jj2 + 1;
    return CLIFFS_OK;
}

}  // extern "C"
