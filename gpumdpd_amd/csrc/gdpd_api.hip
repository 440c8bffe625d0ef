// gdpd_api.hip -- host runtime and C ABI of libgdpd (see include/gdpd.h).
//
// One gdpd_ctx = one device + one HIP stream (default priority).  Each system
// ("replica") owns its device arrays; a device array of RepDev descriptors,
// indexed by system handle, is what the kernels read, so one launch can cover
// the bricks of many replicas.  Per-step scalars (step counter, buffer parity)
// live on the device, which makes a captured step graph replayable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gdpd.h"
#include "gdpd_internal.h"
#include "gdpd_kernels.h"

using namespace gdpd;

namespace {

// LDS of one gfx950 CU, and the share of each of four resident workgroups
constexpr size_t kLdsPerCu = 160 * 1024;
constexpr size_t kLdsFourPerCu = kLdsPerCu / 4;

// Steps between physical re-sorts of the bead arrays into cell order.  A
// fresher order keeps the bricks' gathers local (force pass at 10M beads:
// 1.52 ms re-sorting every 16 steps, 1.48 every 8, 1.44 every 4), while each
// re-sort costs about 0.4 ms there; 6 to 10 balance the two.
constexpr int kDefaultResortEvery = 8;
// automatic brick shapes of small systems shrink until a launch has this many bricks
constexpr long long kMinBricks = 400;

// ---- allocation-free thermo: scratch in a static device array ----------------
// gdpd_thermo_noalloc keeps its per-chunk partial sums and their total here
// instead of in device memory allocated per call.  4 MiB: kThermoParts doubles
// per 256-bead chunk, for systems of up to kNoAllocMaxChunks * 256 =
// 13,421,568 beads, then the kThermoParts-double total at the end.  Every
// call on a device shares that device's copy, which keeps the last call's
// contents; calls that overlap on one device race.  Host code takes the
// array's address per device with hipGetSymbolAddress at context creation
// and passes that address to the kernels; it never uses the name itself.
constexpr size_t kScratchDoubles = 512 * 1024;
constexpr int kNoAllocMaxChunks = (int)(kScratchDoubles / kThermoParts) - 1;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
double *g_scratch_dev[kMaxDevices] = {};  // per device: the address of its copy of scratch
std::mutex g_scratch_mu;

struct Tables {
    // device work tables for one set of systems
    WorkItem *beads = nullptr, *cells = nullptr, *scan = nullptr, *bricks_dpd = nullptr,
             *bricks_mdpd = nullptr;
    int *rep_chunk_off = nullptr, *rep_list = nullptr, *chunk_sums = nullptr;
    int n_beads = 0, n_cells = 0, n_scan = 0, n_dpd = 0, n_mdpd = 0, n_reps = 0;
    BrickLaunch bl_dpd{}, bl_mdpd{};
    int precision = 0;
    std::vector<int> key;
    // defer: where to leave the buffers instead of freeing them now (hipFree
    // waits for the whole device, including other streams' queued work)
    void release(std::vector<void *> *defer = nullptr)
    {
        void *ps[] = {beads, cells, scan, bricks_dpd, bricks_mdpd, rep_chunk_off, rep_list, chunk_sums};
        for (void *p : ps)
            if (p) {
                if (defer) defer->push_back(p);
                else (void)hipFree(p);
            }
        *this = Tables();
    }
};

struct Graph {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int steps = 0;
    int topo = -1;  // bond/angle level the captured integrate kernel was chosen for
    void release()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        steps = 0;
        topo = -1;
    }
};

struct Sys {
    gdpd_system_desc desc{};
    RepDev h{};                 // host copy of the device descriptor
    std::vector<void *> allocs; // device allocations owned by this system
    bool uploaded = false, prepared = false, cells_valid = false, have_pair = false;
    int style = GDPD_STYLE_NONE;
    double temp = 1.0, rc = 1.0, rd = 0.75;
    uint32_t seed = 0;
    std::vector<double> coeff;  // packed, as given
    std::vector<double> masses;
    int tune_b[3] = {0, 0, 0};
    int resort_every = 0, graph_steps = 0;
    int steps_since_resort = 0;
    bool midstep = false;       // between a zero_step/advance and finish
    long long step_host = 0;    // the device step counter once the enqueued work is done
    int cap_s = 0, cap_o = 0;   // LDS capacities set by the caller (0: from the state)
    int regrows = 0;            // LDS capacity increases after a staging overflow
    Tables tab;                 // single-system work tables
    Graph graph;                // single-system step graph
    // reference neighbour list
    int nl_nmax = 0, nl_nb[3] = {0, 0, 0};
    int *nl_binof = nullptr, *nl_count = nullptr, *nl_start = nullptr, *nl_fill = nullptr,
        *nl_content = nullptr, *nl_n = nullptr, *nl_neigh = nullptr, *nl_over = nullptr,
        *nl_flag = nullptr;
    double *nl_last = nullptr;
    bool nl_built = false;
    int64_t bytes = 0;
    // group fixes and computes (the given fixes, then an nve on every type when
    // none is given) and the host mirror of the step counter on their path
    std::vector<gdpd_fix> fixes;
    std::vector<gdpd_compute> computes;
    int type_count[kMaxTypes] = {0};
    bool fix_path = false;
    long long fix_step = 0;
    bool admitted = false;  // created inside the current admission window
};

} // namespace

struct gdpd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;  // the stream calls enqueue on: main, or side while admitting
    hipStream_t main_stream = nullptr;
    // admission window (gdpd_admission): new systems are set up on a second
    // stream while the main stream keeps stepping the others
    hipStream_t side_stream = nullptr;
    hipEvent_t side_done = nullptr;
    bool admitting = false, admit_refused = false;
    // device buffers released inside an admission window, freed at the next
    // point where the host has waited for the main stream anyway
    std::vector<void *> graveyard;
    std::string err;
    std::vector<Sys *> sys;
    RepDev *d_reps = nullptr;
    int d_reps_cap = 0;
    // kernel timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    int64_t force_launches = 0;
    // phase clocks of the timed force launches: 4 sums per work item and wave
    unsigned *d_phase = nullptr;
    size_t phase_items = 0;        // items the buffer holds
    double phase_wave_launches = 0; // sum over timed launches of their waves
    // batch tables cache
    Tables batch;
    Graph batch_graph;
    int max_lds = 65536;
    // systems advanced by enqueued work whose halt flags have not been read yet
    std::vector<int> pending;
    // gather buffers (states, step-time rings) of listed systems
    int *d_ids = nullptr;
    RepState *d_states = nullptr;
    unsigned long long *d_ts = nullptr;
    long long *d_tsteps = nullptr;
    int gather_cap = 0;
    double tick_ns = 10.0;  // device wall clock period
};

namespace {

int fail(gdpd_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) {
        c->err = buf;
        if (c->admit_refused)
            c->err += " (inside an admission window only systems created in it can be used)";
        c->admit_refused = false;
    }
    return code;
}

#define HIPCHK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(ctx, GDPD_ERR_HIP, "%s failed: %s (%s:%d)", #call,                    \
                        hipGetErrorString(e_), __FILE__, __LINE__);                           \
    } while (0)

template <typename T> int dalloc(gdpd_ctx *ctx, Sys *s, T **p, size_t n)
{
    size_t bytes = sizeof(T) * (n > 0 ? n : 1);
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(ctx, GDPD_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    if (s) {
        s->allocs.push_back(*p);
        s->bytes += (int64_t)bytes;
    }
    return GDPD_OK;
}

Sys *get_sys(gdpd_ctx *ctx, int id)
{
    if (!ctx || id < 0 || id >= (int)ctx->sys.size()) return nullptr;
    Sys *s = ctx->sys[id];
    if (s && ctx->admitting && !s->admitted) {
        ctx->admit_refused = true;
        return nullptr;
    }
    return s;
}

// frees the buffers left by admission windows; call only where the host has
// just waited for the main stream
void flush_graveyard(gdpd_ctx *ctx)
{
    if (ctx->admitting) return;
    for (void *p : ctx->graveyard) (void)hipFree(p);
    ctx->graveyard.clear();
}

int push_desc(gdpd_ctx *ctx, int id)
{
    if (id >= ctx->d_reps_cap) {
        int cap = std::max(16, ctx->d_reps_cap * 2);
        while (cap <= id) cap *= 2;
        RepDev *nd = nullptr;
        HIPCHK(hipMalloc((void **)&nd, sizeof(RepDev) * cap));
        if (ctx->d_reps) {
            HIPCHK(hipMemcpyAsync(nd, ctx->d_reps, sizeof(RepDev) * ctx->d_reps_cap,
                                  hipMemcpyDeviceToDevice, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            HIPCHK(hipFree(ctx->d_reps));
        }
        ctx->d_reps = nd;
        ctx->d_reps_cap = cap;
        // every cached graph and table points at the old array
        for (Sys *s : ctx->sys)
            if (s) s->graph.release();
        ctx->batch_graph.release();
        ctx->batch.release();
    }
    HIPCHK(hipMemcpyAsync(ctx->d_reps + id, &ctx->sys[id]->h, sizeof(RepDev), hipMemcpyHostToDevice,
                          ctx->stream));
    return GDPD_OK;
}

// ---- pair coefficient tables -------------------------------------------------

int packed_index(int ti, int tj, int nt)
{
    int a = std::min(ti, tj), b = std::max(ti, tj);
    return ((a - 1) * nt - (a - 1) * (a - 2) / 2 + b - a + 1) - 1;
}

template <typename real> void build_coef(Sys *s, std::vector<PairCoef<real>> &out, double &rc2max,
                                         double &rd2max)
{
    const int nt = s->desc.n_types;
    out.assign((size_t)nt * nt, PairCoef<real>{});
    // force/pair_dpd.cu:577-578 and :315, in the working precision
    const real dtinv_1_2 = (real)1.0 / std::sqrt((real)s->desc.dt);
    const real dt_1_2 = std::sqrt((real)2.0 * (real)s->temp) * dtinv_1_2;
    const real pi = (real)3.14159265358979323846;
    rc2max = 0.0;
    rd2max = 0.0;
    for (int ti = 1; ti <= nt; ++ti)
        for (int tj = 1; tj <= nt; ++tj) {
            const int idx = packed_index(ti, tj, nt);
            PairCoef<real> c{};
            if (s->style == GDPD_STYLE_MDPD) {
                c.A = (real)s->coeff[idx * 5 + 0];
                c.B = (real)s->coeff[idx * 5 + 1];
                c.gamma = (real)s->coeff[idx * 5 + 2];
                c.rc = (real)s->coeff[idx * 5 + 3];
                c.rd = (real)s->coeff[idx * 5 + 4];
            } else {
                c.A = (real)s->coeff[idx * 3 + 0];
                c.B = (real)0;
                c.gamma = (real)s->coeff[idx * 3 + 1];
                c.rc = (real)s->coeff[idx * 3 + 2];
                c.rd = c.rc;
            }
            c.rc2 = c.rc * c.rc;
            c.rcinv = (real)1 / c.rc;
            c.rdinv = (real)1 / c.rd;
            c.sig = std::sqrt(c.gamma) * dt_1_2;
            // force/pair_mdpd.cu:275-276
            c.rhofac = (real)15.0 / (real)2.0 / pi / (c.rd * c.rd * c.rd);
            out[(size_t)(ti - 1) * nt + (tj - 1)] = c;
            rc2max = std::max(rc2max, (double)c.rc2);
            rd2max = std::max(rd2max, (double)(c.rd * c.rd));
        }
}

double max_rc(Sys *s)
{
    const int nt = s->desc.n_types, np = nt * (nt + 1) / 2;
    double m = 0.0;
    for (int k = 0; k < np; ++k)
        m = std::max(m, s->style == GDPD_STYLE_MDPD ? s->coeff[k * 5 + 3] : s->coeff[k * 3 + 2]);
    return m;
}

// ---- work tables -------------------------------------------------------------

int upload_items(gdpd_ctx *ctx, const std::vector<WorkItem> &v, WorkItem **d)
{
    if (v.empty()) return GDPD_OK;
    HIPCHK(hipMalloc((void **)d, sizeof(WorkItem) * v.size()));
    HIPCHK(hipMemcpyAsync(*d, v.data(), sizeof(WorkItem) * v.size(), hipMemcpyHostToDevice, ctx->stream));
    return GDPD_OK;
}

// largest run-list size over the system's bricks: full bricks and the
// shorter ones at the upper end of each axis
int max_run_entries(const RepDev &h)
{
    int e[3][2], ne[3];
    for (int d = 0; d < 3; ++d) {
        ne[d] = 0;
        e[d][ne[d]++] = std::max(1, std::min(h.bdim[d], h.nc[d]));
        const int rem = h.nc[d] % std::max(1, h.bdim[d]);
        if (rem > 0 && rem != e[d][0]) e[d][ne[d]++] = rem;
    }
    int best = 1;
    for (int a = 0; a < ne[0]; ++a)
        for (int b = 0; b < ne[1]; ++b)
            for (int c = 0; c < ne[2]; ++c) best = std::max(best, brick_run_entries(e[0][a], e[1][b], e[2][c]));
    return best;
}

BrickLaunch brick_params_of(const Sys *s)
{
    BrickLaunch b{};
    b.rmax = (s->h.bdim[0] + 2) * (s->h.bdim[1] + 2) * (s->h.bdim[2] + 2);
    b.smax = s->h.smax;
    b.omax = s->h.omax;
    b.nrun = max_run_entries(s->h);
    b.ntypes = s->h.ntypes;
    return b;
}

void bl_max(BrickLaunch &a, const BrickLaunch &b)
{
    a.rmax = std::max(a.rmax, b.rmax);
    a.smax = std::max(a.smax, b.smax);
    a.omax = std::max(a.omax, b.omax);
    a.nrun = std::max(a.nrun, b.nrun);
    a.ntypes = std::max(a.ntypes, b.ntypes);
}

int build_tables(gdpd_ctx *ctx, const std::vector<int> &ids, Tables &T)
{
    T.release(ctx->admitting ? &ctx->graveyard : nullptr);
    std::vector<WorkItem> beads, cells, scan, bd, bm;
    std::vector<int> off(1, 0), reps;
    for (int id : ids) {
        Sys *s = ctx->sys[id];
        const int N = s->h.N, nc = s->h.ncell;
        for (int k = 0; k < (N + kBlock - 1) / kBlock; ++k) beads.push_back({id, k});
        for (int k = 0; k < (nc + kBlock - 1) / kBlock; ++k) cells.push_back({id, k});
        const int nch = (nc + 4 * kBlock - 1) / (4 * kBlock);
        for (int k = 0; k < nch; ++k) scan.push_back({id, k});
        off.push_back(off.back() + nch);
        reps.push_back(id);
        auto &dst = (s->style == GDPD_STYLE_MDPD) ? bm : bd;
        for (int k = 0; k < s->h.nbrick; ++k) dst.push_back({id, k});
        BrickLaunch b = brick_params_of(s);
        bl_max(s->style == GDPD_STYLE_MDPD ? T.bl_mdpd : T.bl_dpd, b);
        T.precision = s->h.precision;
    }
    // a launch whose bricks all belong to one replica, in order, skips the item reads
    auto single = [](const std::vector<WorkItem> &v) {
        for (size_t k = 0; k < v.size(); ++k)
            if (v[k].rep != v[0].rep || v[k].local != (int)k) return -1;
        return v.empty() ? -1 : v[0].rep;
    };
    T.bl_dpd.single_rep = single(bd);
    T.bl_mdpd.single_rep = single(bm);
    int rc;
    if ((rc = upload_items(ctx, beads, &T.beads))) return rc;
    if ((rc = upload_items(ctx, cells, &T.cells))) return rc;
    if ((rc = upload_items(ctx, scan, &T.scan))) return rc;
    if ((rc = upload_items(ctx, bd, &T.bricks_dpd))) return rc;
    if ((rc = upload_items(ctx, bm, &T.bricks_mdpd))) return rc;
    HIPCHK(hipMalloc((void **)&T.rep_chunk_off, sizeof(int) * off.size()));
    HIPCHK(hipMemcpyAsync(T.rep_chunk_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMalloc((void **)&T.rep_list, sizeof(int) * reps.size()));
    HIPCHK(hipMemcpyAsync(T.rep_list, reps.data(), sizeof(int) * reps.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMalloc((void **)&T.chunk_sums, sizeof(int) * std::max<size_t>(1, scan.size())));
    T.n_beads = (int)beads.size();
    T.n_cells = (int)cells.size();
    T.n_scan = (int)scan.size();
    T.n_dpd = (int)bd.size();
    T.n_mdpd = (int)bm.size();
    T.n_reps = (int)reps.size();
    T.key = ids;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GDPD_OK;
}

// ---- kernel timing -------------------------------------------------------------

hipEvent_t next_event(gdpd_ctx *ctx)
{
    if (ctx->ev_used == ctx->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->ev_pool.push_back(e);
    }
    return ctx->ev_pool[ctx->ev_used++];
}

// ---- step sequences ------------------------------------------------------------

int check_launch(gdpd_ctx *ctx)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(ctx, GDPD_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    return GDPD_OK;
}

// rebin: keys for the current positions, then scan/scatter/in-cell sort
int rebin(gdpd_ctx *ctx, Tables &T)
{
    launch_bin_keys(ctx->d_reps, T.beads, T.n_beads, ctx->stream);
    launch_scan(ctx->d_reps, T.scan, T.n_scan, T.rep_chunk_off, T.rep_list, T.n_reps, T.chunk_sums, ctx->stream);
    launch_scatter(ctx->d_reps, T.beads, T.n_beads, ctx->stream);
    launch_cellsort(ctx->d_reps, T.cells, T.n_cells, ctx->stream);
    return check_launch(ctx);
}

// Physical re-sort of every system of T into cell order.  The device swaps the
// id, type and image arrays with their spares (k_resort_swap); the host copies
// of the descriptors swap with them, so later descriptor uploads agree.
int resort_tables(gdpd_ctx *ctx, Tables &T)
{
    launch_resort(ctx->d_reps, T.beads, T.n_beads, T.rep_list, T.n_reps, ctx->stream);
    for (int id : T.key) {
        Sys *s = (id >= 0 && id < (int)ctx->sys.size()) ? ctx->sys[id] : nullptr;
        if (!s) continue;
        std::swap(s->h.id, s->h.id_alt);
        std::swap(s->h.type, s->h.type_alt);
        std::swap(s->h.img, s->h.img_alt);
    }
    return check_launch(ctx);
}

// 0: no system of T has bonds or angles, 1: bonds only, 2: angles too
int topo_of(gdpd_ctx *ctx, const Tables &T)
{
    int t = 0;
    for (int id : T.key) {
        const Sys *s = (id >= 0 && id < (int)ctx->sys.size()) ? ctx->sys[id] : nullptr;
        if (!s) continue;
        if (s->h.nangles > 0) t = 2;
        else if (s->h.nbonds > 0) t = std::max(t, 1);
    }
    return t;
}

int force_launch(gdpd_ctx *ctx, Tables &T, int mode)
{
    T.bl_dpd.perbead = T.bl_mdpd.perbead = ((mode & kModeForceOnly) && (mode & kModeLogForce)) ? 1 : 0;
    const bool timed = ctx->timing;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (T.n_mdpd) launch_density(ctx->d_reps, T.bricks_mdpd, T.n_mdpd, T.bl_mdpd, T.precision, ctx->stream);
    if (timed) {
        e0 = next_event(ctx);
        e1 = next_event(ctx);
        if (!e0 || !e1) return fail(ctx, GDPD_ERR_HIP, "hipEventCreate failed");
        // phase clocks: the buffer holds the items of the largest timed launch
        const size_t items = (size_t)T.n_dpd + (size_t)T.n_mdpd;
        if (items > ctx->phase_items) {
            HIPCHK(hipStreamSynchronize(ctx->stream));
            std::vector<unsigned> keep(ctx->phase_items * 16);
            if (ctx->d_phase) {
                HIPCHK(hipMemcpy(keep.data(), ctx->d_phase, keep.size() * sizeof(unsigned), hipMemcpyDeviceToHost));
                (void)hipFree(ctx->d_phase);
                ctx->d_phase = nullptr;
            }
            keep.resize(items * 16, 0u);
            HIPCHK(hipMalloc((void **)&ctx->d_phase, keep.size() * sizeof(unsigned)));
            HIPCHK(hipMemcpy(ctx->d_phase, keep.data(), keep.size() * sizeof(unsigned), hipMemcpyHostToDevice));
            ctx->phase_items = items;
        }
        T.bl_dpd.phase = ctx->d_phase;
        T.bl_mdpd.phase = ctx->d_phase + (size_t)T.n_dpd * 16;
        ctx->phase_wave_launches += (double)items * (kBlock / 64);
        HIPCHK(hipEventRecord(e0, ctx->stream));
    } else {
        T.bl_dpd.phase = T.bl_mdpd.phase = nullptr;
    }
    if (T.n_dpd) launch_brick(ctx->d_reps, T.bricks_dpd, T.n_dpd, T.bl_dpd, T.precision, GDPD_STYLE_DPD, mode, ctx->stream);
    if (T.n_mdpd) launch_brick(ctx->d_reps, T.bricks_mdpd, T.n_mdpd, T.bl_mdpd, T.precision, GDPD_STYLE_MDPD, mode, ctx->stream);
    if (timed) {
        HIPCHK(hipEventRecord(e1, ctx->stream));
        ctx->force_launches++;
    }
    return check_launch(ctx);
}

// one fused step for every system in T
int step_once(gdpd_ctx *ctx, Tables &T, int mode)
{
    int rc;
    if ((rc = force_launch(ctx, T, kModeLogAuto))) return rc;
    launch_integrate(ctx->d_reps, T.beads, T.n_beads, mode | kModeLogAuto, topo_of(ctx, T), ctx->stream);
    launch_step_end(ctx->d_reps, T.rep_list, T.n_reps, 1, 1, 1, ctx->stream);
    if (mode & kModeDrift) {
        // the order of beads inside a cell is whatever the rank atomics gave:
        // every per-bead and per-brick sum is fixed point and each pair's
        // contribution is the same from either side, so results do not depend
        // on it
        launch_scan(ctx->d_reps, T.scan, T.n_scan, T.rep_chunk_off, T.rep_list, T.n_reps, T.chunk_sums, ctx->stream);
        launch_scatter(ctx->d_reps, T.beads, T.n_beads, ctx->stream);
    }
    return check_launch(ctx);
}

int read_state(gdpd_ctx *ctx, Sys *s, RepState *st)
{
    HIPCHK(hipMemcpyAsync(st, s->h.state, sizeof(RepState), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GDPD_OK;
}

int check_flags(gdpd_ctx *ctx, Sys *s)
{
    RepState st;
    int rc = read_state(ctx, s, &st);
    if (rc) return rc;
    if (st.flags & kFlagStageOverflow)
        return fail(ctx, GDPD_ERR_OVERFLOW, "a brick held more beads than its LDS capacity (smax=%d, omax=%d)",
                    s->h.smax, s->h.omax);
    if (st.flags & kFlagNonFinite)
        return fail(ctx, GDPD_ERR_NAN, "non-finite or runaway bead coordinates");
    if (st.flags & kFlagImageOverflow)
        return fail(ctx, GDPD_ERR_OVERFLOW, "a periodic image count left +-%d box crossings", kImgMax);
    if (st.flags & kFlagScatter)
        return fail(ctx, GDPD_ERR_HIP, "internal error: a binned bead fell outside the bead range");
    return GDPD_OK;
}

// ---- preparation: cell grid, bricks, capacities --------------------------------

int choose_capacities(gdpd_ctx *ctx, Sys *s, const std::vector<int> &cell_start)
{
    RepDev &h = s->h;
    // exact region counts of every brick at this state, then a margin for drift
    int maxreg = 0, maxown = 0;
    for (int b = 0; b < h.nbrick; ++b) {
        const int bz = b % h.nbk[2], by = (b / h.nbk[2]) % h.nbk[1], bx = b / (h.nbk[2] * h.nbk[1]);
        const int c0[3] = {bx * h.bdim[0], by * h.bdim[1], bz * h.bdim[2]};
        int ext[3];
        for (int d = 0; d < 3; ++d) ext[d] = std::min(h.bdim[d], h.nc[d] - c0[d]);
        int cnt = 0, own = 0;
        for (int rx = -1; rx <= ext[0]; ++rx)
            for (int ry = -1; ry <= ext[1]; ++ry)
                for (int rz = -1; rz <= ext[2]; ++rz) {
                    const int gx = (c0[0] + rx + h.nc[0]) % h.nc[0];
                    const int gy = (c0[1] + ry + h.nc[1]) % h.nc[1];
                    const int gz = (c0[2] + rz + h.nc[2]) % h.nc[2];
                    const int c = (gx * h.nc[1] + gy) * h.nc[2] + gz;
                    const int n = cell_start[c + 1] - cell_start[c];
                    cnt += n;
                    if (rx >= 0 && rx < ext[0] && ry >= 0 && ry < ext[1] && rz >= 0 && rz < ext[2]) own += n;
                }
        maxreg = std::max(maxreg, cnt);
        maxown = std::max(maxown, own);
    }
    // Margins for density drift: the largest of a few (up to 12% staged and
    // 15% owned beads) whose LDS layout still lets four workgroups share a
    // CU, the brick kernel's occupancy; a brick that outgrows them holds the
    // run and the capacities are sized again (settle).
    auto set = [&](double f) {
        h.smax = std::min(65000, std::max(64, (int)std::ceil(maxreg * (1.0 + 0.12 * f)) + 16));
        h.omax = std::min(65000, std::max(32, (int)std::ceil(maxown * (1.0 + 0.15 * f)) + 8));
    };
    const double fs[] = {1.0, 0.75, 0.5, 0.25, 0.0};
    for (double f : fs) {
        set(f);
        if (brick_params_of(s).lds_bytes(h.precision, s->style) <= kLdsFourPerCu) return GDPD_OK;
    }
    set(1.0);
    (void)ctx;
    return GDPD_OK;
}

// Capacities for the binned state cs_host, at least smin / omin, then the brick
// shrunk (longest edge first) until its LDS layout fits the device.
int fit_capacities(gdpd_ctx *ctx, Sys *s, const std::vector<int> &cs_host, int smin, int omin)
{
    RepDev &h = s->h;
    int rc;
    auto choose = [&]() {
        int r = choose_capacities(ctx, s, cs_host);
        h.smax = std::max(h.smax, smin);
        h.omax = std::max(h.omax, omin);
        return r;
    };
    if ((rc = choose())) return rc;
    for (int guard = 0; guard < 16; ++guard) {
        BrickLaunch b = brick_params_of(s);
        if (b.lds_bytes(h.precision, s->style) <= (size_t)ctx->max_lds) return GDPD_OK;
        int dmax = 0;
        for (int d = 1; d < 3; ++d)
            if (h.bdim[d] > h.bdim[dmax]) dmax = d;
        if (h.bdim[dmax] <= 1) return fail(ctx, GDPD_ERR_OVERFLOW, "bead density too high for LDS staging");
        h.bdim[dmax]--;
        for (int d = 0; d < 3; ++d) h.nbk[d] = (h.nc[d] + h.bdim[d] - 1) / h.bdim[d];
        h.nbrick = h.nbk[0] * h.nbk[1] * h.nbk[2];
        if (ctx->admitting) ctx->graveyard.push_back(h.thermo_part);
        else (void)hipFree(h.thermo_part);
        s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), (void *)h.thermo_part));
        if ((rc = dalloc(ctx, s, &h.thermo_part, (size_t)h.nbrick * kThermoParts))) return rc;
        if ((rc = choose())) return rc;
    }
    return fail(ctx, GDPD_ERR_OVERFLOW, "bead density too high for LDS staging");
}

// capacities the caller set (gdpd_set_capacity) in place of the automatic ones
int apply_set_capacity(gdpd_ctx *ctx, Sys *s)
{
    RepDev &h = s->h;
    if (s->cap_s > 0) h.smax = s->cap_s;
    if (s->cap_o > 0) h.omax = s->cap_o;
    if (brick_params_of(s).lds_bytes(h.precision, s->style) > (size_t)ctx->max_lds)
        return fail(ctx, GDPD_ERR_ARG, "capacities smax=%d, omax=%d do not fit in LDS with a %dx%dx%d brick", h.smax,
                    h.omax, h.bdim[0], h.bdim[1], h.bdim[2]);
    return GDPD_OK;
}

int sys_index(gdpd_ctx *ctx, const Sys *s)
{
    return (int)(std::find(ctx->sys.begin(), ctx->sys.end(), s) - ctx->sys.begin());
}

int prepare(gdpd_ctx *ctx, Sys *s)
{
    if (!s->uploaded) return fail(ctx, GDPD_ERR_STATE, "upload the system before computing");
    if (!s->have_pair) return fail(ctx, GDPD_ERR_STATE, "no pair_style set");
    if (s->prepared) return GDPD_OK;
    RepDev &h = s->h;
    const double rcm = max_rc(s);
    if (!(rcm > 0.0)) return fail(ctx, GDPD_ERR_ARG, "pair cutoff must be positive");
    for (int d = 0; d < 3; ++d) {
        h.nc[d] = (int)std::floor(h.L[d] / rcm);
        if (h.nc[d] < 3)
            return fail(ctx, GDPD_ERR_ARG,
                        "box edge %g holds fewer than 3 cells of the cutoff %g (the 27-cell stencil would repeat cells)",
                        h.L[d], rcm);
        h.cs[d] = h.L[d] / h.nc[d];
        h.inv_cs[d] = (double)h.nc[d] / h.L[d];
    }
    h.ncell = h.nc[0] * h.nc[1] * h.nc[2];
    // brick edge: aim at ~200 owned beads per 256-thread workgroup
    const double rho = h.N / (h.L[0] * h.L[1] * h.L[2]);
    const double cellvol = h.cs[0] * h.cs[1] * h.cs[2];
    for (int d = 0; d < 3; ++d) h.bdim[d] = s->tune_b[d];
    if (s->tune_b[0] <= 0 || s->tune_b[1] <= 0 || s->tune_b[2] <= 0) {
        const double target = 240.0 / std::max(rho * cellvol, 1e-9);
        int e = std::max(1, (int)std::floor(std::cbrt(target)));
        h.bdim[0] = h.bdim[1] = e;
        h.bdim[2] = std::max(1, (int)std::floor(target / (e * e)));
        h.bdim[2] = std::min(h.bdim[2], 3 * e);
    }
    for (int d = 0; d < 3; ++d) h.bdim[d] = std::max(1, std::min(h.bdim[d], h.nc[d]));
    if (s->tune_b[0] <= 0 || s->tune_b[1] <= 0 || s->tune_b[2] <= 0) {
        // A small system gives few bricks, and a launch of few workgroups runs
        // at the latency of one brick: shrink the longest brick edge (down to
        // 2 cells) until there are kMinBricks.  At 8,000 to 24,000 beads this
        // takes the step from 0.10 to 0.072 ms (mDPD slab) and from 0.047 to
        // 0.040 ms (DPD fluid); larger systems keep the throughput shape.
        auto count = [&]() {
            long long n = 1;
            for (int d = 0; d < 3; ++d) n *= (h.nc[d] + h.bdim[d] - 1) / h.bdim[d];
            return n;
        };
        while (count() < kMinBricks) {
            int dmax = 0;
            for (int d = 1; d < 3; ++d)
                if (h.bdim[d] > h.bdim[dmax]) dmax = d;
            if (h.bdim[dmax] <= 2) break;
            h.bdim[dmax]--;
        }
    }
    for (int d = 0; d < 3; ++d) h.nbk[d] = (h.nc[d] + h.bdim[d] - 1) / h.bdim[d];
    h.nbrick = h.nbk[0] * h.nbk[1] * h.nbk[2];
    if ((h.bdim[0] + 2) * (h.bdim[1] + 2) * (h.bdim[2] + 2) > 4 * kBlock)
        return fail(ctx, GDPD_ERR_ARG, "brick too large: at most %d region cells", 4 * kBlock);
    // the owned-bead info word holds region cell coordinates of 5, 5 and 6 bits
    if (h.bdim[0] > 30 || h.bdim[1] > 30 || h.bdim[2] > 62)
        return fail(ctx, GDPD_ERR_ARG, "brick edges must be at most 30 x 30 x 62 cells");

    int rc;
    if ((rc = dalloc(ctx, s, &h.cell_start, (size_t)h.ncell + 1))) return rc;
    if ((rc = dalloc(ctx, s, &h.cell_count, (size_t)h.ncell))) return rc;
    if ((rc = dalloc(ctx, s, &h.thermo_part, (size_t)h.nbrick * kThermoParts))) return rc;
    HIPCHK(hipMemsetAsync(h.cell_count, 0, sizeof(int) * h.ncell, ctx->stream));

    // coefficient table in the working precision
    double rc2max = 0, rd2max = 0;
    if (h.precision == GDPD_PREC_F64) {
        std::vector<PairCoef<double>> tab;
        build_coef<double>(s, tab, rc2max, rd2max);
        PairCoef<double> *d;
        if ((rc = dalloc(ctx, s, &d, tab.size()))) return rc;
        HIPCHK(hipMemcpyAsync(d, tab.data(), sizeof(tab[0]) * tab.size(), hipMemcpyHostToDevice, ctx->stream));
        h.coef = d;
    } else {
        std::vector<PairCoef<float>> tab;
        build_coef<float>(s, tab, rc2max, rd2max);
        PairCoef<float> *d;
        if ((rc = dalloc(ctx, s, &d, tab.size()))) return rc;
        HIPCHK(hipMemcpyAsync(d, tab.data(), sizeof(tab[0]) * tab.size(), hipMemcpyHostToDevice, ctx->stream));
        h.coef = d;
    }
    h.rcmax2 = rc2max;
    h.rdmax2 = rd2max;
    h.smax = 64;
    h.omax = 32;
    s->prepared = true;
    const int id = (int)(std::find(ctx->sys.begin(), ctx->sys.end(), s) - ctx->sys.begin());
    if ((rc = push_desc(ctx, id))) return rc;
    if ((rc = build_tables(ctx, {id}, s->tab))) return rc;

    // initial binning and physical sort into cell order
    if ((rc = rebin(ctx, s->tab))) return rc;
    if ((rc = resort_tables(ctx, s->tab))) return rc;
    std::vector<int> cs_host((size_t)h.ncell + 1);
    HIPCHK(hipMemcpyAsync(cs_host.data(), h.cell_start, sizeof(int) * cs_host.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if ((rc = fit_capacities(ctx, s, cs_host, 0, 0))) return rc;
    if ((rc = apply_set_capacity(ctx, s))) return rc;
    if ((rc = push_desc(ctx, id))) return rc;
    if ((rc = build_tables(ctx, {id}, s->tab))) return rc;
    s->cells_valid = true;
    s->steps_since_resort = 0;
    return GDPD_OK;
}

int ensure_cells(gdpd_ctx *ctx, Sys *s)
{
    int rc = prepare(ctx, s);
    if (rc) return rc;
    if (!s->cells_valid) {
        if ((rc = rebin(ctx, s->tab))) return rc;
        s->cells_valid = true;
    }
    return GDPD_OK;
}

int set_state_step(gdpd_ctx *ctx, Sys *s, long long step, bool reset_rows)
{
    RepState st;
    int rc = read_state(ctx, s, &st);
    if (rc) return rc;
    st.step = step;
    s->step_host = step;
    if (reset_rows) st.log_rows = 0;
    HIPCHK(hipMemcpyAsync(s->h.state, &st, sizeof(RepState), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GDPD_OK;
}

int maybe_resort(gdpd_ctx *ctx, Sys *s, Tables &T, int steps)
{
    s->steps_since_resort += steps;
    const int every = s->resort_every > 0 ? s->resort_every : kDefaultResortEvery;
    if (s->steps_since_resort >= every) {
        s->steps_since_resort = 0;
        return resort_tables(ctx, T);
    }
    return GDPD_OK;
}

int graph_block(gdpd_ctx *ctx, Graph &G, Tables &T, int steps)
{
    const int topo = topo_of(ctx, T);
    if (G.exec && G.steps == steps && G.topo == topo) {
        HIPCHK(hipGraphLaunch(G.exec, ctx->stream));
        return GDPD_OK;
    }
    G.release();
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = GDPD_OK;
    for (int k = 0; k < steps && !rc; ++k) rc = step_once(ctx, T, kModeFinalKick | kModeDrift);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(ctx, GDPD_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    G.graph = g;
    HIPCHK(hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0));
    G.steps = steps;
    G.topo = topo;
    HIPCHK(hipGraphLaunch(G.exec, ctx->stream));
    return GDPD_OK;
}

int advance_tables(gdpd_ctx *ctx, Tables &T, Graph &G, int nsteps, int graph_steps,
                   const std::vector<Sys *> &members)
{
    int rc;
    int done = 0;
    while (done < nsteps) {
        // steps until the next re-sort of any member
        int chunk = nsteps - done;
        for (Sys *s : members) {
            const int every = s->resort_every > 0 ? s->resort_every : kDefaultResortEvery;
            chunk = std::min(chunk, std::max(1, every - s->steps_since_resort));
        }
        const bool use_graph = !ctx->timing && graph_steps > 1;
        int k = 0;
        while (k < chunk) {
            if (use_graph && chunk - k >= graph_steps) {
                if ((rc = graph_block(ctx, G, T, graph_steps))) return rc;
                k += graph_steps;
            } else {
                if ((rc = step_once(ctx, T, kModeFinalKick | kModeDrift))) return rc;
                ++k;
            }
        }
        // one physical re-sort covers every member of the launch
        bool due = false;
        for (Sys *s : members) {
            s->steps_since_resort += chunk;
            const int every = s->resort_every > 0 ? s->resort_every : kDefaultResortEvery;
            due = due || s->steps_since_resort >= every;
        }
        if (due) {
            if ((rc = resort_tables(ctx, T))) return rc;
            for (Sys *s : members) s->steps_since_resort = 0;
        }
        done += chunk;
    }
    return GDPD_OK;
}


// ---- group fixes and computes --------------------------------------------------

unsigned all_types(const Sys *s) { return s->h.ntypes >= 32 ? ~0u : ((1u << s->h.ntypes) - 1u); }

// bead count and mass of every fix's and compute's group (type masks)
void refresh_groups(Sys *s)
{
    RepDev &h = s->h;
    auto stats = [&](unsigned g, int &cnt, double &mass) {
        cnt = 0;
        mass = 0.0;
        for (int t = 0; t < h.ntypes; ++t)
            if ((g >> t) & 1u) {
                cnt += s->type_count[t];
                mass += s->type_count[t] * h.mass[t];
            }
    };
    for (int k = 0; k < h.nfix; ++k) stats(h.fix[k].group, h.fix[k].gcount, h.fix[k].gmass);
    for (int k = 0; k < h.ncomp; ++k) stats(h.comp[k].group, h.comp[k].gcount, h.comp[k].gmass);
}

// Device fix list = the caller's fixes in order; with none, nve on every bead
// (the default of a system that never set fixes).  Anything beyond that one
// nve, or any compute, runs the step as the hook sequence (fix_iteration).
void finish_fix_setup(Sys *s)
{
    RepDev &h = s->h;
    const unsigned all = all_types(s);
    h.nfix = (int)s->fixes.size();
    for (int k = 0; k < h.nfix; ++k) {
        const gdpd_fix &f = s->fixes[k];
        FixDev &F = h.fix[k];
        F = FixDev{};
        F.kind = f.kind;
        F.group = f.group;
        F.group2 = f.group2;
        F.group3 = f.group3;
        F.every = f.every;
        F.dir = f.dir;
        F.flags = f.flags;
        for (int q = 0; q < 5; ++q) F.p[q] = f.p[q];
    }
    if (h.nfix == 0) {
        h.nfix = 1;
        h.fix[0] = FixDev{};
        h.fix[0].kind = GDPD_FIX_NVE;
        h.fix[0].group = all;
    }
    s->fix_path = !(h.nfix == 1 && h.fix[0].kind == GDPD_FIX_NVE && h.fix[0].group == all) || h.ncomp > 0;
    refresh_groups(s);
    s->graph.release();
}

unsigned fix_mask(const Sys *s, int kind)
{
    unsigned m = 0;
    for (int k = 0; k < s->h.nfix; ++k)
        if (s->h.fix[k].kind == kind) m |= 1u << k;
    return m;
}

// fixes of `kind` whose every-N schedule fires at step n
unsigned fix_firing(const Sys *s, int kind, long long n)
{
    unsigned m = 0;
    for (int k = 0; k < s->h.nfix; ++k) {
        const FixDev &F = s->h.fix[k];
        if (F.kind == kind && F.every > 0 && n % F.every == 0) m |= 1u << k;
    }
    return m;
}

unsigned comp_firing(const Sys *s, long long n, int kind)
{
    unsigned m = 0;
    for (int k = 0; k < s->h.ncomp; ++k) {
        const CompDev &C = s->h.comp[k];
        if ((kind == 0 || C.kind == kind) && C.every > 0 && n % C.every == 0) m |= 1u << k;
    }
    return m;
}

int check_groups(gdpd_ctx *ctx, const Sys *s)
{
    for (int k = 0; k < s->h.nfix; ++k) {
        const FixDev &F = s->h.fix[k];
        if ((F.kind == GDPD_FIX_SPRING || F.kind == GDPD_FIX_RECENTER || F.kind == GDPD_FIX_RVCM) && !(F.gmass > 0.0))
            return fail(ctx, GDPD_ERR_STATE, "fix %d: its group holds no beads", k);
        if (F.kind == GDPD_FIX_TWO_WALL) {
            for (unsigned g : {F.group2, F.group3}) {
                int n = 0;
                for (int t = 0; t < s->h.ntypes; ++t)
                    if ((g >> t) & 1u) n += s->type_count[t];
                if (n == 0) return fail(ctx, GDPD_ERR_STATE, "fix %d (two_wall): a wall group holds no beads", k);
            }
        }
    }
    for (int k = 0; k < s->h.ncomp; ++k)
        if (s->h.comp[k].gcount <= 0) return fail(ctx, GDPD_ERR_STATE, "compute %d: its group holds no beads", k);
    return GDPD_OK;
}

// keys, scan and scatter for the current positions (no in-cell order)
void rebin_fast(gdpd_ctx *ctx, Tables &T)
{
    launch_bin_keys(ctx->d_reps, T.beads, T.n_beads, ctx->stream);
    launch_scan(ctx->d_reps, T.scan, T.n_scan, T.rep_chunk_off, T.rep_list, T.n_reps, T.chunk_sums, ctx->stream);
    launch_scatter(ctx->d_reps, T.beads, T.n_beads, ctx->stream);
}

// One step of a system with group fixes or computes, in the reference's hook
// order (main/utils.cuh:166-314) with every reduction on the device:
//   force at x(n): pair (per-bead PE/virial on log steps), bonds, angles;
//   post_force (spring); final_integrate (nve) unless step 0; end_of_step
//   (rvcm, recenter) unless step 0; thermo record; compute records;
//   then, unless this is the last step, initial_integrate and post_integrate
//   of step n + 1 and its binning.
int fix_iteration(gdpd_ctx *ctx, Sys *s, bool final_kick, bool end_of_step, bool next_initial)
{
    RepDev &h = s->h;
    Tables &T = s->tab;
    hipStream_t st = ctx->stream;
    const long long n = s->fix_step;
    const long long N = h.N;
    const bool log = h.log_every > 0 && n % h.log_every == 0;
    const unsigned spring = fix_mask(s, GDPD_FIX_SPRING), nve = fix_mask(s, GDPD_FIX_NVE),
                   two_wall = fix_mask(s, GDPD_FIX_TWO_WALL);
    const unsigned recenter = end_of_step ? fix_firing(s, GDPD_FIX_RECENTER, n) : 0u;
    const unsigned rvcm = end_of_step ? fix_firing(s, GDPD_FIX_RVCM, n) : 0u;
    const unsigned spring_rec = fix_firing(s, GDPD_FIX_SPRING, n);
    const unsigned cfire = comp_firing(s, n, 0), ctf = comp_firing(s, n, GDPD_COMPUTE_TF),
                   ccontact = comp_firing(s, n, GDPD_COMPUTE_WALL_CONTACT);
    const unsigned cend = cfire & ~ctf;
    launch_fix_clear(ctx->d_reps, T.rep_list, T.n_reps, st);
    if (spring | recenter) launch_fix_reduce(ctx->d_reps, T.beads, T.n_beads, spring | recenter, 0u, 0, st);
    if (log) {
        launch_fill_f64(h.pe, 0.0, N, st);
        launch_fill_f64(h.vir, 0.0, 6 * N, st);
    }
    int rc;
    if ((rc = force_launch(ctx, T, kModeForceOnly | (log ? kModeLogForce : 0)))) return rc;
    if (log) {
        launch_fix_to_double(h.pe, N, h.precision, st);
        launch_fix_to_double(h.vir, 6 * N, h.precision, st);
    }
    if (h.nbonds) launch_bond_hook(ctx->d_reps, T.beads, T.n_beads, log, st);
    if (h.nangles) launch_angle_hook(ctx->d_reps, T.beads, T.n_beads, log, st);
    if (spring) launch_fix_post_force(ctx->d_reps, T.beads, T.n_beads, spring, st);
    if (final_kick) launch_fix_final(ctx->d_reps, T.beads, T.n_beads, nve, st);
    if (rvcm | ctf) launch_fix_reduce(ctx->d_reps, T.beads, T.n_beads, rvcm, ctf, 1, st);
    if (rvcm | recenter) launch_fix_end(ctx->d_reps, T.beads, T.n_beads, rvcm | recenter, st);
    if (log) launch_thermo_arrays(ctx->d_reps, T.beads, T.n_beads, nullptr, st);
    if (cend) launch_fix_reduce(ctx->d_reps, T.beads, T.n_beads, 0u, cend, 2, st);
    if (ccontact) launch_fix_reduce(ctx->d_reps, T.beads, T.n_beads, 0u, ccontact, 3, st);
    if (cfire | spring_rec) launch_fix_records(ctx->d_reps, T.rep_list, T.n_reps, spring_rec, cfire, n, st);
    launch_step_end(ctx->d_reps, T.rep_list, T.n_reps, 0, 1, 1, st, 0);
    s->fix_step = n + 1;
    if (next_initial) {
        launch_fix_initial(ctx->d_reps, T.beads, T.n_beads, (1u << h.nfix) - 1u, st);
        if (two_wall) launch_fix_two_wall(ctx->d_reps, T.beads, T.n_beads, two_wall, st);
        rebin_fast(ctx, T);
        // the physical re-sort moves x, v and the bead tables but not the force:
        // safe here, the next use of the force follows the next force evaluation
        if ((rc = check_launch(ctx))) return rc;
        return maybe_resort(ctx, s, T, 1);
    }
    return check_launch(ctx);
}

int fix_zero_step(gdpd_ctx *ctx, Sys *s, bool next_initial)
{
    int rc;
    if ((rc = check_groups(ctx, s))) return rc;
    if ((rc = ensure_cells(ctx, s))) return rc;
    if ((rc = set_state_step(ctx, s, 0, true))) return rc;
    launch_fill_i32(s->h.rec_n, 0, kRecSlots, ctx->stream);
    s->fix_step = 0;
    return fix_iteration(ctx, s, false, false, next_initial);
}

// ---- staging overflow: hold, grow, resume --------------------------------------
//
// A brick whose beads exceed its LDS capacity halts its system (kFlagHalt):
// every later step kernel leaves that system's state alone, so it stays at the
// last completed step while the rest of a batch goes on.  When the host next
// looks (settle), it sizes the capacities again from the held state's cell
// lists, clears the flag and runs the missing steps for that system alone.  A
// system's arithmetic depends only on its own state and brick shape, so the
// result equals a run that never overflowed, bit for bit, as long as the brick
// shape itself still fits in LDS.

// scalar states of the systems ids: one gather launch and one copy
int read_states(gdpd_ctx *ctx, const std::vector<int> &ids, std::vector<RepState> &out)
{
    const int n = (int)ids.size();
    out.resize(n);
    if (!n) return GDPD_OK;
    if (n > ctx->gather_cap) {
        int cap = std::max(64, ctx->gather_cap);
        while (cap < n) cap *= 2;
        void *old[] = {ctx->d_ids, ctx->d_states, ctx->d_ts, ctx->d_tsteps};
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (void *q : old)
            if (q) (void)hipFree(q);
        ctx->d_ids = nullptr;
        ctx->d_states = nullptr;
        ctx->d_ts = nullptr;
        ctx->d_tsteps = nullptr;
        ctx->gather_cap = 0;
        HIPCHK(hipMalloc((void **)&ctx->d_ids, sizeof(int) * cap));
        HIPCHK(hipMalloc((void **)&ctx->d_states, sizeof(RepState) * cap));
        HIPCHK(hipMalloc((void **)&ctx->d_ts, sizeof(unsigned long long) * kStepTsCap * cap));
        HIPCHK(hipMalloc((void **)&ctx->d_tsteps, sizeof(long long) * cap));
        ctx->gather_cap = cap;
    }
    HIPCHK(hipMemcpyAsync(ctx->d_ids, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    launch_gather_state(ctx->d_reps, ctx->d_ids, n, ctx->d_states, ctx->stream);
    int rc = check_launch(ctx);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out.data(), ctx->d_states, sizeof(RepState) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GDPD_OK;
}

// New LDS capacities from the held state's cell lists (the brick shrinks only
// when they no longer fit); tables and graphs are rebuilt, the halt cleared.
int regrow(gdpd_ctx *ctx, Sys *s)
{
    RepDev &h = s->h;
    std::vector<int> cs((size_t)h.ncell + 1);
    HIPCHK(hipMemcpyAsync(cs.data(), h.cell_start, sizeof(int) * cs.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int s0 = h.smax, o0 = h.omax, b0[3] = {h.bdim[0], h.bdim[1], h.bdim[2]};
    int rc;
    if ((rc = fit_capacities(ctx, s, cs, 0, 0))) return rc;
    const bool same_brick = h.bdim[0] == b0[0] && h.bdim[1] == b0[1] && h.bdim[2] == b0[2];
    if (same_brick && h.smax <= s0 && h.omax <= o0)
        return fail(ctx, GDPD_ERR_OVERFLOW, "a brick holds more beads than the largest LDS capacity (smax=%d, omax=%d)",
                    h.smax, h.omax);
    s->regrows++;
    const int id = sys_index(ctx, s);
    if ((rc = push_desc(ctx, id))) return rc;
    if ((rc = build_tables(ctx, {id}, s->tab))) return rc;
    s->graph.release();
    ctx->batch_graph.release();
    ctx->batch.release();
    RepState st;
    if ((rc = read_state(ctx, s, &st))) return rc;
    st.flags &= ~kFlagHalt;
    HIPCHK(hipMemcpyAsync(h.state, &st, sizeof(st), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GDPD_OK;
}

// The steps a system held at device step `held` still owes: the zero step if
// it halted there, then regular steps up to step_host.  Only zero_step and
// advance leave work pending, so the system is mid-run.
int resume(gdpd_ctx *ctx, Sys *s, long long held)
{
    const long long target = s->step_host;
    int rc = GDPD_OK;
    if (s->fix_path) {
        s->fix_step = held;
        for (long long n = held; n < target && !rc; ++n) rc = fix_iteration(ctx, s, n > 0, n > 0, true);
        s->step_host = target;
        return rc;
    }
    if (held == 0 && target >= 1) {
        if ((rc = step_once(ctx, s->tab, kModeDrift))) return rc;
        if ((rc = maybe_resort(ctx, s, s->tab, 1))) return rc;
        held = 1;
    }
    if (target > held) {
        const int gs = s->graph_steps > 0 ? s->graph_steps : 8;
        if ((rc = advance_tables(ctx, s->tab, s->graph, (int)(target - held), gs, {s}))) return rc;
    }
    return GDPD_OK;
}

void drop_pending(gdpd_ctx *ctx, int id)
{
    ctx->pending.erase(std::remove(ctx->pending.begin(), ctx->pending.end(), id), ctx->pending.end());
}

// Read the halt flags of every system with enqueued steps; grow and resume
// the held ones (each resumed system is checked again).
int settle(gdpd_ctx *ctx)
{
    // inside an admission window only new systems are reachable, and their
    // steps are settled after it, on the main stream
    if (ctx->admitting) return GDPD_OK;
    for (int round = 0; round < 16 && !ctx->pending.empty(); ++round) {
        std::vector<int> ids;
        for (int id : ctx->pending)
            if (get_sys(ctx, id) && std::find(ids.begin(), ids.end(), id) == ids.end()) ids.push_back(id);
        ctx->pending.clear();
        std::vector<RepState> st;
        int rc = read_states(ctx, ids, st);
        if (rc) return rc;
        flush_graveyard(ctx);
        for (size_t k = 0; k < ids.size(); ++k) {
            if (!(st[k].flags & kFlagHalt)) continue;
            Sys *s = ctx->sys[ids[k]];
            if ((rc = regrow(ctx, s))) return rc;
            if ((rc = resume(ctx, s, st[k].step))) return rc;
            ctx->pending.push_back(ids[k]);
        }
    }
    if (!ctx->pending.empty()) {
        ctx->pending.clear();
        return fail(ctx, GDPD_ERR_OVERFLOW, "brick staging kept overflowing after 16 capacity increases");
    }
    return GDPD_OK;
}

// After one synchronous step of s: true (and capacities grown) when staging
// overflowed, so the caller repeats the step on the held state.
int held_step(gdpd_ctx *ctx, Sys *s, bool &again)
{
    again = false;
    RepState st;
    int rc = read_state(ctx, s, &st);
    if (rc) return rc;
    if (!(st.flags & kFlagHalt)) return GDPD_OK;
    if (s->regrows > 64) return fail(ctx, GDPD_ERR_OVERFLOW, "brick staging kept overflowing");
    if ((rc = regrow(ctx, s))) return rc;
    again = true;
    return GDPD_OK;
}
} // namespace

// =============================================================================
// C ABI
// =============================================================================

extern "C" {

int gdpd_create(gdpd_ctx **out, int device)
{
    if (!out) return GDPD_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GDPD_ERR_HIP;
    if (device < 0 || device >= n) return GDPD_ERR_ARG;
    gdpd_ctx *ctx = new gdpd_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return GDPD_ERR_HIP;
    }
    ctx->main_stream = ctx->stream;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.sharedMemPerBlock > 0)
        ctx->max_lds = (int)std::min<size_t>(prop.sharedMemPerBlock, 64 * 1024);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        ctx->tick_ns = 1.0e6 / (double)khz;
    // error messages fit the reserved capacity, so reporting one allocates nothing
    ctx->err.reserve(600);
    if (device < kMaxDevices) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_dev[device]) {
            void *p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_dev[device] = (double *)p;
            else (void)hipGetLastError();
        }
    }
    *out = ctx;
    return GDPD_OK;
}

int gdpd_destroy(gdpd_ctx *ctx)
{
    if (!ctx) return GDPD_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->admitting) (void)gdpd_admission(ctx, 0);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side_stream) (void)hipStreamSynchronize(ctx->side_stream);
    for (size_t i = 0; i < ctx->sys.size(); ++i)
        if (ctx->sys[i]) gdpd_sys_destroy(ctx, (int)i);
    flush_graveyard(ctx);
    ctx->batch.release();
    ctx->batch_graph.release();
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    void *gb[] = {ctx->d_ids, ctx->d_states, ctx->d_ts, ctx->d_tsteps, ctx->d_phase};
    for (void *q : gb)
        if (q) (void)hipFree(q);
    if (ctx->d_reps) (void)hipFree(ctx->d_reps);
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->side_done) (void)hipEventDestroy(ctx->side_done);
    if (ctx->side_stream) (void)hipStreamDestroy(ctx->side_stream);
    delete ctx;
    return GDPD_OK;
}

int gdpd_admission(gdpd_ctx *ctx, int on)
{
    if (!ctx) return GDPD_ERR_ARG;
    if (on) {
        if (ctx->admitting) return GDPD_OK;
        if (!ctx->side_stream) {
            HIPCHK(hipStreamCreateWithFlags(&ctx->side_stream, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&ctx->side_done, hipEventDisableTiming));
        }
        if (ctx->timing) return fail(ctx, GDPD_ERR_STATE, "no admission window while kernel timing is on");
        // New systems touch only their own arrays and their own slot of the
        // descriptor table, which no queued main-stream work reads, so the side
        // stream need not wait for the main one.  (A descriptor table that has
        // to grow is freed by hipFree, which waits for the whole device.)
        ctx->stream = ctx->side_stream;
        ctx->admitting = true;
        return GDPD_OK;
    }
    if (!ctx->admitting) return GDPD_OK;
    // later main-stream work waits for everything the window enqueued
    HIPCHK(hipEventRecord(ctx->side_done, ctx->side_stream));
    HIPCHK(hipStreamWaitEvent(ctx->main_stream, ctx->side_done, 0));
    ctx->stream = ctx->main_stream;
    ctx->admitting = false;
    for (Sys *s : ctx->sys)
        if (s) s->admitted = false;
    return GDPD_OK;
}

const char *gdpd_last_error(const gdpd_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int64_t gdpd_system_bytes(const gdpd_system_desc *d, int n_bonds)
{
    if (!d || d->n_beads <= 0) return 0;
    const int64_t N = d->n_beads;
    double vol = 1.0;
    for (int k = 0; k < 3; ++k) vol *= (d->hi[k] - d->lo[k]);
    const int64_t ncell = (int64_t)std::max(1.0, vol / std::max(d->global_cut * d->global_cut * d->global_cut, 1e-3));
    // state x2 (96) + force/pe/bpe/ape/vir/rho (104) + ints and image word (44) + type
    return N * 249 + ncell * 8 + (int64_t)n_bonds * 16 + (1 << 20);
}

int gdpd_sys_create(gdpd_ctx *ctx, const gdpd_system_desc *d, int *sys_out)
{
    if (!ctx || !d || !sys_out) return fail(ctx, GDPD_ERR_ARG, "null argument");
    if (d->n_beads <= 0) return fail(ctx, GDPD_ERR_ARG, "n_beads must be positive");
    if (d->n_beads > kMaxBeads) return fail(ctx, GDPD_ERR_ARG, "at most %d beads per system", kMaxBeads);
    if (d->n_types < 1 || d->n_types > kMaxTypes)
        return fail(ctx, GDPD_ERR_ARG, "n_types must be in 1..%d", kMaxTypes);
    for (int k = 0; k < 3; ++k)
        if (!(d->hi[k] > d->lo[k])) return fail(ctx, GDPD_ERR_ARG, "box hi must exceed lo");
    if (!(d->dt > 0.0)) return fail(ctx, GDPD_ERR_ARG, "timestep must be positive");
    if (d->precision != GDPD_PREC_F32 && d->precision != GDPD_PREC_F64)
        return fail(ctx, GDPD_ERR_ARG, "unknown precision %d", d->precision);
    HIPCHK(hipSetDevice(ctx->device));
    const bool admitting = ctx->admitting;
    Sys *s = new Sys();
    s->admitted = admitting;
    s->desc = *d;
    RepDev &h = s->h;
    const int N = d->n_beads;
    h.N = N;
    h.ntypes = d->n_types;
    h.precision = d->precision;
    h.dt = d->dt;
    h.log_every = d->log_every > 0 ? d->log_every : 0;
    for (int k = 0; k < 3; ++k) {
        h.lo[k] = d->lo[k];
        h.L[k] = d->hi[k] - d->lo[k];
        h.hL[k] = 0.5 * h.L[k];
    }
    int id = -1;
    for (size_t i = 0; i < ctx->sys.size(); ++i)
        if (!ctx->sys[i]) {
            id = (int)i;
            break;
        }
    if (id < 0) {
        id = (int)ctx->sys.size();
        ctx->sys.push_back(nullptr);
    }
    ctx->sys[id] = s;
    int rc = GDPD_OK;
#define ALLOC(p, n)                                    \
    if ((rc = dalloc(ctx, s, &(p), (size_t)(n)))) {    \
        gdpd_sys_destroy(ctx, id);                     \
        return rc;                                     \
    }
    ALLOC(h.x[0], 3 * (size_t)N);
    ALLOC(h.x[1], 3 * (size_t)N);
    ALLOC(h.v[0], 3 * (size_t)N);
    ALLOC(h.v[1], 3 * (size_t)N);
    ALLOC(h.id, N);
    ALLOC(h.type, N);
    ALLOC(h.img, N);
    ALLOC(h.perm, N);
    ALLOC(h.key, N);
    ALLOC(h.rank, N);
    ALLOC(h.slot_of_id, N);
    ALLOC(h.force, 3 * (size_t)N);
    ALLOC(h.pe, N);
    ALLOC(h.bpe, N);
    ALLOC(h.ape, N);
    ALLOC(h.angle_start, (size_t)N + 1);
    ALLOC(h.vir, 6 * (size_t)N);
    ALLOC(h.rho, N);
    ALLOC(h.id_alt, N);
    ALLOC(h.type_alt, N);
    ALLOC(h.img_alt, N);
    ALLOC(h.state, 1);
    ALLOC(h.step_ts, kStepTsCap);
    ALLOC(h.racc, (size_t)kRecSlots * kAccPer * 2);
    ALLOC(h.rext, (size_t)kRecSlots * 2);
    ALLOC(h.rec_rows, (size_t)kRecSlots * kRecCap * kRecLen);
    ALLOC(h.rec_n, (size_t)kRecSlots);
    h.thermo_cap = 1024;
    ALLOC(h.thermo_rows, (size_t)h.thermo_cap * kThermoRowLen);
    ALLOC(h.thermo_bead, (size_t)((N + kBlock - 1) / kBlock) * kThermoParts);
    // empty bond CSR (bond_start[id] == 0 for all ids)
    ALLOC(h.bond_start, (size_t)N + 1);
#undef ALLOC
    HIPCHK(hipMemsetAsync(h.step_ts, 0, sizeof(unsigned long long) * kStepTsCap, ctx->stream));
    HIPCHK(hipMemsetAsync(h.bond_start, 0, sizeof(int) * ((size_t)N + 1), ctx->stream));
    HIPCHK(hipMemsetAsync(h.rec_n, 0, sizeof(int) * kRecSlots, ctx->stream));
    h.x0uw = nullptr;
    h.nfix = h.ncomp = 0;
    HIPCHK(hipMemsetAsync(h.force, 0, sizeof(double) * 3 * (size_t)N, ctx->stream));
    HIPCHK(hipMemsetAsync(h.pe, 0, sizeof(double) * (size_t)N, ctx->stream));
    HIPCHK(hipMemsetAsync(h.bpe, 0, sizeof(double) * (size_t)N, ctx->stream));
    HIPCHK(hipMemsetAsync(h.ape, 0, sizeof(double) * (size_t)N, ctx->stream));
    HIPCHK(hipMemsetAsync(h.angle_start, 0, sizeof(int) * ((size_t)N + 1), ctx->stream));
    HIPCHK(hipMemsetAsync(h.vir, 0, sizeof(double) * 6 * (size_t)N, ctx->stream));
    HIPCHK(hipMemsetAsync(h.rho, 0, sizeof(double) * (size_t)N, ctx->stream));
    RepState st{};
    HIPCHK(hipMemcpyAsync(h.state, &st, sizeof(st), hipMemcpyHostToDevice, ctx->stream));
    if ((rc = push_desc(ctx, id))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *sys_out = id;
    return GDPD_OK;
}

int gdpd_sys_destroy(gdpd_ctx *ctx, int id)
{
    Sys *s = get_sys(ctx, id);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", id);
    (void)hipStreamSynchronize(ctx->stream);
    flush_graveyard(ctx);
    drop_pending(ctx, id);
    s->tab.release();
    s->graph.release();
    for (void *p : s->allocs) (void)hipFree(p);
    void *nl[] = {s->nl_binof, s->nl_count, s->nl_start, s->nl_fill, s->nl_content, s->nl_n,
                  s->nl_neigh, s->nl_over, s->nl_flag, s->nl_last};
    for (void *p : nl)
        if (p) (void)hipFree(p);
    delete s;
    ctx->sys[id] = nullptr;
    ctx->batch.release();
    ctx->batch_graph.release();
    return GDPD_OK;
}

int gdpd_upload(gdpd_ctx *ctx, int sid, const int *id, const int *type, const double *pos,
                const double *vel, const double *masses)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !id || !type || !pos || !masses) return fail(ctx, GDPD_ERR_ARG, "bad upload arguments");
    RepDev &h = s->h;
    const int N = h.N;
    std::vector<double> X(3 * (size_t)N), V(3 * (size_t)N, 0.0);
    std::vector<int> ids(N), soi(N, -1), perm(N);
    std::vector<long long> img(N, pack_img(0, 0, 0));
    std::vector<unsigned char> ty(N);
    for (int k = 0; k < N; ++k) {
        if (id[k] < 1 || id[k] > N) return fail(ctx, GDPD_ERR_ARG, "atom id %d outside 1..%d", id[k], N);
        if (soi[id[k] - 1] >= 0) return fail(ctx, GDPD_ERR_ARG, "duplicate atom id %d", id[k]);
        if (type[k] < 1 || type[k] > h.ntypes)
            return fail(ctx, GDPD_ERR_ARG, "atom type %d outside 1..%d", type[k], h.ntypes);
        soi[id[k] - 1] = k;
        ids[k] = id[k];
        ty[k] = (unsigned char)type[k];
        perm[k] = k;
        for (int d = 0; d < 3; ++d) {
            double x = pos[3 * (size_t)k + d];
            if (!std::isfinite(x)) return fail(ctx, GDPD_ERR_NAN, "non-finite position of atom %d", id[k]);
            // system/read_data.cu:191-197; a position more than 16 box lengths
            // out is refused (the reference's one-box-at-a-time loop would run
            // for as many iterations as the position is box lengths away)
            const double lo = h.lo[d], hi = h.lo[d] + h.L[d];
            if (x < lo - 16.0 * h.L[d] || x >= hi + 16.0 * h.L[d])
                return fail(ctx, GDPD_ERR_ARG, "atom %d lies more than 16 box lengths outside the box", id[k]);
            while (x >= hi || x < lo) x -= (hi - lo) * (double)((x >= hi) - (x < lo));
            X[(size_t)d * N + k] = x;
            if (vel) V[(size_t)d * N + k] = vel[3 * (size_t)k + d];
        }
    }
    s->masses.assign(masses, masses + h.ntypes);
    for (int t = 0; t < kMaxTypes; ++t) s->type_count[t] = 0;
    for (int k = 0; k < N; ++k) s->type_count[ty[k] - 1] += 1;
    refresh_groups(s);
    for (int t = 0; t < h.ntypes; ++t) {
        if (!(masses[t] > 0.0)) return fail(ctx, GDPD_ERR_ARG, "mass of type %d must be positive", t + 1);
        h.mass[t] = masses[t];
        h.minv[t] = 1.0 / masses[t];
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    RepState st{};
    HIPCHK(hipMemcpyAsync(h.state, &st, sizeof(st), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.x[0], X.data(), sizeof(double) * X.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.x[1], X.data(), sizeof(double) * X.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.v[0], V.data(), sizeof(double) * V.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.v[1], V.data(), sizeof(double) * V.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.id, ids.data(), sizeof(int) * N, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.type, ty.data(), N, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.img, img.data(), sizeof(long long) * N, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.perm, perm.data(), sizeof(int) * N, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.slot_of_id, soi.data(), sizeof(int) * N, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    s->uploaded = true;
    s->cells_valid = false;
    s->midstep = false;
    s->nl_built = false;
    s->step_host = 0;
    drop_pending(ctx, sid);
    int rc = push_desc(ctx, sid);
    if (rc) return rc;
    if (s->prepared) {
        // re-sort the fresh state into cell order, then size the LDS
        // capacities for it (the brick shape is kept while it fits)
        launch_fill_i32(h.cell_count, 0, h.ncell, ctx->stream);
        if ((rc = rebin(ctx, s->tab))) return rc;
        if ((rc = resort_tables(ctx, s->tab))) return rc;
        std::vector<int> cs((size_t)h.ncell + 1);
        HIPCHK(hipMemcpyAsync(cs.data(), h.cell_start, sizeof(int) * cs.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        const RepDev keep = h;
        if ((rc = fit_capacities(ctx, s, cs, 0, 0))) {
            // a state too dense for any brick: keep the brick geometry and
            // require a fresh upload before the next computation
            if (h.nbrick != keep.nbrick) {
                (void)hipFree(h.thermo_part);
                s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), (void *)h.thermo_part));
                h.thermo_part = nullptr;
                const int rc2 = dalloc(ctx, s, &h.thermo_part, (size_t)keep.nbrick * kThermoParts);
                if (rc2) return rc2;
            }
            for (int d = 0; d < 3; ++d) {
                h.bdim[d] = keep.bdim[d];
                h.nbk[d] = keep.nbk[d];
            }
            h.nbrick = keep.nbrick;
            h.smax = keep.smax;
            h.omax = keep.omax;
            s->uploaded = false;
            (void)push_desc(ctx, sid);
            (void)build_tables(ctx, {sid}, s->tab);
            s->graph.release();
            return rc;
        }
        if ((rc = apply_set_capacity(ctx, s))) return rc;
        if ((rc = push_desc(ctx, sid))) return rc;
        if ((rc = build_tables(ctx, {sid}, s->tab))) return rc;
        s->graph.release();
        ctx->batch_graph.release();
        ctx->batch.release();
        s->cells_valid = true;
        s->steps_since_resort = 0;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GDPD_OK;
}

int gdpd_download(gdpd_ctx *ctx, int sid, double *pos, double *vel, int *image, double *force)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !s->uploaded) return fail(ctx, GDPD_ERR_STATE, "system %d has no state", sid);
    RepDev &h = s->h;
    const int N = h.N;
    int rc = settle(ctx);
    if (rc) return rc;
    RepState st;
    if ((rc = read_state(ctx, s, &st))) return rc;
    std::vector<int> ids(N);
    HIPCHK(hipMemcpy(ids.data(), h.id, sizeof(int) * N, hipMemcpyDeviceToHost));
    auto fetch3 = [&](const double *src, double *dst) -> int {
        std::vector<double> tmp(3 * (size_t)N);
        HIPCHK(hipMemcpy(tmp.data(), src, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < N; ++k)
            for (int d = 0; d < 3; ++d) dst[3 * (size_t)(ids[k] - 1) + d] = tmp[(size_t)d * N + k];
        return GDPD_OK;
    };
    if (pos && (rc = fetch3(h.x[st.cur], pos))) return rc;
    if (vel && (rc = fetch3(h.v[st.cur], vel))) return rc;
    if (force && (rc = fetch3(h.force, force))) return rc;
    if (image) {
        std::vector<long long> im(N);
        HIPCHK(hipMemcpy(im.data(), h.img, sizeof(long long) * N, hipMemcpyDeviceToHost));
        for (int k = 0; k < N; ++k) {
            int a, b, c;
            unpack_img(im[k], a, b, c);
            image[3 * (size_t)(ids[k] - 1) + 0] = a;
            image[3 * (size_t)(ids[k] - 1) + 1] = b;
            image[3 * (size_t)(ids[k] - 1) + 2] = c;
        }
    }
    return GDPD_OK;
}

int gdpd_download_pe_virial(gdpd_ctx *ctx, int sid, double *pe, double *virial, double *rho)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !s->uploaded) return fail(ctx, GDPD_ERR_STATE, "system %d has no state", sid);
    RepDev &h = s->h;
    const int N = h.N;
    int src = settle(ctx);
    if (src) return src;
    std::vector<int> ids(N);
    HIPCHK(hipMemcpy(ids.data(), h.id, sizeof(int) * N, hipMemcpyDeviceToHost));
    std::vector<double> tmp(6 * (size_t)N);
    if (pe) {
        HIPCHK(hipMemcpy(tmp.data(), h.pe, sizeof(double) * N, hipMemcpyDeviceToHost));
        for (int k = 0; k < N; ++k) pe[ids[k] - 1] = tmp[k];
    }
    if (rho) {
        HIPCHK(hipMemcpy(tmp.data(), h.rho, sizeof(double) * N, hipMemcpyDeviceToHost));
        for (int k = 0; k < N; ++k) rho[ids[k] - 1] = tmp[k];
    }
    if (virial) {
        HIPCHK(hipMemcpy(tmp.data(), h.vir, sizeof(double) * 6 * N, hipMemcpyDeviceToHost));
        for (int k = 0; k < N; ++k)
            for (int c = 0; c < 6; ++c) virial[6 * (size_t)(ids[k] - 1) + c] = tmp[(size_t)c * N + k];
    }
    return GDPD_OK;
}

static int set_pair(gdpd_ctx *ctx, int sid, int style, double temp, double rc, double rd,
                    uint32_t seed, const double *coeff)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !coeff) return fail(ctx, GDPD_ERR_ARG, "bad pair arguments");
    if (s->prepared) return fail(ctx, GDPD_ERR_STATE, "pair style is fixed once the system has run");
    const int nt = s->desc.n_types, np = nt * (nt + 1) / 2, w = (style == GDPD_STYLE_MDPD) ? 5 : 3;
    // force/pair_dpd.cu:78-81,140-143; pair_mdpd.cu:81-89,154-162
    if (rc > s->desc.global_cut)
        return fail(ctx, GDPD_ERR_ARG, "global rc should not be bigger than the global_cut (neighbor)");
    if (style == GDPD_STYLE_MDPD && rd > rc) return fail(ctx, GDPD_ERR_ARG, "rd should not be bigger than rc");
    for (int k = 0; k < np; ++k) {
        const double rck = (style == GDPD_STYLE_MDPD) ? coeff[k * 5 + 3] : coeff[k * 3 + 2];
        if (!(rck > 0.0)) return fail(ctx, GDPD_ERR_ARG, "pair cutoff of pair %d must be positive", k);
        if (rck > s->desc.global_cut)
            return fail(ctx, GDPD_ERR_ARG, "global rc should not be bigger than the global_cut (neighbor)");
        if (style == GDPD_STYLE_MDPD && coeff[k * 5 + 4] > rck)
            return fail(ctx, GDPD_ERR_ARG, "rd should not be bigger than rc");
        if (style == GDPD_STYLE_MDPD && !(coeff[k * 5 + 4] > 0.0))
            return fail(ctx, GDPD_ERR_ARG, "rd must be positive");
        const double g = (style == GDPD_STYLE_MDPD) ? coeff[k * 5 + 2] : coeff[k * 3 + 1];
        if (g < 0.0) return fail(ctx, GDPD_ERR_ARG, "gamma must be non-negative");
    }
    if (temp < 0.0) return fail(ctx, GDPD_ERR_ARG, "temperature must be non-negative");
    s->style = style;
    s->temp = temp;
    s->rc = rc;
    s->rd = rd;
    s->seed = seed;
    s->coeff.assign(coeff, coeff + (size_t)np * w);
    s->have_pair = true;
    s->h.style = style;
    s->h.seed = seed;
    return push_desc(ctx, sid);
}

int gdpd_set_pair_dpd(gdpd_ctx *ctx, int sid, double temp, double rc, uint32_t seed, const double *coeff)
{
    return set_pair(ctx, sid, GDPD_STYLE_DPD, temp, rc, rc, seed, coeff);
}

int gdpd_set_pair_mdpd(gdpd_ctx *ctx, int sid, double temp, double rc, double rd, uint32_t seed,
                       const double *coeff)
{
    return set_pair(ctx, sid, GDPD_STYLE_MDPD, temp, rc, rd, seed, coeff);
}

int gdpd_set_bonds(gdpd_ctx *ctx, int sid, int n_bonds, int n_bond_types, const int *bondlist,
                   const double *bond_coeff)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || n_bonds < 0 || (n_bonds > 0 && (!bondlist || !bond_coeff)))
        return fail(ctx, GDPD_ERR_ARG, "bad bond arguments");
    if (n_bond_types < 0 || n_bond_types > kMaxBondTypes)
        return fail(ctx, GDPD_ERR_ARG, "n_bond_types must be in 0..%d", kMaxBondTypes);
    RepDev &h = s->h;
    const int N = h.N;
    std::vector<int> deg(N + 1, 0);
    for (int b = 0; b < n_bonds; ++b) {
        const int t = bondlist[3 * b], i = bondlist[3 * b + 1], j = bondlist[3 * b + 2];
        if (t < 1 || t > n_bond_types) return fail(ctx, GDPD_ERR_ARG, "bond %d: type exceeds n_bond_types", b);
        if (i < 1 || i > N || j < 1 || j > N) return fail(ctx, GDPD_ERR_ARG, "bond %d: atom id out of range", b);
        deg[i]++;
        deg[j]++;
    }
    std::vector<int> start(N + 1, 0);
    for (int k = 0; k < N; ++k) start[k + 1] = start[k] + deg[k + 1];
    std::vector<int> fillp(start.begin(), start.end() - 1), partner(std::max(1, 2 * n_bonds));
    std::vector<unsigned char> bt(std::max(1, 2 * n_bonds));
    for (int b = 0; b < n_bonds; ++b) {
        const int t = bondlist[3 * b] - 1, i = bondlist[3 * b + 1], j = bondlist[3 * b + 2];
        partner[fillp[i - 1]] = j;
        bt[fillp[i - 1]++] = (unsigned char)t;
        partner[fillp[j - 1]] = i;
        bt[fillp[j - 1]++] = (unsigned char)t;
    }
    for (int t = 0; t < n_bond_types; ++t) {
        h.bond_k[t] = bond_coeff[2 * t];
        h.bond_r0[t] = bond_coeff[2 * t + 1];
    }
    int rc;
    if ((rc = dalloc(ctx, s, &h.bond_partner, partner.size()))) return rc;
    if ((rc = dalloc(ctx, s, &h.bond_type, bt.size()))) return rc;
    HIPCHK(hipMemcpyAsync(h.bond_start, start.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.bond_partner, partner.data(), sizeof(int) * partner.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.bond_type, bt.data(), bt.size(), hipMemcpyHostToDevice, ctx->stream));
    h.nbonds = n_bonds;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return push_desc(ctx, sid);
}

int gdpd_set_angles(gdpd_ctx *ctx, int sid, int n_angles, int n_angle_types, const int *anglelist,
                    const double *angle_coeff)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || n_angles < 0 || (n_angles > 0 && (!anglelist || !angle_coeff)))
        return fail(ctx, GDPD_ERR_ARG, "bad angle arguments");
    if (n_angle_types < 0 || n_angle_types > kMaxAngleTypes)
        return fail(ctx, GDPD_ERR_ARG, "n_angle_types must be in 0..%d", kMaxAngleTypes);
    RepDev &h = s->h;
    const int N = h.N;
    // every angle is listed once for each of its three beads, with the bead's
    // role and the other two ids in i, j, k order
    std::vector<int> deg(N + 1, 0);
    for (int a = 0; a < n_angles; ++a) {
        const int t = anglelist[4 * a];
        if (t < 1 || t > n_angle_types) return fail(ctx, GDPD_ERR_ARG, "angle %d: type exceeds n_angle_types", a);
        for (int m = 1; m <= 3; ++m) {
            const int id = anglelist[4 * a + m];
            if (id < 1 || id > N) return fail(ctx, GDPD_ERR_ARG, "angle %d: atom id out of range", a);
            deg[id]++;
        }
    }
    std::vector<int> start(N + 1, 0);
    for (int k = 0; k < N; ++k) start[k + 1] = start[k] + deg[k + 1];
    std::vector<int> fillp(start.begin(), start.end() - 1), ent(4 * (size_t)std::max(1, 3 * n_angles), 0);
    for (int a = 0; a < n_angles; ++a) {
        const int t = anglelist[4 * a] - 1, i = anglelist[4 * a + 1], j = anglelist[4 * a + 2],
                  k = anglelist[4 * a + 3];
        const int mem[3] = {i, j, k};
        for (int role = 0; role < 3; ++role) {
            int *e = &ent[4 * (size_t)fillp[mem[role] - 1]++];
            e[0] = t | (role << 8);
            e[1] = role == 0 ? j : i;
            e[2] = role == 2 ? j : k;
            e[3] = 0;
        }
    }
    for (int t = 0; t < n_angle_types; ++t) {
        h.angle_k[t] = angle_coeff[2 * t];
        h.angle_th0[t] = angle_coeff[2 * t + 1] * (M_PI / 180.0);
    }
    int rc;
    if ((rc = dalloc(ctx, s, &h.angle_ent, ent.size()))) return rc;
    HIPCHK(hipMemcpyAsync(h.angle_start, start.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.angle_ent, ent.data(), sizeof(int) * ent.size(), hipMemcpyHostToDevice, ctx->stream));
    h.nangles = n_angles;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return push_desc(ctx, sid);
}

// ---- reference neighbour list ---------------------------------------------------

int gdpd_neigh_build(gdpd_ctx *ctx, int sid)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !s->uploaded) return fail(ctx, GDPD_ERR_STATE, "system %d has no state", sid);
    RepDev &h = s->h;
    const int N = h.N;
    const double bin = s->desc.global_cut + s->desc.skin;
    if (!(bin > 0.0)) return fail(ctx, GDPD_ERR_ARG, "global_cut + skin must be positive");
    if (!s->nl_n) {
        // force/neigh_list.cu:14-17, 64-68
        const double sphere = 4.0 / 3.0 * M_PI * std::pow(bin, 3.0);
        int nmax = (int)(s->desc.max_rho * sphere) + 1;
        s->nl_nmax = ((nmax + 15) / 16) * 16;
        long long nbins = 1;
        for (int d = 0; d < 3; ++d) {
            s->nl_nb[d] = (int)(h.L[d] / bin);
            if (s->nl_nb[d] < 3) return fail(ctx, GDPD_ERR_ARG, "fewer than 3 neighbour bins along axis %d", d);
            nbins *= s->nl_nb[d];
        }
        HIPCHK(hipMalloc((void **)&s->nl_binof, sizeof(int) * N));
        HIPCHK(hipMalloc((void **)&s->nl_count, sizeof(int) * nbins));
        HIPCHK(hipMalloc((void **)&s->nl_start, sizeof(int) * (nbins + 1)));
        HIPCHK(hipMalloc((void **)&s->nl_fill, sizeof(int) * nbins));
        HIPCHK(hipMalloc((void **)&s->nl_content, sizeof(int) * N));
        HIPCHK(hipMalloc((void **)&s->nl_n, sizeof(int) * N));
        HIPCHK(hipMalloc((void **)&s->nl_neigh, sizeof(int) * (size_t)N * s->nl_nmax));
        HIPCHK(hipMalloc((void **)&s->nl_over, sizeof(int)));
        HIPCHK(hipMalloc((void **)&s->nl_flag, sizeof(int)));
        HIPCHK(hipMalloc((void **)&s->nl_last, sizeof(double) * 3 * N));
    }
    const long long nbins = (long long)s->nl_nb[0] * s->nl_nb[1] * s->nl_nb[2];
    HIPCHK(hipMemsetAsync(s->nl_count, 0, sizeof(int) * nbins, ctx->stream));
    HIPCHK(hipMemsetAsync(s->nl_fill, 0, sizeof(int) * nbins, ctx->stream));
    HIPCHK(hipMemsetAsync(s->nl_over, 0, sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(s->nl_neigh, 0xff, sizeof(int) * (size_t)N * s->nl_nmax, ctx->stream));
    RepDev *dev = ctx->d_reps + sid;
    launch_nl_bin(dev, N, bin, s->nl_nb, s->nl_binof, s->nl_count, ctx->stream);
    std::vector<int> cnt(nbins), start(nbins + 1, 0);
    HIPCHK(hipMemcpyAsync(cnt.data(), s->nl_count, sizeof(int) * nbins, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (long long b = 0; b < nbins; ++b) start[b + 1] = start[b] + cnt[b];
    HIPCHK(hipMemcpyAsync(s->nl_start, start.data(), sizeof(int) * (nbins + 1), hipMemcpyHostToDevice, ctx->stream));
    launch_nl_fill(dev, N, s->nl_binof, s->nl_start, s->nl_fill, s->nl_content, ctx->stream);
    launch_nl_build(dev, N, bin, s->nl_nb, s->nl_binof, s->nl_start, s->nl_content, s->nl_nmax,
                    s->nl_n, s->nl_neigh, s->nl_over, s->nl_last, ctx->stream);
    int rc = check_launch(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    s->nl_built = true;
    return GDPD_OK;
}

int gdpd_neigh_check(gdpd_ctx *ctx, int sid, int *flag)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !flag) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    if (!s->nl_built) return fail(ctx, GDPD_ERR_STATE, "no neighbour list built");
    HIPCHK(hipMemsetAsync(s->nl_flag, 0, sizeof(int), ctx->stream));
    launch_nl_check(ctx->d_reps + sid, s->h.N, s->desc.skin, s->nl_last, s->nl_flag, ctx->stream);
    int rc = check_launch(ctx);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(flag, s->nl_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GDPD_OK;
}

int gdpd_neigh_download(gdpd_ctx *ctx, int sid, int *n_neigh, int *neigh, int n_max)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !n_neigh || !neigh || n_max <= 0) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    if (!s->nl_built) return fail(ctx, GDPD_ERR_STATE, "no neighbour list built");
    const int N = s->h.N, M = s->nl_nmax;
    std::vector<int> cnt(N), lst((size_t)N * M);
    int over = 0;
    HIPCHK(hipMemcpy(cnt.data(), s->nl_n, sizeof(int) * N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(lst.data(), s->nl_neigh, sizeof(int) * lst.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&over, s->nl_over, sizeof(int), hipMemcpyDeviceToHost));
    bool trunc = false;
    for (int i = 0; i < N; ++i) {
        std::vector<int> row(cnt[i]);
        for (int k = 0; k < cnt[i]; ++k) row[k] = lst[(size_t)k * N + i];
        std::sort(row.begin(), row.end());
        const int m = std::min((int)row.size(), n_max);
        if ((int)row.size() > n_max) trunc = true;
        n_neigh[i] = m;
        for (int k = 0; k < m; ++k) neigh[(size_t)i * n_max + k] = row[k];
    }
    if (over) return fail(ctx, GDPD_ERR_OVERFLOW, "a bead had more than n_max_neigh=%d list neighbours", M);
    if (trunc) return fail(ctx, GDPD_ERR_OVERFLOW, "row stride %d too small", n_max);
    return GDPD_OK;
}

// ---- plugin-hook path -------------------------------------------------------------

int gdpd_pair_compute(gdpd_ctx *ctx, int sid, uint32_t step, int log)
{
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    if (s->midstep) return fail(ctx, GDPD_ERR_STATE, "finish the fused run before using hooks");
    int rc = settle(ctx);
    if (rc) return rc;
    if ((rc = ensure_cells(ctx, s))) return rc;
    if ((rc = set_state_step(ctx, s, step, false))) return rc;
    const long long N = s->h.N;
    for (int attempt = 0; attempt < 16; ++attempt) {
        if (log) {
            launch_fill_f64(s->h.pe, 0.0, N, ctx->stream);
            launch_fill_f64(s->h.vir, 0.0, 6 * N, ctx->stream);
        }
        if ((rc = force_launch(ctx, s->tab, kModeForceOnly | (log ? kModeLogForce : 0)))) return rc;
        if (log) {
            launch_fix_to_double(s->h.pe, N, s->h.precision, ctx->stream);
            launch_fix_to_double(s->h.vir, 6 * N, s->h.precision, ctx->stream);
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
        bool again = false;
        if ((rc = held_step(ctx, s, again))) return rc;
        if (!again) break;
    }
    return check_flags(ctx, s);
}

int gdpd_bond_compute(gdpd_ctx *ctx, int sid, uint32_t step, int log)
{
    (void)step;
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    int rc = prepare(ctx, s);
    if (rc) return rc;
    if (s->h.nbonds == 0) return GDPD_OK;
    launch_bond_hook(ctx->d_reps, s->tab.beads, s->tab.n_beads, log, ctx->stream);
    if ((rc = check_launch(ctx))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GDPD_OK;
}

int gdpd_angle_compute(gdpd_ctx *ctx, int sid, uint32_t step, int log)
{
    (void)step;
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    int rc = prepare(ctx, s);
    if (rc) return rc;
    if (s->h.nangles == 0) return GDPD_OK;
    launch_angle_hook(ctx->d_reps, s->tab.beads, s->tab.n_beads, log, ctx->stream);
    if ((rc = check_launch(ctx))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GDPD_OK;
}

static int nve_hook(gdpd_ctx *ctx, int sid, int initial)
{
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    if (s->midstep) return fail(ctx, GDPD_ERR_STATE, "finish the fused run before using hooks");
    int rc = prepare(ctx, s);
    if (rc) return rc;
    if (s->fix_path) {
        // the nve fixes of the groups
        const unsigned nve = fix_mask(s, GDPD_FIX_NVE);
        if (initial) launch_fix_initial(ctx->d_reps, s->tab.beads, s->tab.n_beads, nve, ctx->stream);
        else launch_fix_final(ctx->d_reps, s->tab.beads, s->tab.n_beads, nve, ctx->stream);
    } else {
        launch_nve(ctx->d_reps, s->tab.beads, s->tab.n_beads, initial, ctx->stream);
    }
    if ((rc = check_launch(ctx))) return rc;
    if (initial) s->cells_valid = false;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return initial ? check_flags(ctx, s) : GDPD_OK;
}

int gdpd_nve_initial(gdpd_ctx *ctx, int sid) { return nve_hook(ctx, sid, 1); }
int gdpd_nve_final(gdpd_ctx *ctx, int sid) { return nve_hook(ctx, sid, 0); }

int gdpd_thermo(gdpd_ctx *ctx, int sid, double *out)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !out) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    int rc = prepare(ctx, s);
    if (rc) return rc;
    const int n = s->tab.n_beads;
    double *part = nullptr;
    HIPCHK(hipMalloc((void **)&part, sizeof(double) * kThermoParts * n));
    launch_thermo_arrays(ctx->d_reps, s->tab.beads, n, part, ctx->stream);
    std::vector<double> hp((size_t)kThermoParts * n);
    hipError_t e = hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(part);
    if (e != hipSuccess) return fail(ctx, GDPD_ERR_HIP, "thermo reduction failed: %s", hipGetErrorString(e));
    double tot[kThermoParts] = {0};
    for (int b = 0; b < n; ++b)
        for (int k = 0; k < kThermoParts; ++k) tot[k] += hp[(size_t)b * kThermoParts + k];
    const RepDev &h = s->h;
    const double vol = h.L[0] * h.L[1] * h.L[2];
    out[0] = (2.0 * tot[0]) / ((3.0 * h.N - 3.0) * 1.0);
    out[1] = tot[1];
    out[2] = tot[2];
    for (int k = 0; k < 6; ++k) out[4 + k] = tot[3 + k] / vol;
    out[3] = (out[4] + out[7] + out[9]) / 3.0;
    out[10] = tot[0];
    out[11] = -1.0;
    out[12] = tot[9];
    return GDPD_OK;
}

int gdpd_thermo_noalloc(gdpd_ctx *ctx, int sid, double *out)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !out) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    if (!s->prepared || !s->uploaded)
        return fail(ctx, GDPD_ERR_STATE, "system %d: compute it once before the allocation-free thermo", sid);
    const int n = s->tab.n_beads;
    if (n > kNoAllocMaxChunks)
        return fail(ctx, GDPD_ERR_ARG, "the allocation-free thermo takes at most %d beads (system %d has %d)",
                    kNoAllocMaxChunks * kBlock, sid, s->h.N);
    double *ptr = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_dev[ctx->device] : nullptr;
    if (!ptr) return fail(ctx, GDPD_ERR_HIP, "no address for the device scratch array of device %d", ctx->device);
    RepState st;
    HIPCHK(hipMemcpyAsync(&st, s->h.state, sizeof(RepState), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (st.flags & kFlagHalt)
        return fail(ctx, GDPD_ERR_STATE, "system %d is held after a staging overflow: call gdpd_sync first", sid);
    // every chunk's partials are written before they are summed; nothing needs clearing
    double *total = ptr + kScratchDoubles - kThermoParts;
    launch_thermo_arrays(ctx->d_reps, s->tab.beads, n, ptr, ctx->stream);
    launch_thermo_total(ptr, n, total, ctx->stream);
    int rc = check_launch(ctx);
    if (rc) return rc;
    double tot[kThermoParts];
    HIPCHK(hipMemcpyAsync(tot, total, sizeof(tot), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const RepDev &h = s->h;
    const double vol = h.L[0] * h.L[1] * h.L[2];
    out[0] = (2.0 * tot[0]) / ((3.0 * h.N - 3.0) * 1.0);
    out[1] = tot[1];
    out[2] = tot[2];
    for (int k = 0; k < 6; ++k) out[4 + k] = tot[3 + k] / vol;
    out[3] = (out[4] + out[7] + out[9]) / 3.0;
    out[10] = tot[0];
    out[11] = -1.0;
    out[12] = tot[9];
    return GDPD_OK;
}

// ---- fused path ----------------------------------------------------------------------

int gdpd_zero_step(gdpd_ctx *ctx, int sid)
{
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    int rc;
    if ((rc = settle(ctx))) return rc;
    if (s->fix_path) {
        if ((rc = prepare(ctx, s))) return rc;
        if ((rc = fix_zero_step(ctx, s, true))) return rc;
        s->midstep = true;
        s->step_host = 1;
        ctx->pending.push_back(sid);
        return GDPD_OK;
    }
    if ((rc = ensure_cells(ctx, s))) return rc;
    if ((rc = set_state_step(ctx, s, 0, true))) return rc;
    if ((rc = step_once(ctx, s->tab, kModeDrift))) return rc;
    if ((rc = maybe_resort(ctx, s, s->tab, 1))) return rc;
    s->midstep = true;
    s->step_host = 1;
    ctx->pending.push_back(sid);
    return GDPD_OK;
}

int gdpd_advance(gdpd_ctx *ctx, int sid, int nsteps)
{
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    if (!s->midstep) return fail(ctx, GDPD_ERR_STATE, "call gdpd_zero_step first");
    if (nsteps <= 0) return GDPD_OK;
    int rc = settle(ctx);
    if (rc) return rc;
    s->step_host += nsteps;
    ctx->pending.push_back(sid);
    if (s->fix_path) {
        for (int k = 0; k < nsteps; ++k)
            if ((rc = fix_iteration(ctx, s, true, true, true))) return rc;
        return GDPD_OK;
    }
    const int gs = s->graph_steps > 0 ? s->graph_steps : 8;
    return advance_tables(ctx, s->tab, s->graph, nsteps, gs, {s});
}

int gdpd_finish(gdpd_ctx *ctx, int sid)
{
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    if (!s->midstep) return fail(ctx, GDPD_ERR_STATE, "call gdpd_zero_step first");
    int rc = settle(ctx);
    if (rc) return rc;
    const long long n = s->step_host;
    for (int attempt = 0; attempt < 16; ++attempt) {
        if (s->fix_path) s->fix_step = n;
        rc = s->fix_path ? fix_iteration(ctx, s, true, true, false)
                         : step_once(ctx, s->tab, kModeFinalKick | kModeStoreForce);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        bool again = false;
        if ((rc = held_step(ctx, s, again))) return rc;
        if (!again) break;
    }
    s->midstep = false;
    s->step_host = n + 1;
    return check_flags(ctx, s);
}

int gdpd_run(gdpd_ctx *ctx, int sid, int nsteps)
{
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    if (nsteps < 0) return fail(ctx, GDPD_ERR_ARG, "nsteps must be >= 0");
    int rc = ensure_cells(ctx, s);
    if (rc) return rc;
    if (nsteps == 0 && s->fix_path) {
        if ((rc = fix_zero_step(ctx, s, false))) return rc;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return check_flags(ctx, s);
    }
    if (nsteps == 0) {
        if ((rc = settle(ctx))) return rc;
        if ((rc = set_state_step(ctx, s, 0, true))) return rc;
        for (int attempt = 0; attempt < 16; ++attempt) {
            if ((rc = step_once(ctx, s->tab, kModeStoreForce))) return rc;
            HIPCHK(hipStreamSynchronize(ctx->stream));
            bool again = false;
            if ((rc = held_step(ctx, s, again))) return rc;
            if (!again) break;
        }
        s->step_host = 1;
        return check_flags(ctx, s);
    }
    if ((rc = gdpd_zero_step(ctx, sid))) return rc;
    if ((rc = gdpd_advance(ctx, sid, nsteps - 1))) return rc;
    return gdpd_finish(ctx, sid);
}

int gdpd_thermo_rows(gdpd_ctx *ctx, int sid, double *out, int max_rows, int *n_rows)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !n_rows) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    int rc = settle(ctx);
    if (rc) return rc;
    RepState st;
    if ((rc = read_state(ctx, s, &st))) return rc;
    const int cap = s->h.thermo_cap;
    const int have = std::min(st.log_rows, cap);
    const int first = st.log_rows - have;
    std::vector<double> rows((size_t)cap * kThermoRowLen);
    HIPCHK(hipMemcpy(rows.data(), s->h.thermo_rows, sizeof(double) * rows.size(), hipMemcpyDeviceToHost));
    const int n = std::min(have, max_rows);
    for (int k = 0; k < n && out; ++k) {
        const int r = (first + k) % cap;
        memcpy(out + (size_t)k * GDPD_THERMO_LEN, rows.data() + (size_t)r * kThermoRowLen,
               sizeof(double) * kThermoRowLen);
    }
    *n_rows = n;
    return GDPD_OK;
}

int gdpd_get_step(gdpd_ctx *ctx, int sid, int64_t *step)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !step) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    int rc = settle(ctx);
    if (rc) return rc;
    RepState st;
    if ((rc = read_state(ctx, s, &st))) return rc;
    *step = st.step;
    return GDPD_OK;
}


// ---- group fixes and computes --------------------------------------------------

int gdpd_set_fixes(gdpd_ctx *ctx, int sid, const gdpd_fix *fixes, int n)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || n < 0 || (n > 0 && !fixes)) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    if (s->midstep) return fail(ctx, GDPD_ERR_STATE, "finish the fused run before changing fixes");
    RepDev &h = s->h;
    const unsigned all = all_types(s);
    std::vector<gdpd_fix> v(fixes, fixes + n);
    bool has_nve = false;
    unsigned integrated = 0;
    for (int k = 0; k < n; ++k) {
        const gdpd_fix &f = v[k];
        if (f.kind < GDPD_FIX_NVE || f.kind > GDPD_FIX_RVCM) return fail(ctx, GDPD_ERR_ARG, "fix %d: unknown kind %d", k, f.kind);
        if (f.group == 0 || (f.group & ~all)) return fail(ctx, GDPD_ERR_ARG, "fix %d: group names a type outside 1..%d", k, h.ntypes);
        if (f.kind == GDPD_FIX_NVE || f.kind == GDPD_FIX_MOVE) {
            // a bead integrated twice per step is rejected (the reference would apply both)
            if (integrated & f.group) return fail(ctx, GDPD_ERR_ARG, "fix %d: a type is already integrated by another fix", k);
            integrated |= f.group;
            has_nve = has_nve || f.kind == GDPD_FIX_NVE;
        }
        if ((f.kind == GDPD_FIX_WALL_REFLECT || f.kind == GDPD_FIX_TWO_WALL) && (f.dir < 0 || f.dir > 2))
            return fail(ctx, GDPD_ERR_ARG, "fix %d: wall direction must be x, y or z", k);
        if (f.kind == GDPD_FIX_WALL_REFLECT) {
            // fix_wall_reflect.cu:25-27, 47-66
            if (!(f.p[0] < f.p[1])) return fail(ctx, GDPD_ERR_ARG, "fix %d: lo_wall_pos should lower than hi_wall_pos", k);
            if (f.p[0] < h.lo[f.dir] || f.p[1] > h.lo[f.dir] + h.L[f.dir])
                return fail(ctx, GDPD_ERR_ARG, "fix %d: Wall position is out of the box!", k);
        }
        if (f.kind == GDPD_FIX_TWO_WALL && (f.group2 == 0 || f.group3 == 0 || ((f.group2 | f.group3) & ~all)))
            return fail(ctx, GDPD_ERR_ARG, "fix %d: two_wall needs two wall groups of known types", k);
        if ((f.kind == GDPD_FIX_RECENTER || f.kind == GDPD_FIX_RVCM) && f.every <= 0)
            return fail(ctx, GDPD_ERR_ARG, "fix %d: frequency must be positive", k);
        if (f.kind == GDPD_FIX_SPRING && f.every < 0) return fail(ctx, GDPD_ERR_ARG, "fix %d: file frequency must be >= 0", k);
        if (f.kind == GDPD_FIX_SPRING || f.kind == GDPD_FIX_RECENTER) {
            // fix_spring.cu:110-125, fix_recenter.cu:75-90
            const double *c = f.kind == GDPD_FIX_SPRING ? f.p + 1 : f.p;
            for (int d = 0; d < 3; ++d)
                if (c[d] < h.lo[d] || c[d] > h.lo[d] + h.L[d])
                    return fail(ctx, GDPD_ERR_ARG, "fix %d: position is out of the box in %c direction!", k, "xyz"[d]);
        }
    }
    (void)has_nve;
    if ((int)v.size() > kMaxFixes) return fail(ctx, GDPD_ERR_ARG, "at most %d fixes", kMaxFixes);
    s->fixes = v;
    finish_fix_setup(s);
    return push_desc(ctx, sid);
}

int gdpd_set_computes(gdpd_ctx *ctx, int sid, const gdpd_compute *c, int n)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || n < 0 || (n > 0 && !c)) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    if (s->midstep) return fail(ctx, GDPD_ERR_STATE, "finish the fused run before changing computes");
    if (n > kMaxComputes) return fail(ctx, GDPD_ERR_ARG, "at most %d computes", kMaxComputes);
    RepDev &h = s->h;
    const unsigned all = all_types(s);
    bool msd = false;
    for (int k = 0; k < n; ++k) {
        const gdpd_compute &q = c[k];
        if (q.kind < GDPD_COMPUTE_TEMP || q.kind > GDPD_COMPUTE_WALL_CONTACT)
            return fail(ctx, GDPD_ERR_ARG, "compute %d: unknown kind %d", k, q.kind);
        if (q.group == 0 || (q.group & ~all)) return fail(ctx, GDPD_ERR_ARG, "compute %d: group names an unknown type", k);
        if (q.every <= 0) return fail(ctx, GDPD_ERR_ARG, "compute %d: frequency must be positive", k);
        if (q.kind == GDPD_COMPUTE_RHO) {
            // compute_rho.cu:55-85
            for (int d = 0; d < 3; ++d) {
                if (q.p[2 * d] < h.lo[d]) return fail(ctx, GDPD_ERR_ARG, "compute %d: %clo < box.%clo", k, "xyz"[d], "xyz"[d]);
                if (q.p[2 * d + 1] > h.lo[d] + h.L[d])
                    return fail(ctx, GDPD_ERR_ARG, "compute %d: %chi > box.%chi", k, "xyz"[d], "xyz"[d]);
                if (!(q.p[2 * d] < q.p[2 * d + 1])) return fail(ctx, GDPD_ERR_ARG, "compute %d: empty region", k);
            }
        }
        if (q.kind == GDPD_COMPUTE_WALL_CONTACT &&
            (q.group2 == 0 || (q.group2 & ~all) || (q.side != 1 && q.side != -1) || q.dir < 0 || q.dir > 2 || !(q.p[0] > 0)))
            return fail(ctx, GDPD_ERR_ARG, "compute %d: wall_contact needs a wall group, lo|hi, x|y|z and a positive cutoff", k);
        msd = msd || q.kind == GDPD_COMPUTE_MSD;
    }
    if (msd) {
        // reference positions: the unwrapped state now (compute_msd.cu:40-50), by id
        if (!s->uploaded) return fail(ctx, GDPD_ERR_STATE, "upload the state before setting an msd compute");
        const int N = h.N;
        std::vector<double> pos(3 * (size_t)N), x0(3 * (size_t)N);
        std::vector<int> img(3 * (size_t)N);
        int rc = gdpd_download(ctx, sid, pos.data(), nullptr, img.data(), nullptr);
        if (rc) return rc;
        for (int i = 0; i < N; ++i)
            for (int d = 0; d < 3; ++d)
                x0[(size_t)d * N + i] = pos[3 * (size_t)i + d] + img[3 * (size_t)i + d] * h.L[d];
        if (!h.x0uw && (rc = dalloc(ctx, s, &h.x0uw, 3 * (size_t)N))) return rc;
        HIPCHK(hipMemcpy(h.x0uw, x0.data(), sizeof(double) * x0.size(), hipMemcpyHostToDevice));
    }
    s->computes.assign(c, c + n);
    h.ncomp = n;
    for (int k = 0; k < n; ++k) {
        CompDev &C = h.comp[k];
        C = CompDev{};
        C.kind = c[k].kind;
        C.group = c[k].group;
        C.group2 = c[k].group2;
        C.every = c[k].every;
        C.dir = c[k].dir;
        C.side = c[k].side;
        for (int q = 0; q < 6; ++q) C.p[q] = c[k].p[q];
    }
    finish_fix_setup(s);
    return push_desc(ctx, sid);
}

// newest record of a slot, copied out after a sync
static int last_record(gdpd_ctx *ctx, Sys *s, int slot, double *out)
{
    int nrec = 0;
    HIPCHK(hipMemcpy(&nrec, s->h.rec_n + slot, sizeof(int), hipMemcpyDeviceToHost));
    if (nrec <= 0) return fail(ctx, GDPD_ERR_STATE, "no record");
    const int row = (nrec - 1) % kRecCap;
    HIPCHK(hipMemcpy(out, s->h.rec_rows + ((size_t)slot * kRecCap + row) * kRecLen, sizeof(double) * kRecLen,
                     hipMemcpyDeviceToHost));
    return GDPD_OK;
}

int gdpd_fix_hook(gdpd_ctx *ctx, int sid, int index, int hook, uint32_t step, double *out5)
{
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    if (s->midstep) return fail(ctx, GDPD_ERR_STATE, "finish the fused run before using hooks");
    if (index < 0 || index >= s->h.nfix) return fail(ctx, GDPD_ERR_ARG, "no fix %d", index);
    int rc = prepare(ctx, s);
    if (rc) return rc;
    if ((rc = check_groups(ctx, s))) return rc;
    Tables &T = s->tab;
    hipStream_t st = ctx->stream;
    const FixDev &F = s->h.fix[index];
    const unsigned one = 1u << index;
    const bool fire = F.every > 0 && (step % (uint32_t)F.every) == 0;
    bool moved = false;
    switch (hook) {
    case GDPD_HOOK_INITIAL_INTEGRATE:
        if (F.kind == GDPD_FIX_NVE || F.kind == GDPD_FIX_MOVE) {
            launch_fix_initial(ctx->d_reps, T.beads, T.n_beads, one, st);
            moved = true;
        }
        break;
    case GDPD_HOOK_POST_INTEGRATE:
        if (F.kind == GDPD_FIX_WALL_REFLECT) {
            launch_fix_initial(ctx->d_reps, T.beads, T.n_beads, one, st);
            moved = true;
        } else if (F.kind == GDPD_FIX_TWO_WALL) {
            launch_fix_clear(ctx->d_reps, T.rep_list, T.n_reps, st);
            launch_fix_initial(ctx->d_reps, T.beads, T.n_beads, one, st);
            launch_fix_two_wall(ctx->d_reps, T.beads, T.n_beads, one, st);
            moved = true;
        }
        break;
    case GDPD_HOOK_POST_FORCE:
        if (F.kind == GDPD_FIX_SPRING) {
            launch_fix_clear(ctx->d_reps, T.rep_list, T.n_reps, st);
            launch_fix_reduce(ctx->d_reps, T.beads, T.n_beads, one, 0u, 0, st);
            launch_fix_post_force(ctx->d_reps, T.beads, T.n_beads, one, st);
            launch_fix_records(ctx->d_reps, T.rep_list, T.n_reps, one, 0u, (long long)step, st);
        }
        break;
    case GDPD_HOOK_FINAL_INTEGRATE:
        if (F.kind == GDPD_FIX_NVE) launch_fix_final(ctx->d_reps, T.beads, T.n_beads, one, st);
        break;
    case GDPD_HOOK_END_OF_STEP:
        if ((F.kind == GDPD_FIX_RVCM || F.kind == GDPD_FIX_RECENTER) && fire) {
            launch_fix_clear(ctx->d_reps, T.rep_list, T.n_reps, st);
            launch_fix_reduce(ctx->d_reps, T.beads, T.n_beads, one, 0u, F.kind == GDPD_FIX_RVCM ? 1 : 0, st);
            launch_fix_end(ctx->d_reps, T.beads, T.n_beads, one, st);
            moved = F.kind == GDPD_FIX_RECENTER;
        }
        break;
    default:
        return fail(ctx, GDPD_ERR_ARG, "unknown hook %d", hook);
    }
    if ((rc = check_launch(ctx))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (moved) s->cells_valid = false;
    if (out5 && hook == GDPD_HOOK_POST_FORCE && F.kind == GDPD_FIX_SPRING) {
        double r[kRecLen];
        if ((rc = last_record(ctx, s, kMaxComputes + index, r))) return rc;
        for (int q = 0; q < 5; ++q) out5[q] = r[1 + q];
    }
    return check_flags(ctx, s);
}

int gdpd_compute_eval(gdpd_ctx *ctx, int sid, int index, double *out3)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !out3) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    if (index < 0 || index >= s->h.ncomp) return fail(ctx, GDPD_ERR_ARG, "no compute %d", index);
    int rc = prepare(ctx, s);
    if (rc) return rc;
    if ((rc = check_groups(ctx, s))) return rc;
    Tables &T = s->tab;
    hipStream_t st = ctx->stream;
    const unsigned one = 1u << index;
    const int kind = s->h.comp[index].kind;
    launch_fix_clear(ctx->d_reps, T.rep_list, T.n_reps, st);
    if (kind == GDPD_COMPUTE_TF) {
        launch_fix_reduce(ctx->d_reps, T.beads, T.n_beads, 0u, one, 1, st);
    } else {
        launch_fix_reduce(ctx->d_reps, T.beads, T.n_beads, 0u, one, 2, st);
        if (kind == GDPD_COMPUTE_WALL_CONTACT) launch_fix_reduce(ctx->d_reps, T.beads, T.n_beads, 0u, one, 3, st);
    }
    launch_fix_records(ctx->d_reps, T.rep_list, T.n_reps, 0u, one, -1, st);
    if ((rc = check_launch(ctx))) return rc;
    HIPCHK(hipStreamSynchronize(st));
    double r[kRecLen];
    if ((rc = last_record(ctx, s, index, r))) return rc;
    for (int q = 0; q < 3; ++q) out3[q] = r[1 + q];
    return GDPD_OK;
}

int gdpd_records(gdpd_ctx *ctx, int sid, int index, double *out, int max_rows, int *n_rows)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !n_rows || index < 0 || index >= kRecSlots) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    int rc = settle(ctx);
    if (rc) return rc;
    int nrec = 0;
    HIPCHK(hipMemcpy(&nrec, s->h.rec_n + index, sizeof(int), hipMemcpyDeviceToHost));
    const int have = std::min(nrec, kRecCap);
    const int first = nrec - have;
    std::vector<double> rows((size_t)kRecCap * kRecLen);
    HIPCHK(hipMemcpy(rows.data(), s->h.rec_rows + (size_t)index * kRecCap * kRecLen, sizeof(double) * rows.size(),
                     hipMemcpyDeviceToHost));
    const int n = std::min(have, std::max(0, max_rows));
    for (int k = 0; k < n && out; ++k)
        memcpy(out + (size_t)k * kRecLen, rows.data() + (size_t)((first + k) % kRecCap) * kRecLen,
               sizeof(double) * kRecLen);
    *n_rows = n;
    return GDPD_OK;
}

int gdpd_batch_advance(gdpd_ctx *ctx, const int *systems, int n, int nsteps)
{
    if (!ctx || !systems || n <= 0) return fail(ctx, GDPD_ERR_ARG, "bad batch arguments");
    std::vector<int> ids(systems, systems + n);
    std::vector<Sys *> members;
    int prec = -1;
    for (int id : ids) {
        Sys *s = get_sys(ctx, id);
        if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", id);
        if (!s->midstep) return fail(ctx, GDPD_ERR_STATE, "system %d: call gdpd_zero_step first", id);
        if (s->fix_path) return fail(ctx, GDPD_ERR_ARG, "system %d has group fixes or computes: advance it alone", id);
        if (prec >= 0 && s->h.precision != prec) return fail(ctx, GDPD_ERR_ARG, "a batch must share one precision");
        prec = s->h.precision;
        members.push_back(s);
    }
    int rc;
    if (nsteps <= 0) return GDPD_OK;
    if ((rc = settle(ctx))) return rc;
    if (ctx->batch.key != ids) {
        ctx->batch_graph.release();
        if ((rc = build_tables(ctx, ids, ctx->batch))) return rc;
    }
    int gs = 8;
    for (Sys *s : members)
        if (s->graph_steps > 0) gs = s->graph_steps;
    for (size_t k = 0; k < members.size(); ++k) {
        members[k]->step_host += nsteps;
        ctx->pending.push_back(ids[k]);
    }
    return advance_tables(ctx, ctx->batch, ctx->batch_graph, nsteps, gs, members);
}

int gdpd_set_tuning(gdpd_ctx *ctx, int sid, int bx, int by, int bz, int resort_every, int graph_steps)
{
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    if (s->prepared && (bx || by || bz))
        return fail(ctx, GDPD_ERR_STATE, "brick shape is fixed once the system has run");
    s->tune_b[0] = bx;
    s->tune_b[1] = by;
    s->tune_b[2] = bz;
    if (resort_every >= 0) s->resort_every = resort_every;
    if (graph_steps >= 0) {
        s->graph_steps = graph_steps;
        s->graph.release();
    }
    return GDPD_OK;
}

int gdpd_set_capacity(gdpd_ctx *ctx, int sid, int smax, int omax)
{
    Sys *s = get_sys(ctx, sid);
    if (!s) return fail(ctx, GDPD_ERR_ARG, "no system %d", sid);
    if (smax < 0 || omax < 0 || smax > 65000 || omax > 65000)
        return fail(ctx, GDPD_ERR_ARG, "capacities must be in 0..65000");
    int rc = settle(ctx);
    if (rc) return rc;
    s->cap_s = smax;
    s->cap_o = omax;
    if (!s->prepared) return GDPD_OK;
    if ((rc = apply_set_capacity(ctx, s))) return rc;
    if ((rc = push_desc(ctx, sid))) return rc;
    if ((rc = build_tables(ctx, {sid}, s->tab))) return rc;
    s->graph.release();
    ctx->batch_graph.release();
    ctx->batch.release();
    return GDPD_OK;
}

int gdpd_get_tuning(gdpd_ctx *ctx, int sid, int *out8)
{
    Sys *s = get_sys(ctx, sid);
    if (!s || !out8) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    int rc = prepare(ctx, s);
    if (rc) return rc;
    out8[0] = s->h.bdim[0];
    out8[1] = s->h.bdim[1];
    out8[2] = s->h.bdim[2];
    out8[3] = s->h.smax;
    out8[4] = s->h.omax;
    out8[5] = s->resort_every > 0 ? s->resort_every : kDefaultResortEvery;
    out8[6] = s->graph_steps > 0 ? s->graph_steps : 8;
    out8[7] = s->h.nbrick;
    return GDPD_OK;
}

int gdpd_kernel_timing(gdpd_ctx *ctx, int on)
{
    if (!ctx) return GDPD_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->timing = on != 0;
    ctx->ev_used = 0;
    ctx->force_launches = 0;
    ctx->phase_wave_launches = 0;
    if (ctx->d_phase) HIPCHK(hipMemset(ctx->d_phase, 0, ctx->phase_items * 16 * sizeof(unsigned)));
    return GDPD_OK;
}

int gdpd_kernel_phases(gdpd_ctx *ctx, double *us4)
{
    if (!ctx || !us4) return GDPD_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double sum[4] = {0, 0, 0, 0};
    if (ctx->d_phase) {
        std::vector<unsigned> h(ctx->phase_items * 16);
        HIPCHK(hipMemcpy(h.data(), ctx->d_phase, h.size() * sizeof(unsigned), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < h.size(); ++k) sum[k & 3] += h[k];
    }
    for (int k = 0; k < 4; ++k)
        us4[k] = ctx->phase_wave_launches > 0 ? sum[k] * ctx->tick_ns * 1e-3 / ctx->phase_wave_launches : 0.0;
    return GDPD_OK;
}

int gdpd_kernel_timing_get(gdpd_ctx *ctx, double *force_ms, int64_t *force_launches, double *total_ms)
{
    if (!ctx) return GDPD_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double sum = 0.0;
    for (size_t k = 0; k + 1 < ctx->ev_used; k += 2) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev_pool[k], ctx->ev_pool[k + 1]));
        sum += ms;
    }
    double tot = 0.0;
    if (ctx->ev_used >= 2) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev_pool[0], ctx->ev_pool[ctx->ev_used - 1]));
        tot = ms;
    }
    if (force_ms) *force_ms = sum;
    if (force_launches) *force_launches = ctx->force_launches;
    if (total_ms) *total_ms = tot;
    return GDPD_OK;
}

int gdpd_sync(gdpd_ctx *ctx)
{
    if (!ctx) return GDPD_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    flush_graveyard(ctx);
    return settle(ctx);
}

int gdpd_step_times(gdpd_ctx *ctx, const int *systems, int n, uint64_t *times, int64_t *steps, double *tick_ns)
{
    if (!ctx || n < 0 || (n > 0 && (!systems || !times || !steps))) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    int rc = settle(ctx);
    if (rc) return rc;
    if (tick_ns) *tick_ns = ctx->tick_ns;
    if (n == 0) return GDPD_OK;
    std::vector<int> ids(systems, systems + n);
    for (int id : ids)
        if (!get_sys(ctx, id)) return fail(ctx, GDPD_ERR_ARG, "no system %d", id);
    std::vector<RepState> st;
    if ((rc = read_states(ctx, ids, st))) return rc;  // sizes the gather buffers
    HIPCHK(hipMemcpyAsync(ctx->d_ids, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    launch_gather_step_ts(ctx->d_reps, ctx->d_ids, n, ctx->d_ts, ctx->d_tsteps, ctx->stream);
    if ((rc = check_launch(ctx))) return rc;
    static_assert(sizeof(unsigned long long) == sizeof(uint64_t), "64-bit clock words");
    HIPCHK(hipMemcpyAsync(times, ctx->d_ts, sizeof(uint64_t) * kStepTsCap * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(steps, ctx->d_tsteps, sizeof(int64_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GDPD_OK;
}

int gdpd_saru_eval(gdpd_ctx *ctx, int n, const uint32_t *tag1, const uint32_t *tag2, uint32_t seed,
                   const int32_t *steps, uint32_t *bits, float *r32, double *r64)
{
    if (!ctx || n < 0 || (n > 0 && (!tag1 || !tag2 || !steps))) return fail(ctx, GDPD_ERR_ARG, "bad arguments");
    if (n == 0) return GDPD_OK;
    char *buf = nullptr;
    const size_t b4 = sizeof(uint32_t) * (size_t)n;
    HIPCHK(hipMalloc((void **)&buf, b4 * 6 + sizeof(double) * (size_t)n));
    uint32_t *d1 = (uint32_t *)buf, *d2 = d1 + n, *db = d2 + n;
    int32_t *ds = (int32_t *)(db + n);
    float *dr32 = (float *)(ds + n);
    double *dr64 = (double *)(buf + b4 * 6);
    hipError_t e = hipMemcpyAsync(d1, tag1, b4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d2, tag2, b4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ds, steps, b4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launch_saru_eval(n, d1, d2, seed, ds, db, dr32, dr64, ctx->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess && bits) e = hipMemcpyAsync(bits, db, b4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && r32) e = hipMemcpyAsync(r32, dr32, b4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && r64) e = hipMemcpyAsync(r64, dr64, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(ctx, GDPD_ERR_HIP, "saru eval failed: %s", hipGetErrorString(e));
    return GDPD_OK;
}

} // extern "C"
