// gdpd_kernels.hip -- gfx950 kernels of the DPD hot path.
//
// Data layout (per replica, all in HBM):
//   x[2], v[2]   FP64 SoA [X | Y | Z], double-buffered so the fused step can read
//                every neighbour's old state while it writes the new one;
//   id, type     per slot; img packed image counts;
//   cell_start   CSR over cells (cell edge >= the largest pair cutoff, z fastest like
//                the reference's flat bin index, force/neigh_list.cu:280);
//   perm         sorted position -> slot; slots are physically re-sorted into cell
//                order every few steps so perm stays close to the identity.
//
// The hot kernel (k_brick) gives each 256-thread workgroup a brick of cells.  It
// stages the brick plus a one-cell halo from HBM into LDS (positions relative
// to the brick corner, periodic images applied while staging), finds each owned
// bead's in-range partners in LDS and evaluates every pair's DPD/mDPD force
// with the Saru random term.  k_integrate then adds each bead's bonds and angles
// and integrates it (final half-kick of step n, thermo on log steps, half-kick
// and drift of step n+1), writing the new state to the other buffer and its new
// cell.  Per-bead and per-brick sums are 64-bit fixed point, so results do not
// depend on scheduling or on which other replicas share the launch.
#include "gdpd_kernels.h"
#include "gdpd_saru.h"

namespace gdpd {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs; hand each XCD a
// contiguous run of items so neighbouring bricks share that XCD's L2.
__device__ __forceinline__ int xcd_item(int nitems)
{
    const int b = blockIdx.x;
    const int q = nitems / kNumXcd, r = nitems % kNumXcd;
    const int x = b % kNumXcd, k = b / kNumXcd;
    return x * q + min(x, r) + k;
}

// A pointer into HBM as a global-address-space pointer: its loads and stores
// count against vmcnt only, so a kernel's waits for LDS results do not also
// wait for HBM traffic.
template <typename T> __device__ __forceinline__ const __attribute__((address_space(1))) T *gmem(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}
template <typename T> __device__ __forceinline__ __attribute__((address_space(1))) T *gmemw(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// A halted replica (kFlagHalt) keeps its state: the step kernels skip it whole.
// The flag is per replica, so the test is uniform across a workgroup.
__device__ __forceinline__ bool halted(const RepDev &R) { return (R.state->flags & kFlagHalt) != 0; }

// A replica's step counter as a wave-uniform value, so what derives from it
// (the log decision, the Saru step key) stays in scalar registers.
__device__ __forceinline__ long long uniform_step(const RepDev &R)
{
    const long long v = R.state->step;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// step % every == 0 for a log interval every > 0.  Steps below 2^31 (every
// run in practice) take a 32-bit remainder, a few instructions; the general
// 64-bit remainder expands to about 85 vector instructions per wave.
__device__ __forceinline__ bool log_due(long long step, int every)
{
    if (every <= 0) return false;
    if ((unsigned long long)step < 0x80000000ull) return (unsigned)step % (unsigned)every == 0u;
    return step % every == 0;
}


__device__ __forceinline__ int cell_axis(double x, double lo, double inv_cs, int nc)
{
    int c = (int)((x - lo) * inv_cs);
    return min(max(c, 0), nc - 1);
}

__device__ __forceinline__ int cell_of(const RepDev &R, double x, double y, double z)
{
    int cx = cell_axis(x, R.lo[0], R.inv_cs[0], R.nc[0]);
    int cy = cell_axis(y, R.lo[1], R.inv_cs[1], R.nc[1]);
    int cz = cell_axis(z, R.lo[2], R.inv_cs[2], R.nc[2]);
    return (cx * R.nc[1] + cy) * R.nc[2] + cz;
}

// minimum image of force/neigh_list.cu:377-379
__device__ __forceinline__ double min_image(double d, double L, double hL)
{
    return d - L * (double)((d >= hL) - (d < -hL));
}

// wrap into [lo, hi) by whole box lengths like fix/fix_nve.cu:68-76, counting
// images; a non-finite or runaway coordinate is reported through `bad`.
__device__ __forceinline__ double wrap_axis(double x, double lo, double hi, double L, int &img,
                                            bool &bad)
{
    for (int it = 0; it < 16 && (x >= hi || x < lo); ++it) {
        const int up = (x >= hi), dn = (x < lo);
        x -= L * (double)(up - dn);
        img += up - dn;
    }
    if (!(x >= lo && x < hi)) bad = true;
    return x;
}

// wavefront inclusive prefix sum and 64-bit sum (DPP, no LDS round trips)
extern "C" __device__ int __ockl_wfscan_add_i32(int, bool);
extern "C" __device__ long __ockl_wfred_add_i64(long);
__device__ __forceinline__ int wave_incl_scan(int x) { return __ockl_wfscan_add_i32(x, true); }

// exclusive scan over the 256 threads; every thread must call it
__device__ __forceinline__ int block_excl_scan(int v, int *wbuf, int &total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) wbuf[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kBlock / 64; ++k) {
        int s = wbuf[k];
        off += (k < w) ? s : 0;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// fixed-order block sum of n doubles per thread into out[0..n) of thread 0
template <int NV>
__device__ __forceinline__ void block_sum(double (&acc)[NV], double *red, double *out)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = wave_sum(acc[k]);
        if (lane == 0) red[w * NV + k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int ww = 0; ww < kBlock / 64; ++ww) s += red[ww * NV + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// Thermo partial sums as 64-bit fixed point (2^-32 resolution, rounded to
// nearest, |sum| < 2^31): integer sums are exact, so a thermo record does not
// depend on how beads are split over workgroups (slot layout, brick shape or
// batch).
constexpr double kThermoFix = 4294967296.0;


// face classes of a brick's owned cells (see face_class)
constexpr int kSigClasses = 32;

// LDS hash table of the per-block cell grouping in k_integrate
constexpr int kRankBits = 9;
constexpr int kRankSlots = 1 << kRankBits;
static_assert(kRankSlots > kBlock, "the rank table must hold every key of a block");

__device__ __forceinline__ long long wave_sum_i64(long long v) { return (long long)__ockl_wfred_add_i64((long)v); }

__device__ __forceinline__ long long thermo_fix(double v) { return __double2ll_rn(v * kThermoFix); }

// exact block sum of per-thread integer partials; thread k < NV of the block
// gets the total of component k in out[k]
template <int NV>
__device__ __forceinline__ void block_sum_i64(const long long (&acc)[NV], long long *red, long long *out)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const long long s = wave_sum_i64(acc[k]);
        if (lane == 0) red[w * NV + k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        long long s = 0;
        for (int ww = 0; ww < kBlock / 64; ++ww) s += red[ww * NV + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

template <int NV>
__device__ __forceinline__ void block_sum_fix(const double (&acc)[NV], long long *red, long long *out)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const long long s = wave_sum_i64(__double2ll_rn(acc[k] * kThermoFix));
        if (lane == 0) red[w * NV + k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        long long s = 0;
        for (int ww = 0; ww < kBlock / 64; ++ww) s += red[ww * NV + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// binning: keys, scan, scatter, in-cell sort, physical re-sort
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_bin_keys(const RepDev *__restrict__ reps,
                                                     const WorkItem *__restrict__ items, int nitems)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int s = it.local * kBlock + threadIdx.x;
    if (s >= R.N || halted(R)) return;
    const double *X = R.x[R.state->cur];
    const int N = R.N;
    const int key = cell_of(R, X[s], X[N + s], X[2 * N + s]);
    R.key[s] = key;
    R.rank[s] = atomicAdd(&R.cell_count[key], 1);
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const RepDev *__restrict__ reps,
                                                        const WorkItem *__restrict__ items,
                                                        int *__restrict__ chunk_sums)
{
    __shared__ int wbuf[kBlock / 64];
    const WorkItem it = items[blockIdx.x];
    const RepDev &R = reps[it.rep];
    const int base = it.local * 4 * kBlock + 4 * threadIdx.x;
    int loc = 0;
    if (base + 3 < R.ncell) {
        // four counts as one 16-byte load (the array is hipMalloc-aligned and base is a multiple of 4)
        const int4 v = *(const int4 *)gmem(R.cell_count + base);
        loc = v.x + v.y + v.z + v.w;
    } else {
        for (int c = base; c < R.ncell; ++c) loc += R.cell_count[c];
    }
    int total;
    block_excl_scan(loc, wbuf, total);
    if (threadIdx.x == 0) chunk_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_top(const int *__restrict__ rep_chunk_off,
                                                     int *__restrict__ chunk_sums)
{
    __shared__ int wbuf[kBlock / 64];
    const int c0 = rep_chunk_off[blockIdx.x], c1 = rep_chunk_off[blockIdx.x + 1];
    int carry = 0;
    for (int base = c0; base < c1; base += kBlock) {
        const int c = base + threadIdx.x;
        const int v = (c < c1) ? chunk_sums[c] : 0;
        int total;
        const int ex = block_excl_scan(v, wbuf, total);
        if (c < c1) chunk_sums[c] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_down(const RepDev *__restrict__ reps,
                                                      const WorkItem *__restrict__ items,
                                                      const int *__restrict__ chunk_sums)
{
    __shared__ int wbuf[kBlock / 64];
    const WorkItem it = items[blockIdx.x];
    const RepDev &R = reps[it.rep];
    if (halted(R)) return;  // the cell lists stay those of the held state
    const int base = it.local * 4 * kBlock + 4 * threadIdx.x;
    // four cells as one 16-byte access where all four are in range (both
    // arrays are hipMalloc-aligned and base is a multiple of 4)
    const bool whole = base + 3 < R.ncell;
    int v[4], loc = 0;
    if (whole) {
        const int4 q = *(const int4 *)gmem(R.cell_count + base);
        v[0] = q.x;
        v[1] = q.y;
        v[2] = q.z;
        v[3] = q.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (base + k < R.ncell) ? R.cell_count[base + k] : 0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) loc += v[k];
    int total;
    int run = chunk_sums[blockIdx.x] + block_excl_scan(loc, wbuf, total);
    if (whole) {
        int st[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            st[k] = run;
            run += v[k];
        }
        *(int4 *)gmemw(R.cell_start + base) = make_int4(st[0], st[1], st[2], st[3]);
        *(int4 *)gmemw(R.cell_count + base) = make_int4(0, 0, 0, 0);
        if (base + 3 == R.ncell - 1) R.cell_start[R.ncell] = run;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = base + k;
            if (c < R.ncell) {
                R.cell_start[c] = run;
                R.cell_count[c] = 0;
            }
            run += v[k];
            if (c == R.ncell - 1) R.cell_start[R.ncell] = run;
        }
    }
}

// Four bead chunks (work items) per workgroup, their loads issued together:
// each chunk is a short chain of dependent accesses (key and rank, then the
// cell start, then the store), which one chunk per thread leaves exposed.
constexpr int kScatterItems = 4;

__global__ __launch_bounds__(kBlock) void k_scatter(const RepDev *__restrict__ reps,
                                                    const WorkItem *__restrict__ items, int nitems)
{
    const int g0 = xcd_item((nitems + kScatterItems - 1) / kScatterItems) * kScatterItems;
    int key[kScatterItems], rank[kScatterItems], s[kScatterItems];
#pragma unroll
    for (int u = 0; u < kScatterItems; ++u) {
        key[u] = -1;
        rank[u] = 0;
        s[u] = -1;
        if (g0 + u >= nitems) continue;
        const WorkItem it = items[g0 + u];
        const RepDev &R = reps[it.rep];
        const int sl = it.local * kBlock + threadIdx.x;
        if (sl >= R.N) continue;
        s[u] = sl;
        key[u] = gmem(R.key)[sl];
        rank[u] = gmem(R.rank)[sl];
    }
#pragma unroll
    for (int u = 0; u < kScatterItems; ++u) {
        if (g0 + u >= nitems) continue;
        const RepDev &R = reps[items[g0 + u].rep];
        if (s[u] < 0 || halted(R)) continue;
        const int pos = ((unsigned)key[u] < (unsigned)R.ncell) ? gmem(R.cell_start)[key[u]] + rank[u] : -1;
        if ((unsigned)pos < (unsigned)R.N) gmemw(R.perm)[pos] = s[u];
        else atomicOr(&R.state->flags, kFlagScatter);
    }
}

// sort each cell's beads by id.  Used by the initial binning only, so a freshly
// uploaded system starts from one fixed slot layout; the step path needs no
// in-cell order (its sums are fixed point and pair contributions symmetric)
__global__ __launch_bounds__(kBlock) void k_cellsort(const RepDev *__restrict__ reps,
                                                     const WorkItem *__restrict__ items, int nitems)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int c = it.local * kBlock + threadIdx.x;
    if (c >= R.ncell) return;
    const int a = R.cell_start[c], b = R.cell_start[c + 1];
    int *p = R.perm;
    const int *id = R.id;
    for (int i = a + 1; i < b; ++i) {
        const int v = p[i], kv = id[v];
        int j = i - 1;
        while (j >= a && id[p[j]] > kv) {
            p[j + 1] = p[j];
            --j;
        }
        p[j + 1] = v;
    }
}

// Physical re-sort into cell order, one pass: slot d takes the bead at sorted
// position d (perm[d]).  Positions and velocities go to the other state
// buffer, the id, type and image words to the spare arrays, and perm becomes
// the identity; k_resort_swap then flips the buffer parity and swaps the spare
// arrays in.  One pass over the bead arrays instead of a gather into scratch
// and a copy back.
__global__ __launch_bounds__(kBlock) void k_resort(const RepDev *__restrict__ reps,
                                                   const WorkItem *__restrict__ items, int nitems)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int d = it.local * kBlock + threadIdx.x;
    if (d >= R.N) return;
    const int N = R.N;
    const int cur = __builtin_amdgcn_readfirstlane(R.state->cur);
    const int s = gmem(R.perm)[d];
    const auto *Xi = gmem(R.x[cur]), *Vi = gmem(R.v[cur]);
    auto *Xo = gmemw(R.x[1 - cur]), *Vo = gmemw(R.v[1 - cur]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        Xo[k * N + d] = Xi[k * N + s];
        Vo[k * N + d] = Vi[k * N + s];
    }
    const int id = gmem(R.id)[s];
    gmemw(R.id_alt)[d] = id;
    gmemw(R.type_alt)[d] = gmem(R.type)[s];
    gmemw(R.img_alt)[d] = gmem(R.img)[s];
    gmemw(R.perm)[d] = d;
    if ((unsigned)(id - 1) < (unsigned)N) R.slot_of_id[id - 1] = d;
}

// one thread per listed replica: the re-sorted buffers become the current ones
__global__ void k_resort_swap(RepDev *reps, const int *__restrict__ rep_list)
{
    RepDev &R = reps[rep_list[blockIdx.x]];
    R.state->cur ^= 1;
    int *id = R.id;
    R.id = R.id_alt;
    R.id_alt = id;
    unsigned char *ty = R.type;
    R.type = R.type_alt;
    R.type_alt = ty;
    long long *im = R.img;
    R.img = R.img_alt;
    R.img_alt = im;
}

// ---------------------------------------------------------------------------
// brick kernels: LDS staging, per-wave pair queue, fixed-point accumulation
// ---------------------------------------------------------------------------
//
// One 256-thread workgroup per brick of cells.  The brick and a one-cell halo
// are staged into LDS (positions relative to the brick corner, periodic images
// applied while staging).  Each pair of beads is evaluated once inside a brick:
// an owned bead scans only the staged beads after it in LDS order plus every
// halo bead, so a pair of owned beads is found by its lower bead alone and a
// pair with a halo bead once per brick (the neighbouring brick evaluates its
// own side).  Candidate tests run as wave-uniform loops that append in-range
// pairs to a per-wave queue (ballot + mbcnt); full queues are evaluated with
// every lane busy.  Per-bead sums are 64-bit fixed point in LDS: integer
// additions are exact and commute, so the result does not depend on the order
// in which waves and lanes deliver their pairs, nor on the batch or the slot
// layout.  Per-pair arithmetic follows force/pair_dpd.cu:296-399 and
// force/pair_mdpd.cu:219-300, 403-517.

struct BrickParams {
    int rmax, smax, omax, nrun, ntypes, qw, perbead, style;
    // phase clocks (kernel timing on): per work item and wave, the summed
    // device-clock ticks of staging, pair work, the closing barrier and the
    // write-back; null when off
    unsigned *phase;
    // >= 0: every item of the launch is brick `item` of this replica, in
    // order, so the item need not be read (one dependent load less before the
    // first cell-start loads); -1: read items[]
    int single_rep;
};

__device__ __forceinline__ WorkItem brick_item(const WorkItem *items, int item, const BrickParams &bp)
{
    if (bp.single_rep >= 0) return WorkItem{bp.single_rep, item};
    return items[item];
}

// lane 0 of each wave adds its phase durations (t[0..4] device clock) to the
// work item's four counters; each counter has one writer, so plain adds
__device__ __forceinline__ void add_phases(const BrickParams &bp, int item, const unsigned long long (&t)[5])
{
    if ((threadIdx.x & 63) != 0) return;
    unsigned *c = bp.phase + ((size_t)item * (kBlock / 64) + (threadIdx.x >> 6)) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] += (unsigned)(t[k + 1] - t[k]);
}

template <typename real> struct Vec4T;
template <> struct Vec4T<float> { using type = float4; };
template <> struct Vec4T<double> { using type = double4; };

// Staged bead word (pos.w, as integer bits): id | ((type - 1) << 27) | (halo << 31).
// Ids are 1..N with N < 2^27 and types 1..12, so each fits its field.  The
// type is kept zero-based, as it indexes the pair coefficient table.
constexpr int kIdBits = 27;
constexpr int kIdMask = (1 << kIdBits) - 1;

__host__ __device__ __forceinline__ int stage_word(int id, int type, bool halo)
{
    return id | ((type - 1) << kIdBits) | (halo ? (int)0x80000000u : 0);
}
__device__ __forceinline__ int word_id(int w) { return w & kIdMask; }
__device__ __forceinline__ unsigned word_type0(int w) { return ((unsigned)w >> kIdBits) & 15u; }  // type - 1
__device__ __forceinline__ bool word_halo(int w) { return w < 0; }

// Byte offsets of LDS elements from 24-bit multiplies (v_mad_u32_u24): LDS
// indices are small and non-negative, which the compiler cannot see from an
// int, so a plain index costs a full 32-bit multiply.
template <typename T> __device__ __forceinline__ T *lds_elem(T *base, unsigned idx, unsigned stride_bytes)
{
    return (T *)((char *)base + __umul24(idx, stride_bytes));
}

// a * b + c on the low 24 bits of a and b (b uniform), one full-rate VALU
// instruction.  The compiler cannot select it when b comes from another basic
// block, since the operand's range is not visible there; it then emits a 64-bit
// multiply-add.
__device__ __forceinline__ unsigned mad_u24(unsigned a, unsigned b, unsigned c)
{
    unsigned r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

// Owned bead info: staged index k (low 16 bits) and the base of its cell's run
// list (high 16 bits).
__device__ __forceinline__ int owned_info(int k, int rbase) { return k | (rbase << 16); }

template <typename real> struct BrickLds {
    using real4 = typename Vec4T<real>::type;
    PairCoef<real> *coef;
    double *red;               // block sums (in the queue's space, after the pair phase)
    real4 *pos;                // staged beads: xyz relative to the brick, w = stage word
    int *sinfo;                // staged beads: owner index (owned) or slot (halo)
    real *rh;                  // staged density (mDPD force pass)
    int *c_start, *sigcnt, *wbuf, *oinfo;
    unsigned *runs;            // per owned cell: own-column end, candidate runs (js | je << 16), 0
    int *cinfo, *gstart, *cellpre, *rbase; // staging only (in the sums' space)
    real4 *vown;               // owned beads: vxyz, w = slot bits
    long long *facc;           // fixed-point sums per owned bead
    unsigned *queue;           // per wave: li | lj << 16
    unsigned short *cellof;    // staging only: region cell of each staged index (in the queues' space)
};

__host__ __device__ inline size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

// LDS layout as byte offsets; pointers are rebuilt inline from the __shared__
// base so they stay in the LDS address space.  Halo beads keep only a position
// and a word (20 bytes in FP32 mode); velocities are staged for owned beads
// alone, and a pair with a halo bead reads the halo velocity from HBM.  This
// keeps a 4x4x5-cell brick at rho = 3 under 40 KB, four workgroups per CU.
struct LdsOffsets {
    unsigned coef, red, pos, sinfo, rh, c_start, cinfo, gstart, cellpre, rbase, sigcnt, wbuf, oinfo, runs, vown,
        facc, queue, cellof, total;
};

template <typename real>
__host__ __device__ __forceinline__ LdsOffsets lds_offsets(const BrickParams &p)
{
    LdsOffsets o;
    size_t off = 0;
#define GDPD_TAKE(field, bytes)   \
    o.field = (unsigned)off;      \
    off = align16(off + (size_t)(bytes));
    GDPD_TAKE(coef, sizeof(PairCoef<real>) * p.ntypes * p.ntypes)
    GDPD_TAKE(pos, sizeof(real) * 4 * (p.smax + 1))
    GDPD_TAKE(sinfo, sizeof(int) * p.smax)
    GDPD_TAKE(rh, sizeof(real) * (p.style == 2 ? p.smax : 1))
    GDPD_TAKE(c_start, sizeof(int) * (p.rmax + 1))
    GDPD_TAKE(sigcnt, sizeof(int) * (kSigClasses + 1))
    GDPD_TAKE(wbuf, sizeof(int) * 2 * (kBlock / 64))
    GDPD_TAKE(oinfo, sizeof(int) * p.omax)
    GDPD_TAKE(runs, sizeof(unsigned) * p.nrun)
    GDPD_TAKE(vown, sizeof(real) * 4 * p.omax)
    // cinfo, gstart, cellpre and rbase live only while the brick is staged,
    // before the per-bead sums are cleared, so they share the sums' space
    o.cinfo = (unsigned)off;
    o.gstart = (unsigned)(off + sizeof(int) * p.rmax);
    o.cellpre = (unsigned)(off + 2 * sizeof(int) * p.rmax);
    o.rbase = (unsigned)(off + 3 * sizeof(int) * p.rmax);
    GDPD_TAKE(facc, (3 * 8 * (size_t)p.omax > 4 * 4 * (size_t)p.rmax) ? 3 * 8 * (size_t)p.omax : 4 * 4 * (size_t)p.rmax)
    // the pair queues; the region cell of each staged index uses the same
    // space while the brick is staged, the block sums after the pair phase
    o.red = (unsigned)off;
    o.cellof = (unsigned)off;
    {
        size_t qb = sizeof(unsigned) * (kBlock / 64) * p.qw;
        if (qb < sizeof(double) * 64) qb = sizeof(double) * 64;
        if (qb < sizeof(unsigned short) * (size_t)(p.smax + 1)) qb = sizeof(unsigned short) * (size_t)(p.smax + 1);
        GDPD_TAKE(queue, qb)
    }
#undef GDPD_TAKE
    o.total = (unsigned)off;
    return o;
}

template <typename real>
__device__ __forceinline__ BrickLds<real> lds_bind(char *smem, const LdsOffsets &o)
{
    using real4 = typename Vec4T<real>::type;
    BrickLds<real> L;
    L.coef = (PairCoef<real> *)(smem + o.coef);
    L.red = (double *)(smem + o.red);
    L.pos = (real4 *)(smem + o.pos);
    L.sinfo = (int *)(smem + o.sinfo);
    L.rh = (real *)(smem + o.rh);
    L.c_start = (int *)(smem + o.c_start);
    L.cinfo = (int *)(smem + o.cinfo);
    L.gstart = (int *)(smem + o.gstart);
    L.cellpre = (int *)(smem + o.cellpre);
    L.rbase = (int *)(smem + o.rbase);
    L.runs = (unsigned *)(smem + o.runs);
    L.sigcnt = (int *)(smem + o.sigcnt);
    L.wbuf = (int *)(smem + o.wbuf);
    L.oinfo = (int *)(smem + o.oinfo);
    L.vown = (real4 *)(smem + o.vown);
    L.facc = (long long *)(smem + o.facc);
    L.queue = (unsigned *)(smem + o.queue);
    L.cellof = (unsigned short *)(smem + o.cellof);
    return L;
}

struct BrickGeom {
    int c0[3], ext[3], RX, RY, RZ, nreg;
    bool wraps;  // some region cell lies across a periodic face (image shift nonzero)
    double org[3];
};

__device__ __forceinline__ BrickGeom brick_geom(const RepDev &R, int b)
{
    BrickGeom g;
    const int bz = b % R.nbk[2], by = (b / R.nbk[2]) % R.nbk[1], bx = b / (R.nbk[2] * R.nbk[1]);
    const int bb[3] = {bx, by, bz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        g.c0[d] = bb[d] * R.bdim[d];
        g.ext[d] = min(R.bdim[d], R.nc[d] - g.c0[d]);
        g.org[d] = R.lo[d] + g.c0[d] * R.cs[d];
    }
    g.RX = g.ext[0] + 2;
    g.RY = g.ext[1] + 2;
    g.RZ = g.ext[2] + 2;
    g.nreg = g.RX * g.RY * g.RZ;
    g.wraps = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) g.wraps = g.wraps || g.c0[d] == 0 || g.c0[d] + g.ext[d] >= R.nc[d];
    return g;
}

// n / d for 0 <= n < 2^16 and 1 <= d <= 2^10 from a float reciprocal of d
// within 2 ulp: (n + 0.5) / d lies at least 1/(2d) from an integer, far more
// than the float error of the product, so truncation gives the exact quotient
// (checked for every such n and d with the reciprocal 2 ulp either side)
__device__ __forceinline__ int small_div(int n, float dinv) { return (int)(((float)n + 0.5f) * dinv); }

// 1/d for small_div: the hardware reciprocal (1 ulp), one instruction
// instead of the expansion of a correctly rounded division
__device__ __forceinline__ float small_rcp(int d) { return __builtin_amdgcn_rcpf((float)d); }

// region cell rc -> (rx, ry, rz) with rc = (rx * RY + ry) * RZ + rz
__device__ __forceinline__ void split3(int rc, int RY, int RZ, float inv_rz, float inv_ryz, int &rx, int &ry, int &rz)
{
    rx = small_div(rc, inv_ryz);
    const int rem = rc - rx * RY * RZ;
    ry = small_div(rem, inv_rz);
    rz = rem - ry * RZ;
}

// Periodic wrap of a cell coordinate in [-n, 2n): returns the cell and sets
// code to the image shift + 1 (0: one box below, 1: none, 2: one box above).
__device__ __forceinline__ int wrap_code(int c, int n, int &code)
{
    if (c < 0) {
        code = 0;
        return c + n;
    }
    if (c >= n) {
        code = 2;
        return c - n;
    }
    code = 1;
    return c;
}

__device__ __forceinline__ float id_to_real(int id, float) { return __int_as_float(id); }
__device__ __forceinline__ double id_to_real(int id, double) { return __longlong_as_double((long long)(unsigned)id); }
__device__ __forceinline__ int real_to_id(float v) { return __float_as_int(v); }
__device__ __forceinline__ int real_to_id(double v) { return (int)__double_as_longlong(v); }


// The run list of owned cell (rx, ry, rz) at runs[0..]: the end of its own
// column's candidate run, then every non-empty run of staged beads that a bead
// of the cell owns the pair with, as js | je << 16, then 0.  A run is a z-run
// of cells, contiguous in LDS, holding only such partners: later in LDS order,
// or halo.  Per neighbour column (cx, cy) = (rx + r/3 - 1, ry + r%3 - 1):
//   r 0..3  the columns that precede the cell's column: the whole z-run when
//           the column is halo in x or y, else only its bottom halo cell
//           (rz == 1) or its top one (rz == ext_z);
//   r 4     the own column: its bottom halo cell (rz == 1 only); the part
//           after the bead itself starts each bead's walk (the first entry);
//   r 5..8  the four following columns: the whole z-run.
// A brick one cell thick in z has both z-ends halo: an owned preceding column
// then gives its bottom and its top halo cell, as two runs.  A pair of owned
// beads in two columns is found from the bead whose column precedes, through
// its following columns.  Returns the cell's candidate count (the own column
// counted from the cell's first bead).  The entries written never exceed
// cell_run_cap(ext, rx, ry, rz).
__device__ __forceinline__ int cell_runs(const BrickGeom &g, const int *c_start, unsigned *runs, int rx, int ry,
                                         int rz)
{
    // Every cell start the list needs is read before the first entry is
    // written, so the 36 reads go out together instead of one LDS round trip
    // per run.  cs[r][k] = c_start[cb + k] for neighbour column r, with cb its
    // cell below the owned cell's layer; the own column is r = 4.
    int cs[9][4];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int cx = rx + r / 3 - 1, cy = ry + r % 3 - 1;
        const int cb = (cx * g.RY + cy) * g.RZ + rz - 1;
#pragma unroll
        for (int k = 0; k < 4; ++k) cs[r][k] = c_start[cb + k];
    }
    const int dend = cs[4][3];
    runs[0] = (unsigned)dend;
    int n = 1, total = dend - cs[4][1];
    const bool zlo = rz == 1, zhi = rz == g.ext[2], thin = g.ext[2] == 1;
    auto put = [&](int js, int je) {
        if (je > js) {
            runs[n++] = (unsigned)js | ((unsigned)je << 16);
            total += je - js;
        }
    };
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int cx = rx + r / 3 - 1, cy = ry + r % 3 - 1;
        const int A = cs[r][0], D = cs[r][3];
        if (r < 4) {
            const bool xyhalo = cx == 0 || cy == 0 || cy == g.RY - 1;
            if (xyhalo) {
                put(A, D);
            } else {
                if (zlo) put(A, cs[r][1]);
                if (zhi && (thin || !zlo)) put(cs[r][2], D);
            }
        } else if (r == 4) {
            if (zlo) put(A, cs[r][1]);
        } else {
            put(A, D);
        }
    }
    runs[n] = 0u;
    return total;
}

// exclusive scans of two values over the 256 threads (one pair per thread);
// every thread must call it
__device__ __forceinline__ void block_excl_scan2(int v0, int v1, int *wbuf, int &e0, int &e1, int &t0, int &t1)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = wave_incl_scan(v0), i1 = wave_incl_scan(v1);
    if (lane == 63) {
        wbuf[w] = i0;
        wbuf[kBlock / 64 + w] = i1;
    }
    __syncthreads();
    int o0 = 0, o1 = 0, s0 = 0, s1 = 0;
#pragma unroll
    for (int k = 0; k < kBlock / 64; ++k) {
        const int a = wbuf[k], b = wbuf[kBlock / 64 + k];
        o0 += (k < w) ? a : 0;
        o1 += (k < w) ? b : 0;
        s0 += a;
        s1 += b;
    }
    __syncthreads();
    e0 = o0 + i0 - v0;
    e1 = o1 + i1 - v1;
    t0 = s0;
    t1 = s1;
}

// Owned cells are ordered by candidate count, largest first, in classes of
// four candidates: the owned beads are numbered cell by cell in that order,
// so the beads of one wavefront walk runs of about the same total length.
__device__ __forceinline__ int cand_class(int total)
{
    return (kSigClasses - 1) - min(kSigClasses - 1, total >> 2);
}

// Barrier for LDS traffic only: it waits for this wave's LDS operations, not
// for its global loads, so loads issued before it stay in flight across it.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Wave issue priorities of the brick kernels, highest first: staging and the
// force write-back (chains of dependent loads or stores with little arithmetic
// in between), the candidate walk (dependent LDS reads), then the pair
// evaluation (the Saru hash: long runs of VALU work).  A wave of a latency-bound
// phase then issues as soon as its operands are ready, and the waves of the
// other workgroups on the CU fill the VALU in the gaps.  Measured at 10M
// beads: staging 13.3 -> 9.2 us per wave, the force pass 1.373 -> 1.305 ms.
__device__ __forceinline__ void staging_priority() { __builtin_amdgcn_s_setprio(3); }
__device__ __forceinline__ void walk_priority() { __builtin_amdgcn_s_setprio(1); }
__device__ __forceinline__ void normal_priority() { __builtin_amdgcn_s_setprio(0); }

// Stage the brick + halo.  Returns (uniformly) the number of owned beads, or
// -1 when the staged or owned beads exceed the LDS capacity.
//
// Three global round trips are on the critical path: the region cells' starts,
// the permutation, then the bead state.  The pair coefficients are loaded
// beside the first; the permutation loads are issued before the owned-bead
// numbering (LDS atomics and scans behind LDS-only barriers) and the state
// loads in the middle of it, so each overlaps LDS work.
template <typename real>
__device__ __forceinline__ int stage_brick(const RepDev &R, const BrickGeom &g, const BrickLds<real> &L,
                                           const BrickParams &bp, int cur, bool need_vel, bool need_rho)
{
    using real4 = typename Vec4T<real>::type;
    const int t = threadIdx.x;
    // pair coefficients, as 32-bit words, held in registers until the LDS write
    constexpr int kCoefWords = (kMaxTypes * kMaxTypes * (int)sizeof(PairCoef<real>) / 4 + kBlock - 1) / kBlock;
    const int ncw = R.ntypes * R.ntypes * (int)(sizeof(PairCoef<real>) / 4);
    int cw[kCoefWords];
    {
        const auto *cg = gmem((const int *)R.coef);
#pragma unroll
        for (int u = 0; u < kCoefWords; ++u) {
            const int k = t + u * kBlock;
            cw[u] = k < ncw ? cg[k] : 0;
        }
    }
    const float inv_rz = small_rcp(g.RZ), inv_ryz = small_rcp(g.RY * g.RZ);
    const float inv_ez = small_rcp(g.ext[2]), inv_eyz = small_rcp(g.ext[1] * g.ext[2]);
    for (int rc = t; rc < g.nreg; rc += kBlock) {
        int rx, ry, rz;
        split3(rc, g.RY, g.RZ, inv_rz, inv_ryz, rx, ry, rz);
        int wx, wy, wz;
        const int gx = wrap_code(g.c0[0] - 1 + rx, R.nc[0], wx);
        const int gy = wrap_code(g.c0[1] - 1 + ry, R.nc[1], wy);
        const int gz = wrap_code(g.c0[2] - 1 + rz, R.nc[2], wz);
        const int gc = (gx * R.nc[1] + gy) * R.nc[2] + gz;
        const auto *cs = gmem(R.cell_start);
        const int a = cs[gc], b = cs[gc + 1];
        L.cinfo[rc] = rx | (ry << 8) | (rz << 16) | (wx << 24) | (wy << 26) | (wz << 28);
        L.gstart[rc] = a;
        L.c_start[rc + 1] = b - a;
    }
#pragma unroll
    for (int u = 0; u < kCoefWords; ++u) {
        const int k = t + u * kBlock;
        if (k < ncw) ((int *)L.coef)[k] = cw[u];
    }
    __syncthreads();
    const int nown = g.ext[0] * g.ext[1] * g.ext[2];
    {
        // region-cell starts (cells 4t..4t+3 per thread, in order), and the
        // run-list base of every owned cell (cells t, t + 256, ... per thread:
        // each cell needs its own range after the shared empty entry runs[0],
        // in any order), in one pair of scans
        int v[4], c[4], loc = 0, cloc = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int rc = 4 * t + k;
            v[k] = (rc < g.nreg) ? L.c_start[rc + 1] : 0;
            loc += v[k];
            const int oc = t + k * kBlock;
            c[k] = 0;
            if (oc < nown) {
                int rx, ry, rz;
                split3(oc, g.ext[1], g.ext[2], inv_ez, inv_eyz, rx, ry, rz);
                c[k] = cell_run_cap(g.ext[0], g.ext[1], g.ext[2], rx + 1, ry + 1, rz + 1);
            }
            cloc += c[k];
        }
        int run, crun, total, ctotal;
        block_excl_scan2(loc, cloc, L.wbuf, run, crun, total, ctotal);
        crun += 1;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int rc = 4 * t + k;
            if (rc < g.nreg) L.c_start[rc] = run;
            run += v[k];
            const int oc = t + k * kBlock;
            if (oc < nown) L.rbase[oc] = crun;
            crun += c[k];
        }
        if (t == 0) {
            L.c_start[g.nreg] = total;
            L.runs[0] = 0u;
        }
        if (t <= kSigClasses) L.sigcnt[t] = 0;
    }
    __syncthreads();
    const int S = L.c_start[g.nreg];
    const int Sc = min(S, bp.smax);  // indices past the capacity are not staged
    // region cell of each staged index
    for (int rc = t; rc < g.nreg; rc += kBlock) {
        const int e = min(L.c_start[rc + 1], bp.smax);
        for (int i = L.c_start[rc]; i < e; ++i) L.cellof[i] = (unsigned short)rc;
    }
    lds_barrier();

    const int N = R.N;
    const auto *X = gmem(R.x[cur]), *V = gmem(R.v[cur]);
    const auto *perm = gmem(R.perm);
    const auto *gid = gmem(R.id);
    const auto *gtype = gmem(R.type);
    const auto *grho = gmem(R.rho);
    // Four staged beads per thread, 256 apart so each load instruction stays
    // coalesced.  lookup() finds their region cells and issues the permutation
    // loads; fetch() issues the state loads.  Velocities are staged for owned
    // beads only.
    int rcs[4], cis[4], sl[4];
    bool own[4];
    double xs[4], ys[4], zs[4], vxs[4], vys[4], vzs[4];
    int ids[4], tys[4];
    // A slot u whose staged indices all lie past the staged count for the
    // whole wave is skipped (a wave-uniform branch), so no wasted permutation
    // and state loads are issued for it: at rho = 3 a 4x4x5 brick stages about
    // 760 beads, and no wave has a bead in its fourth slot.
    const int lane = t & 63;
    auto wave_has = [&](int base, int u) {
        return __builtin_amdgcn_readfirstlane(base - lane) + u * kBlock < Sc;
    };
    // Stage by stage over the four slots, so each stage's LDS reads go out
    // together: three LDS round trips before the permutation loads, not twelve.
    auto lookup = [&](int base) {
        int lo[4], gs[4], cs0[4];
        bool in[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = base + u * kBlock;
            in[u] = k < Sc;
            lo[u] = in[u] ? (int)L.cellof[k] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            gs[u] = L.gstart[lo[u]];
            cs0[u] = L.c_start[lo[u]];
            cis[u] = L.cinfo[lo[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = base + u * kBlock;
            rcs[u] = lo[u];
            const int ci = cis[u];
            const int rx = ci & 255, ry = (ci >> 8) & 255, rz = (ci >> 16) & 255;
            own[u] = in[u] && rx >= 1 && rx <= g.ext[0] && ry >= 1 && ry <= g.ext[1] && rz >= 1 && rz <= g.ext[2];
            sl[u] = 0;
            if (!wave_has(base, u)) continue;
            if (in[u]) sl[u] = perm[gs[u] + (k - cs0[u])];
        }
    };
    auto fetch = [&](int base) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            xs[u] = ys[u] = zs[u] = 0.0;
            tys[u] = ids[u] = 0;
            vxs[u] = vys[u] = vzs[u] = 0.0;
            if (!wave_has(base, u)) continue;
            sl[u] = ((unsigned)sl[u] < (unsigned)N) ? sl[u] : 0;
            const int sv = sl[u];
            xs[u] = X[sv];
            ys[u] = X[N + sv];
            zs[u] = X[2 * N + sv];
            tys[u] = gtype[sv];
            ids[u] = need_vel ? gid[sv] : 0;
            if (need_vel && own[u]) {
                vxs[u] = V[sv];
                vys[u] = V[N + sv];
                vzs[u] = V[2 * N + sv];
            }
        }
    };
    lookup(t);

    // each owned cell's run list; owned-cell offsets from bead counts per
    // candidate class (LDS atomics; the order of cells inside a class is
    // immaterial), kept in the low half of cellpre with the class above it,
    // then a scan over the classes
    for (int oc = t; oc < nown; oc += kBlock) {
        int rx, ry, rz;
        split3(oc, g.ext[1], g.ext[2], inv_ez, inv_eyz, rx, ry, rz);
        ++rx, ++ry, ++rz;
        const int rc = (rx * g.RY + ry) * g.RZ + rz;
        const int cls = cand_class(cell_runs(g, L.c_start, L.runs + L.rbase[oc], rx, ry, rz));
        L.cellpre[oc] = atomicAdd(&L.sigcnt[cls], L.c_start[rc + 1] - L.c_start[rc]) | (cls << 16);
    }
    lds_barrier();
    if (t < 64) {
        const int v = t < kSigClasses ? L.sigcnt[t] : 0;
        const int inc = wave_incl_scan(v);
        if (t < kSigClasses) L.sigcnt[t] = inc - v;
        if (t == kSigClasses - 1) L.sigcnt[kSigClasses] = inc;
    }
    fetch(t);
    lds_barrier();
    const int O = L.sigcnt[kSigClasses];
    if (S > bp.smax || O > bp.omax) return -1;

    for (int base = t; base < S; base += 4 * kBlock) {
        if (base != t) {
            lookup(base);
            fetch(base);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = base + u * kBlock;
            if (k >= S) continue;
            const int rc2 = rcs[u];
            const int ci = cis[u];
            const int rx = ci & 255, ry = (ci >> 8) & 255, rz = (ci >> 16) & 255;
            // periodic image, then brick-relative: (x + shift) - org, with the
            // shift add left out where every shift is zero (x + 0 is x)
            double px = xs[u], py = ys[u], pz = zs[u];
            if (g.wraps) {
                px += (double)(((ci >> 24) & 3) - 1) * R.L[0];
                py += (double)(((ci >> 26) & 3) - 1) * R.L[1];
                pz += (double)(((ci >> 28) & 3) - 1) * R.L[2];
            }
            real4 p;
            p.x = (real)(px - g.org[0]);
            p.y = (real)(py - g.org[1]);
            p.z = (real)(pz - g.org[2]);
            p.w = id_to_real(stage_word(ids[u], tys[u], !own[u]), (real)0);
            L.pos[k] = p;
            if (own[u]) {
                const int oc = ((rx - 1) * g.ext[1] + (ry - 1)) * g.ext[2] + (rz - 1);
                // the cell's offset inside its class, plus the class's offset
                const int cp = L.cellpre[oc];
                const int o = (cp & 0xFFFF) + L.sigcnt[cp >> 16] + (k - L.c_start[rc2]);
                L.sinfo[k] = o;
                L.oinfo[o] = owned_info(k, L.rbase[oc]);
                real4 q;
                q.x = (real)vxs[u];
                q.y = (real)vys[u];
                q.z = (real)vzs[u];
                q.w = id_to_real(sl[u], (real)0);
                L.vown[o] = q;
            } else {
                L.sinfo[k] = sl[u];
            }
            if (need_rho) L.rh[k] = (real)grho[sl[u]];
        }
    }
    __syncthreads();  // cinfo, gstart and cellpre (aliased by facc) are dead from here
    for (int k = t; k < 3 * O; k += kBlock) L.facc[k] = 0;
    __syncthreads();
    return O;
}

template <typename real> __device__ __forceinline__ real ratio(real a, real b, real binv);
template <> __device__ __forceinline__ float ratio<float>(float a, float b, float binv) { return a * binv; }
template <> __device__ __forceinline__ double ratio<double>(double a, double b, double binv) { return a / b; }

// r and 1/r of a pair distance.  FP32 mode uses the hardware reciprocal
// square root (about 1 ulp, inside the FP32-mode force tolerance); FP64 mode
// the correctly rounded square root and division.
__device__ __forceinline__ void root_and_inverse(float r2, float &r, float &rinv)
{
    rinv = __builtin_amdgcn_rsqf(r2);
    r = r2 * rinv;
}

__device__ __forceinline__ void root_and_inverse(double r2, double &r, double &rinv)
{
    r = sqrt(r2);
    rinv = 1.0 / r;
}

// Fixed-point scale of the per-bead sums: 2^24 for FP32 pair arithmetic (a
// resolution far below the FP32 rounding of a pair force), 2^40 for FP64.
template <typename real> struct FixScale;
template <> struct FixScale<float> { static constexpr double s = 16777216.0, inv = 1.0 / 16777216.0; };
template <> struct FixScale<double> { static constexpr double s = 1099511627776.0, inv = 1.0 / 1099511627776.0; };

template <typename real> __device__ __forceinline__ long long to_fix(double v);

// v * 2^24 rounded to nearest (ties to even) as an integer, for |v| < 2^27:
// v * 2^24 + 1.5 * 2^52 lies in [2^52, 2^53), where the doubles are exactly
// the integers, so the single rounding of the fma is the rounding to an
// integer and the bits of the sum, less those of 1.5 * 2^52, are that integer.
// Three instructions (convert, fma, 64-bit add) instead of the eight of a
// general double-to-int64 conversion.
template <> __device__ __forceinline__ long long to_fix<float>(double v)
{
    const double y = fma(v, FixScale<float>::s, 6755399441055744.0);
    return __double_as_longlong(y) - 0x4338000000000000LL;
}

template <> __device__ __forceinline__ long long to_fix<double>(double v)
{
    return __double2ll_rn(v * FixScale<double>::s);
}

// The same integer from a value already multiplied by the scale in the working
// precision (a power-of-two scaling, so exact): a convert and one add in FP32
// mode.
template <typename real> __device__ __forceinline__ long long fix_scaled(real vs);
template <> __device__ __forceinline__ long long fix_scaled<float>(float vs)
{
    const double y = (double)vs + 6755399441055744.0;
    return __double_as_longlong(y) - 0x4338000000000000LL;
}
template <> __device__ __forceinline__ long long fix_scaled<double>(double vs) { return __double2ll_rn(vs); }

__device__ __forceinline__ void lds_add(long long *p, long long v)
{
    atomicAdd((unsigned long long *)p, (unsigned long long)v);
}

__device__ __forceinline__ void lds_sub(long long *p, long long v)
{
    __hip_atomic_fetch_sub((unsigned long long *)p, (unsigned long long)v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}


__device__ __forceinline__ int lanes_below(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Thermo partial sums per thread: 0 ke, 1 pair pe, 2 bond pe, 3..8 pressure tensor.
// Bead i of a queued pair is always owned; j is owned or halo.  A halo j's
// velocity is read from HBM at its slot and rounded to the working precision
// exactly as staging rounds an owned bead's.
template <typename real, int STYLE, bool DENSITY>
__device__ __forceinline__ void eval_queue(const RepDev &R, const BrickLds<real> &L, const unsigned *q, int qn,
                                           int nt, int cur, uint32_t seed_step, bool log, bool perbead,
                                           long long (&acc)[kThermoParts])
{
    using real4 = typename Vec4T<real>::type;
    const int lane = threadIdx.x & 63;
    const int N = R.N;
    // global address space: the loads count against vmcnt only, so waits for
    // LDS results do not wait for them
    const auto *V = gmem(R.v[cur]);
    for (int e = lane; e < qn; e += 64) {
        const unsigned ent = q[e];
        const int li = ent & 0xFFFF, lj = ent >> 16;
        const real4 pi = L.pos[li], pj = L.pos[lj];
        const int wi = real_to_id(pi.w), wj = real_to_id(pj.w);
        const int oi = L.sinfo[li], sj = L.sinfo[lj];  // sj: owner index of an owned j, slot of a halo j
        const bool jown = !word_halo(wj);
        const int oj = sj;
        const real dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const real r2 = dx * dx + dy * dy + dz * dz;
        // coefficient row (type_i - 1) * nt, column type_j - 1, as an LDS address
        using LdsCoef = __attribute__((address_space(3))) const PairCoef<real>;
        const unsigned cstride = (unsigned)sizeof(PairCoef<real>);
        const unsigned cbase = (unsigned)(size_t)(LdsCoef *)L.coef;
        LdsCoef &C = *(LdsCoef *)(size_t)mad_u24(word_type0(wi), (unsigned)nt * cstride,
                                                 __umul24(word_type0(wj), cstride) + cbase);
        if (DENSITY) {
            const real r = sqrt(r2);
            // force/pair_mdpd.cu:266-289
            real wd = (real)1 - ratio<real>(r, C.rd, C.rdinv);
            wd = wd > (real)0 ? wd : (real)0;
            const long long w = to_fix<real>((double)(C.rhofac * wd * wd));
            lds_add(&L.facc[oi], w);
            if (jown) lds_add(&L.facc[oj], w);
            continue;
        }
        // Every load of the pair is issued before the hash, which needs only
        // the two ids, and the scheduler is kept from moving them below it
        // (sched_barrier): its ~140 instructions then cover the latency of the
        // coefficient, velocity and HBM loads.  The cutoff test masks the sums
        // instead of branching (a lane whose pair fails it idles through the
        // round either way).  The HBM loads of a halo partner's velocity are
        // used after the hash; lanes with an owned partner load slot 0 (one
        // shared address) and ignore the value.
        const int sv = (!jown && (unsigned)sj < (unsigned)N) ? sj : 0;
        const double gvx = V[sv], gvy = V[N + sv], gvz = V[2 * N + sv];
        const real4 vi = L.vown[oi], vo = L.vown[jown ? oj : oi];
        const real cA = C.A, cgamma = C.gamma, csig = C.sig, crc2 = C.rc2;
        const real crc = C.rc, crcinv = C.rcinv;
        real cB = (real)0, crd = (real)0, crdinv = (real)0, rhoi = (real)0, rhoj = (real)0;
        if (STYLE == 2) {
            cB = C.B;
            crd = C.rd;
            crdinv = C.rdinv;
            rhoi = L.rh[li];
            rhoj = L.rh[lj];
        }
        __builtin_amdgcn_sched_barrier(0);
        const int idi = word_id(wi), idj = word_id(wj);
        const uint32_t t1 = (uint32_t)min(idi, idj), t2 = (uint32_t)max(idi, idj);
        const uint32_t hbits = saru_bits(t1, t2, seed_step);
        // force/pair_dpd.cu: pairs inside rc with r >= 1e-10
        const bool in_rc = !(r2 > crc2 || r2 < (real)1.0e-20);
        real r, rinv;
        root_and_inverse(r2, r, rinv);
        const real wc = (real)1 - ratio<real>(r, crc, crcinv);
        const real wr = wc;
        const real rnd = saru_uniform<real>(hbits);
        const real vjx = jown ? vo.x : (real)gvx, vjy = jown ? vo.y : (real)gvy, vjz = jown ? vo.z : (real)gvz;
        const real dot = dx * (vi.x - vjx) + dy * (vi.y - vjy) + dz * (vi.z - vjz);
        real wd = (real)0, fpair;
        if (STYLE == 2) {
            wd = (real)1 - ratio<real>(r, crd, crdinv);
            wd = wd > (real)0 ? wd : (real)0;
            fpair = cA * wc + cB * (rhoi + rhoj) * wd;
        } else {
            fpair = cA * wc;
        }
        fpair -= cgamma * wr * wr * dot * rinv;
        fpair += csig * wr * rnd;
        fpair *= rinv;
        // the force components times the fixed-point scale (exact), straight to
        // fixed point; a pair outside the cutoff adds exact zeros, so the sums
        // need no branch (a branch here would move the loads above behind it)
        const real fps = in_rc ? fpair * (real)FixScale<real>::s : (real)0;
        const long long qx = fix_scaled<real>(dx * fps), qy = fix_scaled<real>(dy * fps),
                        qz = fix_scaled<real>(dz * fps);
        long long *fi = lds_elem(L.facc, (unsigned)oi, 3u * sizeof(long long));
        lds_add(fi + 0, qx);
        lds_add(fi + 1, qy);
        lds_add(fi + 2, qz);
        if (jown) {
            long long *fj = lds_elem(L.facc, (unsigned)oj, 3u * sizeof(long long));
            lds_sub(fj + 0, qx);
            lds_sub(fj + 1, qy);
            lds_sub(fj + 2, qz);
        }
        if (log && in_rc) {
            // per-bead shares as in force/pair_dpd.cu:366-398 and pair_mdpd.cu:479-516
            real pe_i, pe_j;
            if (STYLE == 2) {
                const real u_c = (real)0.5 * C.A * C.rc * wc * wc;
                pe_i = (real)0.5 * u_c + (real)0.25 * C.B * C.rd * rhoi * wd * wd;
                pe_j = (real)0.5 * u_c + (real)0.25 * C.B * C.rd * rhoj * wd * wd;
            } else {
                const real ppe = (real)0.5 * C.A * C.rc * wc * wc;
                pe_i = pe_j = (real)0.5 * ppe;
            }
            const real ffx = dx * fpair, ffy = dy * fpair, ffz = dz * fpair;
            const real w6[6] = {(real)0.5 * dx * ffx, (real)0.5 * dx * ffy, (real)0.5 * dx * ffz,
                                (real)0.5 * dy * ffy, (real)0.5 * dy * ffz, (real)0.5 * dz * ffz};
            // per-pair fixed point: the brick sums do not depend on pair order
            acc[1] += thermo_fix((double)pe_i) + (jown ? thermo_fix((double)pe_j) : 0ll);
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const long long wq = thermo_fix((double)w6[c]);
                acc[3 + c] += jown ? 2 * wq : wq;
            }
            if (perbead) {
                // hook-path log step: per-bead fixed-point sums in the pe/virial arrays
                const int si = real_to_id(vi.w), sjs = jown ? real_to_id(L.vown[oj].w) : 0;
                atomicAdd((unsigned long long *)&R.pe[si], (unsigned long long)to_fix<real>((double)pe_i));
                if (jown) atomicAdd((unsigned long long *)&R.pe[sjs], (unsigned long long)to_fix<real>((double)pe_j));
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    const unsigned long long wq = (unsigned long long)to_fix<real>((double)w6[c]);
                    atomicAdd((unsigned long long *)&R.vir[(size_t)c * N + si], wq);
                    if (jown) atomicAdd((unsigned long long *)&R.vir[(size_t)c * N + sjs], wq);
                }
            }
        }
    }
}

// Generation + evaluation.  The owned beads are split into 4 * nr equal chunks
// of at most 64 (nr = ceil(O / 256) rounds), and wave w takes chunks w, w + 4,
// ...: equal chunks keep the four waves' pair phases about equally long, so
// few wait at the closing barrier.  Each lane walks its bead's candidates as
// one sequence: the rest of its own column after the bead, then the runs of
// its cell's run list (cell_runs), two candidates per step.  The walk does not
// stop at run ends, so the wave's step count follows the longest total walk of
// its lanes, not the longest run per column, and beads are numbered so that a
// wave's beads have about the same total.  A pair of owned beads is found once
// and a pair with a halo bead once per brick.  In-range candidates are
// appended to a per-wave queue (ballot + mbcnt) that is evaluated 64 pairs at a
// time.
template <typename real, int STYLE, bool DENSITY>
__device__ __forceinline__ void brick_pairs(const RepDev &R, const BrickLds<real> &L, const BrickParams &bp, int O,
                                            int nt, real cut2, int cur, uint32_t seed_step, bool log,
                                            long long (&acc)[kThermoParts], unsigned long long *t_done = nullptr)
{
    using real4 = typename Vec4T<real>::type;
    const int t = threadIdx.x, lane = t & 63;
    // the wave index is uniform; readfirstlane keeps it (and the queue base) in SGPRs
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int qw = bp.qw;
    unsigned *q = L.queue + w * qw;
    int qn = 0;
    const int nwave = kBlock / 64;
    const int nchunk = nwave * max(1, (O + kBlock - 1) / kBlock);
    const int per = (O + nchunk - 1) / nchunk;  // <= 64; 0 when the brick owns no bead
    // A step loads candidates j and j + 1 from pos[min(j, jmax)] on: a valid
    // candidate is below the staged count S <= smax, and pos holds smax + 1
    // entries, so both loads stay in bounds.
    const unsigned jmax = (unsigned)bp.smax - 1u;
    walk_priority();
    for (int chunk = w; chunk < nchunk; chunk += nwave) {
        const int o = chunk * per + lane;
        const bool valid = lane < per && o < O;
        // walker: [wj, we) the current run, nxt the next one (js | je << 16;
        // 0 ends the list), rr its index.  A lane without a bead starts on the
        // shared empty entry runs[0] and stays there.
        unsigned li = 0, wj = 0, we = 0, rr = 0, nxt = 0;
        real4 pi;
        pi.x = pi.y = pi.z = pi.w = (real)0;
        if (valid) {
            const int info = L.oinfo[o];
            li = (unsigned)info & 0xFFFFu;
            rr = (unsigned)info >> 16;
            pi = L.pos[li];
            wj = li + 1u;
            we = L.runs[rr];
            rr += 1u;
            nxt = L.runs[rr];
        }
        // Move to the next run once the current one is used up (runs are not
        // empty, so one move per step suffices), then read the entry after it.
        // Branch-free, with the read every step: the LDS reads of a step then
        // come in one fixed order, so the wait for this entry before the next
        // move leaves the newer position loads in flight.
        auto advance = [&]() {
            const bool sw = wj >= we;
            const unsigned nj = nxt & 0xFFFFu, ne = nxt >> 16;
            wj = sw ? nj : wj;
            we = sw ? ne : we;
            rr += (sw && ne != 0u) ? 1u : 0u;
            nxt = L.runs[rr];
        };
        // Candidates j, j + 1 against the queue: the in-range ones are
        // appended (lanes past their walk are dropped by the validity masks).
        // e1: the candidates' run end less one (signed: -1 for a finished
        // walk), so candidate j is valid when j <= e1 and j + 1 when j < e1;
        // the compares are made here, where the branches use them
        auto test_pair = [&](const real4 &pj0, const real4 &pj1, unsigned j, int e1) {
            const bool v0 = (int)j <= e1, v1 = (int)j < e1;
            if (qn > qw - 128) {
                // evaluate whole 64-pair rounds only; the remainder (< 64
                // entries) moves to the front of the queue
                const int ne = qn & ~63;
                normal_priority();
                eval_queue<real, STYLE, DENSITY>(R, L, q, ne, nt, cur, seed_step, log, bp.perbead, acc);
                walk_priority();
                const int rem = qn - ne;
                if (lane < rem) q[lane] = q[ne + lane];
                qn = rem;
            }
            const real dx0 = pi.x - pj0.x, dy0 = pi.y - pj0.y, dz0 = pi.z - pj0.z;
            const real dx1 = pi.x - pj1.x, dy1 = pi.y - pj1.y, dz1 = pi.z - pj1.z;
            const real r20 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0;
            const real r21 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
            const bool ina = DENSITY ? (r20 < cut2) : (r20 <= cut2);
            const bool inb = DENSITY ? (r21 < cut2) : (r21 <= cut2);
            // Ballots of plain compares combine in SALU (a ballot of a
            // combined bool is rematerialised per lane); the branches use the
            // same compares as exec masks.
            const unsigned long long ma = __builtin_amdgcn_ballot_w64(v0) & __builtin_amdgcn_ballot_w64(ina);
            const unsigned long long mb = __builtin_amdgcn_ballot_w64(v1) & __builtin_amdgcn_ballot_w64(inb);
            // append at the queue's uniform end, in lane order, the second
            // candidate after the first
            const unsigned ent = li | (j << 16);
            unsigned *qa = q + qn;
            unsigned *qb = qa + __popcll(ma);
            if (v0 && ina) qa[lanes_below(ma)] = ent;
            if (v1 && inb) qb[lanes_below(mb)] = ent + 0x10000u;
            qn += (int)(__popcll(ma) + __popcll(mb));
        };
        // Software-pipelined over two register sets, so no copies: the next
        // step's walker move and loads come before this step's test.
        advance();  // the own column may hold nothing after the bead
        unsigned ja = wj;
        int ea = (int)we - 1;
        real4 pa0, pa1, pb0, pb1;
        {
            const unsigned lc = min(wj, jmax);
            pa0 = L.pos[lc];
            pa1 = L.pos[lc + 1u];
        }
        wj += 2u;
        if (__builtin_amdgcn_ballot_w64((int)ja <= ea) == 0ull) continue;
        for (;;) {
            advance();
            const unsigned jb = wj;
            const int eb = (int)we - 1;
            {
                const unsigned lc = min(wj, jmax);
                pb0 = L.pos[lc];
                pb1 = L.pos[lc + 1u];
            }
            wj += 2u;
            test_pair(pa0, pa1, ja, ea);
            if (__builtin_amdgcn_ballot_w64((int)jb <= eb) == 0ull) break;
            advance();
            ja = wj;
            ea = (int)we - 1;
            {
                const unsigned lc = min(wj, jmax);
                pa0 = L.pos[lc];
                pa1 = L.pos[lc + 1u];
            }
            wj += 2u;
            test_pair(pb0, pb1, jb, eb);
            if (__builtin_amdgcn_ballot_w64((int)ja <= ea) == 0ull) break;
        }
    }
    normal_priority();
    eval_queue<real, STYLE, DENSITY>(R, L, q, qn, nt, cur, seed_step, log, bp.perbead, acc);
    if (t_done) *t_done = wall_clock64();
    __syncthreads();
}

// ---------------------------------------------------------------------------
// mDPD pass 1: local density (force/pair_mdpd.cu:219-300).  Every pair j != i
// inside rd contributes (15/(2 pi rd^3)) (1 - r/rd)^2 to both beads; there is
// no rc test and no self term.
// ---------------------------------------------------------------------------

template <typename real>
__global__ __launch_bounds__(kBlock) void k_density(const RepDev *__restrict__ reps,
                                                    const WorkItem *__restrict__ items, int nitems,
                                                    BrickParams bp)
{
    extern __shared__ __align__(16) char smem[];
    const BrickLds<real> L = lds_bind<real>(smem, lds_offsets<real>(bp));
    const WorkItem it = brick_item(items, xcd_item(nitems), bp);
    const RepDev &R = reps[it.rep];
    // the halt flag is tested after staging, which only reads, so the brick's
    // first loads do not wait for the state block
    const bool held = halted(R);
    // uniform: the state pointers R.x[cur], R.v[cur] then come from scalar loads
    const int cur = __builtin_amdgcn_readfirstlane(R.state->cur);
    const BrickGeom g = brick_geom(R, it.local);
    staging_priority();
    const int O = stage_brick<real>(R, g, L, bp, cur, false, false);
    normal_priority();
    if (held) return;
    if (O < 0) {
        if (threadIdx.x == 0) atomicOr(&R.state->flags, kFlagStageOverflow);
        return;
    }
    long long acc[kThermoParts];
#pragma unroll
    for (int k = 0; k < kThermoParts; ++k) acc[k] = 0;
    brick_pairs<real, 2, true>(R, L, bp, O, R.ntypes, (real)R.rdmax2, cur, 0u, false, acc);
    staging_priority();
    for (int o = threadIdx.x; o < O; o += kBlock)
        R.rho[real_to_id(L.vown[o].w)] = (double)L.facc[o] * FixScale<real>::inv;
}

// ---------------------------------------------------------------------------
// pair force of every owned bead of a brick, written to the force array, with
// the pair PE/virial of log steps (per brick, or per bead on the hook path)
// ---------------------------------------------------------------------------

template <typename real, int STYLE>
__global__ __launch_bounds__(kBlock) void k_brick(const RepDev *__restrict__ reps,
                                                  const WorkItem *__restrict__ items, int nitems,
                                                  BrickParams bp, int mode)
{
    extern __shared__ __align__(16) char smem[];
    const BrickLds<real> L = lds_bind<real>(smem, lds_offsets<real>(bp));
    const int item = xcd_item(nitems);
    const WorkItem it = brick_item(items, item, bp);
    const RepDev &R = reps[it.rep];
    unsigned long long tp[5] = {0, 0, 0, 0, 0};
    if (bp.phase) tp[0] = wall_clock64();
    // the halt flag is tested after staging, which only reads, so the brick's
    // first loads do not wait for the state block
    const bool held = halted(R);
    // uniform: the state pointers R.x[cur], R.v[cur] then come from scalar loads
    const int cur = __builtin_amdgcn_readfirstlane(R.state->cur);
    const long long step = uniform_step(R);
    const bool log = (mode & kModeLogForce) ||
                     ((mode & kModeLogAuto) && log_due(step, R.log_every));
    const BrickGeom g = brick_geom(R, it.local);
    staging_priority();
    const int O = stage_brick<real>(R, g, L, bp, cur, true, STYLE == 2);
    normal_priority();
    if (held) return;
    if (O < 0) {
        if (threadIdx.x == 0) atomicOr(&R.state->flags, kFlagStageOverflow);
        return;
    }
    const int t = threadIdx.x;
    const int N = R.N;
    const uint32_t seed_step = R.seed + (uint32_t)step;
    long long acc[kThermoParts];
#pragma unroll
    for (int k = 0; k < kThermoParts; ++k) acc[k] = 0;
    if (bp.phase) tp[1] = wall_clock64();
    brick_pairs<real, STYLE, false>(R, L, bp, O, R.ntypes, (real)R.rcmax2, cur, seed_step, log, acc,
                                    bp.phase ? &tp[2] : nullptr);
    if (bp.phase) tp[3] = wall_clock64();
    staging_priority();
    auto *F = gmemw(R.force);
    for (int o = t; o < O; o += kBlock) {
        const int slot = real_to_id(L.vown[o].w);
        F[slot] = (double)L.facc[3 * o + 0] * FixScale<real>::inv;
        F[N + slot] = (double)L.facc[3 * o + 1] * FixScale<real>::inv;
        F[2 * N + slot] = (double)L.facc[3 * o + 2] * FixScale<real>::inv;
        if (log && bp.perbead) {
            R.bpe[slot] = 0.0;
            R.ape[slot] = 0.0;
        }
    }
    if (log && !bp.perbead) {
        long long *red = (long long *)L.red, *out = red + 48;
        block_sum_i64<kThermoParts>(acc, red, out);
        if (t < kThermoParts) ((long long *)R.thermo_part)[(size_t)it.local * kThermoParts + t] = out[t];
    }
    if (bp.phase) {
        tp[4] = wall_clock64();
        add_phases(bp, item, tp);
    }
}

// One harmonic angle's force on bead s (force/angle_harmonic.cu:168-330: cos
// clamped to [-1, 1], sin floored at 0.001, a = -2k (theta - theta0) / sin),
// from the bead's CSR entry {type | role << 8, other ids}.  Adds the force to
// f; with log, adds a third of the angle energy to pe and a third of its
// virial to w.  Every member evaluates the same angle with the same
// arithmetic, so the three forces sum to zero to rounding.
__device__ __forceinline__ void angle_on_bead(const RepDev &R, const double *X, int s, const int *ent, double (&f)[3],
                                              bool log, double &pe, double (&w)[6])
{
    const int N = R.N;
    const int tr = ent[0], role = tr >> 8, at = tr & 255;
    const int sa = R.slot_of_id[ent[1] - 1], sb = R.slot_of_id[ent[2] - 1];
    const int si = role == 0 ? s : sa;
    const int sj = role == 1 ? s : (role == 0 ? sa : sb);
    const int sk = role == 2 ? s : sb;
    double d1[3], d2[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        d1[k] = min_image(X[k * N + si] - X[k * N + sj], R.L[k], R.hL[k]);
        d2[k] = min_image(X[k * N + sk] - X[k * N + sj], R.L[k], R.hL[k]);
    }
    const double rsq1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    const double rsq2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    const double r1 = sqrt(rsq1), r2 = sqrt(rsq2);
    double c = (d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]) / (r1 * r2);
    c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
    double sn = sqrt(1.0 - c * c);
    sn = sn < 0.001 ? 0.001 : sn;
    const double kk = R.angle_k[at];
    const double dth = acos(c) - R.angle_th0[at];
    const double pre = -2.0 * kk * dth / sn;
    const double a11 = pre * c / rsq1, a12 = -pre / (r1 * r2), a22 = pre * c / rsq2;
    double f1[3], f3[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        f1[k] = a11 * d1[k] + a12 * d2[k];
        f3[k] = a22 * d2[k] + a12 * d1[k];
        f[k] += role == 0 ? f1[k] : (role == 2 ? f3[k] : -(f1[k] + f3[k]));
    }
    if (log) {
        pe += (kk * dth * dth) / 3.0;
        w[0] += (d1[0] * f1[0] + d2[0] * f3[0]) / 3.0;
        w[1] += (d1[0] * f1[1] + d2[0] * f3[1]) / 3.0;
        w[2] += (d1[0] * f1[2] + d2[0] * f3[2]) / 3.0;
        w[3] += (d1[1] * f1[1] + d2[1] * f3[1]) / 3.0;
        w[4] += (d1[1] * f1[2] + d2[1] * f3[2]) / 3.0;
        w[5] += (d1[2] * f1[2] + d2[2] * f3[2]) / 3.0;
    }
}

// ---------------------------------------------------------------------------
// Fused bonds + integration, one thread per slot: bond forces
// (force/bond_harmonic.cu:196-268) are added to the pair force, then the final
// half-kick of this step (fix/fix_nve.cu:137-149), the thermo sums of a log step
// (compute/thermo.cu:243-245, :424-433), the half-kick and drift of the next step
// (fix/fix_nve.cu:48-86) and the new cell key.  The new state goes to the other
// buffer.
// ---------------------------------------------------------------------------

// TOPO: 0 no replica of the launch has bonds or angles, 1 bonds only, 2 angles
// too.  Code a launch cannot need is compiled out: the plain-fluid instance
// holds far fewer registers, so more waves keep HBM busy.
template <int TOPO>
__global__ __launch_bounds__(kBlock) void k_integrate(const RepDev *__restrict__ reps,
                                                      const WorkItem *__restrict__ items, int nitems, int mode)
{
    __shared__ long long red[64];
    __shared__ int hkey[kRankSlots], hcnt[kRankSlots], hbase[kRankSlots];
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int N = R.N;
    const int s = it.local * kBlock + threadIdx.x;
    // the force and the type do not depend on the buffer parity: their loads
    // go out before the replica's state block is read
    double fx = 0.0, fy = 0.0, fz = 0.0;
    int tt = 0;
    if (s < N) {
        const auto *F = gmem(R.force);
        fx = F[s];
        fy = F[N + s];
        fz = F[2 * N + s];
        tt = gmem(R.type)[s] - 1;
    }
    if (halted(R)) return;  // the force of this step is incomplete: hold the state
    // uniform: the state pointers R.x[cur], R.v[cur] then come from scalar loads
    const int cur = __builtin_amdgcn_readfirstlane(R.state->cur);
    const long long step = uniform_step(R);
    const bool log = (mode & kModeLogAuto) && log_due(step, R.log_every);
    double acc[kThermoParts];
#pragma unroll
    for (int k = 0; k < kThermoParts; ++k) acc[k] = 0.0;
    bool bad = false;
    int my_key = -1;
    if (s < N) {
        const double *X = R.x[cur];
        const auto *Xg = gmem(R.x[cur]);
        const auto *V = gmem(R.v[cur]);
        double xg = Xg[s], yg = Xg[N + s], zg = Xg[2 * N + s];
        double bpe = 0.0, bw[6] = {0, 0, 0, 0, 0, 0};
        if (TOPO >= 1 && R.nbonds > 0) {
            const int idi = R.id[s];
            const int b1 = R.bond_start[idi];
            for (int bb = R.bond_start[idi - 1]; bb < b1; ++bb) {
                const int ps = R.slot_of_id[R.bond_partner[bb] - 1];
                const double ddx = min_image(xg - Xg[ps], R.L[0], R.hL[0]);
                const double ddy = min_image(yg - Xg[N + ps], R.L[1], R.hL[1]);
                const double ddz = min_image(zg - Xg[2 * N + ps], R.L[2], R.hL[2]);
                const double r = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                const int bt = R.bond_type[bb];
                const double kb = R.bond_k[bt], r0 = R.bond_r0[bt];
                const double fb = (r > 0.0) ? -2.0 * kb * (r - r0) / r : 0.0;
                const double bfx = ddx * fb, bfy = ddy * fb, bfz = ddz * fb;
                fx += bfx;
                fy += bfy;
                fz += bfz;
                if (log) {
                    bpe += 0.5 * (kb * (r - r0) * (r - r0));
                    bw[0] += 0.5 * ddx * bfx;
                    bw[1] += 0.5 * ddx * bfy;
                    bw[2] += 0.5 * ddx * bfz;
                    bw[3] += 0.5 * ddy * bfy;
                    bw[4] += 0.5 * ddy * bfz;
                    bw[5] += 0.5 * ddz * bfz;
                }
            }
        }
        double ape = 0.0;
        if (TOPO >= 2 && R.nangles > 0) {
            double fa[3] = {fx, fy, fz};
            const int idi = R.id[s];
            const int a1 = R.angle_start[idi];
            for (int aa = R.angle_start[idi - 1]; aa < a1; ++aa)
                angle_on_bead(R, X, s, R.angle_ent + 4 * (size_t)aa, fa, log, ape, bw);
            fx = fa[0];
            fy = fa[1];
            fz = fa[2];
        }
        if ((R.nbonds > 0 || R.nangles > 0) && (mode & kModeStoreForce)) {
            R.force[s] = fx;
            R.force[N + s] = fy;
            R.force[2 * N + s] = fz;
        }
        const double h = (0.5 * R.dt) * R.minv[tt];
        double vxg = V[s], vyg = V[N + s], vzg = V[2 * N + s];
        // explicit FMAs: the two half-kicks round exactly as two separate
        // nve kernels do (fix/fix_nve.cu:54-56, :143-145)
        if (mode & kModeFinalKick) {
            vxg = fma(h, fx, vxg);
            vyg = fma(h, fy, vyg);
            vzg = fma(h, fz, vzg);
        }
        if (log) {
            const double m = R.mass[tt];
            acc[0] = 0.5 * m * (vxg * vxg + vyg * vyg + vzg * vzg);
            acc[2] = bpe;
            acc[9] = ape;
            acc[3] = m * vxg * vxg + bw[0];
            acc[4] = m * vxg * vyg + bw[1];
            acc[5] = m * vxg * vzg + bw[2];
            acc[6] = m * vyg * vyg + bw[3];
            acc[7] = m * vyg * vzg + bw[4];
            acc[8] = m * vzg * vzg + bw[5];
        }
        if (mode & kModeDrift) {
            vxg = fma(h, fx, vxg);
            vyg = fma(h, fy, vyg);
            vzg = fma(h, fz, vzg);
            // box crossings of this drift; the image word is read and
            // rewritten only for a bead that crossed
            int jx = 0, jy = 0, jz = 0;
            bool b2 = false;
            xg = wrap_axis(fma(R.dt, vxg, xg), R.lo[0], R.lo[0] + R.L[0], R.L[0], jx, b2);
            yg = wrap_axis(fma(R.dt, vyg, yg), R.lo[1], R.lo[1] + R.L[1], R.L[1], jy, b2);
            zg = wrap_axis(fma(R.dt, vzg, zg), R.lo[2], R.lo[2] + R.L[2], R.L[2], jz, b2);
            if (jx | jy | jz) {
                int ix, iy, iz;
                unpack_img(gmem(R.img)[s], ix, iy, iz);
                ix += jx;
                iy += jy;
                iz += jz;
                if (img_in_range(ix, iy, iz)) gmemw(R.img)[s] = pack_img(ix, iy, iz);
                else atomicOr(&R.state->flags, kFlagImageOverflow);
            }
            if (b2 || !isfinite(vxg + vyg + vzg)) {
                bad = true;
                xg = R.lo[0];
                yg = R.lo[1];
                zg = R.lo[2];
            }
            const int key = cell_of(R, xg, yg, zg);
            auto *Xo = gmemw(R.x[1 - cur]);
            auto *Vo = gmemw(R.v[1 - cur]);
            Xo[s] = xg;
            Xo[N + s] = yg;
            Xo[2 * N + s] = zg;
            Vo[s] = vxg;
            Vo[N + s] = vyg;
            Vo[2 * N + s] = vzg;
            gmemw(R.key)[s] = key;
            my_key = key;
        } else {
            auto *Xo = gmemw(R.x[1 - cur]);
            auto *Vo = gmemw(R.v[1 - cur]);
            Xo[s] = xg;
            Xo[N + s] = yg;
            Xo[2 * N + s] = zg;
            Vo[s] = vxg;
            Vo[N + s] = vyg;
            Vo[2 * N + s] = vzg;
        }
    }
    if (bad) atomicOr(&R.state->flags, kFlagNonFinite);
    if (mode & kModeDrift) {
        // rank of each bead in its new cell.  Neighbouring slots share cells,
        // so the block first groups its beads by cell in an LDS hash table
        // (local ranks from LDS atomics), then takes one global atomic per
        // distinct cell for the group's base.  Ranks within a cell need no
        // particular order.
        int h = 0, lrank = 0;
        for (int e = threadIdx.x; e < kRankSlots; e += kBlock) {
            hkey[e] = -1;
            hcnt[e] = 0;
        }
        __syncthreads();
        if (my_key >= 0) {
            h = (int)(((unsigned)my_key * 2654435761u) >> (32 - kRankBits));
            // at most kBlock distinct keys in kRankSlots > kBlock slots: every
            // probe sequence reaches the key or an empty slot
            for (;;) {
                const int prev = atomicCAS(&hkey[h], -1, my_key);
                if (prev == -1 || prev == my_key) break;
                h = (h + 1) & (kRankSlots - 1);
            }
            lrank = atomicAdd(&hcnt[h], 1);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < kRankSlots; e += kBlock)
            if (hkey[e] >= 0) hbase[e] = atomicAdd(&R.cell_count[hkey[e]], hcnt[e]);
        __syncthreads();
        if (my_key >= 0) gmemw(R.rank)[s] = hbase[h] + lrank;
    }
    if (log) {
        long long *out = red + 48;
        block_sum_fix<kThermoParts>(acc, red, out);
        if (threadIdx.x < kThermoParts)
            ((long long *)R.thermo_bead)[(size_t)it.local * kThermoParts + threadIdx.x] = out[threadIdx.x];
    }
}

// ---------------------------------------------------------------------------
// end of step, per replica: thermo record on log steps (fixed-order sum of the
// brick partials), then flip the buffer parity and advance the step counter.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_step_end(const RepDev *__restrict__ reps,
                                                     const int *__restrict__ rep_list, int flip,
                                                     int advance, int thermo, int bricks)
{
    __shared__ long long red[64];
    __shared__ long long itot[kThermoParts];
    const RepDev &R = reps[rep_list[blockIdx.x]];
    if (halted(R)) return;  // no record, no step count, no buffer flip
    const long long step = uniform_step(R);
    if (thermo && log_due(step, R.log_every)) {
        // exact integer sums of the fixed-point partials
        long long acc[kThermoParts];
#pragma unroll
        for (int k = 0; k < kThermoParts; ++k) acc[k] = 0;
        const long long *tp = (const long long *)R.thermo_part, *tb = (const long long *)R.thermo_bead;
        for (int b = threadIdx.x; bricks && b < R.nbrick; b += kBlock)
#pragma unroll
            for (int k = 0; k < kThermoParts; ++k) acc[k] += tp[(size_t)b * kThermoParts + k];
        const int nchunk = (R.N + kBlock - 1) / kBlock;
        for (int b = threadIdx.x; b < nchunk; b += kBlock)
#pragma unroll
            for (int k = 0; k < kThermoParts; ++k) acc[k] += tb[(size_t)b * kThermoParts + k];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < kThermoParts; ++k) {
            const long long sv = wave_sum_i64(acc[k]);
            if (lane == 0) red[w * kThermoParts + k] = sv;
        }
        __syncthreads();
        if (threadIdx.x < kThermoParts) {
            long long sv = 0;
            for (int ww = 0; ww < kBlock / 64; ++ww) sv += red[ww * kThermoParts + threadIdx.x];
            itot[threadIdx.x] = sv;
        }
        __syncthreads();
        double tot[kThermoParts];
#pragma unroll
        for (int k = 0; k < kThermoParts; ++k) tot[k] = (double)itot[k] / kThermoFix;
        if (threadIdx.x == 0) {
            const double vol = R.L[0] * R.L[1] * R.L[2];
            const int row = R.state->log_rows % R.thermo_cap;
            double *o = R.thermo_rows + (size_t)row * kThermoRowLen;
            const double ndof = 3.0 * R.N - 3.0;
            o[0] = (2.0 * tot[0]) / (ndof * 1.0);
            o[1] = tot[1];
            o[2] = tot[2];
            for (int k = 0; k < 6; ++k) o[4 + k] = tot[3 + k] / vol;
            o[3] = (o[4] + o[7] + o[9]) / 3.0;
            o[10] = tot[0];
            o[11] = (double)step;
            o[12] = tot[9];
            R.state->log_rows += 1;
        }
    }
    if (threadIdx.x == 0) {
        if (advance) R.step_ts[step & (kStepTsCap - 1)] = wall_clock64();
        R.state->step = step + advance;
        R.state->cur ^= flip;
    }
}

// copies of the listed replicas' scalar states (one block per replica)
__global__ void k_gather_state(const RepDev *__restrict__ reps, const int *__restrict__ rep_list,
                               RepState *__restrict__ out)
{
    const RepDev &R = reps[rep_list[blockIdx.x]];
    if (threadIdx.x == 0) out[blockIdx.x] = *R.state;
}

// the listed replicas' step-time rings and step counters (one block per replica)
__global__ __launch_bounds__(kBlock) void k_gather_step_ts(const RepDev *__restrict__ reps,
                                                           const int *__restrict__ rep_list,
                                                           unsigned long long *__restrict__ ts,
                                                           long long *__restrict__ steps)
{
    const RepDev &R = reps[rep_list[blockIdx.x]];
    for (int k = threadIdx.x; k < kStepTsCap; k += kBlock) ts[(size_t)blockIdx.x * kStepTsCap + k] = R.step_ts[k];
    if (threadIdx.x == 0) steps[blockIdx.x] = R.state->step;
}

// ---------------------------------------------------------------------------
// group fixes and computes (fix/fix_*.cu, compute/compute_*.cu).  A group is a
// mask of bead types, so membership is one bit test per bead and no index list
// is kept.  Every reduction is two-word fixed point: v * 2^16 split into its
// integer part and the remaining fraction * 2^32, each summed as int64.  The
// sums are exact and commute, so results do not depend on bead order, block
// count or slot layout; resolution is 2^-48 per term over |v| < 2^40 / N.
// Extents (wall positions) use order-preserving integer keys of the doubles.
// ---------------------------------------------------------------------------

constexpr double kSplitHi = 65536.0, kSplitLo = 4294967296.0;

__device__ __forceinline__ void split2(double v, long long &h, long long &l)
{
    const double t = v * kSplitHi, f = floor(t);
    h = (long long)f;
    l = __double2ll_rn((t - f) * kSplitLo);
}

__host__ __device__ __forceinline__ double join2(const long long *a)
{
    return (double)a[0] * (1.0 / kSplitHi) + (double)a[1] * (1.0 / (kSplitHi * kSplitLo));
}

__device__ __forceinline__ unsigned long long order_key(double d)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__host__ __device__ __forceinline__ double order_key_value(unsigned long long k)
{
    const unsigned long long u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    double d;
    memcpy(&d, &u, sizeof(d));
    return d;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

// Block-level accumulators of one fix/compute kernel: per-wave exact sums
// land in LDS, then one global atomic per nonzero word per block.
struct FixBlockAcc {
    long long *sum;            // [kRecSlots][kAccPer][2]
    unsigned long long *ext;   // [kRecSlots][2]
};

__device__ __forceinline__ void fba_init(FixBlockAcc &A)
{
    for (int i = threadIdx.x; i < kRecSlots * kAccPer * 2; i += kBlock) A.sum[i] = 0;
    for (int i = threadIdx.x; i < kRecSlots; i += kBlock) {
        A.ext[2 * i] = 0ull;
        A.ext[2 * i + 1] = ~0ull;
    }
    __syncthreads();
}

// every lane of the wave calls it; inactive lanes add zero
__device__ __forceinline__ void fba_add(FixBlockAcc &A, int slot, int q, double v, bool active)
{
    long long h = 0, l = 0;
    if (active) split2(v, h, l);
    h = wave_sum_i64(h);
    l = wave_sum_i64(l);
    if ((threadIdx.x & 63) == 0 && (h | l)) {
        atomicAdd((unsigned long long *)&A.sum[(slot * kAccPer + q) * 2], (unsigned long long)h);
        atomicAdd((unsigned long long *)&A.sum[(slot * kAccPer + q) * 2 + 1], (unsigned long long)l);
    }
}

__device__ __forceinline__ void fba_max(FixBlockAcc &A, int slot, double v, bool active)
{
    const unsigned long long k = wave_max_u64(active ? order_key(v) : 0ull);
    if ((threadIdx.x & 63) == 0 && k) atomicMax(&A.ext[2 * slot], k);
}

__device__ __forceinline__ void fba_min(FixBlockAcc &A, int slot, double v, bool active)
{
    const unsigned long long k = wave_min_u64(active ? order_key(v) : ~0ull);
    if ((threadIdx.x & 63) == 0 && ~k) atomicMin(&A.ext[2 * slot + 1], k);
}

__device__ __forceinline__ void fba_flush(FixBlockAcc &A, const RepDev &R)
{
    __syncthreads();
    for (int i = threadIdx.x; i < kRecSlots * kAccPer * 2; i += kBlock)
        if (A.sum[i]) atomicAdd((unsigned long long *)&R.racc[i], (unsigned long long)A.sum[i]);
    for (int i = threadIdx.x; i < kRecSlots; i += kBlock) {
        if (A.ext[2 * i]) atomicMax(&R.rext[2 * i], A.ext[2 * i]);
        if (~A.ext[2 * i + 1]) atomicMin(&R.rext[2 * i + 1], A.ext[2 * i + 1]);
    }
}

__device__ __forceinline__ double racc_value(const RepDev &R, int slot, int q)
{
    return join2(&R.racc[(slot * kAccPer + q) * 2]);
}

__device__ __forceinline__ bool in_group(unsigned mask, int type) { return (mask >> (type - 1)) & 1u; }

// reflection about a plane (fix/fix_wall_reflect.cu:96-118, fix_two_wall.cu:118-140):
// below lo, or else above hi, the coordinate is mirrored and the velocity flipped
__device__ __forceinline__ void reflect_axis(double &x, double &v, double lo, double hi)
{
    if (x < lo) {
        x += 2.0 * (lo - x);
        v = -v;
    } else if (x > hi) {
        x += 2.0 * (hi - x);
        v = -v;
    }
}

// spring of fix/fix_spring.cu:300-350 on the group's centre of mass (from the
// position sums of slot): total force {fx,fy,fz}, signed magnitude, energy
__device__ __forceinline__ void spring_terms(const RepDev &R, const FixDev &F, int slot, double (&o)[5])
{
    double d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) d[a] = racc_value(R, slot, a) / F.gmass - F.p[1 + a];
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double dr = r - F.p[4];
    o[0] = o[1] = o[2] = o[3] = o[4] = 0.0;
    if (r > 1.0e-10) {
#pragma unroll
        for (int a = 0; a < 3; ++a) o[a] = -F.p[0] * dr * d[a] / r;
        o[3] = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
        if (dr < 0.0) o[3] = -o[3];
        o[4] = 0.5 * F.p[0] * dr * dr;
    }
}

// zero the sums and reset the extents of every slot (one block per replica)
__global__ __launch_bounds__(kBlock) void k_fix_clear(const RepDev *__restrict__ reps, const int *__restrict__ rep_list)
{
    const RepDev &R = reps[rep_list[blockIdx.x]];
    if (halted(R)) return;
    for (int i = threadIdx.x; i < kRecSlots * kAccPer * 2; i += kBlock) R.racc[i] = 0;
    for (int i = threadIdx.x; i < kRecSlots; i += kBlock) {
        R.rext[2 * i] = 0ull;
        R.rext[2 * i + 1] = ~0ull;
    }
}

// initial_integrate and post_integrate of the fixes in fmask, in order:
// nve (fix_nve.cu:30-77: kick, drift, wrap), move (fix_move.cu:45-90: velocity
// set, drift, wrap), then wall_reflect; then the wall extents of two_wall
// (its reflection needs the whole group, k_fix_two_wall)
__global__ __launch_bounds__(kBlock) void k_fix_initial(const RepDev *__restrict__ reps,
                                                        const WorkItem *__restrict__ items, int nitems,
                                                        unsigned fmask)
{
    __shared__ long long lsum[kRecSlots * kAccPer * 2];
    __shared__ unsigned long long lext[kRecSlots * 2];
    FixBlockAcc A{lsum, lext};
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    if (halted(R)) return;
    fba_init(A);
    const int s = it.local * kBlock + threadIdx.x;
    const bool live = s < R.N;
    const int N = R.N, cur = R.state->cur;
    double *X = R.x[cur], *V = R.v[cur];
    const int type = live ? (int)R.type[s] : 1;
    double x[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    int im[3] = {0, 0, 0};
    bool moved = false, bad = false;
    if (live) {
        unpack_img(R.img[s], im[0], im[1], im[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            x[d] = X[d * N + s];
            v[d] = V[d * N + s];
        }
    }
    for (int k = 0; k < R.nfix; ++k) {
        if (!((fmask >> k) & 1u)) continue;
        const FixDev &F = R.fix[k];
        if (!live || !in_group(F.group, type)) continue;
        if (F.kind == 1) {  // nve
            const double h = (0.5 * R.dt) * R.minv[type - 1];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                v[d] = fma(h, R.force[d * N + s], v[d]);
                x[d] = wrap_axis(fma(R.dt, v[d], x[d]), R.lo[d], R.lo[d] + R.L[d], R.L[d], im[d], bad);
            }
            moved = true;
        } else if (F.kind == 2) {  // move
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                v[d] = F.p[d];
                x[d] = wrap_axis(x[d] + F.p[d] * R.dt, R.lo[d], R.lo[d] + R.L[d], R.L[d], im[d], bad);
            }
            moved = true;
        }
    }
    for (int k = 0; k < R.nfix; ++k) {
        if (!((fmask >> k) & 1u)) continue;
        const FixDev &F = R.fix[k];
        if (F.kind != 3 || !live || !in_group(F.group, type)) continue;
        const int d = F.dir;
        const double x0 = x[d], v0 = v[d];
        reflect_axis(x[d], v[d], F.p[0], F.p[1]);
        moved = moved || x[d] != x0 || v[d] != v0;
    }
    if (moved) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            X[d * N + s] = x[d];
            V[d * N + s] = v[d];
        }
        if (img_in_range(im[0], im[1], im[2])) R.img[s] = pack_img(im[0], im[1], im[2]);
        else atomicOr(&R.state->flags, kFlagImageOverflow);
        if (bad) atomicOr(&R.state->flags, kFlagNonFinite);
    }
    bool any_ext = false;
    for (int k = 0; k < R.nfix; ++k) {
        if (!((fmask >> k) & 1u)) continue;
        const FixDev &F = R.fix[k];
        if (F.kind != 4) continue;
        any_ext = true;
        fba_max(A, kMaxComputes + k, x[F.dir], live && in_group(F.group2, type));
        fba_min(A, kMaxComputes + k, x[F.dir], live && in_group(F.group3, type));
    }
    if (any_ext) fba_flush(A, R);
}

// two_wall reflection of the group between the top of the lower wall and the
// bottom of the upper wall (fix/fix_two_wall.cu:145-204)
__global__ __launch_bounds__(kBlock) void k_fix_two_wall(const RepDev *__restrict__ reps,
                                                         const WorkItem *__restrict__ items, int nitems,
                                                         unsigned fmask)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int s = it.local * kBlock + threadIdx.x;
    if (s >= R.N || halted(R)) return;
    const int N = R.N, cur = R.state->cur, type = R.type[s];
    double *X = R.x[cur], *V = R.v[cur];
    for (int k = 0; k < R.nfix; ++k) {
        if (!((fmask >> k) & 1u)) continue;
        const FixDev &F = R.fix[k];
        if (F.kind != 4) continue;
        const int slot = kMaxComputes + k;
        const double lo = order_key_value(R.rext[2 * slot]), hi = order_key_value(R.rext[2 * slot + 1]);
        if (s == 0 && !(hi > lo)) atomicOr(&R.state->flags, kFlagWallOrder);
        if (!in_group(F.group, type)) continue;
        const int d = F.dir;
        double x = X[d * N + s], v = V[d * N + s];
        const double x0 = x;
        reflect_axis(x, v, lo, hi);
        if (x != x0) {
            X[d * N + s] = x;
            V[d * N + s] = v;
        }
    }
}

// Group sums of one phase for the fixes in fmask and the computes in cmask.
//   phase 0: spring / recenter: sum m x_unwrapped (fix_spring.cu:240-290)
//   phase 1: rvcm: sum m v (fix_rvcm.cu:95-150); tf: sum f (compute_tf.cu:60-110)
//   phase 2: temp: sum m v^2 / 2 (compute_temp.cu:60-110); msd: sum |x_uw - x0|^2
//            (compute_msd.cu:70-120); rho: count in the region (compute_rho.cu:100-160);
//            wall_contact: extreme of the wall group (compute_wall_contact.cu:200-270)
//   phase 3: wall_contact: count in the contact slab (compute_wall_contact.cu:120-170)
__global__ __launch_bounds__(kBlock) void k_fix_reduce(const RepDev *__restrict__ reps,
                                                       const WorkItem *__restrict__ items, int nitems,
                                                       unsigned fmask, unsigned cmask, int phase)
{
    __shared__ long long lsum[kRecSlots * kAccPer * 2];
    __shared__ unsigned long long lext[kRecSlots * 2];
    FixBlockAcc A{lsum, lext};
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    if (halted(R)) return;
    fba_init(A);
    const int s = it.local * kBlock + threadIdx.x;
    const bool live = s < R.N;
    const int N = R.N, cur = R.state->cur;
    const int sv = live ? s : 0;
    const int type = R.type[sv];
    const double m = R.mass[type - 1];
    const double *X = R.x[cur], *V = R.v[cur];
    int im[3];
    unpack_img(R.img[sv], im[0], im[1], im[2]);
    double x[3], xu[3], v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d] = X[d * N + sv];
        xu[d] = x[d] + (double)im[d] * R.L[d];
        v[d] = V[d * N + sv];
    }
    for (int k = 0; k < R.nfix; ++k) {
        if (!((fmask >> k) & 1u)) continue;
        const FixDev &F = R.fix[k];
        const bool g = live && in_group(F.group, type);
        const int slot = kMaxComputes + k;
        if (phase == 0 && (F.kind == 5 || F.kind == 6)) {
#pragma unroll
            for (int d = 0; d < 3; ++d) fba_add(A, slot, d, m * xu[d], g);
        } else if (phase == 1 && F.kind == 7) {
#pragma unroll
            for (int d = 0; d < 3; ++d) fba_add(A, slot, d, m * v[d], g);
        }
    }
    for (int k = 0; k < R.ncomp; ++k) {
        if (!((cmask >> k) & 1u)) continue;
        const CompDev &C = R.comp[k];
        const bool g = live && in_group(C.group, type);
        if (phase == 1 && C.kind == 4) {
#pragma unroll
            for (int d = 0; d < 3; ++d) fba_add(A, k, d, R.force[d * N + sv], g);
        } else if (phase == 2 && C.kind == 1) {
            fba_add(A, k, 0, 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), g);
        } else if (phase == 2 && C.kind == 2) {
            const int i = R.id[sv] - 1;
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double dd = xu[d] - R.x0uw[(size_t)d * N + i];
                r2 += dd * dd;
            }
            fba_add(A, k, 0, r2, g);
        } else if (phase == 2 && C.kind == 3) {
            const bool in = x[0] >= C.p[0] && x[0] < C.p[1] && x[1] >= C.p[2] && x[1] < C.p[3] && x[2] >= C.p[4] &&
                            x[2] < C.p[5];
            fba_add(A, k, 0, 1.0, g && in);
        } else if (phase == 2 && C.kind == 5) {
            const bool w = live && in_group(C.group2, type);
            if (C.side > 0) fba_max(A, k, x[C.dir], w);
            else fba_min(A, k, x[C.dir], w);
        } else if (phase == 3 && C.kind == 5) {
            const double xd = x[C.dir];
            bool in;
            if (C.side > 0) {
                const double b = order_key_value(R.rext[2 * k]);
                in = xd >= b && xd < b + C.p[0];
            } else {
                const double b = order_key_value(R.rext[2 * k + 1]);
                in = xd >= b - C.p[0] && xd < b;
            }
            fba_add(A, k, 0, 1.0, g && in);
        }
    }
    fba_flush(A, R);
}

// post_force of spring (fix/fix_spring.cu:300-376): the group's spring force,
// shared in proportion to mass, added to the force
__global__ __launch_bounds__(kBlock) void k_fix_post_force(const RepDev *__restrict__ reps,
                                                           const WorkItem *__restrict__ items, int nitems,
                                                           unsigned fmask)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int s = it.local * kBlock + threadIdx.x;
    if (s >= R.N || halted(R)) return;
    const int N = R.N, type = R.type[s];
    for (int k = 0; k < R.nfix; ++k) {
        if (!((fmask >> k) & 1u)) continue;
        const FixDev &F = R.fix[k];
        if (F.kind != 5 || !in_group(F.group, type)) continue;
        double o[5];
        spring_terms(R, F, kMaxComputes + k, o);
        const double m = R.mass[type - 1];
#pragma unroll
        for (int d = 0; d < 3; ++d) R.force[d * N + s] += (o[d] / F.gmass) * m;
    }
}

// final_integrate of nve (fix/fix_nve.cu:120-152)
__global__ __launch_bounds__(kBlock) void k_fix_final(const RepDev *__restrict__ reps,
                                                      const WorkItem *__restrict__ items, int nitems,
                                                      unsigned fmask)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int s = it.local * kBlock + threadIdx.x;
    if (s >= R.N || halted(R)) return;
    const int N = R.N, type = R.type[s];
    double *V = R.v[R.state->cur];
    for (int k = 0; k < R.nfix; ++k) {
        if (!((fmask >> k) & 1u)) continue;
        const FixDev &F = R.fix[k];
        if (F.kind != 1 || !in_group(F.group, type)) continue;
        const double h = (0.5 * R.dt) * R.minv[type - 1];
#pragma unroll
        for (int d = 0; d < 3; ++d) V[d * N + s] = fma(h, R.force[d * N + s], V[d * N + s]);
    }
}

// end_of_step of rvcm (fix/fix_rvcm.cu:220-292: v -= flagged components of the
// group's centre-of-mass velocity) and recenter (fix/fix_recenter.cu:260-340:
// the group is translated so its unwrapped centre of mass sits at the target)
__global__ __launch_bounds__(kBlock) void k_fix_end(const RepDev *__restrict__ reps,
                                                    const WorkItem *__restrict__ items, int nitems,
                                                    unsigned fmask)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int s = it.local * kBlock + threadIdx.x;
    if (s >= R.N || halted(R)) return;
    const int N = R.N, type = R.type[s], cur = R.state->cur;
    double *X = R.x[cur], *V = R.v[cur];
    for (int k = 0; k < R.nfix; ++k) {
        if (!((fmask >> k) & 1u)) continue;
        const FixDev &F = R.fix[k];
        if (!in_group(F.group, type)) continue;
        const int slot = kMaxComputes + k;
        if (F.kind == 7) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if ((F.flags >> d) & 1) V[d * N + s] -= racc_value(R, slot, d) / F.gmass;
        } else if (F.kind == 6) {
            int im[3];
            unpack_img(R.img[s], im[0], im[1], im[2]);
            bool bad = false;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double shift = F.p[d] - racc_value(R, slot, d) / F.gmass;
                X[d * N + s] = wrap_axis(X[d * N + s] + shift, R.lo[d], R.lo[d] + R.L[d], R.L[d], im[d], bad);
            }
            if (img_in_range(im[0], im[1], im[2])) R.img[s] = pack_img(im[0], im[1], im[2]);
            else atomicOr(&R.state->flags, kFlagImageOverflow);
            if (bad) atomicOr(&R.state->flags, kFlagNonFinite);
        }
    }
}

// one record per firing compute and per recording spring (one block per replica)
__global__ __launch_bounds__(64) void k_fix_records(const RepDev *__restrict__ reps, const int *__restrict__ rep_list,
                                                    unsigned fmask, unsigned cmask, long long step)
{
    const RepDev &R = reps[rep_list[blockIdx.x]];
    if (threadIdx.x != 0 || halted(R)) return;
    for (int k = 0; k < R.ncomp; ++k) {
        if (!((cmask >> k) & 1u)) continue;
        const CompDev &C = R.comp[k];
        double o[5] = {0, 0, 0, 0, 0};
        if (C.kind == 1) o[0] = 2.0 * racc_value(R, k, 0) / (3.0 * C.gcount - 3.0);
        else if (C.kind == 2) o[0] = racc_value(R, k, 0) / (double)C.gcount;
        else if (C.kind == 3)
            o[0] = racc_value(R, k, 0) / ((C.p[1] - C.p[0]) * (C.p[3] - C.p[2]) * (C.p[5] - C.p[4]));
        else if (C.kind == 4)
            for (int d = 0; d < 3; ++d) o[d] = racc_value(R, k, d);
        else if (C.kind == 5) o[0] = racc_value(R, k, 0);
        const int row = R.rec_n[k] % kRecCap;
        double *r = R.rec_rows + ((size_t)k * kRecCap + row) * kRecLen;
        r[0] = (double)step;
        for (int q = 0; q < 5; ++q) r[1 + q] = o[q];
        R.rec_n[k] += 1;
    }
    for (int k = 0; k < R.nfix; ++k) {
        if (!((fmask >> k) & 1u)) continue;
        const FixDev &F = R.fix[k];
        if (F.kind != 5) continue;
        const int slot = kMaxComputes + k;
        double o[5];
        spring_terms(R, F, slot, o);
        const int row = R.rec_n[slot] % kRecCap;
        double *r = R.rec_rows + ((size_t)slot * kRecCap + row) * kRecLen;
        r[0] = (double)step;
        for (int q = 0; q < 5; ++q) r[1 + q] = o[q];
        R.rec_n[slot] += 1;
    }
}

// ---------------------------------------------------------------------------
// hook-path kernels (one reference hook each)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_nve(const RepDev *__restrict__ reps,
                                                const WorkItem *__restrict__ items, int nitems,
                                                int initial)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int s = it.local * kBlock + threadIdx.x;
    if (s >= R.N || halted(R)) return;
    const int N = R.N, cur = R.state->cur;
    double *X = R.x[cur], *V = R.v[cur];
    const double h = (0.5 * R.dt) * R.minv[R.type[s] - 1];
    double v[3], x[3];
    int im[3];
    unpack_img(R.img[s], im[0], im[1], im[2]);
    bool bad = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        v[d] = fma(h, R.force[d * N + s], V[d * N + s]);
        V[d * N + s] = v[d];
        if (initial) {
            x[d] = wrap_axis(fma(R.dt, v[d], X[d * N + s]), R.lo[d], R.lo[d] + R.L[d], R.L[d], im[d], bad);
            X[d * N + s] = x[d];
        }
    }
    if (initial) {
        if (img_in_range(im[0], im[1], im[2])) R.img[s] = pack_img(im[0], im[1], im[2]);
        else atomicOr(&R.state->flags, kFlagImageOverflow);
        if (bad) atomicOr(&R.state->flags, kFlagNonFinite);
    }
}

__global__ __launch_bounds__(kBlock) void k_bond_hook(const RepDev *__restrict__ reps,
                                                      const WorkItem *__restrict__ items,
                                                      int nitems, int log)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int s = it.local * kBlock + threadIdx.x;
    if (s >= R.N || halted(R)) return;
    const int N = R.N;
    const double *X = R.x[R.state->cur];
    const int id = R.id[s];
    // added to the pair force one bond at a time, in the order the fused
    // integration kernel uses, so both paths round identically
    double f[3] = {R.force[s], R.force[N + s], R.force[2 * N + s]};
    double w[6] = {0, 0, 0, 0, 0, 0}, bpe = 0.0;
    const int b1 = R.bond_start[id];
    for (int bb = R.bond_start[id - 1]; bb < b1; ++bb) {
        const int ps = R.slot_of_id[R.bond_partner[bb] - 1];
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = min_image(X[k * N + s] - X[k * N + ps], R.L[k], R.hL[k]);
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const int bt = R.bond_type[bb];
        const double kb = R.bond_k[bt], r0 = R.bond_r0[bt];
        const double fb = (r > 0.0) ? -2.0 * kb * (r - r0) / r : 0.0;
        double bf[3];
        for (int k = 0; k < 3; ++k) {
            bf[k] = d[k] * fb;
            f[k] += bf[k];
        }
        bpe += 0.5 * (kb * (r - r0) * (r - r0));
        w[0] += 0.5 * d[0] * bf[0];
        w[1] += 0.5 * d[0] * bf[1];
        w[2] += 0.5 * d[0] * bf[2];
        w[3] += 0.5 * d[1] * bf[1];
        w[4] += 0.5 * d[1] * bf[2];
        w[5] += 0.5 * d[2] * bf[2];
    }
    for (int k = 0; k < 3; ++k) R.force[k * N + s] = f[k];
    if (log) {
        R.bpe[s] = bpe;
        for (int k = 0; k < 6; ++k) R.vir[k * N + s] += w[k];
    }
}

// Angle::compute_force on the hook path: each bead adds its angles to the
// force in the order the fused integration kernel uses
__global__ __launch_bounds__(kBlock) void k_angle_hook(const RepDev *__restrict__ reps,
                                                       const WorkItem *__restrict__ items,
                                                       int nitems, int log)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int s = it.local * kBlock + threadIdx.x;
    if (s >= R.N || halted(R)) return;
    const int N = R.N;
    const double *X = R.x[R.state->cur];
    const int id = R.id[s];
    double f[3] = {R.force[s], R.force[N + s], R.force[2 * N + s]};
    double w[6] = {0, 0, 0, 0, 0, 0}, ape = 0.0;
    const int a1 = R.angle_start[id];
    for (int aa = R.angle_start[id - 1]; aa < a1; ++aa) angle_on_bead(R, X, s, R.angle_ent + 4 * (size_t)aa, f, log != 0, ape, w);
    for (int k = 0; k < 3; ++k) R.force[k * N + s] = f[k];
    if (log) {
        R.ape[s] = ape;
        for (int k = 0; k < 6; ++k) R.vir[k * N + s] += w[k];
    }
}

// Total of nchunk per-chunk thermo partials, summed by one thread in chunk
// order so it rounds exactly as the host's sequential sum of the same
// partials does.
__global__ void k_thermo_total(const double *__restrict__ partials, int nchunk, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    double tot[kThermoParts];
#pragma unroll
    for (int k = 0; k < kThermoParts; ++k) tot[k] = 0.0;
    for (int b = 0; b < nchunk; ++b)
#pragma unroll
        for (int k = 0; k < kThermoParts; ++k) tot[k] += partials[(size_t)b * kThermoParts + k];
#pragma unroll
    for (int k = 0; k < kThermoParts; ++k) out[k] = tot[k];
}

// per-chunk thermo partials from the per-bead arrays: FP64 into partials, or
// (partials == nullptr) fixed point into the replica's thermo_bead like the
// fused integrate writes them
__global__ __launch_bounds__(kBlock) void k_thermo_arrays(const RepDev *__restrict__ reps,
                                                          const WorkItem *__restrict__ items,
                                                          double *__restrict__ partials)
{
    __shared__ double red[16 * (kBlock / 64)];
    __shared__ double out[kThermoParts];
    __shared__ long long fred[64];
    const WorkItem it = items[blockIdx.x];
    const RepDev &R = reps[it.rep];
    if (halted(R)) return;
    const int s = it.local * kBlock + threadIdx.x;
    double acc[kThermoParts];
#pragma unroll
    for (int k = 0; k < kThermoParts; ++k) acc[k] = 0.0;
    if (s < R.N) {
        const int N = R.N;
        const double *V = R.v[R.state->cur];
        const double m = R.mass[R.type[s] - 1];
        const double vx = V[s], vy = V[N + s], vz = V[2 * N + s];
        acc[0] = 0.5 * m * (vx * vx + vy * vy + vz * vz);
        acc[1] = R.pe[s];
        acc[2] = R.bpe[s];
        acc[9] = R.ape[s];
        acc[3] = m * vx * vx + R.vir[s];
        acc[4] = m * vx * vy + R.vir[N + s];
        acc[5] = m * vx * vz + R.vir[2 * N + s];
        acc[6] = m * vy * vy + R.vir[3 * N + s];
        acc[7] = m * vy * vz + R.vir[4 * N + s];
        acc[8] = m * vz * vz + R.vir[5 * N + s];
    }
    if (!partials) {
        long long *fout = fred + 48;
        block_sum_fix<kThermoParts>(acc, fred, fout);
        if (threadIdx.x < kThermoParts)
            ((long long *)R.thermo_bead)[(size_t)it.local * kThermoParts + threadIdx.x] = fout[threadIdx.x];
        return;
    }
    block_sum<kThermoParts>(acc, red, out);
    if (threadIdx.x < kThermoParts) partials[(size_t)blockIdx.x * kThermoParts + threadIdx.x] = out[threadIdx.x];
}

// ---------------------------------------------------------------------------
// reference neighbour list (Neigh_list::build / check_update), single replica.
// Bins of global_cut + skin with the clamp of force/neigh_list.cu:272-280;
// bead index = id - 1; half list i < j; r^2 <= bin^2; column-major rows
// neigh[k * N + i] like neigh_list.cu:384, with a bound check the reference
// lacks.
// ---------------------------------------------------------------------------

struct NlGeom {
    double bin;
    int nb[3];
};

__device__ __forceinline__ int nl_bin_axis(double x, double lo, double bin, int nb)
{
    int b = (int)((x - lo) / bin);
    return min(max(b, 0), nb - 1);
}

__global__ __launch_bounds__(kBlock) void k_nl_bin(const RepDev *__restrict__ rep, NlGeom g,
                                                   int *__restrict__ binof, int *__restrict__ count)
{
    const RepDev &R = *rep;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= R.N) return;
    const int N = R.N;
    const double *X = R.x[R.state->cur];
    const int bx = nl_bin_axis(X[s], R.lo[0], g.bin, g.nb[0]);
    const int by = nl_bin_axis(X[N + s], R.lo[1], g.bin, g.nb[1]);
    const int bz = nl_bin_axis(X[2 * N + s], R.lo[2], g.bin, g.nb[2]);
    const int b = bx * g.nb[1] * g.nb[2] + by * g.nb[2] + bz;
    const int i = R.id[s] - 1;
    binof[i] = b;
    atomicAdd(&count[b], 1);
}

__global__ __launch_bounds__(kBlock) void k_nl_fill(const RepDev *__restrict__ rep,
                                                    const int *__restrict__ binof,
                                                    const int *__restrict__ start,
                                                    int *__restrict__ fillc, int *__restrict__ content)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= rep->N) return;
    const int b = binof[i];
    content[start[b] + atomicAdd(&fillc[b], 1)] = i;
}

__global__ __launch_bounds__(kBlock) void k_nl_build(const RepDev *__restrict__ rep, NlGeom g,
                                                     const int *__restrict__ binof,
                                                     const int *__restrict__ start,
                                                     const int *__restrict__ content, int n_max,
                                                     int *__restrict__ n_neigh,
                                                     int *__restrict__ neigh,
                                                     int *__restrict__ overflow,
                                                     double *__restrict__ last_pos)
{
    const RepDev &R = *rep;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= R.N) return;
    const int N = R.N;
    const double *X = R.x[R.state->cur];
    const int si = R.slot_of_id[i];
    const double xi = X[si], yi = X[N + si], zi = X[2 * N + si];
    last_pos[3 * i + 0] = xi;
    last_pos[3 * i + 1] = yi;
    last_pos[3 * i + 2] = zi;
    const int b = binof[i];
    const int bz = b % g.nb[2], by = (b / g.nb[2]) % g.nb[1], bx = b / (g.nb[2] * g.nb[1]);
    const double cut2 = g.bin * g.bin;
    int cnt = 0;
    for (int oz = -1; oz <= 1; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int cx = (bx + ox + g.nb[0]) % g.nb[0];
                const int cy = (by + oy + g.nb[1]) % g.nb[1];
                const int cz = (bz + oz + g.nb[2]) % g.nb[2];
                const int c = cx * g.nb[1] * g.nb[2] + cy * g.nb[2] + cz;
                for (int k = start[c]; k < start[c + 1]; ++k) {
                    const int j = content[k];
                    if (i >= j) continue;
                    const int sj = R.slot_of_id[j];
                    const double dx = min_image(xi - X[sj], R.L[0], R.hL[0]);
                    const double dy = min_image(yi - X[N + sj], R.L[1], R.hL[1]);
                    const double dz = min_image(zi - X[2 * N + sj], R.L[2], R.hL[2]);
                    if (dx * dx + dy * dy + dz * dz <= cut2) {
                        if (cnt < n_max) neigh[(size_t)cnt * N + i] = j;
                        ++cnt;
                    }
                }
            }
    if (cnt > n_max) {
        atomicOr(overflow, 1);
        cnt = n_max;
    }
    n_neigh[i] = cnt;
}

__global__ __launch_bounds__(kBlock) void k_nl_check(const RepDev *__restrict__ rep, double hs2,
                                                     const double *__restrict__ last_pos,
                                                     int *__restrict__ flag)
{
    const RepDev &R = *rep;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= R.N) return;
    const int N = R.N, i = R.id[s] - 1;
    const double *X = R.x[R.state->cur];
    const double dx = min_image(X[s] - last_pos[3 * i + 0], R.L[0], R.hL[0]);
    const double dy = min_image(X[N + s] - last_pos[3 * i + 1], R.L[1], R.hL[1]);
    const double dz = min_image(X[2 * N + s] - last_pos[3 * i + 2], R.L[2], R.hL[2]);
    if (dx * dx + dy * dy + dz * dz >= hs2) *flag = 1;
}

__global__ __launch_bounds__(kBlock) void k_saru_eval(int n, const uint32_t *__restrict__ t1,
                                                      const uint32_t *__restrict__ t2, uint32_t seed,
                                                      const int32_t *__restrict__ steps,
                                                      uint32_t *__restrict__ bits, float *__restrict__ r32,
                                                      double *__restrict__ r64)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint32_t b = saru_bits(t1[k], t2[k], seed + (uint32_t)steps[k]);
    bits[k] = b;
    r32[k] = saru_uniform<float>(b);
    r64[k] = saru_uniform<double>(b);
}

template <typename real>
__global__ void k_fix_to_double(double *p, long long n)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = (double)__double_as_longlong(p[i]) * FixScale<real>::inv;
}

__global__ void k_fill_i32(int *p, int v, long long n)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double *p, double v, long long n)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

size_t BrickLaunch::lds_bytes(int precision, int style) const
{
    BrickParams p{rmax, smax, omax, nrun, ntypes, qw, perbead, style, nullptr};
    return precision == 1 ? lds_offsets<double>(p).total : lds_offsets<float>(p).total;
}

void launch_bin_keys(const RepDev *reps, const WorkItem *items, int nitems, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_bin_keys, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems);
}
void launch_scatter(const RepDev *reps, const WorkItem *items, int nitems, hipStream_t s)
{
    if (nitems)
        hipLaunchKernelGGL(k_scatter, dim3((nitems + kScatterItems - 1) / kScatterItems), dim3(kBlock), 0, s, reps,
                           items, nitems);
}
void launch_cellsort(const RepDev *reps, const WorkItem *items, int nitems, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_cellsort, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems);
}
void launch_scan(const RepDev *reps, const WorkItem *items, int nitems, const int *rep_chunk_off,
                 const int *rep_list, int nreps, int *chunk_sums, hipStream_t s)
{
    (void)rep_list;
    if (!nitems) return;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nitems), dim3(kBlock), 0, s, reps, items, chunk_sums);
    hipLaunchKernelGGL(k_scan_top, dim3(nreps), dim3(kBlock), 0, s, rep_chunk_off, chunk_sums);
    hipLaunchKernelGGL(k_scan_down, dim3(nitems), dim3(kBlock), 0, s, reps, items,
                       (const int *)chunk_sums);
}
void launch_resort(RepDev *reps, const WorkItem *items, int nitems, const int *rep_list, int nreps, hipStream_t s)
{
    if (!nitems) return;
    hipLaunchKernelGGL(k_resort, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems);
    hipLaunchKernelGGL(k_resort_swap, dim3(nreps), dim3(1), 0, s, reps, rep_list);
}
void launch_density(const RepDev *reps, const WorkItem *items, int nitems, const BrickLaunch &bl,
                    int precision, hipStream_t s)
{
    if (!nitems) return;
    BrickParams p{bl.rmax, bl.smax, bl.omax, bl.nrun, bl.ntypes, bl.qw, 0, 1, nullptr, bl.single_rep};
    const size_t lds = bl.lds_bytes(precision, 1);
    if (precision == 1)
        hipLaunchKernelGGL(k_density<double>, dim3(nitems), dim3(kBlock), lds, s, reps, items, nitems, p);
    else
        hipLaunchKernelGGL(k_density<float>, dim3(nitems), dim3(kBlock), lds, s, reps, items, nitems, p);
}
void launch_brick(const RepDev *reps, const WorkItem *items, int nitems, const BrickLaunch &bl,
                  int precision, int style, int mode, hipStream_t s)
{
    if (!nitems) return;
    BrickParams p{bl.rmax, bl.smax, bl.omax, bl.nrun, bl.ntypes, bl.qw, bl.perbead, style, bl.phase, bl.single_rep};
    const size_t lds = bl.lds_bytes(precision, style);
    if (precision == 1) {
        if (style == 2)
            hipLaunchKernelGGL((k_brick<double, 2>), dim3(nitems), dim3(kBlock), lds, s, reps, items, nitems, p, mode);
        else
            hipLaunchKernelGGL((k_brick<double, 1>), dim3(nitems), dim3(kBlock), lds, s, reps, items, nitems, p, mode);
    } else {
        if (style == 2)
            hipLaunchKernelGGL((k_brick<float, 2>), dim3(nitems), dim3(kBlock), lds, s, reps, items, nitems, p, mode);
        else
            hipLaunchKernelGGL((k_brick<float, 1>), dim3(nitems), dim3(kBlock), lds, s, reps, items, nitems, p, mode);
    }
}
void launch_integrate(const RepDev *reps, const WorkItem *items, int nitems, int mode, int topo, hipStream_t s)
{
    if (!nitems) return;
    if (topo == 0)
        hipLaunchKernelGGL(k_integrate<0>, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, mode);
    else if (topo == 1)
        hipLaunchKernelGGL(k_integrate<1>, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, mode);
    else
        hipLaunchKernelGGL(k_integrate<2>, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, mode);
}
void launch_step_end(const RepDev *reps, const int *rep_list, int nreps, int flip, int advance,
                     int thermo, hipStream_t s, int bricks)
{
    if (nreps)
        hipLaunchKernelGGL(k_step_end, dim3(nreps), dim3(kBlock), 0, s, reps, rep_list, flip, advance, thermo, bricks);
}

void launch_gather_state(const RepDev *reps, const int *rep_list, int nreps, RepState *out, hipStream_t s)
{
    if (nreps) hipLaunchKernelGGL(k_gather_state, dim3(nreps), dim3(64), 0, s, reps, rep_list, out);
}

void launch_gather_step_ts(const RepDev *reps, const int *rep_list, int nreps, unsigned long long *ts,
                           long long *steps, hipStream_t s)
{
    if (nreps) hipLaunchKernelGGL(k_gather_step_ts, dim3(nreps), dim3(kBlock), 0, s, reps, rep_list, ts, steps);
}
void launch_fix_clear(const RepDev *reps, const int *rep_list, int nreps, hipStream_t s)
{
    if (nreps) hipLaunchKernelGGL(k_fix_clear, dim3(nreps), dim3(kBlock), 0, s, reps, rep_list);
}
void launch_fix_initial(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_fix_initial, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, fmask);
}
void launch_fix_two_wall(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_fix_two_wall, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, fmask);
}
void launch_fix_reduce(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, unsigned cmask,
                       int phase, hipStream_t s)
{
    if (nitems)
        hipLaunchKernelGGL(k_fix_reduce, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, fmask, cmask, phase);
}
void launch_fix_post_force(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_fix_post_force, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, fmask);
}
void launch_fix_final(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_fix_final, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, fmask);
}
void launch_fix_end(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_fix_end, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, fmask);
}
void launch_fix_records(const RepDev *reps, const int *rep_list, int nreps, unsigned fmask, unsigned cmask,
                        long long step, hipStream_t s)
{
    if (nreps) hipLaunchKernelGGL(k_fix_records, dim3(nreps), dim3(64), 0, s, reps, rep_list, fmask, cmask, step);
}
void launch_nve(const RepDev *reps, const WorkItem *items, int nitems, int initial, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_nve, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, initial);
}
void launch_bond_hook(const RepDev *reps, const WorkItem *items, int nitems, int log, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_bond_hook, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, log);
}

void launch_angle_hook(const RepDev *reps, const WorkItem *items, int nitems, int log, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_angle_hook, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, log);
}
void launch_thermo_arrays(const RepDev *reps, const WorkItem *items, int nitems, double *partials,
                          hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_thermo_arrays, dim3(nitems), dim3(kBlock), 0, s, reps, items, partials);
}
void launch_thermo_total(const double *partials, int nchunk, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_thermo_total, dim3(1), dim3(64), 0, s, partials, nchunk, out);
}
void launch_nl_bin(const RepDev *rep, int N, double bin, const int *nb, int *binof, int *count,
                   hipStream_t s)
{
    NlGeom g{bin, {nb[0], nb[1], nb[2]}};
    if (N) hipLaunchKernelGGL(k_nl_bin, dim3(nblk(N)), dim3(kBlock), 0, s, rep, g, binof, count);
}
void launch_nl_fill(const RepDev *rep, int N, const int *binof, const int *start, int *fillc,
                    int *content, hipStream_t s)
{
    if (N) hipLaunchKernelGGL(k_nl_fill, dim3(nblk(N)), dim3(kBlock), 0, s, rep, binof, start, fillc, content);
}
void launch_nl_build(const RepDev *rep, int N, double bin, const int *nb, const int *binof,
                     const int *start, const int *content, int n_max, int *n_neigh, int *neigh,
                     int *overflow, double *last_pos, hipStream_t s)
{
    NlGeom g{bin, {nb[0], nb[1], nb[2]}};
    if (N)
        hipLaunchKernelGGL(k_nl_build, dim3(nblk(N)), dim3(kBlock), 0, s, rep, g, binof, start,
                           content, n_max, n_neigh, neigh, overflow, last_pos);
}
void launch_nl_check(const RepDev *rep, int N, double skin, const double *last_pos, int *flag,
                     hipStream_t s)
{
    const double hs2 = (skin * 0.5) * (skin * 0.5);
    if (N) hipLaunchKernelGGL(k_nl_check, dim3(nblk(N)), dim3(kBlock), 0, s, rep, hs2, last_pos, flag);
}
void launch_saru_eval(int n, const uint32_t *t1, const uint32_t *t2, uint32_t seed, const int32_t *steps,
                      uint32_t *bits, float *r32, double *r64, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_saru_eval, dim3(nblk(n)), dim3(kBlock), 0, s, n, t1, t2, seed, steps, bits, r32, r64);
}
void launch_fix_to_double(double *p, long long n, int precision, hipStream_t s)
{
    if (n <= 0) return;
    if (precision == 1) hipLaunchKernelGGL(k_fix_to_double<double>, dim3(nblk(n)), dim3(kBlock), 0, s, p, n);
    else hipLaunchKernelGGL(k_fix_to_double<float>, dim3(nblk(n)), dim3(kBlock), 0, s, p, n);
}
void launch_fill_i32(int *p, int v, long long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(kBlock), 0, s, p, v, n);
}
void launch_fill_f64(double *p, double v, long long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3(nblk(n)), dim3(kBlock), 0, s, p, v, n);
}

} // namespace gdpd
