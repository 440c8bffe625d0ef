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
// to the brick corner, periodic images applied while staging), builds each owned
// bead's in-range list in LDS, evaluates every pair's DPD/mDPD force with the
// Saru random term, adds the bead's bonds, and integrates the bead in the same
// launch (final half-kick of step n, thermo on log steps, half-kick and drift
// of step n+1), writing the new state to the other buffer and binning it.
// Forces are summed per bead in a fixed order in FP64, so the result does not
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

__device__ __forceinline__ int wave_incl_scan(int x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

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

// ---------------------------------------------------------------------------
// binning: keys, scan, scatter, in-cell sort, physical re-sort
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_bin_keys(const RepDev *__restrict__ reps,
                                                     const WorkItem *__restrict__ items, int nitems)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int s = it.local * kBlock + threadIdx.x;
    if (s >= R.N) return;
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
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int c = base + k;
        if (c < R.ncell) loc += R.cell_count[c];
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
    const int base = it.local * 4 * kBlock + 4 * threadIdx.x;
    int v[4], loc = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int c = base + k;
        v[k] = (c < R.ncell) ? R.cell_count[c] : 0;
        loc += v[k];
    }
    int total;
    int run = chunk_sums[blockIdx.x] + block_excl_scan(loc, wbuf, total);
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

__global__ __launch_bounds__(kBlock) void k_scatter(const RepDev *__restrict__ reps,
                                                    const WorkItem *__restrict__ items, int nitems)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int s = it.local * kBlock + threadIdx.x;
    if (s >= R.N) return;
    const int key = R.key[s];
    const int pos = ((unsigned)key < (unsigned)R.ncell) ? R.cell_start[key] + R.rank[s] : -1;
    if ((unsigned)pos < (unsigned)R.N) R.perm[pos] = s;
    else atomicOr(&R.state->flags, kFlagStageOverflow);
}

// sort each cell's beads by id so the in-cell order, and with it every
// per-bead summation order, depends only on the state: not on atomic
// scheduling, on the slot layout (re-sort interval) or on the batch
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

__global__ __launch_bounds__(kBlock) void k_resort_gather(const RepDev *__restrict__ reps,
                                                          const WorkItem *__restrict__ items,
                                                          int nitems)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int d = it.local * kBlock + threadIdx.x;
    if (d >= R.N) return;
    const int N = R.N, cur = R.state->cur;
    const int s = R.perm[d];
    const double *Xi = R.x[cur], *Vi = R.v[cur];
    double *Xo = R.x[1 - cur], *Vo = R.v[1 - cur];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        Xo[k * N + d] = Xi[k * N + s];
        Vo[k * N + d] = Vi[k * N + s];
    }
    R.iscratch[d] = R.id[s];
    R.iscratch[N + d] = R.type[s];
    R.iscratch[2 * N + d] = R.img[s];
}

__global__ __launch_bounds__(kBlock) void k_resort_finish(const RepDev *__restrict__ reps,
                                                          const WorkItem *__restrict__ items,
                                                          int nitems)
{
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int d = it.local * kBlock + threadIdx.x;
    if (it.local == 0 && threadIdx.x == 0) R.state->cur ^= 1;
    if (d >= R.N) return;
    const int N = R.N;
    const int id = R.iscratch[d];
    R.id[d] = id;
    R.type[d] = (unsigned char)R.iscratch[N + d];
    R.img[d] = R.iscratch[2 * N + d];
    R.perm[d] = d;
    if ((unsigned)(id - 1) < (unsigned)N) R.slot_of_id[id - 1] = d;
}

// ---------------------------------------------------------------------------
// brick staging shared by the density and force kernels
// ---------------------------------------------------------------------------

struct BrickParams {
    int rmax, smax, kmax, ncol, ntypes;
};

template <typename real> struct BrickLds {
    PairCoef<real> *coef;
    double *red;
    real *px, *py, *pz, *vx, *vy, *vz, *rh;
    int *c_start, *c_src, *sid, *colpre, *wbuf;
    unsigned short *scell, *list;
    unsigned char *sty;
};

__host__ __device__ inline size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

// LDS layout of the brick kernels as byte offsets.  The pointers are rebuilt
// inline from the __shared__ base in every kernel so they stay in the LDS
// address space (ds_read/ds_write), never in generic pointers or scratch.
struct LdsOffsets {
    unsigned coef, red, px, py, pz, vx, vy, vz, rh, c_start, c_src, sid, colpre, wbuf, scell, list, sty,
        total;
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
    GDPD_TAKE(red, sizeof(double) * 64)
    GDPD_TAKE(px, sizeof(real) * p.smax)
    GDPD_TAKE(py, sizeof(real) * p.smax)
    GDPD_TAKE(pz, sizeof(real) * p.smax)
    GDPD_TAKE(vx, sizeof(real) * p.smax)
    GDPD_TAKE(vy, sizeof(real) * p.smax)
    GDPD_TAKE(vz, sizeof(real) * p.smax)
    GDPD_TAKE(rh, sizeof(real) * p.smax)
    GDPD_TAKE(c_start, sizeof(int) * (p.rmax + 1))
    GDPD_TAKE(c_src, sizeof(int) * p.rmax)
    GDPD_TAKE(sid, sizeof(int) * p.smax)
    GDPD_TAKE(colpre, sizeof(int) * (p.ncol + 1))
    GDPD_TAKE(wbuf, sizeof(int) * (kBlock / 64))
    GDPD_TAKE(scell, sizeof(unsigned short) * p.smax)
    GDPD_TAKE(list, sizeof(unsigned short) * (size_t)p.kmax * kBlock)
    GDPD_TAKE(sty, sizeof(unsigned char) * p.smax)
#undef GDPD_TAKE
    o.total = (unsigned)off;
    return o;
}

template <typename real>
__device__ __forceinline__ BrickLds<real> lds_bind(char *smem, const LdsOffsets &o)
{
    BrickLds<real> L;
    L.coef = (PairCoef<real> *)(smem + o.coef);
    L.red = (double *)(smem + o.red);
    L.px = (real *)(smem + o.px);
    L.py = (real *)(smem + o.py);
    L.pz = (real *)(smem + o.pz);
    L.vx = (real *)(smem + o.vx);
    L.vy = (real *)(smem + o.vy);
    L.vz = (real *)(smem + o.vz);
    L.rh = (real *)(smem + o.rh);
    L.c_start = (int *)(smem + o.c_start);
    L.c_src = (int *)(smem + o.c_src);
    L.sid = (int *)(smem + o.sid);
    L.colpre = (int *)(smem + o.colpre);
    L.wbuf = (int *)(smem + o.wbuf);
    L.scell = (unsigned short *)(smem + o.scell);
    L.list = (unsigned short *)(smem + o.list);
    L.sty = (unsigned char *)(smem + o.sty);
    return L;
}

struct BrickGeom {
    int c0[3], ext[3], RX, RY, RZ, nreg;
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
    return g;
}

__device__ __forceinline__ int wrap_cell(int c, int n, double L, double &shift)
{
    if (c < 0) {
        shift = -L;
        return c + n;
    }
    if (c >= n) {
        shift = L;
        return c - n;
    }
    shift = 0.0;
    return c;
}

// Stage the brick + halo into LDS.  Returns the number of staged beads, or -1
// (uniformly) when it does not fit.  need: 1 = velocities + ids, 2 = density.
template <typename real>
__device__ __forceinline__ int stage_brick(const RepDev &R, const BrickGeom &g, const BrickLds<real> &L,
                           const BrickParams &bp, int cur, bool need_vel, bool need_rho)
{
    const int t = threadIdx.x;
    for (int rc = t; rc < g.nreg; rc += kBlock) {
        const int rz = rc % g.RZ, ry = (rc / g.RZ) % g.RY, rx = rc / (g.RZ * g.RY);
        double sh;
        const int gx = wrap_cell(g.c0[0] - 1 + rx, R.nc[0], R.L[0], sh);
        const int gy = wrap_cell(g.c0[1] - 1 + ry, R.nc[1], R.L[1], sh);
        const int gz = wrap_cell(g.c0[2] - 1 + rz, R.nc[2], R.L[2], sh);
        const int gc = (gx * R.nc[1] + gy) * R.nc[2] + gz;
        L.c_src[rc] = gc;
        L.c_start[rc + 1] = R.cell_start[gc + 1] - R.cell_start[gc];
    }
    __syncthreads();
    {
        int v[4], loc = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int rc = 4 * t + k;
            v[k] = (rc < g.nreg) ? L.c_start[rc + 1] : 0;
            loc += v[k];
        }
        int total;
        int run = block_excl_scan(loc, L.wbuf, total);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int rc = 4 * t + k;
            if (rc < g.nreg) L.c_start[rc] = run;
            run += v[k];
        }
        if (t == 0) L.c_start[g.nreg] = total;
    }
    __syncthreads();
    const int S = L.c_start[g.nreg];
    if (S > bp.smax) return -1;

    const int N = R.N;
    const double *X = R.x[cur], *V = R.v[cur];
    for (int k = t; k < S; k += kBlock) {
        int lo = 0, hi = g.nreg;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (L.c_start[mid] <= k) lo = mid;
            else hi = mid;
        }
        const int rc = lo;
        const int rz = rc % g.RZ, ry = (rc / g.RZ) % g.RY, rx = rc / (g.RZ * g.RY);
        const int gc = L.c_src[rc];
        int slot = R.perm[R.cell_start[gc] + (k - L.c_start[rc])];
        if ((unsigned)slot >= (unsigned)N) slot = 0;
        double shx, shy, shz;
        wrap_cell(g.c0[0] - 1 + rx, R.nc[0], R.L[0], shx);
        wrap_cell(g.c0[1] - 1 + ry, R.nc[1], R.L[1], shy);
        wrap_cell(g.c0[2] - 1 + rz, R.nc[2], R.L[2], shz);
        L.px[k] = (real)((X[slot] + shx) - g.org[0]);
        L.py[k] = (real)((X[N + slot] + shy) - g.org[1]);
        L.pz[k] = (real)((X[2 * N + slot] + shz) - g.org[2]);
        L.sty[k] = R.type[slot];
        L.scell[k] = (unsigned short)rc;
        if (need_vel) {
            L.vx[k] = (real)V[slot];
            L.vy[k] = (real)V[N + slot];
            L.vz[k] = (real)V[2 * N + slot];
            L.sid[k] = R.id[slot];
        }
        if (need_rho) L.rh[k] = (real)R.rho[slot];
    }
    const int ncol = g.ext[0] * g.ext[1];
    for (int k = t; k < R.ntypes * R.ntypes; k += kBlock)
        L.coef[k] = ((const PairCoef<real> *)R.coef)[k];
    if (t == 0) {
        L.colpre[0] = 0;
        for (int q = 0; q < ncol; ++q) {
            const int rx = 1 + q / g.ext[1], ry = 1 + q % g.ext[1];
            const int c = (rx * g.RY + ry) * g.RZ;
            L.colpre[q + 1] = L.colpre[q] + (L.c_start[c + g.ext[2] + 1] - L.c_start[c + 1]);
        }
    }
    __syncthreads();
    return S;
}

// owned bead o -> staged index (brick cells are the region minus its halo)
__device__ __forceinline__ int owned_to_staged(const BrickGeom &g, const int *colpre,
                                               const int *c_start, int o, int &rx, int &ry)
{
    int q = 0;
    while (colpre[q + 1] <= o) ++q;
    rx = 1 + q / g.ext[1];
    ry = 1 + q % g.ext[1];
    return c_start[(rx * g.RY + ry) * g.RZ + 1] + (o - colpre[q]);
}

template <typename real> __device__ __forceinline__ real ratio(real a, real b, real binv);
template <> __device__ __forceinline__ float ratio<float>(float a, float b, float binv) { return a * binv; }
template <> __device__ __forceinline__ double ratio<double>(double a, double b, double binv) { return a / b; }

// ---------------------------------------------------------------------------
// mDPD pass 1: local density (force/pair_mdpd.cu:219-300).  Every pair j != i
// inside rd contributes (15/(2 pi rd^3)) (1 - r/rd)^2; there is no rc test and
// no self term.  Summed per bead in FP64 in a fixed order.
// ---------------------------------------------------------------------------

template <typename real>
__global__ __launch_bounds__(kBlock) void k_density(const RepDev *__restrict__ reps,
                                                    const WorkItem *__restrict__ items, int nitems,
                                                    BrickParams bp)
{
    extern __shared__ __align__(16) char smem[];
    const BrickLds<real> L = lds_bind<real>(smem, lds_offsets<real>(bp));
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int cur = R.state->cur;
    const BrickGeom g = brick_geom(R, it.local);
    const int S = stage_brick<real>(R, g, L, bp, cur, false, false);
    if (S < 0) {
        if (threadIdx.x == 0) atomicOr(&R.state->flags, kFlagStageOverflow);
        return;
    }
    const int O = L.colpre[g.ext[0] * g.ext[1]];
    const int nt = R.ntypes;
    const real rd2 = (real)R.rdmax2;
    for (int o = threadIdx.x; o < O; o += kBlock) {
        int rx, ry;
        const int li = owned_to_staged(g, L.colpre, L.c_start, o, rx, ry);
        const int rc = L.scell[li], rz = rc % g.RZ;
        const real xi = L.px[li], yi = L.py[li], zi = L.pz[li];
        const int ti = L.sty[li] - 1;
        double rho = 0.0;
        for (int ox = -1; ox <= 1; ++ox)
            for (int oy = -1; oy <= 1; ++oy) {
                const int base = ((rx + ox) * g.RY + (ry + oy)) * g.RZ + rz - 1;
                const int j1 = L.c_start[base + 3];
                for (int lj = L.c_start[base]; lj < j1; ++lj) {
                    const real dx = xi - L.px[lj], dy = yi - L.py[lj], dz = zi - L.pz[lj];
                    const real r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 < rd2 && lj != li) {
                        const PairCoef<real> &C = L.coef[ti * nt + (L.sty[lj] - 1)];
                        const real r = sqrt(r2);
                        real wd = (real)1 - ratio<real>(r, C.rd, C.rdinv);
                        wd = wd > (real)0 ? wd : (real)0;
                        rho += (double)(C.rhofac * wd * wd);
                    }
                }
            }
        const int gc = L.c_src[rc];
        const int slot = R.perm[R.cell_start[gc] + (li - L.c_start[rc])];
        R.rho[slot] = rho;
    }
}

// ---------------------------------------------------------------------------
// pair force (+ bonds) and, unless kModeForceOnly, the fused integration.
// Pair arithmetic follows force/pair_dpd.cu:296-399 and pair_mdpd.cu:403-517;
// bonds follow force/bond_harmonic.cu:196-268; integration fix/fix_nve.cu:48-86
// and :137-149; thermo sums compute/thermo.cu:243-245, :328, :424-433.
// ---------------------------------------------------------------------------

template <typename real, int STYLE>
__global__ __launch_bounds__(kBlock) void k_brick(const RepDev *__restrict__ reps,
                                                  const WorkItem *__restrict__ items, int nitems,
                                                  BrickParams bp, int mode)
{
    extern __shared__ __align__(16) char smem[];
    const BrickLds<real> L = lds_bind<real>(smem, lds_offsets<real>(bp));
    const WorkItem it = items[xcd_item(nitems)];
    const RepDev &R = reps[it.rep];
    const int cur = R.state->cur;
    const long long step = R.state->step;
    const bool log = (mode & kModeLogForce) ||
                     ((mode & kModeLogAuto) && R.log_every > 0 && (step % R.log_every) == 0);
    const bool force_only = (mode & kModeForceOnly) != 0;
    const BrickGeom g = brick_geom(R, it.local);
    const int S = stage_brick<real>(R, g, L, bp, cur, true, STYLE == 2);
    if (S < 0) {
        if (threadIdx.x == 0) atomicOr(&R.state->flags, kFlagStageOverflow);
        return;
    }
    const int t = threadIdx.x;
    const int N = R.N, nt = R.ntypes, kmax = max(bp.kmax, 1);
    const int O = L.colpre[g.ext[0] * g.ext[1]];
    const real rcut2 = (real)R.rcmax2;
    const uint32_t seed_step = R.seed + (uint32_t)step;
    const double *X = R.x[cur], *V = R.v[cur];
    double acc[kThermoParts];
#pragma unroll
    for (int k = 0; k < kThermoParts; ++k) acc[k] = 0.0;
    bool bad = false;

    for (int o = t; o < O; o += kBlock) {
        int rx, ry;
        const int li = owned_to_staged(g, L.colpre, L.c_start, o, rx, ry);
        const int rc = L.scell[li], rz = rc % g.RZ;
        const real xi = L.px[li], yi = L.py[li], zi = L.pz[li];

        const int idi = L.sid[li];
        const int ti = L.sty[li] - 1;
        const real vxi = L.vx[li], vyi = L.vy[li], vzi = L.vz[li];
        const real rhoi = (STYLE == 2) ? L.rh[li] : (real)0;
        double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0;
        double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0, w4 = 0.0, w5 = 0.0;
        // The 27 neighbour cells are 9 z-runs of 3 cells, contiguous in LDS.
        // Phase 1 appends in-range neighbours to this thread's list; when the
        // list is full, phase 2 evaluates it and phase 1 resumes at its cursor,
        // so the summation order is the candidate order whatever kmax is.
        int run = 0, lj = 0, j1 = 0;
        {
            const int base = ((rx - 1) * g.RY + (ry - 1)) * g.RZ + rz - 1;
            lj = L.c_start[base];
            j1 = L.c_start[base + 3];
        }
        while (true) {
            int cnt = 0;
            while (run < 9) {
                if (lj >= j1) {
                    if (++run == 9) break;
                    const int base = ((rx - 1 + run / 3) * g.RY + (ry - 1 + run % 3)) * g.RZ + rz - 1;
                    lj = L.c_start[base];
                    j1 = L.c_start[base + 3];
                    continue;
                }
                if (cnt == kmax) break;
                const real dx = xi - L.px[lj], dy = yi - L.py[lj], dz = zi - L.pz[lj];
                const real r2 = dx * dx + dy * dy + dz * dz;
                if (r2 <= rcut2 && r2 >= (real)1.0e-20 && lj != li) L.list[(cnt++) * kBlock + t] = (unsigned short)lj;
                ++lj;
            }

            for (int k = 0; k < cnt; ++k) {
                const int lk = L.list[k * kBlock + t];
                const real dx = xi - L.px[lk], dy = yi - L.py[lk], dz = zi - L.pz[lk];
                const real r2 = dx * dx + dy * dy + dz * dz;
                const PairCoef<real> &C = L.coef[ti * nt + (L.sty[lk] - 1)];
                if (r2 > C.rc2) continue;
                const real r = sqrt(r2);
                const real rinv = (real)1 / r;
                const real dot = dx * (vxi - L.vx[lk]) + dy * (vyi - L.vy[lk]) + dz * (vzi - L.vz[lk]);
                const real wc = (real)1 - ratio<real>(r, C.rc, C.rcinv);
                const real wr = wc;
                const int idj = L.sid[lk];
                const uint32_t t1 = (uint32_t)min(idi, idj), t2 = (uint32_t)max(idi, idj);
                const real rnd = saru_uniform<real>(saru_bits(t1, t2, seed_step));
                real wd = (real)0;
                real fpair;
                if (STYLE == 2) {
                    wd = (real)1 - ratio<real>(r, C.rd, C.rdinv);
                    wd = wd > (real)0 ? wd : (real)0;
                    fpair = C.A * wc + C.B * (rhoi + L.rh[lk]) * wd;
                } else {
                    fpair = C.A * wc;
                }
                fpair -= C.gamma * wr * wr * dot * rinv;
                fpair += C.sig * wr * rnd;
                fpair *= rinv;
                const real ffx = dx * fpair, ffy = dy * fpair, ffz = dz * fpair;
                fx += (double)ffx;
                fy += (double)ffy;
                fz += (double)ffz;
                if (log) {
                    if (STYLE == 2) {
                        const real u_c = (real)0.5 * C.A * C.rc * wc * wc;
                        const real u_d = (real)0.25 * C.B * C.rd * rhoi * wd * wd;
                        pe += (double)((real)0.5 * u_c + u_d);
                    } else {
                        const real ppe = (real)0.5 * C.A * C.rc * wc * wc;
                        pe += (double)((real)0.5 * ppe);
                    }
                    w0 += (double)((real)0.5 * dx * ffx);
                    w1 += (double)((real)0.5 * dx * ffy);
                    w2 += (double)((real)0.5 * dx * ffz);
                    w3 += (double)((real)0.5 * dy * ffy);
                    w4 += (double)((real)0.5 * dy * ffz);
                    w5 += (double)((real)0.5 * dz * ffz);
                }
            }
            if (run == 9) break;
        }

        // phase 3: bonds, then store or integrate
        const int gc = L.c_src[rc];
        int slot = R.perm[R.cell_start[gc] + (li - L.c_start[rc])];
        if ((unsigned)slot >= (unsigned)N) {
            bad = true;
            continue;
        }
        double xg = X[slot], yg = X[N + slot], zg = X[2 * N + slot];
        double bpe = 0.0;
        // bonds belong to Bond::compute_force on the hook path (bond_harmonic.cu:355-402)
        if (R.nbonds > 0 && !force_only) {
            const int b1 = R.bond_start[idi];
            for (int bb = R.bond_start[idi - 1]; bb < b1; ++bb) {
                const int ps = R.slot_of_id[R.bond_partner[bb] - 1];
                const double ddx = min_image(xg - X[ps], R.L[0], R.hL[0]);
                const double ddy = min_image(yg - X[N + ps], R.L[1], R.hL[1]);
                const double ddz = min_image(zg - X[2 * N + ps], R.L[2], R.hL[2]);
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
                    w0 += 0.5 * ddx * bfx;
                    w1 += 0.5 * ddx * bfy;
                    w2 += 0.5 * ddx * bfz;
                    w3 += 0.5 * ddy * bfy;
                    w4 += 0.5 * ddy * bfz;
                    w5 += 0.5 * ddz * bfz;
                }
            }
        }
        if (force_only || (mode & kModeStoreForce)) {
            R.force[slot] = fx;
            R.force[N + slot] = fy;
            R.force[2 * N + slot] = fz;
        }
        if (log) {
            R.pe[slot] = pe;
            R.bpe[slot] = bpe;
            R.vir[slot] = w0;
            R.vir[N + slot] = w1;
            R.vir[2 * N + slot] = w2;
            R.vir[3 * N + slot] = w3;
            R.vir[4 * N + slot] = w4;
            R.vir[5 * N + slot] = w5;
        }
        if (force_only) continue;

        const int tt = L.sty[li] - 1;
        const double h = (0.5 * R.dt) * R.minv[tt];
        double vxg = V[slot], vyg = V[N + slot], vzg = V[2 * N + slot];
        // explicit FMAs: the two half-kicks round exactly as two separate
        // nve kernels do (fix/fix_nve.cu:54-56, :143-145)
        if (mode & kModeFinalKick) {
            vxg = fma(h, fx, vxg);
            vyg = fma(h, fy, vyg);
            vzg = fma(h, fz, vzg);
        }
        if (log) {
            const double m = R.mass[tt];
            acc[0] += 0.5 * m * (vxg * vxg + vyg * vyg + vzg * vzg);
            acc[1] += pe;
            acc[2] += bpe;
            acc[3] += m * vxg * vxg + w0;
            acc[4] += m * vxg * vyg + w1;
            acc[5] += m * vxg * vzg + w2;
            acc[6] += m * vyg * vyg + w3;
            acc[7] += m * vyg * vzg + w4;
            acc[8] += m * vzg * vzg + w5;
        }
        if (mode & kModeDrift) {
            vxg = fma(h, fx, vxg);
            vyg = fma(h, fy, vyg);
            vzg = fma(h, fz, vzg);
            int ix, iy, iz;
            unpack_img(R.img[slot], ix, iy, iz);
            const int ix0 = ix, iy0 = iy, iz0 = iz;
            bool b2 = false;
            xg = wrap_axis(fma(R.dt, vxg, xg), R.lo[0], R.lo[0] + R.L[0], R.L[0], ix, b2);
            yg = wrap_axis(fma(R.dt, vyg, yg), R.lo[1], R.lo[1] + R.L[1], R.L[1], iy, b2);
            zg = wrap_axis(fma(R.dt, vzg, zg), R.lo[2], R.lo[2] + R.L[2], R.L[2], iz, b2);
            if (ix != ix0 || iy != iy0 || iz != iz0) R.img[slot] = pack_img(ix, iy, iz);
            if (b2 || !isfinite(vxg + vyg + vzg)) {
                bad = true;
                xg = R.lo[0];
                yg = R.lo[1];
                zg = R.lo[2];
            }
            const int key = cell_of(R, xg, yg, zg);
            R.key[slot] = key;
            R.rank[slot] = atomicAdd(&R.cell_count[key], 1);
        }
        double *Xo = R.x[1 - cur], *Vo = R.v[1 - cur];
        Xo[slot] = xg;
        Xo[N + slot] = yg;
        Xo[2 * N + slot] = zg;
        Vo[slot] = vxg;
        Vo[N + slot] = vyg;
        Vo[2 * N + slot] = vzg;
    }
    if (bad) atomicOr(&R.state->flags, kFlagNonFinite);
    if (log && !force_only) {
        double *out = L.red + 48;
        block_sum<kThermoParts>(acc, L.red, out);
        if (t < kThermoParts) R.thermo_part[(size_t)it.local * kThermoParts + t] = out[t];
    }
}

// ---------------------------------------------------------------------------
// end of step, per replica: thermo record on log steps (fixed-order sum of the
// brick partials), then flip the buffer parity and advance the step counter.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_step_end(const RepDev *__restrict__ reps,
                                                     const int *__restrict__ rep_list, int flip,
                                                     int advance, int thermo)
{
    __shared__ double red[16 * (kBlock / 64)];
    __shared__ double tot[kThermoParts];
    const RepDev &R = reps[rep_list[blockIdx.x]];
    const long long step = R.state->step;
    if (thermo && R.log_every > 0 && (step % R.log_every) == 0) {
        double acc[kThermoParts];
#pragma unroll
        for (int k = 0; k < kThermoParts; ++k) acc[k] = 0.0;
        for (int b = threadIdx.x; b < R.nbrick; b += kBlock)
#pragma unroll
            for (int k = 0; k < kThermoParts; ++k) acc[k] += R.thermo_part[(size_t)b * kThermoParts + k];
        block_sum<kThermoParts>(acc, red, tot);
        if (threadIdx.x == 0) {
            const double vol = R.L[0] * R.L[1] * R.L[2];
            const int row = R.state->log_rows % R.thermo_cap;
            double *o = R.thermo_rows + (size_t)row * 12;
            const double ndof = 3.0 * R.N - 3.0;
            o[0] = (2.0 * tot[0]) / (ndof * 1.0);
            o[1] = tot[1];
            o[2] = tot[2];
            for (int k = 0; k < 6; ++k) o[4 + k] = tot[3 + k] / vol;
            o[3] = (o[4] + o[7] + o[9]) / 3.0;
            o[10] = tot[0];
            o[11] = (double)step;
            R.state->log_rows += 1;
        }
    }
    if (threadIdx.x == 0) {
        R.state->step = step + advance;
        R.state->cur ^= flip;
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
    if (s >= R.N) return;
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
        R.img[s] = pack_img(im[0], im[1], im[2]);
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
    if (s >= R.N) return;
    const int N = R.N;
    const double *X = R.x[R.state->cur];
    const int id = R.id[s];
    double f[3] = {0, 0, 0}, w[6] = {0, 0, 0, 0, 0, 0}, bpe = 0.0;
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
    for (int k = 0; k < 3; ++k) R.force[k * N + s] += f[k];
    if (log) {
        R.bpe[s] = bpe;
        for (int k = 0; k < 6; ++k) R.vir[k * N + s] += w[k];
    }
}

__global__ __launch_bounds__(kBlock) void k_thermo_arrays(const RepDev *__restrict__ reps,
                                                          const WorkItem *__restrict__ items,
                                                          double *__restrict__ partials)
{
    __shared__ double red[16 * (kBlock / 64)];
    __shared__ double out[kThermoParts];
    const WorkItem it = items[blockIdx.x];
    const RepDev &R = reps[it.rep];
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
        acc[3] = m * vx * vx + R.vir[s];
        acc[4] = m * vx * vy + R.vir[N + s];
        acc[5] = m * vx * vz + R.vir[2 * N + s];
        acc[6] = m * vy * vy + R.vir[3 * N + s];
        acc[7] = m * vy * vz + R.vir[4 * N + s];
        acc[8] = m * vz * vz + R.vir[5 * N + s];
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
    (void)style;
    BrickParams p{rmax, smax, kmax, ncol, ntypes};
    return precision == 1 ? lds_offsets<double>(p).total : lds_offsets<float>(p).total;
}

void launch_bin_keys(const RepDev *reps, const WorkItem *items, int nitems, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_bin_keys, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems);
}
void launch_scatter(const RepDev *reps, const WorkItem *items, int nitems, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_scatter, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems);
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
void launch_resort(const RepDev *reps, const WorkItem *items, int nitems, hipStream_t s)
{
    if (!nitems) return;
    hipLaunchKernelGGL(k_resort_gather, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems);
    hipLaunchKernelGGL(k_resort_finish, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems);
}
void launch_density(const RepDev *reps, const WorkItem *items, int nitems, const BrickLaunch &bl,
                    int precision, hipStream_t s)
{
    if (!nitems) return;
    BrickParams p{bl.rmax, bl.smax, bl.kmax, bl.ncol, bl.ntypes};
    const size_t lds = bl.lds_bytes(precision, 2);
    if (precision == 1)
        hipLaunchKernelGGL(k_density<double>, dim3(nitems), dim3(kBlock), lds, s, reps, items, nitems, p);
    else
        hipLaunchKernelGGL(k_density<float>, dim3(nitems), dim3(kBlock), lds, s, reps, items, nitems, p);
}
void launch_brick(const RepDev *reps, const WorkItem *items, int nitems, const BrickLaunch &bl,
                  int precision, int style, int mode, hipStream_t s)
{
    if (!nitems) return;
    BrickParams p{bl.rmax, bl.smax, bl.kmax, bl.ncol, bl.ntypes};
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
void launch_step_end(const RepDev *reps, const int *rep_list, int nreps, int flip, int advance,
                     int thermo, hipStream_t s)
{
    if (nreps) hipLaunchKernelGGL(k_step_end, dim3(nreps), dim3(kBlock), 0, s, reps, rep_list, flip, advance, thermo);
}
void launch_nve(const RepDev *reps, const WorkItem *items, int nitems, int initial, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_nve, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, initial);
}
void launch_bond_hook(const RepDev *reps, const WorkItem *items, int nitems, int log, hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_bond_hook, dim3(nitems), dim3(kBlock), 0, s, reps, items, nitems, log);
}
void launch_thermo_arrays(const RepDev *reps, const WorkItem *items, int nitems, double *partials,
                          hipStream_t s)
{
    if (nitems) hipLaunchKernelGGL(k_thermo_arrays, dim3(nitems), dim3(kBlock), 0, s, reps, items, partials);
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
void launch_fill_i32(int *p, int v, long long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(kBlock), 0, s, p, v, n);
}
void launch_fill_f64(double *p, double v, long long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3(nblk(n)), dim3(kBlock), 0, s, p, v, n);
}

} // namespace gdpd
