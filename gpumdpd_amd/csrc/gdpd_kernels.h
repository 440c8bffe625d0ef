// gdpd_kernels.h -- launch wrappers for the gfx950 kernels in gdpd_kernels.hip.
#pragma once
#include "gdpd_internal.h"

namespace gdpd {

struct BrickLaunch {
    int rmax;      // max region cells ((bx+2)(by+2)(bz+2)) over the launch
    int smax;      // max staged beads
    int omax;      // max owned beads of a brick
    int nrun;      // max run-list entries of a brick (brick_run_entries)
    int ntypes;    // max types
    int qw = 192;  // pair-queue entries per wave: a flush leaves < 64, one candidate step adds <= 128
    int perbead = 0; // per-bead PE/virial sums (hook path log steps)
    unsigned *phase = nullptr;  // per item and wave: 4 phase-clock sums (kernel timing)
    int single_rep = -1;        // >= 0: the launch's items are bricks 0.. of this replica, in order
    size_t lds_bytes(int precision, int style) const;
};

// elementwise over slots (items: rep, chunk of kBlock slots)
void launch_bin_keys(const RepDev *reps, const WorkItem *items, int nitems, hipStream_t s);
void launch_scatter(const RepDev *reps, const WorkItem *items, int nitems, hipStream_t s);
// cells (items: rep, chunk of kBlock cells)
void launch_cellsort(const RepDev *reps, const WorkItem *items, int nitems, hipStream_t s);
// scan over cell counts: items of 4*kBlock cells; chunk_off[r] = first chunk of rep r
void launch_scan(const RepDev *reps, const WorkItem *items, int nitems, const int *rep_chunk_off,
                 const int *rep_list, int nreps, int *chunk_sums, hipStream_t s);
// physical re-sort of the listed replicas; the host then swaps its copies of
// the id / id_alt, type / type_alt and img / img_alt pointers (resort_tables)
void launch_resort(RepDev *reps, const WorkItem *items, int nitems, const int *rep_list, int nreps, hipStream_t s);
// bricks (items: rep, brick)
void launch_density(const RepDev *reps, const WorkItem *items, int nitems, const BrickLaunch &bl,
                    int precision, hipStream_t s);
void launch_brick(const RepDev *reps, const WorkItem *items, int nitems, const BrickLaunch &bl,
                  int precision, int style, int mode, hipStream_t s);
// bonds + integration + re-binning, one thread per slot; topo: 0 no replica has
// bonds or angles, 1 bonds only, 2 angles too (selects the kernel instance)
void launch_integrate(const RepDev *reps, const WorkItem *items, int nitems, int mode, int topo, hipStream_t s);
// per replica (rep_list: nreps replica indices)
void launch_step_end(const RepDev *reps, const int *rep_list, int nreps, int flip, int advance,
                     int thermo, hipStream_t s, int bricks = 1);
// scalar states and step-time rings of the listed replicas
void launch_gather_state(const RepDev *reps, const int *rep_list, int nreps, RepState *out, hipStream_t s);
void launch_gather_step_ts(const RepDev *reps, const int *rep_list, int nreps, unsigned long long *ts,
                           long long *steps, hipStream_t s);
// group fixes and computes: fmask / cmask select fixes / computes by index
void launch_fix_clear(const RepDev *reps, const int *rep_list, int nreps, hipStream_t s);
void launch_fix_initial(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, hipStream_t s);
void launch_fix_two_wall(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, hipStream_t s);
void launch_fix_reduce(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, unsigned cmask,
                       int phase, hipStream_t s);
void launch_fix_post_force(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, hipStream_t s);
void launch_fix_final(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, hipStream_t s);
void launch_fix_end(const RepDev *reps, const WorkItem *items, int nitems, unsigned fmask, hipStream_t s);
void launch_fix_records(const RepDev *reps, const int *rep_list, int nreps, unsigned fmask, unsigned cmask,
                        long long step, hipStream_t s);
// hook path
void launch_nve(const RepDev *reps, const WorkItem *items, int nitems, int initial, hipStream_t s);
void launch_bond_hook(const RepDev *reps, const WorkItem *items, int nitems, int log, hipStream_t s);
void launch_angle_hook(const RepDev *reps, const WorkItem *items, int nitems, int log, hipStream_t s);
void launch_thermo_arrays(const RepDev *reps, const WorkItem *items, int nitems, double *partials,
                          hipStream_t s);
// one-thread, chunk-ordered total of nchunk partials into out[0..kThermoParts)
void launch_thermo_total(const double *partials, int nchunk, double *out, hipStream_t s);
// reference neighbour list (single replica)
void launch_nl_bin(const RepDev *rep, int N, double bin, const int *nb, int *binof, int *count,
                   hipStream_t s);
void launch_nl_fill(const RepDev *rep, int N, const int *binof, const int *start, int *fillc,
                    int *content, hipStream_t s);
void launch_nl_build(const RepDev *rep, int N, double bin, const int *nb, const int *binof,
                     const int *start, const int *content, int n_max, int *n_neigh, int *neigh,
                     int *overflow, double *last_pos, hipStream_t s);
void launch_nl_check(const RepDev *rep, int N, double skin, const double *last_pos, int *flag,
                     hipStream_t s);
void launch_saru_eval(int n, const uint32_t *t1, const uint32_t *t2, uint32_t seed, const int32_t *steps,
                      uint32_t *bits, float *r32, double *r64, hipStream_t s);
void launch_fix_to_double(double *p, long long n, int precision, hipStream_t s);
// fill helpers (graph-safe replacements for hipMemsetAsync)
void launch_fill_i32(int *p, int v, long long n, hipStream_t s);
void launch_fill_f64(double *p, double v, long long n, hipStream_t s);

} // namespace gdpd
