/*
 * gdpd.h -- C ABI of the MI355X-native DPD hot path (libgdpd.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, integer status codes.  No
 * torch or HIP types appear in any signature.  Host arrays passed in or out are
 * indexed like the reference's host vectors: bead index = id - 1, positions,
 * velocities and forces AoS [i*3 + k], virial [i*6 + {xx,xy,xz,yy,yz,zz}]
 * (system/read_data.cu:265-283, system/system.cuh:6-49).
 *
 * Every entry point returns GDPD_OK (0) or a GDPD_ERR_* code; gdpd_last_error()
 * gives the message.  Nothing calls exit().  A context owns one HIP stream on one
 * device; calls on one context are not thread-safe.
 *
 * Each entry point names the reference interface it replaces (file:line under
 * the reference's src/).
 */
#ifndef GDPD_H
#define GDPD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GDPD_OK 0
#define GDPD_ERR_ARG 1      /* bad argument (the reference's print_error_and_exit cases) */
#define GDPD_ERR_HIP 2      /* HIP runtime error */
#define GDPD_ERR_OVERFLOW 3 /* a capacity bound (staged beads, list length, cell count) was exceeded */
#define GDPD_ERR_STATE 4    /* call out of order (e.g. step before upload) */
#define GDPD_ERR_NOMEM 5    /* device allocation or HBM budget refused */
#define GDPD_ERR_NAN 6      /* non-finite position or velocity detected */

#define GDPD_PREC_F32 0 /* FP32 pair arithmetic, FP64 accumulators and state */
#define GDPD_PREC_F64 1 /* FP64 everywhere */

#define GDPD_STYLE_NONE 0
#define GDPD_STYLE_DPD 1
#define GDPD_STYLE_MDPD 2

/* thermo record layout (compute/thermo.cu:122-214 columns, plus KE, step and angle PE) */
#define GDPD_THERMO_T 0
#define GDPD_THERMO_PAIR_PE 1
#define GDPD_THERMO_BOND_PE 2
#define GDPD_THERMO_P 3
#define GDPD_THERMO_PXX 4 /* 4..9: Pxx Pxy Pxz Pyy Pyz Pzz */
#define GDPD_THERMO_KE 10
#define GDPD_THERMO_STEP 11
#define GDPD_THERMO_ANGLE_PE 12
#define GDPD_THERMO_LEN 13

typedef struct gdpd_ctx gdpd_ctx;

/* One simulated system ("replica").  Mirrors the reference's System, Box and the
   `neighbor`, `timestep` and `thermo` commands (main/common.cuh:75-130,
   main/read_in.cu:93-169, system/system.cuh:53-59). */
typedef struct {
    int n_beads;
    int n_types;
    double lo[3], hi[3];
    double dt;
    double global_cut; /* neighbor global_cut */
    double skin;       /* neighbor skin */
    double max_rho;    /* neighbor max_rho: density bound used to size buffers */
    int check_every;   /* neighbor check_every */
    int log_every;     /* thermo N */
    int precision;     /* GDPD_PREC_* */
} gdpd_system_desc;

/* ---- context -------------------------------------------------------------- */
/* replaces print_header's cudaSetDevice(0) (main/utils.cuh:52-53) */
int gdpd_create(gdpd_ctx **out, int device);
int gdpd_destroy(gdpd_ctx *ctx);
const char *gdpd_last_error(const gdpd_ctx *ctx);
/* bytes of device memory a system of this description would take (admission) */
int64_t gdpd_system_bytes(const gdpd_system_desc *desc, int n_bonds);

/* ---- system state --------------------------------------------------------- */
/* replaces atoms_mem_alloc + atoms_mem_copy_2_gpu (system/system.cu:5-86) */
int gdpd_sys_create(gdpd_ctx *ctx, const gdpd_system_desc *desc, int *sys_out);
int gdpd_sys_destroy(gdpd_ctx *ctx, int sys);
/* id[k] in 1..N (any order), type 1-based, pos/vel AoS; positions are wrapped into
   the box like Read_data::adjust_coordinate (system/read_data.cu:191-197,274-276).
   masses: n_types values (read_data.cu:114-140). vel may be NULL (zero). */
int gdpd_upload(gdpd_ctx *ctx, int sys, const int *id, const int *type, const double *pos,
                const double *vel, const double *masses);
/* id-ordered copies out; any pointer may be NULL.  image: 3 ints per bead
   (unwrapped = pos + image * L).  force is the last computed force. */
int gdpd_download(gdpd_ctx *ctx, int sys, double *pos, double *vel, int *image, double *force);
/* per-bead pair PE, virial (pair + bond) and mDPD density of the last log step */
int gdpd_download_pe_virial(gdpd_ctx *ctx, int sys, double *pe, double *virial, double *rho);

/* ---- force fields ---------------------------------------------------------- */
/* replaces Pair_dpd::isEnabled (force/pair_dpd.cu:19-173): `pair_style dpd T rc seed`
   + `pair_coeff`.  coeff: 3 doubles {A, gamma, rc} per packed pair index
   ((ti-1)*nt - (ti-1)(ti-2)/2 + tj - ti, ti <= tj; pair_dpd.cu:134). */
int gdpd_set_pair_dpd(gdpd_ctx *ctx, int sys, double temp, double rc, uint32_t seed,
                      const double *coeff);
/* replaces Pair_mdpd::isEnabled (force/pair_mdpd.cu:21-192): `pair_style mdpd T rc rd
   seed`; coeff: 5 doubles {A, B, gamma, rc, rd} per packed pair index. */
int gdpd_set_pair_mdpd(gdpd_ctx *ctx, int sys, double temp, double rc, double rd, uint32_t seed,
                       const double *coeff);
/* replaces Bond_harmonic::isEnabled + bonds_mem_copy_2_gpu (force/bond_harmonic.cu:19-147,
   system/system.cu:123-139): bondlist 3 ints {type, id_i, id_j} per bond;
   bond_coeff 2 doubles {k, r0} per type, E = k (r - r0)^2. */
int gdpd_set_bonds(gdpd_ctx *ctx, int sys, int n_bonds, int n_bond_types, const int *bondlist,
                   const double *bond_coeff);

/* replaces Angle_harmonic::isEnabled + angles_mem_copy_2_gpu (force/angle_harmonic.cu:19-151,
   system/system.cu:180-215): anglelist 4 ints per angle {type, id_i, id_j, id_k}, j the vertex;
   angle_coeff 2 doubles per type {k, theta0 in degrees}, E = k (theta - theta0)^2. */
int gdpd_set_angles(gdpd_ctx *ctx, int sys, int n_angles, int n_angle_types, const int *anglelist,
                    const double *angle_coeff);

/* ---- plugin-hook path (one reference hook per call) ----------------------- */
/* Neigh_list::build (force/neigh_list.cu:397-479): half Verlet list, bins of
   global_cut + skin, r^2 <= (global_cut+skin)^2, i < j by index. */
int gdpd_neigh_build(gdpd_ctx *ctx, int sys);
/* Neigh_list::check_update (force/neigh_list.cu:518-551): *flag = 1 when any bead
   moved >= skin/2 since the last build. */
int gdpd_neigh_check(gdpd_ctx *ctx, int sys, int *flag);
/* copies the last built list out, rows indexed by id-1 and sorted by neighbour
   index; n_max is the row stride of neigh.  Returns GDPD_ERR_OVERFLOW when a row
   did not fit the device list (the reference writes out of bounds there,
   neigh_list.cu:384). */
int gdpd_neigh_download(gdpd_ctx *ctx, int sys, int *n_neigh, int *neigh, int n_max);
/* Pair::compute_force (force/pair.cu:77-83 -> pair_dpd.cu:564-633 or
   pair_mdpd.cu:689-779): zeroes the force (and pe/virial when log) and adds the
   pair forces at the current positions for Saru step `step`. */
int gdpd_pair_compute(gdpd_ctx *ctx, int sys, uint32_t step, int log);
/* Bond::compute_force (force/bond_harmonic.cu:355-402): adds bond forces. */
int gdpd_bond_compute(gdpd_ctx *ctx, int sys, uint32_t step, int log);
/* Angle::compute_force (force/angle_harmonic.cu:459-...; kernels :168-458): adds angle forces. */
int gdpd_angle_compute(gdpd_ctx *ctx, int sys, uint32_t step, int log);
/* nve::initial_integrate / final_integrate (fix/fix_nve.cu:93-118, 156-177) */
int gdpd_nve_initial(gdpd_ctx *ctx, int sys);
int gdpd_nve_final(gdpd_ctx *ctx, int sys);
/* Thermo::process reductions (compute/thermo.cu:96-220) on the current state:
   out gets GDPD_THERMO_LEN doubles (step slot = -1). */
int gdpd_thermo(gdpd_ctx *ctx, int sys, double *out);
/* gdpd_thermo without any device or host allocation.  No counterpart in the
   reference: for callers that cannot allocate (serving loops, harnesses that
   forbid hipMalloc).  Its partial sums live in a fixed 4 MiB device array per
   device, so it takes systems of at most 13,421,568 beads (larger ones get
   GDPD_ERR_ARG) that have computed once (GDPD_ERR_STATE before that).  The
   output equals gdpd_thermo's bit for bit.  Calls that overlap on one device
   share the array and race. */
int gdpd_thermo_noalloc(gdpd_ctx *ctx, int sys, double *out);

/* ---- fused step path -------------------------------------------------------- */
/* One `run N` (main/utils.cuh:166-314): zero_step at step 0, then steps 1..N, each
   one fused launch of force (+bonds) + final half-kick + thermo on log steps +
   next half-kick and drift.  Thermo records of every log step are kept on the
   device; read them with gdpd_thermo_rows. */
int gdpd_run(gdpd_ctx *ctx, int sys, int nsteps);
/* The same split in three for timing: zero_step leaves the system between
   steps (after step 1's drift); advance runs n whole steps; finish does the
   last step's force and final half-kick.
   zero_step and advance only enqueue work.  Where the reference sizes its
   neighbour list once from max_rho (force/neigh_list.cu:14-17), the bricks'
   LDS capacities follow the state: a brick that outgrows them holds its system
   at the last completed step (the rest of a batch goes on), and the next call
   that waits for the device (advance, batch_advance, finish, download,
   thermo_rows, get_step, sync, step_times) sizes the capacities again and runs
   the held system's missing steps.  The result equals an uninterrupted run bit
   for bit while the brick shape still fits in LDS (else the brick shrinks and
   the run continues with FP32-rounding differences only). */
int gdpd_zero_step(gdpd_ctx *ctx, int sys);
int gdpd_advance(gdpd_ctx *ctx, int sys, int nsteps);
int gdpd_finish(gdpd_ctx *ctx, int sys);
/* thermo records (GDPD_THERMO_LEN doubles each) of the current run, oldest first */
int gdpd_thermo_rows(gdpd_ctx *ctx, int sys, double *out, int max_rows, int *n_rows);
/* step counter of the next force evaluation */
int gdpd_get_step(gdpd_ctx *ctx, int sys, int64_t *step);

/* ---- group fixes and computes ------------------------------------------------- */
/* A group is a set of bead types (system/system.cu:235-301 selects groups by type
   only), passed as a bitmask: bit (t - 1) set for type t.  Fixes act in the order
   given at the reference's hook points (main/utils.cuh:251-313): initial_integrate
   (nve, move), post_integrate (wall_reflect, two_wall), post_force (spring),
   final_integrate (nve) and end_of_step (rvcm, recenter).  Without any
   GDPD_FIX_NVE entry every bead is integrated by nve, as before. */
#define GDPD_FIX_NVE 1          /* fix ID grp nve                          fix/fix_nve.cu:93-177 */
#define GDPD_FIX_MOVE 2         /* fix ID grp move vx vy vz                fix/fix_move.cu:45-127 */
#define GDPD_FIX_WALL_REFLECT 3 /* fix ID grp wall_reflect lo hi dir       fix/fix_wall_reflect.cu:79-146 */
#define GDPD_FIX_TWO_WALL 4     /* fix ID grp two_wall lo_grp hi_grp dir   fix/fix_two_wall.cu:79-204 */
#define GDPD_FIX_SPRING 5       /* fix ID grp spring k x y z r0            fix/fix_spring.cu:210-376 */
#define GDPD_FIX_RECENTER 6     /* fix ID grp recenter every x y z         fix/fix_recenter.cu:160-340 */
#define GDPD_FIX_RVCM 7         /* fix ID grp rvcm every fx fy fz          fix/fix_rvcm.cu:160-292 */
#define GDPD_MAX_FIXES 8

typedef struct {
    int kind;         /* GDPD_FIX_* */
    uint32_t group;   /* the fix's group */
    uint32_t group2;  /* two_wall: the lower wall's group */
    uint32_t group3;  /* two_wall: the upper wall's group */
    int every;        /* recenter, rvcm: act when step % every == 0 */
    int dir;          /* wall_reflect, two_wall: 0 x, 1 y, 2 z */
    int flags;        /* rvcm: bit 0 x, bit 1 y, bit 2 z */
    double p[5];      /* move {vx,vy,vz}; wall_reflect {lo,hi}; spring {k,x,y,z,r0};
                         recenter {x,y,z} */
} gdpd_fix;

/* replaces preprocessFixes / enableFix (fix/fix.cu:40-120) for the hot-path fixes */
int gdpd_set_fixes(gdpd_ctx *ctx, int sys, const gdpd_fix *fixes, int n_fixes);

#define GDPD_HOOK_INITIAL_INTEGRATE 0
#define GDPD_HOOK_POST_INTEGRATE 1
#define GDPD_HOOK_POST_FORCE 2
#define GDPD_HOOK_FINAL_INTEGRATE 3
#define GDPD_HOOK_END_OF_STEP 4
/* One fix's hook method on the current state (Fix::initial_integrate ... end_of_step,
   fix/fix.cuh:20-25).  step selects the every-N fixes.  out (may be NULL) gets the
   spring's {fx, fy, fz, signed |f|, energy} of that call (fix_spring.cu:330-350). */
int gdpd_fix_hook(gdpd_ctx *ctx, int sys, int index, int hook, uint32_t step, double *out5);

#define GDPD_COMPUTE_TEMP 1         /* compute ID grp temp every file             compute/compute_temp.cu */
#define GDPD_COMPUTE_MSD 2          /* compute ID grp msd every file              compute/compute_msd.cu */
#define GDPD_COMPUTE_RHO 3          /* compute ID grp rho every xlo xhi ylo yhi zlo zhi file */
#define GDPD_COMPUTE_TF 4           /* compute ID grp tf every file               compute/compute_tf.cu */
#define GDPD_COMPUTE_WALL_CONTACT 5 /* compute ID grp wall_contact every wgrp lo|hi dir cutoff file */
#define GDPD_MAX_COMPUTES 8

typedef struct {
    int kind;         /* GDPD_COMPUTE_* */
    uint32_t group;
    uint32_t group2;  /* wall_contact: the wall's group */
    int every;
    int dir;          /* wall_contact axis */
    int side;         /* wall_contact: +1 lo wall, -1 hi wall */
    double p[6];      /* rho {xlo,xhi,ylo,yhi,zlo,zhi}; wall_contact {cutoff} */
} gdpd_compute;

/* replaces preprocessCompute / enableCompute (compute/compute.cu:30-110).  msd takes
   its reference positions (unwrapped) from the state at this call. */
int gdpd_set_computes(gdpd_ctx *ctx, int sys, const gdpd_compute *computes, int n_computes);
/* Compute::compute on the current state: out gets up to 3 values (temp; msd; rho;
   tf x y z; wall_contact count), like the reference's file row for that step. */
int gdpd_compute_eval(gdpd_ctx *ctx, int sys, int index, double *out3);
/* Records of the fused run (gdpd_run / zero_step / advance / finish), oldest first:
   GDPD_RECORD_LEN doubles each, {step, values...}.  index < GDPD_MAX_COMPUTES names
   a compute; GDPD_MAX_COMPUTES + k names fix k (spring: {step, fx, fy, fz, f, e},
   one record per step).  A ring of GDPD_RECORD_CAP records per index. */
#define GDPD_RECORD_LEN 6
#define GDPD_RECORD_CAP 1024
int gdpd_records(gdpd_ctx *ctx, int sys, int index, double *out, int max_rows, int *n_rows);

/* ---- batched replicas (no reference counterpart: the reference runs one system) */
/* Advances every listed system by nsteps whole steps, all systems' bricks in the
   same launches (one launch per phase per step, graph-captured). */
int gdpd_batch_advance(gdpd_ctx *ctx, const int *systems, int n_systems, int nsteps);

/* ---- tuning and measurement (no reference counterpart) ---------------------- */
/* brick edge in cells per axis (0 = automatic), how many steps between
   physical re-sorts of the bead arrays into cell order (0 = automatic, 8) and
   how many steps one captured graph holds (0 = automatic, 8). */
int gdpd_set_tuning(gdpd_ctx *ctx, int sys, int bx, int by, int bz, int resort_every,
                    int graph_steps);
/* out8: brick edges (3), LDS capacities smax and omax, re-sort interval, graph
   length, brick count */
int gdpd_get_tuning(gdpd_ctx *ctx, int sys, int *out8);
/* LDS capacities of the bricks, staged and owned beads per brick (0 = sized
   from the state with a margin).  The reference bounds its per-bead neighbour
   storage by max_rho instead (force/neigh_list.cu:14-17).  A brick that
   outgrows the capacities holds its system and the capacities are sized again
   from the state (see gdpd_advance). */
int gdpd_set_capacity(gdpd_ctx *ctx, int sys, int smax, int omax);
/* Kernel timing: when on, every launch of the fused force kernel is bracketed by
   HIP events on the context stream; get returns the summed milliseconds and the
   launch count since the last reset. */
int gdpd_kernel_timing(gdpd_ctx *ctx, int on);
int gdpd_kernel_timing_get(gdpd_ctx *ctx, double *force_ms, int64_t *force_launches,
                           double *total_ms);
/* Phases of the timed force launches (no counterpart in the reference; a
   measurement aid like gdpd_kernel_timing).  While kernel timing is on, each
   wave of the pair-force kernel reads the device clock at its phase
   boundaries; us4 receives the mean microseconds per wave and launch of:
   staging the brick into LDS, the wave's candidate search and pair
   evaluation, its wait at the closing barrier, and the write-back of forces
   and thermo partial sums.  Reset by gdpd_kernel_timing. */
int gdpd_kernel_phases(gdpd_ctx *ctx, double *us4);
/* waits for the context's stream, then settles held systems (see gdpd_advance) */
int gdpd_sync(gdpd_ctx *ctx);
/* Admission window, for serving loops that admit new systems while others run.
   No counterpart in the reference (it runs one system).  Between
   gdpd_admission(ctx, 1) and gdpd_admission(ctx, 0) the calls that create and
   set up new systems (gdpd_sys_create, gdpd_set_*, gdpd_upload, gdpd_zero_step)
   run on a second stream of the context and do not wait for work already
   queued on its main stream; only systems created inside the window can be
   used there.  Closing the window makes all later main-stream work wait for
   the window's work, without blocking the host.  Results are the same as
   outside a window. */
int gdpd_admission(gdpd_ctx *ctx, int on);
/* Step-completion times of the listed systems, for per-step latency: times gets
   GDPD_STEP_TS_CAP device-clock readings per system, the one of step s at
   [k * GDPD_STEP_TS_CAP + s % GDPD_STEP_TS_CAP], written when the step's last
   kernel ran; steps[k] gets the system's current step counter; tick_ns (may be
   NULL) the clock period.  Read at least every GDPD_STEP_TS_CAP - 1 steps. */
#define GDPD_STEP_TS_CAP 256
int gdpd_step_times(gdpd_ctx *ctx, const int *systems, int n_systems, uint64_t *times, int64_t *steps,
                    double *tick_ns);

/* Evaluates the device Saru hash (the pair-force kernels' own inline function)
   for n host tuples: bits, the FP32-build uniform and the FP64-build uniform.
   Replaces nothing (the reference's saru() is device-inline, force/pair_dpd.cu:212);
   exposed so the hash can be checked bit for bit.  Any output may be NULL. */
int gdpd_saru_eval(gdpd_ctx *ctx, int n, const uint32_t *tag1, const uint32_t *tag2, uint32_t seed,
                   const int32_t *steps, uint32_t *bits, float *r32, double *r64);

#ifdef __cplusplus
}
#endif
#endif
