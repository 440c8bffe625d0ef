/*
 * gdpd_oracle.h -- CPU oracle for the gpumdpd per-step DPD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-threaded (plus an
 * optional OpenMP force loop used as the CPU baseline) restatement of the
 * reference's arithmetic.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or the
 * reported CPU baseline.  The product library (libgdpd) never links it.
 *
 * Conventions follow the reference (force/pair_dpd.cu, force/pair_mdpd.cu, fix/fix_nve.cu, compute/thermo.cu):
 *   - bead arrays are indexed by (id - 1); positions/velocities/forces are AoS
 *     [i*3 + k]; virial is [i*6 + {xx,xy,xz,yy,yz,zz}];
 *   - types are 1-based; pair coefficients are packed upper-triangular with the
 *     index of force/pair_dpd.cu:134 (0-based here);
 *   - a "half" neighbour list holds j > i only (force/neigh_list.cu:370).
 *
 * Parity pins (what checks each function) are listed in the header comment of
 * gdpd_oracle.c next to each function.
 */
#ifndef GDPD_ORACLE_H
#define GDPD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double lo[3];
    double hi[3];
} orc_box;

/* ---- Saru hash (force/pair_dpd.cu:203-236, identical copy pair_mdpd.cu:309-342) ---- */
uint32_t orc_saru_u32(uint32_t seed1, uint32_t seed2, uint32_t seed, int32_t timestep);
float    orc_saru_rand_f32(uint32_t seed1, uint32_t seed2, uint32_t seed, int32_t timestep);
double   orc_saru_rand_f64(uint32_t seed1, uint32_t seed2, uint32_t seed, int32_t timestep);
void     orc_saru_batch(int n, const uint32_t *tag1, const uint32_t *tag2, uint32_t seed,
                        const int32_t *steps, uint32_t *s_out, float *r32_out, double *r64_out);
/* The hash's three stages: out[0..1] = the TEA block (st, ws) after the seed
   pre-mix (pair_dpd.cu:213-226), out[2..3] = the block after the 8 TEA cycles
   (:227-232), out[4] = the finished 32-bit value (:233-234) = orc_saru_u32. */
void     orc_saru_stages(uint32_t seed1, uint32_t seed2, uint32_t seed, int32_t timestep, uint32_t out[5]);

/* ---- packed pair-coefficient index (force/pair_dpd.cu:134, 0-based) ---- */
int orc_pair_index(int ti, int tj, int ntypes);

/* ---- neighbour list (force/neigh_list.cu) ---- */
int orc_nbins(double len, double bin_size);
/* returns the number of list entries dropped for lack of room (0 = no overflow);
   rows are sorted by j so the list is canonical */
int orc_neigh_build(int N, const double *pos, const orc_box *box, double bin_size, int n_max,
                    int *n_neigh, int *neigh);
int orc_neigh_brute(int N, const double *pos, const orc_box *box, double cut, int n_max,
                    int *n_neigh, int *neigh);
int orc_neigh_check(int N, const double *pos, const double *last_pos, const orc_box *box,
                    double skin);
int orc_n_max_neigh(double global_cut, double skin, double max_rho);

/* ---- DPD pair force (force/pair_dpd.cu:240-419, 564-633) ----
   coeff: 3 doubles per packed pair index {A, gamma, rc}.  Zeroes force (and pe,
   virial when non-NULL) first, like Pair_dpd::compute_force. */
void orc_dpd_force(int N, int ntypes, const int *id, const int *type, const double *pos,
                   const double *vel, const orc_box *box, const int *n_neigh, const int *neigh,
                   int n_max, const double *coeff, double temp, double dt, uint32_t seed,
                   uint32_t step, double *force, double *pe, double *virial);

/* ---- mDPD (force/pair_mdpd.cu:219-300 density; 346-537 force; 689-779 driver) ----
   coeff: 5 doubles per packed pair index {A, B, gamma, rc, rd}. */
void orc_mdpd_density(int N, int ntypes, const int *type, const double *pos, const orc_box *box,
                      const int *n_neigh, const int *neigh, int n_max, const double *coeff,
                      double *rho);
void orc_mdpd_force(int N, int ntypes, const int *id, const int *type, const double *pos,
                    const double *vel, const double *rho, const orc_box *box, const int *n_neigh,
                    const int *neigh, int n_max, const double *coeff, double temp, double dt,
                    uint32_t seed, uint32_t step, double *force, double *pe, double *virial);

/* ---- harmonic bonds (force/bond_harmonic.cu:167-351) ----
   bondlist: 3 ints per bond {type, id_i, id_j} (1-based); bond_coeff: {k, r0} per type.
   Accumulates into force (and bond_pe[b], virial when non-NULL). */
void orc_bond_force(int n_bonds, const int *bondlist, const double *bond_coeff, const double *pos,
                    const orc_box *box, double *force, double *bond_pe, double *virial);

/* ---- harmonic angles (force/angle_harmonic.cu:168-330) ----
   anglelist: 4 ints per angle {type, id_i, id_j, id_k} (1-based, j the vertex);
   angle_coeff: {k, theta0 in degrees} per type; E = k (theta - theta0)^2.
   Accumulates into force (and angle_pe[a], virial when non-NULL; each of the
   three beads gets a third of the angle's virial). */
void orc_angle_force(int n_angles, const int *anglelist, const double *angle_coeff, const double *pos,
                     const orc_box *box, double *force, double *angle_pe, double *virial);

/* ---- nve (fix/fix_nve.cu:30-89, 122-152); masses indexed by type-1 ---- */
void orc_nve_initial(int N, double dt, const int *type, const double *masses, const double *force,
                     const orc_box *box, double *pos, double *uwpos, double *vel);
void orc_nve_final(int N, double dt, const int *type, const double *masses, const double *force,
                   double *vel);

/* ---- thermo (compute/thermo.cu:224-520, bond pe 524-601) ----
   out[0]=T, out[1]=pair PE, out[2]=bond PE, out[3]=P, out[4..9]=Pxx Pxy Pxz Pyy Pyz Pzz,
   out[10]=KE */
void orc_thermo(int N, const int *type, const double *masses, const double *vel, const double *pe,
                const double *virial, const orc_box *box, int n_bonds, const double *bond_pe,
                double *out);

/* ---- whole-run driver in the reference step order (main/utils.cuh:166-314) ---- */
typedef struct {
    int N, ntypes, style;            /* style 0 = dpd, 1 = mdpd */
    orc_box box;
    int *id, *type;
    double *pos, *vel, *uwpos;       /* 3N each */
    double masses[16];
    double coeff[5 * 136];           /* packed pair coefficients */
    double temp, dt;
    uint32_t seed;
    double global_cut, skin, max_rho;
    int check_every, log_every;
    int n_bonds;
    int *bondlist;                   /* 3 * n_bonds */
    double bond_coeff[2 * 16];
} orc_system;

/* Runs step 0 and steps 1..nsteps.  thermo_out receives 11 doubles per log step
   (rows in step order, at most max_rows); returns the number of rows written, or
   -1 on neighbour-list overflow.  force_out (3N, may be NULL) receives the last
   step's forces.  n_builds_out (may be NULL) counts list rebuilds. */
int orc_run(orc_system *sys, int nsteps, double *thermo_out, int max_rows, double *force_out,
            int *n_builds_out);
/* The same run calling observer(step, pos, user) after every observe_every-th
   step's final integrate (pos: 3N wrapped positions, AoS). */
typedef void (*orc_observer)(int step, const double *pos, void *user);
int orc_run_observe(orc_system *sys, int nsteps, double *thermo_out, int max_rows, double *force_out,
                    int *n_builds_out, int observe_every, orc_observer observer, void *user);

/* ---- OpenMP CPU baseline: one full step (list check/rebuild, DPD force with
   per-thread force buffers, nve) -- same arithmetic as above.  Returns the number
   of list rebuilds done. ---- */
typedef struct orc_omp_ctx orc_omp_ctx;
orc_omp_ctx *orc_omp_create(orc_system *sys, int nthreads);
int orc_omp_steps(orc_omp_ctx *c, uint32_t first_step, int nsteps);
int orc_omp_threads(orc_omp_ctx *c);
void orc_omp_destroy(orc_omp_ctx *c);

#ifdef __cplusplus
}
#endif
#endif
