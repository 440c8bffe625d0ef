/*
 * gdpd_oracle.c -- FP64 CPU oracle for the gpumdpd per-step DPD hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see gdpd_oracle.h).  Built with -ffp-contract=off so
 * every multiply and add rounds exactly as written.
 *
 * What pins each function (checked in tests/test_oracle_*.py):
 *   orc_saru_*        TEA stage equal to a textbook TEA (Wheeler-Needham 1994) that
 *                     reproduces the published zero-key known answer; pre-mix and
 *                     avalanche: committed table cross-checked against an
 *                     independent numpy uint32 restatement (no second published
 *                     source); range/mean/variance/decorrelation.
 *   orc_neigh_build   set equality with orc_neigh_brute (O(N^2)) incl. clamp and
 *                     box-edge cases; brute force is the definition.
 *   orc_neigh_check   exact threshold cases at skin/2.
 *   orc_dpd_force     2-bead closed forms (conservative, dissipative, random terms
 *                     separately), Newton's third law, zero beyond rc,
 *                     gamma=0 energy conservation of a short NVE run.
 *   orc_mdpd_*        isolated-pair density closed form; 3-bead density sums;
 *                     2-bead force closed form; third law; force = -grad of
 *                     Warren's many-body potential; per-bead PE = half of each
 *                     pair energy + (pi rd^4/30) B rho_i^2 (hand-worked line of
 *                     4 beads with unequal densities, brute force on an uneven
 *                     cluster) and -grad(sum PE) = force.
 *   orc_bond_force    closed form F = -2k(r-r0) r_hat, E = k(r-r0)^2; a 2-bead
 *                     bonded NVE run conserves energy.
 *   orc_angle_force   F = -grad of E = k (theta - theta0)^2 by central differences,
 *                     zero net force and torque, zero force at theta0.
 *   orc_nve_*         free particle x(t) = wrap(x0 + v t), unwrapped drift.
 *   orc_thermo        2-bead closed forms; T = 2KE/(3N-3).
 *   orc_run           step ordering against a hand-sequenced run of the pieces.
 */
#include "gdpd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Saru.  Follows force/pair_dpd.cu:203-236 step by step: seed mixing with    */
/* arithmetic right shifts on signed casts, 8 TEA rounds, the final avalanche */
/* and the conversion u = (float)s * 2^-32 (FP32 even in the DOUBLE build     */
/* because TWO_N32 is a float literal, pair_dpd.cu:203,235).                  */
/* ------------------------------------------------------------------------- */

static inline uint32_t sar32(uint32_t v, int n) { return (uint32_t)((int32_t)v >> n); }

/* Pre-mix of the seeds into the TEA block (pair_dpd.cu:213-226). */
static inline void saru_premix(uint32_t a, uint32_t b, uint32_t seed, int32_t timestep, uint32_t *st_out,
                               uint32_t *ws_out)
{
    uint32_t c = seed + (uint32_t)timestep;
    c ^= (a << 7) ^ (b >> 6);
    b += (a >> 4) ^ (c >> 15);
    a ^= (b << 9) + (c << 8);
    c ^= 0xA5366B4Du * ((b >> 11) ^ (a << 1));
    b += 0x72BE1579u * ((a << 4) ^ (c >> 16));
    a ^= 0x3F38A6EDu * ((c >> 5) ^ sar32(b, 22));
    b += a * c;
    a += c ^ (b >> 2);
    b ^= sar32(b, 17);

    uint32_t st = 0x79dedea3u * (a ^ sar32(a, 14));
    uint32_t ws = (st + b) ^ sar32(st, 8);
    st = st + ws * (ws ^ 0xdddf97f5u);
    ws = 0xABCB96F7u + (ws >> 1);
    *st_out = st;
    *ws_out = ws;
}

/* The 8 TEA cycles on the block (st, ws) (pair_dpd.cu:227-232). */
static inline void saru_tea8(uint32_t *st_io, uint32_t *ws_io)
{
    uint32_t st = *st_io, ws = *ws_io;
    uint32_t sum = 0;
    for (int round = 0; round < 8; ++round) {
        sum += 0x9e3779b9u;
        st += ((ws << 4) + 0xA341316Cu) ^ (ws + sum) ^ ((ws >> 5) + 0xC8013EA4u);
        ws += ((st << 4) + 0xAD90777Du) ^ (st + sum) ^ ((st >> 5) + 0x7E95761Eu);
    }
    *st_io = st;
    *ws_io = ws;
}

/* Output avalanche (pair_dpd.cu:233-234). */
static inline uint32_t saru_finish(uint32_t st, uint32_t ws)
{
    uint32_t v = (st ^ (st >> 26)) + ws;
    return (v ^ (v >> 20)) * 0x6957f5a7u;
}

uint32_t orc_saru_u32(uint32_t a, uint32_t b, uint32_t seed, int32_t timestep)
{
    uint32_t st, ws;
    saru_premix(a, b, seed, timestep, &st, &ws);
    saru_tea8(&st, &ws);
    return saru_finish(st, ws);
}

void orc_saru_stages(uint32_t a, uint32_t b, uint32_t seed, int32_t timestep, uint32_t out[5])
{
    uint32_t st, ws;
    saru_premix(a, b, seed, timestep, &st, &ws);
    out[0] = st;
    out[1] = ws;
    saru_tea8(&st, &ws);
    out[2] = st;
    out[3] = ws;
    out[4] = saru_finish(st, ws);
}

float orc_saru_rand_f32(uint32_t a, uint32_t b, uint32_t seed, int32_t timestep)
{
    uint32_t s = orc_saru_u32(a, b, seed, timestep);
    float u = (float)s * 2.3283064365386962890625e-10f;
    return (float)1.7320508075688772935274463 * (u * 2.0f - 1.0f);
}

double orc_saru_rand_f64(uint32_t a, uint32_t b, uint32_t seed, int32_t timestep)
{
    uint32_t s = orc_saru_u32(a, b, seed, timestep);
    float u = (float)s * 2.3283064365386962890625e-10f;
    return 1.7320508075688772935274463 * ((double)u * 2.0 - 1.0);
}

void orc_saru_batch(int n, const uint32_t *tag1, const uint32_t *tag2, uint32_t seed,
                    const int32_t *steps, uint32_t *s_out, float *r32_out, double *r64_out)
{
    for (int k = 0; k < n; ++k) {
        if (s_out) s_out[k] = orc_saru_u32(tag1[k], tag2[k], seed, steps[k]);
        if (r32_out) r32_out[k] = orc_saru_rand_f32(tag1[k], tag2[k], seed, steps[k]);
        if (r64_out) r64_out[k] = orc_saru_rand_f64(tag1[k], tag2[k], seed, steps[k]);
    }
}

/* packed upper-triangular index, force/pair_dpd.cu:134 / :311 (0-based) */
int orc_pair_index(int ti, int tj, int ntypes)
{
    int a = ti < tj ? ti : tj;
    int b = ti < tj ? tj : ti;
    return ((a - 1) * ntypes - (a - 1) * (a - 2) / 2 + b - a + 1) - 1;
}

/* minimum image, force/neigh_list.cu:377-379 */
static inline double min_image(double d, double len, double half)
{
    return d - len * (double)((d >= half) - (d < -half));
}

/* ------------------------------------------------------------------------- */
/* Neighbour list: force/neigh_list.cu:5-28 (sizes), 146-184 (27-bin stencil),*/
/* 255-318 (binning with clamp), 322-393 (half list, r^2 <= bin_size^2).      */
/* ------------------------------------------------------------------------- */

int orc_nbins(double len, double bin_size) { return (int)(len / bin_size); }

int orc_n_max_neigh(double global_cut, double skin, double max_rho)
{
    double bin = global_cut + skin;
    double sphere = 4.0 / 3.0 * M_PI * pow(bin, 3.0);
    int n = (int)(max_rho * sphere) + 1;
    return ((n + 15) / 16) * 16;
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static int bin_of(double x, double lo, double bin, int nb)
{
    int b = (int)((x - lo) / bin);
    if (b < 0) b = 0;
    if (b > nb - 1) b = nb - 1;
    return b;
}

int orc_neigh_build(int N, const double *pos, const orc_box *box, double bin_size, int n_max,
                    int *n_neigh, int *neigh)
{
    double L[3], hL[3];
    int nb[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = box->hi[d] - box->lo[d];
        hL[d] = 0.5 * L[d];
        nb[d] = orc_nbins(L[d], bin_size);
        if (nb[d] < 3) return -2; /* the 27-bin stencil would repeat bins */
    }
    int nbins = nb[0] * nb[1] * nb[2];
    int *count = (int *)calloc((size_t)nbins + 1, sizeof(int));
    int *start = (int *)calloc((size_t)nbins + 1, sizeof(int));
    int *content = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    int *binof = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    for (int i = 0; i < N; ++i) {
        int bx = bin_of(pos[3 * i + 0], box->lo[0], bin_size, nb[0]);
        int by = bin_of(pos[3 * i + 1], box->lo[1], bin_size, nb[1]);
        int bz = bin_of(pos[3 * i + 2], box->lo[2], bin_size, nb[2]);
        binof[i] = bx * nb[1] * nb[2] + by * nb[2] + bz;
        count[binof[i]]++;
    }
    for (int b = 0; b < nbins; ++b) start[b + 1] = start[b] + count[b];
    memset(count, 0, sizeof(int) * (size_t)nbins);
    for (int i = 0; i < N; ++i) content[start[binof[i]] + count[binof[i]]++] = i;

    double cut2 = bin_size * bin_size;
    int dropped = 0;
    for (int i = 0; i < N; ++i) {
        int b = binof[i];
        int bz = b % nb[2], by = (b / nb[2]) % nb[1], bx = b / (nb[2] * nb[1]);
        int cnt = 0;
        for (int oz = -1; oz <= 1; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    int cx = (bx + ox + nb[0]) % nb[0];
                    int cy = (by + oy + nb[1]) % nb[1];
                    int cz = (bz + oz + nb[2]) % nb[2];
                    int c = cx * nb[1] * nb[2] + cy * nb[2] + cz;
                    for (int k = start[c]; k < start[c + 1]; ++k) {
                        int j = content[k];
                        if (i >= j) continue;
                        double dx = min_image(pos[3 * i + 0] - pos[3 * j + 0], L[0], hL[0]);
                        double dy = min_image(pos[3 * i + 1] - pos[3 * j + 1], L[1], hL[1]);
                        double dz = min_image(pos[3 * i + 2] - pos[3 * j + 2], L[2], hL[2]);
                        double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 <= cut2) {
                            if (cnt < n_max) neigh[(size_t)i * n_max + cnt] = j;
                            else dropped++;
                            cnt++;
                        }
                    }
                }
        if (cnt > n_max) cnt = n_max;
        n_neigh[i] = cnt;
        qsort(neigh + (size_t)i * n_max, (size_t)cnt, sizeof(int), cmp_int);
    }
    free(count);
    free(start);
    free(content);
    free(binof);
    return dropped;
}

int orc_neigh_brute(int N, const double *pos, const orc_box *box, double cut, int n_max,
                    int *n_neigh, int *neigh)
{
    double L[3], hL[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = box->hi[d] - box->lo[d];
        hL[d] = 0.5 * L[d];
    }
    double cut2 = cut * cut;
    int dropped = 0;
    for (int i = 0; i < N; ++i) {
        int cnt = 0;
        for (int j = i + 1; j < N; ++j) {
            double dx = min_image(pos[3 * i + 0] - pos[3 * j + 0], L[0], hL[0]);
            double dy = min_image(pos[3 * i + 1] - pos[3 * j + 1], L[1], hL[1]);
            double dz = min_image(pos[3 * i + 2] - pos[3 * j + 2], L[2], hL[2]);
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 <= cut2) {
                if (cnt < n_max) neigh[(size_t)i * n_max + cnt] = j;
                else dropped++;
                cnt++;
            }
        }
        n_neigh[i] = cnt < n_max ? cnt : n_max;
    }
    return dropped;
}

/* force/neigh_list.cu:483-551: rebuild when any |dx|^2 >= (skin/2)^2 */
int orc_neigh_check(int N, const double *pos, const double *last_pos, const orc_box *box,
                    double skin)
{
    double L[3], hL[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = box->hi[d] - box->lo[d];
        hL[d] = 0.5 * L[d];
    }
    double hs2 = (skin * 0.5) * (skin * 0.5);
    for (int i = 0; i < N; ++i) {
        double dx = min_image(pos[3 * i + 0] - last_pos[3 * i + 0], L[0], hL[0]);
        double dy = min_image(pos[3 * i + 1] - last_pos[3 * i + 1], L[1], hL[1]);
        double dz = min_image(pos[3 * i + 2] - last_pos[3 * i + 2], L[2], hL[2]);
        if (dx * dx + dy * dy + dz * dz >= hs2) return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* DPD pair force: force/pair_dpd.cu:240-419 (log kernel) and 564-633.        */
/* ------------------------------------------------------------------------- */

void orc_dpd_force(int N, int ntypes, const int *id, const int *type, const double *pos,
                   const double *vel, const orc_box *box, const int *n_neigh, const int *neigh,
                   int n_max, const double *coeff, double temp, double dt, uint32_t seed,
                   uint32_t step, double *force, double *pe, double *virial)
{
    double L[3], hL[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = box->hi[d] - box->lo[d];
        hL[d] = 0.5 * L[d];
    }
    memset(force, 0, sizeof(double) * 3 * (size_t)N);
    if (pe) memset(pe, 0, sizeof(double) * (size_t)N);
    if (virial) memset(virial, 0, sizeof(double) * 6 * (size_t)N);
    double dtinv_1_2 = 1.0 / sqrt(dt);
    double dt_1_2 = sqrt(2.0 * temp) * dtinv_1_2;

    for (int i = 0; i < N; ++i) {
        for (int k = 0; k < n_neigh[i]; ++k) {
            int j = neigh[(size_t)i * n_max + k];
            double dx = min_image(pos[3 * i + 0] - pos[3 * j + 0], L[0], hL[0]);
            double dy = min_image(pos[3 * i + 1] - pos[3 * j + 1], L[1], hL[1]);
            double dz = min_image(pos[3 * i + 2] - pos[3 * j + 2], L[2], hL[2]);
            double r2 = dx * dx + dy * dy + dz * dz;
            int idx = orc_pair_index(type[i], type[j], ntypes);
            double A = coeff[idx * 3 + 0], gamma = coeff[idx * 3 + 1], rc = coeff[idx * 3 + 2];
            double sig_dt = sqrt(gamma) * dt_1_2;
            if (!(r2 <= rc * rc)) continue;
            double r = sqrt(r2);
            if (r < 1.0e-10) continue;
            double rinv = 1.0 / r;
            double dvx = vel[3 * i + 0] - vel[3 * j + 0];
            double dvy = vel[3 * i + 1] - vel[3 * j + 1];
            double dvz = vel[3 * i + 2] - vel[3 * j + 2];
            double dot = dx * dvx + dy * dvy + dz * dvz;
            double wc = 1.0 - r / rc;
            double wr = wc;
            uint32_t t1 = (uint32_t)(id[i] < id[j] ? id[i] : id[j]);
            uint32_t t2 = (uint32_t)(id[i] < id[j] ? id[j] : id[i]);
            double rnd = orc_saru_rand_f64(t1, t2, seed, (int32_t)step);
            double fpair = A * wc;
            fpair -= gamma * wr * wr * dot * rinv;
            fpair += sig_dt * wr * rnd;
            fpair *= rinv;
            double fx = dx * fpair, fy = dy * fpair, fz = dz * fpair;
            force[3 * i + 0] += fx;
            force[3 * i + 1] += fy;
            force[3 * i + 2] += fz;
            force[3 * j + 0] -= fx;
            force[3 * j + 1] -= fy;
            force[3 * j + 2] -= fz;
            if (pe) {
                double ppe = 0.5 * A * rc * wc * wc;
                pe[i] += 0.5 * ppe;
                pe[j] += 0.5 * ppe;
            }
            if (virial) {
                double w[6] = {0.5 * dx * fx, 0.5 * dx * fy, 0.5 * dx * fz,
                               0.5 * dy * fy, 0.5 * dy * fz, 0.5 * dz * fz};
                for (int c = 0; c < 6; ++c) {
                    virial[6 * i + c] += w[c];
                    virial[6 * j + c] += w[c];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* mDPD: density pass force/pair_mdpd.cu:219-300 (no rc test, no self term),  */
/* force pass :346-537, PE split :479-485.                                    */
/* ------------------------------------------------------------------------- */

void orc_mdpd_density(int N, int ntypes, const int *type, const double *pos, const orc_box *box,
                      const int *n_neigh, const int *neigh, int n_max, const double *coeff,
                      double *rho)
{
    double L[3], hL[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = box->hi[d] - box->lo[d];
        hL[d] = 0.5 * L[d];
    }
    const double pi_value = 3.14159265358979323846;
    memset(rho, 0, sizeof(double) * (size_t)N);
    for (int i = 0; i < N; ++i) {
        for (int k = 0; k < n_neigh[i]; ++k) {
            int j = neigh[(size_t)i * n_max + k];
            double dx = min_image(pos[3 * i + 0] - pos[3 * j + 0], L[0], hL[0]);
            double dy = min_image(pos[3 * i + 1] - pos[3 * j + 1], L[1], hL[1]);
            double dz = min_image(pos[3 * i + 2] - pos[3 * j + 2], L[2], hL[2]);
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            int idx = orc_pair_index(type[i], type[j], ntypes);
            double rd = coeff[idx * 5 + 4];
            double factor = 15.0 / 2.0 / pi_value / (rd * rd * rd);
            double wd = 1.0 - r / rd;
            if (wd < 0.0) wd = 0.0;
            double w = factor * wd * wd;
            rho[i] += w;
            rho[j] += w;
        }
    }
}

void orc_mdpd_force(int N, int ntypes, const int *id, const int *type, const double *pos,
                    const double *vel, const double *rho, const orc_box *box, const int *n_neigh,
                    const int *neigh, int n_max, const double *coeff, double temp, double dt,
                    uint32_t seed, uint32_t step, double *force, double *pe, double *virial)
{
    double L[3], hL[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = box->hi[d] - box->lo[d];
        hL[d] = 0.5 * L[d];
    }
    memset(force, 0, sizeof(double) * 3 * (size_t)N);
    if (pe) memset(pe, 0, sizeof(double) * (size_t)N);
    if (virial) memset(virial, 0, sizeof(double) * 6 * (size_t)N);
    double dtinv_1_2 = 1.0 / sqrt(dt);
    double dt_1_2 = sqrt(2.0 * temp) * dtinv_1_2;

    for (int i = 0; i < N; ++i) {
        for (int k = 0; k < n_neigh[i]; ++k) {
            int j = neigh[(size_t)i * n_max + k];
            double dx = min_image(pos[3 * i + 0] - pos[3 * j + 0], L[0], hL[0]);
            double dy = min_image(pos[3 * i + 1] - pos[3 * j + 1], L[1], hL[1]);
            double dz = min_image(pos[3 * i + 2] - pos[3 * j + 2], L[2], hL[2]);
            double r2 = dx * dx + dy * dy + dz * dz;
            int idx = orc_pair_index(type[i], type[j], ntypes);
            double A = coeff[idx * 5 + 0], B = coeff[idx * 5 + 1], gamma = coeff[idx * 5 + 2];
            double rc = coeff[idx * 5 + 3], rd = coeff[idx * 5 + 4];
            double sig_dt = sqrt(gamma) * dt_1_2;
            if (!(r2 <= rc * rc)) continue;
            double r = sqrt(r2);
            if (r < 1.0e-10) continue;
            double rinv = 1.0 / r;
            double dvx = vel[3 * i + 0] - vel[3 * j + 0];
            double dvy = vel[3 * i + 1] - vel[3 * j + 1];
            double dvz = vel[3 * i + 2] - vel[3 * j + 2];
            double dot = dx * dvx + dy * dvy + dz * dvz;
            double wc = 1.0 - r / rc;
            double wd = 1.0 - r / rd;
            if (wd < 0.0) wd = 0.0;
            double wr = wc;
            uint32_t t1 = (uint32_t)(id[i] < id[j] ? id[i] : id[j]);
            uint32_t t2 = (uint32_t)(id[i] < id[j] ? id[j] : id[i]);
            double rnd = orc_saru_rand_f64(t1, t2, seed, (int32_t)step);
            double fpair = A * wc + B * (rho[i] + rho[j]) * wd;
            fpair -= gamma * wr * wr * dot * rinv;
            fpair += sig_dt * wr * rnd;
            fpair *= rinv;
            double fx = dx * fpair, fy = dy * fpair, fz = dz * fpair;
            force[3 * i + 0] += fx;
            force[3 * i + 1] += fy;
            force[3 * i + 2] += fz;
            force[3 * j + 0] -= fx;
            force[3 * j + 1] -= fy;
            force[3 * j + 2] -= fz;
            if (pe) {
                double u_c = 0.5 * A * rc * wc * wc;
                double u_d_i = 0.25 * B * rd * rho[i] * wd * wd;
                double u_d_j = 0.25 * B * rd * rho[j] * wd * wd;
                pe[i] += 0.5 * u_c + u_d_i;
                pe[j] += 0.5 * u_c + u_d_j;
            }
            if (virial) {
                double w[6] = {0.5 * dx * fx, 0.5 * dx * fy, 0.5 * dx * fz,
                               0.5 * dy * fy, 0.5 * dy * fz, 0.5 * dz * fz};
                for (int c = 0; c < 6; ++c) {
                    virial[6 * i + c] += w[c];
                    virial[6 * j + c] += w[c];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Harmonic bonds: force/bond_harmonic.cu:167-271.  E = k (r - r0)^2.         */
/* ------------------------------------------------------------------------- */

void orc_bond_force(int n_bonds, const int *bondlist, const double *bond_coeff, const double *pos,
                    const orc_box *box, double *force, double *bond_pe, double *virial)
{
    double L[3], hL[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = box->hi[d] - box->lo[d];
        hL[d] = 0.5 * L[d];
    }
    for (int b = 0; b < n_bonds; ++b) {
        int bt = bondlist[3 * b + 0] - 1;
        int i = bondlist[3 * b + 1] - 1;
        int j = bondlist[3 * b + 2] - 1;
        double k = bond_coeff[2 * bt + 0], r0 = bond_coeff[2 * bt + 1];
        double dx = min_image(pos[3 * i + 0] - pos[3 * j + 0], L[0], hL[0]);
        double dy = min_image(pos[3 * i + 1] - pos[3 * j + 1], L[1], hL[1]);
        double dz = min_image(pos[3 * i + 2] - pos[3 * j + 2], L[2], hL[2]);
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        double fb = 0.0;
        if (r > 0.0) fb = -2.0 * k * (r - r0) / r;
        double fx = dx * fb, fy = dy * fb, fz = dz * fb;
        force[3 * i + 0] += fx;
        force[3 * i + 1] += fy;
        force[3 * i + 2] += fz;
        force[3 * j + 0] -= fx;
        force[3 * j + 1] -= fy;
        force[3 * j + 2] -= fz;
        if (bond_pe) bond_pe[b] = k * (r - r0) * (r - r0);
        if (virial) {
            double w[6] = {0.5 * dx * fx, 0.5 * dx * fy, 0.5 * dx * fz,
                           0.5 * dy * fy, 0.5 * dy * fz, 0.5 * dz * fz};
            for (int c = 0; c < 6; ++c) {
                virial[6 * i + c] += w[c];
                virial[6 * j + c] += w[c];
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* nve: fix/fix_nve.cu:30-89 (half kick, drift, while-wrap, unwrapped drift)  */
/* and :122-152 (second half kick).                                           */
/* ------------------------------------------------------------------------- */

static inline double wrap_coord(double x, double lo, double hi, double len)
{
    while (x >= hi || x < lo) x -= len * (double)((x >= hi) - (x < lo));
    return x;
}

/* Harmonic angles, force/angle_harmonic.cu:168-330: cos(theta) clamped to
   [-1, 1], sin(theta) floored at 0.001, a = -2 k (theta - theta0) / sin(theta). */
void orc_angle_force(int n_angles, const int *anglelist, const double *angle_coeff, const double *pos,
                     const orc_box *box, double *force, double *angle_pe, double *virial)
{
    double L[3], hL[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = box->hi[d] - box->lo[d];
        hL[d] = 0.5 * L[d];
    }
    for (int a = 0; a < n_angles; ++a) {
        const int at = anglelist[4 * a + 0] - 1;
        const int i = anglelist[4 * a + 1] - 1, j = anglelist[4 * a + 2] - 1, k = anglelist[4 * a + 3] - 1;
        const double kk = angle_coeff[2 * at + 0];
        const double theta0 = angle_coeff[2 * at + 1] * (M_PI / 180.0);
        double d1[3], d2[3];
        for (int d = 0; d < 3; ++d) {
            d1[d] = min_image(pos[3 * i + d] - pos[3 * j + d], L[d], hL[d]);
            d2[d] = min_image(pos[3 * k + d] - pos[3 * j + d], L[d], hL[d]);
        }
        const double rsq1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
        const double rsq2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
        const double r1 = sqrt(rsq1), r2 = sqrt(rsq2);
        double c = (d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]) / (r1 * r2);
        if (c > 1.0) c = 1.0;
        if (c < -1.0) c = -1.0;
        double sn = sqrt(1.0 - c * c);
        if (sn < 0.001) sn = 0.001;
        const double dth = acos(c) - theta0;
        const double pre = -2.0 * kk * dth / sn;
        const double a11 = pre * c / rsq1, a12 = -pre / (r1 * r2), a22 = pre * c / rsq2;
        double f1[3], f3[3];
        for (int d = 0; d < 3; ++d) {
            f1[d] = a11 * d1[d] + a12 * d2[d];
            f3[d] = a22 * d2[d] + a12 * d1[d];
            force[3 * i + d] += f1[d];
            force[3 * k + d] += f3[d];
            force[3 * j + d] -= f1[d] + f3[d];
        }
        if (angle_pe) angle_pe[a] = kk * dth * dth;
        if (virial) {
            const double w[6] = {(d1[0] * f1[0] + d2[0] * f3[0]) / 3.0, (d1[0] * f1[1] + d2[0] * f3[1]) / 3.0,
                                 (d1[0] * f1[2] + d2[0] * f3[2]) / 3.0, (d1[1] * f1[1] + d2[1] * f3[1]) / 3.0,
                                 (d1[1] * f1[2] + d2[1] * f3[2]) / 3.0, (d1[2] * f1[2] + d2[2] * f3[2]) / 3.0};
            const int who[3] = {i, j, k};
            for (int m = 0; m < 3; ++m)
                for (int c6 = 0; c6 < 6; ++c6) virial[6 * who[m] + c6] += w[c6];
        }
    }
}

void orc_nve_initial(int N, double dt, const int *type, const double *masses, const double *force,
                     const orc_box *box, double *pos, double *uwpos, double *vel)
{
    for (int i = 0; i < N; ++i) {
        double minv = 1.0 / masses[type[i] - 1];
        double dtm = (0.5 * dt) * minv;
        for (int d = 0; d < 3; ++d) {
            double v = vel[3 * i + d] + dtm * force[3 * i + d];
            vel[3 * i + d] = v;
            double x = pos[3 * i + d] + dt * v;
            pos[3 * i + d] = wrap_coord(x, box->lo[d], box->hi[d], box->hi[d] - box->lo[d]);
            if (uwpos) uwpos[3 * i + d] += dt * v;
        }
    }
}

void orc_nve_final(int N, double dt, const int *type, const double *masses, const double *force,
                   double *vel)
{
    for (int i = 0; i < N; ++i) {
        double minv = 1.0 / masses[type[i] - 1];
        double dtm = (0.5 * dt) * minv;
        for (int d = 0; d < 3; ++d) vel[3 * i + d] = vel[3 * i + d] + dtm * force[3 * i + d];
    }
}

/* ------------------------------------------------------------------------- */
/* Thermo: compute/thermo.cu:224-306 (T), 310-388 (PE), 392-520 (pressure).   */
/* ------------------------------------------------------------------------- */

void orc_thermo(int N, const int *type, const double *masses, const double *vel, const double *pe,
                const double *virial, const orc_box *box, int n_bonds, const double *bond_pe,
                double *out)
{
    double ke = 0.0, pe_sum = 0.0, bpe = 0.0, p[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < N; ++i) {
        double m = masses[type[i] - 1];
        double vx = vel[3 * i + 0], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        ke += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        if (pe) pe_sum += pe[i];
        double w[6] = {0, 0, 0, 0, 0, 0};
        if (virial)
            for (int c = 0; c < 6; ++c) w[c] = virial[6 * i + c];
        p[0] += m * vx * vx + w[0];
        p[1] += m * vx * vy + w[1];
        p[2] += m * vx * vz + w[2];
        p[3] += m * vy * vy + w[3];
        p[4] += m * vy * vz + w[4];
        p[5] += m * vz * vz + w[5];
    }
    for (int b = 0; b < n_bonds; ++b) bpe += bond_pe ? bond_pe[b] : 0.0;
    double vol = (box->hi[0] - box->lo[0]) * (box->hi[1] - box->lo[1]) * (box->hi[2] - box->lo[2]);
    double ndof = 3.0 * N - 3.0;
    out[0] = (2.0 * ke) / (ndof * 1.0);
    out[1] = pe_sum;
    out[2] = bpe;
    for (int c = 0; c < 6; ++c) out[4 + c] = p[c] / vol;
    out[3] = (out[4] + out[7] + out[9]) / 3.0;
    out[10] = ke;
}

/* ------------------------------------------------------------------------- */
/* Whole run in the reference order: zero_step (main/utils.cuh:166-213), then */
/* simulation (:246-314): initial integrate, neigh_update (:217-242), pair,   */
/* bond, final integrate, thermo on step % log_f == 0.                        */
/* ------------------------------------------------------------------------- */

typedef struct {
    int n_max;
    int *n_neigh, *neigh;
    double *last_pos, *force, *pe, *virial, *rho, *bond_pe;
} orc_work;

static void pair_and_bond(orc_system *s, orc_work *w, uint32_t step, int log)
{
    int N = s->N;
    if (s->style == 1) {
        orc_mdpd_density(N, s->ntypes, s->type, s->pos, &s->box, w->n_neigh, w->neigh, w->n_max,
                         s->coeff, w->rho);
        orc_mdpd_force(N, s->ntypes, s->id, s->type, s->pos, s->vel, w->rho, &s->box, w->n_neigh,
                       w->neigh, w->n_max, s->coeff, s->temp, s->dt, s->seed, step, w->force,
                       log ? w->pe : NULL, log ? w->virial : NULL);
    } else {
        orc_dpd_force(N, s->ntypes, s->id, s->type, s->pos, s->vel, &s->box, w->n_neigh, w->neigh,
                      w->n_max, s->coeff, s->temp, s->dt, s->seed, step, w->force,
                      log ? w->pe : NULL, log ? w->virial : NULL);
    }
    if (s->n_bonds > 0)
        orc_bond_force(s->n_bonds, s->bondlist, s->bond_coeff, s->pos, &s->box, w->force,
                       log ? w->bond_pe : NULL, log ? w->virial : NULL);
}

static int build_list(orc_system *s, orc_work *w)
{
    int dropped = orc_neigh_build(s->N, s->pos, &s->box, s->global_cut + s->skin, w->n_max,
                                  w->n_neigh, w->neigh);
    memcpy(w->last_pos, s->pos, sizeof(double) * 3 * (size_t)s->N);
    return dropped;
}

int orc_run(orc_system *s, int nsteps, double *thermo_out, int max_rows, double *force_out,
            int *n_builds_out)
{
    return orc_run_observe(s, nsteps, thermo_out, max_rows, force_out, n_builds_out, 0, 0, 0);
}

int orc_run_observe(orc_system *s, int nsteps, double *thermo_out, int max_rows, double *force_out,
                    int *n_builds_out, int observe_every, orc_observer observer, void *user)
{
    int N = s->N;
    orc_work w;
    w.n_max = orc_n_max_neigh(s->global_cut, s->skin, s->max_rho);
    w.n_neigh = (int *)calloc((size_t)N, sizeof(int));
    w.neigh = (int *)calloc((size_t)N * w.n_max, sizeof(int));
    w.last_pos = (double *)calloc(3 * (size_t)N, sizeof(double));
    w.force = (double *)calloc(3 * (size_t)N, sizeof(double));
    w.pe = (double *)calloc((size_t)N, sizeof(double));
    w.virial = (double *)calloc(6 * (size_t)N, sizeof(double));
    w.rho = (double *)calloc((size_t)N, sizeof(double));
    w.bond_pe = (double *)calloc((size_t)(s->n_bonds > 0 ? s->n_bonds : 1), sizeof(double));
    int rows = 0, builds = 0, bad = 0;

    if (build_list(s, &w) != 0) bad = 1;
    builds++;
    pair_and_bond(s, &w, 0, 1);
    if (thermo_out && rows < max_rows) {
        orc_thermo(N, s->type, s->masses, s->vel, w.pe, w.virial, &s->box, s->n_bonds, w.bond_pe,
                   thermo_out + 11 * rows);
        rows++;
    }
    for (int step = 1; step <= nsteps && !bad; ++step) {
        orc_nve_initial(N, s->dt, s->type, s->masses, w.force, &s->box, s->pos, s->uwpos, s->vel);
        if (step % s->check_every == 0) {
            if (s->skin != 0.0) {
                if (orc_neigh_check(N, s->pos, w.last_pos, &s->box, s->skin)) {
                    if (build_list(s, &w) != 0) bad = 1;
                    builds++;
                }
            } else {
                if (build_list(s, &w) != 0) bad = 1;
                builds++;
            }
        }
        int log = (step % s->log_every) == 0;
        pair_and_bond(s, &w, (uint32_t)step, log);
        orc_nve_final(N, s->dt, s->type, s->masses, w.force, s->vel);
        if (log && thermo_out && rows < max_rows) {
            orc_thermo(N, s->type, s->masses, s->vel, w.pe, w.virial, &s->box, s->n_bonds,
                       w.bond_pe, thermo_out + 11 * rows);
            rows++;
        }
        if (observer && observe_every > 0 && step % observe_every == 0) observer(step, s->pos, user);
    }
    if (force_out) memcpy(force_out, w.force, sizeof(double) * 3 * (size_t)N);
    if (n_builds_out) *n_builds_out = builds;
    free(w.n_neigh);
    free(w.neigh);
    free(w.last_pos);
    free(w.force);
    free(w.pe);
    free(w.virial);
    free(w.rho);
    free(w.bond_pe);
    return bad ? -1 : rows;
}

/* ------------------------------------------------------------------------- */
/* OpenMP CPU baseline: the same arithmetic, parallel over beads, with a      */
/* per-thread force buffer for the j-side of the half list.                   */
/* ------------------------------------------------------------------------- */

struct orc_omp_ctx {
    orc_system *s;
    int nthreads, n_max;
    int *n_neigh, *neigh;
    double *last_pos, *force, *tbuf;
    int *binof, *count, *start, *content;
    int nb[3];
};

orc_omp_ctx *orc_omp_create(orc_system *s, int nthreads)
{
    orc_omp_ctx *c = (orc_omp_ctx *)calloc(1, sizeof(orc_omp_ctx));
    c->s = s;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    c->nthreads = nthreads;
    c->n_max = orc_n_max_neigh(s->global_cut, s->skin, s->max_rho);
    size_t N = (size_t)s->N;
    c->n_neigh = (int *)calloc(N, sizeof(int));
    c->neigh = (int *)calloc(N * c->n_max, sizeof(int));
    c->last_pos = (double *)calloc(3 * N, sizeof(double));
    c->force = (double *)calloc(3 * N, sizeof(double));
    c->tbuf = (double *)calloc(3 * N * (size_t)nthreads, sizeof(double));
    double bin = s->global_cut + s->skin;
    for (int d = 0; d < 3; ++d) c->nb[d] = orc_nbins(s->box.hi[d] - s->box.lo[d], bin);
    size_t nbins = (size_t)c->nb[0] * c->nb[1] * c->nb[2];
    c->binof = (int *)calloc(N, sizeof(int));
    c->count = (int *)calloc(nbins + 1, sizeof(int));
    c->start = (int *)calloc(nbins + 1, sizeof(int));
    c->content = (int *)calloc(N, sizeof(int));
    return c;
}

int orc_omp_threads(orc_omp_ctx *c) { return c->nthreads; }

static void omp_build(orc_omp_ctx *c)
{
    orc_system *s = c->s;
    int N = s->N;
    double bin = s->global_cut + s->skin, cut2 = bin * bin;
    int *nb = c->nb;
    int nbins = nb[0] * nb[1] * nb[2];
    double L[3], hL[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = s->box.hi[d] - s->box.lo[d];
        hL[d] = 0.5 * L[d];
    }
    memset(c->count, 0, sizeof(int) * (size_t)(nbins + 1));
#pragma omp parallel for num_threads(c->nthreads) schedule(static)
    for (int i = 0; i < N; ++i) {
        int bx = bin_of(s->pos[3 * i + 0], s->box.lo[0], bin, nb[0]);
        int by = bin_of(s->pos[3 * i + 1], s->box.lo[1], bin, nb[1]);
        int bz = bin_of(s->pos[3 * i + 2], s->box.lo[2], bin, nb[2]);
        c->binof[i] = bx * nb[1] * nb[2] + by * nb[2] + bz;
    }
    for (int i = 0; i < N; ++i) c->count[c->binof[i]]++;
    c->start[0] = 0;
    for (int b = 0; b < nbins; ++b) c->start[b + 1] = c->start[b] + c->count[b];
    memset(c->count, 0, sizeof(int) * (size_t)nbins);
    for (int i = 0; i < N; ++i) c->content[c->start[c->binof[i]] + c->count[c->binof[i]]++] = i;
#pragma omp parallel for num_threads(c->nthreads) schedule(dynamic, 256)
    for (int i = 0; i < N; ++i) {
        int b = c->binof[i];
        int bz = b % nb[2], by = (b / nb[2]) % nb[1], bx = b / (nb[2] * nb[1]);
        int cnt = 0;
        for (int oz = -1; oz <= 1; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    int cx = (bx + ox + nb[0]) % nb[0];
                    int cy = (by + oy + nb[1]) % nb[1];
                    int cz = (bz + oz + nb[2]) % nb[2];
                    int cc = cx * nb[1] * nb[2] + cy * nb[2] + cz;
                    for (int k = c->start[cc]; k < c->start[cc + 1]; ++k) {
                        int j = c->content[k];
                        if (i >= j) continue;
                        double dx = min_image(s->pos[3 * i] - s->pos[3 * j], L[0], hL[0]);
                        double dy = min_image(s->pos[3 * i + 1] - s->pos[3 * j + 1], L[1], hL[1]);
                        double dz = min_image(s->pos[3 * i + 2] - s->pos[3 * j + 2], L[2], hL[2]);
                        if (dx * dx + dy * dy + dz * dz <= cut2 && cnt < c->n_max)
                            c->neigh[(size_t)i * c->n_max + cnt++] = j;
                    }
                }
        c->n_neigh[i] = cnt;
    }
    memcpy(c->last_pos, s->pos, sizeof(double) * 3 * (size_t)N);
}

static void omp_force(orc_omp_ctx *c, uint32_t step)
{
    orc_system *s = c->s;
    int N = s->N;
    double L[3], hL[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = s->box.hi[d] - s->box.lo[d];
        hL[d] = 0.5 * L[d];
    }
    double dt_1_2 = sqrt(2.0 * s->temp) * (1.0 / sqrt(s->dt));
    memset(c->tbuf, 0, sizeof(double) * 3 * (size_t)N * c->nthreads);
#pragma omp parallel num_threads(c->nthreads)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double *f = c->tbuf + (size_t)t * 3 * N;
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < N; ++i) {
            for (int k = 0; k < c->n_neigh[i]; ++k) {
                int j = c->neigh[(size_t)i * c->n_max + k];
                double dx = min_image(s->pos[3 * i] - s->pos[3 * j], L[0], hL[0]);
                double dy = min_image(s->pos[3 * i + 1] - s->pos[3 * j + 1], L[1], hL[1]);
                double dz = min_image(s->pos[3 * i + 2] - s->pos[3 * j + 2], L[2], hL[2]);
                double r2 = dx * dx + dy * dy + dz * dz;
                int idx = orc_pair_index(s->type[i], s->type[j], s->ntypes);
                double A = s->coeff[idx * 3], gamma = s->coeff[idx * 3 + 1];
                double rc = s->coeff[idx * 3 + 2];
                if (!(r2 <= rc * rc)) continue;
                double r = sqrt(r2);
                if (r < 1.0e-10) continue;
                double rinv = 1.0 / r;
                double dot = dx * (s->vel[3 * i] - s->vel[3 * j]) +
                             dy * (s->vel[3 * i + 1] - s->vel[3 * j + 1]) +
                             dz * (s->vel[3 * i + 2] - s->vel[3 * j + 2]);
                double wc = 1.0 - r / rc;
                uint32_t t1 = (uint32_t)(s->id[i] < s->id[j] ? s->id[i] : s->id[j]);
                uint32_t t2 = (uint32_t)(s->id[i] < s->id[j] ? s->id[j] : s->id[i]);
                double rnd = orc_saru_rand_f64(t1, t2, s->seed, (int32_t)step);
                double fpair = A * wc;
                fpair -= gamma * wc * wc * dot * rinv;
                fpair += sqrt(gamma) * dt_1_2 * wc * rnd;
                fpair *= rinv;
                f[3 * i] += dx * fpair;
                f[3 * i + 1] += dy * fpair;
                f[3 * i + 2] += dz * fpair;
                f[3 * j] -= dx * fpair;
                f[3 * j + 1] -= dy * fpair;
                f[3 * j + 2] -= dz * fpair;
            }
        }
#pragma omp for schedule(static)
        for (int i = 0; i < 3 * N; ++i) {
            double acc = 0.0;
            for (int tt = 0; tt < c->nthreads; ++tt) acc += c->tbuf[(size_t)tt * 3 * N + i];
            c->force[i] = acc;
        }
    }
}

int orc_omp_steps(orc_omp_ctx *c, uint32_t first_step, int nsteps)
{
    orc_system *s = c->s;
    int N = s->N, builds = 0;
    double L[3];
    for (int d = 0; d < 3; ++d) L[d] = s->box.hi[d] - s->box.lo[d];
    if (first_step == 0) {
        omp_build(c);
        builds++;
        omp_force(c, 0);
        first_step = 1;
        nsteps--;
    }
    for (int n = 0; n < nsteps; ++n) {
        uint32_t step = first_step + (uint32_t)n;
#pragma omp parallel for num_threads(c->nthreads) schedule(static)
        for (int i = 0; i < N; ++i) {
            double dtm = (0.5 * s->dt) * (1.0 / s->masses[s->type[i] - 1]);
            for (int d = 0; d < 3; ++d) {
                double v = s->vel[3 * i + d] + dtm * c->force[3 * i + d];
                s->vel[3 * i + d] = v;
                s->pos[3 * i + d] = wrap_coord(s->pos[3 * i + d] + s->dt * v, s->box.lo[d],
                                               s->box.hi[d], L[d]);
            }
        }
        int need = 0;
        double hs2 = (s->skin * 0.5) * (s->skin * 0.5);
#pragma omp parallel for num_threads(c->nthreads) reduction(| : need)
        for (int i = 0; i < N; ++i) {
            double r2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                double dd = min_image(s->pos[3 * i + d] - c->last_pos[3 * i + d], L[d], 0.5 * L[d]);
                r2 += dd * dd;
            }
            if (r2 >= hs2) need |= 1;
        }
        if (need || s->skin == 0.0) {
            omp_build(c);
            builds++;
        }
        omp_force(c, step);
#pragma omp parallel for num_threads(c->nthreads) schedule(static)
        for (int i = 0; i < N; ++i) {
            double dtm = (0.5 * s->dt) * (1.0 / s->masses[s->type[i] - 1]);
            for (int d = 0; d < 3; ++d) s->vel[3 * i + d] += dtm * c->force[3 * i + d];
        }
    }
    return builds;
}

void orc_omp_destroy(orc_omp_ctx *c)
{
    if (!c) return;
    free(c->n_neigh);
    free(c->neigh);
    free(c->last_pos);
    free(c->force);
    free(c->tbuf);
    free(c->binof);
    free(c->count);
    free(c->start);
    free(c->content);
    free(c);
}
