// gdpd_internal.h -- device-side descriptors shared by the kernels and the host
// runtime of libgdpd.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gdpd {

constexpr int kMaxTypes = 12;      // the reference's 256-entry constant tables cap DPD at 12 types
constexpr int kMaxBeads = (1 << 27) - 1;  // ids share a 32-bit word with the type and a halo flag in LDS
constexpr int kMaxBondTypes = 16;
constexpr int kMaxAngleTypes = 16;
constexpr int kBlock = 256;        // threads per workgroup for every kernel
constexpr int kThermoParts = 10;   // ke, pair pe, bond pe, p[6], angle pe
constexpr int kThermoRowLen = 13;  // one device thermo record (GDPD_THERMO_LEN)
constexpr int kNumXcd = 8;         // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kStepTsCap = 256;    // ring of step-completion times per replica (GDPD_STEP_TS_CAP)

// group fixes and computes (GDPD_MAX_FIXES, GDPD_MAX_COMPUTES, GDPD_RECORD_*)
constexpr int kMaxFixes = 8;
constexpr int kMaxComputes = 8;
constexpr int kRecLen = 6;                         // {step, up to 5 values}
constexpr int kRecCap = 1024;                      // ring of records per slot
constexpr int kRecSlots = kMaxComputes + kMaxFixes; // compute k -> k, fix k -> kMaxComputes + k
constexpr int kAccPer = 4;                         // reduced quantities per slot

// device flag bits (RepState::flags)
constexpr int kFlagStageOverflow = 1;
constexpr int kFlagListOverflow = 2;
constexpr int kFlagNonFinite = 4;
constexpr int kFlagNeighOverflow = 8;
constexpr int kFlagWallOrder = 16;   // two_wall: the upper wall is not above the lower wall
constexpr int kFlagImageOverflow = 32;  // a periodic image count would leave its 21-bit field
constexpr int kFlagScatter = 64;     // a binned bead fell outside the bead range (internal error)
// A replica whose brick staging overflowed is halted: every step kernel leaves
// its state alone (no integration, no re-binning, no step count or buffer
// flip), so the state stays at the last good step until the host has grown
// the LDS capacities and cleared the flag.
constexpr int kFlagHalt = kFlagStageOverflow;

// step-kernel mode bits
constexpr int kModeFinalKick = 1;  // v += dt/2 F/m before thermo (not at step 0)
constexpr int kModeDrift = 2;      // v += dt/2 F/m; x += dt v; rebin (not at the last step)
constexpr int kModeStoreForce = 4; // keep F in the force array
constexpr int kModeForceOnly = 8;  // hook path: write F (and pe/virial), no integration
constexpr int kModeLogAuto = 16;   // log step when step % log_every == 0
constexpr int kModeLogForce = 32;  // log step regardless of the step number

// Coefficients of one (type_i, type_j) pair, precomputed on the host in the
// working precision.  sig is sqrt(gamma) * sqrt(2T)/sqrt(dt) evaluated in the
// order of force/pair_dpd.cu:315,577-578.
template <typename real> struct PairCoef {
    real A, B, gamma, rc, rc2, rcinv, rd, rdinv, sig, rhofac;
};

// Mutable per-replica scalars that live on the device so a captured graph can
// be replayed: the step counter, the state-buffer parity and error flags.
struct RepState {
    long long step;   // step number of the next force evaluation
    int cur;          // which of the two x/v buffers is current
    int flags;        // kFlag* bits, sticky
    int log_rows;     // thermo records written in this run
    int pad;
};

// One fix (gdpd_fix) with its group's bead count and mass, set on the host.
struct FixDev {
    int kind;
    unsigned group, group2, group3;  // type masks, bit t-1 for type t
    int every, dir, flags, gcount;
    double p[5];
    double gmass;
};

// One compute (gdpd_compute) with its group's bead count and mass.
struct CompDev {
    int kind;
    unsigned group, group2;
    int every, dir, side, gcount, pad;
    double p[6];
    double gmass;
};

// One replica as the kernels see it.  All bead arrays are in "slot" order; the
// cell lists map cell -> sorted positions -> slots through perm.
struct RepDev {
    double lo[3], L[3], hL[3];
    double cs[3], inv_cs[3];   // cell edge per axis (L / nc) and its inverse
    int nc[3], ncell;
    int bdim[3], nbk[3], nbrick;
    int N, ntypes, style, precision;
    int smax, omax;            // LDS capacities: staged and owned beads per brick
    double rcmax2, rdmax2;     // largest rc^2 and rd^2 over type pairs
    int log_every;
    unsigned seed;
    double dt;
    double mass[kMaxTypes], minv[kMaxTypes];
    double bond_k[kMaxBondTypes], bond_r0[kMaxBondTypes];
    double angle_k[kMaxAngleTypes], angle_th0[kMaxAngleTypes];  // theta0 in radians

    double *x[2], *v[2];       // SoA, 3N each: [X..., Y..., Z...]
    int *id;                   // 1-based ids
    unsigned char *type;       // 1-based types
    long long *img;            // packed image counts, 21 bits per axis (offset 2^20)
    int *perm;                 // sorted position -> slot
    int *cell_start;           // ncell + 1
    int *cell_count;           // ncell, zero between steps
    int *key, *rank;           // per slot: cell of the new position, rank in that cell
    int *slot_of_id;           // id-1 -> slot
    double *force;             // SoA 3N
    double *pe, *vir;          // N and SoA 6N (log steps)
    double *bpe;               // N: per-bead half of its bonds' energy (log steps)
    double *rho;               // N (mDPD)
    int nbonds;
    int *bond_start, *bond_partner;  // CSR by id-1: partner ids (1-based)
    unsigned char *bond_type;        // per CSR entry, 0-based
    int nangles;
    // angles a bead belongs to, CSR by id-1; 4 ints per entry:
    // type-1 | role << 8 (0: end i, 1: vertex j, 2: end k), then the other two
    // member ids in i, j, k order
    int *angle_start, *angle_ent;
    double *ape;                     // N: per-bead third of its angles' energy (log steps)
    const void *coef;          // PairCoef<real>[ntypes * ntypes]
    double *thermo_part;       // [nbrick][kThermoParts]: pair PE and virial of log steps
    double *thermo_bead;       // [ceil(N/256)][kThermoParts]: KE, bond terms, m v v
    double *thermo_rows;       // [thermo_cap][kThermoRowLen]
    int thermo_cap;
    // group fixes and computes, in the order given
    int nfix, ncomp;
    FixDev fix[kMaxFixes];
    CompDev comp[kMaxComputes];
    long long *racc;           // [kRecSlots][kAccPer][2]: two-word fixed-point sums
    unsigned long long *rext;  // [kRecSlots][2]: order-preserving keys of a max and a min
    double *rec_rows;          // [kRecSlots][kRecCap][kRecLen]
    int *rec_n;                // [kRecSlots]: records written
    double *x0uw;              // [3N] by id-1: msd reference (unwrapped) positions
    RepState *state;
    unsigned long long *step_ts;  // [kStepTsCap]: device clock when step s completed, at s % kStepTsCap
    // spare id, type and image arrays: a physical re-sort writes the
    // permuted words into them and swaps them in (k_resort, k_resort_swap)
    int *id_alt;
    unsigned char *type_alt;
    long long *img_alt;
};

// Run-list entries an owned cell (rx, ry, rz), 1-based in a brick whose owned
// extent is ex x ey x ez cells, may need in the brick kernel: the end of its own
// column's run, one entry per candidate run that can hold beads (the four
// following columns always; a preceding column's run when it is halo in x or y
// or the cell touches a z face; the own column's bottom halo cell on the low z
// face; the top halo cells of owned preceding columns in a brick one cell
// thick in z), and a closing empty entry.  Depends on the geometry only.
__host__ __device__ inline int cell_run_cap(int ex, int ey, int ez, int rx, int ry, int rz)
{
    const int RY = ey + 2;
    const bool zlo = rz == 1, zhi = rz == ez, thin = ez == 1;
    int n = 1 + 4 + 1 + (zlo ? 1 : 0);
    for (int r = 0; r < 4; ++r) {
        const int cx = rx + r / 3 - 1, cy = ry + r % 3 - 1;
        const bool xyhalo = cx == 0 || cy == 0 || cy == RY - 1;
        if (xyhalo || zlo || zhi) ++n;
        if (thin && !xyhalo) ++n;
    }
    return n;
}

// Run-list entries of a whole brick: one shared empty entry, then every owned
// cell's.
inline int brick_run_entries(int ex, int ey, int ez)
{
    int n = 1;
    for (int rx = 1; rx <= ex; ++rx)
        for (int ry = 1; ry <= ey; ++ry)
            for (int rz = 1; rz <= ez; ++rz) n += cell_run_cap(ex, ey, ez, rx, ry, rz);
    return n;
}

// One workgroup's work: replica index and the local chunk/brick index.
struct WorkItem {
    int rep;
    int local;
};

// Periodic image counts, 21 bits per axis with an offset of 2^20: +-1,048,575
// box crossings per axis.  A count that would leave that range sets
// kFlagImageOverflow instead of wrapping (the reference keeps the unwrapped
// position in full precision, fix/fix_nve.cu:83-85).
constexpr int kImgBias = 1 << 20;
constexpr int kImgMax = kImgBias - 1;
__host__ __device__ inline long long pack_img(int ix, int iy, int iz)
{
    const long long m = (1ll << 21) - 1;
    return ((long long)(ix + kImgBias) & m) | (((long long)(iy + kImgBias) & m) << 21) |
           (((long long)(iz + kImgBias) & m) << 42);
}
__host__ __device__ inline void unpack_img(long long p, int &ix, int &iy, int &iz)
{
    const long long m = (1ll << 21) - 1;
    ix = (int)(p & m) - kImgBias;
    iy = (int)((p >> 21) & m) - kImgBias;
    iz = (int)((p >> 42) & m) - kImgBias;
}
__host__ __device__ inline bool img_in_range(int ix, int iy, int iz)
{
    return ix >= -kImgMax && ix <= kImgMax && iy >= -kImgMax && iy <= kImgMax && iz >= -kImgMax && iz <= kImgMax;
}

} // namespace gdpd
