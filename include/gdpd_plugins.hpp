// gdpd_plugins.hpp -- C++ host shells that keep the reference's plugin interfaces
// (force/pair.cuh:8-35, force/bond.cuh:8-35, fix/fix.cuh:7-59,
// force/neigh_list.cuh:5-25, compute/thermo.cuh:5-19) as a drop-in over the
// libgdpd C ABI.  A reference driver loop written against these classes runs
// unchanged; each hook forwards to one gdpd_* entry point.
//
// Error behaviour: where the reference prints a message and calls exit(1)
// (main/common.cuh:175-200), these shells throw gdpd_host::Error carrying the
// same message, so a driver can print it and exit(1) itself.
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "gdpd.h"

namespace gdpd_host {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// main/common.cuh:75-130
struct Run_p {
    double global_cut = 1.0, skin = 0.0, max_rho = 10.0;
    int nl_f = 1;
    double dt = 0.01;
    unsigned log_f = 100, nsteps = 0;
    std::string input_data_name, out_data_name;  // read_data / write_data
    bool write_data_flag = false;
    // dump start_step frequency wrapped|unwrapped force_yes|no vel_yes|no
    bool write_dump_flag = false, wrapped_flag = true, dumpforce_flag = false, dumpvel_flag = false;
    unsigned dump_start = 0, dump_f = 1;
};

// system/system.cuh:53-59
struct Box {
    double xlo = 0, ylo = 0, zlo = 0, xhi = 1, yhi = 1, zhi = 1;
    double lx = 1, ly = 1, lz = 1, hlx = 0.5, hly = 0.5, hlz = 0.5;
};

// compute/thermo.cuh / main/common.cuh:136-171
struct Thermo_p {
    double thermo_temp = 0, thermo_pair_pe = 0, thermo_bond_pe = 0, thermo_pressure = 0;
    double pressure_tensor[6] = {0, 0, 0, 0, 0, 0};
};

// The host side of system/system.cuh:112-154; device arrays live in libgdpd.
struct System {
    std::string input;            // run.in text with comments and blank lines removed
    int atom_style = 0;           // 0 atomic, 1 bond, 2 angle
    int n_atoms = 0, n_atomtypes = 0, n_bonds = 0, n_bondtypes = 0, n_angles = 0, n_angletypes = 0;
    Box box;
    Run_p run_p;
    Thermo_p thermo_p;
    std::vector<int> h_id, h_type, h_bondlist, h_anglelist;
    std::vector<int> h_mol_id;    // molecule ids (atom_style bond), index = id - 1
    std::vector<double> h_pos, h_vel, masses;
    bool has_velocities = false;  // the data file had a Velocities section
    int precision = GDPD_PREC_F32;
    int h_update_flag = 0;        // Neigh::h_update_flag
    // system/system.cu:235-301: groups select bead types; a group is kept as a
    // type mask (bit t-1 for type t).  "all" names every type unless a deck
    // defines a group of that name.
    std::map<std::string, uint32_t> groups;
    gdpd_ctx *ctx = nullptr;
    int handle = -1;              // libgdpd system handle
};

void check(System &s, int rc);
// system/system.cu:235-301 (create_groups) from the deck's group commands
void create_groups(System &s);
uint32_t find_group(const System &s, const std::string &name);
// read_in: strips comments (main/read_in.cu:5-29) and parses atom_style,
// neighbor, timestep, thermo, run (main/read_in.cu:40-290)
void read_input(System &s, const std::string &text);
// atoms_mem_alloc + atoms_mem_copy_2_gpu (system/system.cu:5-86): creates the
// device system from the host arrays and uploads them
void atoms_to_device(System &s, int device = 0);
// Read_data::read_data (system/read_data.cu:5-9, 201-381): a LAMMPS-style data
// file with the header counts, box, Masses, Atoms (atomic: id type x y z;
// bond: id mol type x y z), optional Velocities and, for atom_style bond,
// Bonds.  Fills the host arrays id-ordered (index = id - 1); throws Error
// with the reference's messages on malformed input.
void read_data(System &s, const std::string &path);
// Read_data::write_data (system/read_data.cu:469-637) from the host arrays
// (write_data_host) or from the device state (write_data), in the layout the
// reference writes, so a written file reads back with read_data.
void write_data_host(const System &s, const std::string &path);
void write_data(System &s, const std::string &path);
void atoms_from_device(System &s, std::vector<double> *force = nullptr);
void release(System &s);

// ---- Pair (force/pair.cuh:8-35) ---------------------------------------------
class Pair {
public:
    virtual ~Pair() = default;
    virtual bool isEnabled(System &system) = 0;
    virtual void print_pair_info(System &system) = 0;
    virtual std::string getName() = 0;
    virtual void compute_force(System &system, unsigned int step) = 0;
};
std::vector<std::unique_ptr<Pair>> &getPairRegistry();
Pair *&getEnabledPair();
void registerPair(std::unique_ptr<Pair> instance);
void preprocessPairs(System &system);
void postprocessPairs();
void Pair_compute(System &system, unsigned int step);

#define GDPD_REGISTER_PAIR(type)                                         \
    static bool _gdpd_##type##_registered = []() {                       \
        ::gdpd_host::registerPair(std::make_unique<type>());             \
        return true;                                                     \
    }();

class Pair_dpd : public Pair {
public:
    bool isEnabled(System &system) override;
    void print_pair_info(System &system) override;
    std::string getName() override { return "dpd"; }
    void compute_force(System &system, unsigned int step) override;
    double temp = 0, rc = 0;
    unsigned seed = 0;
    std::vector<double> pair_coeff;
};

class Pair_mdpd : public Pair {
public:
    bool isEnabled(System &system) override;
    void print_pair_info(System &system) override;
    std::string getName() override { return "mdpd"; }
    void compute_force(System &system, unsigned int step) override;
    double temp = 0, rc = 0, rd = 0;
    unsigned seed = 0;
    std::vector<double> pair_coeff;
};

// ---- Bond (force/bond.cuh:8-35) ---------------------------------------------
class Bond {
public:
    virtual ~Bond() = default;
    virtual bool isEnabled(System &system) = 0;
    virtual void print_bond_info(System &system) = 0;
    virtual std::string getName() = 0;
    virtual void compute_force(System &system, unsigned int step) = 0;
};
std::vector<std::unique_ptr<Bond>> &getBondRegistry();
Bond *&getEnabledBond();
void registerBond(std::unique_ptr<Bond> instance);
void preprocessBonds(System &system);
void postprocessBonds();
void Bond_compute(System &system, unsigned int step);

class Bond_harmonic : public Bond {
public:
    bool isEnabled(System &system) override;
    void print_bond_info(System &system) override;
    std::string getName() override { return "harmonic"; }
    void compute_force(System &system, unsigned int step) override;
    std::vector<double> bond_coeff;
};

// ---- Angle (force/angle.cuh, force/angle.cu; atom_style angle) ---------------
class Angle {
public:
    virtual ~Angle() = default;
    virtual bool isEnabled(System &system) = 0;
    virtual void print_angle_info(System &system) = 0;
    virtual std::string getName() = 0;
    virtual void compute_force(System &system, unsigned int step) = 0;
};
std::vector<std::unique_ptr<Angle>> &getAngleRegistry();
Angle *&getEnabledAngle();
void registerAngle(std::unique_ptr<Angle> instance);
void preprocessAngles(System &system);
void postprocessAngles();
void Angle_compute(System &system, unsigned int step);

// force/angle_harmonic.cu:19-151, 459-...: `angle_style harmonic`,
// `angle_coeff type k theta0` (degrees), E = k (theta - theta0)^2
class Angle_harmonic : public Angle {
public:
    bool isEnabled(System &system) override;
    void print_angle_info(System &system) override;
    std::string getName() override { return "harmonic"; }
    void compute_force(System &system, unsigned int step) override;
    std::vector<double> angle_coeff;
};

// ---- Fix (fix/fix.cuh:7-59) -------------------------------------------------
class Fix {
public:
    std::string fixID, group_id;
    Fix(std::string id, std::string gid) : fixID(std::move(id)), group_id(std::move(gid)) {}
    virtual ~Fix() = default;
    virtual void validateParams(const std::vector<std::string> &params) = 0;
    virtual std::string getName() = 0;
    virtual void preprocess(System &) {}
    virtual void postprocess(System &) {}
    virtual void initial_integrate(System &, unsigned int) {}
    virtual void post_integrate(System &, unsigned int) {}
    virtual void final_integrate(System &, unsigned int) {}
    virtual void post_force(System &, unsigned int) {}
    virtual void end_of_step(System &, unsigned int) {}
};
using FixConstructor = std::unique_ptr<Fix> (*)(const std::string &, const std::string &,
                                                 const std::vector<std::string> &);
std::unordered_map<std::string, FixConstructor> &getFixRegistry();
std::unordered_map<std::string, std::unique_ptr<Fix>> &getEnabledFixRegistry();
void registerFixType(const std::string &typeName, FixConstructor constructor);
void enableFix(const std::string &id, const std::string &groupid, const std::string &type,
               const std::vector<std::string> &params, System &system);
void preprocessFixes(System &system);
void postprocessFixes(System &system);
void Fix_initial_integrate(System &system, unsigned int step);
void Fix_post_integrate(System &system, unsigned int step);
void Fix_final_integrate(System &system, unsigned int step);
void Fix_post_force(System &system, unsigned int step);
void Fix_end_of_step(System &system, unsigned int step);

// The hot-path fixes (fix/fix_*.cu).  preprocessFixes hands every enabled fix to
// libgdpd in deck order (gdpd_set_fixes); each hook method then forwards to
// gdpd_fix_hook with the fix's index.  Fix_* walk the fixes in deck order.
class GroupFix : public Fix {
public:
    using Fix::Fix;
    int index = -1;          // position in the gdpd_set_fixes array
    virtual gdpd_fix describe(const System &system) const = 0;
    void forward(System &system, int hook, unsigned int step, double *out = nullptr);
};

// fix ID group nve (fix/fix_nve.cu)
class nve : public GroupFix {
public:
    nve(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "nve"; }
    gdpd_fix describe(const System &system) const override;
    void initial_integrate(System &system, unsigned int step) override;
    void final_integrate(System &system, unsigned int step) override;
};

// fix ID group move v_x v_y v_z (fix/fix_move.cu)
class move : public GroupFix {
public:
    move(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "move"; }
    gdpd_fix describe(const System &system) const override;
    void initial_integrate(System &system, unsigned int step) override;
    double v[3] = {0, 0, 0};
};

// fix ID group wall_reflect lo_wall_pos hi_wall_pos x|y|z (fix/fix_wall_reflect.cu)
class wall_reflect : public GroupFix {
public:
    wall_reflect(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "wall_reflect"; }
    gdpd_fix describe(const System &system) const override;
    void preprocess(System &system) override;
    void post_integrate(System &system, unsigned int step) override;
    double lo = 0, hi = 0;
    int dir = 2;
};

// fix ID group two_wall lo_wall_group hi_wall_group x|y|z (fix/fix_two_wall.cu)
class two_wall : public GroupFix {
public:
    two_wall(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "two_wall"; }
    gdpd_fix describe(const System &system) const override;
    void post_integrate(System &system, unsigned int step) override;
    std::string lo_wall, hi_wall;
    int dir = 2;
};

// fix ID group spring k x y z r0 [file frequency] (fix/fix_spring.cu)
class spring : public GroupFix {
public:
    spring(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "spring"; }
    gdpd_fix describe(const System &system) const override;
    void preprocess(System &system) override;
    void postprocess(System &system) override;
    void post_force(System &system, unsigned int step) override;
    void write_row(unsigned int step, const double *f5);
    double k = 0, x = 0, y = 0, z = 0, r0 = 0;
    std::string filename;
    unsigned frequency = 0;
    FILE *file = nullptr;
};

// fix ID group recenter frequency x y z (fix/fix_recenter.cu)
class recenter : public GroupFix {
public:
    recenter(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "recenter"; }
    gdpd_fix describe(const System &system) const override;
    void end_of_step(System &system, unsigned int step) override;
    unsigned frequency = 1;
    double x = 0, y = 0, z = 0;
};

// fix ID group rvcm frequency x_flag y_flag z_flag (fix/fix_rvcm.cu)
class rvcm : public GroupFix {
public:
    rvcm(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "rvcm"; }
    gdpd_fix describe(const System &system) const override;
    void end_of_step(System &system, unsigned int step) override;
    unsigned frequency = 1;
    int flags[3] = {0, 0, 0};
};

// fixes in deck order (the reference walks its unordered_map)
std::vector<GroupFix *> &fixOrder();

// ---- Compute (compute/compute.cuh) --------------------------------------------
class Compute {
public:
    std::string computeID, group_id;
    Compute(std::string id, std::string gid) : computeID(std::move(id)), group_id(std::move(gid)) {}
    virtual ~Compute() = default;
    virtual void validateParams(const std::vector<std::string> &params) = 0;
    virtual std::string getName() = 0;
    virtual void preprocess(System &) {}
    virtual void postprocess(System &);
    virtual void compute(System &system, unsigned int step);
    // device description (compute dump has none: it runs on the host)
    virtual bool on_device() const { return true; }
    virtual gdpd_compute describe(const System &system) const;
    // one file row {v0, v1, v2} of step (reference column layout)
    virtual void write_row(unsigned int step, const double *v) = 0;
    void open(const std::string &header);
    int index = -1;          // position in the gdpd_set_computes array
    int device_kind = 0;     // GDPD_COMPUTE_* (0: host only)
    unsigned frequency = 1;
    std::string filename;
    FILE *file = nullptr;
};
using ComputeConstructor = std::unique_ptr<Compute> (*)(const std::string &, const std::string &,
                                                         const std::vector<std::string> &);
std::unordered_map<std::string, ComputeConstructor> &getComputeRegistry();
std::unordered_map<std::string, std::unique_ptr<Compute>> &getEnabledComputeRegistry();
std::vector<Compute *> &computeOrder();
void registerComputeType(const std::string &typeName, ComputeConstructor constructor);
void enableCompute(const std::string &id, const std::string &groupid, const std::string &type,
                   const std::vector<std::string> &params, System &system);
void preprocessCompute(System &system);
void postprocessCompute(System &system);
void Compute_compute(System &system, unsigned int step);

// compute ID group temp frequency file (compute/compute_temp.cu)
class temp : public Compute {
public:
    temp(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "temp"; }
    void preprocess(System &system) override;
    void write_row(unsigned int step, const double *v) override;
};

// compute ID group msd frequency file (compute/compute_msd.cu)
class msd : public Compute {
public:
    msd(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "msd"; }
    void preprocess(System &system) override;
    void write_row(unsigned int step, const double *v) override;
};

// compute ID group rho frequency xlo xhi ylo yhi zlo zhi file (compute/compute_rho.cu)
class rho : public Compute {
public:
    rho(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "rho"; }
    void preprocess(System &system) override;
    gdpd_compute describe(const System &system) const override;
    void write_row(unsigned int step, const double *v) override;
    double region[6] = {0, 0, 0, 0, 0, 0};
};

// compute ID group tf frequency file (compute/compute_tf.cu)
class tf : public Compute {
public:
    tf(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "tf"; }
    void preprocess(System &system) override;
    void write_row(unsigned int step, const double *v) override;
};

// compute ID group wall_contact frequency wall_group lo|hi x|y|z cutoff file
// (compute/compute_wall_contact.cu)
class wall_contact : public Compute {
public:
    wall_contact(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "wall_contact"; }
    void preprocess(System &system) override;
    gdpd_compute describe(const System &system) const override;
    void write_row(unsigned int step, const double *v) override;
    std::string wall_group;
    int side = 1, dir = 2;
    double cutoff = 1.0;
};

// compute ID group dump frequency file (compute/compute_dump.cu): the group's
// beads as LAMMPS text frames, written on the host from a state download
class dump : public Compute {
public:
    dump(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "dump"; }
    void preprocess(System &system) override;
    void compute(System &system, unsigned int step) override;
    bool on_device() const override { return false; }
    void write_row(unsigned int, const double *) override {}
};

// ---- Neigh_list (force/neigh_list.cuh:5-25) ---------------------------------
class Neigh_list {
public:
    void preprocess(System &system);
    void postprocess(System &system);
    void build(System &system);
    void check_update(System &system);
};

// ---- Thermo (compute/thermo.cuh:5-19) ----------------------------------------
// rows keeps every record; with file_path set, the records also go to that
// file in the reference's thermo.txt columns (compute/thermo.cu:41-62, 122-214)
class Thermo {
public:
    void preprocess(System &system);
    void process(System &system, unsigned int step);
    void postprocess(System &system);
    void write_row(const System &system, const std::vector<double> &row);
    std::vector<std::vector<double>> rows;  // step, T, pair pe, bond pe, P, P6, angle pe
    bool print = false;
    std::string file_path;
    FILE *file = nullptr;
};

// ---- Dump (system/dump.cu) ---------------------------------------------------
// LAMMPS text dump frames (ITEM: TIMESTEP / NUMBER OF ATOMS / BOX BOUNDS pp pp
// pp / ATOMS id type x y z [fx fy fz] [vx vy vz]), id-ordered, positions
// wrapped or unwrapped (x + image * L) as the dump command asks
class Dump {
public:
    void preprocess(System &system, const std::string &path = "all.trj");
    void dump(System &system, unsigned int step);
    void postprocess(System &system);
    FILE *file = nullptr;
};

// The reference's driver (main/utils.cuh:126-314) over the shells: preprocess,
// zero_step, simulation and postprocess.  fused = true runs `run N` through
// gdpd_zero_step / gdpd_advance / gdpd_finish in chunks instead of one call per
// hook; thermo rows and the fix/compute records are collected after each chunk.
void preprocess(System &system, Neigh_list &nl, Thermo &thermo);
void postprocess(System &system, Thermo &thermo);
void zero_step(System &system, Neigh_list &nl, Thermo &thermo, Dump *dump = nullptr);
void simulation(System &system, Neigh_list &nl, Thermo &thermo, Dump *dump = nullptr);
void run_fused(System &system, Thermo &thermo);

// main/main.cu:15-72 for a deck: read_input, read_data (a relative file name is
// taken relative to dir), the device upload, preprocess, the run (fused, or
// one call per hook) and write_data when the deck asks for it.  thermo.txt and
// the dump file all.trj are written into dir; a deck with a dump command runs
// through the hook path (dumps need the state at the end of a step).  The
// system keeps its device state for the caller; release() frees it.
void run_deck(System &system, Thermo &thermo, const std::string &deck_text, const std::string &dir,
              bool fused, int device = 0);

} // namespace gdpd_host
