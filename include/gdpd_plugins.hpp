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
    gdpd_ctx *ctx = nullptr;
    int handle = -1;              // libgdpd system handle
};

void check(System &s, int rc);
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

// fix nve (fix/fix_nve.cu); this build integrates every bead ("group all")
class nve : public Fix {
public:
    nve(std::string id, std::string gid, const std::vector<std::string> &params);
    void validateParams(const std::vector<std::string> &params) override;
    std::string getName() override { return "nve"; }
    void initial_integrate(System &system, unsigned int step) override;
    void final_integrate(System &system, unsigned int step) override;
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
// gdpd_run (one launch per step) instead of one call per hook.
void preprocess(System &system, Neigh_list &nl, Thermo &thermo);
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
