// gdpd_plugins.cpp -- reference plugin interfaces over the libgdpd C ABI.
//
// Each class keeps the reference's method names and argument meaning; the
// device work is one gdpd_* call per hook.  Deck parsing follows the
// reference's commands (main/read_in.cu:40-290, force/pair_dpd.cu:19-173,
// force/pair_mdpd.cu:21-192, force/bond_harmonic.cu:19-147, fix/fix.cu:60-115);
// data files follow system/read_data.cu.
#include "../../include/gdpd_plugins.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>

namespace gdpd_host {

namespace {

std::vector<std::string> split(const std::string &line, std::string &key)
{
    std::istringstream iss(line);
    std::vector<std::string> args;
    std::string a;
    key.clear();
    iss >> key;
    while (iss >> a) args.push_back(a);
    return args;
}

[[noreturn]] void illegal(const std::string &line, const std::string &error, const std::string &format)
{
    throw Error("\n   Error: " + error + "\n   Line: " + line + "\n   Correct format: " + format);
}

double to_f(const std::string &tok, const std::string &line, const std::string &what, const std::string &fmt)
{
    try {
        size_t k = 0;
        double v = std::stod(tok, &k);
        if (k != tok.size()) throw std::invalid_argument(tok);
        return v;
    } catch (...) {
        illegal(line, "invalid value for " + what + ": " + tok, fmt);
    }
}

long to_i(const std::string &tok, const std::string &line, const std::string &what, const std::string &fmt)
{
    try {
        size_t k = 0;
        long v = std::stol(tok, &k);
        if (k != tok.size()) throw std::invalid_argument(tok);
        return v;
    } catch (...) {
        illegal(line, "invalid value for " + what + ": " + tok, fmt);
    }
}

int packed(int ti, int tj, int nt) { return ((ti - 1) * nt - (ti - 1) * (ti - 2) / 2 + tj - ti + 1) - 1; }

template <typename F> void for_lines(const std::string &input, F f)
{
    std::stringstream ss(input);
    std::string line;
    while (std::getline(ss, line)) {
        std::string key;
        auto args = split(line, key);
        if (!key.empty()) f(line, key, args);
    }
}

} // namespace

void check(System &s, int rc)
{
    if (rc != GDPD_OK) throw Error(std::string("libgdpd: ") + gdpd_last_error(s.ctx));
}

void read_input(System &s, const std::string &text)
{
    // main/read_in.cu:5-29: '#' starts a comment, blank lines are dropped
    std::stringstream in(text);
    std::string line;
    s.input.clear();
    while (std::getline(in, line)) {
        const size_t c = line.find('#');
        if (c != std::string::npos) line = line.substr(0, c);
        if (std::all_of(line.begin(), line.end(), [](unsigned char ch) { return std::isspace(ch); })) continue;
        s.input += line + "\n";
    }
    for_lines(s.input, [&](const std::string &l, const std::string &key, const std::vector<std::string> &a) {
        if (key == "atom_style") {
            const char *fmt = "atom_style style_name";
            if (a.size() != 1) illegal(l, "illegal atom_style command", fmt);
            if (a[0] == "atomic") s.atom_style = 0;
            else if (a[0] == "bond") s.atom_style = 1;
            else if (a[0] == "angle") s.atom_style = 2;
            else illegal(l, "illegal atom_style command", fmt);
        } else if (key == "neighbor") {
            const char *fmt = "neighbor global_cut skin max_rho check_frequency";
            if (a.size() != 4) illegal(l, "illegal neighbor command", fmt);
            s.run_p.global_cut = to_f(a[0], l, "global_cut", fmt);
            s.run_p.skin = to_f(a[1], l, "skin", fmt);
            s.run_p.max_rho = to_f(a[2], l, "max_rho", fmt);
            s.run_p.nl_f = (int)to_i(a[3], l, "check_frequency", fmt);
        } else if (key == "timestep") {
            if (a.size() != 1) illegal(l, "illegal timestep command", "timestep dt");
            s.run_p.dt = to_f(a[0], l, "dt", "timestep dt");
        } else if (key == "thermo") {
            if (a.size() != 1) illegal(l, "illegal thermo command", "thermo N");
            s.run_p.log_f = (unsigned)to_i(a[0], l, "N", "thermo N");
        } else if (key == "run") {
            if (a.size() != 1) illegal(l, "illegal run command", "run N");
            s.run_p.nsteps = (unsigned)to_i(a[0], l, "N", "run N");
        } else if (key == "read_data") {
            if (a.size() != 1) illegal(l, "illegal read_data command", "read_data filename");
            s.run_p.input_data_name = a[0];
        } else if (key == "write_data") {
            if (a.size() != 1) illegal(l, "illegal write_data command", "write_data filename");
            s.run_p.out_data_name = a[0];
            s.run_p.write_data_flag = true;
        } else if (key == "dump") {
            const char *fmt = "dump start_step frequency wrapped/unwrapped force_yes/force_no vel_yes/vel_no";
            if (a.size() != 5) illegal(l, "illegal dump command", fmt);
            s.run_p.dump_start = (unsigned)to_i(a[0], l, "start_step", fmt);
            s.run_p.dump_f = (unsigned)to_i(a[1], l, "frequency", fmt);
            if (s.run_p.dump_f == 0) illegal(l, "illegal dump command", fmt);
            if (a[2] != "wrapped" && a[2] != "unwrapped") illegal(l, "illegal dump command", fmt);
            if (a[3] != "force_yes" && a[3] != "force_no") illegal(l, "illegal dump command", fmt);
            if (a[4] != "vel_yes" && a[4] != "vel_no") illegal(l, "illegal dump command", fmt);
            s.run_p.wrapped_flag = a[2] == "wrapped";
            s.run_p.dumpforce_flag = a[3] == "force_yes";
            s.run_p.dumpvel_flag = a[4] == "vel_yes";
            s.run_p.write_dump_flag = true;
        } else if (key != "pair_style" && key != "pair_coeff" && key != "bond_style" && key != "bond_coeff" &&
                   key != "angle_style" && key != "angle_coeff" && key != "group" && key != "fix" &&
                   key != "compute") {
            // main/read_in.cu:295-305
            throw Error("unknown command " + key);
        }
    });
    if (s.run_p.log_f == 0) throw Error("thermo frequency must be positive");
}

// ---- data files (system/read_data.cu) ------------------------------------------

namespace {

std::string strip_comment(const std::string &line)
{
    std::string t = line.substr(0, line.find('#'));
    const size_t a = t.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = t.find_last_not_of(" \t\r\n");
    return t.substr(a, b - a + 1);
}

bool ends_with(const std::string &t, const std::string &suffix)
{
    return t.size() >= suffix.size() && t.compare(t.size() - suffix.size(), suffix.size(), suffix) == 0;
}

const char *kSections[] = {"Masses", "Atoms", "Velocities", "Bonds", "Angles", "Pair Coeffs", "Bond Coeffs",
                           "Angle Coeffs"};

bool is_section(const std::string &t)
{
    for (const char *k : kSections)
        if (t == k) return true;
    return false;
}

// next `count` non-blank rows after line i (comments stripped)
std::vector<std::string> section_rows(const std::vector<std::string> &lines, size_t &i, long count,
                                      const std::string &name)
{
    std::vector<std::string> rows;
    while (i < lines.size() && (long)rows.size() < count) {
        const std::string t = strip_comment(lines[i++]);
        if (t.empty()) continue;
        if (is_section(t)) break;
        rows.push_back(t);
    }
    if ((long)rows.size() != count)
        throw Error("data file: section " + name + " has " + std::to_string(rows.size()) + " rows, expected " +
                    std::to_string(count));
    return rows;
}

} // namespace

void read_data(System &s, const std::string &path)
{
    std::ifstream f(path);
    if (!f) throw Error("Cannot open data file: " + path);
    std::vector<std::string> lines;
    for (std::string l; std::getline(f, l);) lines.push_back(l);
    // line 0 is the title; header counts and box lines follow until a section
    long n_atoms = -1, n_bonds = 0, n_at = -1, n_bt = 0, n_angles = 0, n_ant = 0;
    bool box[3] = {false, false, false};
    size_t i = 1;
    for (; i < lines.size(); ++i) {
        const std::string t = strip_comment(lines[i]);
        if (t.empty()) continue;
        if (is_section(t)) break;
        std::istringstream is(t);
        if (ends_with(t, "atom types")) {
            if (!(is >> n_at)) throw Error("Error reading number of atom types from line: " + lines[i]);
        } else if (ends_with(t, "bond types")) {
            if (!(is >> n_bt)) throw Error("Error reading number of bond types from line: " + lines[i]);
        } else if (ends_with(t, "angle types")) {
            if (!(is >> n_ant)) throw Error("Error reading number of angle types from line: " + lines[i]);
        } else if (ends_with(t, "angles")) {
            if (!(is >> n_angles)) throw Error("Error reading number of angles from line: " + lines[i]);
        } else if (ends_with(t, "atoms")) {
            if (!(is >> n_atoms)) throw Error("Error reading number of atoms from line: " + lines[i]);
        } else if (ends_with(t, "bonds")) {
            if (!(is >> n_bonds)) throw Error("Error reading number of bonds from line: " + lines[i]);
        } else if (ends_with(t, "xlo xhi")) {
            if (!(is >> s.box.xlo >> s.box.xhi)) throw Error("Error reading xlo xhi values.");
            box[0] = true;
        } else if (ends_with(t, "ylo yhi")) {
            if (!(is >> s.box.ylo >> s.box.yhi)) throw Error("Error reading ylo yhi values.");
            box[1] = true;
        } else if (ends_with(t, "zlo zhi")) {
            if (!(is >> s.box.zlo >> s.box.zhi)) throw Error("Error reading zlo zhi values.");
            box[2] = true;
        } else {
            throw Error("data file: unknown header line: " + lines[i]);
        }
    }
    if (n_atoms <= 0) throw Error("data file: number of atoms is missing");
    if (n_at <= 0) throw Error("data file: number of atom types is missing");
    if (!box[0] || !box[1] || !box[2]) throw Error("data file: box bounds are missing");
    if (s.box.xlo >= s.box.xhi || s.box.ylo >= s.box.yhi || s.box.zlo >= s.box.zhi)
        throw Error("Invalid box dimensions.");
    if (s.atom_style == 0 && n_bonds > 0) throw Error("data file has bonds but atom_style is atomic");
    if (s.atom_style != 2 && n_angles > 0) throw Error("data file has angles but atom_style is not angle");
    const bool molecular = s.atom_style >= 1;
    s.n_atoms = (int)n_atoms;
    s.n_atomtypes = (int)n_at;
    s.n_bonds = molecular ? (int)n_bonds : 0;
    s.n_bondtypes = molecular ? (int)n_bt : 0;
    s.n_angles = s.atom_style == 2 ? (int)n_angles : 0;
    s.n_angletypes = s.atom_style == 2 ? (int)n_ant : 0;
    s.box.lx = s.box.xhi - s.box.xlo;
    s.box.ly = s.box.yhi - s.box.ylo;
    s.box.lz = s.box.zhi - s.box.zlo;
    s.box.hlx = 0.5 * s.box.lx;
    s.box.hly = 0.5 * s.box.ly;
    s.box.hlz = 0.5 * s.box.lz;
    const size_t N = (size_t)n_atoms;
    s.masses.assign((size_t)n_at, -1.0);
    s.h_id.assign(N, 0);
    s.h_type.assign(N, 0);
    s.h_pos.assign(3 * N, 0.0);
    s.h_vel.assign(3 * N, 0.0);
    s.h_mol_id.assign(molecular ? N : 0, 0);
    s.h_bondlist.assign(3 * (size_t)s.n_bonds, -1);
    s.h_anglelist.assign(4 * (size_t)s.n_angles, -1);
    s.has_velocities = false;
    bool got_atoms = false, got_masses = false, got_bonds = false, got_angles = false;
    const double lo[3] = {s.box.xlo, s.box.ylo, s.box.zlo}, hi[3] = {s.box.xhi, s.box.yhi, s.box.zhi};
    while (i < lines.size()) {
        const std::string raw = lines[i];
        const std::string t = strip_comment(raw);
        ++i;
        if (t.empty()) continue;
        if (!is_section(t)) throw Error("data file: unexpected line: " + raw);
        if (t == "Masses") {
            for (const std::string &r : section_rows(lines, i, n_at, t)) {
                std::istringstream is(r);
                long ty;
                double m;
                if (!(is >> ty >> m) || ty < 1 || ty > n_at) throw Error("Error reading mass for atom type: " + r);
                s.masses[(size_t)ty - 1] = m;
            }
            got_masses = true;
        } else if (t == "Atoms") {
            // the style hint after '#' (system/read_data.cu:27-58)
            const size_t c = raw.find('#');
            if (c != std::string::npos) {
                std::istringstream hs(raw.substr(c + 1));
                std::string hint;
                hs >> hint;
                const int style = hint == "atomic" ? 0 : hint == "bond" ? 1 : hint == "angle" ? 2 : -1;
                if (style < 0) throw Error("Unknown atom_style in data file.");
                if (style != s.atom_style)
                    throw Error("atom_style in run.in is inconsistent with the style in data file.");
            }
            for (const std::string &r : section_rows(lines, i, n_atoms, t)) {
                std::istringstream is(r);
                long id, mol = 0, ty;
                double x[3];
                bool ok = molecular ? (bool)(is >> id >> mol >> ty >> x[0] >> x[1] >> x[2])
                                    : (bool)(is >> id >> ty >> x[0] >> x[1] >> x[2]);
                if (!ok) throw Error("data file: bad Atoms row: " + r);
                if (id == 0) throw Error("Atom ID cannot be 0.");
                if (id < 1 || id > n_atoms) throw Error("data file: atom id out of range: " + r);
                if (s.h_id[(size_t)id - 1] != 0) throw Error("data file: duplicate atom id: " + r);
                if (ty < 1 || ty > n_at) throw Error("data file: atom type out of range: " + r);
                const size_t k = (size_t)id - 1;
                s.h_id[k] = (int)id;
                s.h_type[k] = (int)ty;
                if (molecular) s.h_mol_id[k] = (int)mol;
                for (int d = 0; d < 3; ++d) {
                    // Read_data::adjust_coordinate (system/read_data.cu:191-197); a
                    // coordinate more than 16 box lengths out is refused rather
                    // than walked back one box length at a time
                    double v = x[d];
                    const double len = hi[d] - lo[d];
                    if (v < lo[d] - 16.0 * len || v >= hi[d] + 16.0 * len)
                        throw Error("data file: atom coordinate more than 16 box lengths outside the box: " + r);
                    while (v >= hi[d] || v < lo[d]) v -= len * ((v >= hi[d]) - (v < lo[d]));
                    s.h_pos[3 * k + d] = v;
                }
            }
            got_atoms = true;
        } else if (t == "Velocities") {
            for (const std::string &r : section_rows(lines, i, n_atoms, t)) {
                std::istringstream is(r);
                long id;
                double v[3];
                if (!(is >> id >> v[0] >> v[1] >> v[2]) || id < 1 || id > n_atoms)
                    throw Error("data file: bad Velocities row: " + r);
                for (int d = 0; d < 3; ++d) s.h_vel[3 * ((size_t)id - 1) + d] = v[d];
            }
            s.has_velocities = true;
        } else if (t == "Bonds") {
            if (!molecular) throw Error("data file has a Bonds section but atom_style is atomic");
            for (const std::string &r : section_rows(lines, i, n_bonds, t)) {
                std::istringstream is(r);
                long bid, bt, a, b;
                if (!(is >> bid >> bt >> a >> b)) throw Error("data file: bad Bonds row: " + r);
                if (bid == 0) throw Error("Bond ID cannot be 0.");
                if (bid < 1 || bid > n_bonds || bt < 1 || bt > std::max(n_bt, 1L) || a < 1 || a > n_atoms ||
                    b < 1 || b > n_atoms)
                    throw Error("data file: bond entry out of range: " + r);
                const size_t k = 3 * ((size_t)bid - 1);
                s.h_bondlist[k] = (int)bt;
                s.h_bondlist[k + 1] = (int)a;
                s.h_bondlist[k + 2] = (int)b;
            }
            got_bonds = true;
        } else if (t == "Angles") {
            if (s.atom_style != 2) throw Error("data file has an Angles section but atom_style is not angle");
            for (const std::string &r : section_rows(lines, i, n_angles, t)) {
                std::istringstream is(r);
                long aid, at, a, b, c;
                if (!(is >> aid >> at >> a >> b >> c)) throw Error("data file: bad Angles row: " + r);
                if (aid == 0) throw Error("Angle ID cannot be 0.");
                if (aid < 1 || aid > n_angles || at < 1 || at > std::max(n_ant, 1L) || a < 1 || a > n_atoms ||
                    b < 1 || b > n_atoms || c < 1 || c > n_atoms)
                    throw Error("data file: angle entry out of range: " + r);
                const size_t k = 4 * ((size_t)aid - 1);
                s.h_anglelist[k] = (int)at;
                s.h_anglelist[k + 1] = (int)a;
                s.h_anglelist[k + 2] = (int)b;
                s.h_anglelist[k + 3] = (int)c;
            }
            got_angles = true;
        } else {
            // coefficient sections: the deck's *_coeff commands set coefficients
            throw Error("data file: section '" + t + "' is not supported (use pair_coeff/bond_coeff in run.in)");
        }
    }
    if (!got_masses) throw Error("Masses section is missing.");
    if (!got_atoms) throw Error("Atoms section is missing.");
    for (double m : s.masses)
        if (!(m > 0.0)) throw Error("All the masses are not set.");
    if (molecular && s.n_bonds > 0 && !got_bonds) throw Error("Bonds section is missing.");
    if (s.atom_style == 2 && s.n_angles > 0 && !got_angles) throw Error("Angles section is missing.");
    for (size_t k = 0; k < N; ++k)
        if (s.h_id[k] == 0) throw Error("data file: atom " + std::to_string(k + 1) + " is missing");
}

void write_data_host(const System &s, const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw Error("Error opening file: " + path);
    const int N = s.n_atoms;
    const bool bonded = s.atom_style >= 1, angled = s.atom_style == 2;
    std::fprintf(f, "# Lammps format data file\n\n");
    std::fprintf(f, "%d atoms\n", N);
    if (bonded) std::fprintf(f, "%d bonds\n", s.n_bonds);
    if (angled) std::fprintf(f, "%d angles\n", s.n_angles);
    std::fprintf(f, "%d atom types\n", s.n_atomtypes);
    if (bonded) std::fprintf(f, "%d bond types\n", s.n_bondtypes);
    if (angled) std::fprintf(f, "%d angle types\n", s.n_angletypes);
    std::fprintf(f, "\n%.9f %.9f xlo xhi\n%.9f %.9f ylo yhi\n%.9f %.9f zlo zhi\n\n", s.box.xlo, s.box.xhi,
                 s.box.ylo, s.box.yhi, s.box.zlo, s.box.zhi);
    std::fprintf(f, "Masses\n\n");
    for (int t = 0; t < s.n_atomtypes; ++t) std::fprintf(f, "%d %.9f\n", t + 1, s.masses[(size_t)t]);
    std::fprintf(f, "\nAtoms  # %s\n\n", angled ? "angle" : bonded ? "bond" : "atomic");
    for (int k = 0; k < N; ++k) {
        const double *x = &s.h_pos[3 * (size_t)k];
        if (bonded)
            std::fprintf(f, "%d %d %d %.9f %.9f %.9f\n", s.h_id[k],
                         s.h_mol_id.empty() ? 0 : s.h_mol_id[(size_t)s.h_id[k] - 1], s.h_type[k], x[0], x[1], x[2]);
        else
            std::fprintf(f, "%d %d %.9f %.9f %.9f\n", s.h_id[k], s.h_type[k], x[0], x[1], x[2]);
    }
    std::fprintf(f, "\nVelocities\n\n");
    for (int k = 0; k < N; ++k) {
        const double *v = s.h_vel.empty() ? nullptr : &s.h_vel[3 * (size_t)k];
        std::fprintf(f, "%d %.9f %.9f %.9f\n", s.h_id[k], v ? v[0] : 0.0, v ? v[1] : 0.0, v ? v[2] : 0.0);
    }
    if (bonded) {
        std::fprintf(f, "\nBonds\n\n");
        for (int b = 0; b < s.n_bonds; ++b)
            std::fprintf(f, "%d %d %d %d\n", b + 1, s.h_bondlist[3 * (size_t)b], s.h_bondlist[3 * (size_t)b + 1],
                         s.h_bondlist[3 * (size_t)b + 2]);
    }
    if (angled) {
        std::fprintf(f, "\nAngles\n\n");
        for (int a = 0; a < s.n_angles; ++a) {
            const int *e = &s.h_anglelist[4 * (size_t)a];
            std::fprintf(f, "%d %d %d %d %d\n", a + 1, e[0], e[1], e[2], e[3]);
        }
    }
    std::fclose(f);
}

void write_data(System &s, const std::string &path)
{
    atoms_from_device(s);
    write_data_host(s, path);
}

void atoms_to_device(System &s, int device)
{
    if (!s.ctx && gdpd_create(&s.ctx, device) != GDPD_OK) throw Error("gdpd_create failed");
    gdpd_system_desc d{};
    d.n_beads = s.n_atoms;
    d.n_types = s.n_atomtypes;
    d.lo[0] = s.box.xlo; d.lo[1] = s.box.ylo; d.lo[2] = s.box.zlo;
    d.hi[0] = s.box.xhi; d.hi[1] = s.box.yhi; d.hi[2] = s.box.zhi;
    d.dt = s.run_p.dt;
    d.global_cut = s.run_p.global_cut;
    d.skin = s.run_p.skin;
    d.max_rho = s.run_p.max_rho;
    d.check_every = s.run_p.nl_f;
    d.log_every = (int)s.run_p.log_f;
    d.precision = s.precision;
    check(s, gdpd_sys_create(s.ctx, &d, &s.handle));
    check(s, gdpd_upload(s.ctx, s.handle, s.h_id.data(), s.h_type.data(), s.h_pos.data(),
                         s.h_vel.empty() ? nullptr : s.h_vel.data(), s.masses.data()));
    s.box.lx = s.box.xhi - s.box.xlo;
    s.box.ly = s.box.yhi - s.box.ylo;
    s.box.lz = s.box.zhi - s.box.zlo;
    s.box.hlx = 0.5 * s.box.lx;
    s.box.hly = 0.5 * s.box.ly;
    s.box.hlz = 0.5 * s.box.lz;
}

void atoms_from_device(System &s, std::vector<double> *force)
{
    const size_t N = (size_t)s.n_atoms;
    s.h_pos.resize(3 * N);
    s.h_vel.resize(3 * N);
    if (force) force->resize(3 * N);
    check(s, gdpd_download(s.ctx, s.handle, s.h_pos.data(), s.h_vel.data(), nullptr,
                           force ? force->data() : nullptr));
    // host arrays are id-ordered after a download: types follow their ids
    std::vector<int> type_by_id(N, 0);
    for (size_t i = 0; i < N; ++i)
        if (s.h_id[i] >= 1 && (size_t)s.h_id[i] <= N) type_by_id[(size_t)s.h_id[i] - 1] = s.h_type[i];
    s.h_type = type_by_id;
    for (size_t i = 0; i < N; ++i) s.h_id[i] = (int)i + 1;
}

void release(System &s)
{
    if (s.ctx) {
        if (s.handle >= 0) gdpd_sys_destroy(s.ctx, s.handle);
        gdpd_destroy(s.ctx);
    }
    s.ctx = nullptr;
    s.handle = -1;
}

// ---- Pair registry (force/pair.cu:5-93) --------------------------------------

std::vector<std::unique_ptr<Pair>> &getPairRegistry()
{
    static std::vector<std::unique_ptr<Pair>> r;
    return r;
}
Pair *&getEnabledPair()
{
    static Pair *p = nullptr;
    return p;
}
void registerPair(std::unique_ptr<Pair> instance) { getPairRegistry().push_back(std::move(instance)); }

void preprocessPairs(System &system)
{
    Pair *&enabled = getEnabledPair();
    enabled = nullptr;
    for (auto &p : getPairRegistry()) {
        if (p->isEnabled(system)) {
            if (enabled) throw Error("more than one enabled pair_style found");
            enabled = p.get();
        }
    }
    if (!enabled) throw Error("no enabled pair_style found");
}
void postprocessPairs() { getEnabledPair() = nullptr; }
void Pair_compute(System &system, unsigned int step)
{
    if (Pair *p = getEnabledPair()) p->compute_force(system, step);
}

// ---- Pair_dpd (force/pair_dpd.cu:19-173, 564-633) ------------------------------

bool Pair_dpd::isEnabled(System &s)
{
    bool enabled = false;
    const int nt = s.n_atomtypes;
    std::vector<int> got(nt * (nt + 1) / 2, 0);
    pair_coeff.assign(got.size() * 3, 0.0);
    for_lines(s.input, [&](const std::string &l, const std::string &key, const std::vector<std::string> &a) {
        if (key != "pair_style" || a.empty() || a[0] != "dpd") return;
        enabled = true;
        const char *fmt = "pair_style dpd temp rc seed";
        if (a.size() != 4) illegal(l, "illegal pair_style dpd command", fmt);
        temp = to_f(a[1], l, "temp", fmt);
        rc = to_f(a[2], l, "rc", fmt);
        seed = (unsigned)to_i(a[3], l, "seed", fmt);
        if (rc > s.run_p.global_cut) illegal(l, "global rc should not be bigger than the global_cut (neighbor)", fmt);
    });
    if (!enabled) return false;
    for_lines(s.input, [&](const std::string &l, const std::string &key, const std::vector<std::string> &a) {
        if (key != "pair_coeff") return;
        const char *fmt = "pair_coeff typei typej A gamma rc(optional)";
        if (a.size() != 4 && a.size() != 6 && a.size() != 5)
            illegal(l, "illegal pair_coeff command for pair dpd", fmt);
        const int ti = (int)to_i(a[0], l, "typei", fmt), tj = (int)to_i(a[1], l, "typej", fmt);
        if (ti < 1 || ti > tj || tj > nt) illegal(l, "illegal pair_coeff command for pair dpd (typei <= typej)", fmt);
        const int k = packed(ti, tj, nt);
        got[k] = 1;
        pair_coeff[k * 3 + 0] = to_f(a[2], l, "A", fmt);
        pair_coeff[k * 3 + 1] = to_f(a[3], l, "gamma", fmt);
        pair_coeff[k * 3 + 2] = (a.size() >= 5) ? to_f(a[4], l, "rc", fmt) : rc;
        if (pair_coeff[k * 3 + 2] > s.run_p.global_cut)
            illegal(l, "global rc should not be bigger than the global_cut (neighbor)", fmt);
    });
    for (int g : got)
        if (!g) throw Error("all pair_coeff are not set");
    check(s, gdpd_set_pair_dpd(s.ctx, s.handle, temp, rc, seed, pair_coeff.data()));
    return true;
}

void Pair_dpd::print_pair_info(System &s)
{
    std::cout << "   pair_style           dpd " << temp << " " << rc << " " << seed << "\n";
    for (int i = 1; i <= s.n_atomtypes; ++i)
        for (int j = i; j <= s.n_atomtypes; ++j) {
            const int k = packed(i, j, s.n_atomtypes);
            std::cout << "   pair_coeff           " << i << " " << j << " " << pair_coeff[k * 3] << " "
                      << pair_coeff[k * 3 + 1] << " " << pair_coeff[k * 3 + 2] << "\n";
        }
}

void Pair_dpd::compute_force(System &s, unsigned int step)
{
    check(s, gdpd_pair_compute(s.ctx, s.handle, step, (step % s.run_p.log_f) == 0));
}

// ---- Pair_mdpd (force/pair_mdpd.cu:21-192, 689-779) -----------------------------

bool Pair_mdpd::isEnabled(System &s)
{
    bool enabled = false;
    const int nt = s.n_atomtypes;
    std::vector<int> got(nt * (nt + 1) / 2, 0);
    pair_coeff.assign(got.size() * 5, 0.0);
    for_lines(s.input, [&](const std::string &l, const std::string &key, const std::vector<std::string> &a) {
        if (key != "pair_style" || a.empty() || a[0] != "mdpd") return;
        enabled = true;
        const char *fmt = "pair_style mdpd temp rc rd seed";
        if (a.size() != 5) illegal(l, "illegal pair_style mdpd command", fmt);
        temp = to_f(a[1], l, "temp", fmt);
        rc = to_f(a[2], l, "rc", fmt);
        rd = to_f(a[3], l, "rd", fmt);
        seed = (unsigned)to_i(a[4], l, "seed", fmt);
        if (rc > s.run_p.global_cut) illegal(l, "global rc should not be bigger than the global_cut (neighbor)", fmt);
        if (rd > rc) illegal(l, "rd should not be bigger than rc", fmt);
    });
    if (!enabled) return false;
    for_lines(s.input, [&](const std::string &l, const std::string &key, const std::vector<std::string> &a) {
        if (key != "pair_coeff") return;
        const char *fmt = "pair_coeff typei typej A B gamma rc(optional) rd(optional)";
        if (a.size() != 5 && a.size() != 7) illegal(l, "illegal pair_coeff command for pair mdpd", fmt);
        const int ti = (int)to_i(a[0], l, "typei", fmt), tj = (int)to_i(a[1], l, "typej", fmt);
        if (ti < 1 || ti > tj || tj > nt) illegal(l, "illegal pair_coeff command for pair mdpd (typei <= typej)", fmt);
        const int k = packed(ti, tj, nt);
        got[k] = 1;
        pair_coeff[k * 5 + 0] = to_f(a[2], l, "A", fmt);
        pair_coeff[k * 5 + 1] = to_f(a[3], l, "B", fmt);
        pair_coeff[k * 5 + 2] = to_f(a[4], l, "gamma", fmt);
        pair_coeff[k * 5 + 3] = (a.size() == 7) ? to_f(a[5], l, "rc", fmt) : rc;
        pair_coeff[k * 5 + 4] = (a.size() == 7) ? to_f(a[6], l, "rd", fmt) : rd;
        if (pair_coeff[k * 5 + 3] > s.run_p.global_cut)
            illegal(l, "global rc should not be bigger than the global_cut (neighbor)", fmt);
        if (pair_coeff[k * 5 + 4] > pair_coeff[k * 5 + 3]) illegal(l, "rd should not be bigger than rc", fmt);
    });
    for (int g : got)
        if (!g) throw Error("all pair_coeff are not set");
    check(s, gdpd_set_pair_mdpd(s.ctx, s.handle, temp, rc, rd, seed, pair_coeff.data()));
    return true;
}

void Pair_mdpd::print_pair_info(System &s)
{
    std::cout << "   pair_style           mdpd " << temp << " " << rc << " " << rd << " " << seed << "\n";
    (void)s;
}

void Pair_mdpd::compute_force(System &s, unsigned int step)
{
    check(s, gdpd_pair_compute(s.ctx, s.handle, step, (step % s.run_p.log_f) == 0));
}

GDPD_REGISTER_PAIR(Pair_dpd)
GDPD_REGISTER_PAIR(Pair_mdpd)

// ---- Bond (force/bond.cu, force/bond_harmonic.cu:19-147, 355-402) --------------

std::vector<std::unique_ptr<Bond>> &getBondRegistry()
{
    static std::vector<std::unique_ptr<Bond>> r;
    return r;
}
Bond *&getEnabledBond()
{
    static Bond *b = nullptr;
    return b;
}
void registerBond(std::unique_ptr<Bond> instance) { getBondRegistry().push_back(std::move(instance)); }
void preprocessBonds(System &system)
{
    Bond *&enabled = getEnabledBond();
    enabled = nullptr;
    for (auto &b : getBondRegistry()) {
        if (b->isEnabled(system)) {
            if (enabled) throw Error("more than one enabled bond_style found");
            enabled = b.get();
        }
    }
    if (!enabled) throw Error("no enabled bond_style found");
}
void postprocessBonds() { getEnabledBond() = nullptr; }
void Bond_compute(System &system, unsigned int step)
{
    if (Bond *b = getEnabledBond()) b->compute_force(system, step);
}

bool Bond_harmonic::isEnabled(System &s)
{
    bool enabled = false;
    for_lines(s.input, [&](const std::string &l, const std::string &key, const std::vector<std::string> &a) {
        if (key != "bond_style" || a.empty() || a[0] != "harmonic") return;
        enabled = true;
        if (a.size() != 1) illegal(l, "illegal bond_style harmonic command", "bond_style harmonic");
    });
    if (!enabled) return false;
    std::vector<int> got(s.n_bondtypes, 0);
    bond_coeff.assign(2 * (size_t)s.n_bondtypes, 0.0);
    for_lines(s.input, [&](const std::string &l, const std::string &key, const std::vector<std::string> &a) {
        if (key != "bond_coeff") return;
        const char *fmt = "bond_coeff type k r0";
        if (a.size() != 3) illegal(l, "illegal bond_coeff command for bond_style harmonic", fmt);
        const int t = (int)to_i(a[0], l, "type", fmt);
        if (t < 1 || t > s.n_bondtypes) throw Error("type exceeds n_bond_types");
        got[t - 1] = 1;
        bond_coeff[2 * (t - 1)] = to_f(a[1], l, "k", fmt);
        bond_coeff[2 * (t - 1) + 1] = to_f(a[2], l, "r0", fmt);
    });
    for (int g : got)
        if (!g) throw Error("all bond_coeff are not set");
    check(s, gdpd_set_bonds(s.ctx, s.handle, s.n_bonds, s.n_bondtypes, s.h_bondlist.data(), bond_coeff.data()));
    return true;
}

void Bond_harmonic::print_bond_info(System &s)
{
    for (int t = 0; t < s.n_bondtypes; ++t)
        std::cout << "   bond_coeff           " << t + 1 << " " << bond_coeff[2 * t] << " " << bond_coeff[2 * t + 1] << "\n";
}

void Bond_harmonic::compute_force(System &s, unsigned int step)
{
    check(s, gdpd_bond_compute(s.ctx, s.handle, step, (step % s.run_p.log_f) == 0));
}

static bool _gdpd_bond_harmonic_registered = []() {
    registerBond(std::make_unique<Bond_harmonic>());
    return true;
}();

// ---- Angle registry (force/angle.cu) ---------------------------------------------

std::vector<std::unique_ptr<Angle>> &getAngleRegistry()
{
    static std::vector<std::unique_ptr<Angle>> r;
    return r;
}
Angle *&getEnabledAngle()
{
    static Angle *p = nullptr;
    return p;
}
void registerAngle(std::unique_ptr<Angle> instance) { getAngleRegistry().push_back(std::move(instance)); }
void preprocessAngles(System &system)
{
    Angle *&en = getEnabledAngle();
    en = nullptr;
    for (auto &a : getAngleRegistry())
        if (a->isEnabled(system)) {
            if (en) throw Error("only one angle style can be enabled");
            en = a.get();
        }
    if (!en) throw Error("no angle style is enabled (angle_style harmonic)");
    en->print_angle_info(system);
}
void postprocessAngles() { getEnabledAngle() = nullptr; }
void Angle_compute(System &system, unsigned int step)
{
    if (getEnabledAngle()) getEnabledAngle()->compute_force(system, step);
}

bool Angle_harmonic::isEnabled(System &s)
{
    bool enabled = false;
    for_lines(s.input, [&](const std::string &l, const std::string &key, const std::vector<std::string> &a) {
        if (key != "angle_style" || a.empty() || a[0] != "harmonic") return;
        enabled = true;
        if (a.size() != 1) illegal(l, "illegal angle_style harmonic command", "angle_style harmonic");
    });
    if (!enabled) return false;
    std::vector<int> got(s.n_angletypes, 0);
    angle_coeff.assign(2 * (size_t)s.n_angletypes, 0.0);
    for_lines(s.input, [&](const std::string &l, const std::string &key, const std::vector<std::string> &a) {
        if (key != "angle_coeff") return;
        const char *fmt = "angle_coeff type k theta0";
        if (a.size() != 3) illegal(l, "illegal angle_coeff command for angle_style harmonic", fmt);
        const int t = (int)to_i(a[0], l, "type", fmt);
        if (t < 1 || t > s.n_angletypes) throw Error("type exceeds n_angle_types");
        got[t - 1] = 1;
        angle_coeff[2 * (t - 1)] = to_f(a[1], l, "k", fmt);
        angle_coeff[2 * (t - 1) + 1] = to_f(a[2], l, "theta0", fmt);
    });
    for (int g : got)
        if (!g) throw Error("all angle_coeff are not set");
    check(s, gdpd_set_angles(s.ctx, s.handle, s.n_angles, s.n_angletypes, s.h_anglelist.data(), angle_coeff.data()));
    return true;
}

void Angle_harmonic::print_angle_info(System &s)
{
    for (int t = 0; t < s.n_angletypes; ++t)
        std::cout << "   angle_coeff           " << t + 1 << " " << angle_coeff[2 * t] << " " << angle_coeff[2 * t + 1]
                  << "\n";
}

void Angle_harmonic::compute_force(System &s, unsigned int step)
{
    check(s, gdpd_angle_compute(s.ctx, s.handle, step, (step % s.run_p.log_f) == 0));
}

static bool _gdpd_angle_harmonic_registered = []() {
    registerAngle(std::make_unique<Angle_harmonic>());
    return true;
}();

// ---- Fix registry (fix/fix.cu:5-195) -------------------------------------------

std::unordered_map<std::string, FixConstructor> &getFixRegistry()
{
    static std::unordered_map<std::string, FixConstructor> r;
    return r;
}
std::unordered_map<std::string, std::unique_ptr<Fix>> &getEnabledFixRegistry()
{
    static std::unordered_map<std::string, std::unique_ptr<Fix>> r;
    return r;
}
void registerFixType(const std::string &typeName, FixConstructor constructor) { getFixRegistry()[typeName] = constructor; }

void enableFix(const std::string &id, const std::string &groupid, const std::string &type,
               const std::vector<std::string> &params, System &system)
{
    (void)system;
    auto &en = getEnabledFixRegistry();
    if (en.count(id)) throw Error("Error: Fix ID '" + id + "' is already registered!");
    auto it = getFixRegistry().find(type);
    if (it == getFixRegistry().end()) throw Error("Error: Fix type '" + type + "' not found!");
    auto f = it->second(id, groupid, params);
    f->validateParams(params);
    en[id] = std::move(f);
}

std::vector<GroupFix *> &fixOrder()
{
    static std::vector<GroupFix *> v;
    return v;
}

void preprocessFixes(System &system)
{
    getEnabledFixRegistry().clear();
    fixOrder().clear();
    std::vector<std::string> ids;
    for_lines(system.input, [&](const std::string &, const std::string &key, const std::vector<std::string> &a) {
        if (key != "fix") return;
        if (a.size() < 3) throw Error("illegal fix command: fix ID groupID style params");
        enableFix(a[0], a[1], a[2], std::vector<std::string>(a.begin() + 3, a.end()), system);
        ids.push_back(a[0]);
    });
    // every fix goes to the device in deck order; its index is its position
    std::vector<gdpd_fix> dev;
    for (const auto &id : ids) {
        Fix *f = getEnabledFixRegistry()[id].get();
        auto *g = dynamic_cast<GroupFix *>(f);
        if (!g) throw Error("fix " + id + " is not one of this build's fixes");
        g->index = (int)dev.size();
        dev.push_back(g->describe(system));
        fixOrder().push_back(g);
    }
    for (GroupFix *g : fixOrder()) g->preprocess(system);
    if (system.handle >= 0) check(system, gdpd_set_fixes(system.ctx, system.handle, dev.data(), (int)dev.size()));
}
void postprocessFixes(System &system)
{
    for (GroupFix *g : fixOrder()) g->postprocess(system);
}
#define GDPD_FIX_HOOK(name)                                                       \
    void Fix_##name(System &system, unsigned int step)                           \
    {                                                                             \
        for (GroupFix *g : fixOrder()) g->name(system, step);                     \
    }
GDPD_FIX_HOOK(initial_integrate)
GDPD_FIX_HOOK(post_integrate)
GDPD_FIX_HOOK(final_integrate)
GDPD_FIX_HOOK(post_force)
GDPD_FIX_HOOK(end_of_step)
#undef GDPD_FIX_HOOK

// ---- groups (system/system.cu:235-301) -----------------------------------------------

void create_groups(System &s)
{
    s.groups.clear();
    const char *fmt = "group group_name type type1 type2 ...";
    for_lines(s.input, [&](const std::string &l, const std::string &key, const std::vector<std::string> &a) {
        if (key != "group") return;
        if (a.size() < 3 || a[1] != "type") illegal(l, "illegal group command", fmt);
        if (s.groups.count(a[0])) throw Error("Group " + a[0] + " already exists");
        uint32_t mask = 0;
        for (size_t k = 2; k < a.size(); ++k) {
            const long t = to_i(a[k], l, "type id", fmt);
            if (t < 1 || t > s.n_atomtypes) throw Error("Type in group is not included in the data file");
            if (mask & (1u << (t - 1))) illegal(l, "Duplicate type found: " + a[k], fmt);
            mask |= 1u << (t - 1);
        }
        s.groups[a[0]] = mask;
    });
}

uint32_t find_group(const System &s, const std::string &name)
{
    auto it = s.groups.find(name);
    if (it != s.groups.end()) return it->second;
    if (name == "all") return s.n_atomtypes >= 32 ? ~0u : ((1u << s.n_atomtypes) - 1u);
    throw Error("Group " + name + " not found");
}

namespace {

int axis_of(const std::string &a, const std::string &line, const std::string &error, const std::string &fmt)
{
    if (a == "x") return 0;
    if (a == "y") return 1;
    if (a == "z") return 2;
    illegal(line, error, fmt);
}

gdpd_fix fix_of(int kind, uint32_t group)
{
    gdpd_fix f{};
    f.kind = kind;
    f.group = group;
    return f;
}

// one line "%15u\t\t%15g..." like the reference's std::setw(15) << ... << "\t\t" rows
void row_line(FILE *file, unsigned step, const double *v, int n)
{
    if (!file) return;
    std::fprintf(file, "%15u", step);
    for (int k = 0; k < n; ++k) std::fprintf(file, "\t\t%15g", v[k]);
    std::fprintf(file, "\n");
    std::fflush(file);
}

template <typename T> void register_fix(const char *name)
{
    registerFixType(name, [](const std::string &id, const std::string &gid,
                             const std::vector<std::string> &p) -> std::unique_ptr<Fix> {
        return std::make_unique<T>(id, gid, p);
    });
}

} // namespace

void GroupFix::forward(System &s, int hook, unsigned int step, double *out)
{
    check(s, gdpd_fix_hook(s.ctx, s.handle, index, hook, step, out));
}

// ---- fix nve (fix/fix_nve.cu) ------------------------------------------------------

nve::nve(std::string id, std::string gid, const std::vector<std::string> &) : GroupFix(std::move(id), std::move(gid)) {}

void nve::validateParams(const std::vector<std::string> &params)
{
    if (!params.empty()) illegal("Invalid fix_nve parameters", "illegal fix_nve command", "fix ID groupID nve");
}
gdpd_fix nve::describe(const System &s) const { return fix_of(GDPD_FIX_NVE, find_group(s, group_id)); }
void nve::initial_integrate(System &s, unsigned int step) { forward(s, GDPD_HOOK_INITIAL_INTEGRATE, step); }
void nve::final_integrate(System &s, unsigned int step) { forward(s, GDPD_HOOK_FINAL_INTEGRATE, step); }

// ---- fix move (fix/fix_move.cu:1-40) ------------------------------------------------

move::move(std::string id, std::string gid, const std::vector<std::string> &) : GroupFix(std::move(id), std::move(gid)) {}

void move::validateParams(const std::vector<std::string> &p)
{
    const char *fmt = "fix ID groupID move v_x v_y v_z";
    if (p.size() != 3) illegal("Invalid fix_move parameters", "illegal fix_move command", fmt);
    for (int d = 0; d < 3; ++d) v[d] = to_f(p[d], "Invalid fix_move parameters", "v_xyz", fmt);
}
gdpd_fix move::describe(const System &s) const
{
    gdpd_fix f = fix_of(GDPD_FIX_MOVE, find_group(s, group_id));
    for (int d = 0; d < 3; ++d) f.p[d] = v[d];
    return f;
}
void move::initial_integrate(System &s, unsigned int step) { forward(s, GDPD_HOOK_INITIAL_INTEGRATE, step); }

// ---- fix wall_reflect (fix/fix_wall_reflect.cu:1-75) -----------------------------------

wall_reflect::wall_reflect(std::string id, std::string gid, const std::vector<std::string> &)
    : GroupFix(std::move(id), std::move(gid)) {}

void wall_reflect::validateParams(const std::vector<std::string> &p)
{
    const char *fmt = "fix ID groupID wall_reflect lo_wall_pos hi_wall_pos wall_direction";
    const char *err = "illegal fix_wall_reflect command";
    if (p.size() != 3) illegal("Invalid fix_wall_reflect parameters", err, fmt);
    lo = to_f(p[0], "Invalid fix_wall_reflect parameters", "lo_wall_pos", fmt);
    hi = to_f(p[1], "Invalid fix_wall_reflect parameters", "hi_wall_pos", fmt);
    if (lo >= hi) illegal("lo_wall_pos should lower than hi_wall_pos", err, fmt);
    dir = axis_of(p[2], "Invalid fix_wall_reflect direction", err, fmt);
}
void wall_reflect::preprocess(System &s)
{
    const double blo[3] = {s.box.xlo, s.box.ylo, s.box.zlo}, bhi[3] = {s.box.xhi, s.box.yhi, s.box.zhi};
    if (lo < blo[dir] || hi > bhi[dir]) throw Error("Wall position is out of the box!");
}
gdpd_fix wall_reflect::describe(const System &s) const
{
    gdpd_fix f = fix_of(GDPD_FIX_WALL_REFLECT, find_group(s, group_id));
    f.dir = dir;
    f.p[0] = lo;
    f.p[1] = hi;
    return f;
}
void wall_reflect::post_integrate(System &s, unsigned int step) { forward(s, GDPD_HOOK_POST_INTEGRATE, step); }

// ---- fix two_wall (fix/fix_two_wall.cu:1-75) -------------------------------------------

two_wall::two_wall(std::string id, std::string gid, const std::vector<std::string> &)
    : GroupFix(std::move(id), std::move(gid)) {}

void two_wall::validateParams(const std::vector<std::string> &p)
{
    const char *fmt = "fix ID groupID two_wall lo_wall_name hi_wall_name direction";
    const char *err = "illegal fix_two_wall command";
    if (p.size() != 3) illegal("Invalid fix_two_wall parameters", err, fmt);
    lo_wall = p[0];
    hi_wall = p[1];
    dir = axis_of(p[2], "Invalid fix_two_wall direction", err, fmt);
}
gdpd_fix two_wall::describe(const System &s) const
{
    gdpd_fix f = fix_of(GDPD_FIX_TWO_WALL, find_group(s, group_id));
    f.group2 = find_group(s, lo_wall);
    f.group3 = find_group(s, hi_wall);
    f.dir = dir;
    return f;
}
void two_wall::post_integrate(System &s, unsigned int step) { forward(s, GDPD_HOOK_POST_INTEGRATE, step); }

// ---- fix spring (fix/fix_spring.cu:1-140) ----------------------------------------------

spring::spring(std::string id, std::string gid, const std::vector<std::string> &) : GroupFix(std::move(id), std::move(gid)) {}

void spring::validateParams(const std::vector<std::string> &p)
{
    const char *fmt = "fix ID groupID spring spring_k spring_x spring_y spring_z spring_r0 "
                      "file_name(optional) file_frequency(optional)";
    const char *bad = "Invalid fix_spring parameters";
    if (p.size() != 5 && p.size() != 7) illegal(bad, "illegal fix_spring command", fmt);
    k = to_f(p[0], bad, "spring_k", fmt);
    x = to_f(p[1], bad, "spring_x", fmt);
    y = to_f(p[2], bad, "spring_y", fmt);
    z = to_f(p[3], bad, "spring_z", fmt);
    r0 = to_f(p[4], bad, "spring_r0", fmt);
    if (p.size() == 7) {
        filename = p[5];
        const long f = to_i(p[6], bad, "file_frequency", fmt);
        if (f <= 0) illegal(bad, "illegal fix_spring command", fmt);
        frequency = (unsigned)f;
    }
}
void spring::preprocess(System &s)
{
    const double c[3] = {x, y, z};
    const double blo[3] = {s.box.xlo, s.box.ylo, s.box.zlo}, bhi[3] = {s.box.xhi, s.box.yhi, s.box.zhi};
    for (int d = 0; d < 3; ++d)
        if (c[d] < blo[d] || c[d] > bhi[d])
            throw Error(std::string("spring position is out of the box in ") + "xyz"[d] + " direction!");
    if (filename.empty()) return;
    file = std::fopen(filename.c_str(), "w");
    if (!file) throw Error("cannot open " + filename);
    std::fprintf(file, "%15s\t\t%15s\t\t%15s\t\t%15s\t\t%15s\t\t%15s\n", "step", "spring_force_x", "spring_force_y",
                 "spring_force_z", "spring_force", "spring_erergy");
}
void spring::postprocess(System &)
{
    if (file) std::fclose(file);
    file = nullptr;
}
gdpd_fix spring::describe(const System &s) const
{
    gdpd_fix f = fix_of(GDPD_FIX_SPRING, find_group(s, group_id));
    f.every = (int)frequency;
    f.p[0] = k;
    f.p[1] = x;
    f.p[2] = y;
    f.p[3] = z;
    f.p[4] = r0;
    return f;
}
void spring::write_row(unsigned int step, const double *f5) { row_line(file, step, f5, 5); }
void spring::post_force(System &s, unsigned int step)
{
    double out[5];
    forward(s, GDPD_HOOK_POST_FORCE, step, out);
    if (frequency && step % frequency == 0) write_row(step, out);
}

// ---- fix recenter (fix/fix_recenter.cu:1-40) -------------------------------------------

recenter::recenter(std::string id, std::string gid, const std::vector<std::string> &)
    : GroupFix(std::move(id), std::move(gid)) {}

void recenter::validateParams(const std::vector<std::string> &p)
{
    const char *fmt = "fix ID groupID recenter frequency recenter_x recenter_y recenter_z";
    const char *bad = "Invalid fix_recenter parameters";
    if (p.size() != 4) illegal(bad, "illegal fix_recenter command", fmt);
    const long f = to_i(p[0], bad, "frequency", fmt);
    if (f <= 0) illegal(bad, "illegal fix_recenter command", fmt);
    frequency = (unsigned)f;
    x = to_f(p[1], bad, "recenter_x", fmt);
    y = to_f(p[2], bad, "recenter_y", fmt);
    z = to_f(p[3], bad, "recenter_z", fmt);
}
gdpd_fix recenter::describe(const System &s) const
{
    gdpd_fix f = fix_of(GDPD_FIX_RECENTER, find_group(s, group_id));
    f.every = (int)frequency;
    f.p[0] = x;
    f.p[1] = y;
    f.p[2] = z;
    return f;
}
void recenter::end_of_step(System &s, unsigned int step) { forward(s, GDPD_HOOK_END_OF_STEP, step); }

// ---- fix rvcm (fix/fix_rvcm.cu:1-40) ---------------------------------------------------

rvcm::rvcm(std::string id, std::string gid, const std::vector<std::string> &) : GroupFix(std::move(id), std::move(gid)) {}

void rvcm::validateParams(const std::vector<std::string> &p)
{
    const char *fmt = "fix ID groupID rvcm frequency x_flag(0/1) y_flag(0/1) z_flag(0/1)";
    const char *bad = "Invalid fix_rvcm parameters";
    if (p.size() != 4) illegal(bad, "illegal fix_rvcm command", fmt);
    const long f = to_i(p[0], bad, "frequency", fmt);
    if (f <= 0) illegal(bad, "illegal fix_rvcm command", fmt);
    frequency = (unsigned)f;
    for (int d = 0; d < 3; ++d) flags[d] = (int)to_i(p[1 + d], bad, "flag", fmt);
}
gdpd_fix rvcm::describe(const System &s) const
{
    gdpd_fix f = fix_of(GDPD_FIX_RVCM, find_group(s, group_id));
    f.every = (int)frequency;
    // fix_rvcm.cu:160-175 multiplies by the flag: any nonzero flag removes that component
    for (int d = 0; d < 3; ++d)
        if (flags[d]) f.flags |= 1 << d;
    return f;
}
void rvcm::end_of_step(System &s, unsigned int step) { forward(s, GDPD_HOOK_END_OF_STEP, step); }

static bool _gdpd_fixes_registered = []() {
    register_fix<nve>("nve");
    register_fix<move>("move");
    register_fix<wall_reflect>("wall_reflect");
    register_fix<two_wall>("two_wall");
    register_fix<spring>("spring");
    register_fix<recenter>("recenter");
    register_fix<rvcm>("rvcm");
    return true;
}();

// ---- computes (compute/compute.cu and compute_*.cu) --------------------------------------

std::unordered_map<std::string, ComputeConstructor> &getComputeRegistry()
{
    static std::unordered_map<std::string, ComputeConstructor> r;
    return r;
}
std::unordered_map<std::string, std::unique_ptr<Compute>> &getEnabledComputeRegistry()
{
    static std::unordered_map<std::string, std::unique_ptr<Compute>> r;
    return r;
}
std::vector<Compute *> &computeOrder()
{
    static std::vector<Compute *> v;
    return v;
}
void registerComputeType(const std::string &typeName, ComputeConstructor constructor)
{
    getComputeRegistry()[typeName] = constructor;
}

void enableCompute(const std::string &id, const std::string &groupid, const std::string &type,
                   const std::vector<std::string> &params, System &system)
{
    auto &en = getEnabledComputeRegistry();
    if (en.count(id)) throw Error("Error: Compute ID '" + id + "' is already registered!");
    auto it = getComputeRegistry().find(type);
    if (it == getComputeRegistry().end()) throw Error("Error: Compute type '" + type + "' not found!");
    (void)find_group(system, groupid);
    auto c = it->second(id, groupid, params);
    c->validateParams(params);
    en[id] = std::move(c);
}

void preprocessCompute(System &system)
{
    getEnabledComputeRegistry().clear();
    computeOrder().clear();
    std::vector<std::string> ids;
    for_lines(system.input, [&](const std::string &, const std::string &key, const std::vector<std::string> &a) {
        if (key != "compute") return;
        if (a.size() < 3) throw Error("illegal compute command: compute ID groupID style params");
        enableCompute(a[0], a[1], a[2], std::vector<std::string>(a.begin() + 3, a.end()), system);
        ids.push_back(a[0]);
    });
    std::vector<gdpd_compute> dev;
    for (const auto &id : ids) {
        Compute *c = getEnabledComputeRegistry()[id].get();
        if (c->on_device()) {
            c->index = (int)dev.size();
            dev.push_back(c->describe(system));
        }
        computeOrder().push_back(c);
    }
    for (Compute *c : computeOrder()) c->preprocess(system);
    if (system.handle >= 0 && !dev.empty())
        check(system, gdpd_set_computes(system.ctx, system.handle, dev.data(), (int)dev.size()));
}

void postprocessCompute(System &system)
{
    for (Compute *c : computeOrder()) c->postprocess(system);
}

void Compute_compute(System &system, unsigned int step)
{
    for (Compute *c : computeOrder()) c->compute(system, step);
}

void Compute::open(const std::string &header)
{
    file = std::fopen(filename.c_str(), "w");
    if (!file) throw Error("cannot open " + filename);
    std::fprintf(file, "%15s\t\t%s\n", "step", header.c_str());
}
void Compute::postprocess(System &)
{
    if (file) std::fclose(file);
    file = nullptr;
}
gdpd_compute Compute::describe(const System &s) const
{
    gdpd_compute c{};
    c.kind = device_kind;
    c.group = find_group(s, group_id);
    c.every = (int)frequency;
    return c;
}
void Compute::compute(System &s, unsigned int step)
{
    if (step % frequency != 0) return;
    double v[3];
    check(s, gdpd_compute_eval(s.ctx, s.handle, index, v));
    write_row(step, v);
}

namespace {

// frequency and file name of "compute ID group style frequency ... file"
void freq_file(Compute &c, const std::vector<std::string> &p, size_t n, const char *name, const char *fmt)
{
    const std::string bad = std::string("Invalid compute_") + name + " parameters";
    const std::string err = std::string("illegal compute_") + name + " command";
    if (p.size() != n) illegal(bad, err, fmt);
    const long f = to_i(p[0], bad, "frequency", fmt);
    if (f <= 0) illegal(bad, err, fmt);
    c.frequency = (unsigned)f;
    c.filename = p[n - 1];
}

template <typename T> void register_compute(const char *name)
{
    registerComputeType(name, [](const std::string &id, const std::string &gid,
                                 const std::vector<std::string> &p) -> std::unique_ptr<Compute> {
        return std::make_unique<T>(id, gid, p);
    });
}

std::string col(const char *name)
{
    char b[32];
    std::snprintf(b, sizeof b, "%15s", name);
    return b;
}

} // namespace

temp::temp(std::string id, std::string gid, const std::vector<std::string> &) : Compute(std::move(id), std::move(gid))
{
    device_kind = GDPD_COMPUTE_TEMP;
}
void temp::validateParams(const std::vector<std::string> &p)
{
    freq_file(*this, p, 2, "temp", "compute ID   groupID temp frequency filename");
}
void temp::preprocess(System &) { open(col("temp")); }
void temp::write_row(unsigned int step, const double *v) { row_line(file, step, v, 1); }

msd::msd(std::string id, std::string gid, const std::vector<std::string> &) : Compute(std::move(id), std::move(gid))
{
    device_kind = GDPD_COMPUTE_MSD;
}
void msd::validateParams(const std::vector<std::string> &p)
{
    freq_file(*this, p, 2, "msd", "compute ID   groupID msd frequency filename");
}
void msd::preprocess(System &) { open(col("msd")); }
void msd::write_row(unsigned int step, const double *v) { row_line(file, step, v, 1); }

rho::rho(std::string id, std::string gid, const std::vector<std::string> &) : Compute(std::move(id), std::move(gid))
{
    device_kind = GDPD_COMPUTE_RHO;
}
void rho::validateParams(const std::vector<std::string> &p)
{
    const char *fmt = "compute ID   groupID rho frequency xlo xhi ylo yhi zlo zhi filename";
    freq_file(*this, p, 8, "rho", fmt);
    for (int q = 0; q < 6; ++q) region[q] = to_f(p[1 + q], "Invalid compute_rho parameters", "region", fmt);
}
void rho::preprocess(System &s)
{
    const char *n[6] = {"xlo < box.xlo", "xhi > box.xhi", "ylo < box.ylo", "yhi > box.yhi", "zlo < box.zlo",
                        "zhi > box.zhi"};
    const double b[6] = {s.box.xlo, s.box.xhi, s.box.ylo, s.box.yhi, s.box.zlo, s.box.zhi};
    for (int q = 0; q < 6; ++q)
        if ((q % 2 == 0 && region[q] < b[q]) || (q % 2 == 1 && region[q] > b[q]))
            illegal(n[q], "illegal compute_rho command", "compute ID   groupID rho frequency xlo xhi ylo yhi zlo zhi filename");
    open(col("rho"));
}
gdpd_compute rho::describe(const System &s) const
{
    gdpd_compute c = Compute::describe(s);
    for (int q = 0; q < 6; ++q) c.p[q] = region[q];
    return c;
}
void rho::write_row(unsigned int step, const double *v) { row_line(file, step, v, 1); }

tf::tf(std::string id, std::string gid, const std::vector<std::string> &) : Compute(std::move(id), std::move(gid))
{
    device_kind = GDPD_COMPUTE_TF;
}
void tf::validateParams(const std::vector<std::string> &p)
{
    freq_file(*this, p, 2, "tf", "compute ID   groupID tf frequency filename");
}
void tf::preprocess(System &) { open(col("tf_x") + "\t\t" + col("tf_y") + "\t\t" + col("tf_z")); }
void tf::write_row(unsigned int step, const double *v) { row_line(file, step, v, 3); }

wall_contact::wall_contact(std::string id, std::string gid, const std::vector<std::string> &) : Compute(std::move(id), std::move(gid))
{
    device_kind = GDPD_COMPUTE_WALL_CONTACT;
}
void wall_contact::validateParams(const std::vector<std::string> &p)
{
    const char *fmt = "compute ID   groupID wall_contact frequency wall_groupname lo/hi wall_direction cutoff filename";
    const char *bad = "Invalid compute_wall_contact parameters";
    freq_file(*this, p, 6, "wall_contact", fmt);
    wall_group = p[1];
    if (p[2] == "lo") side = 1;
    else if (p[2] == "hi") side = -1;
    else illegal(bad, "illegal compute_wall_contact command", fmt);
    dir = axis_of(p[3], bad, "illegal compute_wall_contact command", fmt);
    cutoff = to_f(p[4], bad, "cutoff", fmt);
}
void wall_contact::preprocess(System &) { open(col("wall_contact")); }
gdpd_compute wall_contact::describe(const System &s) const
{
    gdpd_compute c = Compute::describe(s);
    c.group2 = find_group(s, wall_group);
    c.side = side;
    c.dir = dir;
    c.p[0] = cutoff;
    return c;
}
void wall_contact::write_row(unsigned int step, const double *v)
{
    if (!file) return;
    std::fprintf(file, "%15u\t\t%15d\n", step, (int)std::lround(v[0]));
    std::fflush(file);
}

dump::dump(std::string id, std::string gid, const std::vector<std::string> &) : Compute(std::move(id), std::move(gid)) {}
void dump::validateParams(const std::vector<std::string> &p)
{
    freq_file(*this, p, 2, "dump", "compute ID   groupID dump frequency filename");
}
void dump::preprocess(System &)
{
    file = std::fopen(filename.c_str(), "w");
    if (!file) throw Error("cannot open " + filename);
}
void dump::compute(System &s, unsigned int step)
{
    if (step % frequency != 0 || !file) return;
    const uint32_t g = find_group(s, group_id);
    const size_t N = (size_t)s.n_atoms;
    std::vector<double> pos(3 * N);
    check(s, gdpd_download(s.ctx, s.handle, pos.data(), nullptr, nullptr, nullptr));
    std::vector<int> type_by_id(N, 0), mol_by_id(N, 0);
    for (size_t i = 0; i < N; ++i) {
        const int id = s.h_id[i];
        if (id >= 1 && (size_t)id <= N) type_by_id[(size_t)id - 1] = s.h_type[i];
    }
    if (s.atom_style != 0 && s.h_mol_id.size() == N) mol_by_id = s.h_mol_id;
    size_t n = 0;
    for (size_t i = 0; i < N; ++i) n += (g >> (type_by_id[i] - 1)) & 1u;
    // compute/compute_dump.cu:60-130
    std::fprintf(file, "ITEM: TIMESTEP\n%u\nITEM: NUMBER OF ATOMS\n%zu\nITEM: BOX BOUNDS pp pp pp\n", step, n);
    std::fprintf(file, "%.9f %.9f\n%.9f %.9f\n%.9f %.9f\n", s.box.xlo, s.box.xhi, s.box.ylo, s.box.yhi, s.box.zlo,
                 s.box.zhi);
    std::fprintf(file, s.atom_style == 0 ? "ITEM: ATOMS id type x y z\n" : "ITEM: ATOMS id mol type x y z\n");
    for (size_t i = 0; i < N; ++i) {
        if (!((g >> (type_by_id[i] - 1)) & 1u)) continue;
        if (s.atom_style == 0)
            std::fprintf(file, "%d %d %.9f %.9f %.9f\n", (int)i + 1, type_by_id[i], pos[3 * i], pos[3 * i + 1],
                         pos[3 * i + 2]);
        else
            std::fprintf(file, "%d %d %d %.9f %.9f %.9f\n", (int)i + 1, mol_by_id[i], type_by_id[i], pos[3 * i],
                         pos[3 * i + 1], pos[3 * i + 2]);
    }
    std::fflush(file);
}

static bool _gdpd_computes_registered = []() {
    register_compute<temp>("temp");
    register_compute<msd>("msd");
    register_compute<rho>("rho");
    register_compute<tf>("tf");
    register_compute<wall_contact>("wall_contact");
    register_compute<dump>("dump");
    return true;
}();

// ---- Neigh_list (force/neigh_list.cu) ----------------------------------------------

void Neigh_list::preprocess(System &) {}
void Neigh_list::postprocess(System &) {}
void Neigh_list::build(System &s) { check(s, gdpd_neigh_build(s.ctx, s.handle)); }
void Neigh_list::check_update(System &s) { check(s, gdpd_neigh_check(s.ctx, s.handle, &s.h_update_flag)); }

// ---- Thermo (compute/thermo.cu:96-220) -------------------------------------------

void Thermo::preprocess(System &s)
{
    rows.clear();
    if (file_path.empty()) return;
    file = std::fopen(file_path.c_str(), "w");
    if (!file) throw Error("cannot open thermo file: " + file_path);
    if (s.atom_style == 2)
        std::fprintf(file, "%15s\t%10s\t%20s\t%20s\t%20s\t%12s\t%12s\t%12s\t%12s\t%12s\t%12s\t%12s\n", "step",
                     "temp", "pair_pe", "bond_pe", "angle_pe", "pressure", "pxx", "pxy", "pxz", "pyy", "pyz", "pzz");
    else if (s.atom_style == 1)
        std::fprintf(file, "%15s\t%10s\t%20s\t%20s\t%12s\t%12s\t%12s\t%12s\t%12s\t%12s\t%12s\n", "step", "temp",
                     "pair_pe", "bond_pe", "pressure", "pxx", "pxy", "pxz", "pyy", "pyz", "pzz");
    else
        std::fprintf(file, "%15s\t%10s\t%20s\t%12s\t%12s\t%12s\t%12s\t%12s\t%12s\t%12s\n", "step", "temp",
                     "pair_pe", "pressure", "pxx", "pxy", "pxz", "pyy", "pyz", "pzz");
}

void Thermo::postprocess(System &)
{
    if (file) std::fclose(file);
    file = nullptr;
}

void Thermo::write_row(const System &s, const std::vector<double> &r)
{
    // r: step, T, pair PE, bond PE, P, Pxx Pxy Pxz Pyy Pyz Pzz, angle PE
    if (!file) return;
    if (s.atom_style == 2)
        std::fprintf(file,
                     "%15u\t%10.6f\t%20.6f\t%20.6f\t%20.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\n",
                     (unsigned)r[0], r[1], r[2], r[3], r[11], r[4], r[5], r[6], r[7], r[8], r[9], r[10]);
    else if (s.atom_style == 1)
        std::fprintf(file, "%15u\t%10.6f\t%20.6f\t%20.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\n",
                     (unsigned)r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10]);
    else
        std::fprintf(file, "%15u\t%10.6f\t%20.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\t%12.6f\n",
                     (unsigned)r[0], r[1], r[2], r[4], r[5], r[6], r[7], r[8], r[9], r[10]);
    std::fflush(file);
}

void Thermo::process(System &s, unsigned int step)
{
    if (step % s.run_p.log_f != 0) return;
    double o[GDPD_THERMO_LEN];
    check(s, gdpd_thermo(s.ctx, s.handle, o));
    s.thermo_p.thermo_temp = o[GDPD_THERMO_T];
    s.thermo_p.thermo_pair_pe = o[GDPD_THERMO_PAIR_PE];
    s.thermo_p.thermo_bond_pe = o[GDPD_THERMO_BOND_PE];
    s.thermo_p.thermo_pressure = o[GDPD_THERMO_P];
    for (int k = 0; k < 6; ++k) s.thermo_p.pressure_tensor[k] = o[GDPD_THERMO_PXX + k];
    std::vector<double> r = {(double)step, o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], o[8], o[9],
                             o[GDPD_THERMO_ANGLE_PE]};
    rows.push_back(r);
    write_row(s, r);
    if (print)
        std::printf("%15u\t%10.6f\t%20.6f\t%20.6f\t%12.6f\n", step, o[0], o[1], o[2], o[3]);
}

// ---- driver (main/utils.cuh:126-314) -------------------------------------------------

void preprocess(System &system, Neigh_list &nl, Thermo &thermo)
{
    create_groups(system);
    nl.preprocess(system);
    preprocessPairs(system);
    if (system.atom_style >= 1) preprocessBonds(system);
    if (system.atom_style == 2) preprocessAngles(system);
    preprocessFixes(system);
    thermo.preprocess(system);
    preprocessCompute(system);
}

void postprocess(System &system, Thermo &thermo)
{
    postprocessCompute(system);
    postprocessFixes(system);
    thermo.postprocess(system);
}

// ---- Dump (system/dump.cu) ------------------------------------------------------------

void Dump::preprocess(System &, const std::string &path)
{
    file = std::fopen(path.c_str(), "w");
    if (!file) throw Error("cannot open dump file: " + path);
}

void Dump::postprocess(System &)
{
    if (file) std::fclose(file);
    file = nullptr;
}

void Dump::dump(System &s, unsigned int step)
{
    if (!file) return;
    const size_t N = (size_t)s.n_atoms;
    std::vector<double> pos(3 * N), vel(3 * N), force(3 * N);
    std::vector<int> image(3 * N);
    check(s, gdpd_download(s.ctx, s.handle, pos.data(), vel.data(), image.data(), force.data()));
    // types by id (host arrays may be in upload order)
    std::vector<int> type_by_id(N, 0);
    for (size_t i = 0; i < N; ++i)
        if (s.h_id[i] >= 1 && (size_t)s.h_id[i] <= N) type_by_id[(size_t)s.h_id[i] - 1] = s.h_type[i];
    const Run_p &rp = s.run_p;
    const double L[3] = {s.box.xhi - s.box.xlo, s.box.yhi - s.box.ylo, s.box.zhi - s.box.zlo};
    std::fprintf(file, "ITEM: TIMESTEP\n%u\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n", step, s.n_atoms);
    std::fprintf(file, "%.9f %.9f\n%.9f %.9f\n%.9f %.9f\n", s.box.xlo, s.box.xhi, s.box.ylo, s.box.yhi, s.box.zlo,
                 s.box.zhi);
    std::fprintf(file, "ITEM: ATOMS id type x y z%s%s\n", rp.dumpforce_flag ? " fx fy fz" : "",
                 rp.dumpvel_flag ? " vx vy vz" : "");
    for (size_t i = 0; i < N; ++i) {
        double x[3];
        for (int d = 0; d < 3; ++d) x[d] = pos[3 * i + d] + (rp.wrapped_flag ? 0.0 : image[3 * i + d] * L[d]);
        std::fprintf(file, "%d %d %.9f %.9f %.9f", (int)i + 1, type_by_id[i], x[0], x[1], x[2]);
        if (rp.dumpforce_flag) std::fprintf(file, " %.9f %.9f %.9f", force[3 * i], force[3 * i + 1], force[3 * i + 2]);
        if (rp.dumpvel_flag) std::fprintf(file, " %.9f %.9f %.9f", vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        std::fprintf(file, "\n");
    }
    std::fflush(file);
}

void zero_step(System &system, Neigh_list &nl, Thermo &thermo, Dump *dump)
{
    nl.build(system);
    Pair_compute(system, 0);
    if (system.atom_style >= 1) Bond_compute(system, 0);
    if (system.atom_style == 2) Angle_compute(system, 0);
    Fix_post_force(system, 0);
    thermo.process(system, 0);
    Compute_compute(system, 0);
    // main/utils.cuh:200-204
    if (dump && system.run_p.write_dump_flag && system.run_p.dump_start == 0) dump->dump(system, 0);
}

void simulation(System &system, Neigh_list &nl, Thermo &thermo, Dump *dump)
{
    for (unsigned int step = 1; step <= system.run_p.nsteps; ++step) {
        Fix_initial_integrate(system, step);
        Fix_post_integrate(system, step);
        if (step % system.run_p.nl_f == 0) {
            if (system.run_p.skin != 0.0) {
                nl.check_update(system);
                if (system.h_update_flag == 1) nl.build(system);
            } else {
                nl.build(system);
            }
        }
        Pair_compute(system, step);
        if (system.atom_style >= 1) Bond_compute(system, step);
        if (system.atom_style == 2) Angle_compute(system, step);
        Fix_post_force(system, step);
        Fix_final_integrate(system, step);
        Fix_end_of_step(system, step);
        thermo.process(system, step);
        Compute_compute(system, step);
        // main/utils.cuh:296-300
        const Run_p &rp = system.run_p;
        if (dump && rp.write_dump_flag && step >= rp.dump_start && (step - rp.dump_start) % rp.dump_f == 0)
            dump->dump(system, step);
    }
}

void run_deck(System &system, Thermo &thermo, const std::string &deck_text, const std::string &dir, bool fused,
              int device)
{
    read_input(system, deck_text);
    if (system.run_p.input_data_name.empty()) throw Error("read_data command is missing");
    auto resolve = [&](const std::string &name) {
        return (name.empty() || name[0] == '/' || dir.empty()) ? name : dir + "/" + name;
    };
    read_data(system, resolve(system.run_p.input_data_name));
    atoms_to_device(system, device);
    Neigh_list nl;
    if (thermo.file_path.empty()) thermo.file_path = resolve("thermo.txt");
    // compute and fix spring output files are named relative to dir too
    const std::string input = system.input;
    {
        std::stringstream in(input), out;
        std::string line;
        while (std::getline(in, line)) {
            std::string key;
            auto a = split(line, key);
            const size_t n = a.size();
            // the file name is a compute's last argument and a spring's eighth
            if (!dir.empty() && ((key == "compute" && n >= 5) || (key == "fix" && n == 10 && a[2] == "spring"))) {
                const size_t f = key == "compute" ? n - 1 : 8;
                if (a[f][0] != '/') a[f] = dir + "/" + a[f];
                line = key;
                for (const auto &t : a) line += " " + t;
            }
            out << line << "\n";
        }
        system.input = out.str();
    }
    preprocess(system, nl, thermo);
    // dumps need the state at the end of a step: decks with a dump run the hook path
    bool host_dump = system.run_p.write_dump_flag;
    for (Compute *c : computeOrder()) host_dump = host_dump || !c->on_device();
    if (fused && !host_dump) {
        run_fused(system, thermo);
    } else {
        Dump dump;
        if (system.run_p.write_dump_flag) dump.preprocess(system, resolve("all.trj"));
        zero_step(system, nl, thermo, &dump);
        simulation(system, nl, thermo, &dump);
        dump.postprocess(system);
    }
    postprocess(system, thermo);
    system.input = input;
    if (system.run_p.write_data_flag) write_data(system, resolve(system.run_p.out_data_name));
}

void run_fused(System &system, Thermo &thermo)
{
    // thermo rows and fix/compute records wait in device rings of GDPD_RECORD_CAP
    // (thermo: 1024) entries: advance in chunks that fill at most 1000 of them,
    // and keep the rows of steps not seen yet after each chunk
    const unsigned nsteps = system.run_p.nsteps;
    unsigned span = system.run_p.log_f;
    for (Compute *c : computeOrder())
        if (c->on_device()) span = std::min(span, c->frequency);
    for (GroupFix *g : fixOrder()) {
        auto *sp = dynamic_cast<spring *>(g);
        if (sp && sp->frequency) span = std::min(span, sp->frequency);
    }
    const unsigned chunk = std::max(1u, span * 1000u);
    double last_row = -1.0;
    std::vector<double> last_rec(GDPD_MAX_COMPUTES + GDPD_MAX_FIXES, -1.0);
    std::vector<double> buf((size_t)std::max(GDPD_RECORD_CAP * GDPD_RECORD_LEN, 1024 * GDPD_THERMO_LEN));
    auto collect = [&]() {
        int n = 0;
        check(system, gdpd_thermo_rows(system.ctx, system.handle, buf.data(), 1024, &n));
        for (int k = 0; k < n; ++k) {
            const double *o = buf.data() + (size_t)k * GDPD_THERMO_LEN;
            if (o[GDPD_THERMO_STEP] <= last_row) continue;
            last_row = o[GDPD_THERMO_STEP];
            std::vector<double> r = {o[GDPD_THERMO_STEP], o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], o[8],
                                     o[9], o[GDPD_THERMO_ANGLE_PE]};
            thermo.rows.push_back(r);
            thermo.write_row(system, r);
        }
        auto records = [&](int slot, auto emit) {
            check(system, gdpd_records(system.ctx, system.handle, slot, buf.data(), GDPD_RECORD_CAP, &n));
            for (int k = 0; k < n; ++k) {
                const double *o = buf.data() + (size_t)k * GDPD_RECORD_LEN;
                if (o[0] <= last_rec[slot]) continue;
                last_rec[slot] = o[0];
                emit((unsigned)o[0], o + 1);
            }
        };
        for (Compute *c : computeOrder())
            if (c->on_device()) records(c->index, [&](unsigned st, const double *v) { c->write_row(st, v); });
        for (GroupFix *g : fixOrder())
            if (auto *sp = dynamic_cast<spring *>(g))
                if (sp->frequency)
                    records(GDPD_MAX_COMPUTES + g->index, [&](unsigned st, const double *v) { sp->write_row(st, v); });
    };
    thermo.rows.clear();
    if (nsteps == 0) {
        check(system, gdpd_run(system.ctx, system.handle, 0));
        collect();
        return;
    }
    check(system, gdpd_zero_step(system.ctx, system.handle));
    for (unsigned done = 1; done < nsteps;) {
        const unsigned k = std::min(chunk, nsteps - done);
        check(system, gdpd_advance(system.ctx, system.handle, (int)k));
        collect();
        done += k;
    }
    check(system, gdpd_finish(system.ctx, system.handle));
    collect();
}

} // namespace gdpd_host
