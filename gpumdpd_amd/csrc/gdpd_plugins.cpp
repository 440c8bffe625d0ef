// gdpd_plugins.cpp -- reference plugin interfaces over the libgdpd C ABI.
//
// Each class keeps the reference's method names and argument meaning; the
// device work is one gdpd_* call per hook.  Deck parsing follows the
// reference's commands (main/read_in.cu:40-290, force/pair_dpd.cu:19-173,
// force/pair_mdpd.cu:21-192, force/bond_harmonic.cu:19-147, fix/fix.cu:60-115).
#include "../../include/gdpd_plugins.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <iostream>
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
            else illegal(l, "illegal atom_style command (atomic or bond)", fmt);
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
        }
    });
    if (s.run_p.log_f == 0) throw Error("thermo frequency must be positive");
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
    s.h_pos.resize(3 * (size_t)s.n_atoms);
    s.h_vel.resize(3 * (size_t)s.n_atoms);
    if (force) force->resize(3 * (size_t)s.n_atoms);
    check(s, gdpd_download(s.ctx, s.handle, s.h_pos.data(), s.h_vel.data(), nullptr,
                           force ? force->data() : nullptr));
    // host arrays are id-ordered after a download
    for (int i = 0; i < s.n_atoms; ++i) s.h_id[i] = i + 1;
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

void preprocessFixes(System &system)
{
    getEnabledFixRegistry().clear();
    for_lines(system.input, [&](const std::string &, const std::string &key, const std::vector<std::string> &a) {
        if (key != "fix") return;
        if (a.size() < 3) throw Error("illegal fix command: fix ID groupID style params");
        enableFix(a[0], a[1], a[2], std::vector<std::string>(a.begin() + 3, a.end()), system);
    });
    for (auto &kv : getEnabledFixRegistry()) kv.second->preprocess(system);
}
void postprocessFixes(System &system)
{
    for (auto &kv : getEnabledFixRegistry()) kv.second->postprocess(system);
}
#define GDPD_FIX_HOOK(name)                                                       \
    void Fix_##name(System &system, unsigned int step)                           \
    {                                                                             \
        for (auto &kv : getEnabledFixRegistry()) kv.second->name(system, step);   \
    }
GDPD_FIX_HOOK(initial_integrate)
GDPD_FIX_HOOK(post_integrate)
GDPD_FIX_HOOK(final_integrate)
GDPD_FIX_HOOK(post_force)
GDPD_FIX_HOOK(end_of_step)
#undef GDPD_FIX_HOOK

// ---- fix nve (fix/fix_nve.cu) ------------------------------------------------------

nve::nve(std::string id, std::string gid, const std::vector<std::string> &) : Fix(std::move(id), std::move(gid)) {}

void nve::validateParams(const std::vector<std::string> &params)
{
    if (!params.empty()) illegal("Invalid fix_nve parameters", "illegal fix_nve command", "fix ID groupID nve");
}
void nve::initial_integrate(System &s, unsigned int) { check(s, gdpd_nve_initial(s.ctx, s.handle)); }
void nve::final_integrate(System &s, unsigned int) { check(s, gdpd_nve_final(s.ctx, s.handle)); }

static bool _gdpd_nve_registered = []() {
    registerFixType("nve", [](const std::string &id, const std::string &gid,
                              const std::vector<std::string> &p) -> std::unique_ptr<Fix> {
        return std::make_unique<nve>(id, gid, p);
    });
    return true;
}();

// ---- Neigh_list (force/neigh_list.cu) ----------------------------------------------

void Neigh_list::preprocess(System &) {}
void Neigh_list::postprocess(System &) {}
void Neigh_list::build(System &s) { check(s, gdpd_neigh_build(s.ctx, s.handle)); }
void Neigh_list::check_update(System &s) { check(s, gdpd_neigh_check(s.ctx, s.handle, &s.h_update_flag)); }

// ---- Thermo (compute/thermo.cu:96-220) -------------------------------------------

void Thermo::preprocess(System &) { rows.clear(); }
void Thermo::postprocess(System &) {}

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
    std::vector<double> r = {(double)step, o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], o[8], o[9]};
    rows.push_back(r);
    if (print)
        std::printf("%15u\t%10.6f\t%20.6f\t%20.6f\t%12.6f\n", step, o[0], o[1], o[2], o[3]);
}

// ---- driver (main/utils.cuh:126-314) -------------------------------------------------

void preprocess(System &system, Neigh_list &nl, Thermo &thermo)
{
    nl.preprocess(system);
    preprocessPairs(system);
    if (system.atom_style == 1) preprocessBonds(system);
    preprocessFixes(system);
    thermo.preprocess(system);
}

void zero_step(System &system, Neigh_list &nl, Thermo &thermo)
{
    nl.build(system);
    Pair_compute(system, 0);
    if (system.atom_style == 1) Bond_compute(system, 0);
    Fix_post_force(system, 0);
    thermo.process(system, 0);
}

void simulation(System &system, Neigh_list &nl, Thermo &thermo)
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
        if (system.atom_style == 1) Bond_compute(system, step);
        Fix_post_force(system, step);
        Fix_final_integrate(system, step);
        Fix_end_of_step(system, step);
        thermo.process(system, step);
    }
}

void run_fused(System &system, Thermo &thermo)
{
    check(system, gdpd_run(system.ctx, system.handle, (int)system.run_p.nsteps));
    std::vector<double> rows((size_t)GDPD_THERMO_LEN * 4096);
    int n = 0;
    check(system, gdpd_thermo_rows(system.ctx, system.handle, rows.data(), 4096, &n));
    thermo.rows.clear();
    for (int k = 0; k < n; ++k) {
        const double *o = rows.data() + (size_t)k * GDPD_THERMO_LEN;
        thermo.rows.push_back({o[GDPD_THERMO_STEP], o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], o[8], o[9]});
    }
}

} // namespace gdpd_host
