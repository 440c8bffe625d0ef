// Drives the C++ plugin shells (include/gdpd_plugins.hpp) the way the
// reference's main() drives its plugins (main/main.cu:15-72): a run.in deck,
// preprocess, zero_step, simulation.  The same system is then run again through
// the fused path and the two final states are compared bit for bit.
// Prints one line per thermo row and a final "max|dx| max|dv|" line.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <vector>

#include "gdpd_plugins.hpp"

using namespace gdpd_host;

static void make_atoms(System &s, int n, double L, uint64_t seed)
{
    s.n_atoms = n;
    s.n_atomtypes = 1;
    s.box.xlo = s.box.ylo = s.box.zlo = 0.0;
    s.box.xhi = s.box.yhi = s.box.zhi = L;
    s.masses = {1.0};
    s.h_id.resize(n);
    s.h_type.assign(n, 1);
    s.h_pos.resize(3 * (size_t)n);
    s.h_vel.resize(3 * (size_t)n);
    uint64_t x = seed;
    auto u = [&]() {
        x = x * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int i = 0; i < n; ++i) {
        s.h_id[i] = n - i;  // ids in reverse order: upload must key by id
        for (int d = 0; d < 3; ++d) {
            s.h_pos[3 * i + d] = u() * L;
            s.h_vel[3 * i + d] = (u() - 0.5) * 2.0;
        }
    }
}

int main(int argc, char **argv)
{
    const int nsteps = argc > 1 ? std::atoi(argv[1]) : 20;
    const char *deck_fmt =
        "# C0-style deck\n"
        "atom_style atomic\n"
        "neighbor 1.0 0.3 6.0 1\n"
        "timestep 0.04\n"
        "thermo 5\n"
        "run %d\n"
        "pair_style dpd 1.0 1.0 42\n"
        "pair_coeff 1 1 25.0 4.5\n"
        "fix md all nve\n";
    char deck[512];
    std::snprintf(deck, sizeof(deck), deck_fmt, nsteps);
    try {
        System a, b;
        for (System *s : {&a, &b}) {
            read_input(*s, deck);
            s->precision = GDPD_PREC_F64;
            make_atoms(*s, 3000, 10.0, 77);
            atoms_to_device(*s);
        }
        Neigh_list nl;
        Thermo th;
        preprocess(a, nl, th);
        zero_step(a, nl, th);
        simulation(a, nl, th);
        for (auto &r : th.rows) std::printf("hook %g %.12g %.12g %.12g\n", r[0], r[1], r[2], r[4]);
        Thermo tf;
        preprocess(b, nl, tf);
        run_fused(b, tf);
        for (auto &r : tf.rows) std::printf("fused %g %.12g %.12g %.12g\n", r[0], r[1], r[2], r[4]);
        atoms_from_device(a);
        atoms_from_device(b);
        double dx = 0, dv = 0;
        for (size_t k = 0; k < a.h_pos.size(); ++k) {
            dx = std::fmax(dx, std::fabs(a.h_pos[k] - b.h_pos[k]));
            dv = std::fmax(dv, std::fabs(a.h_vel[k] - b.h_vel[k]));
        }
        std::printf("final %.3e %.3e\n", dx, dv);
        release(a);
        release(b);
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
