// Runs a reference deck (run.in + read_data file) through the C++ shells the
// way main/main.cu:15-72 does.  Usage:
//   deck_runner run.in [fused|hook] [f32|f64]
// Prints one "thermo <step> <T> <pair PE> <bond PE> <P> <angle PE>" line per thermo row.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "gdpd_plugins.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: deck_runner run.in [fused|hook] [f32|f64]\n");
        return 2;
    }
    const std::string deck_path = argv[1];
    const bool fused = argc < 3 || std::string(argv[2]) != "hook";
    std::ifstream in(deck_path);
    if (!in) {
        std::fprintf(stderr, "cannot open input file: %s\n", deck_path.c_str());
        return 1;
    }
    std::stringstream text;
    text << in.rdbuf();
    const size_t slash = deck_path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "" : deck_path.substr(0, slash);
    gdpd_host::System s;
    s.precision = (argc > 3 && std::string(argv[3]) == "f64") ? GDPD_PREC_F64 : GDPD_PREC_F32;
    gdpd_host::Thermo thermo;
    try {
        gdpd_host::run_deck(s, thermo, text.str(), dir, fused);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        gdpd_host::release(s);
        return 1;
    }
    for (const auto &r : thermo.rows)
        std::printf("thermo %.0f %.17g %.17g %.17g %.17g %.17g\n", r[0], r[1], r[2], r[3], r[4], r[11]);
    gdpd_host::postprocessPairs();
    gdpd_host::release(s);
    return 0;
}
