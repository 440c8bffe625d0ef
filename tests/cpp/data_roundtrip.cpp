// Reads a LAMMPS-style data file with the shells' read_data and writes it back
// with write_data_host (host only, no device).  Usage:
//   data_roundtrip atomic|bond|angle in.data out.data
// Exit status 0 and "ok <atoms> <bonds>" on success; 1 and the error message
// on stderr when read_data rejects the file.
#include <cstdio>
#include <string>

#include "gdpd_plugins.hpp"

int main(int argc, char **argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: data_roundtrip atomic|bond|angle in.data out.data\n");
        return 2;
    }
    gdpd_host::System s;
    const std::string style = argv[1];
    s.atom_style = style == "angle" ? 2 : style == "bond" ? 1 : 0;
    try {
        gdpd_host::read_data(s, argv[2]);
        gdpd_host::write_data_host(s, argv[3]);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    std::printf("ok %d %d %d\n", s.n_atoms, s.n_bonds, s.has_velocities ? 1 : 0);
    return 0;
}
