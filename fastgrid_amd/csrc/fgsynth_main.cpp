// fgsynth — writes a seeded synthetic receptor (<stem>.pdbqt) and its GPF
// (<stem>.gpf) into a directory.
//   fgsynth DIR STEM SEED N_ATOMS sphere R NX NY NZ "LIGAND TYPES"
//   fgsynth DIR STEM SEED N_ATOMS box A B C NX NY NZ "LIGAND TYPES"
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "fgmaps.h"

int main(int argc, char **argv)
{
    if (argc < 10) {
        std::fprintf(stderr, "usage: %s DIR STEM SEED N (sphere R | box A B C) NX NY NZ \"TYPES\"\n", argv[0]);
        return 1;
    }
    fg_synth_spec s{};
    s.seed = std::strtoull(argv[3], nullptr, 10);
    s.n_atoms = std::atoi(argv[4]);
    int i = 6;
    if (!std::strcmp(argv[5], "sphere")) {
        s.shape = 0;
        s.a = std::atof(argv[i++]);
    } else {
        s.shape = 1;
        s.a = std::atof(argv[i++]);
        s.b = std::atof(argv[i++]);
        s.c = std::atof(argv[i++]);
    }
    if (argc < i + 4) {
        std::fprintf(stderr, "%s: missing grid size or ligand types\n", argv[0]);
        return 1;
    }
    for (int k = 0; k < 3; ++k) s.npts[k] = std::atoi(argv[i++]);
    s.ligand_types = argv[i];
    s.min_sep = 1.2;
    s.spacing = 0.375;
    s.stem = argv[2];
    if (fg_synth_write(&s, argv[1]) != FG_OK) {
        std::fprintf(stderr, "%s: %s\n", argv[0], fg_last_error());
        return 1;
    }
    return 0;
}
