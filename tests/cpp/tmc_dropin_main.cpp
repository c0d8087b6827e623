// The reference's topologically correct MC driver (mc/tmc.cpp) written
// against include/tmc_dmc.hpp: tmc::MarchingCubes on a volume file, OBJ and
// OFF out.  Usage: tmc_dropin <volume> <iso> <c_flag> <obj> <off>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "tmc_dmc.hpp"

int main(int argc, char** argv) {
    if (argc != 6) return 2;
    tmc::MarchingCubes mc;
    mc(std::atof(argv[2]), argv[1], std::atoi(argv[3]) != 0, true, argv[4], true, argv[5]);
    std::printf("{\"nv\": %zu, \"nt\": %zu}\n", mc.points().size(), mc.triangles().size());
    return 0;
}
