// Command-line front end of lbm_latticegen with the reference generator's
// arguments (L/main.cpp:26-39):
//   latticegen <input> <nodes in x> <xseed> <yseed> <zseed> <output> <1|2>
// reads STARCDfiles/<input>.vrt and .bnd and writes <output>_Node.dat,
// <output>_BC.dat and the mesh listing xtecplot.dat in the working directory.
#include "lbm.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

int main(int argc, char** argv) {
    if (argc != 8) {
        std::printf("Usage: meshgen <input_filename> <NoOfNodes in X dir>  <xseed> <yseed> <zseed> <output filename> "
                    "<1- simple, 2 - complex geometry>\n");
        std::printf("%d\n", argc);
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const std::string in = argv[1], out = argv[6];
    const std::string vrt = "STARCDfiles/" + in + ".vrt", bnd = "STARCDfiles/" + in + ".bnd";
    const double seed[3] = {std::atof(argv[3]), std::atof(argv[4]), std::atof(argv[5])};
    int32_t dims[3] = {0, 0, 0};
    int64_t nlinks = 0;
    const lbm_status st = lbm_latticegen(vrt.c_str(), bnd.c_str(), std::atoi(argv[2]), seed, std::atoi(argv[7]),
                                         (out + "_Node.dat").c_str(), (out + "_BC.dat").c_str(), "xtecplot.dat",
                                         dims, &nlinks);
    if (st != LBM_OK) {
        std::fprintf(stderr, "latticegen: %s\n", lbm_latticegen_last_error());
        return 2;
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Complete:  100%%\nlattice %d x %d x %d, %lld boundary links\nwork took: %gs\n", dims[0], dims[1],
                dims[2], (long long)nlinks, s);
    return 0;
}
