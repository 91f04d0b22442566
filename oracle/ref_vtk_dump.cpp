// ref_vtk_dump.cpp -- test infrastructure only.
//
// Links the reference's own src/hdf5_writer.cpp, src/mesh.cpp and src/comm.cpp
// (compiled in place from the reference tree by oracle/Makefile into
// oracle/_ref/, without the HDF5 library: the writer's VTU/PVD/.bin path) and
// writes the snapshot files of a box mesh for the states given, so the
// product's field writer can be compared with the reference's file by file.
//
//   ref_vtk_dump <n1> <nelx> <nely> <nelz> <xmin> <xmax> <ymin> <ymax> <zmin> <zmax>
//                <px> <py> <pz> <continuous> <flags> <states.bin> <out.h5>
//
// flags: 1 E, 2 H, 4 magnitude_E, 8 magnitude_H, 16 energy_density.
// states.bin (little-endian): int32 count; per state int32 step, float64 time,
// float64 state[6 npts] in the uploadState layout.
#include "hdf5_writer.hpp"
#include "mesh.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char** argv) {
    if (argc != 18) return 2;
    const int n1 = std::atoi(argv[1]);
    Mesh mesh(n1, 1);
    mesh.setN(n1);
    mesh.createBoxMesh(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atof(argv[5]),
                       std::atof(argv[6]), std::atof(argv[7]), std::atof(argv[8]), std::atof(argv[9]),
                       std::atof(argv[10]), std::atoi(argv[11]) != 0, std::atoi(argv[12]) != 0,
                       std::atoi(argv[13]) != 0);
    const bool continuous = std::atoi(argv[14]) != 0;
    const int flags = std::atoi(argv[15]);
    Hdf5Writer::FieldSaveOptions opt;
    opt.saveE = flags & 1;
    opt.saveH = flags & 2;
    opt.saveMagnitudeE = flags & 4;
    opt.saveMagnitudeH = flags & 8;
    opt.saveEnergyDensity = flags & 16;
    FILE* f = std::fopen(argv[16], "rb");
    if (!f) return 4;
    int count = 0;
    if (std::fread(&count, sizeof(int), 1, f) != 1) return 5;
    Hdf5Writer w;
    w.setFieldSaveOptions(opt);
    if (!w.initialize(argv[17], mesh, true, continuous)) return 6;
    const int npts = mesh.getTotalPoints();
    std::vector<double> st(6 * size_t(npts));
    for (int r = 0; r < count; ++r) {
        int step = 0;
        double time = 0.0;
        if (std::fread(&step, sizeof(int), 1, f) != 1 || std::fread(&time, sizeof(double), 1, f) != 1 ||
            std::fread(st.data(), sizeof(double), st.size(), f) != st.size())
            return 5;
        if (!w.writeStep(step, time, st.data(), npts, nullptr)) return 7;
    }
    std::fclose(f);
    w.close();
    return 0;
}
