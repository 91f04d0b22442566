// ref_probe_dump.cpp -- test infrastructure only.
//
// Links the reference's own src/mesh.cpp and src/probe.cpp (compiled in place
// from the reference tree by oracle/Makefile into oracle/_ref/) and dumps what
// ProbeManager::init (probe.cpp:33-116) computes for each probe: the element
// offset and the n1^3 interpolation weights, so the oracle's placement and
// weights can be checked against the reference bit for bit.
//
//   ref_probe_dump box <n1> <nelx> <nely> <nelz> <xmin> <xmax> <ymin> <ymax> <zmin> <zmax> <pts> <out>
//   ref_probe_dump rea <n1> <file.rea> <pts> <out>
//
// <pts>: int32 n, then float64 x, y, z per probe.  Output (little-endian): int32
// n, np; then per probe int32 elemOffset and float64 weights[np].
#include "mesh.hpp"
#include "probe.hpp"

#include <array>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    const int n1 = std::atoi(argv[2]);
    Mesh mesh(n1, 1);
    std::string pin, out;
    if (mode == "box" && argc == 14) {
        mesh.setN(n1);
        mesh.createBoxMesh(std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), std::atof(argv[6]),
                           std::atof(argv[7]), std::atof(argv[8]), std::atof(argv[9]), std::atof(argv[10]),
                           std::atof(argv[11]), false, false, false);
        pin = argv[12];
        out = argv[13];
    } else if (mode == "rea" && argc == 6) {
        if (!mesh.loadFromRea(argv[3])) return 3;
        pin = argv[4];
        out = argv[5];
    } else {
        return 2;
    }
    FILE* f = std::fopen(pin.c_str(), "rb");
    if (!f) return 4;
    int n = 0;
    if (std::fread(&n, sizeof(int), 1, f) != 1 || n < 0) return 4;
    std::vector<std::array<double, 3>> pts(n);
    for (int q = 0; q < n; ++q)
        if (std::fread(pts[q].data(), sizeof(double), 3, f) != 3) return 4;
    std::fclose(f);

    ProbeManager pm;
    pm.init(mesh, pts, "none");
    FILE* g = std::fopen(out.c_str(), "wb");
    if (!g) return 5;
    const int np = n1 * n1 * n1;
    const int hdr[2] = {static_cast<int>(pm.getNumProbes()), np};
    std::fwrite(hdr, sizeof(int), 2, g);
    for (const auto& p : pm.getProbes()) {
        std::fwrite(&p.elemOffset, sizeof(int), 1, g);
        std::fwrite(p.interpWeights.data(), sizeof(double), p.interpWeights.size(), g);
    }
    std::fclose(g);
    return 0;
}
