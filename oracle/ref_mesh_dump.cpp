// ref_mesh_dump.cpp -- test infrastructure only.
//
// Links the reference's own src/mesh.cpp (compiled in place from the reference
// tree by oracle/Makefile into oracle/_ref/) and dumps every array that
// DgSolver::initialize (dg_solver.cpp:40-241) uploads, so the oracle's
// restatement of the geometry can be checked against the reference bit for bit.
//
//   ref_mesh_dump box <n1> <nelx> <nely> <nelz> <xmin> <xmax> <ymin> <ymax> <zmin> <zmax> <px> <py> <pz> <out>
//   ref_mesh_dump rea <n1> <file.rea> <out>
//   ref_mesh_dump re2 <n1> <file.re2> <out>
//
// Output (little-endian): int32 n1, nelt, npts, tfp; then float64 z[n1], w[n1],
// D[n1^2], Dt[n1^2], w3[n1^3], corners[nelt*24], x, y, z, jac, rx, sx, tx, ry,
// sy, ty, rz, sz, tz [npts each]; int32 vM, vP, bc [tfp]; float64 nx, ny, nz,
// dA [tfp]; float64 dxmin, autoCfl(0.4), autoDt(1.0, 0.4).
#include "mesh.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static void wd(FILE* f, const std::vector<double>& v) { fwrite(v.data(), sizeof(double), v.size(), f); }
static void wi(FILE* f, const std::vector<int>& v) { fwrite(v.data(), sizeof(int), v.size(), f); }

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::string mode = argv[1];
    int n1 = std::atoi(argv[2]);
    Mesh mesh(n1, 1);
    std::string out;
    if (mode == "box" && argc == 16) {
        mesh.setN(n1);
        mesh.createBoxMesh(std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
                           std::atof(argv[6]), std::atof(argv[7]), std::atof(argv[8]),
                           std::atof(argv[9]), std::atof(argv[10]), std::atof(argv[11]),
                           std::atoi(argv[12]) != 0, std::atoi(argv[13]) != 0, std::atoi(argv[14]) != 0);
        out = argv[15];
    } else if (mode == "rea" && argc == 5) {
        if (!mesh.loadFromRea(argv[3])) return 3;
        out = argv[4];
    } else if (mode == "re2" && argc == 5) {
        if (!mesh.loadFromRe2(argv[3])) return 3;
        out = argv[4];
    } else {
        return 2;
    }
    FILE* f = std::fopen(out.c_str(), "wb");
    if (!f) return 4;
    int hdr[4] = {mesh.getN(), mesh.getNumElements(), mesh.getTotalPoints(), 0};
    const auto& fd = mesh.getFaceData();
    for (const auto& x : fd) hdr[3] += static_cast<int>(x.points.size());
    fwrite(hdr, sizeof(int), 4, f);
    wd(f, mesh.getGllZ()); wd(f, mesh.getGllW()); wd(f, mesh.getD()); wd(f, mesh.getDt()); wd(f, mesh.getW3());
    std::vector<double> corners;
    for (const auto& e : mesh.getElementCorners())
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c) corners.push_back(e[a][c]);
    wd(f, corners);
    wd(f, mesh.getCoordX()); wd(f, mesh.getCoordY()); wd(f, mesh.getCoordZ()); wd(f, mesh.getJac());
    wd(f, mesh.getRx()); wd(f, mesh.getSx()); wd(f, mesh.getTx());
    wd(f, mesh.getRy()); wd(f, mesh.getSy()); wd(f, mesh.getTy());
    wd(f, mesh.getRz()); wd(f, mesh.getSz()); wd(f, mesh.getTz());
    // Flattening and bc codes exactly as dg_solver.cpp:90-112 does.
    std::vector<int> vM, vP, bc;
    std::vector<double> nx, ny, nz, dA;
    for (const auto& x : fd) {
        int code = 0;
        if (x.bcType == "PEC" || x.bcType == "W" || x.bcType == "V" || x.bcType == "v") code = 1;
        else if (x.bcType == "PMC" || x.bcType == "SYM" || x.bcType == "S" || x.bcType == "s") code = 2;
        for (const auto& p : x.points) {
            vM.push_back(p.volIdxMinus);
            vP.push_back(p.volIdxPlus);
            int c = code;
            if (p.volIdxPlus < 0 && c == 0) c = 1;
            bc.push_back(c);
            nx.push_back(p.nx); ny.push_back(p.ny); nz.push_back(p.nz); dA.push_back(p.dA);
        }
    }
    wi(f, vM); wi(f, vP); wi(f, bc);
    wd(f, nx); wd(f, ny); wd(f, nz); wd(f, dA);
    std::vector<double> tail = {mesh.computeMinNodeDistance(), mesh.computeAutomaticCFL(0.4),
                                mesh.computeAutomaticDt(1.0, 0.4)};
    wd(f, tail);
    std::fclose(f);
    return 0;
}
