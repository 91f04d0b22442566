// nw_mesh.hpp -- host mesh description (the data DgSolver::initialize consumes).
#pragma once

#include <vector>

#include "nw_internal.hpp"

namespace nw {

constexpr int NW_FACE_HALO_INTERNAL = 4;

struct Mesh {
    int n1 = 0, np = 0, nelt = 0;
    int box_nel[3] = {0, 0, 0};
    bool box_conforming = false;
    bool all_affine = false;
    bool media = false;
    bool unmatched_is_pec = false;               // E faces with no coincident node become PEC (file readers)
    std::vector<double> corners;                 // nelt * 24
    std::vector<int> kind, nbr_e, nbr_f;         // nelt * 6 (FaceInfo list)
    std::vector<int> fcode, fbc;                 // nelt * 6 (neighbour code, reference bc code)
    std::vector<ElemGeom> geom;
    std::vector<unsigned char> affine;
    std::vector<ElemMat> mat;
    std::vector<double> eps, mu_;                // per element when media
    int nghost_faces = 0;
    std::vector<int> ghost_elem;                 // per ghost face: element owning the partner face
    std::vector<ElemMat> gmat;                   // per ghost face
    std::vector<int> send_vol;                   // per ghost slot: node the pack copies

    static Mesh* box(int n1, int nelx, int nely, int nelz, const double* bounds, int px, int py, int pz,
                     double jitter, uint64_t seed, int* status);
    static Mesh* from_corners(int n1, int nelt, const double* corners, const int* kind, const int* ne,
                              const int* nf, int* status);
    static Mesh* load(const char* path, int n1, int flags, int* status);  // nw_meshio.cpp
    int finish();
    int set_materials(const double* eps, const double* mu);
    int self_halo_x();
    long long npts() const { return (long long)nelt * np; }
    long long tfp() const { return (long long)nelt * 6 * n1 * n1; }
    long long nhalo_points() const;
    void coords(double* x, double* y, double* z) const;
    void jacobian(double* jac) const;
    double min_node_distance() const;
};

// Probe placement of ProbeManager::init (probe.cpp:42-103): per target point,
// the element whose axis-aligned corner box (corners 0, 1, 2, 4) is nearest --
// the first one on ties -- and the point's reference coordinates in it,
// clamped to [-1, 1].
void locate_probes(const Mesh& m, int n, const double* xyz, int* elem, double* rst);
// 1-D Lagrange basis through the GLL nodes at xi (probe.cpp:10-27): the unit
// vector of a node within 1e-12 of xi, else the product formula.
void lagrange_1d(int n1, const double* z, double xi, double* L);

}  // namespace nw

struct nw_mesh {
    nw::Mesh m;
};
