// nw_fields.hpp -- field snapshot files (Hdf5Writer's layout without HDF5).
#pragma once

#include <fstream>
#include <string>
#include <utility>
#include <vector>

#include "nw_mesh.hpp"

namespace nw {

// Which point fields each .vtu carries (Hdf5Writer::FieldSaveOptions,
// hdf5_writer.hpp:56-67, without the derived curl/div fields).
enum { kSaveE = 1, kSaveH = 2, kSaveMagE = 4, kSaveMagH = 8, kSaveEnergyDensity = 16 };

struct FieldWriter {
    int n1 = 0, nelt = 0, save = kSaveE | kSaveH;
    long long npts = 0, ncg = 0;
    bool cont = false, opened = false;
    std::string dir, stem;
    std::vector<double> xyz;              // node coordinates, 3 per node
    std::vector<int> dg2cg;               // continuous export: point of each node
    std::vector<double> cg_xyz, cg_inv;   // continuous points and 1 / their multiplicity
    std::vector<int> cells;               // 8 point indices per linear hexahedron
    std::vector<std::pair<int, double>> steps;
    std::ofstream bin;

    // Hdf5Writer::initialize (hdf5_writer.cpp:482-733): file names from the .h5
    // path, coordinates, continuous points, cells, the .bin header.  0 on success.
    int open(const std::string& h5_path, const Mesh& m, bool continuous, int flags);
    // Hdf5Writer::writeStep (hdf5_writer.cpp:736-1036): one .bin record, the
    // step's .vtu and the .pvd rewritten; state in the uploadState layout.
    int write(int step, double time, const double* state);
    void close();
    ~FieldWriter() { close(); }

  private:
    void group_coincident_nodes();
    void build_cells();
    int write_vtu(int step, const double* st) const;
    int write_pvd() const;
};

}  // namespace nw
