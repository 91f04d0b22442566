// nw_fields.cpp -- field snapshots in the files Hdf5Writer writes when the HDF5
// library is absent (hdf5_writer.cpp:98-119 the .pvd collection, :121-470 one
// .vtu per saved step, :711-733 and :1002-1033 the IEEE-754 .bin archive), so
// ParaView and the reference's post-processing read them unchanged.  The text
// is produced by a std::ofstream with its default formatting, as the
// reference's is (6 significant digits in the .vtu, scientific with 8 digits
// for the .pvd time stamps).  Derived fields (curl, div) need the reference's
// host curl (case.cpp:830-971) and are not offered.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <string>
#include <utility>
#include <vector>

#include "nw_fields.hpp"

namespace nw {

int FieldWriter::open(const std::string& h5_path, const Mesh& m, bool continuous, int flags) {
    close();
    n1 = m.n1;
    nelt = m.nelt;
    npts = m.npts();
    cont = continuous;
    save = flags;
    steps.clear();
    // directory and stem of the .h5 path (hdf5_writer.cpp:492-505): the files are
    // <dir>/<stem>.pvd, <dir>/<stem>.bin and <dir>/<stem>_step_<n>.vtu
    const size_t slash = h5_path.find_last_of("/\\");
    dir = slash == std::string::npos ? "." : h5_path.substr(0, slash);
    stem = slash == std::string::npos ? h5_path : h5_path.substr(slash + 1);
    const size_t dot = stem.find_last_of('.');
    if (dot != std::string::npos) stem = stem.substr(0, dot);

    std::vector<double> x(npts), y(npts), z(npts);
    m.coords(x.data(), y.data(), z.data());
    xyz.resize(3 * npts);
    for (long long i = 0; i < npts; ++i) {
        xyz[3 * i] = x[i];
        xyz[3 * i + 1] = y[i];
        xyz[3 * i + 2] = z[i];
    }
    if (cont) group_coincident_nodes();
    else ncg = npts;
    build_cells();

    bin.open(dir + "/" + stem + ".bin", std::ios::binary | std::ios::trunc);
    if (!bin.is_open()) return -1;
    char magic[16] = {};
    const char tag[] = "NEKWAVE_BIN_V2";
    std::copy(tag, tag + sizeof(tag), magic);
    bin.write(magic, 16);
    const int32_t hdr[4] = {n1, nelt, static_cast<int32_t>(npts), 0};
    bin.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    bin.write(reinterpret_cast<const char*>(xyz.data()), std::streamsize(xyz.size() * sizeof(double)));
    bin.flush();
    opened = true;
    return bin.good() ? 0 : -1;
}

// Coincident DG nodes become one continuous point (hdf5_writer.cpp:538-591):
// nodes sorted by x, then every node not yet placed opens a point and takes the
// later nodes within 1e-6 in each coordinate and in distance.  The sort and the
// sweep are the reference's, so points are numbered as it numbers them.
void FieldWriter::group_coincident_nodes() {
    struct Ref {
        int idx;
        double x, y, z;
    };
    std::vector<Ref> pts(npts);
    for (long long i = 0; i < npts; ++i)
        pts[i] = {static_cast<int>(i), xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    std::sort(pts.begin(), pts.end(), [](const Ref& a, const Ref& b) { return a.x < b.x; });
    dg2cg.assign(npts, -1);
    cg_xyz.clear();
    std::vector<int> count;
    const double tol = 1e-6, tol2 = tol * tol;
    for (size_t s = 0; s < pts.size(); ++s) {
        if (dg2cg[pts[s].idx] != -1) continue;
        const int id = static_cast<int>(count.size());
        dg2cg[pts[s].idx] = id;
        cg_xyz.push_back(pts[s].x);
        cg_xyz.push_back(pts[s].y);
        cg_xyz.push_back(pts[s].z);
        count.push_back(1);
        for (size_t f = s + 1; f < pts.size(); ++f) {
            if (pts[f].x - pts[s].x > tol) break;
            if (dg2cg[pts[f].idx] != -1) continue;
            const double dy = pts[f].y - pts[s].y, dz = pts[f].z - pts[s].z;
            if (std::abs(dy) <= tol && std::abs(dz) <= tol) {
                const double dx = pts[f].x - pts[s].x;
                if (dx * dx + dy * dy + dz * dz <= tol2) {
                    dg2cg[pts[f].idx] = id;
                    count[id]++;
                }
            }
        }
    }
    ncg = static_cast<long long>(count.size());
    cg_inv.resize(ncg);
    for (long long k = 0; k < ncg; ++k) cg_inv[k] = 1.0 / static_cast<double>(count[k]);
}

// (n1 - 1)^3 linear hexahedra per element between neighbouring GLL nodes, VTK
// corner order (hdf5_writer.cpp:596-632).
void FieldWriter::build_cells() {
    const int p = n1 - 1, np = n1 * n1 * n1;
    cells.clear();
    cells.reserve(size_t(nelt) * p * p * p * 8);
    auto node = [&](int e, int i, int j, int k) {
        const long long v = (long long)e * np + i + n1 * (j + n1 * k);
        return cont ? dg2cg[v] : static_cast<int>(v);
    };
    for (int e = 0; e < nelt; ++e)
        for (int k = 0; k < p; ++k)
            for (int j = 0; j < p; ++j)
                for (int i = 0; i < p; ++i) {
                    const int c[8] = {node(e, i, j, k),         node(e, i + 1, j, k),
                                      node(e, i + 1, j + 1, k), node(e, i, j + 1, k),
                                      node(e, i, j, k + 1),     node(e, i + 1, j, k + 1),
                                      node(e, i + 1, j + 1, k + 1), node(e, i, j + 1, k + 1)};
                    cells.insert(cells.end(), c, c + 8);
                }
}

int FieldWriter::write(int step, double time, const double* st) {
    if (!opened) return -1;
    steps.emplace_back(step, time);
    // archive record: int32 step, 0; float64 time; E and H interleaved per node; no derived fields
    const int32_t sh[2] = {step, 0};
    bin.write(reinterpret_cast<const char*>(sh), sizeof(sh));
    bin.write(reinterpret_cast<const char*>(&time), sizeof(double));
    std::vector<double> buf(3 * npts);
    for (int h = 0; h < 2; ++h) {
        for (long long i = 0; i < npts; ++i)
            for (int c = 0; c < 3; ++c) buf[3 * i + c] = st[(3 * h + c) * npts + i];
        bin.write(reinterpret_cast<const char*>(buf.data()), std::streamsize(buf.size() * sizeof(double)));
    }
    const int32_t derived = 0;
    bin.write(reinterpret_cast<const char*>(&derived), sizeof(derived));
    bin.flush();
    if (write_vtu(step, st)) return -1;
    return write_pvd() || !bin.good() ? -1 : 0;
}

int FieldWriter::write_pvd() const {
    std::ofstream pvd(dir + "/" + stem + ".pvd");
    if (!pvd.is_open()) return -1;
    pvd << "<?xml version=\"1.0\"?>\n";
    pvd << "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    pvd << "  <Collection>\n";
    for (const auto& s : steps)
        pvd << "    <DataSet timestep=\"" << std::scientific << std::setprecision(8) << s.second
            << "\" group=\"\" part=\"0\" file=\"" << stem << "_step_" << s.first << ".vtu\"/>\n";
    pvd << "  </Collection>\n";
    pvd << "</VTKFile>\n";
    return pvd.good() ? 0 : -1;
}

int FieldWriter::write_vtu(int step, const double* st) const {
    std::ofstream vtu(dir + "/" + stem + "_step_" + std::to_string(step) + ".vtu");
    if (!vtu.is_open()) return -1;
    const long long ncell = (long long)nelt * (n1 - 1) * (n1 - 1) * (n1 - 1);
    const long long nout = cont ? ncg : npts;
    // the six fields at the output points: the DG nodes, or the averages over
    // each continuous point's coincident nodes.  As in the reference
    // (hdf5_writer.cpp:257-275), the continuous averages of E (of H) are formed
    // only when E (H) is saved, so magnitudes and the energy density of a field
    // that is not saved come out as zeros there.
    std::vector<double> f(6 * nout, 0.0);
    if (cont) {
        const int c0 = (save & kSaveE) ? 0 : 3, c1 = (save & kSaveH) ? 6 : 3;
        for (long long i = 0; i < npts; ++i)
            for (int c = c0; c < c1; ++c) f[c * nout + dg2cg[i]] += st[c * npts + i];
        for (long long k = 0; k < nout; ++k)
            for (int c = 0; c < 6; ++c) f[c * nout + k] *= cg_inv[k];
    } else {
        std::copy(st, st + 6 * npts, f.begin());
    }
    const std::vector<double>& pts = cont ? cg_xyz : xyz;

    vtu << "<?xml version=\"1.0\"?>\n";
    vtu << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    vtu << "  <UnstructuredGrid>\n";
    vtu << "    <Piece NumberOfPoints=\"" << nout << "\" NumberOfCells=\"" << ncell << "\">\n";
    vtu << "      <Points>\n";
    vtu << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (long long i = 0; i < nout; ++i) vtu << pts[3 * i] << " " << pts[3 * i + 1] << " " << pts[3 * i + 2] << "\n";
    vtu << "        </DataArray>\n";
    vtu << "      </Points>\n";

    const bool sE = save & kSaveE, sH = save & kSaveH;
    const std::string vec = sE ? "E" : (sH ? "H" : "");
    const std::string sca = (save & kSaveMagE) ? "magnitude_E" : ((save & kSaveEnergyDensity) ? "energy_density" : "");
    vtu << "      <PointData";
    if (!vec.empty()) vtu << " Vectors=\"" << vec << "\"";
    if (!sca.empty()) vtu << " Scalars=\"" << sca << "\"";
    vtu << ">\n";
    auto vector3 = [&](const char* name, int c0) {
        vtu << "        <DataArray type=\"Float64\" Name=\"" << name << "\" NumberOfComponents=\"3\" format=\"ascii\">\n";
        for (long long i = 0; i < nout; ++i)
            vtu << f[c0 * nout + i] << " " << f[(c0 + 1) * nout + i] << " " << f[(c0 + 2) * nout + i] << "\n";
        vtu << "        </DataArray>\n";
    };
    auto scalar = [&](const char* name, auto&& value) {
        vtu << "        <DataArray type=\"Float64\" Name=\"" << name << "\" NumberOfComponents=\"1\" format=\"ascii\">\n";
        for (long long i = 0; i < nout; ++i) vtu << value(i) << "\n";
        vtu << "        </DataArray>\n";
    };
    auto at = [&](int c, long long i) { return f[c * nout + i]; };
    if (sE) vector3("E", 0);
    if (sH) vector3("H", 3);
    if (save & kSaveMagE)
        scalar("magnitude_E", [&](long long i) {
            const double ex = at(0, i), ey = at(1, i), ez = at(2, i);
            return std::sqrt(ex * ex + ey * ey + ez * ez);
        });
    if (save & kSaveMagH)
        scalar("magnitude_H", [&](long long i) {
            const double hx = at(3, i), hy = at(4, i), hz = at(5, i);
            return std::sqrt(hx * hx + hy * hy + hz * hz);
        });
    if (save & kSaveEnergyDensity)
        scalar("energy_density", [&](long long i) {
            const double ex = at(0, i), ey = at(1, i), ez = at(2, i), hx = at(3, i), hy = at(4, i), hz = at(5, i);
            return 0.5 * (ex * ex + ey * ey + ez * ez + hx * hx + hy * hy + hz * hz);
        });
    vtu << "      </PointData>\n";

    vtu << "      <Cells>\n";
    vtu << "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    for (long long c = 0; c < ncell; ++c) {
        const int* v = &cells[8 * c];
        vtu << v[0] << " " << v[1] << " " << v[2] << " " << v[3] << " " << v[4] << " " << v[5] << " " << v[6] << " "
            << v[7] << "\n";
    }
    vtu << "        </DataArray>\n";
    vtu << "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    for (long long c = 0; c < ncell; ++c) vtu << 8 * (c + 1) << (c + 1 == ncell ? "" : " ");
    vtu << "\n        </DataArray>\n";
    vtu << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    for (long long c = 0; c < ncell; ++c) vtu << 12 << (c + 1 == ncell ? "" : " ");  // VTK_HEXAHEDRON
    vtu << "\n        </DataArray>\n";
    vtu << "      </Cells>\n";
    vtu << "    </Piece>\n";
    vtu << "  </UnstructuredGrid>\n";
    vtu << "</VTKFile>\n";
    return vtu.good() ? 0 : -1;
}

void FieldWriter::close() {
    if (!opened) return;
    bin.close();
    write_pvd();
    opened = false;
}

}  // namespace nw
