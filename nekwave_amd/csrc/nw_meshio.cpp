// nw_meshio.cpp -- NekCEM / Nek5000 mesh files read into nw::Mesh.
//
// Text .rea (Mesh::loadFromRea, mesh.cpp:317-436) and binary .re2
// (Mesh::loadFromRe2, mesh.cpp:439-636), with the reference loader's reading of
// the face list and its boundary-condition tags:
//   * .rea tags are kept raw; the solver's code table (dg_solver.cpp:92-97)
//     makes "PEC" "W" "V" "v" PEC and "PMC" "SYM" "S" "s" PMC (mirrored as PEC);
//     any other tag with a neighbour is matched by coordinates
//     (mesh.cpp:1048-1070), and nodes that do not match become PEC
//     (dg_solver.cpp:102-104).  A raw "P" (periodic) tag therefore ends up PEC
//     unless NW_MESH_REA_PERIODIC asks for (p, q) identity pairing instead;
//   * .re2 tags map to PERIODIC if they contain 'P', else E if they contain 'E',
//     else PEC if they contain 'W' or "PEC", else PERIODIC when a neighbour is
//     given, else PEC (mesh.cpp:541-549); faces without a record pair with the
//     face whose 4-corner centroid rounds to the same 1e-5 grid key
//     (mesh.cpp:567-626).
// Where the reference would go on with an inconsistent description (a face
// missing from the list, a face that matches partly, curved-side records it
// does not skip), this reader stops with NW_ERR_MESH / NW_ERR_UNSUPPORTED.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "nw_mesh.hpp"

namespace nw {
namespace {

struct FaceRec {
    std::string tag;
    int ne = -1, nf = -1;
    bool set = false;
};

// A hex face's four corners in the cyclic order of the Nek face convention
// (face f: 0 s=-1, 1 r=+1, 2 s=+1, 3 r=-1, 4 t=-1, 5 t=+1); the centroid sums
// them in this order.
const int kFaceCorners[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {0, 3, 2, 1}, {4, 5, 6, 7}};

int solver_bc_code(const std::string& t) {
    if (t == "PEC" || t == "W" || t == "V" || t == "v") return 1;
    if (t == "PMC" || t == "SYM" || t == "S" || t == "s") return 2;
    return 0;
}

// The face kind the solver ends up with for a reference bcType string.
int classify(const std::string& bc, int ne, bool rea_periodic) {
    const int code = solver_bc_code(bc);
    if (code == 1) return NW_FACE_PEC;
    if (code == 2) return NW_FACE_PMC;
    if (ne < 0) return NW_FACE_PEC;
    if (bc == "PERIODIC" || (rea_periodic && bc == "P")) return NW_FACE_PERIODIC;
    return NW_FACE_E;
}

bool open_with_fallback(const std::string& path, std::ifstream& f, std::ios::openmode mode) {
    f.open(path, mode);
    if (f.is_open()) return true;
    f.clear();
    f.open("../" + path, mode);  // mesh.cpp:320-325: runs started from a build/ directory
    return f.is_open();
}

Mesh* assemble(int n1, int nel, std::vector<double>& corners, std::vector<FaceRec>& faces, bool rea_periodic,
               int* status) {
    std::vector<int> kind(size_t(nel) * 6), ne(size_t(nel) * 6), nf(size_t(nel) * 6);
    for (size_t x = 0; x < faces.size(); ++x) {
        const FaceRec& r = faces[x];
        if (!r.set) {
            set_error("mesh file: element " + std::to_string(x / 6 + 1) + " face " + std::to_string(x % 6 + 1) +
                      " has no boundary condition and no neighbour");
            *status = NW_ERR_MESH;
            return nullptr;
        }
        if (r.ne >= nel || (r.ne >= 0 && (r.nf < 0 || r.nf > 5))) {
            set_error("mesh file: neighbour out of range at element " + std::to_string(x / 6 + 1));
            *status = NW_ERR_MESH;
            return nullptr;
        }
        kind[x] = classify(r.tag, r.ne, rea_periodic);
        ne[x] = r.ne;
        nf[x] = r.nf;
    }
    Mesh* m = new Mesh();
    m->n1 = n1;
    m->nelt = nel;
    m->corners.swap(corners);
    m->kind.swap(kind);
    m->nbr_e.swap(ne);
    m->nbr_f.swap(nf);
    m->unmatched_is_pec = true;
    *status = m->finish();
    if (*status != NW_OK) {
        delete m;
        return nullptr;
    }
    return m;
}

Mesh* load_re2(const std::string& path, int n1, int nel, int* status) {
    std::ifstream f;
    if (!open_with_fallback(path, f, std::ios::binary)) {
        set_error("cannot open .re2 file " + path);
        *status = NW_ERR_ARG;
        return nullptr;
    }
    char header[81] = {0};
    float tag = 0.0f;
    f.read(header, 80);
    f.read(reinterpret_cast<char*>(&tag), sizeof(float));
    if (!f || std::fabs(tag - 6.54321f) > 1e-5f) {
        set_error(".re2 header or little-endian tag (6.54321) missing in " + path);
        *status = NW_ERR_MESH;
        return nullptr;
    }
    if (nel <= 0) {  // "#v002 <nel> <ndim> <nelv> ..."
        const std::string h(header, 80);
        const size_t p = h.find_first_not_of(" \t", 5);
        if (p != std::string::npos) std::istringstream(h.substr(p)) >> nel;
    }
    if (nel <= 0) {
        set_error("invalid element count in " + path);
        *status = NW_ERR_MESH;
        return nullptr;
    }
    std::vector<double> corners(size_t(nel) * 24);
    std::vector<double> rec(25);
    for (int e = 0; e < nel; ++e) {
        f.read(reinterpret_cast<char*>(rec.data()), 25 * sizeof(double));  // group, x[8], y[8], z[8]
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c) corners[size_t(e) * 24 + a * 3 + c] = rec[1 + c * 8 + a];
    }
    double ncurv = 0.0, nbcd = 0.0;
    f.read(reinterpret_cast<char*>(&ncurv), sizeof(double));
    f.read(reinterpret_cast<char*>(&nbcd), sizeof(double));
    if (!f) {
        set_error("truncated .re2 file " + path);
        *status = NW_ERR_MESH;
        return nullptr;
    }
    if (ncurv != 0.0) {
        set_error(".re2 curved-side records are not supported");
        *status = NW_ERR_UNSUPPORTED;
        return nullptr;
    }
    std::vector<FaceRec> faces(size_t(nel) * 6);
    const long long nbc = static_cast<long long>(nbcd);
    for (long long b = 0; b < nbc; ++b) {
        double v[7];
        char t[9] = {0};
        f.read(reinterpret_cast<char*>(v), sizeof(v));
        f.read(t, 8);
        if (!f) {
            set_error("truncated .re2 boundary records in " + path);
            *status = NW_ERR_MESH;
            return nullptr;
        }
        const int e = static_cast<int>(v[0]) - 1, fc = static_cast<int>(v[1]) - 1;
        if (e < 0 || e >= nel || fc < 0 || fc > 5) {
            set_error(".re2 boundary record out of range");
            *status = NW_ERR_MESH;
            return nullptr;
        }
        FaceRec r;
        r.ne = v[2] > 0.0 ? static_cast<int>(v[2]) - 1 : -1;
        r.nf = v[3] > 0.0 ? static_cast<int>(v[3]) - 1 : -1;
        const std::string raw(t);
        if (raw.find('P') != std::string::npos) r.tag = "PERIODIC";
        else if (raw.find('E') != std::string::npos) r.tag = "E";
        else if (raw.find('W') != std::string::npos || raw.find("PEC") != std::string::npos) r.tag = "PEC";
        else r.tag = r.ne >= 0 ? "PERIODIC" : "PEC";
        r.set = true;
        faces[size_t(e) * 6 + fc] = r;  // a later record for the same face wins
    }
    // faces without a record: pair by centroid key
    std::map<std::tuple<long long, long long, long long>, int> open_faces;
    for (int e = 0; e < nel; ++e)
        for (int fc = 0; fc < 6; ++fc) {
            const size_t x = size_t(e) * 6 + fc;
            if (faces[x].set) continue;
            double c[3] = {0.0, 0.0, 0.0};
            for (int v = 0; v < 4; ++v)
                for (int d = 0; d < 3; ++d) c[d] += corners[size_t(e) * 24 + kFaceCorners[fc][v] * 3 + d] * 0.25;
            const auto key = std::make_tuple(static_cast<long long>(std::round(c[0] * 100000.0)),
                                             static_cast<long long>(std::round(c[1] * 100000.0)),
                                             static_cast<long long>(std::round(c[2] * 100000.0)));
            auto it = open_faces.find(key);
            if (it == open_faces.end()) {
                open_faces.emplace(key, static_cast<int>(x));
                continue;
            }
            const int y = it->second;
            faces[x].tag = "E";
            faces[x].ne = y / 6;
            faces[x].nf = y % 6;
            faces[x].set = true;
            faces[y].tag = "E";
            faces[y].ne = e;
            faces[y].nf = fc;
            faces[y].set = true;
        }
    return assemble(n1, nel, corners, faces, false, status);
}

Mesh* load_rea(const std::string& path, int n1, bool rea_periodic, int* status) {
    std::ifstream f;
    if (!open_with_fallback(path, f, std::ios::in)) {
        set_error("cannot open .rea file " + path);
        *status = NW_ERR_ARG;
        return nullptr;
    }
    std::string line;
    bool found = false;
    while (std::getline(f, line))
        if (line.find("**MESH DATA**") != std::string::npos) {
            found = true;
            break;
        }
    if (!found) {
        set_error("no **MESH DATA** section in " + path);
        *status = NW_ERR_MESH;
        return nullptr;
    }
    int nel = 0, ndim = 3;
    if (std::getline(f, line)) std::istringstream(line) >> nel >> ndim;
    if (nel < 0) {  // binary mesh beside the .rea (mesh.cpp:353-363)
        std::string re2 = path;
        const size_t dot = re2.rfind('.');
        re2 = (dot != std::string::npos ? re2.substr(0, dot) : re2) + ".re2";
        return load_re2(re2, n1, -nel, status);
    }
    if (nel == 0 || ndim != 3) {
        set_error("invalid NEL/NDIM line after **MESH DATA** in " + path);
        *status = NW_ERR_MESH;
        return nullptr;
    }
    std::vector<double> corners(size_t(nel) * 24);
    for (int e = 0; e < nel; ++e) {
        std::getline(f, line);  // "ELEMENT ... [ ... ] GROUP ..."
        // x1..x4, y1..y4, z1..z4, then x5..x8, y5..y8, z5..z8
        for (int half = 0; half < 2; ++half)
            for (int d = 0; d < 3; ++d) {
                if (!std::getline(f, line)) {
                    set_error("truncated element block in " + path);
                    *status = NW_ERR_MESH;
                    return nullptr;
                }
                std::istringstream ss(line);
                for (int a = 0; a < 4; ++a) {
                    double v;
                    if (!(ss >> v)) {
                        set_error("bad coordinate line for element " + std::to_string(e + 1) + " in " + path);
                        *status = NW_ERR_MESH;
                        return nullptr;
                    }
                    corners[size_t(e) * 24 + (half * 4 + a) * 3 + d] = v;
                }
            }
    }
    found = false;
    while (std::getline(f, line))
        if (line.find("FLUID") != std::string::npos && line.find("BOUNDARY CONDITIONS") != std::string::npos) {
            found = true;
            break;
        }
    if (!found) {
        set_error("no FLUID BOUNDARY CONDITIONS section in " + path);
        *status = NW_ERR_MESH;
        return nullptr;
    }
    std::vector<FaceRec> faces(size_t(nel) * 6);
    for (long long x = 0; x < 6LL * nel; ++x) {
        if (!std::getline(f, line)) break;
        if (line.empty()) continue;  // an empty line uses up one record, as in mesh.cpp:408-410
        std::istringstream ss(line);
        std::string tag;
        int e = 0, fc = 0;
        double nbe = 0.0, nbf = 0.0;
        ss >> tag >> e >> fc >> nbe >> nbf;
        if (!ss || e < 1 || e > nel || fc < 1 || fc > 6) {
            set_error("bad boundary-condition line: \"" + line + "\"");
            *status = NW_ERR_MESH;
            return nullptr;
        }
        FaceRec& r = faces[size_t(e - 1) * 6 + (fc - 1)];
        if (r.set) {
            set_error("face listed twice in boundary conditions: \"" + line + "\"");
            *status = NW_ERR_MESH;
            return nullptr;
        }
        r.tag = tag;
        r.ne = nbe > 0.0 ? static_cast<int>(nbe) - 1 : -1;
        r.nf = nbf > 0.0 ? static_cast<int>(nbf) - 1 : -1;
        r.set = true;
    }
    return assemble(n1, nel, corners, faces, rea_periodic, status);
}

}  // namespace

Mesh* Mesh::load(const char* path, int n1, int flags, int* status) {
    *status = NW_OK;
    if (!path || n1 < kMinN1 || n1 > kMaxN1) {
        set_error("mesh load: bad arguments");
        *status = NW_ERR_ARG;
        return nullptr;
    }
    const std::string p(path);
    // Case::loadMesh (case.cpp:149-153): a name containing ".re2" is binary
    if (p.find(".re2") != std::string::npos) return load_re2(p, n1, 0, status);
    return load_rea(p, n1, (flags & NW_MESH_REA_PERIODIC) != 0, status);
}

}  // namespace nw
