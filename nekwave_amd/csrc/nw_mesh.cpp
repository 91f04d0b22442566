// nw_mesh.cpp -- host-side mesh description for the solver (product code).
//
// Holds what the device kernels need and nothing per node: element corners,
// per-element trilinear coefficients, face neighbour codes, media, and the
// ghost-face map of the halo stand-in.  Per-node coordinates are produced on
// request only (the fused path evaluates geometry in-kernel).
#include "nw_mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "nw_fused.hpp"

namespace nw {

// --- GLL (same construction as mesh.cpp:64-145: Newton on L'_p from
// Chebyshev guesses, w = 2/(p(p+1)L_p^2), D_ij = L_p(x_i)/(L_p(x_j)(x_i-x_j)),
// corner diagonals -+ p(p+1)/4; stored column-major) ---
static double legendre(int n, double x) {
    if (n == 0) return 1.0;
    if (n == 1) return x;
    double a = 1.0, b = x;
    for (int k = 1; k < n; ++k) {
        const double c = ((2.0 * k + 1.0) * x * b - k * a) / (k + 1.0);
        a = b;
        b = c;
    }
    return b;
}
static double legendre_d(int n, double x) {
    if (n == 0) return 0.0;
    if (n == 1) return 1.0;
    return (n / (x * x - 1.0)) * (x * legendre(n, x) - legendre(n - 1, x));
}

void gll_host(int n1, double* z, double* w, double* Dcol) {
    const int p = n1 - 1;
    std::fill(Dcol, Dcol + n1 * n1, 0.0);
    if (n1 == 1) {
        z[0] = 0.0;
        w[0] = 2.0;
        return;
    }
    z[0] = -1.0;
    z[n1 - 1] = 1.0;
    for (int i = 1; i < n1 - 1; ++i) {
        double x = -std::cos((M_PI * i) / p);
        for (int it = 0; it < 20; ++it) {
            const double L = legendre(p, x), Ld = legendre_d(p, x);
            const double Ldd = (2.0 * x * Ld - p * (p + 1.0) * L) / (1.0 - x * x);
            const double d = Ld / Ldd;
            x -= d;
            if (std::fabs(d) < 1e-15) break;
        }
        z[i] = x;
    }
    for (int i = 0; i < n1; ++i) {
        const double L = legendre(p, z[i]);
        w[i] = 2.0 / (p * (p + 1.0) * L * L);
    }
    const double d0 = p * (p + 1.0) / 4.0;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n1; ++j) {
            double v = 0.0;
            if (i != j) v = legendre(p, z[i]) / (legendre(p, z[j]) * (z[i] - z[j]));
            else if (i == 0) v = -d0;
            else if (i == p) v = d0;
            Dcol[i + j * n1] = v;
        }
}

const GllTables& host_gll_tables() {
    static GllTables t = [] {
        GllTables g;
        std::memset(&g, 0, sizeof(g));
        for (int n = kMinN1; n <= kMaxN1; ++n) {
            double Dc[kMaxN1 * kMaxN1];
            gll_host(n, g.z[n], g.w[n], Dc);
            for (int i = 0; i < n; ++i) {
                g.iw[n][i] = 1.0 / g.w[n][i];
                for (int l = 0; l < n; ++l) g.D[n][i * n + l] = Dc[i + l * n];
            }
        }
        return g;
    }();
    return t;
}

// Corner sign patterns of the reference's corner order (mesh.cpp:721-728).
static const int kSr[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
static const int kSs[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
static const int kSt[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

static void elem_coeffs(const double* c, ElemGeom& g, unsigned char& affine) {
    for (int d = 0; d < 3; ++d) {
        double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a < 8; ++a) {
            const double v = c[a * 3 + d];
            const int r = kSr[a], q = kSs[a], t = kSt[a];
            s[0] += v; s[1] += r * v; s[2] += q * v; s[3] += t * v;
            s[4] += r * q * v; s[5] += r * t * v; s[6] += q * t * v; s[7] += r * q * t * v;
        }
        for (int m = 0; m < 8; ++m) g.c[d][m] = 0.125 * s[m];
    }
    double scale = 0.0;
    for (int d = 0; d < 3; ++d)
        for (int m = 1; m < 4; ++m) scale = std::max(scale, std::fabs(g.c[d][m]));
    bool aff = true;
    for (int d = 0; d < 3; ++d)
        for (int m = 4; m < 8; ++m)
            if (std::fabs(g.c[d][m]) > 1e-13 * scale) aff = false;
    if (aff)
        for (int d = 0; d < 3; ++d)
            for (int m = 4; m < 8; ++m) g.c[d][m] = 0.0;
    affine = aff ? 1 : 0;
}

// Physical point of reference coordinates (r, s, t) of element e (trilinear map).
static void map_point(const double* c, double r, double s, double t, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    for (int a = 0; a < 8; ++a) {
        const double N = 0.125 * (1.0 + kSr[a] * r) * (1.0 + kSs[a] * s) * (1.0 + kSt[a] * t);
        for (int d = 0; d < 3; ++d) out[d] += N * c[a * 3 + d];
    }
}

static double det_at(const double* c, double r, double s, double t) {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 8; ++a) {
        const double fr = 1.0 + kSr[a] * r, fs = 1.0 + kSs[a] * s, ft = 1.0 + kSt[a] * t;
        const double dr = 0.125 * kSr[a] * fs * ft, ds = 0.125 * kSs[a] * fr * ft, dt = 0.125 * kSt[a] * fr * fs;
        for (int d = 0; d < 3; ++d) {
            J[d][0] += dr * c[a * 3 + d];
            J[d][1] += ds * c[a * 3 + d];
            J[d][2] += dt * c[a * 3 + d];
        }
    }
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

int Mesh::finish() {
    const GllTables& T = host_gll_tables();
    np = n1 * n1 * n1;
    geom.resize(nelt);
    affine.resize(nelt);
    bool all_aff = true;
    for (int e = 0; e < nelt; ++e) {
        elem_coeffs(&corners[size_t(e) * 24], geom[e], affine[e]);
        all_aff = all_aff && affine[e];
        // inverted or degenerate elements fail here (the reference has no check)
        for (int a = 0; a < 8; ++a) {
            const double r = kSr[a] > 0 ? 1.0 : -1.0, s = kSs[a] > 0 ? 1.0 : -1.0, t = kSt[a] > 0 ? 1.0 : -1.0;
            if (!(det_at(&corners[size_t(e) * 24], r, s, t) > 0.0)) {
                set_error("element " + std::to_string(e) + " has a non-positive Jacobian");
                return NW_ERR_MESH;
            }
        }
    }
    this->all_affine = all_aff;
    // face codes
    fcode.assign(size_t(nelt) * 6, -1);
    fbc.assign(size_t(nelt) * 6, 0);
    const double* z = T.z[n1];
    for (int e = 0; e < nelt; ++e) {
        for (int f = 0; f < 6; ++f) {
            const size_t fi = size_t(e) * 6 + f;
            const int k = kind[fi];
            fbc[fi] = (k == NW_FACE_PEC) ? 1 : (k == NW_FACE_PMC) ? 2 : 0;
            const int e2 = nbr_e[fi], f2 = nbr_f[fi];
            // PEC and PMC faces both take the PEC mirror (dg_kernels.cu:244 tests
            // bc != 0), whatever neighbour the face list names
            if (k == NW_FACE_PEC || k == NW_FACE_PMC || e2 < 0) {
                fcode[fi] = -1;
                continue;
            }
            if (e2 >= nelt || f2 < 0 || f2 > 5) {
                set_error("face neighbour out of range at element " + std::to_string(e));
                return NW_ERR_MESH;
            }
            if (k == NW_FACE_PERIODIC || box_conforming) {
                // periodic faces pair (p, q) by identity (mesh.cpp:1043-1047); box
                // meshes are conforming with identity orientation by construction
                fcode[fi] = fcode_pack(e2, f2, 0);
                continue;
            }
            // E faces: coordinate match (mesh.cpp:1048-1070) reduced to finding the
            // square symmetry that maps every node of this face onto the neighbour's
            int found = -1;
            for (int o = 0; o < 8 && found < 0; ++o) {
                bool ok = true;
                for (int q = 0; q < n1 && ok; ++q)
                    for (int p = 0; p < n1 && ok; ++p) {
                        int i, j, kk, p2, q2, i2, j2, k2;
                        face_ijk(f, n1, p, q, i, j, kk);
                        orient_map(o, n1, p, q, p2, q2);
                        face_ijk(f2, n1, p2, q2, i2, j2, k2);
                        double A[3], B[3];
                        map_point(&corners[size_t(e) * 24], z[i], z[j], z[kk], A);
                        map_point(&corners[size_t(e2) * 24], z[i2], z[j2], z[k2], B);
                        const double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
                        ok = dx * dx + dy * dy + dz * dz < 1e-10;
                    }
                if (ok) found = o;
            }
            if (found < 0 && unmatched_is_pec) {
                // mesh.cpp:1048-1070 matches node by node and leaves unmatched nodes
                // at -1, which dg_solver.cpp:102-104 turns into PEC; a face none of
                // whose nodes lies on the neighbour face is therefore a PEC face
                bool any = false;
                for (int q = 0; q < n1 && !any; ++q)
                    for (int p = 0; p < n1 && !any; ++p) {
                        int i, j, kk;
                        face_ijk(f, n1, p, q, i, j, kk);
                        double A[3];
                        map_point(&corners[size_t(e) * 24], z[i], z[j], z[kk], A);
                        for (int q2 = 0; q2 < n1 && !any; ++q2)
                            for (int p2 = 0; p2 < n1 && !any; ++p2) {
                                int i2, j2, k2;
                                face_ijk(f2, n1, p2, q2, i2, j2, k2);
                                double B[3];
                                map_point(&corners[size_t(e2) * 24], z[i2], z[j2], z[k2], B);
                                const double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
                                any = dx * dx + dy * dy + dz * dz < 1e-10;
                            }
                    }
                if (!any) {
                    fcode[fi] = -1;
                    continue;
                }
            }
            if (found < 0) {
                set_error("non-conforming E face at element " + std::to_string(e) + " face " + std::to_string(f));
                return NW_ERR_MESH;
            }
            fcode[fi] = fcode_pack(e2, f2, found);
        }
    }
    mat.assign(nelt, ElemMat{1.0, 1.0, 1.0, 1.0});
    return NW_OK;
}

int Mesh::set_materials(const double* e, const double* m) {
    media = (e != nullptr) || (m != nullptr);
    for (int x = 0; x < nelt; ++x) {
        const double ep = e ? e[x] : 1.0, mu = m ? m[x] : 1.0;
        if (!(ep > 0.0) || !(mu > 0.0)) {
            set_error("eps and mu must be positive");
            return NW_ERR_ARG;
        }
        mat[x] = ElemMat{1.0 / ep, 1.0 / mu, std::sqrt(mu / ep), std::sqrt(ep / mu)};
    }
    eps.assign(e ? e : nullptr, e ? e + nelt : nullptr);
    mu_.assign(m ? m : nullptr, m ? m + nelt : nullptr);
    if (media) {
        if (eps.empty()) eps.assign(nelt, 1.0);
        if (mu_.empty()) mu_.assign(nelt, 1.0);
    }
    for (size_t g = 0; g < ghost_elem.size(); ++g) gmat[g] = mat[ghost_elem[g]];
    return NW_OK;
}

int Mesh::self_halo_x() {
    if (box_nel[0] <= 0) {
        set_error("self halo needs a box mesh");
        return NW_ERR_ARG;
    }
    const int nelx = box_nel[0];
    const int n2 = n1 * n1;
    ghost_elem.clear();
    send_vol.clear();
    int g = 0;
    for (int e = 0; e < nelt; ++e) {
        const int ex = e % nelx;
        for (int f : {1, 3}) {
            const size_t fi = size_t(e) * 6 + f;
            const bool cut = kind[fi] == NW_FACE_PERIODIC && ((f == 1 && ex == nelx - 1) || (f == 3 && ex == 0));
            if (!cut) continue;
            const int code = fcode[fi];
            const int e2 = code >> 6, f2 = (code >> 3) & 7;
            for (int q = 0; q < n1; ++q)
                for (int p = 0; p < n1; ++p) {
                    int i2, j2, k2;
                    face_ijk(f2, n1, p, q, i2, j2, k2);
                    send_vol.push_back(e2 * np + i2 + n1 * (j2 + n1 * k2));
                }
            ghost_elem.push_back(e2);
            fcode[fi] = -(g + 2);
            kind[fi] = NW_FACE_HALO_INTERNAL;
            ++g;
        }
    }
    (void)n2;
    nghost_faces = g;
    gmat.resize(g);
    for (int x = 0; x < g; ++x) gmat[x] = mat[ghost_elem[x]];
    return NW_OK;
}

long long Mesh::nhalo_points() const { return (long long)nghost_faces * n1 * n1; }

void Mesh::coords(double* X, double* Y, double* Z) const {
    const GllTables& T = host_gll_tables();
    const double* z = T.z[n1];
    for (int e = 0; e < nelt; ++e) {
        const double* c = &corners[size_t(e) * 24];
        for (int k = 0; k < n1; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i) {
                    double P[3];
                    map_point(c, z[i], z[j], z[k], P);
                    const size_t id = size_t(e) * np + i + n1 * (j + n1 * k);
                    X[id] = P[0]; Y[id] = P[1]; Z[id] = P[2];
                }
    }
}

void Mesh::jacobian(double* jac) const {
    const GllTables& T = host_gll_tables();
    const double* z = T.z[n1];
    for (int e = 0; e < nelt; ++e) {
        const double* c = &corners[size_t(e) * 24];
        for (int k = 0; k < n1; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i) jac[size_t(e) * np + i + n1 * (j + n1 * k)] = det_at(c, z[i], z[j], z[k]);
    }
}

double Mesh::min_node_distance() const {
    const GllTables& T = host_gll_tables();
    const double* z = T.z[n1];
    double m2 = 1e20;
    std::vector<double> P(size_t(np) * 3);
    for (int e = 0; e < nelt; ++e) {
        const double* c = &corners[size_t(e) * 24];
        for (int k = 0; k < n1; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i) map_point(c, z[i], z[j], z[k], &P[3 * (i + n1 * (j + n1 * k))]);
        for (int k = 0; k < n1; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i) {
                    const double* a = &P[3 * (i + n1 * (j + n1 * k))];
                    const int nb[3] = {i + 1 < n1 ? (i + 1) + n1 * (j + n1 * k) : -1,
                                       j + 1 < n1 ? i + n1 * ((j + 1) + n1 * k) : -1,
                                       k + 1 < n1 ? i + n1 * (j + n1 * (k + 1)) : -1};
                    for (int x = 0; x < 3; ++x) {
                        if (nb[x] < 0) continue;
                        const double* b = &P[3 * nb[x]];
                        const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
                        m2 = std::min(m2, dx * dx + dy * dy + dz * dz);
                    }
                }
    }
    return std::sqrt(m2);
}

Mesh* Mesh::box(int n1, int nelx, int nely, int nelz, const double* b, int px, int py, int pz, double jitter,
                uint64_t seed, int* status) {
    *status = NW_OK;
    if (n1 < kMinN1 || n1 > kMaxN1 || nelx < 1 || nely < 1 || nelz < 1) {
        set_error("box mesh: bad sizes");
        *status = NW_ERR_ARG;
        return nullptr;
    }
    Mesh* m = new Mesh();
    m->n1 = n1;
    m->nelt = nelx * nely * nelz;
    m->box_nel[0] = nelx; m->box_nel[1] = nely; m->box_nel[2] = nelz;
    m->box_conforming = true;
    const double xmin = b[0], xmax = b[1], ymin = b[2], ymax = b[3], zmin = b[4], zmax = b[5];
    const double dx = (xmax - xmin) / nelx, dy = (ymax - ymin) / nely, dz = (zmax - zmin) / nelz;
    std::vector<double> V;
    const int nvx = nelx + 1, nvy = nely + 1, nvz = nelz + 1;
    if (jitter > 0.0) {
        // vertex grid, interior vertices jittered in x-fastest order, 3 draws each
        V.resize(size_t(nvx) * nvy * nvz * 3);
        for (int iz = 0; iz < nvz; ++iz)
            for (int iy = 0; iy < nvy; ++iy)
                for (int ix = 0; ix < nvx; ++ix) {
                    double* v = &V[3 * (size_t(ix) + nvx * (size_t(iy) + nvy * iz))];
                    v[0] = xmin + ix * dx; v[1] = ymin + iy * dy; v[2] = zmin + iz * dz;
                }
        std::mt19937_64 rng(seed);
        auto uni = [&] { return double(rng() >> 11) * (1.0 / 9007199254740992.0); };
        for (int iz = 1; iz < nelz; ++iz)
            for (int iy = 1; iy < nely; ++iy)
                for (int ix = 1; ix < nelx; ++ix) {
                    double* v = &V[3 * (size_t(ix) + nvx * (size_t(iy) + nvy * iz))];
                    const double ux = uni(), uy = uni(), uz = uni();
                    v[0] += jitter * dx * (2.0 * ux - 1.0);
                    v[1] += jitter * dy * (2.0 * uy - 1.0);
                    v[2] += jitter * dz * (2.0 * uz - 1.0);
                }
    }
    m->corners.resize(size_t(m->nelt) * 24);
    m->kind.resize(size_t(m->nelt) * 6);
    m->nbr_e.resize(size_t(m->nelt) * 6);
    m->nbr_f.resize(size_t(m->nelt) * 6);
    auto eid = [&](int x, int y, int z) { return x + nelx * (y + nely * z); };
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int ez = 0; ez < nelz; ++ez) {
        const double z0 = zmin + ez * dz, z1 = z0 + dz;
        for (int ey = 0; ey < nely; ++ey) {
            const double y0 = ymin + ey * dy, y1 = y0 + dy;
            for (int ex = 0; ex < nelx; ++ex) {
                const double x0 = xmin + ex * dx, x1 = x0 + dx;
                const int e = eid(ex, ey, ez);
                double* c = &m->corners[size_t(e) * 24];
                for (int a = 0; a < 8; ++a) {
                    if (!V.empty()) {
                        const double* v =
                            &V[3 * (size_t(ex + off[a][0]) + nvx * (size_t(ey + off[a][1]) + nvy * (ez + off[a][2])))];
                        c[a * 3] = v[0]; c[a * 3 + 1] = v[1]; c[a * 3 + 2] = v[2];
                    } else {
                        c[a * 3] = off[a][0] ? x1 : x0;
                        c[a * 3 + 1] = off[a][1] ? y1 : y0;
                        c[a * 3 + 2] = off[a][2] ? z1 : z0;
                    }
                }
                // face rules of createBoxMesh (mesh.cpp:730-830)
                struct R { bool inside; int ne, nf; bool per; int pe; };
                const R rules[6] = {
                    {ey > 0, ey > 0 ? eid(ex, ey - 1, ez) : -1, 2, py != 0, eid(ex, nely - 1, ez)},
                    {ex < nelx - 1, ex < nelx - 1 ? eid(ex + 1, ey, ez) : -1, 3, px != 0, eid(0, ey, ez)},
                    {ey < nely - 1, ey < nely - 1 ? eid(ex, ey + 1, ez) : -1, 0, py != 0, eid(ex, 0, ez)},
                    {ex > 0, ex > 0 ? eid(ex - 1, ey, ez) : -1, 1, px != 0, eid(nelx - 1, ey, ez)},
                    {ez > 0, ez > 0 ? eid(ex, ey, ez - 1) : -1, 5, pz != 0, eid(ex, ey, nelz - 1)},
                    {ez < nelz - 1, ez < nelz - 1 ? eid(ex, ey, ez + 1) : -1, 4, pz != 0, eid(ex, ey, 0)}};
                for (int f = 0; f < 6; ++f) {
                    const size_t fi = size_t(e) * 6 + f;
                    if (rules[f].inside) {
                        m->kind[fi] = NW_FACE_E; m->nbr_e[fi] = rules[f].ne; m->nbr_f[fi] = rules[f].nf;
                    } else if (rules[f].per) {
                        m->kind[fi] = NW_FACE_PERIODIC; m->nbr_e[fi] = rules[f].pe; m->nbr_f[fi] = rules[f].nf;
                    } else {
                        m->kind[fi] = NW_FACE_PEC; m->nbr_e[fi] = -1; m->nbr_f[fi] = -1;
                    }
                }
            }
        }
    }
    *status = m->finish();
    if (*status != NW_OK) {
        delete m;
        return nullptr;
    }
    return m;
}

Mesh* Mesh::from_corners(int n1, int nelt, const double* corners, const int* kind, const int* ne, const int* nf,
                         int* status) {
    *status = NW_OK;
    if (n1 < kMinN1 || n1 > kMaxN1 || nelt < 1 || !corners || !kind || !ne || !nf) {
        set_error("from_corners: bad arguments");
        *status = NW_ERR_ARG;
        return nullptr;
    }
    Mesh* m = new Mesh();
    m->n1 = n1;
    m->nelt = nelt;
    m->corners.assign(corners, corners + size_t(nelt) * 24);
    m->kind.assign(kind, kind + size_t(nelt) * 6);
    m->nbr_e.assign(ne, ne + size_t(nelt) * 6);
    m->nbr_f.assign(nf, nf + size_t(nelt) * 6);
    for (size_t x = 0; x < m->kind.size(); ++x)
        if (m->kind[x] < NW_FACE_E || m->kind[x] > NW_FACE_PMC) {
            set_error("from_corners: unknown face kind");
            *status = NW_ERR_ARG;
            delete m;
            return nullptr;
        }
    *status = m->finish();
    if (*status != NW_OK) {
        delete m;
        return nullptr;
    }
    return m;
}

void locate_probes(const Mesh& m, int n, const double* xyz, int* elem, double* rst) {
    for (int q = 0; q < n; ++q) {
        const double px = xyz[3 * q], py = xyz[3 * q + 1], pz = xyz[3 * q + 2];
        int best = 0;
        double bd = 1e20, br = 0.0, bs = 0.0, bt = 0.0;
        for (int e = 0; e < m.nelt; ++e) {
            const double* c = &m.corners[size_t(e) * 24];  // corner a, axis d at c[3a + d]
            const double x0 = c[0], x1 = c[3], y0 = c[1], y1 = c[7], z0 = c[2], z1 = c[14];
            const double r = std::fabs(x1 - x0) > 1e-14 ? 2.0 * (px - x0) / (x1 - x0) - 1.0 : 0.0;
            const double s = std::fabs(y1 - y0) > 1e-14 ? 2.0 * (py - y0) / (y1 - y0) - 1.0 : 0.0;
            const double t = std::fabs(z1 - z0) > 1e-14 ? 2.0 * (pz - z0) / (z1 - z0) - 1.0 : 0.0;
            const double dr = std::max(0.0, std::max(-1.0 - r, r - 1.0));
            const double ds = std::max(0.0, std::max(-1.0 - s, s - 1.0));
            const double dt = std::max(0.0, std::max(-1.0 - t, t - 1.0));
            const double d2 = dr * dr + ds * ds + dt * dt;
            if (d2 < bd) {
                bd = d2;
                best = e;
                br = std::max(-1.0, std::min(1.0, r));
                bs = std::max(-1.0, std::min(1.0, s));
                bt = std::max(-1.0, std::min(1.0, t));
            }
        }
        elem[q] = best;
        rst[3 * q] = br;
        rst[3 * q + 1] = bs;
        rst[3 * q + 2] = bt;
    }
}

void lagrange_1d(int n1, const double* z, double xi, double* L) {
    for (int i = 0; i < n1; ++i) {
        if (std::fabs(xi - z[i]) < 1e-12) {
            for (int j = 0; j < n1; ++j) L[j] = 0.0;
            L[i] = 1.0;
            return;
        }
    }
    for (int i = 0; i < n1; ++i) {
        L[i] = 1.0;
        for (int j = 0; j < n1; ++j)
            if (i != j) L[i] *= (xi - z[j]) / (z[i] - z[j]);
    }
}

}  // namespace nw
