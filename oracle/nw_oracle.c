/*
 * nw_oracle.c -- FP64 CPU oracle (test infrastructure only; see nw_oracle.h).
 *
 * Built with -ffp-contract=off so every expression rounds exactly as written;
 * the expression order of the reference is kept wherever the reference has one,
 * so that geometry matches the reference's own mesh.cpp bit for bit.
 */
#include "nw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int g_threads = 1;
void or_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int or_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
/* mt19937_64 (Matsumoto & Nishimura 2004; identical stream to C++'s          */
/* std::mt19937_64) and the project transforms uniform=(x>>11)*2^-53, normal */
/* = Box-Muller first variate with u1 = 1-uniform, u2 = uniform.             */
/* ------------------------------------------------------------------------ */
#define MT_NN 312
#define MT_MM 156
void or_rng_seed(or_rng *r, uint64_t seed) {
    r->mt[0] = seed;
    for (int i = 1; i < MT_NN; ++i)
        r->mt[i] = 6364136223846793005ULL * (r->mt[i - 1] ^ (r->mt[i - 1] >> 62)) + (uint64_t)i;
    r->mti = MT_NN;
}
uint64_t or_rng_next(or_rng *r) {
    static const uint64_t mag01[2] = {0ULL, 0xB5026F5AA96619E9ULL};
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    if (r->mti >= MT_NN) {
        int i;
        uint64_t x;
        for (i = 0; i < MT_NN - MT_MM; ++i) {
            x = (r->mt[i] & UM) | (r->mt[i + 1] & LM);
            r->mt[i] = r->mt[i + MT_MM] ^ (x >> 1) ^ mag01[x & 1ULL];
        }
        for (; i < MT_NN - 1; ++i) {
            x = (r->mt[i] & UM) | (r->mt[i + 1] & LM);
            r->mt[i] = r->mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ mag01[x & 1ULL];
        }
        x = (r->mt[MT_NN - 1] & UM) | (r->mt[0] & LM);
        r->mt[MT_NN - 1] = r->mt[MT_MM - 1] ^ (x >> 1) ^ mag01[x & 1ULL];
        r->mti = 0;
    }
    uint64_t x = r->mt[r->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}
double or_rng_uniform(or_rng *r) { return (double)(or_rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
double or_rng_normal(or_rng *r) {
    double u1 = 1.0 - or_rng_uniform(r);
    double u2 = or_rng_uniform(r);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* ------------------------------------------------------------------------ */
/* GLL nodes, weights, D (mesh.cpp:20-40 legendrePoly/Deriv, :64-145 setupGLL) */
/* ------------------------------------------------------------------------ */
static double leg(int n, double x) {                 /* mesh.cpp:20-32 */
    if (n == 0) return 1.0;
    if (n == 1) return x;
    double p0 = 1.0, p1 = x, p2 = 0.0;
    for (int k = 1; k < n; ++k) {
        p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
        p0 = p1;
        p1 = p2;
    }
    return p2;
}
static double leg_d(int n, double x) {               /* mesh.cpp:36-40 */
    if (n == 0) return 0.0;
    if (n == 1) return 1.0;
    return (n / (x * x - 1.0)) * (x * leg(n, x) - leg(n - 1, x));
}

int or_gll(int n1, double *z, double *w, double *D, double *Dt) {
    if (n1 < 1 || n1 > OR_MAXN1) return -1;
    int N = n1, p = N - 1;
    for (int i = 0; i < N * N; ++i) { D[i] = 0.0; Dt[i] = 0.0; }
    if (N == 1) {                                    /* mesh.cpp:75-79 */
        z[0] = 0.0; w[0] = 2.0;
        return 0;
    }
    z[0] = -1.0;
    z[N - 1] = 1.0;
    for (int i = 1; i < N - 1; ++i) {                /* mesh.cpp:86-97 Newton on L'_p */
        double xi = -cos((M_PI * i) / p);
        for (int it = 0; it < 20; ++it) {
            double L = leg(p, xi);
            double Ld = leg_d(p, xi);
            double Ldd = (2.0 * xi * Ld - p * (p + 1.0) * L) / (1.0 - xi * xi);
            double delta = Ld / Ldd;
            xi -= delta;
            if (fabs(delta) < 1e-15) break;
        }
        z[i] = xi;
    }
    for (int i = 0; i < N; ++i) {                    /* mesh.cpp:101-104 */
        double L = leg(p, z[i]);
        w[i] = 2.0 / (p * (p + 1.0) * L * L);
    }
    double d0 = p * (p + 1.0) / 4.0;                 /* mesh.cpp:113-131 */
    for (int i = 0; i < N; ++i) {
        for (int j = 0; j < N; ++j) {
            int idx = i + j * N;
            if (i != j) {
                double Li = leg(p, z[i]);
                double Lj = leg(p, z[j]);
                D[idx] = Li / (Lj * (z[i] - z[j]));
            } else if (i == 0 && j == 0) {
                D[idx] = -d0;
            } else if (i == p && j == p) {
                D[idx] = d0;
            } else {
                D[idx] = 0.0;
            }
            Dt[j + i * N] = D[idx];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* mesh                                                                      */
/* ------------------------------------------------------------------------ */
static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

void or_mesh_free(or_mesh *m) {
    if (!m) return;
    free(m->w3); free(m->corners); free(m->f_kind); free(m->f_nbr_e); free(m->f_nbr_f);
    free(m->x); free(m->y); free(m->zc); free(m->jac);
    free(m->rx); free(m->sx); free(m->tx); free(m->ry); free(m->sy); free(m->ty);
    free(m->rz); free(m->sz); free(m->tz);
    free(m->vM); free(m->vP); free(m->bc); free(m->nx); free(m->ny); free(m->nz); free(m->dA);
    free(m->eps); free(m->mu); free(m->send_vol); free(m->ghost_elem);
    free(m);
}

/* computeMetricsFromCorners (mesh.cpp:188-314), literal expression order. */
static void metrics(or_mesh *m) {
    int N = m->n1;
    #pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (int e = 0; e < m->nelt; ++e) {
        const double *c = m->corners + (size_t)e * 24;
        int off = e * m->np;
        for (int k = 0; k < N; ++k) {
            double t = m->z[k];
            for (int j = 0; j < N; ++j) {
                double s = m->z[j];
                for (int i = 0; i < N; ++i) {
                    double r = m->z[i];
                    int pt = off + i + N * (j + N * k);
                    double Nf[8], dr[8], ds[8], dt[8];
                    Nf[0] = 0.125 * (1.0 - r) * (1.0 - s) * (1.0 - t);
                    Nf[1] = 0.125 * (1.0 + r) * (1.0 - s) * (1.0 - t);
                    Nf[2] = 0.125 * (1.0 + r) * (1.0 + s) * (1.0 - t);
                    Nf[3] = 0.125 * (1.0 - r) * (1.0 + s) * (1.0 - t);
                    Nf[4] = 0.125 * (1.0 - r) * (1.0 - s) * (1.0 + t);
                    Nf[5] = 0.125 * (1.0 + r) * (1.0 - s) * (1.0 + t);
                    Nf[6] = 0.125 * (1.0 + r) * (1.0 + s) * (1.0 + t);
                    Nf[7] = 0.125 * (1.0 - r) * (1.0 + s) * (1.0 + t);
                    dr[0] = -0.125 * (1.0 - s) * (1.0 - t);
                    dr[1] =  0.125 * (1.0 - s) * (1.0 - t);
                    dr[2] =  0.125 * (1.0 + s) * (1.0 - t);
                    dr[3] = -0.125 * (1.0 + s) * (1.0 - t);
                    dr[4] = -0.125 * (1.0 - s) * (1.0 + t);
                    dr[5] =  0.125 * (1.0 - s) * (1.0 + t);
                    dr[6] =  0.125 * (1.0 + s) * (1.0 + t);
                    dr[7] = -0.125 * (1.0 + s) * (1.0 + t);
                    ds[0] = -0.125 * (1.0 - r) * (1.0 - t);
                    ds[1] = -0.125 * (1.0 + r) * (1.0 - t);
                    ds[2] =  0.125 * (1.0 + r) * (1.0 - t);
                    ds[3] =  0.125 * (1.0 - r) * (1.0 - t);
                    ds[4] = -0.125 * (1.0 - r) * (1.0 + t);
                    ds[5] = -0.125 * (1.0 + r) * (1.0 + t);
                    ds[6] =  0.125 * (1.0 + r) * (1.0 + t);
                    ds[7] =  0.125 * (1.0 - r) * (1.0 + t);
                    dt[0] = -0.125 * (1.0 - r) * (1.0 - s);
                    dt[1] = -0.125 * (1.0 + r) * (1.0 - s);
                    dt[2] = -0.125 * (1.0 + r) * (1.0 + s);
                    dt[3] = -0.125 * (1.0 - r) * (1.0 + s);
                    dt[4] =  0.125 * (1.0 - r) * (1.0 - s);
                    dt[5] =  0.125 * (1.0 + r) * (1.0 - s);
                    dt[6] =  0.125 * (1.0 + r) * (1.0 + s);
                    dt[7] =  0.125 * (1.0 - r) * (1.0 + s);
                    double x = 0.0, y = 0.0, z = 0.0;
                    double xr = 0.0, yr = 0.0, zr = 0.0;
                    double xs = 0.0, ys = 0.0, zs = 0.0;
                    double xt = 0.0, yt = 0.0, zt = 0.0;
                    for (int a = 0; a < 8; ++a) {
                        double vx = c[a * 3 + 0], vy = c[a * 3 + 1], vz = c[a * 3 + 2];
                        x += Nf[a] * vx; y += Nf[a] * vy; z += Nf[a] * vz;
                        xr += dr[a] * vx; yr += dr[a] * vy; zr += dr[a] * vz;
                        xs += ds[a] * vx; ys += ds[a] * vy; zs += ds[a] * vz;
                        xt += dt[a] * vx; yt += dt[a] * vy; zt += dt[a] * vz;
                    }
                    m->x[pt] = x; m->y[pt] = y; m->zc[pt] = z;
                    double detJ = xr * (ys * zt - yt * zs)
                                - xs * (yr * zt - yt * zr)
                                + xt * (yr * zs - ys * zr);
                    m->jac[pt] = detJ;
                    double invJ = 1.0 / detJ;
                    m->rx[pt] = invJ * (ys * zt - yt * zs);
                    m->sx[pt] = invJ * (yt * zr - yr * zt);
                    m->tx[pt] = invJ * (yr * zs - ys * zr);
                    m->ry[pt] = invJ * (xt * zs - xs * zt);
                    m->sy[pt] = invJ * (xr * zt - xt * zr);
                    m->ty[pt] = invJ * (xs * zr - xr * zs);
                    m->rz[pt] = invJ * (xs * yt - xt * ys);
                    m->sz[pt] = invJ * (xt * yr - xr * yt);
                    m->tz[pt] = invJ * (xr * ys - xs * yr);
                }
            }
        }
    }
}

/* getFaceNodeVolIndex (mesh.cpp:955-981). */
static int face_node(int N, int np, int e, int f, int p, int q) {
    int i = 0, j = 0, k = 0;
    switch (f) {
        case 0: i = p; j = 0; k = q; break;
        case 1: i = N - 1; j = p; k = q; break;
        case 2: i = p; j = N - 1; k = q; break;
        case 3: i = 0; j = p; k = q; break;
        case 4: i = p; j = q; k = 0; break;
        case 5: i = p; j = q; k = N - 1; break;
        default: break;
    }
    return e * np + i + N * (j + N * k);
}

/* setupFaceData (mesh.cpp:985-1080) + bc-code flattening (dg_solver.cpp:90-112). */
static void face_data(or_mesh *m) {
    int N = m->n1, N2 = N * N;
    #pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (int fi = 0; fi < m->nelt * 6; ++fi) {
        int eA = fi / 6, fA = fi % 6;
        int kind = m->f_kind[fi];
        for (int qA = 0; qA < N; ++qA) {
            for (int pA = 0; pA < N; ++pA) {
                int P = fi * N2 + pA + N * qA;
                int vM = face_node(N, m->np, eA, fA, pA, qA);
                m->vM[P] = vM;
                double J = m->jac[vM];
                double w2 = m->w[pA] * m->w[qA];
                double Nx = 0.0, Ny = 0.0, Nz = 0.0;
                switch (fA) {
                    case 0: Nx = -J * m->sx[vM]; Ny = -J * m->sy[vM]; Nz = -J * m->sz[vM]; break;
                    case 1: Nx = +J * m->rx[vM]; Ny = +J * m->ry[vM]; Nz = +J * m->rz[vM]; break;
                    case 2: Nx = +J * m->sx[vM]; Ny = +J * m->sy[vM]; Nz = +J * m->sz[vM]; break;
                    case 3: Nx = -J * m->rx[vM]; Ny = -J * m->ry[vM]; Nz = -J * m->rz[vM]; break;
                    case 4: Nx = -J * m->tx[vM]; Ny = -J * m->ty[vM]; Nz = -J * m->tz[vM]; break;
                    case 5: Nx = +J * m->tx[vM]; Ny = +J * m->ty[vM]; Nz = +J * m->tz[vM]; break;
                }
                double L = sqrt(Nx * Nx + Ny * Ny + Nz * Nz);
                m->nx[P] = (L > 1e-14) ? (Nx / L) : 0.0;
                m->ny[P] = (L > 1e-14) ? (Ny / L) : 0.0;
                m->nz[P] = (L > 1e-14) ? (Nz / L) : 0.0;
                m->dA[P] = L * w2;
                int vP;
                int eB = m->f_nbr_e[fi], fB = m->f_nbr_f[fi];
                if (kind == OR_FACE_PEC || (kind != OR_FACE_HALO && eB < 0)) {
                    vP = -1;
                } else if (kind == OR_FACE_PERIODIC) {
                    vP = face_node(N, m->np, eB, fB, pA, qA);
                } else if (kind == OR_FACE_HALO) {
                    vP = -1; /* filled by or_mesh_self_halo_x */
                } else {
                    /* "E" (and PMC with a neighbour): brute-force coordinate match */
                    double xA = m->x[vM], yA = m->y[vM], zA = m->zc[vM];
                    int match = -1;
                    for (int qB = 0; qB < N && match < 0; ++qB) {
                        for (int pB = 0; pB < N; ++pB) {
                            int vB = face_node(N, m->np, eB, fB, pB, qB);
                            double dx = xA - m->x[vB], dy = yA - m->y[vB], dz = zA - m->zc[vB];
                            if (dx * dx + dy * dy + dz * dz < 1e-10) { match = vB; break; }
                        }
                    }
                    vP = match;
                }
                m->vP[P] = vP;
                int code = (kind == OR_FACE_PEC) ? 1 : (kind == OR_FACE_PMC) ? 2 : 0;
                if (vP < 0 && code == 0 && kind != OR_FACE_HALO) code = 1;
                m->bc[P] = code;
            }
        }
    }
}

static or_mesh *alloc_mesh(int n1, int nelt) {
    or_mesh *m = (or_mesh *)xcalloc(1, sizeof(or_mesh));
    m->n1 = n1; m->np = n1 * n1 * n1; m->nelt = nelt;
    m->npts = nelt * m->np; m->tfp = nelt * 6 * n1 * n1;
    or_gll(n1, m->z, m->w, m->D, m->Dt);
    m->w3 = (double *)xcalloc(m->np, sizeof(double));
    for (int k = 0; k < n1; ++k)                     /* mesh.cpp:136-144 */
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n1; ++i)
                m->w3[i + n1 * (j + n1 * k)] = m->w[i] * m->w[j] * m->w[k];
    size_t np = (size_t)m->npts, tf = (size_t)m->tfp;
    m->corners = (double *)xcalloc((size_t)nelt * 24, sizeof(double));
    m->f_kind = (int *)xcalloc((size_t)nelt * 6, sizeof(int));
    m->f_nbr_e = (int *)xcalloc((size_t)nelt * 6, sizeof(int));
    m->f_nbr_f = (int *)xcalloc((size_t)nelt * 6, sizeof(int));
    double **arr[] = {&m->x, &m->y, &m->zc, &m->jac, &m->rx, &m->sx, &m->tx,
                      &m->ry, &m->sy, &m->ty, &m->rz, &m->sz, &m->tz};
    for (size_t a = 0; a < sizeof(arr) / sizeof(arr[0]); ++a) *arr[a] = (double *)xcalloc(np, sizeof(double));
    m->vM = (int *)xcalloc(tf, sizeof(int));
    m->vP = (int *)xcalloc(tf, sizeof(int));
    m->bc = (int *)xcalloc(tf, sizeof(int));
    m->nx = (double *)xcalloc(tf, sizeof(double));
    m->ny = (double *)xcalloc(tf, sizeof(double));
    m->nz = (double *)xcalloc(tf, sizeof(double));
    m->dA = (double *)xcalloc(tf, sizeof(double));
    return m;
}

or_mesh *or_mesh_from_corners(int n1, int nelt, const double *corners, const int *kind,
                              const int *nbr_e, const int *nbr_f) {
    if (n1 < 2 || n1 > OR_MAXN1 || nelt < 1) return NULL;
    or_mesh *m = alloc_mesh(n1, nelt);
    memcpy(m->corners, corners, sizeof(double) * 24 * (size_t)nelt);
    memcpy(m->f_kind, kind, sizeof(int) * 6 * (size_t)nelt);
    memcpy(m->f_nbr_e, nbr_e, sizeof(int) * 6 * (size_t)nelt);
    memcpy(m->f_nbr_f, nbr_f, sizeof(int) * 6 * (size_t)nelt);
    metrics(m);
    face_data(m);
    return m;
}

/*
 * createBoxMesh (mesh.cpp:683-843): element index ex + nelx*(ey + nely*ez),
 * corner order (mesh.cpp:721-728), face list e-major faces 0..5 with the
 * reference's neighbour rules.  Extension: jitter > 0 moves every vertex whose
 * three indices are all interior by U[-jitter*h, +jitter*h] per coordinate
 * (h per direction; draws vertex-major, x fastest, 3 draws per vertex in x,y,z
 * order; uniform = (mt19937_64 >> 11) * 2^-53).  With jitter == 0 the corners
 * are the reference's x0 = xmin + ex*dx, x1 = x0 + dx exactly.
 */
or_mesh *or_mesh_box(int n1, int nelx, int nely, int nelz, const double *b,
                     int px, int py, int pz, double jitter, uint64_t seed) {
    if (nelx < 1 || nely < 1 || nelz < 1) return NULL;
    int nelt = nelx * nely * nelz;
    double *corners = (double *)xcalloc((size_t)nelt * 24, sizeof(double));
    int *kind = (int *)xcalloc((size_t)nelt * 6, sizeof(int));
    int *ne = (int *)xcalloc((size_t)nelt * 6, sizeof(int));
    int *nf = (int *)xcalloc((size_t)nelt * 6, sizeof(int));
    double xmin = b[0], xmax = b[1], ymin = b[2], ymax = b[3], zmin = b[4], zmax = b[5];
    double dx = (xmax - xmin) / nelx, dy = (ymax - ymin) / nely, dz = (zmax - zmin) / nelz;
    int nvx = nelx + 1, nvy = nely + 1, nvz = nelz + 1;
    double *V = NULL;
    if (jitter > 0.0) {
        V = (double *)xcalloc((size_t)nvx * nvy * nvz * 3, sizeof(double));
        for (int iz = 0; iz < nvz; ++iz)
            for (int iy = 0; iy < nvy; ++iy)
                for (int ix = 0; ix < nvx; ++ix) {
                    double *v = V + 3 * ((size_t)ix + nvx * ((size_t)iy + nvy * iz));
                    v[0] = xmin + ix * dx; v[1] = ymin + iy * dy; v[2] = zmin + iz * dz;
                }
        or_rng r;
        or_rng_seed(&r, seed);
        for (int iz = 1; iz < nelz; ++iz)
            for (int iy = 1; iy < nely; ++iy)
                for (int ix = 1; ix < nelx; ++ix) {
                    double *v = V + 3 * ((size_t)ix + nvx * ((size_t)iy + nvy * iz));
                    double ux = or_rng_uniform(&r), uy = or_rng_uniform(&r), uz = or_rng_uniform(&r);
                    v[0] += jitter * dx * (2.0 * ux - 1.0);
                    v[1] += jitter * dy * (2.0 * uy - 1.0);
                    v[2] += jitter * dz * (2.0 * uz - 1.0);
                }
    }
#define EIDX(a, b_, c) ((a) + nelx * ((b_) + nely * (c)))
    for (int ez = 0; ez < nelz; ++ez) {
        double z0 = zmin + ez * dz, z1 = z0 + dz;
        for (int ey = 0; ey < nely; ++ey) {
            double y0 = ymin + ey * dy, y1 = y0 + dy;
            for (int ex = 0; ex < nelx; ++ex) {
                double x0 = xmin + ex * dx, x1 = x0 + dx;
                int e = EIDX(ex, ey, ez);
                double *c = corners + (size_t)e * 24;
                if (V) {
                    static const int off[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1}};
                    for (int a = 0; a < 8; ++a) {
                        const double *v = V + 3 * ((size_t)(ex + off[a][0]) + nvx * ((size_t)(ey + off[a][1]) + nvy * (ez + off[a][2])));
                        c[a * 3 + 0] = v[0]; c[a * 3 + 1] = v[1]; c[a * 3 + 2] = v[2];
                    }
                } else {
                    double cc[8][3] = {{x0,y0,z0},{x1,y0,z0},{x1,y1,z0},{x0,y1,z0},
                                       {x0,y0,z1},{x1,y0,z1},{x1,y1,z1},{x0,y1,z1}};
                    memcpy(c, cc, sizeof(cc));
                }
                int *K = kind + 6 * e, *NE = ne + 6 * e, *NF = nf + 6 * e;
                /* face 0: s=-1 */
                if (ey > 0) { K[0] = OR_FACE_E; NE[0] = EIDX(ex, ey - 1, ez); NF[0] = 2; }
                else if (py) { K[0] = OR_FACE_PERIODIC; NE[0] = EIDX(ex, nely - 1, ez); NF[0] = 2; }
                else { K[0] = OR_FACE_PEC; NE[0] = -1; NF[0] = -1; }
                /* face 1: r=+1 */
                if (ex < nelx - 1) { K[1] = OR_FACE_E; NE[1] = EIDX(ex + 1, ey, ez); NF[1] = 3; }
                else if (px) { K[1] = OR_FACE_PERIODIC; NE[1] = EIDX(0, ey, ez); NF[1] = 3; }
                else { K[1] = OR_FACE_PEC; NE[1] = -1; NF[1] = -1; }
                /* face 2: s=+1 */
                if (ey < nely - 1) { K[2] = OR_FACE_E; NE[2] = EIDX(ex, ey + 1, ez); NF[2] = 0; }
                else if (py) { K[2] = OR_FACE_PERIODIC; NE[2] = EIDX(ex, 0, ez); NF[2] = 0; }
                else { K[2] = OR_FACE_PEC; NE[2] = -1; NF[2] = -1; }
                /* face 3: r=-1 */
                if (ex > 0) { K[3] = OR_FACE_E; NE[3] = EIDX(ex - 1, ey, ez); NF[3] = 1; }
                else if (px) { K[3] = OR_FACE_PERIODIC; NE[3] = EIDX(nelx - 1, ey, ez); NF[3] = 1; }
                else { K[3] = OR_FACE_PEC; NE[3] = -1; NF[3] = -1; }
                /* face 4: t=-1 */
                if (ez > 0) { K[4] = OR_FACE_E; NE[4] = EIDX(ex, ey, ez - 1); NF[4] = 5; }
                else if (pz) { K[4] = OR_FACE_PERIODIC; NE[4] = EIDX(ex, ey, nelz - 1); NF[4] = 5; }
                else { K[4] = OR_FACE_PEC; NE[4] = -1; NF[4] = -1; }
                /* face 5: t=+1 */
                if (ez < nelz - 1) { K[5] = OR_FACE_E; NE[5] = EIDX(ex, ey, ez + 1); NF[5] = 4; }
                else if (pz) { K[5] = OR_FACE_PERIODIC; NE[5] = EIDX(ex, ey, 0); NF[5] = 4; }
                else { K[5] = OR_FACE_PEC; NE[5] = -1; NF[5] = -1; }
            }
        }
    }
#undef EIDX
    or_mesh *m = or_mesh_from_corners(n1, nelt, corners, kind, ne, nf);
    free(V); free(corners); free(kind); free(ne); free(nf);
    return m;
}

/* Extension: per-element eps/mu (case configs with media); NULL = 1. */
int or_mesh_set_materials(or_mesh *m, const double *eps, const double *mu) {
    free(m->eps); free(m->mu); m->eps = NULL; m->mu = NULL;
    if (eps) { m->eps = (double *)xcalloc(m->nelt, sizeof(double)); memcpy(m->eps, eps, sizeof(double) * m->nelt); }
    if (mu) { m->mu = (double *)xcalloc(m->nelt, sizeof(double)); memcpy(m->mu, mu, sizeof(double) * m->nelt); }
    if ((eps || mu) && !m->eps) { m->eps = (double *)xcalloc(m->nelt, sizeof(double)); for (int e = 0; e < m->nelt; ++e) m->eps[e] = 1.0; }
    if ((eps || mu) && !m->mu) { m->mu = (double *)xcalloc(m->nelt, sizeof(double)); for (int e = 0; e < m->nelt; ++e) m->mu[e] = 1.0; }
    return 0;
}

/*
 * Extension (single-GPU share stand-in for the MPI halo of dg_solver.cpp:131-170,
 * 334-441): every x-periodic face (face 1 of ex = nelx-1, face 3 of ex = 0) of a
 * box mesh becomes a cut face whose plus-side trace lives in a ghost slot after
 * npts.  Ghost slots are numbered in face-list order, points q-outer p-inner;
 * send_vol[j] is the node the pack step copies into slot j (the periodic
 * partner), so pack -> copy -> unpack reproduces the periodic run bit for bit.
 */
int or_mesh_self_halo_x(or_mesh *m, int nelx) {
    int N = m->n1, N2 = N * N;
    int nh = 0;
    for (int fi = 0; fi < m->nelt * 6; ++fi) {
        int f = fi % 6, e = fi / 6, ex = e % nelx;
        if (m->f_kind[fi] == OR_FACE_PERIODIC && ((f == 1 && ex == nelx - 1) || (f == 3 && ex == 0))) nh += N2;
    }
    free(m->send_vol); free(m->ghost_elem);
    m->send_vol = (int *)xcalloc(nh, sizeof(int));
    m->ghost_elem = (int *)xcalloc(nh, sizeof(int));
    int g = 0;
    for (int fi = 0; fi < m->nelt * 6; ++fi) {
        int f = fi % 6, e = fi / 6, ex = e % nelx;
        if (!(m->f_kind[fi] == OR_FACE_PERIODIC && ((f == 1 && ex == nelx - 1) || (f == 3 && ex == 0)))) continue;
        m->f_kind[fi] = OR_FACE_HALO;
        for (int q = 0; q < N; ++q)
            for (int p = 0; p < N; ++p) {
                int P = fi * N2 + p + N * q;
                int j = g + p + N * q;
                m->send_vol[j] = m->vP[P];
                m->ghost_elem[j] = m->vP[P] / m->np;
                m->vP[P] = m->npts + j;
                m->bc[P] = 0;
            }
        g += N2;
    }
    m->nhalo = nh;
    return nh;
}

/* computeMinNodeDistance (mesh.cpp:883-930). */
double or_dxmin(const or_mesh *m) {
    double minD2 = 1.0e20;
    int N = m->n1;
    for (int e = 0; e < m->nelt; ++e) {
        int off = e * m->np;
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < N; ++i) {
                    int idx = off + i + N * (j + N * k);
                    double x0 = m->x[idx], y0 = m->y[idx], z0 = m->zc[idx];
                    int nb[3] = {i + 1 < N ? off + (i + 1) + N * (j + N * k) : -1,
                                 j + 1 < N ? off + i + N * ((j + 1) + N * k) : -1,
                                 k + 1 < N ? off + i + N * (j + N * (k + 1)) : -1};
                    for (int a = 0; a < 3; ++a) {
                        if (nb[a] < 0) continue;
                        double dx = m->x[nb[a]] - x0, dy = m->y[nb[a]] - y0, dz = m->zc[nb[a]] - z0;
                        double d2 = dx * dx + dy * dy + dz * dz;
                        if (d2 < minD2) minD2 = d2;
                    }
                }
    }
    return sqrt(minD2);
}
/* computeAutomaticCFL / computeAutomaticDt (mesh.cpp:938-951). */
double or_auto_cfl(double safety) { return (safety * 1.75) / sqrt(3.0); }
double or_auto_dt(const or_mesh *m, double speed, double safety) {
    double dxmin = or_dxmin(m);
    double cfl = or_auto_cfl(safety);
    return (cfl * dxmin) / speed;
}

/* ------------------------------------------------------------------------ */
/* kernels                                                                   */
/* ------------------------------------------------------------------------ */

/* gpu_volume_curl_kernel (dg_kernels.cu:41-155): local_grad3, W*J weight,
 * metric mapping and sign, same expression order. */
void or_volume_curl(const or_mesh *m, const double *u1a, const double *u2a, const double *u3a,
                    double *w1a, double *w2a, double *w3a, double sign) {
    int N = m->n1, N2 = N * N, N3 = m->np;
    const double *D = m->D;
    #pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (int e = 0; e < m->nelt; ++e) {
        const double *s1 = u1a + (size_t)e * N3, *s2 = u2a + (size_t)e * N3, *s3 = u3a + (size_t)e * N3;
        for (int mm = 0; mm < N3; ++mm) {
            int p = e * N3 + mm;
            int i = mm % N, j = (mm / N) % N, k = mm / N2;
            double u1r = 0.0, u1s = 0.0, u1t = 0.0;
            double u2r = 0.0, u2s = 0.0, u2t = 0.0;
            double u3r = 0.0, u3s = 0.0, u3t = 0.0;
            for (int l = 0; l < N; ++l) {
                double d_il = D[i + l * N], d_jl = D[j + l * N], d_kl = D[k + l * N];
                int ir = l + j * N + k * N2, is = i + l * N + k * N2, it = i + j * N + l * N2;
                u1r += d_il * s1[ir]; u2r += d_il * s2[ir]; u3r += d_il * s3[ir];
                u1s += d_jl * s1[is]; u2s += d_jl * s2[is]; u3s += d_jl * s3[is];
                u1t += d_kl * s1[it]; u2t += d_kl * s2[it]; u3t += d_kl * s3[it];
            }
            double w = m->w3[mm] * m->jac[p];
            double u1rw = u1r * w, u1sw = u1s * w, u1tw = u1t * w;
            double u2rw = u2r * w, u2sw = u2s * w, u2tw = u2t * w;
            double u3rw = u3r * w, u3sw = u3s * w, u3tw = u3t * w;
            double rx = m->rx[p], sx = m->sx[p], tx = m->tx[p];
            double ry = m->ry[p], sy = m->sy[p], ty = m->ty[p];
            double rz = m->rz[p], sz = m->sz[p], tz = m->tz[p];
            double o1 = (u3rw * ry + u3sw * sy + u3tw * ty) - (u2rw * rz + u2sw * sz + u2tw * tz);
            double o2 = (u1rw * rz + u1sw * sz + u1tw * tz) - (u3rw * rx + u3sw * sx + u3tw * tx);
            double o3 = (u2rw * rx + u2sw * sx + u2tw * tx) - (u1rw * ry + u1sw * sy + u1tw * ty);
            w1a[p] = sign * o1;
            w2a[p] = sign * o2;
            w3a[p] = sign * o3;
        }
    }
}

/* gpu_restrict_faces_kernel (dg_kernels.cu:170-190). */
void or_restrict(const or_mesh *m, const double *state, int stride, double *fEN, double *fHN) {
    size_t T = (size_t)m->tfp;
    #pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (long p = 0; p < (long)T; ++p) {
        int vM = m->vM[p];
        fEN[0 * T + p] = state[0 * (size_t)stride + vM];
        fEN[1 * T + p] = state[1 * (size_t)stride + vM];
        fEN[2 * T + p] = state[2 * (size_t)stride + vM];
        fHN[0 * T + p] = state[3 * (size_t)stride + vM];
        fHN[1 * T + p] = state[4 * (size_t)stride + vM];
        fHN[2 * T + p] = state[5 * (size_t)stride + vM];
    }
}

static inline int elem_of(const or_mesh *m, int v) {
    return v < m->npts ? v / m->np : m->ghost_elem[v - m->npts];
}

/*
 * gpu_compute_flux_kernel (dg_kernels.cu:220-304): jumps, PEC mirror when
 * bc != 0 or vP < 0, central/upwind flux, H-flux in slots 0..2, E-flux 3..5.
 * Extension (media): with eps/mu set, the upwind flux is impedance weighted,
 *   F_E = ( Z+ n x[[H]] - c0 n x(n x[[E]]) ) / (Z+ + Z-)
 *   F_H = (-Y+ n x[[E]] - c0 n x(n x[[H]]) ) / (Y+ + Y-),
 * the exact Riemann solution of the two-medium problem, which is the
 * reference formula (dg_kernels.cu:288-294) when Z+ = Z- = 1.  A mirror face
 * takes Z+ = Z-.
 */
void or_flux(const or_mesh *m, const double *state, int stride, const double *fEN,
             const double *fHN, double c0, double *flux) {
    size_t T = (size_t)m->tfp, S = (size_t)stride;
    #pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (long p = 0; p < (long)T; ++p) {
        double Ex_m = fEN[0 * T + p], Ey_m = fEN[1 * T + p], Ez_m = fEN[2 * T + p];
        double Hx_m = fHN[0 * T + p], Hy_m = fHN[1 * T + p], Hz_m = fHN[2 * T + p];
        double dEx, dEy, dEz, dHx, dHy, dHz;
        int vP = m->vP[p];
        int mirror = (m->bc[p] || vP < 0);
        if (mirror) {
            dEx = -2.0 * Ex_m; dEy = -2.0 * Ey_m; dEz = -2.0 * Ez_m;
            dHx = 0.0; dHy = 0.0; dHz = 0.0;
        } else {
            dEx = state[0 * S + vP] - Ex_m;
            dEy = state[1 * S + vP] - Ey_m;
            dEz = state[2 * S + vP] - Ez_m;
            dHx = state[3 * S + vP] - Hx_m;
            dHy = state[4 * S + vP] - Hy_m;
            dHz = state[5 * S + vP] - Hz_m;
        }
        double n_x = m->nx[p], n_y = m->ny[p], n_z = m->nz[p];
        double nxE_x = n_y * dEz - n_z * dEy;
        double nxE_y = n_z * dEx - n_x * dEz;
        double nxE_z = n_x * dEy - n_y * dEx;
        double nxH_x = n_y * dHz - n_z * dHy;
        double nxH_y = n_z * dHx - n_x * dHz;
        double nxH_z = n_x * dHy - n_y * dHx;
        double nxnxE_x = n_y * nxE_z - n_z * nxE_y;
        double nxnxE_y = n_z * nxE_x - n_x * nxE_z;
        double nxnxE_z = n_x * nxE_y - n_y * nxE_x;
        double nxnxH_x = n_y * nxH_z - n_z * nxH_y;
        double nxnxH_y = n_z * nxH_x - n_x * nxH_z;
        double nxnxH_z = n_x * nxH_y - n_y * nxH_x;
        double fHx, fHy, fHz, fEx, fEy, fEz;
        if (!m->eps) {
            fHx = -0.5 * nxE_x - 0.5 * c0 * nxnxH_x;
            fHy = -0.5 * nxE_y - 0.5 * c0 * nxnxH_y;
            fHz = -0.5 * nxE_z - 0.5 * c0 * nxnxH_z;
            fEx = +0.5 * nxH_x - 0.5 * c0 * nxnxE_x;
            fEy = +0.5 * nxH_y - 0.5 * c0 * nxnxE_y;
            fEz = +0.5 * nxH_z - 0.5 * c0 * nxnxE_z;
        } else {
            int eM = m->vM[p] / m->np;
            int eP = mirror ? eM : elem_of(m, vP);
            double Zm = sqrt(m->mu[eM] / m->eps[eM]), Zp = sqrt(m->mu[eP] / m->eps[eP]);
            double Ym = 1.0 / Zm, Yp = 1.0 / Zp;
            double iz = 1.0 / (Zp + Zm), iy = 1.0 / (Yp + Ym);
            fHx = (-Yp * nxE_x - c0 * nxnxH_x) * iy;
            fHy = (-Yp * nxE_y - c0 * nxnxH_y) * iy;
            fHz = (-Yp * nxE_z - c0 * nxnxH_z) * iy;
            fEx = (Zp * nxH_x - c0 * nxnxE_x) * iz;
            fEy = (Zp * nxH_y - c0 * nxnxE_y) * iz;
            fEz = (Zp * nxH_z - c0 * nxnxE_z) * iz;
        }
        flux[0 * T + p] = fHx; flux[1 * T + p] = fHy; flux[2 * T + p] = fHz;
        flux[3 * T + p] = fEx; flux[4 * T + p] = fEy; flux[5 * T + p] = fEz;
    }
}

/* gpu_add_flux_kernel (dg_kernels.cu:316-338), atomics replaced by the serial
 * face-point order (face points of element e are contiguous and all land in
 * element e, so the per-element loop is that order exactly). */
void or_add_flux(const or_mesh *m, const double *flux, double *rhs) {
    size_t T = (size_t)m->tfp, np = (size_t)m->npts;
    int per = 6 * m->n1 * m->n1;
    #pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (int e = 0; e < m->nelt; ++e) {
        for (int pp = 0; pp < per; ++pp) {
            size_t p = (size_t)e * per + pp;
            int vM = m->vM[p];
            double a = m->dA[p];
            rhs[0 * np + vM] += a * flux[3 * T + p];
            rhs[1 * np + vM] += a * flux[4 * T + p];
            rhs[2 * np + vM] += a * flux[5 * T + p];
            rhs[3 * np + vM] += a * flux[0 * T + p];
            rhs[4 * np + vM] += a * flux[1 * T + p];
            rhs[5 * np + vM] += a * flux[2 * T + p];
        }
    }
}

/* gpu_inv_mass_kernel (dg_kernels.cu:353-382); eps/mu per element when set. */
void or_inv_mass(const or_mesh *m, double *rhs) {
    size_t np = (size_t)m->npts;
    int nxyz = m->np;
    #pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (long idx = 0; idx < (long)np; ++idx) {
        int e = (int)(idx / nxyz), i = (int)(idx % nxyz);
        double eps = m->eps ? m->eps[e] : 1.0;
        double mu = m->mu ? m->mu[e] : 1.0;
        double iE = 1.0 / (eps * m->jac[idx] * m->w3[i]);
        double iH = 1.0 / (mu * m->jac[idx] * m->w3[i]);
        rhs[0 * np + idx] *= iE; rhs[1 * np + idx] *= iE; rhs[2 * np + idx] *= iE;
        rhs[3 * np + idx] *= iH; rhs[4 * np + idx] *= iH; rhs[5 * np + idx] *= iH;
    }
}

/* gpu_lsrk45_update_kernel (dg_kernels.cu:393-408). */
void or_lsrk_update(const or_mesh *m, double *state, int stride, double *k, const double *rhs,
                    double a, double b, double dt) {
    long n = 6L * m->npts;
    int npts = m->npts;
    #pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (long i = 0; i < n; ++i) {
        long comp = i / npts, p = i % npts;
        double ki = a * k[i] + dt * rhs[i];
        k[i] = ki;
        state[comp * stride + p] += b * ki;
    }
}

/* gpu_pack_halo_kernel / gpu_unpack_halo_kernel (dg_kernels.cu:413-453) with
 * a single neighbour (base 0, M = nhalo). */
void or_pack_halo(const or_mesh *m, const double *state, int stride, double *buf) {
    int M = m->nhalo;
    for (int j = 0; j < M; ++j)
        for (int c = 0; c < 6; ++c) buf[c * M + j] = state[(size_t)c * stride + m->send_vol[j]];
}
void or_unpack_halo(const or_mesh *m, const double *buf, double *state, int stride) {
    int M = m->nhalo;
    for (int j = 0; j < M; ++j)
        for (int c = 0; c < 6; ++c) state[(size_t)c * stride + m->npts + j] = buf[c * M + j];
}

/* LSRK45 a/b (dg_solver.cpp:58-69). */
void or_rk_coeffs(double *a, double *b) {
    a[0] = 0.0;
    a[1] = -567301805773.0 / 1357537059087.0;
    a[2] = -2404267990393.0 / 2016746695238.0;
    a[3] = -3550918686646.0 / 2091501179385.0;
    a[4] = -1275806237668.0 / 842570457699.0;
    b[0] = 1432997174477.0 / 9575080441755.0;
    b[1] = 5161836677717.0 / 13612068292357.0;
    b[2] = 1720146321549.0 / 2090206949498.0;
    b[3] = 3134564353537.0 / 4481467310338.0;
    b[4] = 2277821191437.0 / 14882151754819.0;
}

/* ------------------------------------------------------------------------ */
/* solver (dg_solver.cpp:40-241 allocation, :306-527 step)                   */
/* ------------------------------------------------------------------------ */
or_solver *or_solver_new(const or_mesh *m, double c0) {
    or_solver *s = (or_solver *)xcalloc(1, sizeof(or_solver));
    s->m = m; s->c0 = c0;
    s->stride = m->npts + m->nhalo;
    s->state = (double *)xcalloc((size_t)6 * s->stride, sizeof(double));
    s->k = (double *)xcalloc((size_t)6 * m->npts, sizeof(double));
    s->rhs = (double *)xcalloc((size_t)6 * m->npts, sizeof(double));
    s->fEN = (double *)xcalloc((size_t)3 * m->tfp, sizeof(double));
    s->fHN = (double *)xcalloc((size_t)3 * m->tfp, sizeof(double));
    s->flux = (double *)xcalloc((size_t)6 * m->tfp, sizeof(double));
    s->sendbuf = (double *)xcalloc((size_t)6 * m->nhalo, sizeof(double));
    s->recvbuf = (double *)xcalloc((size_t)6 * m->nhalo, sizeof(double));
    return s;
}
void or_solver_free(or_solver *s) {
    if (!s) return;
    free(s->state); free(s->k); free(s->rhs); free(s->fEN); free(s->fHN); free(s->flux);
    free(s->sendbuf); free(s->recvbuf); free(s);
}
void or_solver_upload(or_solver *s, const double *host) {
    for (int c = 0; c < 6; ++c)
        memcpy(s->state + (size_t)c * s->stride, host + (size_t)c * s->m->npts, sizeof(double) * s->m->npts);
}
void or_solver_download(const or_solver *s, double *host) {
    for (int c = 0; c < 6; ++c)
        memcpy(host + (size_t)c * s->m->npts, s->state + (size_t)c * s->stride, sizeof(double) * s->m->npts);
}
void or_solver_stage(or_solver *s, int stage, double dt) {
    const or_mesh *m = s->m;
    double a[5], b[5];
    or_rk_coeffs(a, b);
    size_t S = (size_t)s->stride, np = (size_t)m->npts;
    if (m->nhalo > 0) {                   /* pack -> "transport" -> unpack */
        or_pack_halo(m, s->state, s->stride, s->sendbuf);
        memcpy(s->recvbuf, s->sendbuf, sizeof(double) * 6 * (size_t)m->nhalo);
    }
    double *st = s->state, *r = s->rhs;
    or_volume_curl(m, st + 3 * S, st + 4 * S, st + 5 * S, r + 0 * np, r + 1 * np, r + 2 * np, 1.0);
    or_volume_curl(m, st + 0 * S, st + 1 * S, st + 2 * S, r + 3 * np, r + 4 * np, r + 5 * np, -1.0);
    if (m->nhalo > 0) or_unpack_halo(m, s->recvbuf, s->state, s->stride);
    or_restrict(m, st, s->stride, s->fEN, s->fHN);
    or_flux(m, st, s->stride, s->fEN, s->fHN, s->c0, s->flux);
    or_add_flux(m, s->flux, r);
    or_inv_mass(m, r);
    or_lsrk_update(m, st, s->stride, s->k, r, a[stage], b[stage], dt);
}
void or_solver_step(or_solver *s, double dt) {
    for (int stage = 0; stage < 5; ++stage) or_solver_stage(s, stage, dt);
}
double or_solver_run(or_solver *s, double dt, int nsteps) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int n = 0; n < nsteps; ++n) or_solver_step(s, dt);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------------------------ */
/* probes                                                                    */
/* ------------------------------------------------------------------------ */
/* evalLagrange1D (probe.cpp:11-30): the delta vector within 1e-12 of a node,
 * else the product over j != i of (xi - z_j) / (z_i - z_j), left to right. */
void or_lagrange_1d(int n1, const double *z, double xi, double *L) {
    for (int i = 0; i < n1; ++i)
        if (fabs(xi - z[i]) < 1e-12) {
            for (int j = 0; j < n1; ++j) L[j] = 0.0;
            L[i] = 1.0;
            return;
        }
    for (int i = 0; i < n1; ++i) {
        L[i] = 1.0;
        for (int j = 0; j < n1; ++j)
            if (j != i) L[i] *= (xi - z[j]) / (z[i] - z[j]);
    }
}

/* ProbeManager::init's search (probe.cpp:64-91): per element the reference
 * coordinates from the axis-aligned box of corners 0, 1 (x), 2 (y), 4 (z);
 * the squared distance of (r,s,t) to [-1,1]^3; the first element with the
 * smallest distance wins (strict <), its (r,s,t) clamped to [-1,1]. */
void or_locate_probes(const or_mesh *m, int n, const double *xyz, int *elem, double *rst) {
    for (int q = 0; q < n; ++q) {
        int best = 0;
        double bestd = 1e20, br = 0.0, bs = 0.0, bt = 0.0;
        for (int e = 0; e < m->nelt; ++e) {
            const double *c = m->corners + (size_t)e * 24;
            double x0 = c[0], x1 = c[1 * 3 + 0], y0 = c[1], y1 = c[2 * 3 + 1], z0 = c[2], z1 = c[4 * 3 + 2];
            double r = fabs(x1 - x0) > 1e-14 ? 2.0 * (xyz[3 * q + 0] - x0) / (x1 - x0) - 1.0 : 0.0;
            double s = fabs(y1 - y0) > 1e-14 ? 2.0 * (xyz[3 * q + 1] - y0) / (y1 - y0) - 1.0 : 0.0;
            double t = fabs(z1 - z0) > 1e-14 ? 2.0 * (xyz[3 * q + 2] - z0) / (z1 - z0) - 1.0 : 0.0;
            double dr = fmax(0.0, fmax(-1.0 - r, r - 1.0));
            double ds = fmax(0.0, fmax(-1.0 - s, s - 1.0));
            double dt = fmax(0.0, fmax(-1.0 - t, t - 1.0));
            double d2 = dr * dr + ds * ds + dt * dt;
            if (d2 < bestd) {
                bestd = d2;
                best = e;
                br = fmax(-1.0, fmin(1.0, r));
                bs = fmax(-1.0, fmin(1.0, s));
                bt = fmax(-1.0, fmin(1.0, t));
            }
        }
        elem[q] = best;
        rst[3 * q + 0] = br;
        rst[3 * q + 1] = bs;
        rst[3 * q + 2] = bt;
    }
}

/* ProbeManager::record (probe.cpp:192-216) with the weights of probe.cpp:98-105:
 * w = (Lr_i Ls_j) Lt_k over the element's nodes in storage order, each
 * component summed as acc += w * u.  out[6q + c]. */
void or_probe_record(const or_mesh *m, int n, const int *elem, const double *rst, const double *state,
                     int stride, double *out) {
    int N = m->n1;
    size_t S = (size_t)stride;
    for (int q = 0; q < n; ++q) {
        double Lr[OR_MAXN1], Ls[OR_MAXN1], Lt[OR_MAXN1];
        or_lagrange_1d(N, m->z, rst[3 * q + 0], Lr);
        or_lagrange_1d(N, m->z, rst[3 * q + 1], Ls);
        or_lagrange_1d(N, m->z, rst[3 * q + 2], Lt);
        size_t off = (size_t)elem[q] * m->np;
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < N; ++i) {
                    double w = Lr[i] * Ls[j] * Lt[k];
                    size_t p = off + (size_t)(i + N * (j + N * k));
                    for (int c = 0; c < 6; ++c) acc[c] += w * state[c * S + p];
                }
        for (int c = 0; c < 6; ++c) out[6 * q + c] = acc[c];
    }
}

/* ------------------------------------------------------------------------ */
/* diagnostics (case.cpp:733-794); eps/mu weighting when media are set       */
/* ------------------------------------------------------------------------ */
double or_energy(const or_mesh *m, const double *u, int stride) {
    int N = m->n1, Np = m->np;
    size_t S = (size_t)stride;
    double energy = 0.0;
    for (int e = 0; e < m->nelt; ++e) {
        double eps = m->eps ? m->eps[e] : 1.0, mu = m->mu ? m->mu[e] : 1.0;
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < N; ++i) {
                    size_t p = (size_t)e * Np + k * N * N + j * N + i;
                    double dV = m->jac[p] * m->w[i] * m->w[j] * m->w[k];
                    double eSq = u[0 * S + p] * u[0 * S + p] + u[1 * S + p] * u[1 * S + p] + u[2 * S + p] * u[2 * S + p];
                    double hSq = u[3 * S + p] * u[3 * S + p] + u[4 * S + p] * u[4 * S + p] + u[5 * S + p] * u[5 * S + p];
                    if (m->eps) energy += 0.5 * (eps * eSq + mu * hSq) * dV;
                    else energy += 0.5 * (eSq + hSq) * dV;
                }
    }
    return energy;
}
static double max_mag(const or_mesh *m, const double *u, int stride, int c0) {
    double mx = 0.0;
    size_t S = (size_t)stride;
    for (int i = 0; i < m->npts; ++i) {
        double a = u[(c0 + 0) * S + i], b = u[(c0 + 1) * S + i], c = u[(c0 + 2) * S + i];
        double mag = sqrt(a * a + b * b + c * c);
        if (mag > mx) mx = mag;
    }
    return mx;
}
double or_max_e(const or_mesh *m, const double *u, int stride) { return max_mag(m, u, stride, 0); }
double or_max_h(const or_mesh *m, const double *u, int stride) { return max_mag(m, u, stride, 3); }
