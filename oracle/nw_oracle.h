/*
 * nw_oracle.h -- FP64 CPU oracle for NekWave's DG-SEM Maxwell step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (NekWave src/mesh.cpp, src/dg_solver.cpp, src/device/dg_kernels.cu,
 * src/case.cpp, src/probe.cpp) used as the checker for the HIP product and as the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Extensions that
 * the reference does not have (seeded vertex jitter, per-element eps/mu with an
 * impedance-weighted upwind flux, the single-GPU self-halo) are marked
 * "extension" and reduce exactly to the reference when switched off.
 *
 * Layouts follow the reference exactly:
 *   - state: 6 component blocks [Ex|Ey|Ez|Hx|Hy|Hz], block stride = npts+nhalo,
 *     element-major, node m = i + n1*(j + n1*k) (mesh.cpp:216);
 *   - D column-major, D[i + j*n1] = D_ij (mesh.cpp:116);
 *   - face points element-major, faces 0..5, q outer, p inner (mesh.cpp:993-1004).
 */
#ifndef NW_ORACLE_H
#define NW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXN1 12

/* Face kinds (the bcType strings of mesh.hpp:18 / mesh.cpp:730-830). */
enum { OR_FACE_E = 0, OR_FACE_PERIODIC = 1, OR_FACE_PEC = 2, OR_FACE_PMC = 3, OR_FACE_HALO = 4 };

typedef struct or_mesh {
    int n1, np, nelt, npts, tfp, nhalo;
    double z[OR_MAXN1], w[OR_MAXN1];
    double D[OR_MAXN1 * OR_MAXN1];   /* column-major */
    double Dt[OR_MAXN1 * OR_MAXN1];
    double *w3;                      /* np */
    double *corners;                 /* nelt*8*3 */
    int *f_kind, *f_nbr_e, *f_nbr_f; /* nelt*6 */
    double *x, *y, *zc, *jac;        /* npts */
    double *rx, *sx, *tx, *ry, *sy, *ty, *rz, *sz, *tz;
    int *vM, *vP, *bc;               /* tfp */
    double *nx, *ny, *nz, *dA;       /* tfp */
    double *eps, *mu;                /* nelt, NULL = vacuum (reference) */
    int *send_vol;                   /* nhalo: volume node packed into ghost slot j */
    int *ghost_elem;                 /* nhalo: element owning the packed node */
} or_mesh;

/* ---- RNG (mt19937_64 + project transforms) ---- */
typedef struct { uint64_t mt[312]; int mti; } or_rng;
void     or_rng_seed(or_rng *r, uint64_t seed);
uint64_t or_rng_next(or_rng *r);
double   or_rng_uniform(or_rng *r);            /* (x >> 11) * 2^-53 */
double   or_rng_normal(or_rng *r);             /* Box-Muller, first variate */

/* ---- GLL (mesh.cpp:20-145) ---- */
int or_gll(int n1, double *z, double *w, double *D, double *Dt);

/* ---- mesh construction ---- */
or_mesh *or_mesh_box(int n1, int nelx, int nely, int nelz, const double *bounds6,
                     int px, int py, int pz, double jitter, uint64_t seed);
or_mesh *or_mesh_from_corners(int n1, int nelt, const double *corners, const int *kind,
                              const int *nbr_e, const int *nbr_f);
void or_mesh_free(or_mesh *m);
int  or_mesh_set_materials(or_mesh *m, const double *eps, const double *mu);
int  or_mesh_self_halo_x(or_mesh *m, int nelx);
double or_dxmin(const or_mesh *m);
double or_auto_cfl(double safety);
double or_auto_dt(const or_mesh *m, double speed, double safety);

/* ---- kernels (dg_kernels.cu) ---- */
void or_volume_curl(const or_mesh *m, const double *u1, const double *u2, const double *u3,
                    double *w1, double *w2, double *w3o, double sign);
void or_restrict(const or_mesh *m, const double *state, int stride, double *fEN, double *fHN);
void or_flux(const or_mesh *m, const double *state, int stride, const double *fEN,
             const double *fHN, double c0, double *flux);
void or_add_flux(const or_mesh *m, const double *flux, double *rhs);
void or_inv_mass(const or_mesh *m, double *rhs);
void or_lsrk_update(const or_mesh *m, double *state, int stride, double *k, const double *rhs,
                    double a, double b, double dt);
void or_pack_halo(const or_mesh *m, const double *state, int stride, double *buf);
void or_unpack_halo(const or_mesh *m, const double *buf, double *state, int stride);

/* ---- LSRK45 coefficients (dg_solver.cpp:58-69) ---- */
void or_rk_coeffs(double *a5, double *b5);

/* ---- solver (dg_solver.cpp:306-527) ---- */
typedef struct or_solver {
    const or_mesh *m;
    double c0;
    int stride;
    double *state, *k, *rhs, *fEN, *fHN, *flux, *sendbuf, *recvbuf;
} or_solver;
or_solver *or_solver_new(const or_mesh *m, double c0);
void or_solver_free(or_solver *s);
void or_solver_upload(or_solver *s, const double *host);   /* 6*npts */
void or_solver_download(const or_solver *s, double *host);
void or_solver_stage(or_solver *s, int stage, double dt);
void or_solver_step(or_solver *s, double dt);
double or_solver_run(or_solver *s, double dt, int nsteps); /* returns wall seconds */

/* ---- diagnostics (case.cpp:733-794) ---- */
double or_energy(const or_mesh *m, const double *state, int stride);
double or_max_e(const or_mesh *m, const double *state, int stride);
double or_max_h(const or_mesh *m, const double *state, int stride);

/* probes (probe.cpp:11-30, 57-105, 192-216): element and clamped (r,s,t) of
 * each point, and one record (6 values per probe) of a state */
void or_lagrange_1d(int n1, const double *z, double xi, double *L);
void or_locate_probes(const or_mesh *m, int n, const double *xyz, int *elem, double *rst);
void or_probe_record(const or_mesh *m, int n, const int *elem, const double *rst, const double *state,
                     int stride, double *out);

void or_set_threads(int n);
int  or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
