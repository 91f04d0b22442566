/*
 * nekwave_amd.h -- C ABI of the MI355X-native DG-SEM Maxwell step.
 *
 * Drop-in boundary for NekWave's solver path.  The reference exposes it as the
 * C++ class DgSolver (src/dg_solver.hpp:35-102) fed by a Mesh
 * (src/mesh.hpp:51-218), with the device work behind nekwave::device::launch_*
 * (src/device/dg_kernels.hpp:13-88).  Each entry point below names the
 * reference interface it replaces.  Plain pointers and sizes only; every call
 * returns an NW_* status (never aborts) and nw_last_error() gives the text.
 *
 * Threading: one nw_dg per device, externally synchronized (the reference's
 * single host thread + single stream, dg_solver.cpp:114).  The library owns all
 * device memory, its stream and its graphs; the caller owns host arrays and no
 * host pointer is retained after a call returns.
 */
#ifndef NEKWAVE_AMD_H
#define NEKWAVE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NW_ABI_VERSION 1

/* status codes */
enum {
    NW_OK = 0,
    NW_ERR_ARG = 1,         /* bad argument (the reference asserts, dg_solver.cpp:244,265,308) */
    NW_ERR_HIP = 2,         /* HIP runtime error (the reference aborts, gpu_runtime.hpp:113-121) */
    NW_ERR_ALLOC = 3,       /* device allocation failed */
    NW_ERR_STATE = 4,       /* call out of order (e.g. step before upload) */
    NW_ERR_MESH = 5,        /* invalid mesh (inverted element, non-conforming face) */
    NW_ERR_UNSUPPORTED = 6  /* n1/precision/pipeline combination not built */
};

/* arithmetic of the stage (fields / LSRK register).  NW_PREC_BF16X3: FP32
 * fields and registers with the volume contractions (dg_kernels.cu:101-124) as
 * split-precision BF16 products on the matrix cores -- each operand as
 * hi + lo round-to-nearest BF16 halves, hi*hi + hi*lo + lo*hi accumulated in
 * FP32 -- fused pipeline, n1 = 8 only.  Its stated error bound is larger than
 * FP32's (relative L2 against FP64 <= 1e-4 after 1000 steps of the 16^3
 * workload, against 1e-5 for FP32). */
enum { NW_PREC_FP64 = 0, NW_PREC_FP32 = 1, NW_PREC_FP32_K64 = 2, NW_PREC_BF16X3 = 3 };
/* NW_PIPE_PLAIN: the reference's 7-kernel stage (dg_solver.cpp:393-505), FP64 only.
 * NW_PIPE_FUSED: one fused kernel per stage. */
enum { NW_PIPE_PLAIN = 0, NW_PIPE_FUSED = 1 };
/* face kinds: the bcType strings of FaceInfo (mesh.hpp:15-21, mesh.cpp:730-830) */
enum { NW_FACE_E = 0, NW_FACE_PERIODIC = 1, NW_FACE_PEC = 2, NW_FACE_PMC = 3 };

const char* nw_last_error(void);
int nw_abi_version(void);
/* Number of visible devices and properties of `device` (name may be NULL). */
int nw_device_info(int device, int* ndev, char* name, size_t name_len, int* cus, size_t* mem_bytes);

/* ------------------------------------------------------------------------- */
/* Mesh -- replaces Mesh (src/mesh.hpp:51-218)                               */
/* ------------------------------------------------------------------------- */
typedef struct nw_mesh nw_mesh;

typedef struct nw_mesh_info {
    int n1;          /* GLL points per direction (Mesh::getN) */
    int np;          /* n1^3 (Mesh::getNumPointsPerElement) */
    int nelt;        /* Mesh::getNumElements */
    int nhalo;       /* ghost trace slots after npts (Mesh::getNumHaloPoints) */
    int64_t npts;    /* Mesh::getTotalPoints */
    int64_t tfp;     /* total face points, 6 n1^2 nelt */
    int has_media;   /* per-element eps/mu set */
    int all_affine;  /* every element is a parallelepiped */
} nw_mesh_info;

/* Mesh::setN + Mesh::createBoxMesh (mesh.cpp:56-61, 683-843): element order
 * ex + nelx (ey + nely ez), corner order mesh.cpp:721-728, faces PEC/PERIODIC/E.
 * bounds = {xmin,xmax,ymin,ymax,zmin,zmax}.  Extension: jitter > 0 moves every
 * vertex whose three indices are interior by U[-jitter h, +jitter h] per
 * coordinate (mt19937_64(seed), uniform = (x >> 11) 2^-53, vertices x-fastest,
 * draws x,y,z). */
int nw_mesh_box(int n1, int nelx, int nely, int nelz, const double bounds[6], int periodic_x, int periodic_y,
                int periodic_z, double jitter, uint64_t seed, nw_mesh** out);
/* The mesh Mesh::loadFromRea produces (mesh.cpp:317-445): corners [nelt][8][3]
 * and the FaceInfo list (kind, neighbour element, neighbour face; -1 = none).
 * E faces are matched by node coordinates like setupFaceData (mesh.cpp:1048-1070)
 * and must be conforming (one of the 8 square symmetries). */
int nw_mesh_from_corners(int n1, int nelt, const double* corners, const int* face_kind, const int* nbr_elem,
                         const int* nbr_face, nw_mesh** out);
/* Mesh::loadFromRea / Mesh::loadFromRe2 (mesh.cpp:317-636) as Case::loadMesh
 * picks them (case.cpp:149-153): a path containing ".re2" is read as binary
 * .re2, anything else as text .rea (NEL < 0 there redirects to the .re2 of the
 * same stem).  Boundary tags follow the reference's reading: on .rea, "PEC" "W"
 * "V" "v" are PEC, "PMC" "SYM" "S" "s" PMC (mirrored as PEC), other tags with a
 * neighbour are matched by coordinates and faces that match nowhere are PEC, so
 * a raw periodic "P" tag is PEC; flags NW_MESH_REA_PERIODIC pairs "P" faces by
 * (p, q) identity instead.  On .re2, tags containing 'P' are periodic and faces
 * without a record pair by centroid.  Inconsistent files (a face missing from
 * the list, partly matching faces, curved-side records) fail with NW_ERR_MESH
 * or NW_ERR_UNSUPPORTED where the reference would continue. */
enum { NW_MESH_REA_PERIODIC = 1 };
int nw_mesh_load(const char* path, int n1, int flags, nw_mesh** out);
/* The face list the mesh was built from, nelt*6 each (e-major): kind
 * (NW_FACE_*), neighbour element and face (-1 = none). */
int nw_mesh_faces(const nw_mesh* m, int* face_kind, int* nbr_elem, int* nbr_face);
/* Extension: per-element permittivity/permeability (NULL = 1).  The inverse
 * mass uses them (dg_kernels.cu:365-370 has eps = mu = 1) and the upwind flux
 * becomes impedance weighted; eps = mu = 1 reproduces the reference. */
int nw_mesh_set_materials(nw_mesh* m, const double* eps, const double* mu);
/* Extension: single-GPU share of a partitioned domain (Mesh::partition,
 * mesh.cpp:1083-1406): every x-periodic face becomes a cut face whose plus
 * trace is a ghost slot after npts, filled each stage by the on-device
 * pack -> copy -> unpack that stands in for the MPI exchange
 * (dg_solver.cpp:334-441).  Returns the number of ghost trace slots. */
int nw_mesh_self_halo_x(nw_mesh* m, int* nhalo);
int nw_mesh_get_info(const nw_mesh* m, nw_mesh_info* out);
/* getGllZ/getGllW (n1), getD (column-major n1*n1, D[i + j n1] = D_ij), getW3 (n1^3). */
int nw_mesh_gll(const nw_mesh* m, double* z, double* w, double* D, double* w3);
/* getCoordX/Y/Z (npts each), computed on the host. */
int nw_mesh_coords(const nw_mesh* m, double* x, double* y, double* z);
/* Jacobian determinant per node (getJac), npts values, computed on the host. */
int nw_mesh_jacobian(const nw_mesh* m, double* jac);
/* Element corners [nelt][8][3] (getElementCorners). */
int nw_mesh_corners(const nw_mesh* m, double* corners);
/* computeMinNodeDistance (mesh.cpp:883-930). */
int nw_mesh_min_node_distance(const nw_mesh* m, double* out);
/* computeAutomaticDt (mesh.cpp:947-951): cfl(safety) * dxmin / speed. */
int nw_mesh_auto_dt(const nw_mesh* m, double speed, double safety, double* out);
void nw_mesh_destroy(nw_mesh* m);

/* ------------------------------------------------------------------------- */
/* Solver -- replaces DgSolver (src/dg_solver.hpp:35-102)                    */
/* ------------------------------------------------------------------------- */
typedef struct nw_dg nw_dg;

typedef struct nw_opts {
    int precision;    /* NW_PREC_* (default FP64) */
    int pipeline;     /* NW_PIPE_* (default FUSED) */
    int use_graph;    /* capture each step as a hipGraph (default 1) */
    int device;       /* HIP device ordinal (default 0) */
    int tile_option;  /* fused tile option 0/1/2 (elements per workgroup; at n1 = 8 in FP32,
                       * 1 = matrix-core volume), -1 = tuned default */
    int groups_per_cu; /* extension, no counterpart in the reference: at most this many
                        * workgroups of a fused launch per CU (0 = as many as fit, the
                        * default), leaving room to other work on the device.  Results do
                        * not depend on it. */
} nw_opts;
void nw_opts_default(nw_opts* o);

/* DgSolver::initialize (dg_solver.cpp:40-241): all geometry to the device.  c0 is
 * the flux penalty (0 central, 1 upwind).  The mesh is only read during the call. */
int nw_dg_create(const nw_mesh* m, double c0, const nw_opts* opts, nw_dg** out);
/* DgSolver::uploadState (dg_solver.cpp:243-262): host [Ex|Ey|Ez|Hx|Hy|Hz], n = 6 npts. */
int nw_dg_upload(nw_dg* s, const double* host, size_t n);
/* DgSolver::downloadState (dg_solver.cpp:264-283). */
int nw_dg_download(nw_dg* s, double* host, size_t n);
/* DgSolver::step (dg_solver.cpp:306-527): one LSRK45 step, synchronous like the
 * reference.  `time` is unused, as in the reference (dg_solver.cpp:307). */
int nw_dg_step(nw_dg* s, double dt, double time);
/* nsteps steps with one synchronization at the end. */
int nw_dg_run(nw_dg* s, double dt, int nsteps);
/* nsteps steps timed with HIP events on the solver's stream (elapsed_ms). */
int nw_dg_time(nw_dg* s, double dt, int nsteps, double* elapsed_ms);
/* `warmup` untimed steps, then `repeats` timings of exactly nsteps steps each
 * (elapsed_ms[0..repeats-1]), all queued back to back on the solver's stream
 * with one host wait at the end. */
int nw_dg_time_repeats(nw_dg* s, double dt, int warmup, int nsteps, int repeats, double* elapsed_ms);
int nw_dg_sync(nw_dg* s);

typedef struct nw_diag {
    double energy;   /* 1/2 sum (eps|E|^2 + mu|H|^2) J w_i w_j w_k, FP64 (case.cpp:733-766) */
    double max_e;    /* max |E| (case.cpp:768-780) */
    double max_h;    /* max |H| (case.cpp:782-794) */
} nw_diag;
/* Diagnostics reduced on the device (no state download). */
int nw_dg_diag(nw_dg* s, nw_diag* out);

typedef struct nw_dg_info {
    int n1, nelt, precision, pipeline, use_graph, epw;  /* n1 = 0 for a batch of mixed n1 */
    int64_t npts, stride, device_bytes;
    int kernels_per_step, instances, segments;
} nw_dg_info;
int nw_dg_get_info(const nw_dg* s, nw_dg_info* out);
/* Extension (batched cavities): one solver advancing `count` independent meshes of
 * any mix of n1 with a common dt.  Instances of equal n1 form one segment; each
 * stage launches one persistent fused kernel per segment, one after another.
 * Instance i is addressed by its index in `meshes`. */
int nw_batch_create(int count, const nw_mesh* const* meshes, double c0, const nw_opts* opts, nw_dg** out);
int nw_dg_upload_instance(nw_dg* s, int instance, const double* host, size_t n);
int nw_dg_download_instance(nw_dg* s, int instance, double* host, size_t n);
int nw_dg_diag_instance(nw_dg* s, int instance, nw_diag* out);
/* Extension: relative L2 distance of an instance's current state to a host
 * reference state in the uploadState layout (6 component blocks of npts,
 * FP64): sqrt(sum (u - ref)^2 dV / sum ref^2 dV), dV = J w_i w_j w_k -- the
 * error norm of 3dboxpec_test.cpp:76-119 -- reduced on the device in FP64,
 * without downloading the state. */
int nw_dg_rel_l2(nw_dg* s, int instance, const double* ref, size_t n, double* rel_l2);
/* Extension (on-device ProbeManager, probe.cpp:33-216): places `n` probes at
 * xyz[3q..3q+2] exactly as ProbeManager::init does (first element whose
 * axis-aligned corner box is nearest, reference coordinates clamped to [-1,1],
 * 1-D Lagrange weights on the GLL nodes), and from then on every step appends one
 * record -- 6 field values per probe, summed in ProbeManager::record's order in
 * FP64 -- to a device history of `capacity` records.  `m` is the mesh the solver
 * was created from (single-mesh solvers only); n = 0 removes the probes. */
int nw_dg_set_probes(nw_dg* s, const nw_mesh* m, int n, const double* xyz, int capacity);
/* Appends one record of the current state (the reference records t = 0 before
 * stepping, case.cpp:473-474). */
int nw_dg_probe_sample(nw_dg* s);
/* Copies the records taken since the last call, oldest first, as
 * out[(r * n + q) * 6 + c] (c = Ex Ey Ez Hx Hy Hz), sets *nrecords, and starts the
 * history again.  NW_ERR_STATE when more than `capacity` records were taken. */
int nw_dg_probe_history(nw_dg* s, double* out, int max_records, int* nrecords);
/* Host half of ProbeManager::init (probe.cpp:57-96): element and clamped (r,s,t)
 * of each probe. */
int nw_mesh_locate_probes(const nw_mesh* m, int n, const double* xyz, int* elem, double* rst);
/* Autotuner: time every distinct tile option per segment -- with and without
 * non-temporal k accesses where the working set exceeds the 256 MiB Infinity
 * Cache -- and keep the fastest (runs by itself before the first step when
 * tile_option = -1).  Choices are kept per process, keyed by device, n1,
 * precision, element count, single mesh or batch, whether the working set
 * exceeds the Infinity Cache, and groups_per_cu. */
int nw_dg_tune(nw_dg* s);
/* n1, element count and chosen elements per workgroup of segment `seg`. */
int nw_dg_segment_info(const nw_dg* s, int seg, int* n1, int* nelt, int* epw);
/* Extension: the tile option segment `seg` runs (0/1/2, as tuned or set),
 * whether that option computes the volume contractions on the matrix cores
 * (n1 = 8 with FP32 fields, option 1; results equal the vector path's) and
 * whether its LSRK register accesses are non-temporal (tuned, or by default
 * where the working set exceeds the 256 MiB Infinity Cache and every wave
 * covers whole cache lines).  Any output pointer may be NULL. */
int nw_dg_segment_option(const nw_dg* s, int seg, int* option, int* matrix_volume, int* stream_k);
/* Extension: set segment `seg`'s tile option (0/1/2) and non-temporal k
 * accesses (1 on, 0 off, -1 the default rule) -- e.g. choices an earlier
 * process tuned and kept (nw_dg_segment_option).  The solver then counts as
 * tuned: autotuning does not override the choice until nw_dg_tune is called.
 * Results do not depend on either choice. */
int nw_dg_set_segment_option(nw_dg* s, int seg, int option, int stream_k);
/* Extension, no counterpart in the reference: the walk direction of the fused
 * stages.  mode 1: every other stage visits its element groups backwards, so it
 * starts on the groups the stage before it wrote last (some of them still in
 * the Infinity Cache; a batch then also takes its segments in reverse order);
 * 0: every stage forwards; -1 (default): alternate where the largest segment's
 * working set exceeds the 256 MiB Infinity Cache by at most 8x.
 * nw_dg_walk reports whether the solver alternates.  Results do not depend
 * on it. */
int nw_dg_set_walk(nw_dg* s, int mode);
int nw_dg_walk(const nw_dg* s, int* alternating);
/* DgSolver::finalize (dg_solver.cpp:529-582) + destructor. */
void nw_dg_destroy(nw_dg* s);

/* ------------------------------------------------------------------------- */
/* Field snapshots -- replaces Hdf5Writer (src/hdf5_writer.hpp:22-120) as it  */
/* writes without the HDF5 library (hdf5_writer.cpp:98-119, 121-470,          */
/* 711-733, 1002-1033)                                                        */
/* ------------------------------------------------------------------------- */
typedef struct nw_fields nw_fields;
/* point fields of each .vtu (FieldSaveOptions, hdf5_writer.hpp:56-67; the
 * default is E and H; the derived curl/div fields are not offered) */
enum { NW_SAVE_E = 1, NW_SAVE_H = 2, NW_SAVE_MAGNITUDE_E = 4, NW_SAVE_MAGNITUDE_H = 8, NW_SAVE_ENERGY_DENSITY = 16 };
/* Hdf5Writer::initialize(h5Path, mesh, enableXdmf, exportContinuous): from
 * "<dir>/<stem>.h5" it writes <dir>/<stem>.bin (the "NEKWAVE_BIN_V2" archive:
 * header and node coordinates) and, per saved step, <dir>/<stem>_step_<n>.vtu
 * plus the <dir>/<stem>.pvd collection.  export_continuous averages coincident
 * nodes into continuous points (within 1e-6) as the reference does.  The mesh
 * is read during the call only. */
int nw_fields_open(const char* h5_path, const nw_mesh* m, int export_continuous, int save_flags, nw_fields** out);
/* Hdf5Writer::writeStep(step, time, state, npts): state in the uploadState layout,
 * n = 6 npts. */
int nw_fields_write(nw_fields* f, int step, double time, const double* state, size_t n);
/* Hdf5Writer::close + destructor: the final .pvd, the archive closed. */
void nw_fields_close(nw_fields* f);

/* ------------------------------------------------------------------------- */
/* Kernel level -- replaces nekwave::device::launch_* (dg_kernels.hpp:13-88) */
/* Device pointers; stream NULL = default stream.                            */
/* ------------------------------------------------------------------------- */
/* launch_volume_curl (dg_kernels.hpp:13-21; kernel dg_kernels.cu:41-155). */
int nw_launch_volume_curl(const double* u1, const double* u2, const double* u3, double* w1, double* w2, double* w3,
                          const double* D, const double* w3_gll, const double* jac, const double* rx,
                          const double* sx, const double* tx, const double* ry, const double* sy, const double* ty,
                          const double* rz, const double* sz, const double* tz, double sign, int n1, int nelt,
                          void* stream);
/* launch_restrict_faces (dg_kernels.hpp:26-30; dg_kernels.cu:170-190). */
int nw_launch_restrict_faces(const double* state, double* fEN, double* fHN, const int* volIdxMinus,
                             int64_t totalFacePoints, int64_t stateStride, void* stream);
/* launch_compute_flux (dg_kernels.hpp:35-41; dg_kernels.cu:220-304).  Extension:
 * eps/mu per element (NULL = reference flux), with volIdxMinus, ghost_elem (the
 * element owning each ghost slot) and np to find both sides' media. */
int nw_launch_compute_flux(const double* state, const double* fEN, const double* fHN, const int* volIdxPlus,
                           const int* isPEC, const double* nx, const double* ny, const double* nz, double* flux,
                           double c0, int64_t totalFacePoints, int64_t stateStride, const int* volIdxMinus,
                           const double* eps, const double* mu, const int* ghost_elem, int np, int64_t npts,
                           void* stream);
/* launch_add_flux (dg_kernels.hpp:46-50; dg_kernels.cu:316-338), the reference's
 * arguments: rhs[c npts + volIdxMinus[p]] += dA[p] flux[...] (E rows from the
 * flux's H-half order of dg_kernels.cu:329-336).  Where the face points follow
 * the element-major layout of dg_solver.cpp:76-112 (sizes tfp = 6 n1^2 nelt,
 * npts = n1^3 nelt and volIdxMinus as that layout puts them), each node gathers
 * its points in face order 0..5: deterministic, no atomics.  Any point whose
 * volIdxMinus departs from that layout is added with FP64 atomics as the
 * reference adds every point. */
int nw_launch_add_flux(const double* flux, const int* volIdxMinus, const double* dA, double* rhs,
                       int64_t totalFacePoints, int64_t npts, void* stream);
/* launch_inv_mass (dg_kernels.hpp:55-59; dg_kernels.cu:353-382); eps/mu per element or NULL. */
int nw_launch_inv_mass(double* rhs, const double* jac, const double* w3, int nelt, int nxyz, int64_t npts,
                       const double* eps, const double* mu, void* stream);
/* launch_lsrk45_update (dg_kernels.hpp:64-68; dg_kernels.cu:393-408). */
int nw_launch_lsrk45_update(double* state, double* k, const double* rhs, double rk4a, double rk4b, double dt,
                            int64_t npts, int64_t stateStride, void* stream);
/* launch_pack_halo / launch_unpack_halo (dg_kernels.hpp:73-90; dg_kernels.cu:413-453). */
int nw_launch_pack_halo(const double* state, double* sendBuf, const int* sendVolIndices, const int* sendBufOffsets,
                        const int* exchangeSizes, int numHaloPoints, int64_t stateStride, void* stream);
int nw_launch_unpack_halo(const double* recvBuf, double* state, const int* recvBufOffsets, const int* exchangeSizes,
                          int numHaloPoints, int64_t npts, int64_t stateStride, void* stream);

/* Case::computeTotalEnergy, Case::getMaxE, Case::getMaxH (case.cpp:733-794) on a
 * device state instead of the downloaded host copy: `state` in the reference
 * layout (component c of node p at c stateStride + p, stateStride >= npts =
 * nelt n1^3), jac per node (getJac), w the n1 1-D GLL weights (getGllW),
 * eps/mu per element or NULL (NULL, NULL is the reference's energy).  Writes
 * out[0] = 1/2 sum (eps |E|^2 + mu |H|^2) J w_i w_j w_k, out[1] = max |E|,
 * out[2] = max |H| to the device array `out` (3 doubles), asynchronously on
 * `stream`.  Partial sums go to a buffer the library allocates once per device
 * (on that device's first call).  FP64, deterministic for a given npts. */
int nw_launch_diagnostics(const double* state, int64_t stateStride, const double* jac, const double* w,
                          const double* eps, const double* mu, int n1, int nelt, double* out, void* stream);

/* No counterpart in the reference: nw_launch_diagnostics for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results, bit
 * for bit; the partial sums live in a fixed 24 KiB array per GPU (1024 partial
 * triples, held until the process ends) and the call allocates nothing.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * streams, other host threads) race.  NW_ERR_ARG if the state would need more
 * partial sums than the array holds. */
int nw_launch_diagnostics_noalloc(const double* state, int64_t stateStride, const double* jac, const double* w,
                                  const double* eps, const double* mu, int n1, int nelt, double* out, void* stream);

/* Seeded synthetic inputs: std::mt19937_64(seed) with uniform = (x >> 11) 2^-53,
 * normal = Box-Muller first variate of (1 - uniform, uniform); n draws. */
int nw_rand_uniform(uint64_t seed, size_t n, double* out);
int nw_rand_normal(uint64_t seed, size_t n, double* out);

/* Device memory helpers for callers without their own allocator (tests, FFI). */
int nw_dev_alloc(size_t bytes, void** p);
int nw_dev_free(void* p);
int nw_dev_h2d(void* dst, const void* src, size_t bytes);
int nw_dev_d2h(void* dst, const void* src, size_t bytes);
int nw_dev_sync(void);
/* Plain device-to-device copy kernel of `bytes` (HBM calibration), ms per copy. */
int nw_bench_copy(size_t bytes, int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* NEKWAVE_AMD_H */
