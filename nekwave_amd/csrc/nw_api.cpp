// nw_api.cpp -- extern "C" boundary (include/nekwave_amd.h).
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "nw_fields.hpp"
#include "nw_fused.hpp"
#include "nw_plain.hpp"
#include "nw_solver.hpp"

namespace nw {
static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const char* last_error() { return g_err.c_str(); }
}  // namespace nw

using nw::set_error;

#define NW_API extern "C" __attribute__((visibility("default")))

#define NW_GUARD_BEGIN try {
#define NW_GUARD_END                                      \
    }                                                     \
    catch (const std::bad_alloc&) {                       \
        set_error("host allocation failed");              \
        return NW_ERR_ALLOC;                              \
    }                                                     \
    catch (const std::exception& ex) {                    \
        set_error(std::string("exception: ") + ex.what()); \
        return NW_ERR_ARG;                                \
    }

NW_API const char* nw_last_error(void) { return nw::last_error(); }
NW_API int nw_abi_version(void) { return NW_ABI_VERSION; }

NW_API int nw_device_info(int device, int* ndev, char* name, size_t name_len, int* cus, size_t* mem) {
    int n = 0;
    NW_TRY_HIP(hipGetDeviceCount(&n));
    if (ndev) *ndev = n;
    if (device < 0 || device >= n) {
        set_error("no such device");
        return NW_ERR_ARG;
    }
    hipDeviceProp_t p;
    NW_TRY_HIP(hipGetDeviceProperties(&p, device));
    if (name && name_len) {
        std::strncpy(name, p.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cus) *cus = p.multiProcessorCount;
    if (mem) *mem = p.totalGlobalMem;
    return NW_OK;
}

// ---------------------------------------------------------------- mesh ----
static int wrap_mesh(nw::Mesh* m, int st, nw_mesh** out) {
    if (!m) return st ? st : NW_ERR_MESH;
    nw_mesh* w = new nw_mesh;
    w->m = std::move(*m);
    delete m;
    *out = w;
    return NW_OK;
}

NW_API int nw_mesh_box(int n1, int nelx, int nely, int nelz, const double bounds[6], int px, int py, int pz,
                       double jitter, uint64_t seed, nw_mesh** out) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(out && bounds, "nw_mesh_box: null argument");
    NW_CHECK_ARG(jitter >= 0.0 && jitter < 0.5, "nw_mesh_box: jitter must be in [0, 0.5)");
    *out = nullptr;
    int st = NW_OK;
    nw::Mesh* m = nw::Mesh::box(n1, nelx, nely, nelz, bounds, px, py, pz, jitter, seed, &st);
    return wrap_mesh(m, st, out);
    NW_GUARD_END
}

NW_API int nw_mesh_from_corners(int n1, int nelt, const double* corners, const int* kind, const int* ne,
                                const int* nf, nw_mesh** out) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(out, "nw_mesh_from_corners: null out");
    *out = nullptr;
    int st = NW_OK;
    nw::Mesh* m = nw::Mesh::from_corners(n1, nelt, corners, kind, ne, nf, &st);
    return wrap_mesh(m, st, out);
    NW_GUARD_END
}

NW_API int nw_mesh_load(const char* path, int n1, int flags, nw_mesh** out) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(out && path, "nw_mesh_load: null argument");
    *out = nullptr;
    int st = NW_OK;
    nw::Mesh* m = nw::Mesh::load(path, n1, flags, &st);
    return wrap_mesh(m, st, out);
    NW_GUARD_END
}

NW_API int nw_mesh_faces(const nw_mesh* m, int* kind, int* ne, int* nf) {
    NW_CHECK_ARG(m && kind && ne && nf, "null argument");
    const size_t n = m->m.kind.size();
    for (size_t x = 0; x < n; ++x)  // cut faces of the halo stand-in are periodic faces fed through ghosts
        kind[x] = m->m.kind[x] == nw::NW_FACE_HALO_INTERNAL ? NW_FACE_PERIODIC : m->m.kind[x];
    std::memcpy(ne, m->m.nbr_e.data(), sizeof(int) * n);
    std::memcpy(nf, m->m.nbr_f.data(), sizeof(int) * n);
    return NW_OK;
}

NW_API int nw_mesh_set_materials(nw_mesh* m, const double* eps, const double* mu) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(m, "null mesh");
    return m->m.set_materials(eps, mu);
    NW_GUARD_END
}

NW_API int nw_mesh_self_halo_x(nw_mesh* m, int* nhalo) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(m, "null mesh");
    int st = m->m.self_halo_x();
    if (nhalo) *nhalo = static_cast<int>(m->m.nhalo_points());
    return st;
    NW_GUARD_END
}

NW_API int nw_mesh_get_info(const nw_mesh* m, nw_mesh_info* o) {
    NW_CHECK_ARG(m && o, "null argument");
    o->n1 = m->m.n1;
    o->np = m->m.np;
    o->nelt = m->m.nelt;
    o->nhalo = static_cast<int>(m->m.nhalo_points());
    o->npts = m->m.npts();
    o->tfp = m->m.tfp();
    o->has_media = m->m.media ? 1 : 0;
    o->all_affine = m->m.all_affine ? 1 : 0;
    return NW_OK;
}

NW_API int nw_mesh_gll(const nw_mesh* m, double* z, double* w, double* D, double* w3) {
    NW_CHECK_ARG(m, "null mesh");
    const int n1 = m->m.n1;
    const nw::GllTables& T = nw::host_gll_tables();
    if (z) std::memcpy(z, T.z[n1], sizeof(double) * n1);
    if (w) std::memcpy(w, T.w[n1], sizeof(double) * n1);
    if (D)
        for (int i = 0; i < n1; ++i)
            for (int l = 0; l < n1; ++l) D[i + l * n1] = T.D[n1][i * n1 + l];
    if (w3)
        for (int k = 0; k < n1; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i) w3[i + n1 * (j + n1 * k)] = T.w[n1][i] * T.w[n1][j] * T.w[n1][k];
    return NW_OK;
}

NW_API int nw_mesh_coords(const nw_mesh* m, double* x, double* y, double* z) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(m && x && y && z, "null argument");
    m->m.coords(x, y, z);
    return NW_OK;
    NW_GUARD_END
}

NW_API int nw_mesh_jacobian(const nw_mesh* m, double* jac) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(m && jac, "null argument");
    m->m.jacobian(jac);
    return NW_OK;
    NW_GUARD_END
}

NW_API int nw_mesh_corners(const nw_mesh* m, double* c) {
    NW_CHECK_ARG(m && c, "null argument");
    std::memcpy(c, m->m.corners.data(), sizeof(double) * m->m.corners.size());
    return NW_OK;
}

NW_API int nw_mesh_min_node_distance(const nw_mesh* m, double* out) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(m && out, "null argument");
    *out = m->m.min_node_distance();
    return NW_OK;
    NW_GUARD_END
}

NW_API int nw_mesh_auto_dt(const nw_mesh* m, double speed, double safety, double* out) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(m && out && speed > 0.0, "bad argument");
    const double cfl = (safety * 1.75) / std::sqrt(3.0);  // computeAutomaticCFL (mesh.cpp:938-943)
    *out = (cfl * m->m.min_node_distance()) / speed;
    return NW_OK;
    NW_GUARD_END
}

NW_API void nw_mesh_destroy(nw_mesh* m) { delete m; }

// -------------------------------------------------------------- solver ----
NW_API void nw_opts_default(nw_opts* o) {
    if (!o) return;
    o->precision = NW_PREC_FP64;
    o->pipeline = NW_PIPE_FUSED;
    o->use_graph = 1;
    o->device = 0;
    o->tile_option = -1;
    o->groups_per_cu = 0;
}

NW_API int nw_dg_create(const nw_mesh* m, double c0, const nw_opts* opts, nw_dg** out) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(m && out, "null argument");
    *out = nullptr;
    nw_opts o;
    nw_opts_default(&o);
    if (opts) o = *opts;
    nw_dg* d = new nw_dg;
    const nw::Mesh* one = &m->m;
    int st = d->s.init(&one, 1, c0, o);
    if (st) {
        delete d;
        return st;
    }
    *out = d;
    return NW_OK;
    NW_GUARD_END
}

NW_API int nw_batch_create(int count, const nw_mesh* const* meshes, double c0, const nw_opts* opts, nw_dg** out) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(meshes && out && count >= 1, "bad argument");
    *out = nullptr;
    nw_opts o;
    nw_opts_default(&o);
    if (opts) o = *opts;
    std::vector<const nw::Mesh*> ms(count);
    for (int x = 0; x < count; ++x) {
        NW_CHECK_ARG(meshes[x], "null mesh in batch");
        ms[x] = &meshes[x]->m;
    }
    nw_dg* d = new nw_dg;
    int st = d->s.init(ms.data(), count, c0, o);
    if (st) {
        delete d;
        return st;
    }
    *out = d;
    return NW_OK;
    NW_GUARD_END
}

NW_API int nw_dg_upload(nw_dg* s, const double* host, size_t n) {
    NW_CHECK_ARG(s, "null solver");
    return s->s.upload(0, host, n);
}
NW_API int nw_dg_download(nw_dg* s, double* host, size_t n) {
    NW_CHECK_ARG(s, "null solver");
    return s->s.download(0, host, n);
}
NW_API int nw_dg_upload_instance(nw_dg* s, int instance, const double* host, size_t n) {
    NW_CHECK_ARG(s, "null solver");
    return s->s.upload(instance, host, n);
}
NW_API int nw_dg_download_instance(nw_dg* s, int instance, double* host, size_t n) {
    NW_CHECK_ARG(s, "null solver");
    return s->s.download(instance, host, n);
}
NW_API int nw_dg_rel_l2(nw_dg* s, int instance, const double* ref, size_t n, double* rel_l2) {
    NW_CHECK_ARG(s, "null solver");
    return s->s.rel_l2(instance, ref, n, rel_l2);
}
NW_API int nw_dg_set_probes(nw_dg* s, const nw_mesh* m, int n, const double* xyz, int capacity) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(s && m, "null argument");
    return s->s.set_probes(m->m, n, xyz, capacity);
    NW_GUARD_END
}
NW_API int nw_dg_probe_sample(nw_dg* s) {
    NW_CHECK_ARG(s, "null solver");
    return s->s.probe_sample();
}
NW_API int nw_dg_probe_history(nw_dg* s, double* out, int max_records, int* nrecords) {
    NW_CHECK_ARG(s, "null solver");
    return s->s.probe_history(out, max_records, nrecords);
}
NW_API int nw_mesh_locate_probes(const nw_mesh* m, int n, const double* xyz, int* elem, double* rst) {
    NW_CHECK_ARG(m && n >= 0 && (n == 0 || (xyz && elem && rst)), "null argument");
    nw::locate_probes(m->m, n, xyz, elem, rst);
    return NW_OK;
}
NW_API int nw_dg_diag_instance(nw_dg* s, int instance, nw_diag* out) {
    NW_CHECK_ARG(s && out, "null argument");
    return s->s.diag(instance, out);
}
NW_API int nw_dg_tune(nw_dg* s) {
    NW_CHECK_ARG(s, "null solver");
    if (!s->s.uploaded) {
        nw::set_error("tune needs an uploaded state");
        return NW_ERR_STATE;
    }
    return s->s.tune();
}
NW_API int nw_dg_set_segment_option(nw_dg* s, int seg, int option, int stream_k) {
    NW_CHECK_ARG(s, "null solver");
    return s->s.set_segment_option(seg, option, stream_k);
}
NW_API int nw_dg_set_walk(nw_dg* s, int mode) {
    NW_CHECK_ARG(s, "null solver");
    return s->s.set_walk(mode);
}
NW_API int nw_dg_walk(const nw_dg* s, int* alternating) {
    NW_CHECK_ARG(s && alternating, "null argument");
    *alternating = s->s.alternating_walk() ? 1 : 0;
    return NW_OK;
}
NW_API int nw_dg_segment_info(const nw_dg* s, int seg, int* n1, int* nelt, int* epw) {
    NW_CHECK_ARG(s && seg >= 0 && seg < (int)s->s.seg.size(), "bad segment");
    const nw::Segment& g = s->s.seg[seg];
    if (n1) *n1 = g.n1;
    if (nelt) *nelt = g.nelt;
    if (epw) *epw = nw::fused_epw(g.n1, g.option, s->s.prec);
    return NW_OK;
}
NW_API int nw_dg_segment_option(const nw_dg* s, int seg, int* option, int* matrix_volume, int* stream_k) {
    NW_CHECK_ARG(s && seg >= 0 && seg < (int)s->s.seg.size(), "bad segment");
    const nw::Segment& g = s->s.seg[seg];
    const bool fused = s->s.pipe == NW_PIPE_FUSED;
    if (option) *option = g.option;
    if (matrix_volume) *matrix_volume = fused && nw::fused_matrix_volume(g.n1, g.option, s->s.prec) ? 1 : 0;
    if (stream_k) *stream_k = fused ? s->s.stream_of(g, g.option) : 0;
    return NW_OK;
}
NW_API int nw_dg_step(nw_dg* s, double dt, double time) {
    (void)time;
    NW_CHECK_ARG(s, "null solver");
    return s->s.step(dt, 1);
}
NW_API int nw_dg_run(nw_dg* s, double dt, int nsteps) {
    NW_CHECK_ARG(s, "null solver");
    return s->s.step(dt, nsteps);
}
NW_API int nw_dg_time(nw_dg* s, double dt, int nsteps, double* ms) {
    NW_CHECK_ARG(s && ms, "null argument");
    return s->s.time(dt, 0, nsteps, 1, ms);
}
NW_API int nw_dg_time_repeats(nw_dg* s, double dt, int warmup, int nsteps, int repeats, double* ms) {
    NW_CHECK_ARG(s && ms, "null argument");
    return s->s.time(dt, warmup, nsteps, repeats, ms);
}
NW_API int nw_dg_sync(nw_dg* s) {
    NW_CHECK_ARG(s, "null solver");
    NW_TRY_HIP(hipSetDevice(s->s.device));
    NW_TRY_HIP(hipStreamSynchronize(s->s.stream));
    return NW_OK;
}
NW_API int nw_dg_diag(nw_dg* s, nw_diag* out) {
    NW_CHECK_ARG(s && out, "null argument");
    return s->s.diag(0, out);
}
NW_API int nw_dg_get_info(const nw_dg* s, nw_dg_info* o) {
    NW_CHECK_ARG(s && o, "null argument");
    o->n1 = s->s.n1;
    o->nelt = s->s.nelt;
    o->precision = s->s.prec;
    o->pipeline = s->s.pipe;
    o->use_graph = s->s.use_graph ? 1 : 0;
    o->epw = s->s.pipe == NW_PIPE_FUSED ? nw::fused_epw(s->s.seg[0].n1, s->s.seg[0].option, s->s.prec) : 1;
    o->instances = static_cast<int>(s->s.inst.size());
    o->segments = static_cast<int>(s->s.seg.size());
    o->npts = s->s.npts;
    o->stride = s->s.stride;
    o->device_bytes = s->s.device_bytes;
    o->kernels_per_step = s->s.kernels_per_step;
    return NW_OK;
}
NW_API void nw_dg_destroy(nw_dg* s) { delete s; }

// -------------------------------------------------------- kernel level ----
static hipStream_t hs(void* s) { return static_cast<hipStream_t>(s); }

static int ensure_tables() {
    int dev = 0;
    NW_TRY_HIP(hipGetDevice(&dev));
    return nw::upload_gll_tables(dev);
}

NW_API int nw_launch_volume_curl(const double* u1, const double* u2, const double* u3, double* w1, double* w2,
                                 double* w3, const double* D, const double* w3g, const double* jac, const double* rx,
                                 const double* sx, const double* tx, const double* ry, const double* sy,
                                 const double* ty, const double* rz, const double* sz, const double* tz, double sign,
                                 int n1, int nelt, void* stream) {
    NW_CHECK_ARG(n1 >= nw::kMinN1 && n1 <= nw::kMaxN1 && nelt >= 0, "volume_curl: n1 must be in [2, 10]");
    nw::PlainVolumeArgs a{u1, u2, u3, w1, w2, w3, D, w3g, jac, rx, sx, tx, ry, sy, ty, rz, sz, tz, sign, n1, nelt};
    NW_TRY_HIP(nw::launch_volume_curl(a, hs(stream)));
    return NW_OK;
}
NW_API int nw_launch_restrict_faces(const double* st, double* fEN, double* fHN, const int* vM, int64_t tfp,
                                    int64_t S, void* stream) {
    NW_CHECK_ARG(tfp >= 0 && S >= 0, "restrict: negative size");
    NW_TRY_HIP(nw::launch_restrict(st, fEN, fHN, vM, tfp, S, hs(stream)));
    return NW_OK;
}
NW_API int nw_launch_compute_flux(const double* st, const double* fEN, const double* fHN, const int* vP,
                                  const int* isPEC, const double* nx, const double* ny, const double* nz,
                                  double* flux, double c0, int64_t tfp, int64_t S, const int* vM, const double* eps,
                                  const double* mu, const int* ghost_elem, int np, int64_t npts, void* stream) {
    NW_CHECK_ARG(tfp >= 0, "flux: negative size");
    NW_CHECK_ARG(!eps == !mu, "flux: eps and mu must both be given or both NULL");
    NW_CHECK_ARG(!eps || (vM && np > 0), "flux: media need volIdxMinus and np");
    NW_TRY_HIP(nw::launch_flux(st, fEN, fHN, vP, isPEC, nx, ny, nz, flux, c0, tfp, S, vM, eps, mu, ghost_elem, np,
                               npts, hs(stream)));
    return NW_OK;
}
NW_API int nw_launch_add_flux(const double* flux, const int* volIdxMinus, const double* dA, double* rhs, int64_t tfp,
                              int64_t npts, void* stream) {
    NW_CHECK_ARG(tfp >= 0 && npts >= 0, "add_flux: negative size");
    NW_CHECK_ARG(tfp == 0 || (flux && volIdxMinus && dA && rhs), "add_flux: null pointer");
    NW_TRY_HIP(nw::launch_add_flux(flux, volIdxMinus, dA, rhs, tfp, npts, hs(stream)));
    return NW_OK;
}
NW_API int nw_launch_inv_mass(double* rhs, const double* jac, const double* w3, int nelt, int nxyz, int64_t npts,
                              const double* eps, const double* mu, void* stream) {
    NW_CHECK_ARG(npts == (int64_t)nelt * nxyz, "inv_mass: npts != nelt*nxyz");
    NW_TRY_HIP(nw::launch_inv_mass(rhs, jac, w3, nxyz, npts, eps, mu, hs(stream)));
    return NW_OK;
}
NW_API int nw_launch_lsrk45_update(double* st, double* k, const double* rhs, double a, double b, double dt,
                                   int64_t npts, int64_t S, void* stream) {
    NW_CHECK_ARG(S >= npts, "lsrk: stride < npts");
    NW_TRY_HIP(nw::launch_lsrk(st, k, rhs, a, b, nullptr, dt, npts, S, hs(stream)));
    return NW_OK;
}
NW_API int nw_launch_pack_halo(const double* st, double* buf, const int* vol, const int* off, const int* sizes,
                               int nh, int64_t S, void* stream) {
    NW_TRY_HIP(nw::launch_pack<double>(st, buf, vol, off, sizes, nh, S, 1, hs(stream)));
    return NW_OK;
}
NW_API int nw_launch_unpack_halo(const double* buf, double* st, const int* off, const int* sizes, int nh,
                                 int64_t npts, int64_t S, void* stream) {
    NW_TRY_HIP(nw::launch_unpack<double>(buf, st, off, sizes, nh, npts, S, 1, hs(stream)));
    return NW_OK;
}

// Partial-sum buffer of nw_launch_diagnostics, one per device, allocated on the
// device's first call and kept until the process ends.
static std::mutex g_part_mu;
static double* g_part[64] = {};

static int check_diag_args(const double* st, int64_t S, const double* jac, const double* w, int n1, int nelt,
                           const double* out) {
    NW_CHECK_ARG(n1 >= nw::kMinN1 && n1 <= nw::kMaxN1, "diagnostics: n1 must be in [2, 10]");
    NW_CHECK_ARG(nelt > 0, "diagnostics: nelt must be positive");
    NW_CHECK_ARG(st && jac && w && out, "diagnostics: null pointer");
    NW_CHECK_ARG(S >= (int64_t)nelt * n1 * n1 * n1, "diagnostics: stateStride < npts");
    return NW_OK;
}

NW_API int nw_launch_diagnostics(const double* st, int64_t S, const double* jac, const double* w, const double* eps,
                                 const double* mu, int n1, int nelt, double* out, void* stream) {
    NW_GUARD_BEGIN
    int rc = check_diag_args(st, S, jac, w, n1, nelt, out);
    if (rc) return rc;
    int dev = 0;
    NW_TRY_HIP(hipGetDevice(&dev));
    NW_CHECK_ARG(dev >= 0 && dev < 64, "diagnostics: device ordinal beyond 63");
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_part_mu);
        if (!g_part[dev]) {
            void* p = nullptr;
            if (hipMalloc(&p, sizeof(double) * 3 * nw::kEnergyMaxBlocks) != hipSuccess) {
                set_error("diagnostics: hipMalloc of the partial sums failed");
                return NW_ERR_ALLOC;
            }
            g_part[dev] = static_cast<double*>(p);
        }
        part = g_part[dev];
    }
    const long long npts = (long long)nelt * n1 * n1 * n1;
    NW_TRY_HIP(nw::launch_energy(st, S, npts, n1, jac, w, eps, mu, part, out, hs(stream)));
    return NW_OK;
    NW_GUARD_END
}

// The same reduction with its partial sums in the library's fixed per-device
// array (nw_plain.hip `scratch`): no allocation of any kind in the call.
NW_API int nw_launch_diagnostics_noalloc(const double* st, int64_t S, const double* jac, const double* w,
                                         const double* eps, const double* mu, int n1, int nelt, double* out,
                                         void* stream) {
    int rc = check_diag_args(st, S, jac, w, n1, nelt, out);
    if (rc) return rc;
    const long long npts = (long long)nelt * n1 * n1 * n1;
    const int nb = nw::energy_blocks(npts);
    if (nb > nw::kEnergyScratchBlocks) {
        set_error("diagnostics_noalloc: " + std::to_string(nb) + " partial sums exceed the scratch's " +
                  std::to_string(nw::kEnergyScratchBlocks));
        return NW_ERR_ARG;
    }
    double* part = nullptr;
    if (nw::energy_scratch(&part) != 0) {
        set_error("diagnostics_noalloc: no scratch address for the current device");
        return NW_ERR_HIP;
    }
    NW_TRY_HIP(nw::launch_energy(st, S, npts, n1, jac, w, eps, mu, part, out, hs(stream)));
    return NW_OK;
}

// ------------------------------------------------------ field snapshots ----
struct nw_fields {
    nw::FieldWriter w;
};

NW_API int nw_fields_open(const char* h5_path, const nw_mesh* m, int export_continuous, int save_flags,
                          nw_fields** out) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(h5_path && m && out, "fields_open: null argument");
    NW_CHECK_ARG((save_flags & ~31) == 0, "fields_open: unknown save flag (derived curl/div fields are not offered)");
    *out = nullptr;
    nw_fields* f = new nw_fields;
    if (f->w.open(h5_path, m->m, export_continuous != 0, save_flags)) {
        delete f;
        set_error(std::string("fields_open: cannot create the snapshot files for ") + h5_path);
        return NW_ERR_ARG;
    }
    *out = f;
    return NW_OK;
    NW_GUARD_END
}

NW_API int nw_fields_write(nw_fields* f, int step, double time, const double* state, size_t n) {
    NW_GUARD_BEGIN
    NW_CHECK_ARG(f && state, "fields_write: null argument");
    NW_CHECK_ARG(n == size_t(6 * f->w.npts), "fields_write: state size must be 6*npts");
    if (f->w.write(step, time, state)) {
        set_error("fields_write: writing the snapshot files failed");
        return NW_ERR_ARG;
    }
    return NW_OK;
    NW_GUARD_END
}

NW_API void nw_fields_close(nw_fields* f) {
    if (!f) return;
    try {
        f->w.close();
    } catch (...) {
    }
    delete f;
}

// ---------------------------------------------------------------- rng ----
NW_API int nw_rand_uniform(uint64_t seed, size_t n, double* out) {
    NW_CHECK_ARG(out || n == 0, "null out");
    std::mt19937_64 g(seed);
    for (size_t x = 0; x < n; ++x) out[x] = double(g() >> 11) * (1.0 / 9007199254740992.0);
    return NW_OK;
}
NW_API int nw_rand_normal(uint64_t seed, size_t n, double* out) {
    NW_CHECK_ARG(out || n == 0, "null out");
    std::mt19937_64 g(seed);
    for (size_t x = 0; x < n; ++x) {
        const double u1 = 1.0 - double(g() >> 11) * (1.0 / 9007199254740992.0);
        const double u2 = double(g() >> 11) * (1.0 / 9007199254740992.0);
        out[x] = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    return NW_OK;
}

// ------------------------------------------------------------- device ----
NW_API int nw_dev_alloc(size_t bytes, void** p) {
    NW_CHECK_ARG(p, "null out");
    int st = ensure_tables();
    if (st) return st;
    NW_TRY_HIP(hipMalloc(p, bytes ? bytes : 16));
    return NW_OK;
}
NW_API int nw_dev_free(void* p) {
    NW_TRY_HIP(hipFree(p));
    return NW_OK;
}
NW_API int nw_dev_h2d(void* dst, const void* src, size_t bytes) {
    NW_TRY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return NW_OK;
}
NW_API int nw_dev_d2h(void* dst, const void* src, size_t bytes) {
    NW_TRY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return NW_OK;
}
NW_API int nw_dev_sync(void) {
    NW_TRY_HIP(hipDeviceSynchronize());
    return NW_OK;
}
NW_API int nw_bench_copy(size_t bytes, int reps, double* ms) {
    NW_CHECK_ARG(ms && reps > 0 && bytes >= 16, "bad argument");
    void *a = nullptr, *b = nullptr;
    NW_TRY_HIP(hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e != hipSuccess) {
        (void)hipFree(a);
        set_error("hipMalloc failed");
        return NW_ERR_ALLOC;
    }
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = NW_OK;
    float f = 0.0f;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        set_error("stream/event creation failed");
        st = NW_ERR_HIP;
    }
    if (!st) {
        (void)hipMemsetAsync(a, 0, bytes, s);
        (void)nw::launch_copy(a, b, (long long)bytes, s);  // warm-up
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) (void)nw::launch_copy(a, b, (long long)bytes, s);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&f, e0, e1) != hipSuccess) {
            set_error("copy benchmark failed");
            st = NW_ERR_HIP;
        }
        *ms = f / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return st;
}
