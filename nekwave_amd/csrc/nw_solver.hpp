// nw_solver.hpp -- DgSolver replacement (device residency and stage sequencing)
// for one mesh or a batch of independent meshes.
#pragma once

#include <vector>

#include "nw_fused.hpp"
#include "nw_mesh.hpp"

namespace nw {

// One independent mesh inside the solver's buffers.
struct Instance {
    int n1 = 0, np = 0, nelt = 0, elem_base = 0;
    long long node_base = 0, npts = 0;
    bool media = false;
};

// A run of elements sharing n1: one fused launch per stage.
struct Segment {
    int n1 = 0, nelt = 0, elem_base = 0, option = 0;
    int stream = -1;  // non-temporal k accesses: 1 / 0 as tuned, -1 the default rule (fused_stream_k)
    bool affine = false;  // every element a parallelepiped: records hold the geometry
    long long node_base = 0;
    long long kbase = 0;                   // first k chunk (the most any tile option needs, after earlier segments)
    long long goff[kNumTileOptions] = {};  // offset of this segment's group order, per tile option
    long long soff[kNumTileOptions] = {};  // offset of its workgroup sequence offsets, per tile option
    int nseq[kNumTileOptions] = {};        // its workgroup sequences, per tile option
};

// FNV-1a hash of what makes a mesh this mesh: n1, corner coordinates (bit
// patterns) and the face links.  Two meshes of the same shape but other bounds,
// jitter or connectivity differ.
unsigned long long mesh_signature(const Mesh& m);

struct Solver {
    int device = -1;
    int groups_per_cu = 0;  // cap on fused workgroups per CU (0: as many as fit)
    int walk_mode = -1;     // walk direction of the fused stages (set_walk)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int prec = 0, pipe = 1;
    bool use_graph = true, initialized = false, uploaded = false, media = false, autotune = false, tuned = false;
    int n1 = 0, np = 0, nelt = 0, kernels_per_step = 0;
    unsigned long long mesh_sig = 0;  // mesh_signature of a single-mesh solver's mesh
    long long npts = 0, tfp = 0, nhalo = 0, stride = 0, device_bytes = 0;
    long long kchunks = 0;  // fused pipeline: k chunks of 6 x 64 values (Segment::kbase)
    double c0 = 1.0, last_dt = -1.0;
    int cur = 0;
    std::vector<Instance> inst;  // in caller order
    std::vector<Segment> seg;
    std::vector<void*> allocs;

    // shared
    double* d_dt = nullptr;
    double* d_diag_part = nullptr;
    double* d_diag_out = nullptr;
    double* d_stage = nullptr;
    ElemMat* d_mat = nullptr;
    ElemGeom* d_geom = nullptr;
    // halo stand-in (single mesh)
    int *d_send_vol = nullptr, *d_off = nullptr, *d_sizes = nullptr, *d_ghost_elem = nullptr;
    ElemMat* d_gmat = nullptr;
    void *d_sendbuf = nullptr, *d_recvbuf = nullptr;
    // fused pipeline
    void* d_u[2] = {nullptr, nullptr};
    void* d_k = nullptr;
    int4* d_fnb = nullptr;
    void* d_fw = nullptr;  // flux weights per element face, 4 values of the field type
    double* d_rec = nullptr;  // per element: 24 trilinear coefficients, 1/eps, 1/mu
    int* d_gorder = nullptr;  // per segment and tile option: the order the fused kernel visits groups in
    int* d_ghead = nullptr;   // per segment and tile option: kFusedHead words per workgroup sequence
    int* d_gorder_rev = nullptr;  // the same walks with every sequence visited backwards
    int* d_ghead_rev = nullptr;
    // plain pipeline (reference residency, single mesh)
    double *d_state = nullptr, *d_kk = nullptr, *d_rhs = nullptr, *d_D = nullptr, *d_w3 = nullptr;
    double* d_met[13] = {};  // x y z jac rx sx tx ry sy ty rz sz tz
    int *d_vM = nullptr, *d_vP = nullptr, *d_bc = nullptr;
    double* d_face[4] = {};  // nx ny nz dA
    double *d_fEN = nullptr, *d_fHN = nullptr, *d_flux = nullptr, *d_eps = nullptr, *d_mu = nullptr;
    // probes (device-side ProbeManager): per probe the node base of its element and
    // the 1-D Lagrange weights Lr, Ls, Lt; a history of `probe_cap` records
    int nprobes = 0, probe_cap = 0;
    long long* d_probe_base = nullptr;
    double* d_probe_w = nullptr;
    double* d_probe_hist = nullptr;
    int* d_probe_slot = nullptr;
    // graphs per starting parity
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};

    ~Solver();
    void release();
    int init(const Mesh* const* meshes, int count, double c0, const nw_opts& o);
    int upload(int instance, const double* host, size_t n);
    int download(int instance, double* host, size_t n);
    int step(double dt, int nsteps);
    int launch_steps(double dt, int nsteps);
    int time(double dt, int warmup, int nsteps, int repeats, double* ms);
    int diag(int instance, nw_diag* out);
    int rel_l2(int instance, const double* ref, size_t n, double* out);
    int set_probes(const Mesh& m, int n, const double* xyz, int capacity);
    int probe_sample();
    int probe_history(double* host, int max_records, int* nrecords);
    int tune();
    int set_segment_option(int s, int option, int stream_k);
    int set_walk(int mode);  // -1 auto (fused_alternate_walk), 0 forwards, 1 alternating
    bool alternating_walk() const;
    // non-temporal k accesses of segment sg under a tile option: the tuned
    // choice, else the default rule (fused_stream_k)
    int stream_of(const Segment& sg, int option) const;

   private:
    int dalloc(void** p, size_t bytes);
    void dfree(void* p);
    template <typename T>
    int upload_vec(T** dst, const std::vector<T>& v);
    int init_plain(const Mesh& m);
    int set_dt(double dt);
    int launch_segment(const Segment& sg, int s, int parity, int option, int stream_k, hipStream_t st,
                       bool backwards = false);
    long long working_set_bytes() const;
    int enqueue_fused_stage(int s, int parity);
    int enqueue_plain_stage(int s);
    int enqueue_step(int parity);
    int build_graph(int parity);
    int enqueue_probe(int state_parity);
    void drop_graphs();
};

}  // namespace nw

struct nw_dg {
    nw::Solver s;
};
