// nw_solver.hpp -- DgSolver replacement (device residency and stage sequencing).
#pragma once

#include <vector>

#include "nw_mesh.hpp"

namespace nw {

struct Solver {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int prec = 0, pipe = 1, option = 0;
    bool use_graph = true, initialized = false, uploaded = false, media = false;
    int n1 = 0, np = 0, nelt = 0, kernels_per_step = 0;
    long long npts = 0, tfp = 0, nhalo = 0, stride = 0, device_bytes = 0;
    double c0 = 1.0, last_dt = -1.0;
    int cur = 0;
    std::vector<void*> allocs;

    // shared
    double* d_dt = nullptr;
    double* d_diag_part = nullptr;
    double* d_diag_out = nullptr;
    double* d_stage = nullptr;
    ElemMat* d_mat = nullptr;
    ElemGeom* d_geom = nullptr;
    // halo stand-in
    int *d_send_vol = nullptr, *d_off = nullptr, *d_sizes = nullptr, *d_ghost_elem = nullptr;
    ElemMat* d_gmat = nullptr;
    void *d_sendbuf = nullptr, *d_recvbuf = nullptr;
    // fused pipeline
    void* d_u[2] = {nullptr, nullptr};
    void* d_k = nullptr;
    unsigned char* d_affine = nullptr;
    int* d_fcode = nullptr;
    // plain pipeline (reference residency)
    double *d_state = nullptr, *d_kk = nullptr, *d_rhs = nullptr, *d_D = nullptr, *d_w3 = nullptr;
    double* d_met[13] = {};  // x y z jac rx sx tx ry sy ty rz sz tz
    int *d_vM = nullptr, *d_vP = nullptr, *d_bc = nullptr;
    double* d_face[4] = {};  // nx ny nz dA
    double *d_fEN = nullptr, *d_fHN = nullptr, *d_flux = nullptr, *d_eps = nullptr, *d_mu = nullptr;
    // graphs per starting parity
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};

    ~Solver();
    void release();
    int init(const Mesh& m, double c0, const nw_opts& o);
    int upload(const double* host, size_t n);
    int download(double* host, size_t n);
    int step(double dt, int nsteps);
    int launch_steps(double dt, int nsteps);
    int time(double dt, int nsteps, double* ms);
    int diag(nw_diag* out);

   private:
    int dalloc(void** p, size_t bytes);
    template <typename T>
    int upload_vec(T** dst, const std::vector<T>& v);
    int set_dt(double dt);
    int enqueue_fused_stage(int s, int parity);
    int enqueue_plain_stage(int s);
    int enqueue_step(int parity);
    int build_graph(int parity);
};

}  // namespace nw

struct nw_dg {
    nw::Solver s;
};
