// nw_plain.hpp -- launchers of the reference-decomposed FP64 kernels and the
// device-side helpers (geometry setup, diagnostics, conversion, halo).
#pragma once

#include "nw_internal.hpp"

namespace nw {

struct PlainVolumeArgs {
    const double *u1, *u2, *u3;
    double *w1, *w2, *w3;
    const double *D, *w3g, *jac;
    const double *rx, *sx, *tx, *ry, *sy, *ty, *rz, *sz, *tz;
    double sign;
    int n1, nelt;
};

int plain_upload_tables(const GllTables& t);

hipError_t launch_volume_curl(const PlainVolumeArgs& a, hipStream_t s);
hipError_t launch_restrict(const double* st, double* fEN, double* fHN, const int* vM, long long tfp, long long S,
                           hipStream_t s);
hipError_t launch_flux(const double* st, const double* fEN, const double* fHN, const int* vP, const int* bc,
                       const double* nx, const double* ny, const double* nz, double* flux, double c0, long long tfp,
                       long long S, const int* vM, const double* eps, const double* mu, const int* ghost_elem, int np,
                       long long npts, hipStream_t s);
hipError_t launch_lift(const double* flux, const double* dA, double* rhs, long long tfp, long long npts, int n1,
                       hipStream_t s);
// launch_add_flux with the reference's arguments (dg_kernels.hpp:46-50): nodes of
// the element-major face layout gather their points in face order
// (deterministic); points whose volIdxMinus departs from that layout are added
// with atomics as the reference adds all of them.
hipError_t launch_add_flux(const double* flux, const int* vM, const double* dA, double* rhs, long long tfp,
                           long long npts, hipStream_t s);
hipError_t launch_inv_mass(double* rhs, const double* jac, const double* w3, int nxyz, long long npts,
                           const double* eps, const double* mu, hipStream_t s);
hipError_t launch_lsrk(double* st, double* k, const double* rhs, double a, double b, const double* dtp, double dt,
                       long long npts, long long S, hipStream_t s);
// State layouts: component c of node p at c * cs + p * ns (the reference's
// blocks [Ex|..|Hz]: cs = stride, ns = 1; node-major: cs = 1, ns = 6).
template <typename R>
hipError_t launch_pack(const R* st, R* buf, const int* vol, const int* off, const int* sizes, int nh, long long cs,
                       long long ns, hipStream_t s);
template <typename R>
hipError_t launch_unpack(const R* buf, R* st, const int* off, const int* sizes, int nh, long long npts, long long cs,
                         long long ns, hipStream_t s);
hipError_t launch_metrics(const double* corners, int n1, long long npts, double* const* out13, hipStream_t s);
hipError_t launch_face_data(const int* fcode, const int* fbc, int n1, long long tfp, long long npts,
                            const double* const* met10, int* vM, int* vP, int* bc, double* nx, double* ny, double* nz,
                            double* dA, hipStream_t s);
int diag_blocks();
template <typename R>
hipError_t launch_diag(const R* u, long long cs, long long ns, long long npts, int n1, const double* jac,
                       const ElemGeom* geom, const ElemMat* mat, double* part, double* out, hipStream_t s);
// Energy, max |E|, max |H| of a reference-layout FP64 state into out[0..2]
// (Case::computeTotalEnergy / getMaxE / getMaxH, case.cpp:733-794): one partial
// triple per block into part (3 energy_blocks(npts) doubles, every one written,
// none read before), then one reduction in block order.
constexpr int kEnergyMaxBlocks = 1024;
int energy_blocks(long long npts);
// The device address of this GPU's fixed partial-sum array (3 kEnergyScratchBlocks
// doubles, 24 KiB): looked up once per device with that device current; 0 on
// success, negative when the lookup failed.
constexpr int kEnergyScratchBlocks = 1024;
constexpr int kMaxDevices = 64;
int energy_scratch(double** out);
hipError_t launch_energy(const double* st, long long S, long long npts, int n1, const double* jac, const double* w,
                         const double* eps, const double* mu, double* part, double* out, hipStream_t s);
// Sums of (u - ref)^2 dV and ref^2 dV over all components and nodes into
// out[0], out[1] (FP64; ref in component blocks of npts).
template <typename R>
hipError_t launch_l2(const R* u, long long cs, long long ns, const double* ref, long long npts, int n1, const double* jac,
                     const ElemGeom* geom, double* part, double* out, hipStream_t s);
// One probe record of the state: history[(*slot mod cap)][probe][component],
// then ++*slot (one block; weights lw = Lr, Ls, Lt per probe, 3 n1 doubles).
template <typename R>
hipError_t launch_probe(const R* u, long long cs, long long ns, int nprobes, const long long* base, const double* lw,
                        int n1, double* hist, int* slot, int cap, hipStream_t s);
hipError_t launch_zero_int(int* p, hipStream_t s);
template <typename D, typename S>
hipError_t launch_convert(D* dst, long long dcs, long long dns, const S* src, long long scs, long long sns, long long n,
                          int ncomp, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, long long bytes, hipStream_t s);
hipError_t launch_fill(double* p, long long n, double v, hipStream_t s);

}  // namespace nw
