// nw_internal.hpp -- shared definitions of the nekwave_amd library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/nekwave_amd.h"

namespace nw {

constexpr int kMaxN1 = 10;          // reference limit: Np <= 1024 (dg_kernels.cu:478)
constexpr int kMinN1 = 2;

// Thread-local error text behind nw_last_error().
void set_error(const std::string& msg);
const char* last_error();

#define NW_TRY_HIP(expr)                                                              \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            ::nw::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
            return NW_ERR_HIP;                                                        \
        }                                                                             \
    } while (0)

#define NW_CHECK_ARG(cond, msg)                                                       \
    do {                                                                              \
        if (!(cond)) {                                                                \
            ::nw::set_error(msg);                                                     \
            return NW_ERR_ARG;                                                        \
        }                                                                             \
    } while (0)

// Per-n1 GLL tables uploaded once per device into constant memory.
struct GllTables {
    double z[kMaxN1 + 1][kMaxN1];
    double w[kMaxN1 + 1][kMaxN1];
    double iw[kMaxN1 + 1][kMaxN1];               // 1 / w
    double D[kMaxN1 + 1][kMaxN1 * kMaxN1];       // row-major here: D[i*n1 + l] = D_il
};

// Host-side GLL (nodes, weights, column-major D as the reference stores it).
void gll_host(int n1, double* z, double* w, double* Dcol);

// Per-element geometry record consumed by the fused kernels.  The trilinear
// map x(r,s,t) = c0 + c1 r + c2 s + c3 t + c4 rs + c5 rt + c6 st + c7 rst per
// coordinate (24 doubles), plus a flag telling that c4..c7 vanish (affine).
struct ElemGeom {
    double c[3][8];
};

// Per-element material record: 1/eps, 1/mu, Z = sqrt(mu/eps), Y = 1/Z.
struct ElemMat {
    double ieps, imu, Z, Y;
};

// Face neighbour code per (element, face): >= 0 neighbour element * 8 + face,
// -1 mirror (PEC/PMC/unmatched), <= -2 ghost face g = -(code + 2).
inline int face_code_neighbor(int e, int f) { return e * 8 + f; }

int upload_gll_tables(int device);
const GllTables& host_gll_tables();

}  // namespace nw
