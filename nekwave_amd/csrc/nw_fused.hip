// nw_fused.hip -- one fused LSRK45 stage per element group (gfx950).
//
// Computes, for every node of every element, in one pass over HBM:
//   rhs_E = (1/eps) [ curl H + M^-1 sum_faces dA F_E ]      (dg_kernels.cu:41-155, 316-382)
//   rhs_H = (1/mu)  [-curl E + M^-1 sum_faces dA F_H ]      (dg_kernels.cu:220-304 flux)
//   k     = a k + dt rhs ;  u_out = u_in + b k              (dg_kernels.cu:393-408)
// which is the reference's 7-kernel stage (dg_solver.cpp:393-505) with the
// restrict/flux/lift/inverse-mass/update passes folded into the element's
// workgroup.  u ping-pongs (neighbours read the pre-stage trace of u_in);
// k is element-local and updated in place.
//
// Work layout: a workgroup owns EPW elements; thread (i, j) of an element owns
// the k-column of nodes (i, j, 0..n1-1).  The element's six fields sit in LDS
// for the r/s contractions; the t contraction runs from the thread's own
// registers.  Each thread also evaluates point (p=i, q=j) of all six faces:
// every face point of the element is computed exactly once, with all lanes
// busy.  The face terms dA*F go through LDS and each node gathers them in
// fixed face order 0..5, so the lift is deterministic and atomic-free.
//
// Geometry stays in FP64: the trilinear map's Jacobian, its cofactors (J J^-T,
// mesh.cpp:291-309) and the Nanson normals/areas (mesh.cpp:1011-1038) are
// evaluated in-kernel from 24 per-element coefficients instead of streaming
// 10 metric arrays per node; affine elements take per-element constants.
#include "nw_fused.hpp"

namespace nw {
namespace {

__constant__ GllTables c_gll;
__constant__ float c_Df[kMaxN1 + 1][kMaxN1 * kMaxN1];
__constant__ float c_iwf[kMaxN1 + 1][kMaxN1];

template <typename R> __device__ __forceinline__ R gD(int n1, int idx);
template <> __device__ __forceinline__ double gD<double>(int n1, int idx) { return c_gll.D[n1][idx]; }
template <> __device__ __forceinline__ float gD<float>(int n1, int idx) { return c_Df[n1][idx]; }

// Jacobian entries at reference point (r, s, t) from the tensor coefficients.
struct Jac {
    double xr, yr, zr, xs, ys, zs, xt, yt, zt;
};
__device__ __forceinline__ Jac eval_jac(const ElemGeom* __restrict__ G, double r, double s, double t) {
    Jac J;
    double v[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double* c = G->c[d];
        v[d][0] = c[1] + c[4] * s + c[5] * t + c[7] * (s * t);
        v[d][1] = c[2] + c[4] * r + c[6] * t + c[7] * (r * t);
        v[d][2] = c[3] + c[5] * r + c[6] * s + c[7] * (r * s);
    }
    J.xr = v[0][0]; J.xs = v[0][1]; J.xt = v[0][2];
    J.yr = v[1][0]; J.ys = v[1][1]; J.yt = v[1][2];
    J.zr = v[2][0]; J.zs = v[2][1]; J.zt = v[2][2];
    return J;
}

// Cofactor matrix C = det(J) J^-T, rows r, s, t (mesh.cpp:298-309 times J).
struct Cof {
    double rx, ry, rz, sx, sy, sz, tx, ty, tz, det;
};
__device__ __forceinline__ Cof cofactors(const Jac& J) {
    Cof C;
    C.rx = J.ys * J.zt - J.yt * J.zs;
    C.sx = J.yt * J.zr - J.yr * J.zt;
    C.tx = J.yr * J.zs - J.ys * J.zr;
    C.ry = J.xt * J.zs - J.xs * J.zt;
    C.sy = J.xr * J.zt - J.xt * J.zr;
    C.ty = J.xs * J.zr - J.xr * J.zs;
    C.rz = J.xs * J.yt - J.xt * J.ys;
    C.sz = J.xt * J.yr - J.xr * J.yt;
    C.tz = J.xr * J.ys - J.xs * J.yr;
    C.det = J.xr * C.rx + J.xs * C.sx + J.xt * C.tx;
    return C;
}

// Outward Nanson vector of face f (mesh.cpp:1013-1032): +-(cofactor row).
__device__ __forceinline__ void face_normal(int f, const Cof& C, double& Nx, double& Ny, double& Nz) {
    switch (f) {
        case 0: Nx = -C.sx; Ny = -C.sy; Nz = -C.sz; break;
        case 1: Nx = C.rx; Ny = C.ry; Nz = C.rz; break;
        case 2: Nx = C.sx; Ny = C.sy; Nz = C.sz; break;
        case 3: Nx = -C.rx; Ny = -C.ry; Nz = -C.rz; break;
        case 4: Nx = -C.tx; Ny = -C.ty; Nz = -C.tz; break;
        default: Nx = C.tx; Ny = C.ty; Nz = C.tz; break;
    }
}

template <int N1, int EPW, typename R, typename KR>
__global__ __launch_bounds__(EPW* N1* N1) void k_fused_stage(FusedArgs A) {
    constexpr int N2 = N1 * N1, NP = N1 * N2;
    __shared__ R sU[EPW][6][NP];
    __shared__ R sF[EPW][6][N2][6];

    // XCD-aware remap: hardware dispatches blocks round-robin over the 8 XCDs;
    // give each XCD a contiguous run of elements so face neighbours share its L2.
    const int hb = blockIdx.x;
    const int L = (hb & 7) * A.nchunk + (hb >> 3);
    if (L * EPW >= A.nelt) return;  // whole block: uniform exit before any barrier

    const int tid = threadIdx.x;
    const int es = tid / N2;
    const int l = tid - es * N2;
    const int i = l % N1, j = l / N1;
    const int e = L * EPW + es;
    const bool live = e < A.nelt;
    const int ee = live ? e : A.nelt - 1;  // clamped: every read stays in bounds

    const R* __restrict__ uin = static_cast<const R*>(A.uin);
    R* __restrict__ uout = static_cast<R*>(A.uout);
    KR* __restrict__ kb = static_cast<KR*>(A.kbuf);
    const long long S = A.stride, KS = A.kstride;
    const long long nb = A.node_base + (long long)ee * NP + l;
    const ElemGeom* __restrict__ G = A.geom + ee;
    const bool aff = A.affine[ee] != 0;
    const ElemMat M = A.mat[ee];

    // ---- phase L: the element's fields -> registers (own column) + LDS ----
    R col[6][N1];
#pragma unroll
    for (int k = 0; k < N1; ++k) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const R v = uin[c * S + nb + N2 * k];
            col[c][k] = v;
            sU[es][c][l + N2 * k] = v;
        }
    }
    __syncthreads();

    // ---- phase F: point (i, j) of each of the six faces ----
    Cof Ca;  // affine elements: one cofactor matrix for the whole element
    if (aff) Ca = cofactors(eval_jac(G, 0.0, 0.0, 0.0));
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        int fi, fj, fk;
        face_ijk(f, N1, i, j, fi, fj, fk);
        const int ml = fi + N1 * fj + N2 * fk;
        R um[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) um[c] = sU[es][c][ml];
        const int code = A.fcode[ee * 6 + f];
        R jE[3], jH[3];
        double Zp = M.Z, Yp = M.Y;
        if (code == -1) {  // PEC mirror (dg_kernels.cu:244-251)
#pragma unroll
            for (int c = 0; c < 3; ++c) { jE[c] = R(-2) * um[c]; jH[c] = R(0); }
        } else {
            long long idx;
            if (code >= 0) {
                const int e2 = code >> 6, f2 = (code >> 3) & 7, o = code & 7;
                int p2, q2, i2, j2, k2;
                orient_map(o, N1, i, j, p2, q2);
                face_ijk(f2, N1, p2, q2, i2, j2, k2);
                idx = A.node_base + (long long)e2 * NP + i2 + N1 * j2 + N2 * k2;
                const ElemMat M2 = A.mat[e2];
                Zp = M2.Z; Yp = M2.Y;
            } else {
                const int g = -(code + 2);
                idx = A.ghost_base + (long long)g * N2 + i + N1 * j;
                const ElemMat M2 = A.gmat[g];
                Zp = M2.Z; Yp = M2.Y;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                jE[c] = uin[c * S + idx] - um[c];
                jH[c] = uin[(c + 3) * S + idx] - um[c + 3];
            }
        }
        // Nanson normal and area in FP64
        Cof C = aff ? Ca : cofactors(eval_jac(G, c_gll.z[N1][fi], c_gll.z[N1][fj], c_gll.z[N1][fk]));
        double Nx, Ny, Nz;
        face_normal(f, C, Nx, Ny, Nz);
        const double Ln = sqrt(Nx * Nx + Ny * Ny + Nz * Nz);
        const double il = Ln > 1e-14 ? 1.0 / Ln : 0.0;
        const R nx = R(Nx * il), ny = R(Ny * il), nz = R(Nz * il);
        const R dA = R(Ln * c_gll.w[N1][i] * c_gll.w[N1][j]);
        // central / upwind flux (dg_kernels.cu:265-294), impedance weighted with media
        const R nxEx = ny * jE[2] - nz * jE[1], nxEy = nz * jE[0] - nx * jE[2], nxEz = nx * jE[1] - ny * jE[0];
        const R nxHx = ny * jH[2] - nz * jH[1], nxHy = nz * jH[0] - nx * jH[2], nxHz = nx * jH[1] - ny * jH[0];
        const R nnEx = ny * nxEz - nz * nxEy, nnEy = nz * nxEx - nx * nxEz, nnEz = nx * nxEy - ny * nxEx;
        const R nnHx = ny * nxHz - nz * nxHy, nnHy = nz * nxHx - nx * nxHz, nnHz = nx * nxHy - ny * nxHx;
        const R c0 = R(A.c0);
        R wEh, wEe, wHe, wHh;  // F_E = wEh nx[[H]] - wEe nxnx[[E]];  F_H = -wHe nx[[E]] - wHh nxnx[[H]]
        if (A.weighted) {
            const double iz = 1.0 / (Zp + M.Z), iy = 1.0 / (Yp + M.Y);
            wEh = R(Zp * iz); wEe = R(A.c0 * iz);
            wHe = R(Yp * iy); wHh = R(A.c0 * iy);
        } else {
            wEh = R(0.5); wEe = R(0.5) * c0; wHe = R(0.5); wHh = R(0.5) * c0;
        }
        R* out = sF[es][f][i + N1 * j];
        out[0] = dA * (wEh * nxHx - wEe * nnEx);
        out[1] = dA * (wEh * nxHy - wEe * nnEy);
        out[2] = dA * (wEh * nxHz - wEe * nnEz);
        out[3] = dA * (-wHe * nxEx - wHh * nnHx);
        out[4] = dA * (-wHe * nxEy - wHh * nnHy);
        out[5] = dA * (-wHe * nxEz - wHh * nnHz);
    }
    __syncthreads();

    // ---- phase V: volume curl + lift + inverse mass + LSRK update ----
    R Di[N1], Dj[N1];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
        Di[q] = gD<R>(N1, i * N1 + q);
        Dj[q] = gD<R>(N1, j * N1 + q);
    }
    const double dt = *A.dt;
    const double r0 = c_gll.z[N1][i], s0 = c_gll.z[N1][j];
    const R ieps = R(M.ieps), imu = R(M.imu);
    const R iwij = R(c_gll.iw[N1][i] * c_gll.iw[N1][j]);
#pragma unroll
    for (int k = 0; k < N1; ++k) {
        const Cof C = aff ? Ca : cofactors(eval_jac(G, r0, s0, c_gll.z[N1][k]));
        const R Crx = R(C.rx), Cry = R(C.ry), Crz = R(C.rz);
        const R Csx = R(C.sx), Csy = R(C.sy), Csz = R(C.sz);
        const R Ctx = R(C.tx), Cty = R(C.ty), Ctz = R(C.tz);
        const R iJ = R(1.0 / C.det);
        // reference derivatives of the six fields (local_grad3, dg_kernels.cu:101-124)
        R gr[6], gs[6], gt[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            R ar = R(0), as = R(0), at = R(0);
#pragma unroll
            for (int q = 0; q < N1; ++q) {
                ar += Di[q] * sU[es][c][q + N1 * j + N2 * k];
                as += Dj[q] * sU[es][c][i + N1 * q + N2 * k];
                at += gD<R>(N1, k * N1 + q) * col[c][q];
            }
            gr[c] = ar; gs[c] = as; gt[c] = at;
        }
        // J * curl u in cofactor form (dg_kernels.cu:140-149 with w*J*J^-T = W*C)
        auto dx = [&](int c) { return gr[c] * Crx + gs[c] * Csx + gt[c] * Ctx; };
        auto dy = [&](int c) { return gr[c] * Cry + gs[c] * Csy + gt[c] * Cty; };
        auto dz = [&](int c) { return gr[c] * Crz + gs[c] * Csz + gt[c] * Ctz; };
        R cH[3] = {dy(5) - dz(4), dz(3) - dx(5), dx(4) - dy(3)};
        R cE[3] = {dy(2) - dz(1), dz(0) - dx(2), dx(1) - dy(0)};
        // lift: face contributions in face order 0..5
        R LE[3] = {R(0), R(0), R(0)}, LH[3] = {R(0), R(0), R(0)};
        auto add = [&](int f, int pq) {
            const R* s = sF[es][f][pq];
            LE[0] += s[0]; LE[1] += s[1]; LE[2] += s[2];
            LH[0] += s[3]; LH[1] += s[4]; LH[2] += s[5];
        };
        if (j == 0) add(0, i + N1 * k);
        if (i == N1 - 1) add(1, j + N1 * k);
        if (j == N1 - 1) add(2, i + N1 * k);
        if (i == 0) add(3, j + N1 * k);
        if (k == 0) add(4, i + N1 * j);
        if (k == N1 - 1) add(5, i + N1 * j);
        const R Winv = iwij * R(c_gll.iw[N1][k]);
        R rhs[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rhs[c] = ieps * iJ * (cH[c] + LE[c] * Winv);
            rhs[c + 3] = imu * iJ * (-cE[c] + LH[c] * Winv);
        }
        if (live) {
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const long long ix = nb + N2 * k;
                const KR kold = A.first_stage ? KR(0) : kb[c * KS + ix];
                const KR kn = KR(A.a) * kold + KR(dt) * KR(rhs[c]);
                kb[c * KS + ix] = kn;
                uout[c * S + ix] = R(KR(col[c][k]) + KR(A.b) * kn);
            }
        }
    }
}

// Elements per workgroup, two options per n1 (index 0 is the default), capped
// so one workgroup's LDS stays within 64 KiB (two or more workgroups per CU).
constexpr int kEpw[kMaxN1 + 1][2] = {
    {1, 1}, {1, 1}, {16, 32}, {7, 14}, {4, 8}, {5, 10}, {7, 2}, {4, 2}, {1, 2}, {3, 1}, {1, 1}};

constexpr int lds_per_elem(int n1, int rsz) { return (6 * n1 * n1 * n1 + 36 * n1 * n1) * rsz; }
constexpr int epw_of(int n1, int opt, int rsz) {
    const int want = kEpw[n1][opt];
    const int cap = 65536 / lds_per_elem(n1, rsz);
    return want < cap ? want : (cap > 0 ? cap : 1);
}

template <int N1, int OPT, typename R, typename KR>
hipError_t launch_one(const FusedArgs& a, hipStream_t s) {
    constexpr int EPW = epw_of(N1, OPT, sizeof(R));
    const int nblocks = (a.nelt + EPW - 1) / EPW;
    FusedArgs b = a;
    b.nchunk = (nblocks + 7) / 8;
    hipLaunchKernelGGL((k_fused_stage<N1, EPW, R, KR>), dim3(8 * b.nchunk), dim3(EPW * N1 * N1), 0, s, b);
    return hipGetLastError();
}

using LaunchFn = hipError_t (*)(const FusedArgs&, hipStream_t);

template <int N1, typename R, typename KR>
LaunchFn pick(int option) {
    return option == 0 ? &launch_one<N1, 0, R, KR> : &launch_one<N1, 1, R, KR>;
}

template <typename R, typename KR>
LaunchFn pick_n1(int n1, int option) {
    switch (n1) {
        case 2: return pick<2, R, KR>(option);
        case 3: return pick<3, R, KR>(option);
        case 4: return pick<4, R, KR>(option);
        case 5: return pick<5, R, KR>(option);
        case 6: return pick<6, R, KR>(option);
        case 7: return pick<7, R, KR>(option);
        case 8: return pick<8, R, KR>(option);
        case 9: return pick<9, R, KR>(option);
        default: return nullptr;
    }
}

}  // namespace

int fused_epw(int n1, int option, int prec) {
    if (n1 < 0 || n1 > kMaxN1) return 0;
    return epw_of(n1, option ? 1 : 0, prec == kFP64 ? 8 : 4);
}

bool fused_supported(int n1, int prec) { return n1 >= 2 && n1 <= 9 && prec >= 0 && prec < kNumPrec; }

hipError_t launch_fused(int n1, int prec, int option, const FusedArgs& a, hipStream_t s) {
    LaunchFn fn = nullptr;
    if (prec == kFP64) fn = pick_n1<double, double>(n1, option);
    else if (prec == kFP32) fn = pick_n1<float, float>(n1, option);
    else if (prec == kFP32K64) fn = pick_n1<float, double>(n1, option);
    if (!fn) return hipErrorInvalidValue;
    if (a.nelt <= 0) return hipSuccess;
    return fn(a, s);
}

int fused_upload_tables(const GllTables& t) {
    static float df[kMaxN1 + 1][kMaxN1 * kMaxN1];
    static float iwf[kMaxN1 + 1][kMaxN1];
    for (int n = 0; n <= kMaxN1; ++n) {
        for (int x = 0; x < kMaxN1 * kMaxN1; ++x) df[n][x] = static_cast<float>(t.D[n][x]);
        for (int x = 0; x < kMaxN1; ++x) iwf[n][x] = static_cast<float>(t.iw[n][x]);
    }
    NW_TRY_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gll), &t, sizeof(GllTables)));
    NW_TRY_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_Df), df, sizeof(df)));
    NW_TRY_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_iwf), iwf, sizeof(iwf)));
    return NW_OK;
}

}  // namespace nw
