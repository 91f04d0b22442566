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
// Work layout: a workgroup owns EPW elements, one thread per node.  Every HBM
// read of the stage -- the node's six field values, its six LSRK registers and
// the neighbour trace of the thread's face point -- is issued before the first
// barrier, so a wave has all its memory in flight at once.  The element's
// fields then sit in LDS for the three tensor contractions.  Face point
// (f, p, q) is computed by one thread (face points outnumber nodes only for
// n1 < 6, where a thread takes two); the terms dA*F go through LDS and each
// node gathers them in fixed face order 0..5, so the lift is deterministic and
// atomic-free.
//
// Geometry stays in FP64: the trilinear map's Jacobian, its cofactors (J J^-T,
// mesh.cpp:291-309) and the Nanson normals/areas (mesh.cpp:1011-1038) are
// evaluated in-kernel from 24 per-element coefficients instead of streaming
// ten metric arrays per node.
#include "nw_fused.hpp"

namespace nw {
namespace {

__constant__ GllTables c_gll;
__constant__ float c_Df[kMaxN1 + 1][kMaxN1 * kMaxN1];
__constant__ float c_iwf[kMaxN1 + 1][kMaxN1];

template <typename R> __device__ __forceinline__ R gD(int n1, int idx);
template <> __device__ __forceinline__ double gD<double>(int n1, int idx) { return c_gll.D[n1][idx]; }
template <> __device__ __forceinline__ float gD<float>(int n1, int idx) { return c_Df[n1][idx]; }

// Tangent vectors d x/dr, d x/ds, d x/dt of the trilinear map at (r, s, t).
struct Tangents {
    double r[3], s[3], t[3];
};
__device__ __forceinline__ Tangents tangents(const ElemGeom* G, double r, double s, double t) {
    Tangents T;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double* c = G->c[d];
        T.r[d] = c[1] + c[4] * s + c[5] * t + c[7] * (s * t);
        T.s[d] = c[2] + c[4] * r + c[6] * t + c[7] * (r * t);
        T.t[d] = c[3] + c[5] * r + c[6] * s + c[7] * (r * s);
    }
    return T;
}

__device__ __forceinline__ void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Local node of face point (p, q) of face f (getFaceNodeVolIndex, mesh.cpp:955-981)
// as base + p*sp + q*sq.
template <int N1>
__device__ __forceinline__ int face_local_node(int f, int p, int q) {
    constexpr int N2 = N1 * N1;
    const int base = f == 1 ? N1 - 1 : f == 2 ? N1 * (N1 - 1) : f == 5 ? N2 * (N1 - 1) : 0;
    const int sp = (f == 1 || f == 3) ? N1 : 1;
    const int sq = (f >= 4) ? N1 : N2;
    return base + p * sp + q * sq;
}

// An empty volatile asm that takes and returns v: the value must be computed
// before this point in program order (the scheduler otherwise sinks pure
// arithmetic below later loads and barriers, stretching live ranges).
template <typename T>
__device__ __forceinline__ void pin(T& v) {
    asm volatile("" : "+v"(v));
}

// Occupancy target: enough waves per SIMD to cover HBM latency with a whole
// stage's reads in flight per wave.
template <int NT>
struct WavesPerEU {
    static constexpr int value = NT >= 512 ? 4 : 3;
};

template <int N1, int EPW, bool FIRST, typename R, typename KR>
__global__ __launch_bounds__(EPW* N1* N1* N1) __attribute__((amdgpu_waves_per_eu(WavesPerEU<EPW * N1 * N1 * N1>::value)))
void k_fused_stage(FusedArgs A) {
    constexpr int N2 = N1 * N1, NP = N1 * N2, NF = 6 * N2, NT = EPW * NP;
    constexpr int FPT = (EPW * NF + NT - 1) / NT;  // face points per thread (1 for n1 >= 6)
    __shared__ R sU[EPW][6][NP];
    __shared__ R sF[EPW][NF][6];
    __shared__ ElemGeom sG[EPW];

    // XCD-aware remap: the hardware deals workgroups round-robin over the 8
    // XCDs; each XCD gets a contiguous run of elements so face neighbours are
    // read from its own L2.
    const int hb = blockIdx.x;
    const int L = (hb & 7) * A.nchunk + (hb >> 3);
    if (L * EPW >= A.nelt) return;  // uniform over the block, before any barrier

    const int tid = threadIdx.x;
    const int es = tid / NP;
    const int m = tid - es * NP;
    const int i = m % N1, j = (m / N1) % N1, k = m / N2;
    const int e = L * EPW + es;
    const bool live = e < A.nelt;
    int ee = live ? e : A.nelt - 1;  // clamped: every read stays in bounds
    if (EPW == 1) ee = __builtin_amdgcn_readfirstlane(ee);
    const R* __restrict__ uin = static_cast<const R*>(A.uin);
    R* __restrict__ uout = static_cast<R*>(A.uout);
    KR* __restrict__ kb = static_cast<KR*>(A.kbuf);
    const long long S = A.stride, KS = A.kstride;
    const long long node = A.node_base + (long long)ee * NP + m;

    // ---- every HBM read of the stage is issued before the first barrier ----
    R u[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) u[c] = uin[c * S + node];
    KR kold[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) kold[c] = FIRST ? KR(0) : kb[c * KS + node];
    R up[FPT][6];
    int4 nbr[FPT];
#pragma unroll
    for (int t = 0; t < FPT; ++t) {
        const int fp = tid + t * NT;
        const int efs = fp / NF, r = fp - efs * NF;
        const int f = r / N2, pq = r - f * N2, p = pq % N1, q = pq / N1;
        const int ef = (fp < EPW * NF) ? min(L * EPW + efs, A.nelt - 1) : ee;
        const int4 nb = A.fnb[ef * 6 + f];
        nbr[t] = nb;
        const long long idx = nb.x >= 0 ? (long long)nb.x + p * nb.y + q * nb.z : node;
#pragma unroll
        for (int c = 0; c < 6; ++c) up[t][c] = uin[c * S + idx];
    }
    // rows of D for this node (constant memory), loaded with the fields
    R Di[N1], Dj[N1], Dk[N1];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
        Di[q] = gD<R>(N1, i * N1 + q);
        Dj[q] = gD<R>(N1, j * N1 + q);
        Dk[q] = gD<R>(N1, k * N1 + q);
    }
    if (EPW > 1) {
        for (int x = tid; x < EPW * 24; x += NT) {
            const int el = min(L * EPW + x / 24, A.nelt - 1);
            (&sG[0].c[0][0])[x] = (&A.geom[el].c[0][0])[x % 24];
        }
        __syncthreads();
    }

    // ---- FP64 geometry while the reads are in flight ----
    // volume: cofactor matrix C = J J^-T (mesh.cpp:291-309 times J)
    R Crx, Cry, Crz, Csx, Csy, Csz, Ctx, Cty, Ctz, iJ;
    {
        const Tangents T = tangents((EPW == 1) ? A.geom + ee : sG + es, c_gll.z[N1][i], c_gll.z[N1][j], c_gll.z[N1][k]);
        double Cr[3], Cs[3], Ct[3];
        cross3(T.s, T.t, Cr);  // (rx, ry, rz) * J
        cross3(T.t, T.r, Cs);  // (sx, sy, sz) * J
        cross3(T.r, T.s, Ct);  // (tx, ty, tz) * J
        const double det = T.r[0] * Cr[0] + T.r[1] * Cr[1] + T.r[2] * Cr[2];
        Crx = R(Cr[0]); Cry = R(Cr[1]); Crz = R(Cr[2]);
        Csx = R(Cs[0]); Csy = R(Cs[1]); Csz = R(Cs[2]);
        Ctx = R(Ct[0]); Cty = R(Ct[1]); Ctz = R(Ct[2]);
        iJ = R(1.0 / det);
    }
    // computed now, while the reads are in flight, not after the barrier
    pin(Crx); pin(Cry); pin(Crz); pin(Csx); pin(Csy); pin(Csz); pin(Ctx); pin(Cty); pin(Ctz); pin(iJ);
    // faces: Nanson vector (mesh.cpp:1011-1038) = +-(a x b) of the face's two
    // tangents: r-faces ds x dt, s-faces dt x dr, t-faces dr x ds
    R fn[FPT][4];  // nx, ny, nz, dA
#pragma unroll
    for (int t = 0; t < FPT; ++t) {
        const int fp = tid + t * NT;
        const int efs = min(fp / NF, EPW - 1), r = fp - (fp / NF) * NF;
        const int f = r / N2, pq = r - f * N2, p = pq % N1, q = pq / N1;
        const int ml = face_local_node<N1>(f, p, q);
        const int fi = ml % N1, fj = (ml / N1) % N1, fk = ml / N2;
        const Tangents T = tangents((EPW == 1) ? A.geom + ee : sG + efs, c_gll.z[N1][fi], c_gll.z[N1][fj], c_gll.z[N1][fk]);
        const bool rf = (f == 1 || f == 3), sf = (f == 0 || f == 2);
        double a3[3], b3[3], Nv[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a3[d] = rf ? T.s[d] : sf ? T.t[d] : T.r[d];
            b3[d] = rf ? T.t[d] : sf ? T.r[d] : T.s[d];
        }
        cross3(a3, b3, Nv);
        const double sg = (f == 1 || f == 2 || f == 5) ? 1.0 : -1.0;
        const double Ln = sqrt(Nv[0] * Nv[0] + Nv[1] * Nv[1] + Nv[2] * Nv[2]);
        const double il = Ln > 1e-14 ? sg / Ln : 0.0;
        fn[t][0] = R(Nv[0] * il); fn[t][1] = R(Nv[1] * il); fn[t][2] = R(Nv[2] * il);
        fn[t][3] = R(Ln * c_gll.w[N1][p] * c_gll.w[N1][q]);
        pin(fn[t][0]); pin(fn[t][1]); pin(fn[t][2]); pin(fn[t][3]);
    }

#pragma unroll
    for (int c = 0; c < 6; ++c) sU[es][c][m] = u[c];
    __syncthreads();

    // ---- face terms dA * F of this thread's face point(s) -> LDS ----
#pragma unroll
    for (int t = 0; t < FPT; ++t) {
        const int fp = tid + t * NT;
        if (fp < EPW * NF) {
            const int efs = fp / NF, r = fp - efs * NF;
            const int f = r / N2, pq = r - f * N2, p = pq % N1, q = pq / N1;
            const int ef = min(L * EPW + efs, A.nelt - 1);
            const int ml = face_local_node<N1>(f, p, q);
            const int4 nb = nbr[t];
            const ElemMat Mm = A.mat[A.elem_base + ef];
            const ElemMat Mp = nb.x < 0 ? Mm : (nb.w >= 0 ? A.mat[nb.w] : A.gmat[-(nb.w + 1)]);
            R jE[3], jH[3];
            if (nb.x < 0) {  // PEC mirror: [[E]] = -2 E-, [[H]] = 0 (dg_kernels.cu:244-251)
#pragma unroll
                for (int c = 0; c < 3; ++c) { jE[c] = R(-2) * sU[efs][c][ml]; jH[c] = R(0); }
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    jE[c] = up[t][c] - sU[efs][c][ml];
                    jH[c] = up[t][c + 3] - sU[efs][c + 3][ml];
                }
            }
            const R nx = fn[t][0], ny = fn[t][1], nz = fn[t][2], dA = fn[t][3];
            // central / upwind flux (dg_kernels.cu:265-294), impedance weighted with media
            const R nxEx = ny * jE[2] - nz * jE[1], nxEy = nz * jE[0] - nx * jE[2], nxEz = nx * jE[1] - ny * jE[0];
            const R nxHx = ny * jH[2] - nz * jH[1], nxHy = nz * jH[0] - nx * jH[2], nxHz = nx * jH[1] - ny * jH[0];
            const R nnEx = ny * nxEz - nz * nxEy, nnEy = nz * nxEx - nx * nxEz, nnEz = nx * nxEy - ny * nxEx;
            const R nnHx = ny * nxHz - nz * nxHy, nnHy = nz * nxHx - nx * nxHz, nnHz = nx * nxHy - ny * nxHx;
            R wEh, wEe, wHe, wHh;  // F_E = wEh nx[[H]] - wEe nxnx[[E]];  F_H = -wHe nx[[E]] - wHh nxnx[[H]]
            if (A.weighted) {
                const double iz = 1.0 / (Mp.Z + Mm.Z), iy = 1.0 / (Mp.Y + Mm.Y);
                wEh = R(Mp.Z * iz); wEe = R(A.c0 * iz);
                wHe = R(Mp.Y * iy); wHh = R(A.c0 * iy);
            } else {
                wEh = R(0.5); wEe = R(0.5 * A.c0); wHe = R(0.5); wHh = R(0.5 * A.c0);
            }
            R* out = sF[efs][r];
            out[0] = dA * (wEh * nxHx - wEe * nnEx);
            out[1] = dA * (wEh * nxHy - wEe * nnEy);
            out[2] = dA * (wEh * nxHz - wEe * nnEz);
            out[3] = dA * (-wHe * nxEx - wHh * nnHx);
            out[4] = dA * (-wHe * nxEy - wHh * nnHy);
            out[5] = dA * (-wHe * nxEz - wHh * nnHz);
        }
    }

    // ---- volume: each field's reference derivatives (local_grad3,
    // dg_kernels.cu:101-124) folded straight into J * curl ----
    R cE[3] = {R(0), R(0), R(0)}, cH[3] = {R(0), R(0), R(0)};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        R ar = R(0), as = R(0), at = R(0);
#pragma unroll
        for (int q = 0; q < N1; ++q) {
            ar += Di[q] * sU[es][c][q + N1 * j + N2 * k];
            as += Dj[q] * sU[es][c][i + N1 * q + N2 * k];
            at += Dk[q] * sU[es][c][i + N1 * j + N2 * q];
        }
        const R dxc = ar * Crx + as * Csx + at * Ctx;  // J d/dx (dg_kernels.cu:140-149, W*J*J^-T = W*C)
        const R dyc = ar * Cry + as * Csy + at * Cty;
        const R dzc = ar * Crz + as * Csz + at * Ctz;
        R* cu = c < 3 ? cE : cH;
        switch (c % 3) {  // curl u = (dy u_z - dz u_y, dz u_x - dx u_z, dx u_y - dy u_x)
            case 0: cu[1] += dzc; cu[2] -= dyc; break;
            case 1: cu[0] -= dzc; cu[2] += dxc; break;
            default: cu[0] += dyc; cu[1] -= dxc; break;
        }
        // pin this field's arithmetic here: one field's LDS reads live at a time
        pin(cE[0]); pin(cE[1]); pin(cE[2]); pin(cH[0]); pin(cH[1]); pin(cH[2]);
    }
    __syncthreads();  // face terms complete

    // ---- lift: this node's face terms in face order 0..5 (deterministic) ----
    R LE[3] = {R(0), R(0), R(0)}, LH[3] = {R(0), R(0), R(0)};
    auto add = [&](int f, int pq) {
        const R* s = sF[es][f * N2 + pq];
        LE[0] += s[0]; LE[1] += s[1]; LE[2] += s[2];
        LH[0] += s[3]; LH[1] += s[4]; LH[2] += s[5];
    };
    if (j == 0) add(0, i + N1 * k);
    if (i == N1 - 1) add(1, j + N1 * k);
    if (j == N1 - 1) add(2, i + N1 * k);
    if (i == 0) add(3, j + N1 * k);
    if (k == 0) add(4, i + N1 * j);
    if (k == N1 - 1) add(5, i + N1 * j);

    // ---- inverse mass + LSRK45 update (dg_kernels.cu:353-408) ----
    const ElemMat M = A.mat[A.elem_base + ee];
    const R Winv = R(c_gll.iw[N1][i] * c_gll.iw[N1][j] * c_gll.iw[N1][k]);
    const R sE = R(M.ieps) * iJ, sH = R(M.imu) * iJ;
    if (live) {
        const double dt = *A.dt;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const R rhs = c < 3 ? sE * (cH[c] + LE[c] * Winv) : sH * (-cE[c - 3] + LH[c - 3] * Winv);
            const KR kn = FIRST ? KR(dt) * KR(rhs) : KR(A.a) * kold[c] + KR(dt) * KR(rhs);
            kb[c * KS + node] = kn;
            uout[c * S + node] = R(KR(u[c]) + KR(A.b) * kn);
        }
    }
}

// Elements per workgroup, two options per n1 (index 0 is the default), capped
// so one workgroup's LDS stays within 64 KiB (two or more workgroups per CU).
constexpr int kEpw[kMaxN1 + 1][2] = {
    {1, 1}, {1, 1}, {8, 32}, {9, 4}, {4, 1}, {2, 1}, {1, 3}, {1, 2}, {1, 2}, {1, 1}, {1, 1}};

constexpr int lds_per_elem(int n1, int rsz) { return (6 * n1 * n1 * n1 + 36 * n1 * n1) * rsz; }
constexpr int epw_of(int n1, int opt, int rsz) {
    const int want = kEpw[n1][opt];
    const int cap_lds = 65536 / lds_per_elem(n1, rsz);
    const int cap_thr = 1024 / (n1 * n1 * n1);
    const int cap = cap_lds < cap_thr ? cap_lds : cap_thr;
    return want < cap ? want : (cap > 0 ? cap : 1);
}

template <int N1, int OPT, typename R, typename KR>
hipError_t launch_one(const FusedArgs& a, hipStream_t s) {
    constexpr int EPW = epw_of(N1, OPT, sizeof(R));
    const int nblocks = (a.nelt + EPW - 1) / EPW;
    FusedArgs b = a;
    b.nchunk = (nblocks + 7) / 8;
    if (a.first_stage)
        hipLaunchKernelGGL((k_fused_stage<N1, EPW, true, R, KR>), dim3(8 * b.nchunk), dim3(EPW * N1 * N1 * N1), 0, s, b);
    else
        hipLaunchKernelGGL((k_fused_stage<N1, EPW, false, R, KR>), dim3(8 * b.nchunk), dim3(EPW * N1 * N1 * N1), 0, s, b);
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

int make_face_nbr(const int* fcode, int nelt, int n1, long long node_base, long long ghost_base, int elem_base,
                  int4* out) {
    const int n2 = n1 * n1, np = n2 * n1;
    auto face_node = [&](int f, int p, int q) {
        int i, j, k;
        face_ijk(f, n1, p, q, i, j, k);
        return i + n1 * (j + n1 * k);
    };
    for (long long x = 0; x < (long long)nelt * 6; ++x) {
        const int code = fcode[x];
        const int e = static_cast<int>(x / 6);
        int4 v;
        if (code == -1) {
            v = make_int4(-1, 0, 0, elem_base + e);
        } else if (code >= 0) {
            const int e2 = code >> 6, f2 = (code >> 3) & 7, o = code & 7;
            int p2, q2;
            orient_map(o, n1, 0, 0, p2, q2);
            const int n00 = face_node(f2, p2, q2);
            orient_map(o, n1, 1, 0, p2, q2);
            const int n10 = face_node(f2, p2, q2);
            orient_map(o, n1, 0, 1, p2, q2);
            const int n01 = face_node(f2, p2, q2);
            const long long base = node_base + (long long)e2 * np + n00;
            if (base > 0x7fffffffLL) return NW_ERR_UNSUPPORTED;
            v = make_int4(static_cast<int>(base), n10 - n00, n01 - n00, elem_base + e2);
        } else {
            const int g = -(code + 2);
            const long long base = ghost_base + (long long)g * n2;
            if (base > 0x7fffffffLL) return NW_ERR_UNSUPPORTED;
            v = make_int4(static_cast<int>(base), 1, n1, -(g + 1));
        }
        out[x] = v;
    }
    return NW_OK;
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
