// nw_plain.hip -- the reference's stage decomposition, one gfx950 kernel per
// reference kernel (FP64), plus device-side geometry setup, diagnostics,
// precision conversion and the halo stand-in.
//
// The stage kernels take the same arrays as the reference launchers
// (dg_kernels.hpp:13-88) so every reference kernel has a counterpart that can
// be checked in isolation.  Design differences from the reference:
//   * volume curl: one thread per (i, j) column with the t-contraction in
//     registers and EPW elements per workgroup (not one thread per node);
//   * lift: one thread per node gathering its face contributions in face order
//     0..5 (deterministic; the reference uses FP64 atomics, dg_kernels.cu:329);
//   * elementwise kernels are grid-stride loops sized for 256 CUs.
#include <mutex>

#include "nw_plain.hpp"

#include "nw_fused.hpp"

namespace nw {
namespace {

__constant__ GllTables c_gll;

constexpr int kBlock = 256;

inline int grid_for(long long n, int block = kBlock) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return static_cast<int>(g);
}

// ---------------------------------------------------------------------------
// volume curl (dg_kernels.cu:41-155): w = sign * W*J * (grad_ref u . J^-T) x
// ---------------------------------------------------------------------------
template <int N1, int EPW>
__global__ __launch_bounds__(EPW* N1* N1) void k_volume_curl(
    const double* __restrict__ u1, const double* __restrict__ u2, const double* __restrict__ u3,
    double* __restrict__ w1, double* __restrict__ w2, double* __restrict__ w3o,
    const double* __restrict__ D, const double* __restrict__ w3g, const double* __restrict__ jac,
    const double* __restrict__ rx, const double* __restrict__ sx, const double* __restrict__ tx,
    const double* __restrict__ ry, const double* __restrict__ sy, const double* __restrict__ ty,
    const double* __restrict__ rz, const double* __restrict__ sz, const double* __restrict__ tz,
    double sign, int nelt) {
    constexpr int N2 = N1 * N1, NP = N1 * N2;
    __shared__ double sU[EPW][3][NP];
    __shared__ double sD[N2];
    const int tid = threadIdx.x;
    const int es = tid / N2, l = tid - es * N2, i = l % N1, j = l / N1;
    const int e = blockIdx.x * EPW + es;
    const bool live = e < nelt;
    const int ee = live ? e : nelt - 1;
    for (int x = tid; x < N2; x += EPW * N2) sD[x] = D[x];  // column-major D[i + l*N1]
    const long long nb = (long long)ee * NP + l;
    double c1[N1], c2[N1], c3[N1];
#pragma unroll
    for (int k = 0; k < N1; ++k) {
        c1[k] = u1[nb + N2 * k]; c2[k] = u2[nb + N2 * k]; c3[k] = u3[nb + N2 * k];
        sU[es][0][l + N2 * k] = c1[k]; sU[es][1][l + N2 * k] = c2[k]; sU[es][2][l + N2 * k] = c3[k];
    }
    __syncthreads();
    if (!live) return;
#pragma unroll
    for (int k = 0; k < N1; ++k) {
        double a1r = 0, a1s = 0, a1t = 0, a2r = 0, a2s = 0, a2t = 0, a3r = 0, a3s = 0, a3t = 0;
#pragma unroll
        for (int q = 0; q < N1; ++q) {
            const double di = sD[i + q * N1], dj = sD[j + q * N1], dk = sD[k + q * N1];
            const int ir = q + N1 * j + N2 * k, is = i + N1 * q + N2 * k;
            a1r += di * sU[es][0][ir]; a2r += di * sU[es][1][ir]; a3r += di * sU[es][2][ir];
            a1s += dj * sU[es][0][is]; a2s += dj * sU[es][1][is]; a3s += dj * sU[es][2][is];
            a1t += dk * c1[q]; a2t += dk * c2[q]; a3t += dk * c3[q];
        }
        const long long p = nb + N2 * k;
        const double w = w3g[l + N2 * k] * jac[p];
        const double u1rw = a1r * w, u1sw = a1s * w, u1tw = a1t * w;
        const double u2rw = a2r * w, u2sw = a2s * w, u2tw = a2t * w;
        const double u3rw = a3r * w, u3sw = a3s * w, u3tw = a3t * w;
        const double Rx = rx[p], Sx = sx[p], Tx = tx[p], Ry = ry[p], Sy = sy[p], Ty = ty[p];
        const double Rz = rz[p], Sz = sz[p], Tz = tz[p];
        const double o1 = (u3rw * Ry + u3sw * Sy + u3tw * Ty) - (u2rw * Rz + u2sw * Sz + u2tw * Tz);
        const double o2 = (u1rw * Rz + u1sw * Sz + u1tw * Tz) - (u3rw * Rx + u3sw * Sx + u3tw * Tx);
        const double o3 = (u2rw * Rx + u2sw * Sx + u2tw * Tx) - (u1rw * Ry + u1sw * Sy + u1tw * Ty);
        w1[p] = sign * o1; w2[p] = sign * o2; w3o[p] = sign * o3;
    }
}

constexpr int kVolEpw[kMaxN1 + 1] = {1, 1, 16, 7, 4, 2, 1, 1, 1, 1, 1};

template <int N1>
hipError_t launch_vc(const PlainVolumeArgs& a, hipStream_t s) {
    constexpr int EPW = kVolEpw[N1];
    const int grid = (a.nelt + EPW - 1) / EPW;
    hipLaunchKernelGGL((k_volume_curl<N1, EPW>), dim3(grid), dim3(EPW * N1 * N1), 0, s,
                       a.u1, a.u2, a.u3, a.w1, a.w2, a.w3, a.D, a.w3g, a.jac, a.rx, a.sx, a.tx,
                       a.ry, a.sy, a.ty, a.rz, a.sz, a.tz, a.sign, a.nelt);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// restrict / flux / lift / inverse mass / LSRK / halo
// ---------------------------------------------------------------------------
__global__ void k_restrict(const double* __restrict__ st, double* __restrict__ fEN, double* __restrict__ fHN,
                           const int* __restrict__ vM, long long tfp, long long S) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < tfp; p += (long long)gridDim.x * blockDim.x) {
        const long long v = vM[p];
        fEN[p] = st[v]; fEN[tfp + p] = st[S + v]; fEN[2 * tfp + p] = st[2 * S + v];
        fHN[p] = st[3 * S + v]; fHN[tfp + p] = st[4 * S + v]; fHN[2 * tfp + p] = st[5 * S + v];
    }
}

__global__ void k_flux(const double* __restrict__ st, const double* __restrict__ fEN, const double* __restrict__ fHN,
                       const int* __restrict__ vP, const int* __restrict__ bc, const double* __restrict__ nxa,
                       const double* __restrict__ nya, const double* __restrict__ nza, double* __restrict__ flux,
                       double c0, long long tfp, long long S, const int* __restrict__ vM, const double* __restrict__ eps,
                       const double* __restrict__ mu, const int* __restrict__ ghost_elem, int np, long long npts) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < tfp; p += (long long)gridDim.x * blockDim.x) {
        const double Exm = fEN[p], Eym = fEN[tfp + p], Ezm = fEN[2 * tfp + p];
        const double Hxm = fHN[p], Hym = fHN[tfp + p], Hzm = fHN[2 * tfp + p];
        const int v = vP[p];
        const bool mirror = bc[p] != 0 || v < 0;
        double dEx, dEy, dEz, dHx, dHy, dHz;
        if (mirror) {
            dEx = -2.0 * Exm; dEy = -2.0 * Eym; dEz = -2.0 * Ezm; dHx = 0.0; dHy = 0.0; dHz = 0.0;
        } else {
            dEx = st[v] - Exm; dEy = st[S + v] - Eym; dEz = st[2 * S + v] - Ezm;
            dHx = st[3 * S + v] - Hxm; dHy = st[4 * S + v] - Hym; dHz = st[5 * S + v] - Hzm;
        }
        const double n_x = nxa[p], n_y = nya[p], n_z = nza[p];
        const double nxEx = n_y * dEz - n_z * dEy, nxEy = n_z * dEx - n_x * dEz, nxEz = n_x * dEy - n_y * dEx;
        const double nxHx = n_y * dHz - n_z * dHy, nxHy = n_z * dHx - n_x * dHz, nxHz = n_x * dHy - n_y * dHx;
        const double nnEx = n_y * nxEz - n_z * nxEy, nnEy = n_z * nxEx - n_x * nxEz, nnEz = n_x * nxEy - n_y * nxEx;
        const double nnHx = n_y * nxHz - n_z * nxHy, nnHy = n_z * nxHx - n_x * nxHz, nnHz = n_x * nxHy - n_y * nxHx;
        double fHx, fHy, fHz, fEx, fEy, fEz;
        if (!eps) {
            fHx = -0.5 * nxEx - 0.5 * c0 * nnHx; fHy = -0.5 * nxEy - 0.5 * c0 * nnHy; fHz = -0.5 * nxEz - 0.5 * c0 * nnHz;
            fEx = 0.5 * nxHx - 0.5 * c0 * nnEx; fEy = 0.5 * nxHy - 0.5 * c0 * nnEy; fEz = 0.5 * nxHz - 0.5 * c0 * nnEz;
        } else {
            const int em = vM[p] / np;
            const int ep = mirror ? em : (v < npts ? v / np : ghost_elem[v - npts]);
            const double Zm = sqrt(mu[em] / eps[em]), Zp = sqrt(mu[ep] / eps[ep]);
            const double Ym = 1.0 / Zm, Yp = 1.0 / Zp;
            const double iz = 1.0 / (Zp + Zm), iy = 1.0 / (Yp + Ym);
            fHx = (-Yp * nxEx - c0 * nnHx) * iy; fHy = (-Yp * nxEy - c0 * nnHy) * iy; fHz = (-Yp * nxEz - c0 * nnHz) * iy;
            fEx = (Zp * nxHx - c0 * nnEx) * iz; fEy = (Zp * nxHy - c0 * nnEy) * iz; fEz = (Zp * nxHz - c0 * nnEz) * iz;
        }
        flux[p] = fHx; flux[tfp + p] = fHy; flux[2 * tfp + p] = fHz;
        flux[3 * tfp + p] = fEx; flux[4 * tfp + p] = fEy; flux[5 * tfp + p] = fEz;
    }
}

// Deterministic lift: node (i,j,k) of element e gathers dA*flux from the face
// points that map to it, faces in order 0..5 (the serial face-point order).
// vM == nullptr: the element-major face layout is known to hold (the solver's
// own tables).  Otherwise a face point counts here only if volIdxMinus names the
// node the layout puts it at; k_lift_scatter adds the others.
__global__ void k_lift(const double* __restrict__ flux, const double* __restrict__ dA, double* __restrict__ rhs,
                       long long tfp, long long npts, int n1, const int* __restrict__ vM) {
    const int n2 = n1 * n1, np = n2 * n1;
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < npts; v += (long long)gridDim.x * blockDim.x) {
        const long long e = v / np;
        const int m = static_cast<int>(v - e * np);
        const int i = m % n1, j = (m / n1) % n1, k = m / n2;
        int pts[6], nf = 0;
        const long long fb = e * 6 * n2;
        if (j == 0) pts[nf++] = 0 * n2 + i + n1 * k;
        if (i == n1 - 1) pts[nf++] = 1 * n2 + j + n1 * k;
        if (j == n1 - 1) pts[nf++] = 2 * n2 + i + n1 * k;
        if (i == 0) pts[nf++] = 3 * n2 + j + n1 * k;
        if (k == 0) pts[nf++] = 4 * n2 + i + n1 * j;
        if (k == n1 - 1) pts[nf++] = 5 * n2 + i + n1 * j;
        if (nf == 0) continue;
        double r0 = rhs[v], r1 = rhs[npts + v], r2 = rhs[2 * npts + v];
        double r3 = rhs[3 * npts + v], r4 = rhs[4 * npts + v], r5 = rhs[5 * npts + v];
        for (int t = 0; t < nf; ++t) {
            const long long p = fb + pts[t];
            if (vM && vM[p] != v) continue;
            const double a = dA[p];
            r0 += a * flux[3 * tfp + p]; r1 += a * flux[4 * tfp + p]; r2 += a * flux[5 * tfp + p];
            r3 += a * flux[p]; r4 += a * flux[tfp + p]; r5 += a * flux[2 * tfp + p];
        }
        rhs[v] = r0; rhs[npts + v] = r1; rhs[2 * npts + v] = r2;
        rhs[3 * npts + v] = r3; rhs[4 * npts + v] = r4; rhs[5 * npts + v] = r5;
    }
}

// The face points k_lift did not take (n1 > 0: those whose volIdxMinus is not
// the node the element-major layout puts them at; n1 == 0: all of them), added
// as the reference adds every point: rhs[vM] += dA * flux with FP64 atomics
// (dg_kernels.cu:316-338).  Points with vM outside [0, npts) are skipped.
__global__ void k_lift_scatter(const double* __restrict__ flux, const double* __restrict__ dA, double* __restrict__ rhs,
                               long long tfp, long long npts, int n1, const int* __restrict__ vM) {
    const int n2 = n1 * n1, nf2 = 6 * n2;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < tfp; p += (long long)gridDim.x * blockDim.x) {
        const long long v = vM[p];
        if (n1 > 0) {
            const long long e = p / nf2;
            const int r = static_cast<int>(p - e * nf2), f = r / n2, pq = r - f * n2;
            int i, j, k;
            face_ijk(f, n1, pq % n1, pq / n1, i, j, k);
            if (v == e * n2 * n1 + i + n1 * (j + n1 * k)) continue;
        }
        if (v < 0 || v >= npts) continue;
        const double a = dA[p];
        atomicAdd(&rhs[v], a * flux[3 * tfp + p]);
        atomicAdd(&rhs[npts + v], a * flux[4 * tfp + p]);
        atomicAdd(&rhs[2 * npts + v], a * flux[5 * tfp + p]);
        atomicAdd(&rhs[3 * npts + v], a * flux[p]);
        atomicAdd(&rhs[4 * npts + v], a * flux[tfp + p]);
        atomicAdd(&rhs[5 * npts + v], a * flux[2 * tfp + p]);
    }
}

__global__ void k_inv_mass(double* __restrict__ rhs, const double* __restrict__ jac, const double* __restrict__ w3,
                           int nxyz, long long npts, const double* __restrict__ eps, const double* __restrict__ mu) {
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < npts; idx += (long long)gridDim.x * blockDim.x) {
        const long long e = idx / nxyz;
        const int i = static_cast<int>(idx - e * nxyz);
        const double ep = eps ? eps[e] : 1.0, m = mu ? mu[e] : 1.0;
        const double iE = 1.0 / (ep * jac[idx] * w3[i]);
        const double iH = 1.0 / (m * jac[idx] * w3[i]);
        rhs[idx] *= iE; rhs[npts + idx] *= iE; rhs[2 * npts + idx] *= iE;
        rhs[3 * npts + idx] *= iH; rhs[4 * npts + idx] *= iH; rhs[5 * npts + idx] *= iH;
    }
}

__global__ void k_lsrk(double* __restrict__ st, double* __restrict__ k, const double* __restrict__ rhs, double a,
                       double b, const double* __restrict__ dtp, double dtv, long long npts, long long S) {
    const double dt = dtp ? *dtp : dtv;
    const long long n = 6 * npts;
    for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < n; x += (long long)gridDim.x * blockDim.x) {
        const long long c = x / npts, p = x - c * npts;
        const double ki = a * k[x] + dt * rhs[x];
        k[x] = ki;
        st[c * S + p] += b * ki;
    }
}

template <typename R>
__global__ void k_pack(const R* __restrict__ st, R* __restrict__ buf, const int* __restrict__ vol,
                       const int* __restrict__ off, const int* __restrict__ sizes, int nh, long long cs, long long ns) {
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < nh; x += gridDim.x * blockDim.x) {
        const long long v = vol[x];
        const int base = off[x], M = sizes[x];
#pragma unroll
        for (int c = 0; c < 6; ++c) buf[base + c * M] = st[c * cs + v * ns];
    }
}

template <typename R>
__global__ void k_unpack(const R* __restrict__ buf, R* __restrict__ st, const int* __restrict__ off,
                         const int* __restrict__ sizes, int nh, long long npts, long long cs, long long ns) {
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < nh; x += gridDim.x * blockDim.x) {
        const int base = off[x], M = sizes[x];
#pragma unroll
        for (int c = 0; c < 6; ++c) st[c * cs + (npts + x) * ns] = buf[base + c * M];
    }
}

// ---------------------------------------------------------------------------
// device geometry setup (computeMetricsFromCorners mesh.cpp:188-314 and
// setupFaceData mesh.cpp:985-1080, in the reference's expression order)
// ---------------------------------------------------------------------------
__global__ void k_metrics(const double* __restrict__ corners, int n1, long long npts, double* __restrict__ X,
                          double* __restrict__ Y, double* __restrict__ Z, double* __restrict__ jac, double* __restrict__ rx,
                          double* __restrict__ sx, double* __restrict__ tx, double* __restrict__ ry,
                          double* __restrict__ sy, double* __restrict__ ty, double* __restrict__ rz,
                          double* __restrict__ sz, double* __restrict__ tz) {
    const int n2 = n1 * n1, np = n2 * n1;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < npts; p += (long long)gridDim.x * blockDim.x) {
        const long long e = p / np;
        const int m = static_cast<int>(p - e * np);
        const double r = c_gll.z[n1][m % n1], s = c_gll.z[n1][(m / n1) % n1], t = c_gll.z[n1][m / n2];
        const double* c = corners + e * 24;
        const double sr[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, ss[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
        const double st[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
        double x = 0, y = 0, z = 0, xr = 0, yr = 0, zr = 0, xs = 0, ys = 0, zs = 0, xt = 0, yt = 0, zt = 0;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const double fr = 1.0 + sr[a] * r, fs = 1.0 + ss[a] * s, ft = 1.0 + st[a] * t;
            const double Nf = 0.125 * fr * fs * ft;
            const double dr = sr[a] * 0.125 * fs * ft, ds = ss[a] * 0.125 * fr * ft, dt = st[a] * 0.125 * fr * fs;
            const double vx = c[a * 3], vy = c[a * 3 + 1], vz = c[a * 3 + 2];
            x += Nf * vx; y += Nf * vy; z += Nf * vz;
            xr += dr * vx; yr += dr * vy; zr += dr * vz;
            xs += ds * vx; ys += ds * vy; zs += ds * vz;
            xt += dt * vx; yt += dt * vy; zt += dt * vz;
        }
        X[p] = x; Y[p] = y; Z[p] = z;
        const double det = xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
        jac[p] = det;
        const double iJ = 1.0 / det;
        rx[p] = iJ * (ys * zt - yt * zs); sx[p] = iJ * (yt * zr - yr * zt); tx[p] = iJ * (yr * zs - ys * zr);
        ry[p] = iJ * (xt * zs - xs * zt); sy[p] = iJ * (xr * zt - xt * zr); ty[p] = iJ * (xs * zr - xr * zs);
        rz[p] = iJ * (xs * yt - xt * ys); sz[p] = iJ * (xt * yr - xr * yt); tz[p] = iJ * (xr * ys - xs * yr);
    }
}

__global__ void k_face_data(const int* __restrict__ fcode, const int* __restrict__ fbc, int n1, long long tfp,
                            long long npts, const double* __restrict__ jac, const double* __restrict__ rx,
                            const double* __restrict__ sx, const double* __restrict__ tx, const double* __restrict__ ry,
                            const double* __restrict__ sy, const double* __restrict__ ty, const double* __restrict__ rz,
                            const double* __restrict__ sz, const double* __restrict__ tz, int* __restrict__ vM,
                            int* __restrict__ vP, int* __restrict__ bc, double* __restrict__ nx, double* __restrict__ ny,
                            double* __restrict__ nz, double* __restrict__ dA) {
    const int n2 = n1 * n1, np = n2 * n1;
    for (long long P = blockIdx.x * (long long)blockDim.x + threadIdx.x; P < tfp; P += (long long)gridDim.x * blockDim.x) {
        const long long fi = P / n2;
        const int pq = static_cast<int>(P - fi * n2), p = pq % n1, q = pq / n1;
        const int e = static_cast<int>(fi / 6), f = static_cast<int>(fi % 6);
        int i, j, k;
        face_ijk(f, n1, p, q, i, j, k);
        const long long v = (long long)e * np + i + n1 * (j + n1 * k);
        vM[P] = static_cast<int>(v);
        const double J = jac[v], w2 = c_gll.w[n1][p] * c_gll.w[n1][q];
        double Nx, Ny, Nz;
        switch (f) {
            case 0: Nx = -J * sx[v]; Ny = -J * sy[v]; Nz = -J * sz[v]; break;
            case 1: Nx = J * rx[v]; Ny = J * ry[v]; Nz = J * rz[v]; break;
            case 2: Nx = J * sx[v]; Ny = J * sy[v]; Nz = J * sz[v]; break;
            case 3: Nx = -J * rx[v]; Ny = -J * ry[v]; Nz = -J * rz[v]; break;
            case 4: Nx = -J * tx[v]; Ny = -J * ty[v]; Nz = -J * tz[v]; break;
            default: Nx = J * tx[v]; Ny = J * ty[v]; Nz = J * tz[v]; break;
        }
        const double L = sqrt(Nx * Nx + Ny * Ny + Nz * Nz);
        nx[P] = L > 1e-14 ? Nx / L : 0.0; ny[P] = L > 1e-14 ? Ny / L : 0.0; nz[P] = L > 1e-14 ? Nz / L : 0.0;
        dA[P] = L * w2;
        const int code = fcode[fi];
        int plus;
        if (code >= 0) {
            const int e2 = code >> 6, f2 = (code >> 3) & 7, o = code & 7;
            int p2, q2, i2, j2, k2;
            orient_map(o, n1, p, q, p2, q2);
            face_ijk(f2, n1, p2, q2, i2, j2, k2);
            plus = e2 * np + i2 + n1 * (j2 + n1 * k2);
        } else if (code == -1) {
            plus = -1;
        } else {
            plus = static_cast<int>(npts) + (-(code + 2)) * n2 + pq;
        }
        vP[P] = plus;
        int b = fbc[fi];
        if (plus < 0 && b == 0) b = 1;
        bc[P] = b;
    }
}

// ---------------------------------------------------------------------------
// diagnostics: energy (FP64), max |E|, max |H|; two-pass deterministic reduce
// ---------------------------------------------------------------------------
// Quadrature weight of node p: J w_i w_j w_k, J from the metric array when
// given, else from the element's trilinear map (mesh.cpp:291-295).
__device__ __forceinline__ double node_dv(long long p, int n1, const double* __restrict__ jac,
                                          const ElemGeom* __restrict__ geom) {
    const int n2 = n1 * n1, np = n2 * n1;
    const long long e = p / np;
    const int m = static_cast<int>(p - e * np);
    const int i = m % n1, j = (m / n1) % n1, k = m / n2;
    double J;
    if (jac) {
        J = jac[p];
    } else {
        const double r = c_gll.z[n1][i], s = c_gll.z[n1][j], t = c_gll.z[n1][k];
        const double* c0 = geom[e].c[0];
        const double* c1 = geom[e].c[1];
        const double* c2 = geom[e].c[2];
        const double xr = c0[1] + c0[4] * s + c0[5] * t + c0[7] * (s * t);
        const double xs = c0[2] + c0[4] * r + c0[6] * t + c0[7] * (r * t);
        const double xt = c0[3] + c0[5] * r + c0[6] * s + c0[7] * (r * s);
        const double yr = c1[1] + c1[4] * s + c1[5] * t + c1[7] * (s * t);
        const double ys = c1[2] + c1[4] * r + c1[6] * t + c1[7] * (r * t);
        const double yt = c1[3] + c1[5] * r + c1[6] * s + c1[7] * (r * s);
        const double zr = c2[1] + c2[4] * s + c2[5] * t + c2[7] * (s * t);
        const double zs = c2[2] + c2[4] * r + c2[6] * t + c2[7] * (r * t);
        const double zt = c2[3] + c2[5] * r + c2[6] * s + c2[7] * (r * s);
        J = xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
    }
    return J * c_gll.w[n1][i] * c_gll.w[n1][j] * c_gll.w[n1][k];
}

template <typename R>
__global__ void k_diag_partial(const R* __restrict__ u, long long cs, long long ns, long long npts, int n1,
                               const double* __restrict__ jac,
                               const ElemGeom* __restrict__ geom, const ElemMat* __restrict__ mat,
                               double* __restrict__ part) {
    __shared__ double sh[3][kBlock];
    const int np = n1 * n1 * n1;
    double en = 0.0, me = 0.0, mh = 0.0;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < npts; p += (long long)gridDim.x * blockDim.x) {
        const long long e = p / np;
        const double dV = node_dv(p, n1, jac, geom);
        const R* up = u + p * ns;
        const double ex = up[0], ey = up[cs], ez = up[2 * cs];
        const double hx = up[3 * cs], hy = up[4 * cs], hz = up[5 * cs];
        const double e2 = ex * ex + ey * ey + ez * ez, h2 = hx * hx + hy * hy + hz * hz;
        const double ep = mat ? 1.0 / mat[e].ieps : 1.0, mu = mat ? 1.0 / mat[e].imu : 1.0;
        en += 0.5 * (ep * e2 + mu * h2) * dV;
        me = fmax(me, e2);
        mh = fmax(mh, h2);
    }
    sh[0][threadIdx.x] = en; sh[1][threadIdx.x] = me; sh[2][threadIdx.x] = mh;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + w]);
            sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = sh[0][0]; part[3 * blockIdx.x + 1] = sh[1][0]; part[3 * blockIdx.x + 2] = sh[2][0];
    }
}

__global__ void k_diag_final(const double* __restrict__ part, int n, double* __restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double en = 0.0, me = 0.0, mh = 0.0;
        for (int b = 0; b < n; ++b) {
            en += part[3 * b]; me = fmax(me, part[3 * b + 1]); mh = fmax(mh, part[3 * b + 2]);
        }
        out[0] = en; out[1] = sqrt(me); out[2] = sqrt(mh);
    }
}

// Case::computeTotalEnergy, getMaxE, getMaxH (case.cpp:733-794) on a device
// state in the reference layout (component c of node p at c S + p): per block
// the sum of 1/2 (eps |E|^2 + mu |H|^2) J w_i w_j w_k (the reference's
// J[p] * w[i] * w[j] * w[k]) and the largest |E|^2, |H|^2; eps/mu per element
// or NULL (the reference's 1/2 (|E|^2 + |H|^2) dV).
__global__ void k_energy_partial(const double* __restrict__ st, long long S, long long npts, int n1,
                                 const double* __restrict__ jac, const double* __restrict__ w,
                                 const double* __restrict__ eps, const double* __restrict__ mu,
                                 double* __restrict__ part) {
    __shared__ double sh[3][kBlock];
    const int n2 = n1 * n1, np = n2 * n1;
    double en = 0.0, me = 0.0, mh = 0.0;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < npts; p += (long long)gridDim.x * blockDim.x) {
        const long long e = p / np;
        const int m = static_cast<int>(p - e * np);
        const int i = m % n1, j = (m / n1) % n1, k = m / n2;
        const double dV = jac[p] * w[i] * w[j] * w[k];
        const double ex = st[p], ey = st[S + p], ez = st[2 * S + p];
        const double hx = st[3 * S + p], hy = st[4 * S + p], hz = st[5 * S + p];
        const double e2 = ex * ex + ey * ey + ez * ez, h2 = hx * hx + hy * hy + hz * hz;
        if (eps || mu)
            en += 0.5 * ((eps ? eps[e] : 1.0) * e2 + (mu ? mu[e] : 1.0) * h2) * dV;
        else
            en += 0.5 * (e2 + h2) * dV;
        me = fmax(me, e2);
        mh = fmax(mh, h2);
    }
    sh[0][threadIdx.x] = en; sh[1][threadIdx.x] = me; sh[2][threadIdx.x] = mh;
    __syncthreads();
    for (int r = blockDim.x / 2; r > 0; r >>= 1) {
        if (threadIdx.x < r) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + r];
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + r]);
            sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + r]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = sh[0][0]; part[3 * blockIdx.x + 1] = sh[1][0]; part[3 * blockIdx.x + 2] = sh[2][0];
    }
}

// Relative L2 distance to a reference state (3dboxpec_test.cpp:76-119):
// sum over components and nodes of (u - ref)^2 dV and ref^2 dV, in FP64.
// ref is in the reference's host layout (component blocks of npts).
template <typename R>
__global__ void k_l2_partial(const R* __restrict__ u, long long cs, long long ns, const double* __restrict__ ref,
                             long long npts, int n1, const double* __restrict__ jac,
                             const ElemGeom* __restrict__ geom, double* __restrict__ part) {
    __shared__ double sh[2][kBlock];
    double num = 0.0, den = 0.0;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < npts; p += (long long)gridDim.x * blockDim.x) {
        const double dV = node_dv(p, n1, jac, geom);
        const R* up = u + p * ns;
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double r = ref[c * npts + p], d = double(up[c * cs]) - r;
            a += d * d;
            b += r * r;
        }
        num += a * dV;
        den += b * dV;
    }
    sh[0][threadIdx.x] = num; sh[1][threadIdx.x] = den;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sh[0][0]; part[2 * blockIdx.x + 1] = sh[1][0];
    }
}

__global__ void k_l2_final(const double* __restrict__ part, int n, double* __restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double num = 0.0, den = 0.0;
        for (int b = 0; b < n; ++b) {
            num += part[2 * b]; den += part[2 * b + 1];
        }
        out[0] = num; out[1] = den;
    }
}

// Probe record (ProbeManager::record, probe.cpp:192-216) on the device, one
// workgroup per probe: the threads form the rounded products w_p * u_c(p),
// w_p = (Lr[i] Ls[j]) Lt[k], of all the element's nodes at once into LDS; then
// one thread per component adds them in node order -- the reference's
// operations in its order, in FP64, so the record is bit for bit the host's --
// into record slot (*slot mod cap).  k_probe_advance moves the slot afterwards
// (stream order), so every workgroup reads the same slot.
template <typename R>
__global__ void __launch_bounds__(256) k_probe(const R* __restrict__ u, long long cs, long long ns, int nprobes,
                                               const long long* __restrict__ base, const double* __restrict__ lw,
                                               int n1, double* __restrict__ hist, const int* __restrict__ slot,
                                               int cap) {
    __shared__ double prod[6][kMaxN1 * kMaxN1 * kMaxN1];
    const int q = blockIdx.x;
    const int np = n1 * n1 * n1;
    const double* Lr = lw + size_t(q) * 3 * n1;
    const double* Ls = Lr + n1;
    const double* Lt = Ls + n1;
    const R* up = u + base[q] * ns;
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
        const int i = p % n1, j = (p / n1) % n1, kk = p / (n1 * n1);
        const double w = __dmul_rn(__dmul_rn(Lr[i], Ls[j]), Lt[kk]);
#pragma unroll
        for (int c = 0; c < 6; ++c) prod[c][p] = __dmul_rn(w, double(up[p * ns + c * cs]));
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int c = threadIdx.x;
        double acc = 0.0;
        for (int p = 0; p < np; ++p) acc = __dadd_rn(acc, prod[c][p]);  // no contraction: the host sum's roundings
        hist[(size_t(*slot % cap) * nprobes + q) * 6 + c] = acc;
    }
}

// Counts records up to cap + 1 (the overflow the host reports), so a run that
// never fetches its history cannot wrap the counter.
__global__ void k_probe_advance(int* p, int cap) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && *p <= cap) *p = *p + 1;
}

// Sets *slot = 0 with a kernel (no memset node if a caller captures it).
__global__ void k_zero_int(int* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}
// ---------------------------------------------------------------------------
// conversions and plain copy (HBM calibration)
// ---------------------------------------------------------------------------
template <typename D, typename S>
__global__ void k_convert(D* __restrict__ dst, long long dcs, long long dns, const S* __restrict__ src, long long scs,
                          long long sns, long long n, int ncomp) {
    const long long tot = n * ncomp;
    for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < tot; x += (long long)gridDim.x * blockDim.x) {
        const long long c = x / n, p = x - c * n;
        dst[c * dcs + p * dns] = static_cast<D>(src[c * scs + p * sns]);
    }
}

__global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, long long n4) {
    for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < n4; x += (long long)gridDim.x * blockDim.x)
        dst[x] = src[x];
}

__global__ void k_fill(double* __restrict__ p, long long n, double v) {
    for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < n; x += (long long)gridDim.x * blockDim.x)
        p[x] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
int plain_upload_tables(const GllTables& t) {
    NW_TRY_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gll), &t, sizeof(GllTables)));
    return NW_OK;
}

hipError_t launch_volume_curl(const PlainVolumeArgs& a, hipStream_t s) {
    if (a.nelt <= 0) return hipSuccess;
    switch (a.n1) {
        case 2: return launch_vc<2>(a, s);
        case 3: return launch_vc<3>(a, s);
        case 4: return launch_vc<4>(a, s);
        case 5: return launch_vc<5>(a, s);
        case 6: return launch_vc<6>(a, s);
        case 7: return launch_vc<7>(a, s);
        case 8: return launch_vc<8>(a, s);
        case 9: return launch_vc<9>(a, s);
        case 10: return launch_vc<10>(a, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_restrict(const double* st, double* fEN, double* fHN, const int* vM, long long tfp, long long S,
                           hipStream_t s) {
    if (tfp <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_restrict, dim3(grid_for(tfp)), dim3(kBlock), 0, s, st, fEN, fHN, vM, tfp, S);
    return hipGetLastError();
}

hipError_t launch_flux(const double* st, const double* fEN, const double* fHN, const int* vP, const int* bc,
                       const double* nx, const double* ny, const double* nz, double* flux, double c0, long long tfp,
                       long long S, const int* vM, const double* eps, const double* mu, const int* ghost_elem, int np,
                       long long npts, hipStream_t s) {
    if (tfp <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_flux, dim3(grid_for(tfp)), dim3(kBlock), 0, s, st, fEN, fHN, vP, bc, nx, ny, nz, flux, c0,
                       tfp, S, vM, eps, mu, ghost_elem, np, npts);
    return hipGetLastError();
}

hipError_t launch_lift(const double* flux, const double* dA, double* rhs, long long tfp, long long npts, int n1,
                       hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lift, dim3(grid_for(npts)), dim3(kBlock), 0, s, flux, dA, rhs, tfp, npts, n1,
                       static_cast<const int*>(nullptr));
    return hipGetLastError();
}

hipError_t launch_add_flux(const double* flux, const int* vM, const double* dA, double* rhs, long long tfp,
                           long long npts, hipStream_t s) {
    if (tfp <= 0) return hipSuccess;
    // n1 of the element-major layout, when the sizes admit one: tfp = 6 n1^2 nelt,
    // npts = n1^3 nelt
    int n1 = 0;
    if (npts > 0 && (6 * npts) % tfp == 0) {
        const long long c = 6 * npts / tfp;
        if (c >= kMinN1 && c <= kMaxN1 && npts % (c * c * c) == 0 && tfp == 6 * c * c * (npts / (c * c * c)))
            n1 = static_cast<int>(c);
    }
    if (n1 > 0) {
        hipLaunchKernelGGL(k_lift, dim3(grid_for(npts)), dim3(kBlock), 0, s, flux, dA, rhs, tfp, npts, n1, vM);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_lift_scatter, dim3(grid_for(tfp)), dim3(kBlock), 0, s, flux, dA, rhs, tfp, npts, n1, vM);
    return hipGetLastError();
}

hipError_t launch_inv_mass(double* rhs, const double* jac, const double* w3, int nxyz, long long npts,
                           const double* eps, const double* mu, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_inv_mass, dim3(grid_for(npts)), dim3(kBlock), 0, s, rhs, jac, w3, nxyz, npts, eps, mu);
    return hipGetLastError();
}

hipError_t launch_lsrk(double* st, double* k, const double* rhs, double a, double b, const double* dtp, double dt,
                       long long npts, long long S, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lsrk, dim3(grid_for(6 * npts)), dim3(kBlock), 0, s, st, k, rhs, a, b, dtp, dt, npts, S);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_pack(const R* st, R* buf, const int* vol, const int* off, const int* sizes, int nh, long long cs,
                       long long ns, hipStream_t s) {
    if (nh <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack<R>, dim3(grid_for(nh)), dim3(kBlock), 0, s, st, buf, vol, off, sizes, nh, cs, ns);
    return hipGetLastError();
}
template <typename R>
hipError_t launch_unpack(const R* buf, R* st, const int* off, const int* sizes, int nh, long long npts, long long cs,
                         long long ns, hipStream_t s) {
    if (nh <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack<R>, dim3(grid_for(nh)), dim3(kBlock), 0, s, buf, st, off, sizes, nh, npts, cs, ns);
    return hipGetLastError();
}
template hipError_t launch_pack<double>(const double*, double*, const int*, const int*, const int*, int, long long,
                                        long long, hipStream_t);
template hipError_t launch_pack<float>(const float*, float*, const int*, const int*, const int*, int, long long, long long,
                                       hipStream_t);
template hipError_t launch_unpack<double>(const double*, double*, const int*, const int*, int, long long, long long,
                                          long long, hipStream_t);
template hipError_t launch_unpack<float>(const float*, float*, const int*, const int*, int, long long, long long,
                                         long long, hipStream_t);

hipError_t launch_metrics(const double* corners, int n1, long long npts, double* const* out13, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_metrics, dim3(grid_for(npts)), dim3(kBlock), 0, s, corners, n1, npts, out13[0], out13[1],
                       out13[2], out13[3], out13[4], out13[5], out13[6], out13[7], out13[8], out13[9], out13[10],
                       out13[11], out13[12]);
    return hipGetLastError();
}

hipError_t launch_face_data(const int* fcode, const int* fbc, int n1, long long tfp, long long npts,
                            const double* const* met10, int* vM, int* vP, int* bc, double* nx, double* ny, double* nz,
                            double* dA, hipStream_t s) {
    if (tfp <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_face_data, dim3(grid_for(tfp)), dim3(kBlock), 0, s, fcode, fbc, n1, tfp, npts, met10[0],
                       met10[1], met10[2], met10[3], met10[4], met10[5], met10[6], met10[7], met10[8], met10[9], vM, vP,
                       bc, nx, ny, nz, dA);
    return hipGetLastError();
}

int diag_blocks() { return 1024; }

template <typename R>
hipError_t launch_diag(const R* u, long long cs, long long ns, long long npts, int n1, const double* jac,
                       const ElemGeom* geom, const ElemMat* mat, double* part, double* out, hipStream_t s) {
    const int nb = diag_blocks();
    hipLaunchKernelGGL(k_diag_partial<R>, dim3(nb), dim3(kBlock), 0, s, u, cs, ns, npts, n1, jac, geom, mat, part);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s, part, nb, out);
    return hipGetLastError();
}
// Partial sums of nw_launch_diagnostics_noalloc: a fixed device array (one
// copy per GPU, each at its own address, held until the process ends) instead
// of an allocation.  Host code must use the address energy_scratch() returns,
// never this name.  Every call on a GPU shares that GPU's copy: calls that
// overlap on one GPU race.
static __device__ __align__(256) double scratch[3 * kEnergyScratchBlocks];
static_assert(kEnergyScratchBlocks >= kEnergyMaxBlocks, "the scratch holds every block of the largest grid");

int energy_scratch(double** out) {
    static std::mutex mu;
    static double* addr[kMaxDevices] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -2;
        addr[dev] = static_cast<double*>(p);
    }
    *out = addr[dev];
    return 0;
}

int energy_blocks(long long npts) {
    const long long b = (npts + kBlock - 1) / kBlock;
    return static_cast<int>(b < 1 ? 1 : (b > kEnergyMaxBlocks ? kEnergyMaxBlocks : b));
}

hipError_t launch_energy(const double* st, long long S, long long npts, int n1, const double* jac, const double* w,
                         const double* eps, const double* mu, double* part, double* out, hipStream_t s) {
    const int nb = energy_blocks(npts);
    hipLaunchKernelGGL(k_energy_partial, dim3(nb), dim3(kBlock), 0, s, st, S, npts, n1, jac, w, eps, mu, part);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s, part, nb, out);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_l2(const R* u, long long cs, long long ns, const double* ref, long long npts, int n1, const double* jac,
                     const ElemGeom* geom, double* part, double* out, hipStream_t s) {
    const int nb = diag_blocks();
    hipLaunchKernelGGL(k_l2_partial<R>, dim3(nb), dim3(kBlock), 0, s, u, cs, ns, ref, npts, n1, jac, geom, part);
    hipLaunchKernelGGL(k_l2_final, dim3(1), dim3(64), 0, s, part, nb, out);
    return hipGetLastError();
}
template hipError_t launch_l2<double>(const double*, long long, long long, const double*, long long, int, const double*,
                                      const ElemGeom*, double*, double*, hipStream_t);
template hipError_t launch_l2<float>(const float*, long long, long long, const double*, long long, int, const double*,
                                     const ElemGeom*, double*, double*, hipStream_t);
template hipError_t launch_diag<double>(const double*, long long, long long, long long, int, const double*,
                                        const ElemGeom*, const ElemMat*, double*, double*, hipStream_t);
template hipError_t launch_diag<float>(const float*, long long, long long, long long, int, const double*, const ElemGeom*,
                                       const ElemMat*, double*, double*, hipStream_t);

template <typename D, typename S>
hipError_t launch_convert(D* dst, long long dcs, long long dns, const S* src, long long scs, long long sns, long long n,
                          int ncomp, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_convert<D, S>), dim3(grid_for(n * ncomp)), dim3(kBlock), 0, s, dst, dcs, dns, src, scs, sns, n,
                       ncomp);
    return hipGetLastError();
}
#define NW_CONVERT(D, S)                                                                                        \
    template hipError_t launch_convert<D, S>(D*, long long, long long, const S*, long long, long long, long long, int, \
                                             hipStream_t);
NW_CONVERT(float, double)
NW_CONVERT(double, float)
NW_CONVERT(double, double)
NW_CONVERT(float, float)
#undef NW_CONVERT

hipError_t launch_copy(const void* src, void* dst, long long bytes, hipStream_t s) {
    const long long n4 = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n4, 256) > 8192 ? 8192 : grid_for(n4, 256)), dim3(256), 0, s,
                       static_cast<const float4*>(src), static_cast<float4*>(dst), n4);
    return hipGetLastError();
}

hipError_t launch_fill(double* p, long long n, double v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n, v);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_probe(const R* u, long long cs, long long ns, int nprobes, const long long* base, const double* lw,
                        int n1, double* hist, int* slot, int cap, hipStream_t s) {
    if (nprobes <= 0) return hipSuccess;
    if (n1 < 2 || n1 > kMaxN1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_probe<R>, dim3(nprobes), dim3(256), 0, s, u, cs, ns, nprobes, base, lw, n1, hist, slot, cap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_probe_advance, dim3(1), dim3(64), 0, s, slot, cap);
    return hipGetLastError();
}
template hipError_t launch_probe<double>(const double*, long long, long long, int, const long long*, const double*, int,
                                         double*, int*, int, hipStream_t);
template hipError_t launch_probe<float>(const float*, long long, long long, int, const long long*, const double*, int,
                                        double*, int*, int, hipStream_t);

hipError_t launch_zero_int(int* p, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, p);
    return hipGetLastError();
}


}  // namespace nw
