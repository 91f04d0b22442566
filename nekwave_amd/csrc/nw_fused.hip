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
// Work layout: a persistent workgroup walks element groups of EPW elements,
// one thread per node.  The HBM reads of a group -- the node's six field
// values and the neighbour trace of the thread's face point -- are issued one
// iteration ahead, while the previous group computes; its LSRK registers at
// the top of its own iteration.  The element's fields sit in LDS for the
// three tensor contractions (vector ALUs, or at n1 = 8 optionally the matrix
// cores, with identical results).  Face point
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

// Tangent vectors d x/dr, d x/ds, d x/dt of the trilinear map at (r, s, t), in
// arithmetic T, from the 24 coefficients c[d * 8 + x] (ElemGeom order).
template <typename T>
struct Tangents {
    T r[3], s[3], t[3];
};
template <typename T>
__device__ __forceinline__ Tangents<T> tangents(const T* cf, T r, T s, T t) {
    Tangents<T> G;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T* c = cf + 8 * d;
        G.r[d] = c[1] + c[4] * s + c[5] * t + c[7] * (s * t);
        G.s[d] = c[2] + c[4] * r + c[6] * t + c[7] * (r * t);
        G.t[d] = c[3] + c[5] * r + c[6] * s + c[7] * (r * s);
    }
    return G;
}

template <typename T>
__device__ __forceinline__ void cross3(const T* a, const T* b, T* o) {
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

// Occupancy target (waves per SIMD) that the register budget is sized for.
// FP32 fields: a workgroup of two waves per SIMD gets 4 (two workgroups per CU
// in 128 VGPRs), any other shape 3 (168 VGPRs).  FP64 fields get 2 (256
// VGPRs): the doubled prefetch registers spill otherwise, and a spill reload
// inside the loop waits on every prefetch in flight (measured 1.8x slower).
template <int NT, int RBYTES>
struct WavesPerEU {
    static constexpr int value = RBYTES == 8 ? 2 : (((NT + 63) / 64 + 3) / 4 == 2 ? 4 : 3);
};

// Global traffic goes through raw buffer instructions: the base address lives
// in a buffer resource (SGPRs), each lane supplies a 32-bit byte offset and the
// component stride rides in the scalar offset, so no lane computes a 64-bit
// address.  Offsets past num_records read 0 instead of faulting.  The
// launcher checks that every offset fits in 32 bits.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p) {
    // dword3 0x00020000: the raw-buffer format word of the gfx9 family
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, 0xffffffffu, 0x00020000);
}
__device__ __forceinline__ float bload(float*, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
__device__ __forceinline__ double bload(double*, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
template <typename T>
__device__ __forceinline__ T bld(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
    return bload(static_cast<T*>(nullptr), r, vo, so);
}
__device__ __forceinline__ int4 bld_int4(__amdgpu_buffer_rsrc_t r, unsigned vo) {
    typedef __attribute__((ext_vector_type(4))) int i32x4;
    const i32x4 v = __builtin_bit_cast(i32x4, __builtin_amdgcn_raw_buffer_load_b128(r, vo, 0, 0));
    return make_int4(v[0], v[1], v[2], v[3]);
}
// The six field values of one node (node-major state: 6 consecutive values
// per node) in as few accesses as the node's alignment allows: FP32 nodes sit
// 24 bytes apart (one 16-byte plus one 8-byte access; buffer accesses need only
// dword alignment), FP64 nodes 48 bytes apart (three 16-byte accesses).  Two
// instructions instead of three matter most for the neighbour-trace gathers,
// whose lanes each touch a different cache line.
typedef __attribute__((ext_vector_type(2))) unsigned u32x2;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4;
typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(2))) double f64x2;
// (Each 8-byte word is reinterpreted whole as two floats: extracting single
// 32-bit lanes of the loaded vector before the reinterpretation let the
// compiler narrow the access to its first lane.)
// CP: the access's cache-policy bits (kNT: non-temporal -- no reuse expected,
// so the line is a first candidate for eviction from L2; kSC1: system scope --
// a store is written through L2 instead of left dirty in it).
constexpr int kNT = 2;
constexpr int kSC1 = 16;
template <int CP = 0>
__device__ __forceinline__ void ld6(__amdgpu_buffer_rsrc_t r, unsigned vo, float (&o)[6]) {
    const f32x4 a = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, vo, 0, CP));
    const f32x2 b = __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(r, vo + 16u, 0, CP));
    o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; o[3] = a[3]; o[4] = b[0]; o[5] = b[1];
}
template <int CP = 0>
__device__ __forceinline__ void ld6(__amdgpu_buffer_rsrc_t r, unsigned vo, double (&o)[6]) {
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const f64x2 a = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, vo + 16u * x, 0, CP));
        o[2 * x] = a[0];
        o[2 * x + 1] = a[1];
    }
}
__device__ __forceinline__ void ld4(__amdgpu_buffer_rsrc_t r, unsigned vo, float (&o)[4]) {
    const f32x4 a = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, vo, 0, 0));
#pragma unroll
    for (int c = 0; c < 4; ++c) o[c] = a[c];
}
__device__ __forceinline__ void ld4(__amdgpu_buffer_rsrc_t r, unsigned vo, double (&o)[4]) {
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        const f64x2 a = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, vo + 16u * x, 0, 0));
        o[2 * x] = a[0];
        o[2 * x + 1] = a[1];
    }
}
// The LSRK register k of one node.  k is private to the thread that updates
// the node (it lives only within one step), so its layout serves the accesses:
// per wave of a group, six component blocks of 64 values (a chunk), component c
// of lane l at chunk offset 64 c + l -- every wave reads and writes whole cache
// lines of each component, whatever the shape of the group.  vo: the lane's
// byte offset in its wave's chunk (loop-invariant); so: the chunk's byte offset
// (wave-uniform, in the instruction's scalar offset).
template <int CP>
__device__ __forceinline__ void ldk(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, float (&o)[6]) {
#pragma unroll
    for (int c = 0; c < 6; ++c)
        o[c] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo + 256u * c, so, CP));
}
template <int CP>
__device__ __forceinline__ void ldk(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, double (&o)[6]) {
#pragma unroll
    for (int c = 0; c < 6; ++c)
        o[c] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo + 512u * c, so, CP));
}
template <int CP>
__device__ __forceinline__ void stk(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, const float (&v)[6]) {
#pragma unroll
    for (int c = 0; c < 6; ++c)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v[c]), r, vo + 256u * c, so, CP);
}
template <int CP>
__device__ __forceinline__ void stk(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, const double (&v)[6]) {
#pragma unroll
    for (int c = 0; c < 6; ++c)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v[c]), r, vo + 512u * c, so, CP);
}
template <int CP = 0>
__device__ __forceinline__ void st6(__amdgpu_buffer_rsrc_t r, unsigned vo, const float (&v)[6]) {
    const f32x4 a = {v[0], v[1], v[2], v[3]};
    const f32x2 b = {v[4], v[5]};
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, a), r, vo, 0, CP);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, b), r, vo + 16u, 0, CP);
}
template <int CP = 0>
__device__ __forceinline__ void st6(__amdgpu_buffer_rsrc_t r, unsigned vo, const double (&v)[6]) {
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const f64x2 a = {v[2 * x], v[2 * x + 1]};
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, a), r, vo + 16u * x, 0, CP);
    }
}

// Split BF16: four FP32 values as round-to-nearest BF16 halves hi = bf16(x),
// lo = bf16(x - hi), packed as the 4-element operands of the BF16 matrix
// instructions (x = hi + lo to 2^-17 relative; the product of two such pairs
// without lo*lo to about 2^-16).
typedef __attribute__((ext_vector_type(4))) short v4s;
typedef __attribute__((ext_vector_type(2))) __bf16 v2bf;
__device__ __forceinline__ void split4(const float* x, v4s& hi, v4s& lo) {
    unsigned h[2], l[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const v2bf a = {(__bf16)x[2 * t], (__bf16)x[2 * t + 1]};
        const v2bf b = {(__bf16)(x[2 * t] - (float)a[0]), (__bf16)(x[2 * t + 1] - (float)a[1])};
        h[t] = __builtin_bit_cast(unsigned, a);
        l[t] = __builtin_bit_cast(unsigned, b);
    }
    hi = __builtin_bit_cast(v4s, (unsigned long long)h[0] | ((unsigned long long)h[1] << 32));
    lo = __builtin_bit_cast(v4s, (unsigned long long)l[0] | ((unsigned long long)l[1] << 32));
}

// 1/J: in FP32 the hardware reciprocal (1 ulp); in FP64 the IEEE quotient.
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
// |N| and 1/|N| from |N|^2 with the reference's guard |N| > 1e-14
// (mesh.cpp:1032-1035): FP32 through the hardware reciprocal square root,
// FP64 through the IEEE square root and quotient.
__device__ __forceinline__ void norm_and_inverse(float n2, float& len, float& inv) {
    inv = n2 > 1e-28f ? __builtin_amdgcn_rsqf(n2) : 0.0f;
    len = n2 * inv;
}
__device__ __forceinline__ void norm_and_inverse(double n2, double& len, double& inv) {
    len = sqrt(n2);
    inv = len > 1e-14 ? 1.0 / len : 0.0;
}

// Geometry at reference point (r, s, t) of a trilinear element: the cofactors
// C = J J^-T (mesh.cpp:291-309 times J) and 1/J, in the field precision R (FP64
// fields: FP64 arithmetic; FP32 fields: FP32 arithmetic on the coefficients
// rounded once, relative error ~1e-7, below the FP32 fields' own rounding).  On
// an affine element (c4..c7 exactly 0) every product with them is an exact
// zero, so the result is the same bits at every node.
template <typename R>
__device__ __forceinline__ void node_geometry(const R* cf, R r, R s, R t, R (&C)[10]) {
    const Tangents<R> T = tangents(cf, r, s, t);
    R Cr[3], Cs[3], Ct[3];
    cross3(T.s, T.t, Cr);  // (rx, ry, rz) * J
    cross3(T.t, T.r, Cs);  // (sx, sy, sz) * J
    cross3(T.r, T.s, Ct);  // (tx, ty, tz) * J
    const R det = T.r[0] * Cr[0] + T.r[1] * Cr[1] + T.r[2] * Cr[2];
    C[0] = Cr[0]; C[1] = Cr[1]; C[2] = Cr[2];
    C[3] = Cs[0]; C[4] = Cs[1]; C[5] = Cs[2];
    C[6] = Ct[0]; C[7] = Ct[1]; C[8] = Ct[2];
    C[9] = recip(det);
}

// Records of an affine segment: each element's geometry computed once (at the
// reference origin, equal to every node's) into the first ten words of its
// 26-word record, as doubles holding the R values exactly.
template <typename R>
__global__ void k_affine_records(double* rec, int nelt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelt) return;
    double* q = rec + size_t(e) * 26;
    R cf[24];
#pragma unroll
    for (int x = 0; x < 24; ++x) cf[x] = R(q[x]);
    R C[10];
    node_geometry(cf, R(0), R(0), R(0), C);
#pragma unroll
    for (int x = 0; x < 10; ++x) q[x] = double(C[x]);
}

// Persistent grid: workgroup W walks its own sequence of element groups,
// gorder[s0 .. s0 + tl] from its header (built on the host: along the x-chains of the
// mesh, neighbouring sequences on one XCD, so face neighbours' traces are read
// while still in that XCD's L2, and an x face whose neighbour is the previous
// or next group in LDS needs no gather).  HBM reads are issued one or two
// iterations before they are needed, so they travel while earlier groups
// compute.  Per group:
//   1. the node's face (Nanson) vectors from its geometry (computed during the
//      previous iteration) and its field values go to LDS;           barrier
//   2. the next group's reads are issued; volume: the three reference
//      derivatives of each field from LDS folded straight into J * curl
//      (matrix-core variant: derivatives to LDS, folded after the barrier);
//   3. face point (f, p, q): jumps, flux, dA -> LDS;                barrier
//   4. lift: each node gathers its (at most three) face terms, one per
//      direction, in fixed face order; inverse mass; LSRK update; stores;
//      the next group's geometry.
// AFF: the segment is affine and its records hold the geometry (FusedArgs::affine).
template <int N1, int EPW, bool FIRST, typename R, typename KR, int MX, bool AFF>
__global__ __launch_bounds__(EPW* N1* N1* N1) __attribute__((amdgpu_waves_per_eu(WavesPerEU<EPW * N1 * N1 * N1, sizeof(R)>::value)))
void k_fused_stage(FusedArgs A) {
    constexpr int N2 = N1 * N1, NP = N1 * N2, NF = 6 * N2, NT = EPW * NP;
    constexpr int FPT = (EPW * NF + NT - 1) / NT;  // face points per thread (1 for n1 >= 6)
    constexpr int NW = (NT + 63) / 64;             // waves per workgroup (k chunks per group)
    constexpr int ERW = 26;                         // doubles per element record: geometry (24), 1/eps, 1/mu
    constexpr int GQ = (EPW * ERW + NT - 1) / NT;   // element-record words per thread
    struct ElemRec {
        ElemGeom g;
        double ieps, imu;
    };
    static_assert(sizeof(ElemRec) == ERW * sizeof(double), "element record layout");
    // MX: the volume contractions on the matrix cores (n1 = 8, FP32 fields,
    // one element per workgroup) -- 1: FP32 products, 2: split BF16 products
    // (hi*hi + hi*lo + lo*hi with FP32 accumulation) -- else (0) on the vector ALUs.
    static_assert(!MX || (N1 == 8 && EPW == 1 && sizeof(R) == 4), "matrix-core volume shape");
    __shared__ R sU[EPW][6][NP];
    __shared__ R sF[EPW * NF + 1][6];  // face terms dA*F; the last row stays zero
    constexpr int SPARE = sizeof(R) == 4 ? 64 : 0;  // spare face rows (branch-free writes, FP32)
    __shared__ R sC[EPW * NF + SPARE][9];  // per face point: Nanson vector, minus-side trace (6)
    __shared__ R sD[N2];               // D, row-major: sD[i * N1 + l] = D_il
    __shared__ ElemRec sE[2][EPW];     // element records, double-buffered over groups
    // FP32 fields: the records' 24 trilinear coefficients rounded to R once per
    // group (the per-node geometry reads them); FP64 fields read sE itself
    constexpr bool CFR = sizeof(R) == 4 && !AFF;
    __shared__ R sEc[CFR ? 2 : 1][CFR ? EPW : 1][24];
    // the i = n1-1 plane of the last slot of the previous group (its pre-stage
    // fields), rows of 7 words so a wave's reads of consecutive rows do not conflict
    __shared__ R sP3[N2][7];
    // n1 = 8, 9, FP32, vector-ALU volume: the fields again with the s-lines
    // (i, k) and t-lines (i, j) contiguous, lines 12 words apart (16-byte
    // aligned; at n1 = 8 eight lines cover all 32 banks), so those directions
    // read their lines with 16-byte accesses instead of 4-byte ones at strides
    // of n1 and n1^2 words (measured faster at n1 = 8, 9; slower at 6 and 7)
    constexpr bool TL = (N1 == 8 || N1 == 9) && EPW == 1 && sizeof(R) == 4 && !MX;
    constexpr int RW = 12;
    __shared__ __attribute__((aligned(16))) R sUs[TL ? 6 : 1][TL ? RW * N2 : 1];
    __shared__ __attribute__((aligned(16))) R sUt[TL ? 6 : 1][TL ? RW * N2 : 1];
    // matrix-core path: the 18 reference derivatives of every node, by direction
    // r (node order), s (line (i, k) major, j fastest), t (line (i, j) major, k
    // fastest, lines 9 words apart so the node-order reads are conflict-free)
    __shared__ R sA[MX ? 2 : 1][MX ? 6 : 1][MX ? NP : 1];
    __shared__ R sAt[MX ? 6 : 1][MX ? 9 * N2 : 1];

    // This workgroup's groups: gorder[s0 .. s0 + tl], in that order.  The
    // headers are stored by hardware workgroup id, already remapped XCD-aware
    // (fused_sequence_of_block: the hardware deals workgroups round-robin over
    // the 8 XCDs, and neighbouring sequences, with most face neighbours, go to
    // one XCD's L2); a header also holds the first three groups, so the first
    // field reads wait on one scalar read.  Workgroups past the sequences have
    // tl = -1 and return (uniform over the block, before any barrier).
    typedef const __attribute__((address_space(4))) int CInt;
    CInt* const hd = (CInt*)A.ghead + kFusedHead * blockIdx.x;
    const int s0 = hd[0];
    const int tl = hd[1];  // last iteration
    if (tl < 0) return;

    const int tid = threadIdx.x;
    const int es = tid / NP;
    const int m = tid - es * NP;
    const int i = m % N1, j = (m / N1) % N1, k = m / N2;
    // FP32 at n1 = 8, one element per workgroup: each wave holds one k-plane, so
    // whether its nodes lie on a t-face is wave-uniform and the t-face row
    // writes and t-lift reads of the six interior planes are skipped by a
    // scalar branch (same arithmetic: the lift still adds the zero row's 0)
    const bool tface = (N2 == 64 && EPW == 1 && sizeof(R) == 4)
                           ? (__builtin_amdgcn_readfirstlane(k) == 0 || __builtin_amdgcn_readfirstlane(k) == N1 - 1)
                           : true;
    const __amdgpu_buffer_rsrc_t rU = make_rsrc(A.uin), rO = make_rsrc(A.uout), rK = make_rsrc(A.kbuf);
    const __amdgpu_buffer_rsrc_t rW = make_rsrc(A.fw), rN = make_rsrc(A.fnb);
    const __amdgpu_buffer_rsrc_t rR = make_rsrc(A.rec);
    constexpr unsigned UB = 6 * sizeof(R);  // bytes per node of u
    const unsigned kv = unsigned(384 * (tid >> 6) + (tid & 63)) * unsigned(sizeof(KR));  // lane's k offset in the group's chunks

    // loop-invariant: node coordinates, inverse mass weight
    const R zr = R(c_gll.z[N1][i]), zs = R(c_gll.z[N1][j]), zt = R(c_gll.z[N1][k]);
    const R Winv = R(c_gll.iw[N1][i] * c_gll.iw[N1][j] * c_gll.iw[N1][k]);
    R wpq[FPT];  // w_p w_q of the thread's face point(s)
#pragma unroll
    for (int t = 0; t < FPT; ++t) {
        const int r = (tid + t * NT) % NF, pq = r % N2;
        wpq[t] = R(c_gll.w[N1][pq % N1] * c_gll.w[N1][pq / N1]);
    }
    const double dt = *A.dt;  // read once: a load inside the loop would wait on the prefetch
    for (int x = tid; x < N2; x += NT) sD[x] = gD<R>(N1, x);
    for (int x = tid; x < 6; x += NT) sF[EPW * NF][x] = R(0);

    auto clamp_e = [&](int g, int slot) { return min(g * EPW + slot, A.nelt - 1); };
    // group of iteration t (past the last iteration: the last group again)
    // Wave-uniform reads go through the scalar unit instead of the vector
    // memory pipeline (a 64-lane load of one 16-byte row costs the address
    // unit as much as 1 KiB of field data): the group order always, and at
    // n1 = 8 with one element per workgroup, where wave w owns exactly the
    // face points of face w, the face-neighbour row and the flux weights.
    // They are issued at the top of an iteration and first waited on at its
    // first barrier, before which the iteration reads no LDS.
    constexpr bool SF = N2 == 64 && EPW == 1;
    const int wface = min(__builtin_amdgcn_readfirstlane(tid >> 6), 5);  // SF: this wave's face
    CInt* const sGo = (CInt*)A.gorder;
    CInt* const sNb = (CInt*)A.fnb;
    auto order_s = [&](int t) -> int { return sGo[s0 + min(t, tl)]; };
    auto node_of = [&](int g) -> unsigned {  // clamped: every read stays in bounds
        return static_cast<unsigned>(A.node_base + (long long)clamp_e(g, es) * NP + m);
    };
    // Every read an iteration depends on is issued ahead of it, so no wave
    // waits on memory latency inside an iteration:
    //   three iterations ahead: the group order entry;
    //   two groups ahead: the face-neighbour table rows (they address the
    //   neighbour-trace gather) and the element records (staged into LDS one
    //   iteration ahead);
    //   one group ahead, after the iteration's first barrier: the fields and
    //   the neighbour traces;
    //   at the top of the iteration: LSRK registers and flux weights, needed
    //   only at its end.
    struct Nbr {
        int4 b[FPT];
    };
    struct Rec {
        double w[GQ];
    };
    // Where a face point's plus-side trace comes from (an x face -- face 1 or
    // 3 -- whose neighbour element is already in LDS needs no gather):
    //   kGather: the gathered registers up[t];
    //   kSlot:   the neighbouring slot of this group (face 1 <-> slot + 1, face 3 <-> slot - 1);
    //   kCarry:  face 3 of slot 0, whose neighbour is the last slot of this
    //            workgroup's previous group: its i = n1-1 plane, kept in sP3
    //            (walking backwards: face 1 of the last slot, whose neighbour
    //            is slot 0 of the previous group: its i = 0 plane);
    //   kMirror: PEC mirror face, no plus side.
    enum { kGather = 0, kSlot = 1, kCarry = 2, kMirror = 3 };
    struct Pre {
        R u[6];          // this node's fields
        R up[FPT][6];    // plus-side trace of each face point (kGather only)
        int kind[FPT];
    };
    auto fetch_nbr = [&](int g, Nbr& N) {
#pragma unroll
        for (int t = 0; t < FPT; ++t) {
            const int fp = tid + t * NT;
            const int efs = min(fp / NF, EPW - 1), f = (fp - (fp / NF) * NF) / N2;
            if constexpr (SF)
            {
                CInt* r = sNb + (clamp_e(g, 0) * 6 + wface) * 4;
                N.b[t] = make_int4(r[0], r[1], r[2], r[3]);
            }
            else
                N.b[t] = bld_int4(rN, unsigned(clamp_e(g, efs) * 6 + f) * 16u);
        }
    };
    auto fetch_rec = [&](int g, Rec& Q) {
#pragma unroll
        for (int t = 0; t < GQ; ++t) {
            const int x = min(tid + t * NT, EPW * ERW - 1);  // word x of the group's records
            const unsigned el = unsigned(clamp_e(g, x / ERW));
            Q.w[t] = bld<double>(rR, (el * unsigned(ERW) + unsigned(x - (x / ERW) * ERW)) * 8u, 0);
        }
    };
    auto put_rec = [&](const Rec& Q, int buf) {
#pragma unroll
        for (int t = 0; t < GQ; ++t) {
            const int x = tid + t * NT;
            if (x < EPW * ERW) (&sE[buf][0].g.c[0][0])[x] = Q.w[t];
            if constexpr (CFR) {
                const int el = x / ERW, w = x - el * ERW;
                if (x < EPW * ERW && w < 24) sEc[buf][el][w] = R(Q.w[t]);
            }
        }
    };
    // gprev: the group this workgroup processes just before g (-1: none)
    auto fetch = [&](int g, const Nbr& N, Pre& P, int gprev) {
        const unsigned nd = node_of(g);
        ld6(rU, nd * UB, P.u);
        // issued by every thread (threads without a face point, and x faces
        // served from LDS, read their own node): the same number of loads on
        // every path keeps the compiler's memory-counter waits exact instead of
        // conservative
#pragma unroll
        for (int t = 0; t < FPT; ++t) {
            const int fp = tid + t * NT;
            const int efs = min(fp / NF, EPW - 1);
            const int r = fp - (fp / NF) * NF;
            const int f = r / N2;
            const int pq = r - f * N2, p = pq % N1, q = pq / N1;
            const int4 b = N.b[t];
            int kind = b.x < 0 ? kMirror : kGather;
            // an x face paired with the opposite x face of a neighbour in the
            // same orientation (plus node of (p, q) = base + p n1 + q n1^2)
            // (node indices fit in 32 bits: the launcher checks the byte offsets)
            if (fp < EPW * NF && b.y == N1 && b.z == N2) {
                const unsigned nb = unsigned(A.node_base), bx = unsigned(b.x);
                if (f == 1 && efs + 1 < EPW && bx == nb + unsigned(clamp_e(g, efs + 1)) * NP) kind = kSlot;
                if (f == 3 && efs > 0 && bx == nb + unsigned(clamp_e(g, efs - 1)) * NP + (N1 - 1)) kind = kSlot;
                if (A.backwards) {
                    if (f == 1 && efs == EPW - 1 && gprev >= 0 && bx == nb + unsigned(clamp_e(gprev, 0)) * NP)
                        kind = kCarry;
                } else if (f == 3 && efs == 0 && gprev >= 0 &&
                           bx == nb + unsigned(clamp_e(gprev, EPW - 1)) * NP + (N1 - 1)) {
                    kind = kCarry;
                }
            }
            // at n1 = 8 with one element per workgroup a wave holds one face: uniform
            if constexpr (SF) kind = __builtin_amdgcn_readfirstlane(kind);
            P.kind[t] = kind;
            const unsigned idx = (kind == kGather && fp < EPW * NF) ? unsigned(b.x + p * b.y + q * b.z) : nd;
            ld6(rU, idx * UB, P.up[t]);
        }
    };

    // prologue: records of the first group in LDS, its fields in flight,
    // two-ahead reads of the second
    Pre cur;
    Nbr nb1;
    Rec rec1;
    // groups of iterations t, t + 1, t + 2 (wave-uniform)
    int gA = hd[2], gB = hd[3], gC = hd[4];
    {
        Nbr nb0;
        Rec rec0;
        fetch_nbr(gA, nb0);
        fetch_rec(gA, rec0);
        fetch_nbr(gB, nb1);
        fetch_rec(gB, rec1);
        // the fields first: staging the record waits for it, and the field
        // reads then travel during that wait instead of after it
        fetch(gA, nb0, cur, -1);
        put_rec(rec0, 0);
    }
    // Wait for the prologue's reads here: left in flight, the compiler's
    // memory-counter analysis merges them into the loop and then waits, at the
    // first use of each iteration's prefetched values, on that iteration's
    // most recent stores as well.
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
    __syncthreads();

    // Geometry of the thread's node from an element record (node_geometry, in
    // the field precision): cofactors C = J J^-T (mesh.cpp:291-309 times J) and
    // 1/J.  A group's geometry is computed at the end of the previous
    // iteration, where the waves otherwise wait for the prefetched fields to
    // land.  Affine segments (every element a parallelepiped) carry these ten
    // values, precomputed once per element by k_affine_records with the same
    // arithmetic, in the record's first ten words: read them instead.
    auto geometry = [&](int buf, R(&C)[10]) {
        if constexpr (AFF) {
#pragma unroll
            for (int x = 0; x < 10; ++x) C[x] = R((&sE[buf][es].g.c[0][0])[x]);
        } else if constexpr (CFR) {
            node_geometry(&sEc[buf][es][0], zr, zs, zt, C);
        } else {
            node_geometry(&sE[buf][es].g.c[0][0], zr, zs, zt, C);
        }
    };
    R geo[10];
    geometry(0, geo);

    for (int t = 0; t <= tl; ++t) {
        const int g = gA;
        Nbr nb2;
        Rec rec2;
        fetch_nbr(gC, nb2);
        fetch_rec(gC, rec2);
        const int gnext = order_s(t + 3);
        // reads of this group needed late (LSRK registers, flux weights) go
        // before the prefetch: waiting on them does not wait on it
        const unsigned nd = node_of(g);
        KR kold[6] = {KR(0), KR(0), KR(0), KR(0), KR(0), KR(0)};
        // group g's waves hold k chunks kbase + NW g + (0 .. NW-1)
        const unsigned ks = (unsigned(A.kbase) + unsigned(g) * unsigned(NW)) * (384u * sizeof(KR));
        if (!FIRST) {  // non-temporal where the caller asks (stream_k: no L2 reuse)
            if (A.stream_k)
                ldk<kNT>(rK, kv, ks, kold);
            else
                ldk<0>(rK, kv, ks, kold);
        }
        R fw[FPT][4];
#pragma unroll
        for (int t = 0; t < FPT; ++t) {
            const int fp = tid + t * NT;
            const int efs = min(fp / NF, EPW - 1), f = (fp - (fp / NF) * NF) / N2;
            const unsigned face = unsigned(clamp_e(g, efs) * 6 + f);
            if constexpr (SF) {
                typedef const __attribute__((address_space(4))) R CR;
                CR* w = (CR*)A.fw + size_t(clamp_e(g, 0) * 6 + wface) * 4;
#pragma unroll
                for (int c = 0; c < 4; ++c) fw[t][c] = w[c];
            } else {
                ld4(rW, face * unsigned(4 * sizeof(R)), fw[t]);
            }
        }
        Pre nxt;

        const ElemRec& ER = sE[t & 1][es];

        // ---- 1. this group's geometry (computed at the end of the previous iteration) ----
        const R Crx = geo[0], Cry = geo[1], Crz = geo[2], Csx = geo[3], Csy = geo[4], Csz = geo[5];
        const R Ctx = geo[6], Cty = geo[7], Ctz = geo[8], iJ = geo[9];
        // The Nanson vector of a face point is the cofactor row of its face's
        // direction at that node (mesh.cpp:1011-1038: -+J (rx, ry, rz) etc.),
        // outward: r-faces +-C_r, s-faces +-C_s, t-faces +-C_t.
        // The node also hands its field values (the face point's minus-side
        // trace) to the face point: rows of 9 values are read without LDS bank
        // conflicts, which strided reads of the face nodes out of sU are not.
        // FP32: branch-free, every node writes one row per direction -- its
        // face point's row where it lies on that direction's face, else a
        // spare row of its lane that nobody reads (no branches around the LDS
        // stores; measured faster).  FP64: only the face nodes write (the
        // spare rows would double the LDS bytes; measured slower).
        auto row = [&](int x, R a, R b, R c) {
            R* o = sC[x];
            o[0] = a; o[1] = b; o[2] = c;
#pragma unroll
            for (int c2 = 0; c2 < 6; ++c2) o[3 + c2] = cur.u[c2];
        };
        if constexpr (SPARE > 0) {
            const int spare = EPW * NF + (tid & 63);
            const int rs = j == 0 ? es * NF + 0 * N2 + i + N1 * k : (j == N1 - 1 ? es * NF + 2 * N2 + i + N1 * k : spare);
            const int rr = i == N1 - 1 ? es * NF + 1 * N2 + j + N1 * k : (i == 0 ? es * NF + 3 * N2 + j + N1 * k : spare);
            const int rt = k == 0 ? es * NF + 4 * N2 + i + N1 * j : (k == N1 - 1 ? es * NF + 5 * N2 + i + N1 * j : spare);
            const R ss = j == 0 ? R(-1) : R(1), sr = i == 0 ? R(-1) : R(1), st = k == 0 ? R(-1) : R(1);
            row(rs, ss * Csx, ss * Csy, ss * Csz);
            row(rr, sr * Crx, sr * Cry, sr * Crz);
            if (tface) row(rt, st * Ctx, st * Cty, st * Ctz);
        } else {
            const int b = es * NF;
            if (j == 0) row(b + 0 * N2 + i + N1 * k, -Csx, -Csy, -Csz);
            if (i == N1 - 1) row(b + 1 * N2 + j + N1 * k, Crx, Cry, Crz);
            if (j == N1 - 1) row(b + 2 * N2 + i + N1 * k, Csx, Csy, Csz);
            if (i == 0) row(b + 3 * N2 + j + N1 * k, -Crx, -Cry, -Crz);
            if (k == 0) row(b + 4 * N2 + i + N1 * j, -Ctx, -Cty, -Ctz);
            if (k == N1 - 1) row(b + 5 * N2 + i + N1 * j, Ctx, Cty, Ctz);
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) sU[es][c][m] = cur.u[c];
        if constexpr (TL) {
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                sUs[c][(i + N1 * k) * RW + j] = cur.u[c];
                sUt[c][(i + N1 * j) * RW + k] = cur.u[c];
            }
        }
        __syncthreads();
        put_rec(rec1, (t + 1) & 1);  // read from the next iteration on, after the barrier below
        // The next group's fields and neighbour traces: issued after the barrier
        // so the burst of vector-memory instructions at the top of the
        // iteration, which every wave of the CU issues at once, stays short
        // and the barrier comes sooner; they still land before the iteration
        // ends (and in flight while this group computes).
        fetch(gB, nb1, nxt, gA);

        // ---- 2. volume: reference derivatives (local_grad3, dg_kernels.cu:101-124)
        // folded straight into J * curl (dg_kernels.cu:140-149) ----
        R cE[3] = {R(0), R(0), R(0)}, cH[3] = {R(0), R(0), R(0)};
        if constexpr (MX) {
            // 36 tasks (direction d, field c, half of the 64 lines), in pairs:
            // one v_mfma_f32_4x4x1_16b per column q of D and task.  Lane 4b + x
            // supplies A = D[4h + x][q] and B = u(line 4(b >> 1) + x, q) of block
            // b (h = b & 1 picks the output rows) and receives the derivative of
            // that line at rows 4h..4h+3; each step is one fused multiply-add,
            // so the result equals the vector contraction bit for bit.  Waves
            // 0-5, which also compute face terms, take two pairs, waves 6, 7 three.
            typedef __attribute__((ext_vector_type(4))) float v4f;
            const int lane = tid & 63, wv = tid >> 6;
            const int h = (lane >> 2) & 1;
            const int lq = 4 * (lane >> 3) + (lane & 3);
            R Ar[N1];
#pragma unroll
            for (int q = 0; q < N1; ++q) Ar[q] = sD[(4 * h + (lane & 3)) * N1 + q];
            // split BF16: the lane's row of D as hi and lo halves, q = 0..3 and 4..7
            v4s Ahi[2], Alo[2];
            if constexpr (MX == 2) {
#pragma unroll
                for (int kh = 0; kh < 2; ++kh) split4(&Ar[4 * kh], Ahi[kh], Alo[kh]);
            }
            const int p0 = wv < 6 ? 2 * wv : 12 + 3 * (wv - 6), np = wv < 6 ? 2 : 3;
            for (int pp = 0; pp < np; ++pp) {
                const int P = p0 + pp, d = P / 6, c = P - 6 * d;
                const R* u = sU[0][c];
                R B0[N1], B1[N1];
                const int l0 = lq, l1 = lq + 32;
                if (d == 0) {
#pragma unroll
                    for (int q = 0; q < N1; ++q) { B0[q] = u[8 * l0 + q]; B1[q] = u[8 * l1 + q]; }
                } else if (d == 1) {
#pragma unroll
                    for (int q = 0; q < N1; ++q) { B0[q] = u[(l0 & 7) + 64 * (l0 >> 3) + 8 * q]; B1[q] = u[(l1 & 7) + 64 * (l1 >> 3) + 8 * q]; }
                } else {
#pragma unroll
                    for (int q = 0; q < N1; ++q) { B0[q] = u[l0 + 64 * q]; B1[q] = u[l1 + 64 * q]; }
                }
                v4f a0 = {0.0f, 0.0f, 0.0f, 0.0f}, a1 = {0.0f, 0.0f, 0.0f, 0.0f};
                if constexpr (MX == 1) {
#pragma unroll
                    for (int q = 0; q < N1; ++q) {
                        a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(Ar[q], B0[q], a0, 0, 0, 0);
                        a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(Ar[q], B1[q], a1, 0, 0, 0);
                    }
                } else {
                    // v_mfma_f32_4x4x4_16b_bf16: lane 4b + r holds row r of block b's
                    // A (k = 0..3), lane 4b + j column j of its B, the same lane
                    // placement as the FP32 form; the small products first
#pragma unroll
                    for (int kh = 0; kh < 2; ++kh) {
                        v4s Bh0, Bl0, Bh1, Bl1;
                        split4(&B0[4 * kh], Bh0, Bl0);
                        split4(&B1[4 * kh], Bh1, Bl1);
                        a0 = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(Alo[kh], Bh0, a0, 0, 0, 0);
                        a1 = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(Alo[kh], Bh1, a1, 0, 0, 0);
                        a0 = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(Ahi[kh], Bl0, a0, 0, 0, 0);
                        a1 = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(Ahi[kh], Bl1, a1, 0, 0, 0);
                        a0 = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(Ahi[kh], Bh0, a0, 0, 0, 0);
                        a1 = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(Ahi[kh], Bh1, a1, 0, 0, 0);
                    }
                }
                if (d < 2) {
                    *reinterpret_cast<v4f*>(&sA[d][c][8 * l0 + 4 * h]) = a0;
                    *reinterpret_cast<v4f*>(&sA[d][c][8 * l1 + 4 * h]) = a1;
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) { sAt[c][9 * l0 + 4 * h + r] = a0[r]; sAt[c][9 * l1 + 4 * h + r] = a1[r]; }
                }
            }
        } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {  // reference direction r, s, t: one row of D live at a time
            const int row = d == 0 ? i : (d == 1 ? j : k);
            R Dr[N1];
#pragma unroll
            for (int q = 0; q < N1; ++q) Dr[q] = sD[row * N1 + q];
            const R Cx = d == 0 ? Crx : (d == 1 ? Csx : Ctx);
            const R Cy = d == 0 ? Cry : (d == 1 ? Csy : Cty);
            const R Cz = d == 0 ? Crz : (d == 1 ? Csz : Ctz);
            // E then H: three fields' contractions as independent chains, interleaved
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                R a[3] = {R(0), R(0), R(0)};
#pragma unroll
                for (int q = 0; q < N1; ++q) {
                    const int at = d == 0 ? q + N1 * j + N2 * k : (d == 1 ? i + N1 * q + N2 * k : i + N1 * j + N2 * q);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const R v = !TL || d == 0 ? sU[es][3 * h + c][at]
                                                  : (d == 1 ? sUs[3 * h + c][(i + N1 * k) * RW + q] : sUt[3 * h + c][(i + N1 * j) * RW + q]);
                        a[c] += Dr[q] * v;
                    }
                }
                // J d/dx_e of component c gets a * C_d,e;
                // curl u = (dy u_z - dz u_y, dz u_x - dx u_z, dx u_y - dy u_x)
                R* cv = h == 0 ? cE : cH;
                cv[1] += a[0] * Cz; cv[2] -= a[0] * Cy;
                cv[0] -= a[1] * Cz; cv[2] += a[1] * Cx;
                cv[0] += a[2] * Cy; cv[1] -= a[2] * Cx;
            }
            // keep each direction's contraction here: sinking it into the update
            // below would hold the fields' LDS values in registers
            pin(cE[0]); pin(cE[1]); pin(cE[2]); pin(cH[0]); pin(cH[1]); pin(cH[2]);
        }
        }

        // ---- 3. face terms dA * F of this thread's face point(s) ----
#pragma unroll
        for (int t = 0; t < FPT; ++t) {
            const int fp = tid + t * NT;
            if (fp < EPW * NF) {
                const int efs = fp / NF, r = fp - efs * NF;
                const int f = r / N2, pq = r - f * N2, p = pq % N1, q = pq / N1;
                const R* mine = &sC[fp][3];
                R jE[3], jH[3];
                const int kind = cur.kind[t];
                if (kind == kMirror) {  // PEC mirror: [[E]] = -2 E-, [[H]] = 0 (dg_kernels.cu:244-251)
#pragma unroll
                    for (int c = 0; c < 3; ++c) { jE[c] = R(-2) * mine[c]; jH[c] = R(0); }
                } else {
                    // plus trace: kSlot -- face 1: node (0, p, q) of slot + 1, face 3:
                    // node (n1-1, p, q) of slot - 1; kCarry -- the kept plane
                    const R* lp = &sP3[pq][0];
                    int ls = 1;
                    if constexpr (EPW > 1) {
                        if (kind == kSlot) {
                            lp = &sU[f == 1 ? efs + 1 : efs - 1][0][(f == 1 ? 0 : N1 - 1) + N1 * p + N2 * q];
                            ls = NP;
                        }
                    }
                    // the plus trace as one block (kind is wave-uniform at n1 = 8:
                    // one branch instead of one per component)
                    R plus[6];
                    if (kind != kGather) {
#pragma unroll
                        for (int c = 0; c < 6; ++c) plus[c] = lp[c * ls];
                    } else {
#pragma unroll
                        for (int c = 0; c < 6; ++c) plus[c] = cur.up[t][c];
                    }
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        jE[c] = plus[c] - mine[c];
                        jH[c] = plus[c + 3] - mine[c + 3];
                    }
                }
                // unit normal and area element (mesh.cpp:1032-1038)
                const R Nx = sC[fp][0], Ny = sC[fp][1], Nz = sC[fp][2];
                R Ln, il;
                norm_and_inverse(Nx * Nx + Ny * Ny + Nz * Nz, Ln, il);
                const R nx = Nx * il, ny = Ny * il, nz = Nz * il;
                const R dA = Ln * wpq[t];
                // central / upwind flux (dg_kernels.cu:265-294), impedance weighted with media
                const R nxEx = ny * jE[2] - nz * jE[1], nxEy = nz * jE[0] - nx * jE[2], nxEz = nx * jE[1] - ny * jE[0];
                const R nxHx = ny * jH[2] - nz * jH[1], nxHy = nz * jH[0] - nx * jH[2], nxHz = nx * jH[1] - ny * jH[0];
                const R nnEx = ny * nxEz - nz * nxEy, nnEy = nz * nxEx - nx * nxEz, nnEz = nx * nxEy - ny * nxEx;
                const R nnHx = ny * nxHz - nz * nxHy, nnHy = nz * nxHx - nx * nxHz, nnHz = nx * nxHy - ny * nxHx;
                // F_E = wEh nx[[H]] - wEe nxnx[[E]];  F_H = -wHe nx[[E]] - wHh nxnx[[H]]
                const R wEh = fw[t][0], wEe = fw[t][1], wHe = fw[t][2], wHh = fw[t][3];
                R* out = sF[fp];
                out[0] = dA * (wEh * nxHx - wEe * nnEx);
                out[1] = dA * (wEh * nxHy - wEe * nnEy);
                out[2] = dA * (wEh * nxHz - wEe * nnEz);
                out[3] = dA * (-wHe * nxEx - wHh * nnHx);
                out[4] = dA * (-wHe * nxEy - wHh * nnHy);
                out[5] = dA * (-wHe * nxEz - wHh * nnHz);
            }
        }
        __syncthreads();  // face terms complete
        if constexpr (MX) {  // fold the node's 18 derivatives into J * curl
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const R Cx = d == 0 ? Crx : (d == 1 ? Csx : Ctx);
                const R Cy = d == 0 ? Cry : (d == 1 ? Csy : Cty);
                const R Cz = d == 0 ? Crz : (d == 1 ? Csz : Ctz);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    R a[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        a[c] = d == 0 ? sA[0][3 * h + c][m] : (d == 1 ? sA[1][3 * h + c][j + 8 * i + 64 * k] : sAt[3 * h + c][k + 9 * (i + 8 * j)]);
                    R* cv = h == 0 ? cE : cH;
                    cv[1] += a[0] * Cz; cv[2] -= a[0] * Cy;
                    cv[0] -= a[1] * Cz; cv[2] += a[1] * Cx;
                    cv[0] += a[2] * Cy; cv[1] -= a[2] * Cx;
                }
            }
        }

        // ---- 4. lift: one face term per direction (the zero row off the faces).
        // A node meets at most one face per direction; the sum (s + r) + t equals
        // the reference's face-order gather 0..5 bit for bit (its first two
        // terms commute exactly and a t-face always comes last) ----
        R LE[3], LH[3];
        {
            const int zero = EPW * NF, base = es * NF;
            const int sj = j == 0 ? base + 0 * N2 + i + N1 * k : (j == N1 - 1 ? base + 2 * N2 + i + N1 * k : zero);
            const int si = i == N1 - 1 ? base + 1 * N2 + j + N1 * k : (i == 0 ? base + 3 * N2 + j + N1 * k : zero);
            const int sk = k == 0 ? base + 4 * N2 + i + N1 * j : (k == N1 - 1 ? base + 5 * N2 + i + N1 * j : zero);
            if (tface) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    LE[c] = (sF[sj][c] + sF[si][c]) + sF[sk][c];
                    LH[c] = (sF[sj][c + 3] + sF[si][c + 3]) + sF[sk][c + 3];
                }
            } else {  // the zero row's value, added as the gather does
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    LE[c] = (sF[sj][c] + sF[si][c]) + R(0);
                    LH[c] = (sF[sj][c + 3] + sF[si][c + 3]) + R(0);
                }
            }
        }

        // ---- inverse mass + LSRK45 update (dg_kernels.cu:353-408) ----
        // Stored by every thread: the spare slots of a last, partly filled
        // group repeat its last element with the same inputs and arithmetic,
        // so they store the same values to the same places (no branch: a
        // store on one path only would make the compiler's memory-counter
        // waits conservative and stall on the stores).
        // the last slot's i = n1-1 plane (pre-stage fields) for the next group's
        // face 3 (walking backwards: slot 0's i = 0 plane for its face 1); its
        // readers in this iteration are past the barrier above
        if (A.backwards ? (es == 0 && i == 0) : (es == EPW - 1 && i == N1 - 1)) {
#pragma unroll
            for (int c = 0; c < 6; ++c) sP3[j + N1 * k][c] = cur.u[c];
        }
        {
            const R sE_ = R(ER.ieps) * iJ, sH_ = R(ER.imu) * iJ;
            KR kn[6];
            R un[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const R rhs = c < 3 ? sE_ * (cH[c] + LE[c] * Winv) : sH_ * (-cE[c - 3] + LH[c - 3] * Winv);
                kn[c] = FIRST ? KR(dt) * KR(rhs) : KR(A.a) * kold[c] + KR(dt) * KR(rhs);
                un[c] = R(KR(sU[es][c][m]) + KR(A.b) * kn[c]);
            }
            // written through L2 (whole lines): the kernel boundary, which writes
            // every dirty L2 line back for the other XCDs, has k's lines off its list.
            // The last stage's k is dead (the next step's first stage has a = 0 and
            // does not read k): not stored.
            if (!A.last_stage) {
                if (A.stream_k)
                    stk<kNT>(rK, kv, ks, kn);
                else
                    stk<kSC1>(rK, kv, ks, kn);
            }
            st6(rO, nd * UB, un);
        }
        geometry((t + 1) & 1, geo);  // the next group's record, staged in this iteration
        cur = nxt;
        nb1 = nb2;
        rec1 = rec2;
        gA = gB;
        gB = gC;
        gC = gnext;
    }
}

// Elements per workgroup, three options per n1 (index 0 is the default),
// capped so one workgroup's LDS stays within 64 KiB (two or more workgroups per
// CU).  At n1 = 8 option 1 is instead the matrix-core volume for FP32 fields.
// An option whose shape repeats an earlier one is that option (fused_option_of).
constexpr int kEpw[kMaxN1 + 1][kNumTileOptions] = {
    {1, 1, 1}, {1, 1, 1}, {8, 32, 16}, {9, 4, 9}, {4, 1, 2}, {2, 1, 4}, {1, 3, 2}, {1, 2, 1}, {1, 1, 1}, {1, 1, 1},
    {1, 1, 1}};
constexpr bool matrix_volume(int n1, int opt, int rsz) { return n1 == 8 && opt == 1 && rsz == 4; }
// volume contraction mode of a launch: 0 vector ALUs, 1 FP32 matrix cores, 2 split BF16 matrix cores
constexpr int volume_mode(int n1, int opt, int rsz, bool bf16x3) {
    return bf16x3 ? 2 : (matrix_volume(n1, opt, rsz) ? 1 : 0);
}

constexpr int lds_per_elem(int n1, int rsz) { return (6 * n1 * n1 * n1 + 36 * n1 * n1) * rsz; }
constexpr int epw_of(int n1, int opt, int rsz) {
    const int want = kEpw[n1][opt];
    const int cap_lds = 65536 / lds_per_elem(n1, rsz);
    const int cap_thr = 1024 / (n1 * n1 * n1);
    const int cap = cap_lds < cap_thr ? cap_lds : cap_thr;
    return want < cap ? want : (cap > 0 ? cap : 1);
}
// The option that builds the same kernel as `opt` (itself unless it repeats an
// earlier option's elements per workgroup and volume mode).
constexpr int canonical_option(int n1, int opt, int rsz) {
    for (int o = 0; o < opt; ++o)
        if (epw_of(n1, o, rsz) == epw_of(n1, opt, rsz) && matrix_volume(n1, o, rsz) == matrix_volume(n1, opt, rsz))
            return o;
    return opt;
}

// Workgroups resident per CU for one kernel instance (both FIRST variants
// share the resource footprint), times the CU count: the persistent grid.
template <typename K>
int resident_groups(K kernel, int threads) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess) return 0;
    return (per_cu > 0 ? per_cu : 1) * cus;
}

// Workgroups of one kernel shape resident on the whole device at once.
template <int N1, int OPT, typename R, typename KR, bool BX = false>
int resident_of() {
    constexpr int EPW = BX ? 1 : epw_of(N1, OPT, sizeof(R));
    constexpr int MX = volume_mode(N1, OPT, sizeof(R), BX);
    constexpr int NT = EPW * N1 * N1 * N1;
    static int resident = 0;
    if (resident <= 0) resident = resident_groups(k_fused_stage<N1, EPW, false, R, KR, MX, false>, NT);
    return resident;
}

template <int N1, int OPT, typename R, typename KR, bool BX = false>
hipError_t launch_one(const FusedArgs& a, hipStream_t s) {
    constexpr int EPW = BX ? 1 : epw_of(N1, OPT, sizeof(R));
    constexpr int MX = volume_mode(N1, OPT, sizeof(R), BX);
    constexpr int NT = EPW * N1 * N1 * N1;
    if (resident_of<N1, OPT, R, KR, BX>() <= 0) return hipErrorInvalidValue;
    // every byte offset (component block included) of a field, register or
    // ghost slot fits in 32 bits
    const unsigned long long span = 6ull * (unsigned long long)(a.stride > a.kstride ? a.stride : a.kstride);
    if (span * sizeof(double) >= (1ull << 32)) return hipErrorInvalidValue;
    FusedArgs b = a;
    b.ngroups = (a.nelt + EPW - 1) / EPW;
    // the segment's k chunks (one per wave and group) lie inside the k buffer
    if (a.kbase < 0 || (a.kbase + (long long)b.ngroups * fused_k_chunks_per_group(NT)) * 64 > a.kstride)
        return hipErrorInvalidValue;
    if (!a.gorder || !a.ghead || a.nseq <= 0 || a.nseq > b.ngroups) return hipErrorInvalidValue;
    b.gstride = fused_grid(a.nseq);  // headers past the sequences return at once
    const dim3 grid(b.gstride), block(NT);
    if (a.affine) {
        if (a.first_stage)
            hipLaunchKernelGGL((k_fused_stage<N1, EPW, true, R, KR, MX, true>), grid, block, 0, s, b);
        else
            hipLaunchKernelGGL((k_fused_stage<N1, EPW, false, R, KR, MX, true>), grid, block, 0, s, b);
    } else {
        if (a.first_stage)
            hipLaunchKernelGGL((k_fused_stage<N1, EPW, true, R, KR, MX, false>), grid, block, 0, s, b);
        else
            hipLaunchKernelGGL((k_fused_stage<N1, EPW, false, R, KR, MX, false>), grid, block, 0, s, b);
    }
    return hipGetLastError();
}

using LaunchFn = hipError_t (*)(const FusedArgs&, hipStream_t);
using ResidentFn = int (*)();

template <int N1, typename R, typename KR>
ResidentFn pick_resident(int option) {
    constexpr int O2 = canonical_option(N1, 2, sizeof(R));
    return option == 1 ? &resident_of<N1, 1, R, KR> : (option == 2 ? &resident_of<N1, O2, R, KR> : &resident_of<N1, 0, R, KR>);
}

template <typename R, typename KR>
ResidentFn pick_resident_n1(int n1, int option) {
    switch (n1) {
        case 2: return pick_resident<2, R, KR>(option);
        case 3: return pick_resident<3, R, KR>(option);
        case 4: return pick_resident<4, R, KR>(option);
        case 5: return pick_resident<5, R, KR>(option);
        case 6: return pick_resident<6, R, KR>(option);
        case 7: return pick_resident<7, R, KR>(option);
        case 8: return pick_resident<8, R, KR>(option);
        case 9: return pick_resident<9, R, KR>(option);
        default: return nullptr;
    }
}

template <int N1, typename R, typename KR>
LaunchFn pick(int option) {
    constexpr int O2 = canonical_option(N1, 2, sizeof(R));
    return option == 1 ? &launch_one<N1, 1, R, KR> : (option == 2 ? &launch_one<N1, O2, R, KR> : &launch_one<N1, 0, R, KR>);
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

static int clamp_option(int option) { return option < 0 ? 0 : (option >= kNumTileOptions ? kNumTileOptions - 1 : option); }

int fused_epw(int n1, int option, int prec) {
    if (n1 < 0 || n1 > kMaxN1) return 0;
    if (prec == kBF16X3) return 1;
    return epw_of(n1, clamp_option(option), prec == kFP64 ? 8 : 4);
}

bool fused_matrix_volume(int n1, int option, int prec) {
    return prec == kBF16X3 || matrix_volume(n1, clamp_option(option), prec == kFP64 ? 8 : 4);
}

int fused_option_of(int n1, int option, int prec) {
    if (n1 < 0 || n1 > kMaxN1 || prec == kBF16X3) return 0;
    return canonical_option(n1, clamp_option(option), prec == kFP64 ? 8 : 4);
}

bool fused_beyond_infinity_cache(long long ws_bytes) { return ws_bytes > (256ll << 20); }

bool fused_alternate_walk(long long ws_bytes) {
    return fused_beyond_infinity_cache(ws_bytes) && ws_bytes <= 8 * (256ll << 20);
}

bool fused_stream_k(long long ws_bytes, int threads, int ksize) {
    return fused_beyond_infinity_cache(ws_bytes) && ((threads % 64) * ksize) % 128 == 0;
}

long long fused_k_chunks(int n1, int option, int prec, int nelt) {
    const int epw = fused_epw(n1, option, prec);
    if (epw <= 0) return 0;
    return (long long)((nelt + epw - 1) / epw) * fused_k_chunks_per_group(epw * n1 * n1 * n1);
}

bool fused_supported(int n1, int prec) {
    if (prec == kBF16X3) return n1 == 8;  // the 4x4x4 BF16 tiles cover 8-point lines
    return n1 >= 2 && n1 <= 9 && prec >= 0 && prec < kNumPrec;
}

int fused_resident(int n1, int prec, int option, int max_per_cu) {
    ResidentFn fn = nullptr;
    if (prec == kFP64) fn = pick_resident_n1<double, double>(n1, option);
    else if (prec == kFP32) fn = pick_resident_n1<float, float>(n1, option);
    else if (prec == kFP32K64) fn = pick_resident_n1<float, double>(n1, option);
    else if (prec == kBF16X3 && n1 == 8) fn = &resident_of<8, 0, float, float, true>;
    const int all = fn ? fn() : 0;
    if (all <= 0 || max_per_cu <= 0) return all;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) return 0;
    return std::min(all / cus, max_per_cu) * cus;
}

hipError_t launch_fused(int n1, int prec, int option, const FusedArgs& a, hipStream_t s) {
    LaunchFn fn = nullptr;
    if (prec == kFP64) fn = pick_n1<double, double>(n1, option);
    else if (prec == kFP32) fn = pick_n1<float, float>(n1, option);
    else if (prec == kFP32K64) fn = pick_n1<float, double>(n1, option);
    else if (prec == kBF16X3 && n1 == 8) fn = &launch_one<8, 0, float, float, true>;  // one shape: both options
    if (!fn) return hipErrorInvalidValue;
    if (a.nelt <= 0) return hipSuccess;
    return fn(a, s);
}

hipError_t launch_affine_records(int prec, double* rec, int nelt, hipStream_t s) {
    if (nelt <= 0) return hipSuccess;
    const dim3 grid((nelt + 255) / 256), block(256);
    if (prec == kFP64)
        hipLaunchKernelGGL(k_affine_records<double>, grid, block, 0, s, rec, nelt);
    else
        hipLaunchKernelGGL(k_affine_records<float>, grid, block, 0, s, rec, nelt);
    return hipGetLastError();
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
