// nw_fused.hpp -- launch interface of the fused per-element LSRK45 stage kernel.
#pragma once

#include "nw_internal.hpp"

namespace nw {

// kBF16X3: FP32 fields and registers, volume contractions as split BF16
// products on the matrix cores (n1 = 8 only).
enum Prec { kFP64 = 0, kFP32 = 1, kFP32K64 = 2, kBF16X3 = 3, kNumPrec = 4 };

// Everything one launch of the fused stage kernel needs.  One launch covers a
// contiguous range of elements that share n1 (a whole mesh, or one n1 segment
// of a batched set of independent meshes).
// Words of a workgroup's sequence header (FusedArgs::ghead): start s0 in the
// group order, last iteration tl (-1: empty), the groups of iterations 0, 1, 2
// (the last one repeated past tl), padding.
constexpr int kFusedHead = 8;
// The persistent grid of nseq sequences (a multiple of the 8 XCDs), and the
// sequence hardware workgroup hb runs: the hardware deals workgroups
// round-robin over the XCDs, so sequences W .. W + nseq/8 go to one XCD
// (>= nseq: none).  The headers are stored in hardware workgroup order.
constexpr int fused_grid(int nseq) { return 8 * ((nseq + 7) / 8); }
constexpr int fused_sequence_of_block(int hb, int nseq) { return (hb & 7) * ((nseq + 7) / 8) + (hb >> 3); }

struct FusedArgs {
    const void* uin;          // node-major: the 6 values (R) of each node consecutive; ghosts after npts
    void* uout;               // same layout
    void* kbuf;               // LSRK register (KR): per wave of a group, 6 component blocks of 64 lanes
    long long stride;         // npts_total + nghost (ghost slots after the local nodes)
    long long kstride;        // k values per component in the whole buffer (64 per chunk)
    long long kbase;          // first k chunk of this segment (fused_k_chunks)
    long long node_base;      // first node of this segment
    long long ghost_base;     // first ghost slot (node index) of this segment
    int nelt;                 // elements in this segment
    int elem_base;            // global index of the segment's first element (into mat)
    int ngroups;              // element groups (EPW elements each); set by the launcher
    int gstride;              // workgroups of the persistent grid; set by the launcher
    const int* gorder;        // ngroups entries: the groups in visiting order (a permutation)
    const int* ghead;         // kFusedHead words per hardware workgroup (fused_grid(nseq) of them):
                              // s0, tl, the first three groups; it visits gorder[s0 .. s0 + tl]
    int nseq;                 // workgroup sequences (<= the resident workgroups, fused_resident)
    const ElemGeom* geom;     // per element, segment-local
    const double* rec;        // per element, segment-local: 24 trilinear coefficients, 1/eps, 1/mu
    const ElemMat* mat;       // per element, global index (always present; vacuum = 1)
    const ElemMat* gmat;      // per ghost face
    const int4* fnb;          // per element face (nelt*6): neighbour addressing, see FaceNbr
    const double* dt;         // device scalar
    double a, b, c0;
    int first_stage;          // a == 0: k is not read
    int last_stage;           // k is not stored: the next step's first stage (a == 0) does not read it
    int backwards;            // gorder/ghead hold the backward walk: x-chains are walked towards -x,
                              // the carried x face is face 1 of the last slot
    const void* fw;           // per element face (nelt*6, segment-local), 4 R: flux weights
                              // (wEh, wEe, wHe, wHh), see face_flux_weights
    int affine;               // every element affine: records hold the geometry (launch_affine_records)
    int stream_k;             // k accesses non-temporal (set by the caller: fused_stream_k)
};

// Whether the LSRK register accesses should be non-temporal, by default (the
// autotuner times both): where the stage's working set (u in, u out, k:
// `ws_bytes`) is larger than the 256 MiB Infinity Cache, which a non-temporal
// stream would bypass, and where the last wave of a group of `threads` threads
// also covers whole cache lines of k (elements `ksize` bytes): a partly written
// line streamed out is written again when its other part is.
bool fused_stream_k(long long ws_bytes, int threads, int ksize);
// Whether a stage's working set is larger than the 256 MiB Infinity Cache.
bool fused_beyond_infinity_cache(long long ws_bytes);
// Whether stages alternate the walk's direction: a stage that walks back over
// the groups the stage before it wrote last finds some of them still in the
// Infinity Cache -- worth it where the working set exceeds that cache but the
// cache still holds a fair share (an eighth) of it.
bool fused_alternate_walk(long long ws_bytes);
// k chunks (6 x 64 values each) of one group of `threads` threads, and of a
// segment of `nelt` elements under a tile option: a step's stages all use one
// option, so the k buffer holds, per segment, the most any option needs.
constexpr long long fused_k_chunks_per_group(int threads) { return (threads + 63) / 64; }
long long fused_k_chunks(int n1, int option, int prec, int nelt);

// Tile options per n1: elements per workgroup (and at n1 = 8 in FP32, option 1,
// the matrix-core volume).
constexpr int kNumTileOptions = 3;
int fused_epw(int n1, int option, int prec);
// The lowest option that builds the same kernel as `option` (options that
// repeat a shape are not timed twice by the tuner).
int fused_option_of(int n1, int option, int prec);
// Workgroups of that shape resident on the current device at once (0: not built).
int fused_resident(int n1, int prec, int option, int max_per_cu = 0);
// Whether tile option `option` computes the volume contractions on the matrix cores.
bool fused_matrix_volume(int n1, int option, int prec);
bool fused_supported(int n1, int prec);
hipError_t launch_fused(int n1, int prec, int option, const FusedArgs& a, hipStream_t s);
int fused_upload_tables(const GllTables& t);
// Rewrites the records (26 doubles per element) of an affine segment so their
// first ten words hold the element's geometry as the fused kernel computes it
// per node (cofactors, 1/J in the precision's R); the kernel then reads them
// (FusedArgs::affine).
hipError_t launch_affine_records(int prec, double* rec, int nelt, hipStream_t s);

// Face orientation: (p, q) on a face -> (p', q') on the neighbour face.
// bit0: swap p,q; bit1: p' = n1-1-p'; bit2: q' = n1-1-q'.
__host__ __device__ inline void orient_map(int o, int n1, int p, int q, int& p2, int& q2) {
    int a = (o & 1) ? q : p, b = (o & 1) ? p : q;
    if (o & 2) a = n1 - 1 - a;
    if (o & 4) b = n1 - 1 - b;
    p2 = a;
    q2 = b;
}

// getFaceNodeVolIndex (mesh.cpp:955-981) as local (i, j, k).
__host__ __device__ inline void face_ijk(int f, int n1, int p, int q, int& i, int& j, int& k) {
    switch (f) {
        case 0: i = p; j = 0; k = q; break;
        case 1: i = n1 - 1; j = p; k = q; break;
        case 2: i = p; j = n1 - 1; k = q; break;
        case 3: i = 0; j = p; k = q; break;
        case 4: i = p; j = q; k = 0; break;
        default: i = p; j = q; k = n1 - 1; break;
    }
}

// Per element face: the plus-side node of face point (p, q) is
//   x + p * y + q * z   (x < 0: mirror face, no plus side),
// and w names the plus side's medium: >= 0 element index, < 0 ghost face -(w+1).
int make_face_nbr(const int* fcode, int nelt, int n1, long long node_base, long long ghost_base, int elem_base,
                  int4* out);

// Flux weights of one face for the minus/plus media (Z = sqrt(mu/eps), Y = 1/Z):
//   F_E = wEh n x [[H]] - wEe n x n x [[E]],  F_H = -wHe n x [[E]] - wHh n x n x [[H]]
// with wEh = Z+/(Z+ + Z-), wEe = c0/(Z+ + Z-), wHe = Y+/(Y+ + Y-), wHh = c0/(Y+ + Y-);
// without media the reference's 1/2, c0/2 (dg_kernels.cu:265-294).
inline void face_flux_weights(const ElemMat& mm, const ElemMat& mp, double c0, bool media, double w[4]) {
    if (media) {
        const double iz = 1.0 / (mp.Z + mm.Z), iy = 1.0 / (mp.Y + mm.Y);
        w[0] = mp.Z * iz;
        w[1] = c0 * iz;
        w[2] = mp.Y * iy;
        w[3] = c0 * iy;
    } else {
        w[0] = 0.5;
        w[1] = 0.5 * c0;
        w[2] = 0.5;
        w[3] = 0.5 * c0;
    }
}

// fcode encoding: >= 0 : nbr_elem * 64 + nbr_face * 8 + orient; -1 : mirror; <= -2 : ghost face -(c+2)
__host__ __device__ inline int fcode_pack(int e, int f, int o) { return e * 64 + f * 8 + o; }

}  // namespace nw
