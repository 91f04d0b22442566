// nw_solver.cpp -- the DgSolver replacement: device residency, stage
// sequencing, graphs, diagnostics (dg_solver.cpp:40-582 is the reference),
// for one mesh or a batch of independent meshes of mixed n1.
#include "nw_solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "nw_fused.hpp"
#include "nw_plain.hpp"

namespace nw {

namespace {
std::mutex g_tables_mu;
bool g_tables_done[64] = {false};
}  // namespace

int upload_gll_tables(int device) {
    std::lock_guard<std::mutex> lk(g_tables_mu);
    if (device < 0 || device >= 64) return NW_ERR_ARG;
    if (g_tables_done[device]) return NW_OK;
    const GllTables& t = host_gll_tables();
    int st = fused_upload_tables(t);
    if (st != NW_OK) return st;
    st = plain_upload_tables(t);
    if (st != NW_OK) return st;
    g_tables_done[device] = true;
    return NW_OK;
}

// LSRK45 coefficients, Carpenter & Kennedy (1994) (dg_solver.cpp:58-69).
constexpr int kStages = 5;  // LSRK45 stages per step
static const double kRkA[kStages] = {0.0, -567301805773.0 / 1357537059087.0,
                                     -2404267990393.0 / 2016746695238.0, -3550918686646.0 / 2091501179385.0,
                                     -1275806237668.0 / 842570457699.0};
static const double kRkB[kStages] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                                     1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                                     2277821191437.0 / 14882151754819.0};

// Visiting order of the element groups (EPW consecutive elements each) of one
// fused-kernel launch, as `nseq` workgroup sequences: workgroup W visits
// order[seq[W] .. seq[W+1]) one group per iteration.
//
// Chains: runs of groups g, g+1, ... where the last element of g meets the
// first element of g+1 through its face 1 in the same orientation (the x-lines
// of a box mesh).  Along a chain the kernel takes the face-3 traces of a group
// from the plane its predecessor left in LDS instead of gathering them, so when
// the mesh has chains every workgroup walks along them: equal chains are cut
// into equal segments (as many per chain as makes the busiest workgroup's share
// smallest: the balanced share, in the longest segments that reach it) dealt
// round-robin in Morton (Z-order) order of their first group's centroid;
// chains of different lengths (a batch: storage order) are concatenated and cut
// into nseq pieces of nearly equal length, neighbouring pieces on one XCD.  Without chains (single groups), the grid's iteration t
// covers positions t nseq .. t nseq + nseq - 1 of the Morton curve, a compact
// block of elements whose face neighbours are read while still in L2.
static bool uniform_chains(const std::vector<int>& cstart, int nchain) {
    for (int c = 1; c < nchain; ++c)
        if (cstart[c + 1] - cstart[c] != cstart[1] - cstart[0]) return false;
    return nchain > 0;
}

static void append_walk(const Mesh* m, const int4* fnb, int n1, int epw, int nelt, long long node_base, int nseq,
                        std::vector<int>& order, std::vector<int>& seq) {
    const int ng = (nelt + epw - 1) / epw;
    const long long np = (long long)n1 * n1 * n1;
    nseq = std::max(1, std::min(nseq, ng));
    const int obase = static_cast<int>(order.size());
    auto first = [&](int g) { return g * epw; };
    auto last = [&](int g) { return std::min(g * epw + epw - 1, nelt - 1); };
    // chains
    std::vector<int> cstart;
    for (int g = 0; g < ng; ++g) {
        bool joins = false;
        if (g > 0) {
            const int4 b = fnb[size_t(last(g - 1)) * 6 + 1];
            joins = b.y == n1 && b.z == n1 * n1 && (long long)b.x == node_base + first(g) * np;
        }
        if (!joins) cstart.push_back(g);
    }
    const int nchain = static_cast<int>(cstart.size());
    cstart.push_back(ng);
    // Morton key of a group's first element centroid (single mesh)
    std::vector<uint64_t> key;
    if (m) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int g = 0; g < ng; ++g)
            for (int d = 0; d < 3; ++d) {
                const double x = m->geom[size_t(first(g))].c[d][0];
                lo[d] = std::min(lo[d], x);
                hi[d] = std::max(hi[d], x);
            }
        auto spread = [](uint64_t v) {  // 10 bits -> every third bit
            v &= 0x3ff;
            v = (v | (v << 16)) & 0x030000ffull;
            v = (v | (v << 8)) & 0x0300f00full;
            v = (v | (v << 4)) & 0x030c30c3ull;
            v = (v | (v << 2)) & 0x09249249ull;
            return v;
        };
        key.resize(size_t(ng));
        for (int g = 0; g < ng; ++g) {
            uint64_t code = 0;
            for (int d = 0; d < 3; ++d) {
                const double span = hi[d] - lo[d];
                const double x = span > 0 ? (m->geom[size_t(first(g))].c[d][0] - lo[d]) / span : 0.0;
                const uint64_t q = static_cast<uint64_t>(std::min(1023.0, std::max(0.0, x * 1024.0)));
                code |= spread(q) << d;
            }
            key[size_t(g)] = code;
        }
    }
    // equal chains cut into d pieces each (d dividing the chain length): the
    // busiest workgroup walks ceil(nchain d / nseq) pieces of lc / d groups;
    // the d with the fewest groups on the busiest workgroup (the longest pieces
    // among equals) -- a launch lasts as long as its busiest workgroup.  d = lc
    // always reaches the balanced share ceil(ng / nseq), so the choice never
    // leaves a workgroup with more than that.
    int mseg = 0;
    const bool uniform = 2 * nchain <= ng && uniform_chains(cstart, nchain);
    if (uniform) {
        const int lc = cstart[1] - cstart[0];
        long long best = -1;
        for (int d = 1; d <= lc; ++d) {
            if (lc % d) continue;
            const long long load = ((long long)nchain * d + nseq - 1) / nseq * (lc / d);
            if (best < 0 || load < best) {
                best = load;
                mseg = d;
            }
        }
    }
    if (2 * nchain > ng) {
        // (almost) no chains: round-robin positions of the Morton curve (storage order for a batch)
        std::vector<int> curve(static_cast<size_t>(ng));
        std::iota(curve.begin(), curve.end(), 0);
        if (m) std::stable_sort(curve.begin(), curve.end(), [&](int a, int b) { return key[size_t(a)] < key[size_t(b)]; });
        for (int W = 0; W < nseq; ++W) {
            seq.push_back(static_cast<int>(order.size()) - obase);
            for (long long p = W; p < ng; p += nseq) order.push_back(curve[size_t(p)]);
        }
    } else if (uniform) {
        // equal chains (the x-lines of a box mesh): each cut into mseg equal
        // segments (chosen above); segments in Morton order dealt round-robin, so
        // in every round the workgroups of one XCD walk neighbouring segments side
        // by side and the y and z neighbours of a group are visited at the same time
        const int lc = cstart[1] - cstart[0];
        const int ls = lc / mseg;
        std::vector<int> sgs;  // first group of each segment
        for (int c = 0; c < nchain; ++c)
            for (int x = 0; x < mseg; ++x) sgs.push_back(cstart[c] + x * ls);
        if (m) std::stable_sort(sgs.begin(), sgs.end(), [&](int a, int b) { return key[size_t(a)] < key[size_t(b)]; });
        const int nsg = static_cast<int>(sgs.size());
        nseq = std::min(nseq, nsg);
        for (int W = 0; W < nseq; ++W) {
            seq.push_back(static_cast<int>(order.size()) - obase);
            for (int x = W; x < nsg; x += nseq)
                for (int g = sgs[size_t(x)]; g < sgs[size_t(x)] + ls; ++g) order.push_back(g);
        }
    } else {
        std::vector<int> co(static_cast<size_t>(nchain));
        std::iota(co.begin(), co.end(), 0);
        if (m)
            std::stable_sort(co.begin(), co.end(),
                             [&](int a, int b) { return key[size_t(cstart[a])] < key[size_t(cstart[b])]; });
        std::vector<int> stream;
        stream.reserve(static_cast<size_t>(ng));
        for (int c : co)
            for (int g = cstart[c]; g < cstart[c + 1]; ++g) stream.push_back(g);
        const int base = ng / nseq, extra = ng % nseq;
        size_t at = 0;
        for (int W = 0; W < nseq; ++W) {
            seq.push_back(static_cast<int>(order.size()) - obase);
            const int len = base + (W < extra ? 1 : 0);
            order.insert(order.end(), stream.begin() + at, stream.begin() + at + len);
            at += static_cast<size_t>(len);
        }
    }
    seq.push_back(static_cast<int>(order.size()) - obase);
}

unsigned long long mesh_signature(const Mesh& m) {
    unsigned long long h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t x = 0; x < n; ++x) h = (h ^ b[x]) * 1099511628211ull;
    };
    mix(&m.n1, sizeof(m.n1));
    mix(&m.nelt, sizeof(m.nelt));
    mix(m.corners.data(), m.corners.size() * sizeof(double));
    mix(m.fcode.data(), m.fcode.size() * sizeof(int));
    return h;
}

Solver::~Solver() { release(); }

void Solver::release() {
    if (device < 0) return;
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    ev0 = ev1 = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    device = -1;
}

int Solver::dalloc(void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(" + std::to_string(bytes) + " B) failed: " + hipGetErrorString(e));
        *p = nullptr;
        return NW_ERR_ALLOC;
    }
    allocs.push_back(*p);
    device_bytes += (long long)bytes;
    return NW_OK;
}

// Frees one buffer obtained from dalloc (no-op for null).
void Solver::dfree(void* p) {
    if (!p) return;
    for (size_t i = 0; i < allocs.size(); ++i)
        if (allocs[i] == p) {
            allocs.erase(allocs.begin() + long(i));
            (void)hipFree(p);
            return;
        }
}

template <typename T>
int Solver::upload_vec(T** dst, const std::vector<T>& v) {
    int st = dalloc(reinterpret_cast<void**>(dst), v.size() * sizeof(T));
    if (st) return st;
    if (!v.empty()) NW_TRY_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return NW_OK;
}

int Solver::init(const Mesh* const* meshes, int count, double c0_, const nw_opts& o) {
    NW_CHECK_ARG(meshes && count >= 1, "need at least one mesh");
    NW_CHECK_ARG(o.precision >= 0 && o.precision < kNumPrec, "bad precision");
    NW_CHECK_ARG(o.pipeline == NW_PIPE_PLAIN || o.pipeline == NW_PIPE_FUSED, "bad pipeline");
    if (o.pipeline == NW_PIPE_PLAIN && o.precision != NW_PREC_FP64) {
        set_error("the plain (reference-decomposed) pipeline is FP64 only");
        return NW_ERR_UNSUPPORTED;
    }
    if (o.pipeline == NW_PIPE_PLAIN && count != 1) {
        set_error("the plain pipeline runs one mesh");
        return NW_ERR_UNSUPPORTED;
    }
    for (int x = 0; x < count; ++x) {
        NW_CHECK_ARG(meshes[x] != nullptr, "null mesh in batch");
        if (o.pipeline == NW_PIPE_FUSED && !fused_supported(meshes[x]->n1, o.precision)) {
            set_error("fused pipeline supports n1 in [2, 9]");
            return NW_ERR_UNSUPPORTED;
        }
        if (count > 1 && meshes[x]->nhalo_points() > 0) {
            set_error("the halo stand-in runs a single mesh");
            return NW_ERR_UNSUPPORTED;
        }
    }
    int ndev = 0;
    NW_TRY_HIP(hipGetDeviceCount(&ndev));
    NW_CHECK_ARG(o.device >= 0 && o.device < ndev, "bad device ordinal");
    device = o.device;
    NW_TRY_HIP(hipSetDevice(device));
    int st = upload_gll_tables(device);
    if (st) return st;
    prec = o.precision;
    pipe = o.pipeline;
    use_graph = o.use_graph != 0;
    NW_CHECK_ARG(o.groups_per_cu >= 0, "groups_per_cu must be >= 0");
    groups_per_cu = o.groups_per_cu;
    autotune = (o.tile_option < 0) && pipe == NW_PIPE_FUSED;
    const int opt = o.tile_option < 0 ? 0 : std::min(o.tile_option, kNumTileOptions - 1);
    c0 = c0_;
    NW_TRY_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    NW_TRY_HIP(hipEventCreate(&ev0));
    NW_TRY_HIP(hipEventCreate(&ev1));

    // layout: instances ordered by n1 so each n1 forms one contiguous segment
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return meshes[a]->n1 < meshes[b]->n1; });
    inst.assign(count, Instance{});
    long long nodes = 0;
    int elems = 0;
    media = false;
    for (int r : order) {
        const Mesh& m = *meshes[r];
        Instance& I = inst[r];
        I.n1 = m.n1; I.np = m.np; I.nelt = m.nelt; I.elem_base = elems; I.node_base = nodes; I.npts = m.npts();
        I.media = m.media;
        media = media || m.media;
        elems += m.nelt;
        nodes += m.npts();
        if (seg.empty() || seg.back().n1 != m.n1) {
            Segment sg;
            sg.n1 = m.n1;
            sg.elem_base = I.elem_base;
            sg.node_base = I.node_base;
            sg.option = opt;
            sg.affine = true;
            seg.push_back(sg);
        }
        seg.back().nelt += m.nelt;
        seg.back().affine = seg.back().affine && m.all_affine;
    }
    nelt = elems;
    npts = nodes;
    mesh_sig = count == 1 ? mesh_signature(*meshes[0]) : 0;
    n1 = count == 1 ? meshes[0]->n1 : 0;
    np = count == 1 ? meshes[0]->np : 0;
    tfp = count == 1 ? meshes[0]->tfp() : 0;
    nhalo = count == 1 ? meshes[0]->nhalo_points() : 0;
    stride = npts + nhalo;

    const size_t rsz = (prec == kFP64) ? 8 : 4;
    const size_t ksz = (prec == kFP32 || prec == kBF16X3) ? 4 : 8;
    if ((st = dalloc(reinterpret_cast<void**>(&d_dt), 8))) return st;
    if ((st = dalloc(reinterpret_cast<void**>(&d_diag_part), sizeof(double) * 3 * diag_blocks()))) return st;
    if ((st = dalloc(reinterpret_cast<void**>(&d_diag_out), sizeof(double) * 3))) return st;
    long long max_inst = 0;
    for (const Instance& I : inst) max_inst = std::max(max_inst, I.npts);
    if ((st = dalloc(reinterpret_cast<void**>(&d_stage), sizeof(double) * 6 * max_inst))) return st;
    {
        std::vector<ElemMat> mat(nelt);
        std::vector<ElemGeom> geom(nelt);
        for (int r = 0; r < count; ++r) {
            std::copy(meshes[r]->mat.begin(), meshes[r]->mat.end(), mat.begin() + inst[r].elem_base);
            std::copy(meshes[r]->geom.begin(), meshes[r]->geom.end(), geom.begin() + inst[r].elem_base);
        }
        if ((st = upload_vec(&d_mat, mat))) return st;
        if ((st = upload_vec(&d_geom, geom))) return st;
        std::vector<double> rec(size_t(nelt) * 26);
        for (int e = 0; e < nelt; ++e) {
            std::copy(&geom[e].c[0][0], &geom[e].c[0][0] + 24, &rec[size_t(e) * 26]);
            rec[size_t(e) * 26 + 24] = mat[e].ieps;
            rec[size_t(e) * 26 + 25] = mat[e].imu;
        }
        if ((st = upload_vec(&d_rec, rec))) return st;
        // NW_GENERAL_GEOMETRY=1 keeps the per-node geometry on affine meshes too
        // (the two give the same bits; tests compare them)
        const char* gen = std::getenv("NW_GENERAL_GEOMETRY");
        const bool per_node = gen && gen[0] == '1';
        if (pipe == NW_PIPE_FUSED)
            for (Segment& sg : seg) {
                if (per_node) sg.affine = false;
                if (sg.affine)
                    NW_TRY_HIP(launch_affine_records(prec, d_rec + size_t(sg.elem_base) * 26, sg.nelt, stream));
            }
    }

    // halo stand-in metadata (dg_solver.cpp:131-170 with one neighbour: self)
    if (nhalo > 0) {
        const Mesh& m = *meshes[0];
        const int nh = static_cast<int>(nhalo);
        std::vector<int> off(nh), sizes(nh, nh);
        for (int j = 0; j < nh; ++j) off[j] = j;
        if ((st = upload_vec(&d_send_vol, m.send_vol))) return st;
        if ((st = upload_vec(&d_off, off))) return st;
        if ((st = upload_vec(&d_sizes, sizes))) return st;
        if ((st = upload_vec(&d_gmat, m.gmat))) return st;
        // element owning each ghost slot (the plain flux kernel looks media up per point)
        std::vector<int> ge(nh);
        const int n2 = m.n1 * m.n1;
        for (int j = 0; j < nh; ++j) ge[j] = m.ghost_elem[j / n2];
        if ((st = upload_vec(&d_ghost_elem, ge))) return st;
        if ((st = dalloc(&d_sendbuf, rsz * 6 * nh))) return st;
        if ((st = dalloc(&d_recvbuf, rsz * 6 * nh))) return st;
    }

    if (pipe == NW_PIPE_PLAIN) {
        st = init_plain(*meshes[0]);
        if (st) return st;
    } else {
        if ((st = dalloc(&d_u[0], rsz * 6 * stride))) return st;
        if ((st = dalloc(&d_u[1], rsz * 6 * stride))) return st;
        kchunks = 0;
        for (Segment& sg : seg) {
            sg.kbase = kchunks;
            long long most = 0;
            for (int opt = 0; opt < kNumTileOptions; ++opt)
                most = std::max(most, fused_k_chunks(sg.n1, opt, prec, sg.nelt));
            kchunks += most;
        }
        if ((st = dalloc(&d_k, ksz * 384 * kchunks))) return st;
        NW_TRY_HIP(hipMemsetAsync(d_u[0], 0, rsz * 6 * stride, stream));
        NW_TRY_HIP(hipMemsetAsync(d_u[1], 0, rsz * 6 * stride, stream));
        NW_TRY_HIP(hipMemsetAsync(d_k, 0, ksz * 384 * kchunks, stream));
        std::vector<int4> fnb(size_t(nelt) * 6);
        for (int r = 0; r < count; ++r) {
            const Instance& I = inst[r];
            if (make_face_nbr(meshes[r]->fcode.data(), I.nelt, I.n1, I.node_base, npts, I.elem_base,
                              fnb.data() + size_t(I.elem_base) * 6)) {
                set_error("node index beyond 2^31");
                return NW_ERR_UNSUPPORTED;
            }
        }
        if ((st = upload_vec(&d_fnb, fnb))) return st;
        std::vector<int> gorder, ghead, gorder_rev, ghead_rev;
        for (int opt = 0; opt < kNumTileOptions; ++opt)
            for (Segment& sg : seg) {
                const int epw = fused_epw(sg.n1, opt, prec);
                const int resident = fused_resident(sg.n1, prec, opt, groups_per_cu);
                if (resident <= 0) {
                    set_error("no resident workgroups for the fused kernel");
                    return NW_ERR_HIP;
                }
                sg.goff[opt] = static_cast<long long>(gorder.size());
                sg.soff[opt] = static_cast<long long>(ghead.size());
                std::vector<int> seq;
                append_walk(count == 1 ? meshes[0] : nullptr, fnb.data() + size_t(sg.elem_base) * 6, sg.n1, epw,
                            sg.nelt, sg.node_base, resident, gorder, seq);
                sg.nseq[opt] = static_cast<int>(seq.size()) - 1;
                // per hardware workgroup of the grid, the header of the sequence
                // it runs: start, last iteration and first three groups
                // (kFusedHead words; tl = -1: none), one scalar read for the kernel
                const size_t go = static_cast<size_t>(sg.goff[opt]);
                const int ns = sg.nseq[opt];
                // the backward walk: every sequence reversed, so a stage starts on the
                // groups the stage before it (walking forwards) wrote last
                gorder_rev.resize(gorder.size());
                for (int W = 0; W < ns; ++W)
                    for (int p = seq[size_t(W)], q = seq[size_t(W) + 1] - 1; q >= seq[size_t(W)]; ++p, --q)
                        gorder_rev[go + size_t(p)] = gorder[go + size_t(q)];
                for (int hb = 0; hb < fused_grid(ns); ++hb) {
                    const int W = fused_sequence_of_block(hb, ns);
                    int head[kFusedHead] = {0, -1, 0, 0, 0, 0, 0, 0};
                    int hrev[kFusedHead] = {0, -1, 0, 0, 0, 0, 0, 0};
                    if (W < ns) {
                        const int s0 = seq[size_t(W)], tl = seq[size_t(W) + 1] - s0 - 1;
                        head[0] = hrev[0] = s0;
                        head[1] = hrev[1] = tl;
                        for (int t = 0; t < 3 && tl >= 0; ++t) {
                            head[2 + t] = gorder[go + size_t(s0 + std::min(t, tl))];
                            hrev[2 + t] = gorder_rev[go + size_t(s0 + std::min(t, tl))];
                        }
                    }
                    ghead.insert(ghead.end(), head, head + kFusedHead);
                    ghead_rev.insert(ghead_rev.end(), hrev, hrev + kFusedHead);
                }
            }
        if ((st = upload_vec(&d_gorder, gorder))) return st;
        if ((st = upload_vec(&d_ghead, ghead))) return st;
        if ((st = upload_vec(&d_gorder_rev, gorder_rev))) return st;
        if ((st = upload_vec(&d_ghead_rev, ghead_rev))) return st;
        // flux weights per element face from the two media (the plus side of a
        // mirror face is the element itself; of a cut face, its ghost medium)
        std::vector<double> w64(size_t(nelt) * 24);
        for (int r = 0; r < count; ++r) {
            const Instance& I = inst[r];
            const Mesh& m = *meshes[r];
            for (int e = 0; e < I.nelt; ++e)
                for (int f = 0; f < 6; ++f) {
                    const int code = m.fcode[size_t(e) * 6 + f];
                    const ElemMat& mm = m.mat[e];
                    const ElemMat& mp = code >= 0 ? m.mat[code >> 6] : (code == -1 ? mm : m.gmat[-(code + 2)]);
                    face_flux_weights(mm, mp, c0, media, &w64[(size_t(I.elem_base + e) * 6 + f) * 4]);
                }
        }
        if (rsz == 8) {
            if ((st = upload_vec(reinterpret_cast<double**>(&d_fw), w64))) return st;
        } else {
            std::vector<float> w32(w64.begin(), w64.end());
            if ((st = upload_vec(reinterpret_cast<float**>(&d_fw), w32))) return st;
        }
        kernels_per_step = 5 * (static_cast<int>(seg.size()) + (nhalo > 0 ? 2 : 0));
    }
    NW_TRY_HIP(hipStreamSynchronize(stream));
    initialized = true;
    return NW_OK;
}

// Reference residency (dg_solver.cpp:114-236): state, k, rhs, D, w3, metrics,
// face arrays, face scratch -- with the geometry generated on the device.
int Solver::init_plain(const Mesh& m) {
    int st;
    if ((st = dalloc(reinterpret_cast<void**>(&d_state), 8 * 6 * stride))) return st;
    if ((st = dalloc(reinterpret_cast<void**>(&d_kk), 8 * 6 * npts))) return st;
    if ((st = dalloc(reinterpret_cast<void**>(&d_rhs), 8 * 6 * npts))) return st;
    NW_TRY_HIP(hipMemsetAsync(d_state, 0, 8 * 6 * stride, stream));
    NW_TRY_HIP(hipMemsetAsync(d_kk, 0, 8 * 6 * npts, stream));
    std::vector<double> Dcol(n1 * n1), w3(np);
    const GllTables& T = host_gll_tables();
    for (int i = 0; i < n1; ++i)
        for (int l = 0; l < n1; ++l) Dcol[i + l * n1] = T.D[n1][i * n1 + l];
    for (int k = 0; k < n1; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n1; ++i) w3[i + n1 * (j + n1 * k)] = T.w[n1][i] * T.w[n1][j] * T.w[n1][k];
    if ((st = upload_vec(&d_D, Dcol))) return st;
    if ((st = upload_vec(&d_w3, w3))) return st;
    for (int a = 0; a < 13; ++a)
        if ((st = dalloc(reinterpret_cast<void**>(&d_met[a]), 8 * npts))) return st;
    double* corners = nullptr;
    int* fcode = nullptr;
    int* fbc = nullptr;
    if ((st = upload_vec(&corners, m.corners))) return st;
    if ((st = upload_vec(&fcode, m.fcode))) return st;
    if ((st = upload_vec(&fbc, m.fbc))) return st;
    NW_TRY_HIP(launch_metrics(corners, n1, npts, d_met, stream));
    if ((st = dalloc(reinterpret_cast<void**>(&d_vM), 4 * tfp))) return st;
    if ((st = dalloc(reinterpret_cast<void**>(&d_vP), 4 * tfp))) return st;
    if ((st = dalloc(reinterpret_cast<void**>(&d_bc), 4 * tfp))) return st;
    for (int a = 0; a < 4; ++a)
        if ((st = dalloc(reinterpret_cast<void**>(&d_face[a]), 8 * tfp))) return st;
    const double* met10[10] = {d_met[3], d_met[4], d_met[5], d_met[6],  d_met[7],
                               d_met[8], d_met[9], d_met[10], d_met[11], d_met[12]};
    NW_TRY_HIP(launch_face_data(fcode, fbc, n1, tfp, npts, met10, d_vM, d_vP, d_bc, d_face[0], d_face[1], d_face[2],
                                d_face[3], stream));
    if ((st = dalloc(reinterpret_cast<void**>(&d_fEN), 8 * 3 * tfp))) return st;
    if ((st = dalloc(reinterpret_cast<void**>(&d_fHN), 8 * 3 * tfp))) return st;
    if ((st = dalloc(reinterpret_cast<void**>(&d_flux), 8 * 6 * tfp))) return st;
    if (media) {
        if ((st = upload_vec(&d_eps, m.eps))) return st;
        if ((st = upload_vec(&d_mu, m.mu_))) return st;
    }
    kernels_per_step = 5 * (7 + (nhalo > 0 ? 2 : 0));
    return NW_OK;
}

int Solver::upload(int ix, const double* host, size_t n) {
    if (!initialized) { set_error("upload before create"); return NW_ERR_STATE; }
    NW_CHECK_ARG(ix >= 0 && ix < (int)inst.size(), "bad instance index");
    const Instance& I = inst[ix];
    NW_CHECK_ARG(host != nullptr && n == size_t(6 * I.npts), "upload: size must be 6*npts");
    NW_TRY_HIP(hipSetDevice(device));
    const long long P = I.npts, S = stride;
    if (pipe == NW_PIPE_PLAIN) {
        NW_TRY_HIP(hipMemcpy2DAsync(d_state, S * 8, host, P * 8, P * 8, 6, hipMemcpyHostToDevice, stream));
    } else {
        // host blocks [Ex|..|Hz] -> node-major device layout (6 values per node)
        NW_TRY_HIP(hipMemcpyAsync(d_stage, host, 8 * 6 * P, hipMemcpyHostToDevice, stream));
        if (prec == kFP64) {
            double* dst = static_cast<double*>(d_u[cur]) + 6 * I.node_base;
            NW_TRY_HIP((launch_convert<double, double>(dst, 1, 6, d_stage, P, 1, P, 6, stream)));
        } else {
            float* dst = static_cast<float*>(d_u[cur]) + 6 * I.node_base;
            NW_TRY_HIP((launch_convert<float, double>(dst, 1, 6, d_stage, P, 1, P, 6, stream)));
        }
    }
    NW_TRY_HIP(hipStreamSynchronize(stream));
    uploaded = true;
    return NW_OK;
}

int Solver::download(int ix, double* host, size_t n) {
    if (!initialized) { set_error("download before create"); return NW_ERR_STATE; }
    NW_CHECK_ARG(ix >= 0 && ix < (int)inst.size(), "bad instance index");
    const Instance& I = inst[ix];
    NW_CHECK_ARG(host != nullptr && n == size_t(6 * I.npts), "download: size must be 6*npts");
    NW_TRY_HIP(hipSetDevice(device));
    const long long P = I.npts, S = stride;
    if (pipe == NW_PIPE_PLAIN) {
        NW_TRY_HIP(hipMemcpy2DAsync(host, P * 8, d_state, S * 8, P * 8, 6, hipMemcpyDeviceToHost, stream));
    } else {
        if (prec == kFP64) {
            const double* src = static_cast<const double*>(d_u[cur]) + 6 * I.node_base;
            NW_TRY_HIP((launch_convert<double, double>(d_stage, P, 1, src, 1, 6, P, 6, stream)));
        } else {
            const float* src = static_cast<const float*>(d_u[cur]) + 6 * I.node_base;
            NW_TRY_HIP((launch_convert<double, float>(d_stage, P, 1, src, 1, 6, P, 6, stream)));
        }
        NW_TRY_HIP(hipMemcpyAsync(host, d_stage, 8 * 6 * P, hipMemcpyDeviceToHost, stream));
    }
    NW_TRY_HIP(hipStreamSynchronize(stream));
    return NW_OK;
}

int Solver::set_dt(double dt) {
    if (dt == last_dt) return NW_OK;
    NW_TRY_HIP(hipStreamSynchronize(stream));
    NW_TRY_HIP(hipMemcpy(d_dt, &dt, sizeof(double), hipMemcpyHostToDevice));
    last_dt = dt;
    return NW_OK;
}

long long Solver::working_set_bytes() const {
    const long long su = prec == kFP64 ? 8 : 4;
    const long long sk = (prec == kFP32 || prec == kBF16X3) ? 4 : 8;
    return npts * 6 * (2 * su + sk);
}

int Solver::stream_of(const Segment& sg, int option) const {
    if (sg.stream >= 0) return sg.stream;
    const int threads = fused_epw(sg.n1, option, prec) * sg.n1 * sg.n1 * sg.n1;
    const int ksize = (prec == kFP32 || prec == kBF16X3) ? 4 : 8;
    return fused_stream_k(working_set_bytes(), threads, ksize) ? 1 : 0;
}

int Solver::launch_segment(const Segment& sg, int s, int parity, int option, int stream_k, hipStream_t st,
                           bool backwards) {
    FusedArgs a{};
    a.uin = d_u[parity];
    a.uout = d_u[1 - parity];
    a.kbuf = d_k;
    a.stride = stride;
    a.kstride = 64 * kchunks;
    a.kbase = sg.kbase;
    a.node_base = sg.node_base;
    a.ghost_base = npts;
    a.nelt = sg.nelt;
    a.elem_base = sg.elem_base;
    a.geom = d_geom + sg.elem_base;
    a.rec = d_rec + size_t(sg.elem_base) * 26;
    a.mat = d_mat;
    a.gmat = d_gmat;
    a.fnb = d_fnb + size_t(sg.elem_base) * 6;
    a.gorder = (backwards ? d_gorder_rev : d_gorder) + sg.goff[option];
    a.ghead = (backwards ? d_ghead_rev : d_ghead) + sg.soff[option];
    a.backwards = backwards ? 1 : 0;
    a.nseq = sg.nseq[option];
    a.dt = d_dt;
    a.a = kRkA[s];
    a.b = kRkB[s];
    a.c0 = c0;
    a.first_stage = (s == 0);
    a.last_stage = (s == kStages - 1);
    a.affine = sg.affine ? 1 : 0;
    a.stream_k = stream_k;
    a.fw = static_cast<const char*>(d_fw) + size_t(sg.elem_base) * 24 * (prec == kFP64 ? 8 : 4);
    NW_TRY_HIP(launch_fused(sg.n1, prec, option, a, st));
    return NW_OK;
}

// One stage of the fused pipeline, reading d_u[parity] and writing d_u[1-parity].
// Segments of different n1 are independent within a stage.
int Solver::enqueue_fused_stage(int s, int parity) {
    void* uin = d_u[parity];
    if (nhalo > 0) {
        const int nh = static_cast<int>(nhalo);
        if (prec == kFP64) {
            NW_TRY_HIP(launch_pack<double>(static_cast<const double*>(uin), static_cast<double*>(d_sendbuf), d_send_vol,
                                           d_off, d_sizes, nh, 1, 6, stream));
        } else {
            NW_TRY_HIP(launch_pack<float>(static_cast<const float*>(uin), static_cast<float*>(d_sendbuf), d_send_vol,
                                          d_off, d_sizes, nh, 1, 6, stream));
        }
        const size_t rsz = (prec == kFP64) ? 8 : 4;
        NW_TRY_HIP(hipMemcpyAsync(d_recvbuf, d_sendbuf, rsz * 6 * nhalo, hipMemcpyDeviceToDevice, stream));
        if (prec == kFP64) {
            NW_TRY_HIP(launch_unpack<double>(static_cast<const double*>(d_recvbuf), static_cast<double*>(uin), d_off,
                                             d_sizes, nh, npts, 1, 6, stream));
        } else {
            NW_TRY_HIP(launch_unpack<float>(static_cast<const float*>(d_recvbuf), static_cast<float*>(uin), d_off,
                                            d_sizes, nh, npts, 1, 6, stream));
        }
    }
    // Segments (one per n1) run one after another on the stream: each launch is
    // a persistent grid sized to fill the GPU, so concurrent segment launches
    // would only compete for the same CUs.
    int st;
    // the ping-pong parity alternates from stage to stage, across steps too:
    // with an alternating walk every other stage walks backwards, and a batch
    // then takes its segments in reverse order (the segment at the turn starts
    // on its own last writes)
    const bool back = parity != 0 && alternating_walk();
    for (size_t i = 0; i < seg.size(); ++i) {
        const size_t x = back ? seg.size() - 1 - i : i;
        if ((st = launch_segment(seg[x], s, parity, seg[x].option, stream_of(seg[x], seg[x].option), stream, back)))
            return st;
    }
    return NW_OK;
}

int Solver::enqueue_plain_stage(int s) {
    double* st = d_state;
    const long long S = stride;
    if (nhalo > 0) {
        const int nh = static_cast<int>(nhalo);
        NW_TRY_HIP(launch_pack<double>(st, static_cast<double*>(d_sendbuf), d_send_vol, d_off, d_sizes, nh, S, 1, stream));
        NW_TRY_HIP(hipMemcpyAsync(d_recvbuf, d_sendbuf, 8 * 6 * nhalo, hipMemcpyDeviceToDevice, stream));
    }
    PlainVolumeArgs v{};
    v.D = d_D; v.w3g = d_w3; v.jac = d_met[3];
    v.rx = d_met[4]; v.sx = d_met[5]; v.tx = d_met[6]; v.ry = d_met[7]; v.sy = d_met[8]; v.ty = d_met[9];
    v.rz = d_met[10]; v.sz = d_met[11]; v.tz = d_met[12];
    v.n1 = n1; v.nelt = nelt;
    // +curl H -> rhs_E ; -curl E -> rhs_H (dg_solver.cpp:393-413)
    v.u1 = st + 3 * S; v.u2 = st + 4 * S; v.u3 = st + 5 * S;
    v.w1 = d_rhs; v.w2 = d_rhs + npts; v.w3 = d_rhs + 2 * npts; v.sign = 1.0;
    NW_TRY_HIP(launch_volume_curl(v, stream));
    v.u1 = st; v.u2 = st + S; v.u3 = st + 2 * S;
    v.w1 = d_rhs + 3 * npts; v.w2 = d_rhs + 4 * npts; v.w3 = d_rhs + 5 * npts; v.sign = -1.0;
    NW_TRY_HIP(launch_volume_curl(v, stream));
    if (nhalo > 0) {
        NW_TRY_HIP(launch_unpack<double>(static_cast<const double*>(d_recvbuf), st, d_off, d_sizes,
                                         static_cast<int>(nhalo), npts, S, 1, stream));
    }
    NW_TRY_HIP(launch_restrict(st, d_fEN, d_fHN, d_vM, tfp, S, stream));
    NW_TRY_HIP(launch_flux(st, d_fEN, d_fHN, d_vP, d_bc, d_face[0], d_face[1], d_face[2], d_flux, c0, tfp, S, d_vM,
                           media ? d_eps : nullptr, media ? d_mu : nullptr, d_ghost_elem, np, npts, stream));
    NW_TRY_HIP(launch_lift(d_flux, d_face[3], d_rhs, tfp, npts, n1, stream));
    NW_TRY_HIP(launch_inv_mass(d_rhs, d_met[3], d_w3, np, npts, media ? d_eps : nullptr, media ? d_mu : nullptr, stream));
    NW_TRY_HIP(launch_lsrk(st, d_kk, d_rhs, kRkA[s], kRkB[s], d_dt, 0.0, npts, S, stream));
    return NW_OK;
}

int Solver::enqueue_step(int parity) {
    int st;
    for (int s = 0; s < kStages; ++s) {
        if (pipe == NW_PIPE_FUSED) st = enqueue_fused_stage(s, (parity + s) & 1);
        else st = enqueue_plain_stage(s);
        if (st) return st;
    }
    // the step's result sits in d_u[parity ^ 1] (fused) or d_state (plain)
    return nprobes > 0 ? enqueue_probe(parity ^ 1) : NW_OK;
}

// One probe record of the state in d_u[state_parity] (fused) or d_state (plain).
int Solver::enqueue_probe(int state_parity) {
    if (pipe == NW_PIPE_PLAIN) {
        NW_TRY_HIP(launch_probe<double>(d_state, stride, 1, nprobes, d_probe_base, d_probe_w, n1, d_probe_hist,
                                        d_probe_slot, probe_cap, stream));
    } else if (prec == kFP64) {
        NW_TRY_HIP(launch_probe<double>(static_cast<const double*>(d_u[state_parity]), 1, 6, nprobes, d_probe_base,
                                        d_probe_w, n1, d_probe_hist, d_probe_slot, probe_cap, stream));
    } else {
        NW_TRY_HIP(launch_probe<float>(static_cast<const float*>(d_u[state_parity]), 1, 6, nprobes, d_probe_base,
                                       d_probe_w, n1, d_probe_hist, d_probe_slot, probe_cap, stream));
    }
    return NW_OK;
}

void Solver::drop_graphs() {
    for (int p = 0; p < 2; ++p) {
        if (gexec[p]) (void)hipGraphExecDestroy(gexec[p]);
        if (graph[p]) (void)hipGraphDestroy(graph[p]);
        gexec[p] = nullptr;
        graph[p] = nullptr;
    }
}

// ProbeManager::init (probe.cpp:42-103) for the device: place the probes on the
// host exactly as the reference does, keep per probe its element's first node
// and the 1-D Lagrange weights, and a history of `capacity` records that every
// step appends to (the step graphs are rebuilt to include the record).
int Solver::set_probes(const Mesh& m, int n, const double* xyz, int capacity) {
    if (!initialized) { set_error("set_probes before create"); return NW_ERR_STATE; }
    NW_CHECK_ARG(n >= 0 && (n == 0 || xyz != nullptr), "bad probe list");
    NW_CHECK_ARG(capacity >= 1, "probe history capacity must be >= 1");
    NW_CHECK_ARG(inst.size() == 1, "probes need a single-mesh solver");
    NW_CHECK_ARG(m.n1 == n1 && m.nelt == nelt && mesh_signature(m) == mesh_sig,
                 "probes: mesh is not the one the solver was created from (corners or faces differ)");
    NW_CHECK_ARG(n <= 65535, "too many probes (at most 65535)");
    NW_TRY_HIP(hipSetDevice(device));
    NW_TRY_HIP(hipStreamSynchronize(stream));
    drop_graphs();
    nprobes = 0;
    dfree(d_probe_base);
    dfree(d_probe_w);
    dfree(d_probe_hist);
    d_probe_base = nullptr;
    d_probe_w = nullptr;
    d_probe_hist = nullptr;
    if (n == 0) return NW_OK;
    std::vector<int> elem(n);
    std::vector<double> rst(3 * size_t(n));
    locate_probes(m, n, xyz, elem.data(), rst.data());
    const GllTables& t = host_gll_tables();
    std::vector<long long> base(n);
    std::vector<double> lw(size_t(n) * 3 * n1);
    for (int q = 0; q < n; ++q) {
        base[q] = (long long)elem[q] * np;
        for (int d = 0; d < 3; ++d) lagrange_1d(n1, t.z[n1], rst[3 * q + d], &lw[(size_t(q) * 3 + d) * n1]);
    }
    int st;
    if ((st = upload_vec(&d_probe_base, base))) return st;
    if ((st = upload_vec(&d_probe_w, lw))) return st;
    if ((st = dalloc(reinterpret_cast<void**>(&d_probe_hist), sizeof(double) * 6 * size_t(n) * capacity))) return st;
    if (!d_probe_slot && (st = dalloc(reinterpret_cast<void**>(&d_probe_slot), sizeof(int)))) return st;
    NW_TRY_HIP(launch_zero_int(d_probe_slot, stream));
    NW_TRY_HIP(hipStreamSynchronize(stream));
    nprobes = n;
    probe_cap = capacity;
    return NW_OK;
}

// Records the current state once (the reference records t = 0 before stepping,
// case.cpp:473-474).
int Solver::probe_sample() {
    if (!initialized || !uploaded) { set_error("probe_sample before create/upload"); return NW_ERR_STATE; }
    if (nprobes <= 0) { set_error("no probes set"); return NW_ERR_STATE; }
    NW_TRY_HIP(hipSetDevice(device));
    int st = enqueue_probe(cur);
    if (st) return st;
    NW_TRY_HIP(hipStreamSynchronize(stream));
    return NW_OK;
}

// Downloads the records taken since the last call (oldest first) and starts
// the history again; more records than the capacity is an error (the oldest
// were overwritten).
int Solver::probe_history(double* host, int max_records, int* nrecords) {
    if (!initialized) { set_error("probe_history before create"); return NW_ERR_STATE; }
    NW_CHECK_ARG(nrecords != nullptr, "null record count");
    if (nprobes <= 0) { *nrecords = 0; return NW_OK; }
    NW_TRY_HIP(hipSetDevice(device));
    int taken = 0;
    NW_TRY_HIP(hipMemcpyAsync(&taken, d_probe_slot, sizeof(int), hipMemcpyDeviceToHost, stream));
    NW_TRY_HIP(hipStreamSynchronize(stream));
    if (taken > probe_cap) {
        set_error("probe history overflow: fetch at least every capacity steps");
        return NW_ERR_STATE;
    }
    NW_CHECK_ARG(host != nullptr || taken == 0, "null host buffer");
    NW_CHECK_ARG(max_records >= taken, "host buffer holds fewer records than were taken");
    if (taken > 0)
        NW_TRY_HIP(hipMemcpyAsync(host, d_probe_hist, sizeof(double) * 6 * size_t(nprobes) * taken,
                                  hipMemcpyDeviceToHost, stream));
    NW_TRY_HIP(launch_zero_int(d_probe_slot, stream));
    NW_TRY_HIP(hipStreamSynchronize(stream));
    *nrecords = taken;
    return NW_OK;
}

int Solver::build_graph(int parity) {
    hipGraph_t g = nullptr;
    NW_TRY_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int st = enqueue_step(parity);
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return NW_ERR_HIP;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        return NW_ERR_HIP;
    }
    graph[parity] = g;
    gexec[parity] = x;
    return NW_OK;
}

// Tuned tile options and k policies kept for the life of the process, keyed by
// what decides them: device, n1, precision, element count of the segment,
// whether it is a single mesh (Morton walk) or a batch segment (storage order),
// whether the working set exceeds the Infinity Cache (non-temporal k a
// candidate) and the cap on workgroups per CU.  A later solver of the same
// shape takes the choice without timing again.
namespace {
struct TuneKey {
    int device, n1, prec, single, streamable, per_cu;
    long long nelt;
    bool operator<(const TuneKey& o) const {
        if (device != o.device) return device < o.device;
        if (per_cu != o.per_cu) return per_cu < o.per_cu;
        if (n1 != o.n1) return n1 < o.n1;
        if (prec != o.prec) return prec < o.prec;
        if (single != o.single) return single < o.single;
        if (streamable != o.streamable) return streamable < o.streamable;
        return nelt < o.nelt;
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tune;  // (tile option, non-temporal k)
}  // namespace

// Autotuner: per n1 segment, time every distinct tile option (elements per
// workgroup, or at n1 = 8 the matrix-core volume) on a regular stage (it reads k
// and writes u_out and k, all rewritten by the first real stage, which does not
// read k) and keep the fastest.  The options alternate twice and each keeps its
// best time, so a slow moment of the GPU does not decide.  Runs at a step
// boundary, before graphs are captured.
int Solver::tune() {
    if (pipe != NW_PIPE_FUSED) return NW_OK;
    NW_TRY_HIP(hipSetDevice(device));
    // non-temporal k is a candidate only where the working set exceeds the
    // Infinity Cache (below it the hint only costs)
    const int nstream = fused_beyond_infinity_cache(working_set_bytes()) ? 2 : 1;
    for (Segment& sg : seg) {
        const TuneKey key{device, sg.n1, prec, inst.size() == 1 ? 1 : 0, nstream - 1, groups_per_cu, sg.nelt};
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune.find(key);
            if (it != g_tune.end()) {
                sg.option = it->second.first;
                sg.stream = it->second.second;
                continue;
            }
        }
        float best[kNumTileOptions][2];
        for (int opt = 0; opt < kNumTileOptions; ++opt) {  // warm each (first launch sets up the kernel)
            best[opt][0] = best[opt][1] = 1e30f;
            if (fused_option_of(sg.n1, opt, prec) != opt) continue;  // the same kernel as an earlier option
            const int st = launch_segment(sg, 1, cur, opt, 0, stream);
            if (st) return st;
        }
        // candidates: distinct options x k policies
        std::vector<std::pair<int, int>> cand;
        for (int opt = 0; opt < kNumTileOptions; ++opt)
            if (fused_option_of(sg.n1, opt, prec) == opt)
                for (int str = 0; str < nstream; ++str) cand.push_back({opt, str});
        auto timed = [&](int opt, int str, int n, float* ms) -> int {
            NW_TRY_HIP(hipEventRecord(ev0, stream));
            for (int r = 0; r < n; ++r) {
                const int st = launch_segment(sg, 1, cur, opt, str, stream);
                if (st) return st;
            }
            NW_TRY_HIP(hipEventRecord(ev1, stream));
            NW_TRY_HIP(hipEventSynchronize(ev1));
            NW_TRY_HIP(hipEventElapsedTime(ms, ev0, ev1));
            return NW_OK;
        };
        // bring the clocks up first (after host-side set-up they ramp for tens
        // of milliseconds, which would favour whichever candidate runs last):
        // about 30 ms of launches
        float one = 0.0f;
        int st = timed(0, 0, 1, &one);
        if (st) return st;
        const int nsettle = std::max(2, std::min(400, static_cast<int>(30.0f / std::max(one, 1e-3f))));
        if ((st = timed(0, 0, nsettle, &one))) return st;
        one /= static_cast<float>(nsettle);
        // three rounds, forward and backward in turn (a drift of the clocks
        // then does not favour one end), best of the rounds per candidate;
        // each measurement at least ~2 ms of launches (a small segment's
        // launch takes tens of microseconds)
        const int nlaunch = std::max(3, std::min(200, static_cast<int>(2.0f / std::max(one, 1e-3f))));
        for (int rep = 0; rep < 3; ++rep)
            for (size_t x = 0; x < cand.size(); ++x) {
                const auto& c = cand[rep & 1 ? cand.size() - 1 - x : x];
                float ms = 0.0f;
                if ((st = timed(c.first, c.second, nlaunch, &ms))) return st;
                best[c.first][c.second] = std::min(best[c.first][c.second], ms);
            }
        sg.option = 0;
        sg.stream = 0;
        for (int opt = 0; opt < kNumTileOptions; ++opt)
            for (int str = 0; str < nstream; ++str)
                if (best[opt][str] < best[sg.option][sg.stream]) {
                    sg.option = opt;
                    sg.stream = str;
                }
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = {sg.option, sg.stream};
    }
    drop_graphs();  // captured launches carry the previous choices
    tuned = true;
    return NW_OK;
}

int Solver::set_segment_option(int s, int option, int stream_k) {
    NW_CHECK_ARG(pipe == NW_PIPE_FUSED, "tile options belong to the fused pipeline");
    NW_CHECK_ARG(s >= 0 && s < static_cast<int>(seg.size()), "bad segment");
    NW_CHECK_ARG(option >= 0 && option < kNumTileOptions, "tile option must be 0, 1 or 2");
    NW_CHECK_ARG(stream_k >= -1 && stream_k <= 1, "stream_k must be -1, 0 or 1");
    NW_TRY_HIP(hipSetDevice(device));
    NW_TRY_HIP(hipStreamSynchronize(stream));
    seg[size_t(s)].option = option;
    seg[size_t(s)].stream = stream_k;
    tuned = true;  // an explicit choice stands until tune() is called
    drop_graphs();
    return NW_OK;
}

// Default rule on the largest segment's working set: segments run one after
// another, and with the order reversed on backward stages the largest one (a
// single mesh's only one) follows its own previous stage at the turn.
bool Solver::alternating_walk() const {
    if (pipe != NW_PIPE_FUSED) return false;
    if (walk_mode >= 0) return walk_mode == 1;
    long long most = 0;
    for (const Segment& sg : seg)
        most = std::max(most, (long long)sg.nelt * sg.n1 * sg.n1 * sg.n1);
    return fused_alternate_walk(working_set_bytes() / std::max(1ll, npts) * most);
}

int Solver::set_walk(int mode) {
    NW_CHECK_ARG(pipe == NW_PIPE_FUSED, "the walk belongs to the fused pipeline");
    NW_CHECK_ARG(mode >= -1 && mode <= 1, "walk mode must be -1, 0 or 1");
    NW_TRY_HIP(hipSetDevice(device));
    NW_TRY_HIP(hipStreamSynchronize(stream));
    walk_mode = mode;
    drop_graphs();  // captured launches carry the previous direction
    return NW_OK;
}

int Solver::launch_steps(double dt, int nsteps) {
    if (!initialized) { set_error("step before create"); return NW_ERR_STATE; }
    if (!uploaded) { set_error("step before uploadState"); return NW_ERR_STATE; }
    NW_CHECK_ARG(nsteps >= 0, "nsteps must be >= 0");
    NW_TRY_HIP(hipSetDevice(device));
    int st = set_dt(dt);
    if (st) return st;
    if (autotune && !tuned && nsteps > 0 && (st = tune())) return st;
    for (int n = 0; n < nsteps; ++n) {
        const int parity = (pipe == NW_PIPE_FUSED) ? cur : 0;
        if (use_graph) {
            if (!gexec[parity] && (st = build_graph(parity))) return st;
            NW_TRY_HIP(hipGraphLaunch(gexec[parity], stream));
        } else {
            if ((st = enqueue_step(parity))) return st;
        }
        if (pipe == NW_PIPE_FUSED) cur ^= 1;  // five stages flip the ping-pong parity
    }
    return NW_OK;
}

int Solver::step(double dt, int nsteps) {
    int st = launch_steps(dt, nsteps);
    if (st) return st;
    NW_TRY_HIP(hipStreamSynchronize(stream));
    return NW_OK;
}

int Solver::time(double dt, int warmup, int nsteps, int repeats, double* ms) {
    if (!initialized || !uploaded) { set_error("time before create/upload"); return NW_ERR_STATE; }
    NW_CHECK_ARG(warmup >= 0 && nsteps >= 0 && repeats >= 1 && repeats <= 1024, "bad warmup/steps/repeats");
    NW_TRY_HIP(hipSetDevice(device));
    int st = set_dt(dt);
    if (st) return st;
    // tune and capture outside the timed region
    if (autotune && !tuned && (st = tune())) return st;
    if (use_graph) {
        for (int p = 0; p < 2; ++p)
            if (!gexec[p] && (pipe == NW_PIPE_FUSED || p == 0) && (st = build_graph(p))) return st;
    }
    // warm-up steps, then `repeats` timed runs of exactly nsteps steps, all
    // queued back to back: no host wait between them, one at the end
    std::vector<hipEvent_t> ev(size_t(repeats) + 1, nullptr);
    auto destroy = [&]() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    for (hipEvent_t& e : ev) {
        const hipError_t he = hipEventCreate(&e);
        if (he != hipSuccess) {
            destroy();
            set_error(std::string("hipEventCreate: ") + hipGetErrorString(he));
            return NW_ERR_HIP;
        }
    }
    hipError_t he = hipSuccess;
    if ((st = launch_steps(dt, warmup))) { destroy(); return st; }
    he = hipEventRecord(ev[0], stream);
    for (int r = 0; r < repeats && he == hipSuccess && st == NW_OK; ++r) {
        st = launch_steps(dt, nsteps);
        if (st == NW_OK) he = hipEventRecord(ev[size_t(r) + 1], stream);
    }
    if (he == hipSuccess && st == NW_OK) he = hipEventSynchronize(ev[size_t(repeats)]);
    for (int r = 0; r < repeats && he == hipSuccess && st == NW_OK; ++r) {
        float f = 0.0f;
        he = hipEventElapsedTime(&f, ev[size_t(r)], ev[size_t(r) + 1]);
        ms[r] = f;
    }
    destroy();
    if (st) return st;
    if (he != hipSuccess) {
        set_error(std::string("timing events: ") + hipGetErrorString(he));
        return NW_ERR_HIP;
    }
    return NW_OK;
}

int Solver::diag(int ix, nw_diag* out) {
    if (!initialized) { set_error("diag before create"); return NW_ERR_STATE; }
    NW_CHECK_ARG(ix >= 0 && ix < (int)inst.size(), "bad instance index");
    const Instance& I = inst[ix];
    NW_TRY_HIP(hipSetDevice(device));
    const ElemMat* mat = I.media ? d_mat + I.elem_base : nullptr;
    const ElemGeom* geom = d_geom + I.elem_base;
    if (pipe == NW_PIPE_PLAIN) {
        NW_TRY_HIP(launch_diag<double>(d_state, stride, 1, npts, n1, d_met[3], nullptr, mat, d_diag_part, d_diag_out,
                                       stream));
    } else if (prec == kFP64) {
        NW_TRY_HIP(launch_diag<double>(static_cast<const double*>(d_u[cur]) + 6 * I.node_base, 1, 6, I.npts, I.n1,
                                       nullptr, geom, mat, d_diag_part, d_diag_out, stream));
    } else {
        NW_TRY_HIP(launch_diag<float>(static_cast<const float*>(d_u[cur]) + 6 * I.node_base, 1, 6, I.npts, I.n1,
                                      nullptr, geom, mat, d_diag_part, d_diag_out, stream));
    }
    double h[3];
    NW_TRY_HIP(hipMemcpyAsync(h, d_diag_out, sizeof(h), hipMemcpyDeviceToHost, stream));
    NW_TRY_HIP(hipStreamSynchronize(stream));
    out->energy = h[0];
    out->max_e = h[1];
    out->max_h = h[2];
    return NW_OK;
}

// Relative L2 distance of instance ix's state to a host reference state (the
// reference's layout, 6 component blocks of npts), weighted by J w_i w_j w_k
// (3dboxpec_test.cpp:76-119), reduced on the device in FP64.
int Solver::rel_l2(int ix, const double* ref, size_t n, double* out) {
    if (!initialized) { set_error("rel_l2 before create"); return NW_ERR_STATE; }
    NW_CHECK_ARG(ix >= 0 && ix < (int)inst.size(), "bad instance index");
    const Instance& I = inst[ix];
    NW_CHECK_ARG(ref != nullptr && out != nullptr && n == size_t(6 * I.npts), "rel_l2: reference size must be 6*npts");
    NW_TRY_HIP(hipSetDevice(device));
    NW_TRY_HIP(hipMemcpyAsync(d_stage, ref, 8 * 6 * I.npts, hipMemcpyHostToDevice, stream));
    const ElemGeom* geom = d_geom + I.elem_base;
    if (pipe == NW_PIPE_PLAIN) {
        NW_TRY_HIP(launch_l2<double>(d_state, stride, 1, d_stage, npts, n1, d_met[3], nullptr, d_diag_part, d_diag_out,
                                     stream));
    } else if (prec == kFP64) {
        NW_TRY_HIP(launch_l2<double>(static_cast<const double*>(d_u[cur]) + 6 * I.node_base, 1, 6, d_stage, I.npts,
                                     I.n1, nullptr, geom, d_diag_part, d_diag_out, stream));
    } else {
        NW_TRY_HIP(launch_l2<float>(static_cast<const float*>(d_u[cur]) + 6 * I.node_base, 1, 6, d_stage, I.npts, I.n1,
                                    nullptr, geom, d_diag_part, d_diag_out, stream));
    }
    double h[2];
    NW_TRY_HIP(hipMemcpyAsync(h, d_diag_out, sizeof(h), hipMemcpyDeviceToHost, stream));
    NW_TRY_HIP(hipStreamSynchronize(stream));
    *out = h[1] > 0.0 ? std::sqrt(h[0] / h[1]) : (h[0] > 0.0 ? HUGE_VAL : 0.0);
    return NW_OK;
}

}  // namespace nw
