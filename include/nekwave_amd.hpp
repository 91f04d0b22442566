// nekwave_amd.hpp -- C++ host interface over the C ABI (nekwave_amd.h).
//
// Same classes, method names and argument meanings as the reference's host
// API, so code written against NekWave's Mesh / DgSolver / Case compiles
// against this header:
//   Mesh      src/mesh.hpp:51-218   (createBoxMesh, GLL data, dt helpers)
//   DgSolver  src/dg_solver.hpp:35-102 (initialize, uploadState,
//             downloadState, step, finalize)
//   Config    src/config.hpp:21-57  (the keys that shape the step)
//   Case      src/case.hpp:48-153   (preprocess, simulate, postprocess, run,
//             initial-condition and post-processing hooks, energy and max
//             diagnostics)
// Errors: the reference aborts or asserts; here every failed call throws
// nekwave_amd::Error carrying the NW_* status and message.  initialize()
// keeps the reference's bool return.
#ifndef NEKWAVE_AMD_HPP
#define NEKWAVE_AMD_HPP

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <array>
#include <vector>

#include "nekwave_amd.h"

namespace nekwave_amd {

class Error : public std::runtime_error {
   public:
    Error(int status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
    int status() const { return status_; }

   private:
    int status_;
};

inline void check(int st) {
    if (st != NW_OK) throw Error(st, nw_last_error());
}

using StateVector = std::vector<double>;  // [Ex|Ey|Ez|Hx|Hy|Hz], dg_solver.hpp:14

class Mesh {
   public:
    explicit Mesh(int N = 3, int numElements = 1) : N_(N), nel0_(numElements) {}
    ~Mesh() { nw_mesh_destroy(m_); }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;

    void setN(int N) { N_ = N; }
    bool createBoxMesh(int nelx, int nely, int nelz, double xmin = -1.0, double xmax = 1.0, double ymin = -1.0,
                       double ymax = 1.0, double zmin = -1.0, double zmax = 1.0, bool periodicX = false,
                       bool periodicY = false, bool periodicZ = false, double jitter = 0.0, uint64_t seed = 0) {
        nw_mesh_destroy(m_);
        m_ = nullptr;
        const double b[6] = {xmin, xmax, ymin, ymax, zmin, zmax};
        check(nw_mesh_box(N_, nelx, nely, nelz, b, periodicX, periodicY, periodicZ, jitter, seed, &m_));
        return true;
    }
    void setMaterials(const std::vector<double>& eps, const std::vector<double>& mu) {
        check(nw_mesh_set_materials(m_, eps.empty() ? nullptr : eps.data(), mu.empty() ? nullptr : mu.data()));
    }
    int selfHaloX() {
        int n = 0;
        check(nw_mesh_self_halo_x(m_, &n));
        return n;
    }

    int getN() const { return N_; }
    int getNumPointsPerElement() const { return N_ * N_ * N_; }
    int getNumElements() const { return m_ ? info().nelt : nel0_; }
    int getTotalPoints() const { return static_cast<int>(info().npts); }
    int getNumHaloPoints() const { return info().nhalo; }
    bool isPartitioned() const { return getNumHaloPoints() > 0; }
    double computeMinNodeDistance() const {
        double v = 0.0;
        check(nw_mesh_min_node_distance(m_, &v));
        return v;
    }
    double computeAutomaticCFL(double safetyFactor = 0.4) const { return (safetyFactor * 1.75) / std::sqrt(3.0); }
    double computeAutomaticDt(double waveSpeed = 1.0, double safetyFactor = 0.4) const {
        double v = 0.0;
        check(nw_mesh_auto_dt(m_, waveSpeed, safetyFactor, &v));
        return v;
    }
    std::vector<double> getGllZ() const { return gll(0); }
    std::vector<double> getGllW() const { return gll(1); }
    std::vector<double> getD() const { return gll(2); }
    std::vector<double> getW3() const { return gll(3); }
    std::vector<double> getCoordX() const { return coord(0); }
    std::vector<double> getCoordY() const { return coord(1); }
    std::vector<double> getCoordZ() const { return coord(2); }
    std::vector<double> getJac() const {
        std::vector<double> j(getTotalPoints());
        check(nw_mesh_jacobian(m_, j.data()));
        return j;
    }
    const nw_mesh* handle() const { return m_; }

   private:
    nw_mesh_info info() const {
        nw_mesh_info i;
        check(nw_mesh_get_info(m_, &i));
        return i;
    }
    std::vector<double> gll(int which) const {
        const int n = N_;
        std::vector<double> z(n), w(n), D(n * n), w3(n * n * n);
        check(nw_mesh_gll(m_, z.data(), w.data(), D.data(), w3.data()));
        return which == 0 ? z : which == 1 ? w : which == 2 ? D : w3;
    }
    std::vector<double> coord(int which) const {
        const size_t n = getTotalPoints();
        std::vector<double> x(n), y(n), z(n);
        check(nw_mesh_coords(m_, x.data(), y.data(), z.data()));
        return which == 0 ? x : which == 1 ? y : z;
    }
    int N_;
    int nel0_;
    nw_mesh* m_ = nullptr;
};

class DgSolver {
   public:
    DgSolver() { nw_opts_default(&opts_); }
    explicit DgSolver(const nw_opts& o) : opts_(o) {}
    ~DgSolver() { finalize(); }
    DgSolver(const DgSolver&) = delete;
    DgSolver& operator=(const DgSolver&) = delete;

    bool initialize(const Mesh& mesh, double c0) {
        finalize();
        if (nw_dg_create(mesh.handle(), c0, &opts_, &s_) != NW_OK) {
            s_ = nullptr;
            return false;
        }
        return true;
    }
    void uploadState(const double* hostState, size_t size) { check(nw_dg_upload(s_, hostState, size)); }
    void downloadState(double* hostState, size_t size) const { check(nw_dg_download(s_, hostState, size)); }
    void step(double dt, double time) { check(nw_dg_step(s_, dt, time)); }
    void finalize() {
        nw_dg_destroy(s_);
        s_ = nullptr;
    }
    // extensions
    void run(double dt, int nsteps) { check(nw_dg_run(s_, dt, nsteps)); }
    nw_diag diagnostics() const {
        nw_diag d;
        check(nw_dg_diag(s_, &d));
        return d;
    }
    // Extension: relative L2 distance of the state to a host reference state
    // (uploadState layout), reduced on the device (nw_dg_rel_l2).
    double relL2(const double* ref, size_t n) const {
        double r = 0.0;
        check(nw_dg_rel_l2(s_, 0, ref, n, &r));
        return r;
    }
    // Extension: on-device ProbeManager (probe.cpp:33-216).  Every step after
    // setProbes appends one record (6 values per probe) to a device history of
    // `capacity` records; probeHistory copies and clears it.
    void setProbes(const Mesh& mesh, const std::vector<std::array<double, 3>>& xyz, int capacity) {
        check(nw_dg_set_probes(s_, mesh.handle(), static_cast<int>(xyz.size()),
                               xyz.empty() ? nullptr : xyz[0].data(), capacity));
        nprobes_ = static_cast<int>(xyz.size());
        probe_cap_ = capacity;
    }
    void sampleProbes() { check(nw_dg_probe_sample(s_)); }
    // records x probes x 6 values, oldest record first
    std::vector<double> probeHistory(int* nrecords) {
        std::vector<double> out(size_t(probe_cap_) * size_t(nprobes_ > 0 ? nprobes_ : 1) * 6);
        int n = 0;
        check(nw_dg_probe_history(s_, out.data(), probe_cap_, &n));
        out.resize(size_t(n) * size_t(nprobes_) * 6);
        if (nrecords) *nrecords = n;
        return out;
    }
    nw_opts& options() { return opts_; }
    // Extension: launch-shape autotuning (nw_dg_tune, nw_dg_segment_option,
    // nw_dg_set_segment_option).  Results never depend on the choices.
    void tune() { check(nw_dg_tune(s_)); }
    void segmentOption(int seg, int* option, bool* matrixVolume = nullptr, bool* streamK = nullptr) const {
        int o = 0, mv = 0, sk = 0;
        check(nw_dg_segment_option(s_, seg, &o, &mv, &sk));
        if (option) *option = o;
        if (matrixVolume) *matrixVolume = mv != 0;
        if (streamK) *streamK = sk != 0;
    }
    void setSegmentOption(int seg, int option, int streamK = -1) {
        check(nw_dg_set_segment_option(s_, seg, option, streamK));
    }
    // Extension: walk direction of the fused stages (nw_dg_set_walk / nw_dg_walk).
    void setWalk(int mode) { check(nw_dg_set_walk(s_, mode)); }
    bool alternatingWalk() const {
        int a = 0;
        check(nw_dg_walk(s_, &a));
        return a != 0;
    }

   private:
    nw_opts opts_;
    nw_dg* s_ = nullptr;
    int nprobes_ = 0, probe_cap_ = 1;
};

using InitialConditionFn = std::function<void(double x, double y, double z, double& Ex, double& Ey, double& Ez,
                                              double& Hx, double& Hy, double& Hz)>;  // case.hpp:21-23

// Hdf5Writer (hdf5_writer.hpp:22-120) as the reference runs it without the
// HDF5 library: <dir>/<stem>.bin, <stem>_step_<n>.vtu and <stem>.pvd.
class FieldWriter {
   public:
    struct FieldSaveOptions {  // hdf5_writer.hpp:56-67 without the derived curl/div fields
        bool saveE = true, saveH = true, saveMagnitudeE = false, saveMagnitudeH = false, saveEnergyDensity = false;
        int flags() const {
            return (saveE ? NW_SAVE_E : 0) | (saveH ? NW_SAVE_H : 0) | (saveMagnitudeE ? NW_SAVE_MAGNITUDE_E : 0) |
                   (saveMagnitudeH ? NW_SAVE_MAGNITUDE_H : 0) | (saveEnergyDensity ? NW_SAVE_ENERGY_DENSITY : 0);
        }
    };
    FieldWriter() = default;
    FieldWriter(const FieldWriter&) = delete;
    FieldWriter& operator=(const FieldWriter&) = delete;
    ~FieldWriter() { close(); }
    void setFieldSaveOptions(const FieldSaveOptions& o) { opt_ = o; }
    bool initialize(const std::string& h5Path, const Mesh& mesh, bool enableXdmf = true, bool exportContinuous = false) {
        (void)enableXdmf;  // the XDMF descriptor belongs to the HDF5 archive
        close();
        return nw_fields_open(h5Path.c_str(), mesh.handle(), exportContinuous, opt_.flags(), &f_) == NW_OK;
    }
    bool writeStep(int step, double time, const double* state, int npts) {
        return f_ && nw_fields_write(f_, step, time, state, 6 * size_t(npts)) == NW_OK;
    }
    void close() {
        nw_fields_close(f_);
        f_ = nullptr;
    }

   private:
    FieldSaveOptions opt_;
    nw_fields* f_ = nullptr;
};

struct Config {
    int order = 3;           // GLL points per direction (config.hpp:24)
    double cfl = -1.0;       // < 0: automatic
    double dt = -1.0;        // < 0: from cfl
    int numSteps = 100;
    int maxSteps = 100;
    bool hasExplicitMaxSteps = false;  // a step count was given (config.hpp:29, 182-185)
    double finalTime = -1.0;
    int outputFreq = 10;
    double c0 = 1.0;         // 0 central, 1 upwind
    int nelx = 0, nely = 0, nelz = 0;
    double xmin = -1.0, xmax = 1.0, ymin = -1.0, ymax = 1.0, zmin = -1.0, zmax = 1.0;
    bool periodicX = false, periodicY = false, periodicZ = false;
    // extensions
    int precision = NW_PREC_FP64;
    int pipeline = NW_PIPE_FUSED;
    double jitter = 0.0;
    uint64_t seed = 0;
    // probes (config.hpp:46, 57, 291-322): absolute points, points relative to
    // the box centre in units of its edge lengths, or num_probes generated ones
    std::vector<std::array<double, 3>> probes, probeRel;
    int numProbes = 0;
    int numModes = 1;                  // config "num_modes" (probes.json; modes of the bloch wave)
    // built-in initial conditions (case.cpp:247-372); NaN = the reference's default
    double pulseSigma = 20.0;
    double carrierK = std::nan("");
    double angleDeg = 0.0, kStart = 0.0, kEnd = 0.0;
    double packetSigma = 0.30;
    double packetX0 = std::nan("");
    std::string outputDir = "output";  // "none" keeps probe records in memory only
    // field snapshots (config.hpp:32-48, 193-226): Hdf5Writer's VTU/PVD/.bin files
    // when exportFields is on and exportFormat is "hdf5" or "vtk"; saveFreq <= 0
    // means outputFreq
    bool exportFields = false;
    std::string exportFormat = "hdf5";
    int saveFreq = -1;
    bool exportContinuousVtk = false;
    // the registry entry "wave_type" (Config::set, config.hpp:127-129): the built-in
    // initial condition when Case::setWaveType was not called (case.cpp:163)
    std::string waveType;

    // Step limit as Case::preprocess settles it (case.cpp:429-436 after
    // config.hpp:446-450): a non-default or explicit step count wins; a positive
    // finalTime with no step count given runs to finalTime (-1: no limit).
    int effectiveMaxSteps() const {
        if (maxSteps != 100) return maxSteps;
        if (numSteps != 100) return numSteps;
        if (hasExplicitMaxSteps || finalTime <= 0.0) return 100;
        return -1;
    }
};

class Case;
using PostprocessingFn = std::function<void(Case& c)>;  // case.hpp:34

class Case {
   public:
    Case() = default;
    void setConfig(const Config& c) { config_ = c; }
    const Config& config() const { return config_; }
    Config& config() { return config_; }
    void setInitialCondition(InitialConditionFn fn) { ic_ = std::move(fn); }
    void setInitialConditionHook(InitialConditionFn fn) { setInitialCondition(std::move(fn)); }
    void setPostprocessingHook(PostprocessingFn fn) { post_ = std::move(fn); }
    // Built-in initial condition used when no hook is set (case.cpp:59-62, 164-176,
    // 247-372): "gaussian", "bloch"/"periodic", "multimode", "packet"; "3dboxper"
    // builds a [-1, 1]^3 box on [0, 2 pi]^3 instead (Mesh::rescale, box meshes).
    void setWaveType(const std::string& w) { waveType_ = w; }
    // setWaveType's value, else the config's (case.cpp:163)
    const std::string& waveType() const { return waveType_.empty() ? config_.waveType : waveType_; }
    void setFieldSaveOptions(const FieldWriter::FieldSaveOptions& o) { saveOptions_ = o; }
    void addProbe(double x, double y, double z) { customProbes_.push_back({{x, y, z}}); }
    void addRelativeProbe(double rx, double ry, double rz) { customRel_.push_back({{rx, ry, rz}}); }

    // Probe coordinates in the order of case.cpp:383-422.
    std::vector<std::array<double, 3>> probeList() const {
        const Config& c = config_;
        std::vector<std::array<double, 3>> p = c.probes;
        p.insert(p.end(), customProbes_.begin(), customProbes_.end());
        const double lx = c.xmax - c.xmin, ly = c.ymax - c.ymin, lz = c.zmax - c.zmin;
        const double xc = 0.5 * (c.xmin + c.xmax), yc = 0.5 * (c.ymin + c.ymax), zc = 0.5 * (c.zmin + c.zmax);
        std::vector<std::array<double, 3>> rel = c.probeRel;
        rel.insert(rel.end(), customRel_.begin(), customRel_.end());
        for (const auto& r : rel) p.push_back({{xc + r[0] * lx, yc + r[1] * ly, zc + r[2] * lz}});
        if (p.empty() && c.numProbes > 0) {
            if (c.nelx == 1 && c.nely == 1 && c.nelz == 1) {
                const std::vector<double> z = mesh_->getGllZ();
                for (int q = 0; q < std::min(c.numProbes, mesh_->getN()); ++q)
                    p.push_back({{c.xmin + 0.5 * (1.0 + z[q]) * lx, yc, zc}});
            } else {
                const double x0 = c.xmin + 0.05 * lx, x1 = c.xmax - 0.05 * lx;
                const double dx = c.numProbes > 1 ? (x1 - x0) / (c.numProbes - 1) : 0.0;
                for (int q = 0; q < c.numProbes; ++q) p.push_back({{x0 + q * dx, yc, zc}});
            }
        }
        return p;
    }

    // case.cpp:98-524 (box meshes; dt rule :196-210; IC hook :227-246)
    void preprocess() {
        mesh_.reset(new Mesh(config_.order));
        const bool boxGiven = config_.nelx > 0 && config_.nely > 0 && config_.nelz > 0;
        double b[6] = {config_.xmin, config_.xmax, config_.ymin, config_.ymax, config_.zmin, config_.zmax};
        if (!boxGiven) b[0] = b[2] = b[4] = -1.0, b[1] = b[3] = b[5] = 1.0;
        if (waveType() == "3dboxper" && std::abs(b[0] + 1.0) < 1e-3 && std::abs(b[1] - 1.0) < 1e-3)
            b[0] = b[2] = b[4] = 0.0, b[1] = b[3] = b[5] = 2.0 * M_PI;  // case.cpp:165-176
        if (boxGiven)
            mesh_->createBoxMesh(config_.nelx, config_.nely, config_.nelz, b[0], b[1], b[2], b[3], b[4], b[5],
                                 config_.periodicX, config_.periodicY, config_.periodicZ, config_.jitter,
                                 config_.seed);
        else
            mesh_->createBoxMesh(3, 3, 3, b[0], b[1], b[2], b[3], b[4], b[5]);
        const double dxmin = mesh_->computeMinNodeDistance();
        if (config_.dt > 0.0) {
            dt_ = config_.dt;
            cfl_ = dt_ / dxmin;
        } else if (config_.cfl > 0.0) {
            cfl_ = config_.cfl;
            dt_ = cfl_ * dxmin;
        } else {
            cfl_ = mesh_->computeAutomaticCFL(0.4);
            dt_ = mesh_->computeAutomaticDt(1.0, 0.4);
        }
        const int npts = mesh_->getTotalPoints();
        state_.assign(6 * size_t(npts), 0.0);
        numModesUsed_ = config_.numModes;
        modeK_.clear();
        if (ic_) {
            const auto x = mesh_->getCoordX(), y = mesh_->getCoordY(), z = mesh_->getCoordZ();
            for (int k = 0; k < npts; ++k)
                ic_(x[k], y[k], z[k], state_[k], state_[npts + k], state_[2 * size_t(npts) + k],
                    state_[3 * size_t(npts) + k], state_[4 * size_t(npts) + k], state_[5 * size_t(npts) + k]);
        } else if (!waveType().empty()) {
            builtinIc();
        }
        currentTime_ = 0.0;
        currentStep_ = 0;
    }

    // case.cpp:529-660 without the per-step download: probes are recorded on the
    // device every step, diagnostics are reduced on the device at output
    // frequency, and the state comes back once.
    void simulate() {
        nw_opts o;
        nw_opts_default(&o);
        o.precision = config_.precision;
        o.pipeline = config_.pipeline;
        solver_.reset(new DgSolver(o));
        if (!solver_->initialize(*mesh_, config_.c0)) throw Error(NW_ERR_STATE, nw_last_error());
        solver_->uploadState(state_.data(), state_.size());
        const int maxSteps = config_.effectiveMaxSteps(), outFreq = config_.outputFreq > 1 ? config_.outputFreq : 1;
        const double finalTime = config_.finalTime;
        int step = 0;
        history_.clear();
        probeSteps_.clear();
        probeTimes_.clear();
        probeValues_.clear();
        probePoints_ = probeList();
        std::vector<std::pair<int, double>> pending;
        const int probeCap = std::min(outFreq + 1, kProbeCapacity);
        std::FILE* csv = nullptr;
        if (!probePoints_.empty()) {
            solver_->setProbes(*mesh_, probePoints_, probeCap);
            const std::string& d = config_.outputDir;
            if (!d.empty() && d != "none" && d != "None" && d != "NONE") {  // probe.cpp:43-48, 117-133
                ::mkdir(d.c_str(), 0755);
                csv = std::fopen((d + "/probe_history.csv").c_str(), "w");
                if (csv) {
                    std::fputs("step,time", csv);
                    for (size_t q = 1; q <= probePoints_.size(); ++q)
                        std::fprintf(csv, ",Ex_%zu,Ey_%zu,Ez_%zu,Hx_%zu,Hy_%zu,Hz_%zu", q, q, q, q, q, q);
                    std::fputs("\n", csv);
                    writeProbesJson(d + "/probes.json");
                }
            }
            solver_->sampleProbes();  // t = 0 (case.cpp:473-474)
            pending.push_back({0, 0.0});
        }
        // field snapshots at t = 0, every saveFreq steps and at the last step
        // (case.cpp:438-464, 603-625); the state is downloaded for those steps only
        writer_.reset();
        {
            const std::string& d = config_.outputDir;
            const bool saveDir = !d.empty() && d != "none" && d != "None" && d != "NONE";
            if (saveDir && config_.exportFields && (config_.exportFormat == "hdf5" || config_.exportFormat == "vtk")) {
                ::mkdir(d.c_str(), 0755);
                writer_.reset(new FieldWriter());
                writer_->setFieldSaveOptions(saveOptions_);
                if (!writer_->initialize(d + "/fields.h5", *mesh_, true, config_.exportContinuousVtk))
                    throw Error(NW_ERR_ARG, nw_last_error());
                writer_->writeStep(0, 0.0, state_.data(), mesh_->getTotalPoints());
            }
        }
        const int saveFreq = std::max(1, config_.saveFreq > 0 ? config_.saveFreq : outFreq);
        StateVector snap(writer_ ? state_.size() : 0);
        // The reference's step schedule (case.cpp:562-583); the steps between two
        // outputs with equal dt go to the device as one multi-step launch.
        double chunkDt = 0.0;
        int chunkN = 0;
        auto flush = [&]() {
            if (chunkN) solver_->run(chunkDt, chunkN);
            chunkN = 0;
        };
        while (true) {
            if (maxSteps > 0 && step >= maxSteps) break;
            if (finalTime > 0.0 && currentTime_ >= finalTime - 1e-13) break;
            double dt = dt_;
            if (finalTime > 0.0 && currentTime_ + dt > finalTime) dt = finalTime - currentTime_;
            if (chunkN && dt != chunkDt) flush();
            chunkDt = dt;
            ++chunkN;
            ++step;
            currentTime_ += dt;
            currentStep_ = step;
            if (!probePoints_.empty()) {
                pending.push_back({step, currentTime_});
                if (static_cast<int>(pending.size()) == probeCap) {  // device history full: fetch now
                    flush();
                    fetchProbes(pending, csv);
                }
            }
            const bool last =
                (maxSteps > 0 && step >= maxSteps) || (finalTime > 0.0 && currentTime_ >= finalTime - 1e-13);
            if (writer_ && (step % saveFreq == 0 || last)) {
                flush();
                solver_->downloadState(snap.data(), snap.size());
                if (!writer_->writeStep(step, currentTime_, snap.data(), mesh_->getTotalPoints()))
                    throw Error(NW_ERR_ARG, nw_last_error());
            }
            if (step % outFreq == 0 || last) {
                flush();
                history_.push_back(solver_->diagnostics());
                fetchProbes(pending, csv);
            }
        }
        flush();
        fetchProbes(pending, csv);
        if (csv) std::fclose(csv);
        solver_->downloadState(state_.data(), state_.size());
    }

    void postprocess() {
        writer_.reset();  // case.cpp:689-693
        if (post_) post_(*this);
        if (solver_) solver_->finalize();
    }
    void run() {
        preprocess();
        simulate();
        postprocess();
    }

    double computeTotalEnergy() const { return solver_->diagnostics().energy; }
    double getMaxE() const { return solver_->diagnostics().max_e; }
    double getMaxH() const { return solver_->diagnostics().max_h; }

    const Mesh& mesh() const { return *mesh_; }
    DgSolver& solver() { return *solver_; }
    const StateVector& state() const { return state_; }
    double dt() const { return dt_; }
    double cfl() const { return cfl_; }
    double currentTime() const { return currentTime_; }
    int currentStep() const { return currentStep_; }
    const std::vector<nw_diag>& history() const { return history_; }
    // probe records: step, time and probes x 6 values (Ex Ey Ez Hx Hy Hz) each
    const std::vector<std::array<double, 3>>& probePoints() const { return probePoints_; }
    const std::vector<int>& probeSteps() const { return probeSteps_; }
    const std::vector<double>& probeTimes() const { return probeTimes_; }
    const std::vector<double>& probeValues() const { return probeValues_; }

    static constexpr int kProbeCapacity = 4096;  // device probe records kept between fetches

   private:
    // The reference's built-in fields at the nodes (case.cpp:247-372).
    void builtinIc() {
        const Config& c = config_;
        const auto x = mesh_->getCoordX(), y = mesh_->getCoordY(), z = mesh_->getCoordZ();
        const size_t n = x.size();
        double *Ez = &state_[2 * n], *Hx = &state_[3 * n], *Hy = &state_[4 * n];
        const double lx = c.xmax - c.xmin, ly = c.ymax - c.ymin;
        const std::string& w = waveType();
        if (w == "bloch" || w == "periodic") {
            const double a = c.angleDeg, rad = a * M_PI / 180.0;
            double kx, ky;
            if (std::abs(a - 0.0) < 1e-4) kx = 2.0 * M_PI / lx, ky = 0.0;
            else if (std::abs(a - 45.0) < 1e-4) kx = 2.0 * M_PI / lx, ky = 2.0 * M_PI / ly;
            else if (std::abs(a - 90.0) < 1e-4) kx = 0.0, ky = 2.0 * M_PI / ly;
            else if (std::abs(a - 26.6) < 1.0 || std::abs(a - 30.0) < 5.0) kx = 4.0 * M_PI / lx, ky = 2.0 * M_PI / ly;
            else if (std::abs(a - 18.4) < 1.0 || std::abs(a - 15.0) < 5.0) kx = 6.0 * M_PI / lx, ky = 2.0 * M_PI / ly;
            else {
                const int mx = std::max(1, (int)std::round(std::cos(rad) * 4.0));
                const int my = (int)std::round(std::sin(rad) * 4.0);
                kx = 2.0 * M_PI * mx / lx, ky = 2.0 * M_PI * my / ly;
            }
            const double kMag = std::sqrt(kx * kx + ky * ky), ca = kx / kMag, sa = ky / kMag;
            const int nm = std::min(c.numModes, std::max(1, c.nelx * (mesh_->getN() - 1) / 2));
            if (c.kStart > 0.0 && c.kEnd > 0.0)
                for (int m = 0; m < nm; ++m) modeK_.push_back(nm > 1 ? c.kStart + m * (c.kEnd - c.kStart) / (nm - 1) : c.kStart);
            else
                for (int m = 1; m <= nm; ++m) modeK_.push_back(m * kMag);
            const bool carrier = !std::isnan(c.carrierK);
            const double ck = carrier ? c.carrierK : 2.0 * M_PI / lx;
            for (size_t k = 0; k < n; ++k) {
                const double xt = x[k] - c.xmin, yt = y[k] - c.ymin;
                double e = 0.0;
                if (nm > 1) {
                    for (double km : modeK_) e += std::cos(km * (ca * xt + sa * yt));
                    e /= std::sqrt((double)nm);
                } else {
                    e = std::cos((carrier && std::abs(a) < 1e-4) ? ck * xt : kx * xt + ky * yt);
                }
                Ez[k] = e;
                Hx[k] = sa * e;
                Hy[k] = -ca * e;
            }
            numModesUsed_ = nm;
        } else if (w == "multimode") {
            for (size_t k = 0; k < n; ++k) {
                double e = 0.0;
                for (int m = 1; m <= 8; ++m) e += (1.0 / m) * std::sin(m * M_PI / lx * (x[k] - c.xmin));
                Ez[k] = e * std::sin(M_PI * (y[k] - c.ymin) / ly);
            }
        } else if (w == "packet") {
            const double ck = std::isnan(c.carrierK) ? 4.0 * M_PI : c.carrierK;
            const double x0 = std::isnan(c.packetX0) ? c.xmin + 0.25 * lx : c.packetX0;
            for (size_t k = 0; k < n; ++k) {
                const double d = x[k] - x0;
                Ez[k] = std::exp(-(d * d) / (2.0 * c.packetSigma * c.packetSigma)) * std::sin(ck * (x[k] - c.xmin)) *
                        std::sin(M_PI * (y[k] - c.ymin) / ly);
            }
        } else if (w == "gaussian") {
            for (size_t k = 0; k < n; ++k) Ez[k] = std::exp(-c.pulseSigma * (x[k] * x[k] + y[k] * y[k] + z[k] * z[k]));
        }
    }

    // probes.json as ProbeManager::init writes it (probe.cpp:135-173): placed by
    // interpolation, so nodeIdx -1, actual = target, offset 0; the stream turns
    // scientific (precision 8) at the separation line, which exists only for two
    // or more probes, so a single probe prints in the default format.
    void writeProbesJson(const std::string& path) const {
        std::FILE* f = std::fopen(path.c_str(), "w");
        if (!f) return;
        const auto& p = probePoints_;
        const bool sci = p.size() >= 2;
        const char* num = sci ? "%.8e" : "%g";
        std::fprintf(f, "{\n  \"order\": %d,\n  \"num_modes\": %d,\n  \"num_probes\": %zu,\n", mesh_->getN(),
                     numModesUsed_, p.size());
        if (sci) {
            const double dx = p[1][0] - p[0][0], dy = p[1][1] - p[0][1], dz = p[1][2] - p[0][2];
            const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
            std::fprintf(f,
                         "  \"separation_distance\": %.8e,\n  \"delta_x\": %.8e,\n  \"delta_y\": %.8e,\n"
                         "  \"delta_z\": %.8e,\n  \"domain_Lx\": %.8e,\n  \"elements_x\": %d,\n",
                         dist, dx, dy, dz, config_.xmax - config_.xmin, config_.nelx);
            if (!modeK_.empty()) {
                std::fputs("  \"k_modes\": [", f);
                for (size_t m = 0; m < modeK_.size(); ++m)
                    std::fprintf(f, "%.8e%s", modeK_[m], m + 1 < modeK_.size() ? ", " : "");
                std::fputs("],\n", f);
            }
        }
        std::fputs("  \"probes\": [\n", f);
        for (size_t q = 0; q < p.size(); ++q) {
            std::fprintf(f, "    {\n      \"id\": %zu,\n      \"nodeIdx\": -1,\n", q + 1);
            for (const char* key : {"target", "actual"}) {
                std::fprintf(f, "      \"%s\": [", key);
                for (int d = 0; d < 3; ++d) {
                    std::fprintf(f, num, p[q][d]);
                    std::fputs(d < 2 ? ", " : "],\n", f);
                }
            }
            std::fputs("      \"offset\": ", f);
            std::fprintf(f, num, 0.0);
            std::fprintf(f, "\n    }%s\n", q + 1 < p.size() ? "," : "");
        }
        std::fputs("  ]\n}\n", f);
        std::fclose(f);
    }

    void fetchProbes(std::vector<std::pair<int, double>>& pending, std::FILE* csv) {
        if (pending.empty()) return;
        int n = 0;
        const std::vector<double> h = solver_->probeHistory(&n);
        if (n != static_cast<int>(pending.size())) throw Error(NW_ERR_STATE, "probe history is short");
        const size_t per = probePoints_.size() * 6;
        for (int r = 0; r < n; ++r) {
            probeSteps_.push_back(pending[r].first);
            probeTimes_.push_back(pending[r].second);
            probeValues_.insert(probeValues_.end(), h.begin() + long(r * per), h.begin() + long((r + 1) * per));
            if (csv) {  // std::scientific, setprecision(8) (probe.cpp:133, 196-214)
                std::fprintf(csv, "%d,%.8e", pending[r].first, pending[r].second);
                for (size_t v = 0; v < per; ++v) std::fprintf(csv, ",%.8e", h[r * per + v]);
                std::fputs("\n", csv);
            }
        }
        pending.clear();
    }

    Config config_;
    std::unique_ptr<Mesh> mesh_;
    std::unique_ptr<DgSolver> solver_;
    std::unique_ptr<FieldWriter> writer_;
    FieldWriter::FieldSaveOptions saveOptions_;
    StateVector state_;
    InitialConditionFn ic_;
    PostprocessingFn post_;
    std::vector<nw_diag> history_;
    std::vector<std::array<double, 3>> customProbes_, customRel_, probePoints_;
    std::vector<int> probeSteps_;
    std::string waveType_;
    int numModesUsed_ = 1;
    std::vector<double> modeK_;
    std::vector<double> probeTimes_, probeValues_;
    double dt_ = 0.0, cfl_ = 0.0, currentTime_ = 0.0;
    int currentStep_ = 0;
};

}  // namespace nekwave_amd

#endif  // NEKWAVE_AMD_HPP
