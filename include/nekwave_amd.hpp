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

#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
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
    nw_opts& options() { return opts_; }

   private:
    nw_opts opts_;
    nw_dg* s_ = nullptr;
};

using InitialConditionFn = std::function<void(double x, double y, double z, double& Ex, double& Ey, double& Ez,
                                              double& Hx, double& Hy, double& Hz)>;  // case.hpp:21-23

struct Config {
    int order = 3;           // GLL points per direction (config.hpp:24)
    double cfl = -1.0;       // < 0: automatic
    double dt = -1.0;        // < 0: from cfl
    int numSteps = 100;
    int maxSteps = 100;
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

    // case.cpp:98-524 (box meshes; dt rule :196-210; IC hook :227-246)
    void preprocess() {
        mesh_.reset(new Mesh(config_.order));
        if (config_.nelx > 0 && config_.nely > 0 && config_.nelz > 0)
            mesh_->createBoxMesh(config_.nelx, config_.nely, config_.nelz, config_.xmin, config_.xmax, config_.ymin,
                                 config_.ymax, config_.zmin, config_.zmax, config_.periodicX, config_.periodicY,
                                 config_.periodicZ, config_.jitter, config_.seed);
        else
            mesh_->createBoxMesh(3, 3, 3, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0);
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
        if (ic_) {
            const auto x = mesh_->getCoordX(), y = mesh_->getCoordY(), z = mesh_->getCoordZ();
            for (int k = 0; k < npts; ++k)
                ic_(x[k], y[k], z[k], state_[k], state_[npts + k], state_[2 * size_t(npts) + k],
                    state_[3 * size_t(npts) + k], state_[4 * size_t(npts) + k], state_[5 * size_t(npts) + k]);
        }
        currentTime_ = 0.0;
        currentStep_ = 0;
    }

    // case.cpp:529-660 without the per-step download: diagnostics are reduced
    // on the device at output frequency and the state comes back once.
    void simulate() {
        nw_opts o;
        nw_opts_default(&o);
        o.precision = config_.precision;
        o.pipeline = config_.pipeline;
        solver_.reset(new DgSolver(o));
        if (!solver_->initialize(*mesh_, config_.c0)) throw Error(NW_ERR_STATE, nw_last_error());
        solver_->uploadState(state_.data(), state_.size());
        const int maxSteps = config_.maxSteps, outFreq = config_.outputFreq > 1 ? config_.outputFreq : 1;
        const double finalTime = config_.finalTime;
        int step = 0;
        history_.clear();
        while (true) {
            if (maxSteps > 0 && step >= maxSteps) break;
            if (finalTime > 0.0 && currentTime_ >= finalTime - 1e-13) break;
            double dt = dt_;
            if (finalTime > 0.0 && currentTime_ + dt > finalTime) dt = finalTime - currentTime_;
            ++step;
            solver_->step(dt, currentTime_);
            currentTime_ += dt;
            currentStep_ = step;
            const bool last =
                (maxSteps > 0 && step >= maxSteps) || (finalTime > 0.0 && currentTime_ >= finalTime - 1e-13);
            if (step % outFreq == 0 || last) history_.push_back(solver_->diagnostics());
        }
        solver_->downloadState(state_.data(), state_.size());
    }

    void postprocess() {
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

   private:
    Config config_;
    std::unique_ptr<Mesh> mesh_;
    std::unique_ptr<DgSolver> solver_;
    StateVector state_;
    InitialConditionFn ic_;
    PostprocessingFn post_;
    std::vector<nw_diag> history_;
    double dt_ = 0.0, cfl_ = 0.0, currentTime_ = 0.0;
    int currentStep_ = 0;
};

}  // namespace nekwave_amd

#endif  // NEKWAVE_AMD_HPP
