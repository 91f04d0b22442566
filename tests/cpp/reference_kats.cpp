// reference_kats.cpp -- the reference's own C++ test cases, restated against
// the C++ host interface (include/nekwave_amd.hpp) and run on the device.
//
//   MaxwellStabilityTest  tests/nekwave_test.cpp:212-293 (DgSolver directly)
//   3dboxpec              tests/3dboxpec_test/3dboxpec_test.cpp:11-151 (Case
//                         lifecycle with IC and post-processing hooks; the
//                         .rea mesh restated as a 3x3x3 box on [-1,1]^3)
//
// Exit status 0 when every check passes.  Usage: reference_kats [fp64|fp32|host]
// ("host": only the checks that need no device -- built-in fields, step limit).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nekwave_amd.hpp"

using namespace nekwave_amd;

static int g_failed = 0;
#define EXPECT_TRUE(c)                                                        \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "  FAILED: %s at line %d\n", #c, __LINE__); \
            ++g_failed;                                                       \
            return;                                                           \
        }                                                                     \
    } while (0)

static int g_prec = NW_PREC_FP64;

static void MaxwellStabilityTest() {
    std::printf("[RUN] MaxwellStabilityTest\n");
    const int N = 4;
    Mesh mesh(N, 1);
    mesh.createBoxMesh(2, 2, 2, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5);
    nw_opts o;
    nw_opts_default(&o);
    o.precision = g_prec;
    DgSolver solver(o);
    EXPECT_TRUE(solver.initialize(mesh, 1.0));
    const int npts = mesh.getTotalPoints();
    StateVector state(6 * npts, 0.0);
    const auto x = mesh.getCoordX(), y = mesh.getCoordY(), z = mesh.getCoordZ();
    for (int i = 0; i < npts; ++i) state[2 * npts + i] = std::exp(-50.0 * (x[i] * x[i] + y[i] * y[i] + z[i] * z[i]));
    const auto J = mesh.getJac();
    const auto w = mesh.getGllW();
    auto energy = [&](const StateVector& u) {
        double E = 0.0;
        const int Np = N * N * N;
        for (int e = 0; e < mesh.getNumElements(); ++e)
            for (int k = 0; k < N; ++k)
                for (int j = 0; j < N; ++j)
                    for (int i = 0; i < N; ++i) {
                        const int p = e * Np + k * N * N + j * N + i;
                        const double dV = J[p] * w[i] * w[j] * w[k];
                        double s = 0.0;
                        for (int c = 0; c < 6; ++c) s += u[c * npts + p] * u[c * npts + p];
                        E += 0.5 * s * dV;
                    }
        return E;
    };
    const double e0 = energy(state);
    EXPECT_TRUE(e0 > 0.0);
    solver.uploadState(state.data(), state.size());
    const double slack = g_prec == NW_PREC_FP64 ? 1e-12 : 1e-6 * e0;
    double t = 0.0;
    for (int s = 0; s < 5; ++s) {
        solver.step(0.005, t);
        t += 0.005;
        solver.downloadState(state.data(), state.size());
        const double e = energy(state);
        std::printf("  step %d energy %.15e\n", s + 1, e);
        EXPECT_TRUE(e <= e0 + slack);
        EXPECT_TRUE(!std::isnan(e) && !std::isinf(e));
    }
    solver.finalize();
    std::printf("  PASSED\n");
}

static void BoxPecCase() {
    std::printf("[RUN] 3dboxpec (Case lifecycle)\n");
    Case c;
    Config cfg;
    cfg.order = 4;
    cfg.nelx = cfg.nely = cfg.nelz = 3;
    cfg.c0 = 0.0;
    cfg.maxSteps = 50;
    cfg.outputFreq = 10;
    cfg.precision = g_prec;
    char dirTemplate[] = "/tmp/nw_kat_probesXXXXXX";
    const char* dir = mkdtemp(dirTemplate);
    cfg.outputDir = dir ? dir : "none";
    cfg.probeRel = {{{0.1, 0.2, -0.1}}};
    c.setConfig(cfg);
    c.addProbe(0.3, -0.4, 0.25);
    c.setInitialConditionHook([](double x, double y, double z, double& ex, double& ey, double& ez, double& hx,
                                 double& hy, double& hz) {
        const double is6 = 1.0 / std::sqrt(6.0);
        ex = ey = ez = 0.0;
        hx = -std::sin(M_PI * x) * std::cos(M_PI * y) * std::cos(M_PI * z) * is6;
        hy = -std::cos(M_PI * x) * std::sin(M_PI * y) * std::cos(M_PI * z) * is6;
        hz = 2.0 * std::cos(M_PI * x) * std::cos(M_PI * y) * std::sin(M_PI * z) * is6;
    });
    double rel = 1.0;
    c.setPostprocessingHook([&](Case& cs) {
        const double t = cs.currentTime();
        const auto& m = cs.mesh();
        const auto& st = cs.state();
        const int npts = m.getTotalPoints(), Np = m.getNumPointsPerElement();
        const auto x = m.getCoordX(), y = m.getCoordY(), z = m.getCoordZ(), w3 = m.getW3(), jac = m.getJac();
        const double om = M_PI * std::sqrt(3.0), th = std::cos(om * t) / std::sqrt(6.0), te = std::sin(om * t) / std::sqrt(2.0);
        double num = 0.0, den = 0.0;
        for (int idx = 0; idx < npts; ++idx) {
            const double dV = jac[idx] * w3[idx % Np];
            const double sx = std::sin(M_PI * x[idx]), cx = std::cos(M_PI * x[idx]);
            const double sy = std::sin(M_PI * y[idx]), cy = std::cos(M_PI * y[idx]);
            const double sz = std::sin(M_PI * z[idx]), cz = std::cos(M_PI * z[idx]);
            const double ex[6] = {-cx * sy * sz * te, sx * cy * sz * te, 0.0, -sx * cy * cz * th, -cx * sy * cz * th,
                                  2.0 * cx * cy * sz * th};
            for (int k = 0; k < 6; ++k) {
                const double d = st[k * size_t(npts) + idx] - ex[k];
                num += d * d * dV;
                den += ex[k] * ex[k] * dV;
            }
        }
        rel = std::sqrt(num / den);
        std::printf("  t = %.5e  relative L2 = %.5e\n", t, rel);
    });
    c.run();
    EXPECT_TRUE(c.currentStep() == 50);
    EXPECT_TRUE(c.history().size() == 5);
    EXPECT_TRUE(rel <= 0.05);  // 3dboxpec_test.cpp:136
    // probes: t = 0 and every step recorded on the device, fetched every 10 steps
    EXPECT_TRUE(c.probePoints().size() == 2);
    EXPECT_TRUE(c.probeSteps().size() == 51 && c.probeSteps().back() == 50);
    EXPECT_TRUE(c.probeValues().size() == 51 * 2 * 6);
    EXPECT_TRUE(std::isfinite(c.probeValues().back()) && c.probeValues()[5] != 0.0);
    // probe_history.csv as ProbeManager writes it (probe.cpp:117-133, 192-216)
    EXPECT_TRUE(dir != nullptr);
    const std::string csvPath = std::string(dir) + "/probe_history.csv";
    std::FILE* f = std::fopen(csvPath.c_str(), "r");
    EXPECT_TRUE(f != nullptr);
    std::vector<std::string> lines;
    char buf[4096];
    while (std::fgets(buf, sizeof(buf), f)) lines.push_back(buf);
    std::fclose(f);
    std::remove(csvPath.c_str());
    const std::string jsonPath = std::string(dir) + "/probes.json";
    std::FILE* jf = std::fopen(jsonPath.c_str(), "r");
    EXPECT_TRUE(jf != nullptr);
    std::vector<std::string> jl;
    while (std::fgets(buf, sizeof(buf), jf)) jl.push_back(buf);
    std::fclose(jf);
    std::remove(jsonPath.c_str());
    std::remove(dir);
    EXPECT_TRUE(jl.size() > 10 && jl[1] == "  \"order\": 4,\n" && jl[3] == "  \"num_probes\": 2,\n");
    EXPECT_TRUE(jl[4].rfind("  \"separation_distance\": ", 0) == 0);
    EXPECT_TRUE(lines.size() == 52);
    EXPECT_TRUE(lines[0].rfind("step,time,Ex_1,Ey_1,Ez_1,Hx_1,Hy_1,Hz_1,Ex_2,", 0) == 0);
    char expect[64];
    std::snprintf(expect, sizeof(expect), "50,%.8e,", c.probeTimes().back());
    EXPECT_TRUE(lines.back().rfind(expect, 0) == 0);
    std::printf("  PASSED\n");
}

// Built-in Bloch wave (case.cpp:247-329): a plane wave (|H| = |E|), modes clamped
// to the grid's Nyquist limit, every cosine 1 at the box's corner line.
static void BuiltinBlochIc() {
    std::printf("[RUN] built-in bloch initial condition\n");
    Case c;
    Config cfg;
    cfg.order = 5;
    cfg.nelx = 4;
    cfg.nely = cfg.nelz = 2;
    cfg.xmin = 0.0, cfg.xmax = 2.0, cfg.ymin = 0.0, cfg.ymax = 1.0, cfg.zmin = 0.0, cfg.zmax = 1.0;
    cfg.numModes = 50;
    c.setConfig(cfg);
    c.setWaveType("bloch");
    c.preprocess();
    const auto& st = c.state();
    const size_t n = st.size() / 6;
    const auto x = c.mesh().getCoordX(), y = c.mesh().getCoordY();
    int corner = 0;
    for (size_t k = 0; k < n; ++k) {
        const double ez = st[2 * n + k], hx = st[3 * n + k], hy = st[4 * n + k];
        EXPECT_TRUE(std::abs(hx * hx + hy * hy - ez * ez) < 1e-13);
        if (x[k] == 0.0 && y[k] == 0.0) {
            EXPECT_TRUE(std::abs(ez - std::sqrt(8.0)) < 1e-13);  // 8 = 4 elements * 4 / 2 modes
            ++corner;
        }
    }
    EXPECT_TRUE(corner > 0);
    std::printf("  PASSED\n");
}

// Built-in fields by value (case.cpp:330-372), host only: -ln(Ez) / r^2 is the
// Gaussian's sigma at every node off the origin; the multimode sum at x = Lx/2,
// y = Ly/2 is 1 - 1/3 + 1/5 - 1/7.  The wave type may come from the config
// (Config::set "wave_type", case.cpp:163).
static void BuiltinFieldValues() {
    std::printf("[RUN] built-in gaussian and multimode values\n");
    Config cfg;
    cfg.order = 5;
    cfg.nelx = 4;
    cfg.nely = cfg.nelz = 2;
    cfg.pulseSigma = 3.0;
    cfg.waveType = "gaussian";
    Case g;
    g.setConfig(cfg);
    g.preprocess();
    EXPECT_TRUE(g.waveType() == "gaussian");
    {
        const auto& st = g.state();
        const size_t n = st.size() / 6;
        const auto x = g.mesh().getCoordX(), y = g.mesh().getCoordY(), z = g.mesh().getCoordZ();
        int checked = 0;
        for (size_t k = 0; k < n; ++k) {
            const double r2 = x[k] * x[k] + y[k] * y[k] + z[k] * z[k];
            if (r2 < 0.05) continue;
            EXPECT_TRUE(std::abs(-std::log(st[2 * n + k]) / r2 - 3.0) < 1e-12);
            ++checked;
        }
        EXPECT_TRUE(checked > 100);
    }
    cfg.waveType.clear();
    cfg.xmin = 0.0, cfg.xmax = 2.0, cfg.ymin = 0.0, cfg.ymax = 1.0, cfg.zmin = 0.0, cfg.zmax = 1.0;
    Case mm;
    mm.setConfig(cfg);
    mm.setWaveType("multimode");
    mm.preprocess();
    const auto& st = mm.state();
    const size_t n = st.size() / 6;
    const auto x = mm.mesh().getCoordX(), y = mm.mesh().getCoordY();
    int at = 0;
    for (size_t k = 0; k < n; ++k)
        if (std::abs(x[k] - 1.0) < 1e-12 && std::abs(y[k] - 0.5) < 1e-12) {
            EXPECT_TRUE(std::abs(st[2 * n + k] - (1.0 - 1.0 / 3 + 1.0 / 5 - 1.0 / 7)) < 1e-14);
            ++at;
        }
    EXPECT_TRUE(at > 0);
    std::printf("  PASSED\n");
}

// Step limit of Case::preprocess (case.cpp:429-436, config.hpp:446-450).
static void StepLimit() {
    std::printf("[RUN] step limit from num_steps / final_time\n");
    Config c;
    EXPECT_TRUE(c.effectiveMaxSteps() == 100);
    c.finalTime = 1.0;
    EXPECT_TRUE(c.effectiveMaxSteps() == -1);  // runs to finalTime
    c.hasExplicitMaxSteps = true;
    EXPECT_TRUE(c.effectiveMaxSteps() == 100);
    c.hasExplicitMaxSteps = false;
    c.maxSteps = 7;
    EXPECT_TRUE(c.effectiveMaxSteps() == 7);
    std::printf("  PASSED\n");
}

// Extension API of the C++ header: every launch shape (tile option, LSRK
// register policy), the autotuner's choice and backward walks give the bits of
// tile option 0.
static void LaunchShapesGiveTheSameBits() {
    std::printf("[RUN] LaunchShapesGiveTheSameBits\n");
    const int N = 6;
    Mesh mesh(N, 1);
    mesh.createBoxMesh(4, 3, 2, 0.0, 1.0, 0.0, 0.8, 0.0, 0.5);
    const int npts = mesh.getTotalPoints();
    const auto x = mesh.getCoordX(), y = mesh.getCoordY(), z = mesh.getCoordZ();
    StateVector init(6 * size_t(npts));
    for (int c = 0; c < 6; ++c)
        for (int i = 0; i < npts; ++i) init[size_t(c) * npts + i] = std::sin(3.0 * x[i] + (c + 1) * y[i]) * std::cos(2.0 * z[i] - c);
    const double dt = 0.002;
    StateVector ref;
    for (int run = 0; run < 8; ++run) {  // options 0..2 x k policy 0/1, the tuner's pick, backward walks
        nw_opts o;
        nw_opts_default(&o);
        o.precision = g_prec;
        DgSolver solver(o);
        EXPECT_TRUE(solver.initialize(mesh, 1.0));
        solver.uploadState(init.data(), init.size());
        const int opt = run / 2, sk = run % 2;
        if (run < 6) {
            solver.setSegmentOption(0, opt, sk);
            int got = -1;
            bool gotk = !sk;
            solver.segmentOption(0, &got, nullptr, &gotk);
            EXPECT_TRUE(got == opt && gotk == (sk != 0));
        } else if (run == 7) {
            solver.setWalk(1);
            EXPECT_TRUE(solver.alternatingWalk());
        } else {
            solver.tune();
            int got = -1;
            solver.segmentOption(0, &got);
            EXPECT_TRUE(got >= 0 && got < 3);
        }
        solver.run(dt, 5);
        StateVector out(6 * size_t(npts));
        solver.downloadState(out.data(), out.size());
        solver.finalize();
        if (run == 0) {
            ref = out;
            EXPECT_TRUE(out != init);
        } else {
            EXPECT_TRUE(out == ref);
        }
    }
    std::printf("  PASSED\n");
}

int main(int argc, char** argv) {
    if (argc > 1 && std::strcmp(argv[1], "host") == 0) {  // host-side checks only (no device)
        try {
            BuiltinFieldValues();
            StepLimit();
        } catch (const Error& e) {
            std::fprintf(stderr, "error %d: %s\n", e.status(), e.what());
            return 2;
        }
        std::printf("%s\n", g_failed ? "FAILED" : "ALL PASSED");
        return g_failed ? 1 : 0;
    }
    if (argc > 1 && std::strcmp(argv[1], "fp32") == 0) g_prec = NW_PREC_FP32;
    try {
        BuiltinFieldValues();
        StepLimit();
        MaxwellStabilityTest();
        BuiltinBlochIc();
        BoxPecCase();
        LaunchShapesGiveTheSameBits();
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status(), e.what());
        return 2;
    }
    std::printf("%s\n", g_failed ? "FAILED" : "ALL PASSED");
    return g_failed ? 1 : 0;
}
