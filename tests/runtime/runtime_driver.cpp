// Runs the reference's PortableRuntime -- the forcing engine, the
// constant-stratification integration system, WVFixedStepRK4 and
// WVAdaptiveRK23 -- over whichever WVTransformConstantStratificationKernel it
// is linked with, and writes the results as named float64 arrays.
//
// TEST INFRASTRUCTURE.  Built twice from the same source:
//   oracle/_ref/runtime_driver_ref   with the reference kernel
//                                    (CompiledKernel/src, direct-DFT engine),
//                                    on the CPU; its output is committed as
//                                    tests/golden/runtime_*.npz;
//   oracle/_ref/hip_runtime/runtime_driver_hip with the MI355X kernel (the drop-in
//                                    adapter over libwvamd), on the GPU.
// The runtime sources are compiled unchanged in both; tests/test_gpu_runtime.py
// compares the two.
//
// usage: runtime_driver <output file>
// Output: records of (int32 name length, name, int64 count, count doubles).
#include "WaveVortexKernel/WVTransformConstantStratificationKernel.hpp"
#include "WaveVortexRuntime/WVConstantStratificationIntegrationSystem.hpp"
#include "WaveVortexRuntime/WVExtensionCatalog.hpp"
#include "WaveVortexRuntime/WVForcingEngine.hpp"
#include "WaveVortexRuntime/WVIntegrationState.hpp"
#include "WaveVortexRuntime/WVObserverContracts.hpp"
#include "WaveVortexRuntime/WVRungeKutta.hpp"
#include "WVReferenceFFTEngine.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

using namespace wavevortex;
using namespace wavevortex::runtime;

namespace {

void require(bool ok, const std::string& what) {
    if (!ok) throw std::runtime_error(what);
}
void require(const WVKernelStatus& st, const std::string& what) {
    if (!st) throw std::runtime_error(what + ": " + st.message);
}

class Output {
public:
    explicit Output(const char* path) : f_(std::fopen(path, "wb")) { require(f_ != nullptr, "cannot open output"); }
    ~Output() {
        if (f_) std::fclose(f_);
    }
    void put(const std::string& name, const double* data, std::size_t n) {
        const std::int32_t len = static_cast<std::int32_t>(name.size());
        const std::int64_t cnt = static_cast<std::int64_t>(n);
        std::fwrite(&len, sizeof(len), 1, f_);
        std::fwrite(name.data(), 1, name.size(), f_);
        std::fwrite(&cnt, sizeof(cnt), 1, f_);
        std::fwrite(data, sizeof(double), n, f_);
    }
    void put(const std::string& name, double v) { put(name, &v, 1); }
    void put(const std::string& name, const std::vector<WVComplex64>& v) {
        put(name, reinterpret_cast<const double*>(v.data()), 2 * v.size());
    }

private:
    std::FILE* f_;
};

WVTransformConstantStratificationConfiguration configuration(std::size_t Nx, std::size_t Ny, std::size_t Nz,
                                                             std::size_t Nj, bool hydrostatic) {
    WVTransformConstantStratificationConfiguration c;
    c.Nx = Nx; c.Ny = Ny; c.Nz = Nz; c.Nj = Nj;
    c.Lx = 50e3; c.Ly = 40e3; c.Lz = 1300.0;
    c.N0 = 5.2e-3; c.rho0 = 1025.0; c.g = 9.81;
    c.planetaryRadius = 6.371e6; c.rotationRate = 7.2921e-5; c.latitude = 33.0;
    c.isHydrostatic = hydrostatic; c.shouldAntialias = true;
    return c;
}

// A deterministic flow of equal energy per coefficient (total energy
// density e2 * count, about Lz u^2 / 2 for u ~ 1 cm/s) with smooth
// pseudo-random phases: |A| = sqrt(e2 / factor), factor the wave or
// geostrophic total-energy factor (the forms of the reference's adaptive
// tolerance, PortableRuntime/src/WVForcingEngine.cpp:126-160); at K = 0 the
// self-conjugacy of the transforms: Am = conj(Ap), A0 real
// (@WVTransformConstantStratification/transformUVEtaToWaveVortex.m:38-39).
struct State {
    WVShape2D shape;
    std::vector<WVComplex64> values;
    double t = 0.0, t0 = 0.0;
    State(const WVTransformConstantStratificationDescriptor& d, double u) : shape(d.spectralShape()), values(3 * shape.elementCount()) {
        const auto& c = d.configuration();
        const auto& vm = d.verticalModes();
        const std::size_t n = shape.elementCount(), Nj = shape.rows;
        const double e2 = 0.5 * c.Lz * u * u / static_cast<double>(3 * n);
        const double f2 = vm.coriolisFrequency * vm.coriolisFrequency, N02 = c.N0 * c.N0;
        for (std::size_t i = 0; i < n; ++i) {
            const std::size_t j = i % Nj, mode = i / Nj;
            const double Kh = d.fourierModes()[mode].Kh, m = vm.verticalWavenumber[j], h0 = vm.h0[j];
            double wave = 0.0, vortex = 0.0;
            if (Kh == 0.0)
                wave = j == 0 ? c.Lz : (c.isHydrostatic ? N02 : N02 - f2) / (c.g * m * m);
            else if (j > 0)
                wave = 2.0 * (c.isHydrostatic ? N02 / (c.g * m * m) : (N02 - f2) / (c.g * (m * m + Kh * Kh)));
            if (Kh > 0.0)
                vortex = h0 / (Kh * Kh + (j == 0 ? 0.0 : f2 / (c.g * h0)));
            else if (j > 0)
                vortex = c.g / 2.0;
            const double a = wave > 0.0 ? std::sqrt(e2 / wave) : 0.0;
            const double g = vortex > 0.0 ? std::sqrt(e2 / vortex) : 0.0;
            values[i] = {a * std::sin(0.17 * (i + 1)), a * std::cos(0.13 * (i + 2))};
            values[n + i] = {a * std::cos(0.11 * (i + 3)), a * std::sin(0.07 * (i + 4))};
            values[2 * n + i] = {g * std::sin(0.19 * (i + 5)), g * std::cos(0.05 * (i + 6))};
            if (Kh == 0.0) {
                values[n + i] = {values[i].real, -values[i].imag};
                values[2 * n + i].imag = 0.0;
            }
        }
    }
    WVMutableIntegrationState integration() {
        const std::size_t n = shape.elementCount();
        return {{t, t0, {{values.data(), shape}, {values.data() + n, shape}, {values.data() + 2 * n, shape}}}, nullptr, 0};
    }
    WVState view() const {
        const std::size_t n = shape.elementCount();
        return {t, t0, {{values.data(), shape}, {values.data() + n, shape}, {values.data() + 2 * n, shape}}};
    }
};

WVFrozenForcingEntry entry(const char* type, const char* name, WVForcingStage stage, std::uint8_t priority,
                           std::vector<WVPortableNamedValue> values = {}) {
    return {type, WVPortablePairContractVersion, name, stage, priority, 0, "driver",
            {"wave-vortex-forcing-configuration-v1", 1, std::move(values)}};
}

WVFrozenForcingSchedule nonlinearSchedule() {
    WVFrozenForcingSchedule s;
    s.entries.push_back(entry("WVNonlinearAdvection", "nonlinear", WVForcingStage::spatial, 127));
    return s;
}

// nonlinear advection with the other built-in spectral and bottom forcings
WVFrozenForcingSchedule forcedSchedule() {
    WVFrozenForcingSchedule s = nonlinearSchedule();
    s.entries.push_back(entry("WVAdaptiveDamping", "adaptive damping", WVForcingStage::spectral, 100));
    s.entries.push_back(entry("WVBetaPlanePVAdvection", "beta plane", WVForcingStage::spectral, 110));
    s.entries.push_back(entry("WVBottomFrictionLinear", "linear bottom friction", WVForcingStage::spatial, 200,
                              {{"r", {}, std::vector<double>{2.5e-7}}}));
    s.entries.push_back(entry("WVBottomFrictionQuadratic", "quadratic bottom friction", WVForcingStage::spatial, 210,
                              {{"Cd", {}, std::vector<double>{2e-3}}}));
    return s;
}

// nonlinear advection with pseudo-topographic wave generation (tidal flow
// over a seamount, ramped in) and fixed amplitudes held on a few
// coefficients of each family
WVFrozenForcingSchedule topographicSchedule(std::size_t Nx, std::size_t Ny) {
    WVFrozenForcingSchedule s = nonlinearSchedule();
    std::vector<double> height(Nx * Ny);
    for (std::size_t y = 0; y < Ny; ++y)
        for (std::size_t x = 0; x < Nx; ++x) {
            const double dx = (static_cast<double>(x) - 0.4 * Nx) / Nx, dy = (static_cast<double>(y) - 0.55 * Ny) / Ny;
            height[x + Nx * y] = 120.0 * std::exp(-(dx * dx + dy * dy) / 0.02);
        }
    s.entries.push_back(entry("WVPseudoTopographicWaveGeneration", "seamount tide", WVForcingStage::spectral, 120,
                              {{"topographicHeight", {Ny, Nx}, height},
                               {"barotropicVelocityAmplitudeReal", {2}, std::vector<double>{0.04, 0.01}},
                               {"barotropicVelocityAmplitudeImag", {2}, std::vector<double>{0.0, 0.015}},
                               {"frequency", {}, std::vector<double>{1.4052e-4}},
                               {"darwinSymbol", {}, std::vector<std::string>{"M2"}},
                               {"rampDuration", {}, std::vector<double>{1200.0}},
                               {"startTime", {}, std::vector<double>{300.0}},
                               {"shouldAvoidAdaptiveDamping", {}, std::vector<std::uint8_t>{0}},
                               {"maximumForcedHorizontalWavenumber", {}, std::vector<double>{1e-3}},
                               {"maximumForcedVerticalMode", {}, std::vector<double>{3.0}}}));
    s.entries.push_back(entry("WVFixedAmplitudeForcing", "held modes", WVForcingStage::spectralAmplitude, 130,
                              {{"ApIndices", {2}, std::vector<std::int64_t>{6, 13}},
                               {"ApValuesReal", {2}, std::vector<double>{2e-3, -1e-3}},
                               {"ApValuesImag", {2}, std::vector<double>{1e-3, 5e-4}},
                               {"A0Indices", {1}, std::vector<std::int64_t>{11}},
                               {"A0ValuesReal", {1}, std::vector<double>{4e-3}},
                               {"A0ValuesImag", {1}, std::vector<double>{-2e-3}}}));
    return s;
}

std::shared_ptr<const WVExtensionCatalog> catalog() {
    std::shared_ptr<const WVExtensionCatalog> c;
    require(makeBuiltInExtensionCatalog(c), "built-in extension catalog");
    return c;
}

std::unique_ptr<WVConstantStratificationIntegrationSystem> system(const WVTransformConstantStratificationConfiguration& c,
                                                                  const WVFrozenForcingSchedule& s) {
    std::unique_ptr<WVConstantStratificationIntegrationSystem> out;
    require(WVConstantStratificationIntegrationSystem::create(c, s, catalog(), std::make_unique<WVReferenceFFTEngine>(), out),
            "integration system");
    return out;
}

// Fixed-step RK4 through the integration system (WVRungeKutta.cpp:335-480):
// advanceToTime with a clipped final step.
void fixedStep(Output& out, bool hydrostatic) {
    const std::string p = hydrostatic ? "rk4_h_" : "rk4_nh_";
    auto sys = system(configuration(16, 16, 9, 5, hydrostatic), nonlinearSchedule());
    State st(sys->kernel().descriptor(), 0.01);
    auto state = st.integration();
    WVFixedStepRK4 rk4(*sys);
    require(rk4.prepareStateAfterRestart(state), "rk4 restart preparation");
    require(rk4.advanceToTime(state, 3300.0, 600.0), "rk4 advanceToTime");
    out.put(p + "state", st.values);
    out.put(p + "t", state.waveVortex.t);
    out.put(p + "steps", static_cast<double>(rk4.metrics().stepCount));
    out.put(p + "last_step", rk4.metrics().lastStepSize);
}

// Adaptive RK23 (WVRungeKutta.hpp:112-203) through the integration system.
void adaptive(Output& out) {
    auto sys = system(configuration(8, 6, 7, 4, false), nonlinearSchedule());
    State st(sys->kernel().descriptor(), 0.01);
    auto state = st.integration();
    WVAdaptiveRK23Options options;
    options.relativeTolerance = 1e-6;
    WVAdaptiveRK23 rk23(*sys, options);
    require(rk23.prepareStateAfterRestart(state), "rk23 restart preparation");
    require(rk23.advanceToTime(state, 20000.0, 500.0), "rk23 advanceToTime");
    out.put("rk23_state", st.values);
    out.put("rk23_t", state.waveVortex.t);
    out.put("rk23_accepted", static_cast<double>(rk23.metrics().acceptedStepCount));
    out.put("rk23_rejected", static_cast<double>(rk23.metrics().rejectedStepCount));
    out.put("rk23_rhs", static_cast<double>(rk23.metrics().rightHandSideEvaluationCount));
}

// The forced schedule: one right-hand side from the forcing engine, and RK4
// steps through the integration system.
void forced(Output& out, bool hydrostatic) {
    const std::string p = hydrostatic ? "forced_h_" : "forced_nh_";
    const auto c = configuration(8, 6, 7, 5, hydrostatic);
    std::unique_ptr<WVConstantStratificationForcingEngine> engine;
    require(WVConstantStratificationForcingEngine::create(c, forcedSchedule(), catalog(),
                                                          std::make_unique<WVReferenceFFTEngine>(), engine),
            "forcing engine");
    State st(engine->kernel().descriptor(), 0.01);
    st.t = 1800.0;
    const std::size_t n = st.shape.elementCount();
    std::vector<WVComplex64> flux(3 * n);
    WVFlux fv{{flux.data(), st.shape}, {flux.data() + n, st.shape}, {flux.data() + 2 * n, st.shape}};
    require(engine->nonlinearFlux(st.view(), fv), "forced right-hand side");
    out.put(p + "rhs", flux);
    auto sys = system(c, forcedSchedule());
    State s2(sys->kernel().descriptor(), 0.01);
    auto state = s2.integration();
    WVFixedStepRK4 rk4(*sys);
    require(rk4.prepareStateAfterRestart(state), "forced rk4 restart preparation");
    require(rk4.advanceToTime(state, 2400.0, 600.0), "forced rk4");
    out.put(p + "state", s2.values);
}

// The topographic schedule: one right-hand side after the ramp has begun,
// and RK4 steps through the integration system with the fixed amplitudes
// restored after every step (WVRungeKutta.cpp, applyConstraint).
void topographic(Output& out, bool hydrostatic) {
    const std::string p = hydrostatic ? "topo_h_" : "topo_nh_";
    const auto c = configuration(8, 6, 7, 5, hydrostatic);
    std::unique_ptr<WVConstantStratificationForcingEngine> engine;
    require(WVConstantStratificationForcingEngine::create(c, topographicSchedule(c.Nx, c.Ny), catalog(),
                                                          std::make_unique<WVReferenceFFTEngine>(), engine),
            "topographic forcing engine");
    State st(engine->kernel().descriptor(), 0.01);
    st.t = 900.0;
    const std::size_t n = st.shape.elementCount();
    std::vector<WVComplex64> flux(3 * n);
    WVFlux fv{{flux.data(), st.shape}, {flux.data() + n, st.shape}, {flux.data() + 2 * n, st.shape}};
    require(engine->nonlinearFlux(st.view(), fv), "topographic right-hand side");
    out.put(p + "rhs", flux);
    // the same state under nonlinear advection alone (the forcing's own
    // contribution is the difference)
    std::unique_ptr<WVConstantStratificationForcingEngine> plain;
    require(WVConstantStratificationForcingEngine::create(c, nonlinearSchedule(), catalog(),
                                                          std::make_unique<WVReferenceFFTEngine>(), plain),
            "nonlinear forcing engine");
    require(plain->nonlinearFlux(st.view(), fv), "nonlinear right-hand side");
    out.put(p + "rhs_nonlinear", flux);
    auto sys = system(c, topographicSchedule(c.Nx, c.Ny));
    State s2(sys->kernel().descriptor(), 0.01);
    auto state = s2.integration();
    WVFixedStepRK4 rk4(*sys);
    require(rk4.prepareStateAfterRestart(state), "topographic rk4 restart preparation");
    require(rk4.advanceToTime(state, 3000.0, 600.0), "topographic rk4");
    out.put(p + "state", s2.values);
}

// Observers coupled to the right-hand side
// (WVConstantStratificationIntegrationSystem.cpp:294-406): Lagrangian floats
// sampled with spline interpolation in 3-D, surface drifters (x, y only) with
// linear interpolation, and an antialiased F-grid dye; every RHS shares the
// kernel's advection fields (nonlinearFluxWithAdvectionFields) among them.
// Fixed-step RK4 over 0 -> 1800 s.
void observers(Output& out) {
    const auto c = configuration(16, 16, 9, 5, false);
    WVTransformConstantStratificationDescriptor td;
    require(WVTransformConstantStratificationDescriptor::create(c, td), "transform descriptor");
    WVPortableObserverRecord rec;
    auto block = [&](const char* id, WVStateScalarType type, std::vector<std::size_t> dims, WVToleranceKind kind,
                     double tol) {
        rec.stateBlocks.push_back({id, type, std::move(dims), kind, tol, WVStateOwnership::integratorOwned,
                                   WVRestartRequirement::requiredDynamicState});
    };
    for (const char* id : {"Ap", "Am", "A0"})
        block(id, WVStateScalarType::complex64, {c.Nj, td.Nkl()}, WVToleranceKind::coefficientEnergyScaled, 1e-6);
    for (const char* id : {"float-x", "float-y"}) block(id, WVStateScalarType::real64, {4}, WVToleranceKind::uniformAbsolute, 1e-5);
    block("float-z", WVStateScalarType::real64, {4}, WVToleranceKind::uniformAbsolute, 1e-6);
    for (const char* id : {"drifter-x", "drifter-y"}) block(id, WVStateScalarType::real64, {3}, WVToleranceKind::uniformAbsolute, 1e-5);
    block("dye", WVStateScalarType::real64, {c.Nx, c.Ny, c.Nz}, WVToleranceKind::uniformAbsolute, 1e-8);

    WVObserverRecord coefficients;
    coefficients.identifier = "coefficients";
    coefficients.name = "coefficients";
    coefficients.typeIdentifier = "WVCoefficients";
    coefficients.stateBlockIdentifiers = {"Ap", "Am", "A0"};
    rec.observers.push_back(coefficients);
    WVObserverRecord floats;
    floats.identifier = "floats";
    floats.name = "floats";
    floats.typeIdentifier = "WVLagrangianParticles";
    floats.stateBlockIdentifiers = {"float-x", "float-y", "float-z"};
    floats.x = {1200.0, 17500.0, 33100.0, 48800.0};
    floats.y = {2500.0, 39000.0, 11000.0, 26700.0};
    floats.z = {-80.0, -420.0, -777.0, -1210.0};
    floats.advectionInterpolation = WVPositionInterpolation::spline;
    floats.horizontalAbsoluteTolerance = 1e-5;
    floats.verticalAbsoluteTolerance = 1e-6;
    rec.observers.push_back(floats);
    WVObserverRecord drifters;
    drifters.identifier = "drifters";
    drifters.name = "drifters";
    drifters.typeIdentifier = "WVLagrangianParticles";
    drifters.stateBlockIdentifiers = {"drifter-x", "drifter-y"};
    drifters.x = {-150.0, 20000.0, c.Lx + 40.0};
    drifters.y = {700.0, -35.0, 31000.0};
    drifters.z = {0.0, -10.0, -650.0};
    drifters.isXYOnly = true;
    drifters.advectionInterpolation = WVPositionInterpolation::linear;
    drifters.horizontalAbsoluteTolerance = 1e-5;
    rec.observers.push_back(drifters);
    WVObserverRecord dye;
    dye.identifier = "dye";
    dye.name = "dye";
    dye.typeIdentifier = "WVTracer";
    dye.stateBlockIdentifiers = {"dye"};
    dye.shouldAntialias = true;
    rec.observers.push_back(dye);
    const auto cat = catalog();  // the descriptor and the system share one catalog
    WVPortableObserverDescriptor desc;
    require(WVPortableObserverDescriptor::create(rec, cat, desc), "observer descriptor");

    std::unique_ptr<WVConstantStratificationIntegrationSystem> sys;
    require(WVConstantStratificationIntegrationSystem::create(c, nonlinearSchedule(), desc, cat,
                                                              std::make_unique<WVReferenceFFTEngine>(), sys),
            "integration system with observers");
    State st(sys->kernel().descriptor(), 0.01);
    WVAdditionalStateStorage storage;
    require(storage.initialize(sys->stateLayout()), "additional state storage");
    WVMutableIntegrationState state = st.integration();
    state.additionalBlocks = storage.mutableBlocks();
    state.additionalBlockCount = storage.blockCount();
    require(sys->initializeParticleState(state), "particle initial positions");
    auto find = [&](const std::string& id) -> WVAdditionalStateBlockView& {
        for (std::size_t b = 0; b < state.additionalBlockCount; ++b)
            if (state.additionalBlocks[b].layout->identifier == id) return state.additionalBlocks[b];
        throw std::runtime_error("missing state block " + id);
    };
    // dye: a smooth, band-limited blob pattern
    double* d = find("dye").realData;
    const double pi = std::acos(-1.0);
    for (std::size_t z = 0; z < c.Nz; ++z)
        for (std::size_t y = 0; y < c.Ny; ++y)
            for (std::size_t x = 0; x < c.Nx; ++x)
                d[x + c.Nx * (y + c.Ny * z)] = std::cos(2.0 * pi * x / c.Nx) * std::sin(2.0 * pi * y / c.Ny) +
                                               0.3 * std::cos(pi * z / (c.Nz - 1.0));
    WVFixedStepRK4 rk4(*sys);
    require(rk4.prepareStateAfterRestart(state), "observer rk4 restart preparation");
    require(rk4.advanceToTime(state, 1800.0, 600.0), "observer rk4");
    out.put("obs_state", st.values);
    for (const char* id : {"float-x", "float-y", "float-z", "drifter-x", "drifter-y"}) {
        const auto& b = find(id);
        out.put(std::string("obs_") + id, b.realData, b.layout->elementCount);
    }
    out.put("obs_dye", d, c.Nx * c.Ny * c.Nz);
    out.put("obs_rhs_count", static_cast<double>(sys->metrics().rightHandSideEvaluationCount));
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <output file> [scenario]\n", argv[0]);
        return 2;
    }
    try {
        Output out(argv[1]);
        const std::string only = argc > 2 ? argv[2] : "";
        auto run = [&](const char* name, auto&& body) {
            if (!only.empty() && only != name) return;
            const auto t0 = std::chrono::steady_clock::now();
            body();
            std::printf("%s done (%.2f s)\n", name,
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
            std::fflush(stdout);
        };
        run("rk4-nh", [&] { fixedStep(out, false); });
        run("rk4-h", [&] { fixedStep(out, true); });
        run("rk23", [&] { adaptive(out); });
        run("forced-nh", [&] { forced(out, false); });
        run("forced-h", [&] { forced(out, true); });
        run("topo-nh", [&] { topographic(out, false); });
        run("topo-h", [&] { topographic(out, true); });
        run("observers", [&] { observers(out); });
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "runtime driver failed: %s\n", e.what());
        return 1;
    }
}
