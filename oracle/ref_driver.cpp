// extern "C" driver around the REFERENCE kernel, built from the reference's
// own sources in place (see oracle/Makefile; output goes to oracle/_ref/).
//
// TEST INFRASTRUCTURE ONLY.  It uses the reference's public API as a caller
// would: WVTransformConstantStratificationKernel::create with the shipped
// WVReferenceFFTEngine (direct DFT, CompiledKernel/adapters/reference), then
// nonlinearFlux / transforms.  It is used in this container to write golden
// vectors under tests/golden/ and to pin the oracle restatement; it is slow
// (O(N^4) transforms) and therefore only meant for tiny grids.
//
// The API mirrors oracle/wv_oracle.h (prefix wvr_ instead of wvo_).
//
// Built a second time with WVR_FAST_ENGINE defined (oracle/_ref/libwvref_fast.so):
// the same reference kernel over the build-authored fast engine of
// fast_fft_engine.cpp, to time the reference's CPU path at production sizes;
// wvr_engine_compare then checks that engine against WVReferenceFFTEngine.
#include "WaveVortexKernel/WVTransformConstantStratificationKernel.hpp"
#include "WVReferenceFFTEngine.hpp"
#ifdef WVR_FAST_ENGINE
#include "fast_fft_engine.hpp"

#include <algorithm>
#include <cmath>
#include <random>
#endif

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace wavevortex;

namespace {

struct RefConfig {
    size_t Nx, Ny, Nz, Nj;
    double Lx, Ly, Lz, N0, rho0, g, planetaryRadius, rotationRate, latitude;
    int32_t isHydrostatic;
    int32_t shouldAntialias;
};

struct Handle {
    std::unique_ptr<WVTransformConstantStratificationKernel> kernel;
};

WVComplexConstView cview(const double* p, WVShape2D s) { return {reinterpret_cast<const WVComplex64*>(p), s}; }
WVComplexView mview(double* p, WVShape2D s) { return {reinterpret_cast<WVComplex64*>(p), s}; }

int report(const WVKernelStatus& st, char* msg, size_t len) {
    if (msg && len) std::snprintf(msg, len, "%s", st.message.c_str());
    return static_cast<int>(st.code);
}

}  // namespace

#ifdef WVR_FAST_ENGINE
int gThreads = 0;  // threads of the fast engine (0: OpenMP's default)
#endif

extern "C" {

#ifdef WVR_FAST_ENGINE
void wvr_set_threads(int n) { gThreads = n; }

// The fast engine against WVReferenceFFTEngine on one plan of the given kind
// (0 r2c 2-D [n0 = Ny, n1 = Nx], 1 c2r 2-D, 2 DCT-I and 3 DST-I of length n0),
// `batch` contiguous batches, seeded random input (c2r: a Hermitian-consistent
// half spectrum, the r2c of random data).  Returns max|fast - ref| / max|ref|.
double wvr_engine_compare(int kind, size_t n0, size_t n1, size_t batch, unsigned seed) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd;
    WVFFTPlanSpecification sp;
    size_t inCount = 0, outCount = 0;  // doubles
    if (kind <= 1) {
        const size_t Ny = n0, Nx = n1, H = Nx / 2 + 1;
        const std::ptrdiff_t real = static_cast<std::ptrdiff_t>(Nx * Ny), half = static_cast<std::ptrdiff_t>(H * Ny);
        sp.kind = kind == 0 ? WVFFTPlanKind::horizontalRealToComplex2D : WVFFTPlanKind::horizontalComplexToReal2D;
        if (kind == 0) {
            sp.transformDimensions = {{Ny, static_cast<std::ptrdiff_t>(Nx), static_cast<std::ptrdiff_t>(H)}, {Nx, 1, 1}};
            sp.batchDimensions = {{batch, real, half}};
            inCount = batch * real;
            outCount = 2 * batch * half;
        } else {
            sp.transformDimensions = {{Ny, static_cast<std::ptrdiff_t>(H), static_cast<std::ptrdiff_t>(Nx)}, {Nx, 1, 1}};
            sp.batchDimensions = {{batch, half, real}};
            inCount = 2 * batch * half;
            outCount = batch * real;
        }
    } else {
        const size_t N = n0;
        sp.kind = kind == 2 ? WVFFTPlanKind::verticalDCTI : WVFFTPlanKind::verticalDSTI;
        sp.transformDimensions = {{N, 2, 2}};
        sp.batchDimensions = {{2, 1, 1}, {batch, static_cast<std::ptrdiff_t>(2 * N), static_cast<std::ptrdiff_t>(2 * N)}};
        inCount = outCount = 2 * N * batch;
    }
    sp.inputBytes = inCount * sizeof(double);
    sp.outputBytes = outCount * sizeof(double);
    std::vector<double> in(inCount);
    for (auto& v : in) v = nd(rng);
    if (kind == 1) {  // a consistent half spectrum: the r2c of random real data
        WVFFTPlanSpecification f = sp;
        f.kind = WVFFTPlanKind::horizontalRealToComplex2D;
        for (auto& d : f.transformDimensions) std::swap(d.inputStride, d.outputStride);
        for (auto& d : f.batchDimensions) std::swap(d.inputStride, d.outputStride);
        std::vector<double> real(outCount);
        for (auto& v : real) v = nd(rng);
        std::unique_ptr<WVFFTPlan> p;
        wvfast::FastFFTEngine(gThreads).createPlan(f, p);
        p->execute(real.data(), in.data());
    }
    std::vector<double> a(outCount), b(outCount);
    std::unique_ptr<WVFFTPlan> pf, pr;
    wvfast::FastFFTEngine fast(gThreads);
    WVReferenceFFTEngine ref;
    if (!fast.createPlan(sp, pf) || !ref.createPlan(sp, pr)) return -1.0;
    std::vector<double> in2 = in;
    if (!pf->execute(in.data(), a.data()) || !pr->execute(in2.data(), b.data())) return -1.0;
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < outCount; ++i) {
        num = std::max(num, std::fabs(a[i] - b[i]));
        den = std::max(den, std::fabs(b[i]));
    }
    return den > 0.0 ? num / den : num;
}
#endif

void* wvr_create(const RefConfig* c, char* msg, size_t msglen) {
    WVTransformConstantStratificationConfiguration cfg;
    cfg.Nx = c->Nx; cfg.Ny = c->Ny; cfg.Nz = c->Nz; cfg.Nj = c->Nj;
    cfg.Lx = c->Lx; cfg.Ly = c->Ly; cfg.Lz = c->Lz; cfg.N0 = c->N0; cfg.rho0 = c->rho0; cfg.g = c->g;
    cfg.planetaryRadius = c->planetaryRadius; cfg.rotationRate = c->rotationRate; cfg.latitude = c->latitude;
    cfg.isHydrostatic = c->isHydrostatic != 0;
    cfg.shouldAntialias = c->shouldAntialias != 0;
    auto h = std::make_unique<Handle>();
#ifdef WVR_FAST_ENGINE
    auto engine = std::make_unique<wvfast::FastFFTEngine>(gThreads);
#else
    auto engine = std::make_unique<WVReferenceFFTEngine>();
#endif
    const auto st = WVTransformConstantStratificationKernel::create(cfg, std::move(engine), h->kernel);
    if (!st) { report(st, msg, msglen); return nullptr; }
    return h.release();
}

void wvr_destroy(void* h) { delete static_cast<Handle*>(h); }

size_t wvr_nkl(void* h) { return static_cast<Handle*>(h)->kernel->descriptor().Nkl(); }

double wvr_coriolis(void* h) { return static_cast<Handle*>(h)->kernel->descriptor().verticalModes().coriolisFrequency; }

void wvr_modes(void* h, int64_t* kMode, int64_t* lMode, double* k, double* l, double* Kh, double* cosA, double* sinA, int64_t* row, uint8_t* conj) {
    const auto& d = static_cast<Handle*>(h)->kernel->descriptor();
    const auto& md = d.fourierModes();
    const auto& map = d.halfSpectrumMappings();
    for (size_t i = 0; i < md.size(); ++i) {
        kMode[i] = md[i].kMode; lMode[i] = md[i].lMode; k[i] = md[i].k; l[i] = md[i].l; Kh[i] = md[i].Kh;
        cosA[i] = md[i].cosAlpha; sinA[i] = md[i].sinAlpha;
        row[i] = static_cast<int64_t>(map.storageRowsByWVIndex[i]);
        conj[i] = map.conjugatesStoredValueByWVIndex[i];
    }
}

long wvr_table(void* h, const char* name, double* out) {
    const auto& vm = static_cast<Handle*>(h)->kernel->descriptor().verticalModes();
    const std::string n(name);
    auto real = [&](const std::vector<double>& v) { std::memcpy(out, v.data(), v.size() * sizeof(double)); return long(v.size()); };
    auto cpx = [&](const std::vector<WVComplex64>& v) { std::memcpy(out, v.data(), v.size() * sizeof(WVComplex64)); return long(2 * v.size()); };
    if (n == "z") return real(vm.z);
    if (n == "h0") return real(vm.h0);
    if (n == "verticalWavenumber") return real(vm.verticalWavenumber);
    if (n == "Fg") return real(vm.Fg);
    if (n == "Gg") return real(vm.Gg);
    if (n == "inertialScale") return real(vm.inertialScale);
    if (n == "gWaveScale") return real(vm.gWaveScale);
    if (n == "apmWProjectionPrefactor") return real(vm.apmWProjectionPrefactor);
    if (n == "omega") return real(vm.omega);
    if (n == "fWaveScale") return real(vm.fWaveScale);
    if (n == "NApField") return real(vm.NApField);
    if (n == "NA0Field") return real(vm.NA0Field);
    if (n == "A0FromVorticity") return real(vm.A0FromVorticity);
    if (n == "A0FromBuoyancy") return real(vm.A0FromBuoyancy);
    if (n == "ApmNProjection") return real(vm.ApmNProjection);
    if (n == "ApmDScaled") return real(vm.ApmDScaled);
    if (n == "UApField") return cpx(vm.UApField);
    if (n == "VApField") return cpx(vm.VApField);
    if (n == "WApField") return cpx(vm.WApField);
    if (n == "UA0Field") return cpx(vm.UA0Field);
    if (n == "VA0Field") return cpx(vm.VA0Field);
    if (n == "ApmDProjection") return cpx(vm.ApmDProjection);
    return -1;
}

int wvr_nonlinear_flux(void* h, double t, double t0, const double* Ap, const double* Am, const double* A0, double* Fp, double* Fm, double* F0) {
    auto& k = *static_cast<Handle*>(h)->kernel;
    const auto s = k.descriptor().spectralShape();
    WVState state{t, t0, {cview(Ap, s), cview(Am, s), cview(A0, s)}};
    WVFlux flux{mview(Fp, s), mview(Fm, s), mview(F0, s)};
    return static_cast<int>(k.nonlinearFlux(state, flux).code);
}

int wvr_inverse_uvweta(void* h, double t, double t0, const double* Ap, const double* Am, const double* A0, double* fields) {
    auto& k = *static_cast<Handle*>(h)->kernel;
    const auto s = k.descriptor().spectralShape();
    const auto sp = k.descriptor().spatialShape();
    WVState state{t, t0, {cview(Ap, s), cview(Am, s), cview(A0, s)}};
    WVRealFieldBundleView out{fields, {sp.first, sp.second, sp.third, 4}};
    return static_cast<int>(k.transformWaveVortexToUVWEta(state, out).code);
}

int wvr_inverse_uvw(void* h, double t, double t0, const double* Ap, const double* Am, const double* A0, double* fields) {
    auto& k = *static_cast<Handle*>(h)->kernel;
    const auto s = k.descriptor().spectralShape();
    const auto sp = k.descriptor().spatialShape();
    WVState state{t, t0, {cview(Ap, s), cview(Am, s), cview(A0, s)}};
    WVRealFieldBundleView out{fields, {sp.first, sp.second, sp.third, 3}};
    return static_cast<int>(k.transformWaveVortexToUVW(state, out).code);
}

int wvr_forward_uvweta(void* h, const double* fields, double t, double t0, double* Ap, double* Am, double* A0) {
    auto& k = *static_cast<Handle*>(h)->kernel;
    const auto s = k.descriptor().spectralShape();
    const auto sp = k.descriptor().spatialShape();
    WVRealFieldBundleConstView in{fields, {sp.first, sp.second, sp.third, 4}};
    WVMutableCoefficients out{mview(Ap, s), mview(Am, s), mview(A0, s)};
    return static_cast<int>(k.transformUVWEtaToWaveVortex(in, t, t0, out).code);
}

int wvr_forward_uveta(void* h, const double* fields, double t, double t0, double* Ap, double* Am, double* A0) {
    auto& k = *static_cast<Handle*>(h)->kernel;
    const auto s = k.descriptor().spectralShape();
    const auto sp = k.descriptor().spatialShape();
    WVRealFieldBundleConstView in{fields, {sp.first, sp.second, sp.third, 3}};
    WVMutableCoefficients out{mview(Ap, s), mview(Am, s), mview(A0, s)};
    return static_cast<int>(k.transformUVEtaToWaveVortex(in, t, t0, out).code);
}

int wvr_fg_all_derivatives(void* h, int fFamily, const double* Apm, const double* A0, double* fields) {
    auto& k = *static_cast<Handle*>(h)->kernel;
    const auto s = k.descriptor().spectralShape();
    const auto sp = k.descriptor().spatialShape();
    WVRealFieldBundleView out{fields, {sp.first, sp.second, sp.third, 4}};
    const auto st = fFamily ? k.transformToSpatialDomainWithFAllDerivatives(cview(Apm, s), cview(A0, s), out)
                            : k.transformToSpatialDomainWithGAllDerivatives(cview(Apm, s), cview(A0, s), out);
    return static_cast<int>(st.code);
}

// nonlinearFluxUsingAdvectionFields (...Kernel.cpp:1160-1171).
int wvr_nonlinear_flux_using_advection_fields(void* h, double t, double t0, const double* Ap, const double* Am,
                                              const double* A0, const double* uvw, double* Fp, double* Fm,
                                              double* F0) {
    auto& k = *static_cast<Handle*>(h)->kernel;
    const auto s = k.descriptor().spectralShape();
    const auto sp = k.descriptor().spatialShape();
    WVState state{t, t0, {cview(Ap, s), cview(Am, s), cview(A0, s)}};
    WVFlux flux{mview(Fp, s), mview(Fm, s), mview(F0, s)};
    WVRealFieldBundleConstView adv{uvw, {sp.first, sp.second, sp.third, 3}};
    return static_cast<int>(k.nonlinearFluxUsingAdvectionFields(state, flux, adv).code);
}

// transformStateFieldDerivatives (...Kernel.cpp:719-742).
int wvr_state_field_derivatives(void* h, double t, double t0, const double* Ap, const double* Am, const double* A0,
                                int field, double* out) {
    auto& k = *static_cast<Handle*>(h)->kernel;
    const auto s = k.descriptor().spectralShape();
    const auto sp = k.descriptor().spatialShape();
    WVState state{t, t0, {cview(Ap, s), cview(Am, s), cview(A0, s)}};
    WVRealFieldBundleView d{out, {sp.first, sp.second, sp.third, 3}};
    return static_cast<int>(k.transformStateFieldDerivatives(state, static_cast<WVDynamicalField>(field), d).code);
}

// advectFGridScalar (...Kernel.cpp:1032-1138): rhs = -(u, v, w) . grad(scalar)
// of an F-grid scalar, optionally antialiased.
int wvr_advect_f_grid_scalar(void* h, const double* scalar, const double* uvw, int antialias, double* rhs) {
    auto& k = *static_cast<Handle*>(h)->kernel;
    const auto sp = k.descriptor().spatialShape();
    WVRealVolumeConstView in{scalar, sp};
    WVRealFieldBundleConstView adv{uvw, {sp.first, sp.second, sp.third, 3}};
    WVRealVolumeView out{rhs, sp};
    return static_cast<int>(k.advectFGridScalar(in, adv, antialias != 0, out).code);
}

}  // extern "C"
