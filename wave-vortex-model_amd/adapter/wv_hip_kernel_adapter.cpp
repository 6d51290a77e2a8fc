// wavevortex::WVTransformConstantStratificationKernel (contract version 4,
// CompiledKernel/include/WaveVortexKernel/WVTransformConstantStratificationKernel.hpp:47-110)
// implemented over libwvamd's C-ABI (include/wv_hip.h).
//
// Each operation validates its views as the contract's implementation does
// (shapes, null pointers, overlap, reentrancy; CompiledKernel/src/
// WVTransformConstantStratificationKernel.cpp:86-135 and the entry points),
// copies its host inputs into device staging allocated at create, runs the
// libwvamd device entry point on the default stream and copies the outputs
// back (synchronous, as the host-memory contract is).  The descriptor is the
// contract's own (WVTransformConstantStratificationDescriptor::create,
// CompiledKernel/src/WVKernelTypes.cpp); libwvamd builds a bitwise-identical
// copy on the device.
#include "WaveVortexKernel/WVTransformConstantStratificationKernel.hpp"

#include "wv_hip.h"

#include <hip/hip_runtime_api.h>

#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>

namespace wavevortex {
namespace {

using Kernel = WVTransformConstantStratificationKernel;

WVKernelStatus status(WVKernelStatusCode code, std::string message) { return {code, std::move(message)}; }

// One operation at a time per kernel (ExecutionGuard, ...Kernel.cpp:76-84).
class Busy {
public:
    explicit Busy(bool& flag) : flag_(flag), entered_(!flag) {
        if (entered_) flag_ = true;
    }
    ~Busy() {
        if (entered_) flag_ = false;
    }
    bool entered() const noexcept { return entered_; }

private:
    bool& flag_;
    bool entered_;
};

WVKernelStatus checkBundle(const double* data, const WVShape4D& shape, WVShape3D spatial, std::size_t channels,
                           const char* name) {
    if (shape.first != spatial.first || shape.second != spatial.second || shape.third != spatial.third ||
        shape.fourth != channels)
        return status(WVKernelStatusCode::invalidShape, std::string(name) + " must have shape [Nx,Ny,Nz,Nfield].");
    if (data == nullptr) return status(WVKernelStatusCode::invalidPointer, std::string(name) + " has a null data pointer.");
    return WVKernelStatus::ok();
}

template <typename T>
WVKernelStatus checkSpectral(const WVMatrixView<T>& view, WVShape2D spectral, const char* name) {
    if (view.shape.rows != spectral.rows || view.shape.columns != spectral.columns)
        return status(WVKernelStatusCode::invalidShape, std::string(name) + " must have shape [Nj,Nkl].");
    if (view.data == nullptr) return status(WVKernelStatusCode::invalidPointer, std::string(name) + " has a null data pointer.");
    return WVKernelStatus::ok();
}

bool overlaps(const void* a, std::size_t an, const void* b, std::size_t bn) {
    if (a == nullptr || b == nullptr || an == 0 || bn == 0) return false;
    const auto x = reinterpret_cast<std::uintptr_t>(a), y = reinterpret_cast<std::uintptr_t>(b);
    return x < y + bn && y < x + an;
}

WVKernelStatus checkState(const WVState& state, WVShape2D spectral) {
    const WVKernelStatus s[] = {checkSpectral(state.coefficients.Ap, spectral, "Ap"),
                                checkSpectral(state.coefficients.Am, spectral, "Am"),
                                checkSpectral(state.coefficients.A0, spectral, "A0")};
    for (const auto& v : s)
        if (!v) return v;
    return WVKernelStatus::ok();
}

WVKernelStatus checkInverseOwnership(const WVState& state, const void* fields, std::size_t fieldBytes,
                                     std::size_t coefficientBytes) {
    for (const void* in : {static_cast<const void*>(state.coefficients.Ap.data),
                           static_cast<const void*>(state.coefficients.Am.data),
                           static_cast<const void*>(state.coefficients.A0.data)})
        if (overlaps(fields, fieldBytes, in, coefficientBytes))
            return status(WVKernelStatusCode::overlappingArrays, "Inverse spatial output must not overlap coefficient inputs.");
    return WVKernelStatus::ok();
}

int envInt(const char* name, int fallback) {
    const char* v = std::getenv(name);
    return v ? std::atoi(v) : fallback;
}

}  // namespace

WVKernelStatus Kernel::create(const WVTransformConstantStratificationConfiguration& configuration,
                              std::unique_ptr<WVFFTEngine> engine, std::unique_ptr<Kernel>& kernel) {
    if (!engine) return status(WVKernelStatusCode::invalidPointer, "An FFT engine is required.");
    try {
        std::unique_ptr<Kernel> candidate(new Kernel());
        auto st = WVTransformConstantStratificationDescriptor::create(configuration, candidate->descriptor_);
        if (!st) return st;
        candidate->engineIdentifier_ = engine->identifier();
        candidate->engineLibraryIdentity_ = engine->libraryIdentity();
        candidate->engine_ = std::move(engine);
        wv_config c{};
        c.contractVersion = configuration.contractVersion;
        c.Nx = configuration.Nx;
        c.Ny = configuration.Ny;
        c.Nz = configuration.Nz;
        c.Nj = configuration.Nj;
        c.Lx = configuration.Lx;
        c.Ly = configuration.Ly;
        c.Lz = configuration.Lz;
        c.N0 = configuration.N0;
        c.rho0 = configuration.rho0;
        c.g = configuration.g;
        c.planetaryRadius = configuration.planetaryRadius;
        c.rotationRate = configuration.rotationRate;
        c.latitude = configuration.latitude;
        c.isHydrostatic = configuration.isHydrostatic ? 1 : 0;
        c.shouldAntialias = configuration.shouldAntialias ? 1 : 0;
        const char* prec = std::getenv("WVAMD_PRECISION");
        candidate->precision_ = (prec && std::strcmp(prec, "mixed") == 0) ? WV_PRECISION_MIXED : WV_PRECISION_FP64;
        const int device = envInt("WVAMD_DEVICE", 0);
        const int rc = wv_create(&c, static_cast<wv_precision>(candidate->precision_), device, &candidate->device_);
        if (rc != WV_SUCCESS) return candidate->fromDevice(rc);
        candidate->coefficientBytes_ = candidate->descriptor_.spectralShape().elementCount() * sizeof(WVComplex64);
        candidate->volumeBytes_ = candidate->descriptor_.spatialShape().elementCount() * sizeof(double);
        if (hipSetDevice(device) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&candidate->dCoefficients_), 6 * candidate->coefficientBytes_) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&candidate->dVolumes_), 5 * candidate->volumeBytes_) != hipSuccess)
            return status(WVKernelStatusCode::allocationFailure, "Unable to allocate the device staging buffers.");
        kernel = std::move(candidate);
        return WVKernelStatus::ok();
    } catch (const std::bad_alloc&) {
        return status(WVKernelStatusCode::allocationFailure, "Unable to allocate the bounded kernel scratch arena.");
    } catch (const std::overflow_error& error) {
        return status(WVKernelStatusCode::sizeOverflow, error.what());
    }
}

Kernel::~WVTransformConstantStratificationKernel() {
    if (dCoefficients_) (void)hipFree(dCoefficients_);
    if (dVolumes_) (void)hipFree(dVolumes_);
    if (device_) wv_destroy(device_);
}

WVKernelStatus Kernel::fromDevice(int code) const {
    if (code == WV_SUCCESS) return WVKernelStatus::ok();
    return status(static_cast<WVKernelStatusCode>(code), wv_last_error());
}

WVKernelStatus Kernel::toDevice(const void* host, std::size_t bytes, void* device) {
    if (hipMemcpy(device, host, bytes, hipMemcpyHostToDevice) != hipSuccess)
        return status(WVKernelStatusCode::fftExecutionFailure, "Host-to-device copy failed.");
    metrics_.bytesCopied += bytes;
    return WVKernelStatus::ok();
}

WVKernelStatus Kernel::toHost(const void* device, std::size_t bytes, void* host) {
    if (hipMemcpy(host, device, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return status(WVKernelStatusCode::fftExecutionFailure, "Device-to-host copy failed.");
    metrics_.bytesCopied += bytes;
    return WVKernelStatus::ok();
}

WVKernelStatus Kernel::stageState(const WVState& state) {
    const std::size_t n = coefficientBytes_ / sizeof(double);
    auto st = toDevice(state.coefficients.Ap.data, coefficientBytes_, dCoefficients_);
    if (st) st = toDevice(state.coefficients.Am.data, coefficientBytes_, dCoefficients_ + n);
    if (st) st = toDevice(state.coefficients.A0.data, coefficientBytes_, dCoefficients_ + 2 * n);
    return st;
}

const WVKernelMetrics& Kernel::metrics() const noexcept {
    wv_kernel_metrics m{};
    if (wv_metrics(device_, &m) != WV_SUCCESS) return metrics_;
    const std::size_t staging = 6 * coefficientBytes_ + 5 * volumeBytes_;
    metrics_.descriptorBytes = m.descriptorBytes;
    metrics_.planBytes = m.planBytes;
    metrics_.engineBytes = engine_ ? engine_->persistentBytes() : 0;
    metrics_.kernelManagementBytes = sizeof(*this) + engineIdentifier_.capacity() + engineLibraryIdentity_.capacity() +
                                     m.kernelManagementBytes;
    metrics_.halfSpectrumScratchCapacityBytes = m.halfSpectrumScratchCapacityBytes;
    metrics_.realScratchCapacityBytes = staging;
    metrics_.scratchCapacityBytes = m.scratchCapacityBytes + staging;
    metrics_.scratchHighWaterBytes = metrics_.scratchCapacityBytes;
    metrics_.planCount = m.planCount;
    metrics_.executionCount = m.executionCount;
    metrics_.horizontalExecutionCount = m.horizontalExecutionCount;
    metrics_.verticalExecutionCount = m.verticalExecutionCount;
    metrics_.nonlinearFluxCallCount = m.nonlinearFluxCallCount;
    metrics_.nonlinearFluxPhaseEvaluationCount = m.nonlinearFluxPhaseEvaluationCount;
    metrics_.advectionVelocityReconstructionCount = m.advectionVelocityReconstructionCount;
    metrics_.scalarAdvectionCount = m.scalarAdvectionCount;
    metrics_.scalarAntialiasCount = m.scalarAntialiasCount;
    metrics_.phaseSeconds = m.phaseSeconds;
    metrics_.reconstructionSeconds = m.reconstructionSeconds;
    metrics_.productSeconds = m.productSeconds;
    metrics_.projectionSeconds = m.projectionSeconds;
    return metrics_;
}

const char* Kernel::nonlinearFluxScheduleIdentifier() const noexcept { return wv_nonlinear_flux_schedule_identifier(device_); }
const char* Kernel::phaseImplementationIdentifier() const noexcept { return wv_phase_implementation_identifier(device_); }
const char* Kernel::coefficientArithmeticModeIdentifier() const noexcept {
    return wv_coefficient_arithmetic_mode_identifier(device_);
}
const char* Kernel::inverseNormalizationPlacementIdentifier() const noexcept {
    return wv_inverse_normalization_placement_identifier(device_);
}
const char* Kernel::optimizationImplementationIdentifier() const noexcept {
    return wv_optimization_implementation_identifier(device_);
}
std::size_t Kernel::coefficientWorkerCount() const noexcept { return wv_coefficient_worker_count(device_); }
std::size_t Kernel::phaseReservationBytes() const noexcept { return wv_phase_reservation_bytes(device_); }
void Kernel::setStageInstrumentation(bool enabled) noexcept { (void)wv_set_stage_instrumentation(device_, enabled ? 1 : 0); }

std::size_t Kernel::scratchBytes() const noexcept {
    return wv_scratch_bytes(device_) + 6 * coefficientBytes_ + 5 * volumeBytes_;
}

std::size_t Kernel::persistentBytes() const noexcept {
    const auto& m = metrics();
    return m.kernelManagementBytes + m.engineBytes + m.descriptorBytes + m.planBytes + m.scratchCapacityBytes;
}

// ---- forward transforms (...Kernel.cpp:583-684) -------------------------------
WVKernelStatus Kernel::forwardCall(const WVRealFieldBundleConstView& fields, std::size_t channels, double t, double t0,
                                   WVMutableCoefficients& coefficients) {
    const auto spectral = descriptor_.spectralShape();
    const WVKernelStatus s[] = {checkSpectral(coefficients.Ap, spectral, "Ap"), checkSpectral(coefficients.Am, spectral, "Am"),
                                checkSpectral(coefficients.A0, spectral, "A0")};
    for (const auto& v : s)
        if (!v) return v;
    const std::size_t fieldBytes = channels * volumeBytes_;
    const void* outs[] = {coefficients.Ap.data, coefficients.Am.data, coefficients.A0.data};
    for (std::size_t i = 0; i < 3; ++i) {
        if (overlaps(outs[i], coefficientBytes_, fields.data, fieldBytes))
            return status(WVKernelStatusCode::overlappingArrays, "Forward coefficient outputs must not overlap the spatial input.");
        for (std::size_t j = i + 1; j < 3; ++j)
            if (overlaps(outs[i], coefficientBytes_, outs[j], coefficientBytes_))
                return status(WVKernelStatusCode::overlappingArrays, "Forward coefficient outputs must not overlap each other.");
    }
    Busy busy(executing_);
    if (!busy.entered()) return status(WVKernelStatusCode::reentrantExecution, "Kernel operations are not reentrant.");
    auto st = toDevice(fields.data, fieldBytes, dVolumes_);
    if (!st) return st;
    const std::size_t n = coefficientBytes_ / sizeof(double);
    double* o = dCoefficients_ + 3 * n;
    const int rc = channels == 4 ? wv_transform_uvweta_to_wave_vortex_device(device_, dVolumes_, t, t0, o, o + n, o + 2 * n, nullptr)
                                 : wv_transform_uveta_to_wave_vortex_device(device_, dVolumes_, t, t0, o, o + n, o + 2 * n, nullptr);
    if (rc != WV_SUCCESS) return fromDevice(rc);
    st = toHost(o, coefficientBytes_, coefficients.Ap.data);
    if (st) st = toHost(o + n, coefficientBytes_, coefficients.Am.data);
    if (st) st = toHost(o + 2 * n, coefficientBytes_, coefficients.A0.data);
    return st;
}

WVKernelStatus Kernel::transformUVEtaToWaveVortex(const WVRealFieldBundleConstView& fields, double t, double t0,
                                                  WVMutableCoefficients& coefficients) {
    if (!descriptor_.configuration().isHydrostatic)
        return status(WVKernelStatusCode::invalidConfiguration, "transformUVEtaToWaveVortex requires a hydrostatic kernel.");
    auto st = checkBundle(fields.data, fields.shape, descriptor_.spatialShape(), 3, "Hydrostatic fields");
    if (!st) return st;
    return forwardCall(fields, 3, t, t0, coefficients);
}

WVKernelStatus Kernel::transformUVWEtaToWaveVortex(const WVRealFieldBundleConstView& fields, double t, double t0,
                                                   WVMutableCoefficients& coefficients) {
    if (descriptor_.configuration().isHydrostatic)
        return status(WVKernelStatusCode::invalidConfiguration, "transformUVWEtaToWaveVortex requires a nonhydrostatic kernel.");
    auto st = checkBundle(fields.data, fields.shape, descriptor_.spatialShape(), 4, "Nonhydrostatic fields");
    if (!st) return st;
    return forwardCall(fields, 4, t, t0, coefficients);
}

// ---- inverse transforms (...Kernel.cpp:686-742) -------------------------------
WVKernelStatus Kernel::transformWaveVortexToUVWEta(const WVState& state, WVRealFieldBundleView& fields) {
    auto st = checkBundle(fields.data, fields.shape, descriptor_.spatialShape(), 4, "Reconstructed fields");
    if (!st) return st;
    if (!(st = checkState(state, descriptor_.spectralShape()))) return st;
    if (!(st = checkInverseOwnership(state, fields.data, 4 * volumeBytes_, coefficientBytes_))) return st;
    Busy busy(executing_);
    if (!busy.entered()) return status(WVKernelStatusCode::reentrantExecution, "Kernel operations are not reentrant.");
    if (!(st = stageState(state))) return st;
    const std::size_t n = coefficientBytes_ / sizeof(double);
    const int rc = wv_transform_wave_vortex_to_uvweta_device(device_, state.t, state.t0, dCoefficients_, dCoefficients_ + n,
                                                             dCoefficients_ + 2 * n, dVolumes_, nullptr);
    if (rc != WV_SUCCESS) return fromDevice(rc);
    return toHost(dVolumes_, 4 * volumeBytes_, fields.data);
}

WVKernelStatus Kernel::transformWaveVortexToUVW(const WVState& state, WVRealFieldBundleView& fields) {
    auto st = checkBundle(fields.data, fields.shape, descriptor_.spatialShape(), 3, "Advection fields");
    if (!st) return st;
    if (!(st = checkState(state, descriptor_.spectralShape()))) return st;
    if (!(st = checkInverseOwnership(state, fields.data, 3 * volumeBytes_, coefficientBytes_))) return st;
    Busy busy(executing_);
    if (!busy.entered()) return status(WVKernelStatusCode::reentrantExecution, "Kernel operations are not reentrant.");
    if (!(st = stageState(state))) return st;
    const std::size_t n = coefficientBytes_ / sizeof(double);
    const int rc = wv_transform_wave_vortex_to_uvw_device(device_, state.t, state.t0, dCoefficients_, dCoefficients_ + n,
                                                          dCoefficients_ + 2 * n, dVolumes_, nullptr);
    if (rc != WV_SUCCESS) return fromDevice(rc);
    return toHost(dVolumes_, 3 * volumeBytes_, fields.data);
}

WVKernelStatus Kernel::transformStateFieldDerivatives(const WVState& state, WVDynamicalField field,
                                                      WVRealFieldBundleView& derivatives) {
    const auto target = static_cast<std::size_t>(field);
    if (target > static_cast<std::size_t>(WVDynamicalField::eta))
        return status(WVKernelStatusCode::invalidConfiguration, "Unknown dynamical field derivative target.");
    auto st = checkBundle(derivatives.data, derivatives.shape, descriptor_.spatialShape(), 3, "Field derivatives");
    if (!st) return st;
    if (!(st = checkState(state, descriptor_.spectralShape()))) return st;
    if (!(st = checkInverseOwnership(state, derivatives.data, 3 * volumeBytes_, coefficientBytes_))) return st;
    Busy busy(executing_);
    if (!busy.entered()) return status(WVKernelStatusCode::reentrantExecution, "Kernel operations are not reentrant.");
    if (!(st = stageState(state))) return st;
    const std::size_t n = coefficientBytes_ / sizeof(double);
    const int rc = wv_transform_state_field_derivatives_device(device_, state.t, state.t0, dCoefficients_,
                                                               dCoefficients_ + n, dCoefficients_ + 2 * n,
                                                               static_cast<int>(target), dVolumes_, nullptr);
    if (rc != WV_SUCCESS) return fromDevice(rc);
    return toHost(dVolumes_, 3 * volumeBytes_, derivatives.data);
}

// ---- F / G all-derivatives (...Kernel.cpp:819-863) ----------------------------
WVKernelStatus Kernel::familyDerivatives(bool fFamily, const WVComplexConstView& Apm, const WVComplexConstView& A0,
                                         WVRealFieldBundleView& fields) {
    const auto spectral = descriptor_.spectralShape();
    auto st = checkSpectral(Apm, spectral, "Apm");
    if (!st) return st;
    if (!(st = checkSpectral(A0, spectral, "A0"))) return st;
    if (!(st = checkBundle(fields.data, fields.shape, descriptor_.spatialShape(), 4, "Field derivatives"))) return st;
    if (overlaps(fields.data, 4 * volumeBytes_, Apm.data, coefficientBytes_) ||
        overlaps(fields.data, 4 * volumeBytes_, A0.data, coefficientBytes_))
        return status(WVKernelStatusCode::overlappingArrays, "Derivative output must not overlap coefficient inputs.");
    Busy busy(executing_);
    if (!busy.entered()) return status(WVKernelStatusCode::reentrantExecution, "Kernel operations are not reentrant.");
    const std::size_t n = coefficientBytes_ / sizeof(double);
    if (!(st = toDevice(Apm.data, coefficientBytes_, dCoefficients_))) return st;
    if (!(st = toDevice(A0.data, coefficientBytes_, dCoefficients_ + n))) return st;
    const int rc = fFamily ? wv_transform_to_spatial_domain_with_f_all_derivatives_device(device_, dCoefficients_,
                                                                                          dCoefficients_ + n, dVolumes_, nullptr)
                           : wv_transform_to_spatial_domain_with_g_all_derivatives_device(device_, dCoefficients_,
                                                                                          dCoefficients_ + n, dVolumes_, nullptr);
    if (rc != WV_SUCCESS) return fromDevice(rc);
    return toHost(dVolumes_, 4 * volumeBytes_, fields.data);
}

WVKernelStatus Kernel::transformToSpatialDomainWithFAllDerivatives(const WVComplexConstView& Apm, const WVComplexConstView& A0,
                                                                   WVRealFieldBundleView& fields) {
    return familyDerivatives(true, Apm, A0, fields);
}

WVKernelStatus Kernel::transformToSpatialDomainWithGAllDerivatives(const WVComplexConstView& Apm, const WVComplexConstView& A0,
                                                                   WVRealFieldBundleView& fields) {
    return familyDerivatives(false, Apm, A0, fields);
}

// ---- the flux (...Kernel.cpp:1140-1171) -----------------------------------------
WVKernelStatus Kernel::fluxCall(const WVState& state, WVFlux& flux, double* uvwOut, const double* uvwIn) {
    Busy busy(executing_);
    if (!busy.entered()) return status(WVKernelStatusCode::reentrantExecution, "Kernel operations are not reentrant.");
    auto st = stageState(state);
    if (!st) return st;
    const std::size_t n = coefficientBytes_ / sizeof(double);
    const double* y = dCoefficients_;
    double* f = dCoefficients_ + 3 * n;
    int rc;
    if (uvwIn) {
        if (!(st = toDevice(uvwIn, 3 * volumeBytes_, dVolumes_))) return st;
        rc = wv_nonlinear_flux_using_advection_fields_device(device_, state.t, state.t0, y, y + n, y + 2 * n, dVolumes_, f,
                                                             f + n, f + 2 * n, nullptr);
    } else if (uvwOut) {
        rc = wv_nonlinear_flux_with_advection_fields_device(device_, state.t, state.t0, y, y + n, y + 2 * n, f, f + n,
                                                            f + 2 * n, dVolumes_, nullptr);
    } else {
        rc = wv_nonlinear_flux_device(device_, state.t, state.t0, y, y + n, y + 2 * n, f, f + n, f + 2 * n, nullptr);
    }
    if (rc != WV_SUCCESS) return fromDevice(rc);
    st = toHost(f, coefficientBytes_, flux.Fp.data);
    if (st) st = toHost(f + n, coefficientBytes_, flux.Fm.data);
    if (st) st = toHost(f + 2 * n, coefficientBytes_, flux.F0.data);
    if (st && uvwOut) st = toHost(dVolumes_, 3 * volumeBytes_, uvwOut);
    return st;
}

WVKernelStatus Kernel::nonlinearFlux(const WVState& state, WVFlux& flux) {
    auto st = validateStateAndFlux(descriptor_, state, flux);
    if (!st) return st;
    return fluxCall(state, flux, nullptr, nullptr);
}

WVKernelStatus Kernel::nonlinearFluxWithAdvectionFields(const WVState& state, WVFlux& flux, WVRealFieldBundleView& advectionFields) {
    auto st = validateStateAndFlux(descriptor_, state, flux);
    if (!st) return st;
    if (!(st = checkBundle(advectionFields.data, advectionFields.shape, descriptor_.spatialShape(), 3, "Advection fields")))
        return st;
    if (!(st = checkInverseOwnership(state, advectionFields.data, 3 * volumeBytes_, coefficientBytes_))) return st;
    return fluxCall(state, flux, advectionFields.data, nullptr);
}

WVKernelStatus Kernel::nonlinearFluxUsingAdvectionFields(const WVState& state, WVFlux& flux,
                                                         const WVRealFieldBundleConstView& advectionFields) {
    auto st = validateStateAndFlux(descriptor_, state, flux);
    if (!st) return st;
    if (!(st = checkBundle(advectionFields.data, advectionFields.shape, descriptor_.spatialShape(), 3,
                           "Prepared advection fields")))
        return st;
    if (!(st = checkInverseOwnership(state, advectionFields.data, 3 * volumeBytes_, coefficientBytes_))) return st;
    return fluxCall(state, flux, nullptr, advectionFields.data);
}

// ---- tracer advection (...Kernel.cpp:992-1138) ----------------------------------
WVKernelStatus Kernel::prepareScalarAdvection() {
    Busy busy(executing_);
    if (!busy.entered()) return status(WVKernelStatusCode::reentrantExecution, "Kernel operations are not reentrant.");
    return fromDevice(wv_prepare_scalar_advection(device_));
}

WVKernelStatus Kernel::advectFGridScalar(const WVRealVolumeConstView& scalar, const WVRealFieldBundleConstView& advectionFields,
                                         bool shouldAntialias, WVRealVolumeView& rightHandSide) {
    const auto spatial = descriptor_.spatialShape();
    if (scalar.shape.first != spatial.first || scalar.shape.second != spatial.second || scalar.shape.third != spatial.third ||
        rightHandSide.shape.first != spatial.first || rightHandSide.shape.second != spatial.second ||
        rightHandSide.shape.third != spatial.third)
        return status(WVKernelStatusCode::invalidShape, "F-grid scalar input and output must have shape [Nx,Ny,Nz].");
    if (scalar.data == nullptr || rightHandSide.data == nullptr)
        return status(WVKernelStatusCode::invalidPointer, "F-grid scalar input and output pointers must be nonnull.");
    auto st = checkBundle(advectionFields.data, advectionFields.shape, spatial, 3, "Advection fields");
    if (!st) return st;
    const std::size_t bytes = volumeBytes_;
    if (overlaps(scalar.data, bytes, rightHandSide.data, bytes) || overlaps(advectionFields.data, 3 * bytes, scalar.data, bytes) ||
        overlaps(advectionFields.data, 3 * bytes, rightHandSide.data, bytes))
        return status(WVKernelStatusCode::overlappingArrays, "Tracer scalar, velocity, and RHS arrays must not overlap.");
    Busy busy(executing_);
    if (!busy.entered()) return status(WVKernelStatusCode::reentrantExecution, "Kernel operations are not reentrant.");
    const std::size_t n = volumeBytes_ / sizeof(double);
    double* dScalar = dVolumes_;
    double* dAdv = dVolumes_ + n;
    double* dRhs = dVolumes_ + 4 * n;
    if (!(st = toDevice(scalar.data, bytes, dScalar))) return st;
    if (!(st = toDevice(advectionFields.data, 3 * bytes, dAdv))) return st;
    const int rc = wv_advect_f_grid_scalar_device(device_, dScalar, dAdv, shouldAntialias ? 1 : 0, dRhs, nullptr);
    if (rc != WV_SUCCESS) return fromDevice(rc);
    return toHost(dRhs, bytes, rightHandSide.data);
}

}  // namespace wavevortex
