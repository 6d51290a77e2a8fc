// Drop-in replacement of the kernel-contract header
// CompiledKernel/include/WaveVortexKernel/WVTransformConstantStratificationKernel.hpp
// (contract version 4) backed by libwvamd.so on one MI355X.
//
// The public members, their signatures and WVKernelMetrics / WVDynamicalField
// are the contract's (...Kernel.hpp:10-83), so that its callers
// (PortableRuntime's WVConstantStratificationForcingEngine and integration
// system, the MEX gateway) compile against this header unchanged, with the
// contract's own WVKernelTypes.hpp and WVFFTEngine.hpp.  The private section
// is this implementation's: a libwvamd handle and device staging buffers
// allocated at create (the calls take host-memory views, as the contract
// does, and copy them to and from the device).
//
// Placing this directory before CompiledKernel/include on the include path
// and linking wv_hip_kernel_adapter.cpp + libwvamd.so in place of
// WVTransformConstantStratificationKernel.cpp is the whole integration.
#pragma once

#include "WaveVortexKernel/WVFFTEngine.hpp"

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

struct wv_kernel;

namespace wavevortex {

enum class WVDynamicalField : std::uint8_t { u, v, w, eta };

struct WVKernelMetrics {
    std::size_t descriptorBytes = 0;
    std::size_t planBytes = 0;
    std::size_t engineBytes = 0;
    std::size_t kernelManagementBytes = 0;
    std::size_t scratchCapacityBytes = 0;
    std::size_t scratchHighWaterBytes = 0;
    std::size_t halfSpectrumScratchCapacityBytes = 0;
    std::size_t realScratchCapacityBytes = 0;
    std::size_t planCount = 0;
    std::size_t executionCount = 0;
    std::size_t horizontalExecutionCount = 0;
    std::size_t verticalExecutionCount = 0;
    std::size_t nonlinearFluxCallCount = 0;
    std::size_t nonlinearFluxPhaseEvaluationCount = 0;
    std::size_t advectionVelocityReconstructionCount = 0;
    std::size_t scalarAdvectionCount = 0;
    std::size_t scalarAntialiasCount = 0;
    std::size_t bytesCopied = 0;
    double phaseSeconds = 0.0;
    double reconstructionSeconds = 0.0;
    double derivativeReconstructionSeconds = 0.0;
    double productSeconds = 0.0;
    double projectionSeconds = 0.0;
    double coefficientAssemblySeconds = 0.0;
    double derivativeCoefficientAssemblySeconds = 0.0;
    double coefficientProjectionSeconds = 0.0;
    double scalarForwardSeconds = 0.0;
    double scalarDerivativeAssemblySeconds = 0.0;
    double scalarVerticalDerivativeSeconds = 0.0;
    double scalarInverseSeconds = 0.0;
    double scalarProductSeconds = 0.0;
    double scalarAntialiasSeconds = 0.0;
};

class WVTransformConstantStratificationKernel {
public:
    // The engine is kept for its identity (engineIdentifier,
    // engineLibraryIdentity, engineBytes); the transforms are libwvamd's.
    // Environment: WVAMD_DEVICE selects the HIP device (default 0),
    // WVAMD_PRECISION=mixed the mixed-precision flux (default FP64).
    static WVKernelStatus create(const WVTransformConstantStratificationConfiguration& configuration,
                                 std::unique_ptr<WVFFTEngine> engine,
                                 std::unique_ptr<WVTransformConstantStratificationKernel>& kernel);

    ~WVTransformConstantStratificationKernel();
    WVTransformConstantStratificationKernel(const WVTransformConstantStratificationKernel&) = delete;
    WVTransformConstantStratificationKernel& operator=(const WVTransformConstantStratificationKernel&) = delete;
    WVTransformConstantStratificationKernel(WVTransformConstantStratificationKernel&&) = delete;
    WVTransformConstantStratificationKernel& operator=(WVTransformConstantStratificationKernel&&) = delete;

    const WVTransformConstantStratificationDescriptor& descriptor() const noexcept { return descriptor_; }
    const WVKernelMetrics& metrics() const noexcept;
    const std::string& engineIdentifier() const noexcept { return engineIdentifier_; }
    const std::string& engineLibraryIdentity() const noexcept { return engineLibraryIdentity_; }
    const char* nonlinearFluxScheduleIdentifier() const noexcept;
    const char* phaseImplementationIdentifier() const noexcept;
    const char* coefficientArithmeticModeIdentifier() const noexcept;
    const char* inverseNormalizationPlacementIdentifier() const noexcept;
    const char* optimizationImplementationIdentifier() const noexcept;
    std::size_t coefficientWorkerCount() const noexcept;
    std::size_t phaseReservationBytes() const noexcept;
    void setStageInstrumentation(bool enabled) noexcept;
    std::size_t persistentBytes() const noexcept;
    std::size_t scratchBytes() const noexcept;

    WVKernelStatus transformUVEtaToWaveVortex(const WVRealFieldBundleConstView& fields, double t, double t0, WVMutableCoefficients& coefficients);
    WVKernelStatus transformUVWEtaToWaveVortex(const WVRealFieldBundleConstView& fields, double t, double t0, WVMutableCoefficients& coefficients);
    WVKernelStatus transformWaveVortexToUVWEta(const WVState& state, WVRealFieldBundleView& fields);
    WVKernelStatus transformWaveVortexToUVW(const WVState& state, WVRealFieldBundleView& fields);
    WVKernelStatus transformStateFieldDerivatives(const WVState& state, WVDynamicalField field, WVRealFieldBundleView& derivatives);
    WVKernelStatus transformToSpatialDomainWithFAllDerivatives(const WVComplexConstView& Apm, const WVComplexConstView& A0, WVRealFieldBundleView& fields);
    WVKernelStatus transformToSpatialDomainWithGAllDerivatives(const WVComplexConstView& Apm, const WVComplexConstView& A0, WVRealFieldBundleView& fields);
    WVKernelStatus nonlinearFlux(const WVState& state, WVFlux& flux);
    WVKernelStatus nonlinearFluxWithAdvectionFields(const WVState& state, WVFlux& flux, WVRealFieldBundleView& advectionFields);
    WVKernelStatus nonlinearFluxUsingAdvectionFields(const WVState& state, WVFlux& flux, const WVRealFieldBundleConstView& advectionFields);
    WVKernelStatus prepareScalarAdvection();
    WVKernelStatus advectFGridScalar(const WVRealVolumeConstView& scalar, const WVRealFieldBundleConstView& advectionFields, bool shouldAntialias, WVRealVolumeView& rightHandSide);

private:
    WVTransformConstantStratificationKernel() = default;
    WVKernelStatus fromDevice(int code) const;
    WVKernelStatus toDevice(const void* host, std::size_t bytes, void* device);
    WVKernelStatus toHost(const void* device, std::size_t bytes, void* host);
    WVKernelStatus stageState(const WVState& state);
    WVKernelStatus fluxCall(const WVState& state, WVFlux& flux, double* uvwOut, const double* uvwIn);
    WVKernelStatus familyDerivatives(bool fFamily, const WVComplexConstView& Apm, const WVComplexConstView& A0, WVRealFieldBundleView& fields);
    WVKernelStatus forwardCall(const WVRealFieldBundleConstView& fields, std::size_t channels, double t, double t0, WVMutableCoefficients& coefficients);

    WVTransformConstantStratificationDescriptor descriptor_;
    std::unique_ptr<WVFFTEngine> engine_;
    std::string engineIdentifier_;
    std::string engineLibraryIdentity_;
    wv_kernel* device_ = nullptr;
    int precision_ = 0;
    // device staging: six coefficient arrays (state, flux) and five spatial
    // volumes (fields, tracer, its right-hand side)
    double* dCoefficients_ = nullptr;
    double* dVolumes_ = nullptr;
    std::size_t coefficientBytes_ = 0;
    std::size_t volumeBytes_ = 0;
    mutable WVKernelMetrics metrics_;
    bool executing_ = false;
};

} // namespace wavevortex
