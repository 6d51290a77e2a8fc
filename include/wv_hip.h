/*
 * wv_hip.h -- C-ABI of the MI355X (gfx950) wave-vortex nonlinear-flux / RK4
 * library (libwvamd.so).
 *
 * The reference exposes this path as the C++ class contract
 * wavevortex::WVTransformConstantStratificationKernel (contract version 4,
 * CompiledKernel/include/WaveVortexKernel/WVTransformConstantStratificationKernel.hpp:47-110)
 * and the fixed-step integrator wavevortex::runtime::WVFixedStepRK4
 * (PortableRuntime/include/WaveVortexRuntime/WVRungeKutta.hpp).  Every entry
 * point below names the reference member it replaces.  All entry points use
 * plain pointers and sizes, return a wv_status code (never throw) and leave a
 * message readable with wv_last_error().
 *
 * Layouts are the reference's (WVKernelTypes.hpp:66-127):
 *   coefficients  [Nj, Nkl] column-major complex (offset j + Nj*mode), stored as
 *                 interleaved (re, im) doubles (WVComplex64);
 *   spatial fields [Nx, Ny, Nz, Nfield] column-major doubles (x fastest).
 *
 * "_device" entry points take device pointers (hipMalloc'd, resident in HBM)
 * and enqueue on `stream` (a hipStream_t; NULL = the default stream); they
 * return once the work is enqueued, except where a result is returned on
 * the host (energy, rk4 time), which synchronizes the stream.  The entry
 * points without the suffix take host pointers, copy in and out, and
 * synchronize (the reference's host-memory semantics).
 *
 * Like the reference kernel, a wv_kernel is not reentrant: a call made while
 * another call on the same handle is still inside the library returns
 * WV_REENTRANT_EXECUTION.  Calls on one handle share its device workspaces;
 * each call records an event on its stream once its work is enqueued, and a
 * later call on a different stream makes that stream wait for it, so the
 * calls on one handle run on the device in the order they were made,
 * whatever streams they use.  Distinct handles are independent.
 */
#ifndef WV_HIP_H
#define WV_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WV_CONTRACT_VERSION 4u /* WVKernelTypes.hpp:11 WVKernelContractVersion */

/* WVKernelStatusCode, WVKernelTypes.hpp:22-35 (same values and order). */
typedef enum wv_status {
    WV_SUCCESS = 0,
    WV_INVALID_CONFIGURATION,
    WV_INVALID_SHAPE,
    WV_INVALID_POINTER,
    WV_OVERLAPPING_ARRAYS,
    WV_SIZE_OVERFLOW,
    WV_ALLOCATION_FAILURE,
    WV_FFT_PLAN_FAILURE,
    WV_FFT_EXECUTION_FAILURE,
    WV_NUMERICAL_FAILURE,
    WV_UNSUPPORTED_OPERATION,
    WV_REENTRANT_EXECUTION
} wv_status;

/* Arithmetic used inside the flux.  FP64 everywhere (WV_PRECISION_FP64), or
 * FP32 spatial fields, spectra and FFTs with FP64 coefficients, RK4 state and
 * reductions, the vertical transforms on FP32 MFMA (WV_PRECISION_MIXED) or,
 * for the constant-N synthesis, on split-FP16 MFMA products: operands scaled
 * by powers of two and held as FP16 hi/lo halves, three products summed in
 * FP32 (WV_PRECISION_MIXED_SPLIT16). */
typedef enum wv_precision {
    WV_PRECISION_FP64 = 0,
    WV_PRECISION_MIXED = 1,
    WV_PRECISION_MIXED_SPLIT16 = 2
} wv_precision;

/* WVTransformConstantStratificationConfiguration, WVKernelTypes.hpp:129-146. */
typedef struct wv_config {
    uint32_t contractVersion;
    size_t Nx, Ny, Nz, Nj;
    double Lx, Ly, Lz, N0, rho0, g, planetaryRadius, rotationRate, latitude;
    int32_t isHydrostatic;
    int32_t shouldAntialias;
} wv_config;

typedef struct wv_kernel wv_kernel;

/* WVKernelMetrics (...Kernel.hpp:12-45): same members, order and meaning.
 *   descriptorBytes   device bytes of the descriptor tables and column maps
 *   planBytes         device bytes of the transform operators and twiddles
 *   engineBytes, kernelManagementBytes  0 / host bytes of the handle
 *   scratchCapacityBytes = halfSpectrumScratchCapacityBytes (the half-spectrum
 *                     and coefficient workspaces) + realScratchCapacityBytes
 *                     (0: no real field is ever stored by the flux)
 *   planCount         the 17 logical transform plans of the contract's
 *                     schedule (PlanIndex, ...Kernel.cpp:47-55), all prepared
 *                     at create, + 1 once scalar advection is prepared
 *                     (ensureScalarInversePlan, :992-1002)
 *   execution counts  the reference's plan executions that each call stands
 *                     for (per flux: 1 + 2 per target horizontal, 3 + 4 per
 *                     target vertical), counted per call
 *   *Seconds          device time of the last flux (phase = evolved spectra,
 *                     reconstruction = synthesis + inverse y transform,
 *                     product = fused x transforms and products, projection =
 *                     forward y transform + analysis + projection) when stage
 *                     instrumentation is on, else 0 */
typedef struct wv_kernel_metrics {
    size_t descriptorBytes, planBytes, engineBytes, kernelManagementBytes, scratchCapacityBytes,
        scratchHighWaterBytes, halfSpectrumScratchCapacityBytes, realScratchCapacityBytes, planCount, executionCount,
        horizontalExecutionCount, verticalExecutionCount, nonlinearFluxCallCount, nonlinearFluxPhaseEvaluationCount,
        advectionVelocityReconstructionCount, scalarAdvectionCount, scalarAntialiasCount, bytesCopied;
    double phaseSeconds, reconstructionSeconds, derivativeReconstructionSeconds, productSeconds, projectionSeconds,
        coefficientAssemblySeconds, derivativeCoefficientAssemblySeconds, coefficientProjectionSeconds,
        scalarForwardSeconds, scalarDerivativeAssemblySeconds, scalarVerticalDerivativeSeconds, scalarInverseSeconds,
        scalarProductSeconds, scalarAntialiasSeconds;
} wv_kernel_metrics;

/* Per-stage device time of the last calls (WVKernelMetrics, ...Kernel.hpp:12-45),
 * filled from HIP events when stage instrumentation is on. */
typedef struct wv_stage_timing {
    char name[32];
    double total_ms;
    uint64_t launches;
    double bytes_per_launch; /* algorithmic HBM bytes one launch moves */
} wv_stage_timing;

/* --- lifetime ----------------------------------------------------------- */
/* WVTransformConstantStratificationKernel::create (...Kernel.cpp:442-491). The
 * FFT engine argument of the reference has no counterpart: the transforms are
 * this library's own HIP kernels.  `device` selects the HIP device.  Any Nx and
 * Ny (the reference's engine takes any length) up to 4096, bounded by the LDS
 * of one workgroup: Nx <= 1024 in FP64 and <= 2048 in the mixed forms;
 * larger grids return WV_UNSUPPORTED_OPERATION. */
wv_status wv_create(const wv_config* config, wv_precision precision, int device, wv_kernel** kernel);
void wv_destroy(wv_kernel* kernel);
/* Vertical operators of an arbitrary stratification N^2(z) with a rigid lid, as
 * WVGeometryDoublyPeriodicStratified holds them
 * (@WVGeometryDoublyPeriodicStratified/WVGeometryDoublyPeriodicStratified.m:117,
 * @WVStratification/verticalProjectionOperatorsWithRigidLid.m:17-72); matrices
 * column-major as in MATLAB.  h[0] (the barotropic entry) is not used. */
typedef struct wv_vertical_operators {
    const double* z;      /* [Nz] quadrature points, -Lz .. 0 */
    const double* N2;     /* [Nz] N^2 at z */
    const double* dLnN2;  /* [Nz] d ln N^2 / dz at z */
    const double* h;      /* [Nj] equivalent depths h_0 */
    const double* P;      /* [Nj] F preconditioner P0 */
    const double* Q;      /* [Nj] G preconditioner Q0 */
    const double* PFinv;  /* [Nz x Nj] PF0inv */
    const double* PF;     /* [Nj x Nz] PF0 */
    const double* QGinv;  /* [Nz x Nj] QG0inv */
    const double* QG;     /* [Nj x Nz] QG0 */
} wv_vertical_operators;

/* The hydrostatic transform of an arbitrary stratification
 * (WVTransformHydrostatic, @WVTransformHydrostatic/WVTransformHydrostatic.m;
 * nonlinearFlux :381-393 with WVNonlinearAdvection.addHydrostaticSpatialForcing,
 * Forcing/WVNonlinearAdvection.m:103-107).  config->isHydrostatic must be set;
 * config->N0 is a representative buoyancy frequency (scaling only).  The
 * handle supports every entry point below with the dense F/G operators in
 * place of the DCT/DST. */
wv_status wv_create_hydrostatic(const wv_config* config, const wv_vertical_operators* operators,
                                wv_precision precision, int device, wv_kernel** kernel);

/* Message of the last failing call on this thread (or of the last create). */
const char* wv_last_error(void);
uint32_t wv_contract_version(void);

/* --- descriptor (WVTransformConstantStratificationDescriptor, WVKernelTypes.hpp:214-233) */
size_t wv_nkl(const wv_kernel* kernel);
double wv_coriolis_frequency(const wv_kernel* kernel);
/* fourierModes() + halfSpectrumMappings() per WV mode; any pointer may be NULL. */
wv_status wv_descriptor_modes(const wv_kernel* kernel, int64_t* kMode, int64_t* lMode, double* k, double* l,
                              double* Kh, double* cosAlpha, double* sinAlpha, int64_t* storageRow,
                              uint8_t* conjugated);
/* verticalModes() tables by name ("omega", "UApField", ..., plus the total-energy
 * factors "ApmTE"/"A0TE"); complex tables return 2*count doubles.  Returns the
 * number of doubles written, -1 if unknown.  out may be NULL to query the size. */
long wv_descriptor_table(const wv_kernel* kernel, const char* name, double* out);
/* bytes held on the device by this kernel (descriptor tables + workspaces);
 * WVTransformConstantStratificationKernel::persistentBytes (...Kernel.cpp:514-519). */
size_t wv_persistent_bytes(const wv_kernel* kernel);
/* scratchBytes (...Kernel.hpp:70): bytes of the kernel's workspaces (half
 * spectra, coefficient work arrays). */
size_t wv_scratch_bytes(const wv_kernel* kernel);
/* metrics() (...Kernel.cpp:522-537): copies the counters into *out. */
wv_status wv_metrics(const wv_kernel* kernel, wv_kernel_metrics* out);
/* Identifier accessors (...Kernel.hpp:59-67, ...Kernel.cpp:539-568): static
 * strings naming this implementation; NULL for a null kernel. */
const char* wv_engine_identifier(const wv_kernel* kernel);
const char* wv_engine_library_identity(const wv_kernel* kernel);
const char* wv_nonlinear_flux_schedule_identifier(const wv_kernel* kernel);
const char* wv_phase_implementation_identifier(const wv_kernel* kernel);
const char* wv_coefficient_arithmetic_mode_identifier(const wv_kernel* kernel);
const char* wv_inverse_normalization_placement_identifier(const wv_kernel* kernel);
const char* wv_optimization_implementation_identifier(const wv_kernel* kernel);
/* coefficientWorkerCount (:563): 0, no host worker threads. */
size_t wv_coefficient_worker_count(const wv_kernel* kernel);
/* phaseReservationBytes (:565-568): NxHalf * Ny * Nz complex values. */
size_t wv_phase_reservation_bytes(const wv_kernel* kernel);

/* --- the flux ---------------------------------------------------------------
 * WVTransformConstantStratificationKernel::nonlinearFlux (...Kernel.cpp:1140-1146,
 * 1173-1244): F = the wave-vortex projection of -(u . grad)(u,v,w,eta), evaluated
 * at time t and rotated back to t0.  Outputs may not overlap each other or the
 * inputs (validateStateAndFlux, WVKernelTypes.cpp:343-362). */
wv_status wv_nonlinear_flux_device(wv_kernel* kernel, double t, double t0, const double* Ap, const double* Am,
                                   const double* A0, double* Fp, double* Fm, double* F0, void* stream);
wv_status wv_nonlinear_flux(wv_kernel* kernel, double t, double t0, const double* Ap, const double* Am,
                            const double* A0, double* Fp, double* Fm, double* F0);

/* nonlinearFluxWithAdvectionFields (...Kernel.cpp:1148-1158): also returns the
 * advecting (u,v,w) as a [Nx,Ny,Nz,3] bundle.  Every argument is validated
 * before anything is written. */
wv_status wv_nonlinear_flux_with_advection_fields_device(wv_kernel* kernel, double t, double t0, const double* Ap,
                                                         const double* Am, const double* A0, double* Fp,
                                                         double* Fm, double* F0, double* uvw, void* stream);

/* nonlinearFluxUsingAdvectionFields (...Kernel.cpp:1160-1171): the flux with
 * the caller's advecting (u,v,w) [Nx,Ny,Nz,3] in the products instead of the
 * state's own (the fields a previous nonlinearFluxWithAdvectionFields
 * returned, shared with tracer advection).  uvw may not overlap the state or
 * the flux. */
wv_status wv_nonlinear_flux_using_advection_fields_device(wv_kernel* kernel, double t, double t0, const double* Ap,
                                                          const double* Am, const double* A0, const double* uvw,
                                                          double* Fp, double* Fm, double* F0, void* stream);

/* prepareScalarAdvection (...Kernel.cpp:1004-1006): allocates the scalar
 * advection path's workspace now, so that advectFGridScalar calls allocate
 * nothing.  Idempotent. */
wv_status wv_prepare_scalar_advection(wv_kernel* kernel);

/* advectFGridScalar (...Kernel.cpp:1032-1138): rightHandSide = -(u, v, w) .
 * grad(scalar) for an F-grid scalar [Nx,Ny,Nz] advected by the velocity bundle
 * advectionFields [Nx,Ny,Nz,3] (u, v, w), then antialiased to the retained
 * modes' rows when shouldAntialias is nonzero (:1008-1030).  FP64 in either
 * precision.  Without a prior wv_prepare_scalar_advection the first call
 * allocates the path's scratch (three half spectra), as the reference's
 * first call creates its plan (:1009); later calls allocate nothing.  Same
 * null-pointer, overlap and reentrancy errors as the reference. */
wv_status wv_advect_f_grid_scalar_device(wv_kernel* kernel, const double* scalar, const double* advectionFields,
                                         int shouldAntialias, double* rightHandSide, void* stream);

/* --- transforms ------------------------------------------------------------- */
/* transformWaveVortexToUVWEta (...Kernel.cpp:686-694,745-776): fields [Nx,Ny,Nz,4]. */
wv_status wv_transform_wave_vortex_to_uvweta_device(wv_kernel* kernel, double t, double t0, const double* Ap,
                                                    const double* Am, const double* A0, double* fields,
                                                    void* stream);
/* transformWaveVortexToUVW (...Kernel.cpp:696-717): fields [Nx,Ny,Nz,3]. */
wv_status wv_transform_wave_vortex_to_uvw_device(wv_kernel* kernel, double t, double t0, const double* Ap,
                                                 const double* Am, const double* A0, double* fields,
                                                 void* stream);
/* transformUVWEtaToWaveVortex (nonhydrostatic, fields [..,4]) and
 * transformUVEtaToWaveVortex (hydrostatic, fields [..,3]) (...Kernel.cpp:583-684). */
wv_status wv_transform_uvweta_to_wave_vortex_device(wv_kernel* kernel, const double* fields, double t, double t0,
                                                    double* Ap, double* Am, double* A0, void* stream);
wv_status wv_transform_uveta_to_wave_vortex_device(wv_kernel* kernel, const double* fields, double t, double t0,
                                                   double* Ap, double* Am, double* A0, void* stream);

/* transformStateFieldDerivatives (...Kernel.cpp:719-742): d/dx, d/dy, d/dz of
 * the dynamical field `field` (0 u, 1 v, 2 w, 3 eta; WVDynamicalField,
 * ...Kernel.hpp:10) of the state at time t, derivatives [Nx,Ny,Nz,3]. */
wv_status wv_transform_state_field_derivatives_device(wv_kernel* kernel, double t, double t0, const double* Ap,
                                                      const double* Am, const double* A0, int field,
                                                      double* derivatives, void* stream);

/* Available potential vorticity of the state at time t, apv [Nx,Ny,Nz]:
 *   zeta_z - zeta_x d/dx eta - zeta_y d/dy eta - (zeta_z + f) d/dz eta
 * (Operations/APVOperation.m:15-26; zeta as in
 * @WVTransform/classDefinedOperationForKnownVariable.m:126-137).  The
 * reference forms it from eta_true (Operations/EtaTrueOperation.m), which for
 * the constant-N background used as the no-motion profile (the default,
 * @WVTransformConstantStratification/WVTransformConstantStratification.m:110)
 * is eta itself; arbitrary-stratification kernels return
 * WV_UNSUPPORTED_OPERATION. */
wv_status wv_apv_device(wv_kernel* kernel, double t, double t0, const double* Ap, const double* Am,
                        const double* A0, double* apv, void* stream);

/* transformToSpatialDomainWithFAllDerivatives / ...GAllDerivatives (...Kernel.cpp:819-863):
 * the F-family (cosine) or G-family (sine) field Apm*waveScale + A0*geostrophicScale
 * and its x, y, z derivatives, fields [Nx,Ny,Nz,4] = (value, d/dx, d/dy, d/dz). */
wv_status wv_transform_to_spatial_domain_with_f_all_derivatives_device(wv_kernel* kernel, const double* Apm,
                                                                       const double* A0, double* fields,
                                                                       void* stream);
wv_status wv_transform_to_spatial_domain_with_g_all_derivatives_device(wv_kernel* kernel, const double* Apm,
                                                                       const double* A0, double* fields,
                                                                       void* stream);

/* --- fixed-step RK4 ----------------------------------------------------------
 * WVFixedStepRK4::step / advanceToTime (PortableRuntime/src/WVRungeKutta.cpp:335-480)
 * for the nonlinear-advection system: nsteps classic RK4 steps of size h, state
 * (Ap, Am, A0) updated in place on the device, *t advanced on the host.  The
 * stage combinations are fused into the flux projection. */
wv_status wv_rk4_advance_device(wv_kernel* kernel, double* Ap, double* Am, double* A0, double* t, double t0,
                                double h, int nsteps, void* stream);

/* WVFixedStepRK4::advanceToTime (WVRungeKutta.cpp:464-480): steps of size
 * min(h, finalTime - t) while t < finalTime (the last one clipped), then
 * t = finalTime.  finalTime below *t or not finite: WV_INVALID_CONFIGURATION.
 * Each step is committed by its last launch (WVRungeKutta.cpp:441): if a step
 * fails, the state and *t are those after the last whole step. */
wv_status wv_rk4_advance_to_time_device(wv_kernel* kernel, double* Ap, double* Am, double* A0, double* t, double t0,
                                        double finalTime, double h, void* stream);

/* WVFixedStepRK4 metrics (WVRungeKutta.hpp): steps taken on this handle and
 * the size of the last one. */
wv_status wv_rk4_counters(const wv_kernel* kernel, uint64_t* stepCount, double* lastStepSize);

/* Total energy sum(ApmTE*(|Ap|^2+|Am|^2) + A0TE*|A0|^2), FP64 reduction
 * (@WVTransformConstantStratification/WVTransformConstantStratification.m:374-376). */
wv_status wv_total_energy_device(wv_kernel* kernel, const double* Ap, const double* Am, const double* A0,
                                 double* energy, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate or
 * synchronize (work under HIP graph capture, serving loops): the reduction of
 * wv_total_energy_device with its partial sums in a library-wide static device
 * array (1024 doubles per GPU, held until the process ends) instead of the
 * handle's workspace, and the result written to the device pointer `energy`.
 * Nothing is allocated, copied to the host or synchronized; the work is
 * enqueued on `stream`.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (different streams, no ordering between them) race.
 * Coefficient counts above INT_MAX - 2^18 return WV_SIZE_OVERFLOW. */
wv_status wv_total_energy_no_alloc_device(wv_kernel* kernel, const double* Ap, const double* Am, const double* A0,
                                          double* energy, void* stream);

/* --- adjoint (vector-Jacobian products) ---------------------------------------
 * No counterpart in the reference, which has no adjoint.  Cotangents use the
 * real inner product <a, b> = Re sum conj(a) b over the coefficient arrays (and
 * sum a b over real fields), so that <J dy, L> = <dy, J^T L>; the gradient of a
 * real cost with respect to a complex coefficient is dJ/dRe + i dJ/dIm.  The
 * first adjoint call on a handle allocates the adjoint workspace (transposed
 * operators, three spectral buffers and seven coefficient triples). */

/* (Gp, Gm, G0) = J^T (Lp, Lm, L0), J the Jacobian of nonlinearFlux at (t, state). */
wv_status wv_nonlinear_flux_vjp_device(wv_kernel* kernel, double t, double t0, const double* Ap, const double* Am,
                                       const double* A0, const double* Lp, const double* Lm, const double* L0,
                                       double* Gp, double* Gm, double* G0, void* stream);

/* Transpose of transformWaveVortexToUVWEta: field cotangents [Nx,Ny,Nz,4] ->
 * coefficient cotangents. */
wv_status wv_transform_wave_vortex_to_uvweta_vjp_device(wv_kernel* kernel, const double* fieldsBar, double t,
                                                        double t0, double* Gp, double* Gm, double* G0, void* stream);

/* One RK4 step of size h from the state (Ap, Am, A0) at time t: on entry
 * (Lp, Lm, L0) is the cotangent of the state after the step, on return that of
 * the state before it.  The stages are recomputed from the given state. */
wv_status wv_rk4_step_vjp_device(wv_kernel* kernel, const double* Ap, const double* Am, const double* A0, double t,
                                 double t0, double h, double* Lp, double* Lm, double* L0, void* stream);

/* nsteps RK4 steps from (Ap, Am, A0) at time t (the state is not modified): on
 * entry (Lp, Lm, L0) is the cotangent of the final state, on return that of the
 * initial state.  Forward states are recomputed from checkpoints kept every
 * `interval` steps; device memory for ceil(nsteps/interval) + interval states
 * is held for the duration of the call. */
wv_status wv_rk4_adjoint_device(wv_kernel* kernel, const double* Ap, const double* Am, const double* A0, double t,
                                double t0, double h, int nsteps, int interval, double* Lp, double* Lm, double* L0,
                                void* stream);

/* --- instrumentation (setStageInstrumentation, ...Kernel.hpp:68) ----------- */
wv_status wv_set_stage_instrumentation(wv_kernel* kernel, int enabled);
/* Copies up to `capacity` stage records; returns the number of stages. */
int wv_stage_timings(wv_kernel* kernel, wv_stage_timing* out, int capacity);
void wv_reset_stage_timings(wv_kernel* kernel);

/* Test hook, the counterpart of the reference tests' failing FFT engine
 * (PortableRuntime/tests/TestWVForcingAndRK4.cpp:313-320): with n > 0, the
 * n-th next kernel launch of the flux pipeline is not made and its call
 * fails with WV_FFT_EXECUTION_FAILURE; n = 0 disarms it. */
wv_status wv_debug_inject_launch_failure(wv_kernel* kernel, long n);

#ifdef __cplusplus
}
#endif
#endif /* WV_HIP_H */
