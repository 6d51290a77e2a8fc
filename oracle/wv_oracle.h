/*
 * CPU oracle for the wave-vortex nonlinear-flux / RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain FP64 CPU restatement of
 * the reference algorithm (CompiledKernel + WVFixedStepRK4).  It exists so the
 * tests, the smoke check and the CPU-baseline leg of bench.py can check the
 * HIP product path against it.  Nothing in the product links, loads or calls
 * it.
 *
 * Complex arrays are interleaved (re, im) doubles.  Coefficient arrays are
 * column-major [Nj, Nkl] (offset j + Nj*mode); spatial bundles are
 * [Nx, Ny, Nz, Nfield] with x fastest.
 *
 * Parity status per function (what pins it; see the test files):
 *   wvo_create / descriptor tables ..... pinned: golden vectors from oracle/_ref
 *                                         (reference descriptor) + closed forms
 *   wvo_nonlinear_flux ................. pinned: oracle/_ref golden fluxes,
 *                                         zero-state, energy-conservation invariant
 *   wvo_inverse_uvweta / wvo_inverse_uvw pinned: oracle/_ref golden fields,
 *                                         single-mode closed form
 *   wvo_forward_uvweta / wvo_forward_uveta pinned: oracle/_ref golden coefficients,
 *                                         forward(inverse(A)) = A round trip
 *   wvo_fg_all_derivatives ............. pinned: oracle/_ref golden fields
 *   wvo_nonlinear_flux_using_advection_fields, wvo_state_field_derivatives
 *                                        pinned: oracle/_ref golden vectors
 *                                         (tests/golden/contract_*.npz)
 *   wvo_rk4_advance .................... pinned: linear test y'=lambda*y closed
 *                                         form (via wvo_rk4_linear_test) and the
 *                                         reference step arithmetic
 *   wvo_energy ......................... pinned: spectral energy equals the
 *                                         spatially integrated energy
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    size_t Nx, Ny, Nz, Nj;
    double Lx, Ly, Lz, N0, rho0, g, planetaryRadius, rotationRate, latitude;
    int32_t isHydrostatic;
    int32_t shouldAntialias;
} wvo_config;

/* Returns NULL and fills msg on an invalid configuration. */
void* wvo_create(const wvo_config* config, char* msg, size_t msglen);
void wvo_destroy(void* h);
void wvo_set_threads(int threads);

size_t wvo_nkl(void* h);
/* kMode,lMode[Nkl]; k,l,Kh,cosAlpha,sinAlpha[Nkl]; storageRow[Nkl]; conj[Nkl] */
void wvo_modes(void* h, int64_t* kMode, int64_t* lMode, double* k, double* l, double* Kh,
               double* cosAlpha, double* sinAlpha, int64_t* storageRow, uint8_t* conj);
/* Named [Nj,Nkl] or [Nj] or [Nz] table; complex tables return 2*count doubles.
 * Returns the number of doubles written, or -1 for an unknown name. */
long wvo_table(void* h, const char* name, double* out);
double wvo_coriolis(void* h);

int wvo_nonlinear_flux(void* h, double t, double t0, const double* Ap, const double* Am,
                       const double* A0, double* Fp, double* Fm, double* F0);
/* nonlinearFluxUsingAdvectionFields: the products use uvw [Nx,Ny,Nz,3]. */
int wvo_nonlinear_flux_using_advection_fields(void* h, double t, double t0, const double* Ap, const double* Am,
                                              const double* A0, const double* uvw, double* Fp, double* Fm,
                                              double* F0);
/* transformStateFieldDerivatives: d/dx, d/dy, d/dz of field (0 u,1 v,2 w,3 eta), [Nx,Ny,Nz,3]. */
int wvo_state_field_derivatives(void* h, double t, double t0, const double* Ap, const double* Am, const double* A0,
                                int field, double* out);
int wvo_inverse_uvweta(void* h, double t, double t0, const double* Ap, const double* Am,
                       const double* A0, double* fields);
int wvo_inverse_uvw(void* h, double t, double t0, const double* Ap, const double* Am,
                    const double* A0, double* fields);
int wvo_forward_uvweta(void* h, const double* fields, double t, double t0, double* Ap,
                       double* Am, double* A0);
int wvo_forward_uveta(void* h, const double* fields, double t, double t0, double* Ap,
                      double* Am, double* A0);
int wvo_fg_all_derivatives(void* h, int fFamily, const double* Apm, const double* A0,
                           double* fields);
/* Fixed-step RK4 (reference step arithmetic), nsteps steps of size h from *t. */
int wvo_rk4_advance(void* h, double* Ap, double* Am, double* A0, double* t, double t0,
                    double dt, int nsteps);
/* Same stage arithmetic applied to y' = lambda*y (no flux); for the closed-form check. */
void wvo_rk4_linear_test(double lambda_re, double lambda_im, double* y, double dt, int nsteps);
double wvo_energy(void* h, const double* Ap, const double* Am, const double* A0);
/* sum_z w_z * mean_xy(u^2+v^2[+w^2]+N0^2 eta^2)/2 from [Nx,Ny,Nz,4] fields */
double wvo_energy_spatial(void* h, const double* fields);

#ifdef __cplusplus
}
#endif
