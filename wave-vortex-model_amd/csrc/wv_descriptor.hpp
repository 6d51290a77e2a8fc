// Host-side descriptor of the constant-stratification wave-vortex transform:
// retained Fourier modes, coefficient tables, the kx/iy column lists the HIP
// kernels iterate over, and the dense vertical transform matrices.
//
// The mode enumeration, canonical order and every coefficient table follow
// WVTransformConstantStratificationDescriptor::create
// (CompiledKernel/src/WVKernelTypes.cpp:119-330) value for value; the column
// lists and matrices are this library's own layout for the GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "wv_hip.h"

namespace wvamd {

struct FourierMode {
    int64_t kMode = 0, lMode = 0;
    double k = 0, l = 0, Kh = 0, cosAlpha = 0, sinAlpha = 0;
    size_t storageRow = 0;   // half-spectrum row kx + NxHalf*iy
    bool conjugated = false; // stored as conj at (-k,-l)
};

// One nonzero row of the half spectrum (kx >= 0) that the vertical transforms
// and the y-FFTs must populate: either the storage row of a WV mode or a
// Hermitian-completion row on kx = 0.
enum ColumnFlags : uint32_t { kColConj = 1u, kColSelfConj = 2u, kColCompletion = 4u };
struct Column {
    int32_t iy = 0, kx = 0, wv = 0;
    uint32_t flags = 0;
};

// Vertical structure of an arbitrary stratification N^2(z) (rigid lid,
// hydrostatic): the operators WVGeometryDoublyPeriodicStratified holds
// (@WVGeometryDoublyPeriodicStratified/WVGeometryDoublyPeriodicStratified.m:117,
// @WVStratification/verticalProjectionOperatorsWithRigidLid.m:17-72), all
// column-major as in MATLAB.
struct Stratification {
    std::vector<double> z, N2, dLnN2;  // [Nz] quadrature points and N^2, d ln N^2 / dz there
    std::vector<double> h, P, Q;       // [Nj]
    std::vector<double> PFinv, QGinv;  // [Nz x Nj]
    std::vector<double> PF, QG;        // [Nj x Nz]
};

struct Descriptor {
    wv_config c{};
    bool stratified = false;  // dense F/G operators of an arbitrary N^2(z) instead of the DCT/DST
    size_t NxHalf = 0, Nkl = 0, M = 0, R = 0;
    size_t nkx = 0;   // 1 + largest kx present
    size_t nkxP = 0;  // nkx padded to the y-FFT tile width
    std::vector<double> kxv;  // x wavenumber of half-spectrum column kx < nkx
    double f = 0;
    // N0^2, f^2 and the square roots the per-coefficient tables share:
    // waveNorm = sqrt(2 g_modal / (Lz (N0^2 - f^2))), projPrefactor =
    // sqrt(g Lz / (2 (N0^2 - f^2)))
    double N02 = 0, f2 = 0, waveNorm = 0, projPrefactor = 0;

    std::vector<FourierMode> modes;
    std::vector<Column> columns;      // sorted by (iy, kx)
    std::vector<int32_t> wvColumn;    // column index of each WV mode
    std::vector<size_t> completionRows, completionSources, selfConjugateRows;

    // vertical-only [Nj]
    std::vector<double> z, h0, m, Fg, Gg, inertialScale, gWaveScale, apmW;
    // [Nj, Nkl] column-major
    std::vector<double> omega, fWaveScale, NAp, NA0, A0Vort, A0Buoy, ApmN, ApmDScaled, ApmTE, A0TE;
    std::vector<double> UAp, VAp, WAp, UA0, VA0, ApmD;  // complex, interleaved

    // Dense vertical operators, zero padded:
    //   syn[C|S]  [Zs x Js] row-major (z, j): spectral value at j -> physical z,
    //             the unnormalized DCT-I / DST-I restricted to j < Nj.
    //   ana[C|S]  [Ja x Za] row-major (j, z): physical z -> coefficient j, the
    //             DCT-I / DST-I followed by the forward normalization.
    // For an arbitrary stratification synC/synS hold the F and G mode values
    // (PFinv*P, QGinv*Q), anaC/anaS the projections (PF/P, QG/Q), and
    //   synGz = -(N^2/g) G            (d/dz of an F-family field, diffZF.m:30-35)
    //   synFh = F / h                 (d/dz of a G-family field, diffZG.m:26-28)
    //   synEz = F / h + dLnN2 * G     (d/dz eta + eta dLnN2, WVNonlinearAdvection.m:106)
    size_t Zs = 0, Js = 0, Ja = 0, Za = 0;
    std::vector<double> synC, synS, anaC, anaS, synGz, synFh, synEz;
    std::vector<double> twX, twY;  // FFT twiddles (Stockham pass order for powers of two), interleaved

    // Returns an empty string on success, otherwise the reference's message.
    // strat != nullptr selects the arbitrary-stratification vertical structure.
    std::string build(const wv_config& config, size_t kxTile, const Stratification* strat = nullptr);
    // Named table export (same names as the reference descriptor fields).
    long table(const std::string& name, double* out) const;
};

}  // namespace wvamd
