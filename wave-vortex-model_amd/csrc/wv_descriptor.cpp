// Host-side descriptor; see wv_descriptor.hpp.  Compiled with
// -ffp-contract=off so every table rounds exactly as the reference's
// (CompiledKernel/src/WVKernelTypes.cpp is built without FMA contraction).
#include "wv_descriptor.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace wvamd {
namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
// F/G normalizations use a fixed 9.81 regardless of the configured g
// (WVKernelTypes.cpp:13-16, 206-208).
constexpr double kModalGravity = 9.81;

// sqrt(a*a + b*b) with both squares rounded separately, so tied shells sort
// exactly as MATLAB's sortrows([Kh,K,L]) (WVKernelTypes.cpp:22-29).
double shellRadius(double a, double b) {
    volatile double a2 = a * a;
    volatile double b2 = b * b;
    return std::sqrt(a2 + b2);
}

std::vector<int64_t> signedDftModes(size_t n) {  // WVKernelTypes.cpp:38-45
    std::vector<int64_t> out;
    out.reserve(n);
    for (size_t q = 0; q < (n + 1) / 2; ++q) out.push_back(static_cast<int64_t>(q));
    for (size_t q = n / 2; q > 0; --q) out.push_back(-static_cast<int64_t>(q));
    return out;
}

size_t dftIndex(int64_t mode, size_t n) { return static_cast<size_t>(mode >= 0 ? mode : static_cast<int64_t>(n) + mode); }

bool selfConjugateMode(int64_t mode, size_t n) { return mode == 0 || (n % 2 == 0 && mode == -static_cast<int64_t>(n / 2)); }

size_t mulChecked(size_t a, size_t b) {
    if (a != 0 && b > std::numeric_limits<size_t>::max() / a) throw std::overflow_error("shape element count overflow");
    return a * b;
}

size_t roundUp(size_t v, size_t q) { return (v + q - 1) / q * q; }

// cos / sin of pi*p/q with the exact reduction of the integer p modulo 2q.
double cosPiRatio(size_t p, size_t q) {
    const size_t r = p % (2 * q);
    return std::cos(kPi * static_cast<double>(r) / static_cast<double>(q));
}
double sinPiRatio(size_t p, size_t q) {
    const size_t r = p % (2 * q);
    return std::sin(kPi * static_cast<double>(r) / static_cast<double>(q));
}

}  // namespace

std::string Descriptor::build(const wv_config& cfg, size_t kxTile, const Stratification* strat) {
    c = cfg;
    stratified = strat != nullptr;
    if (c.contractVersion != WV_CONTRACT_VERSION) return "Unsupported kernel contract version.";
    if (c.Nx < 2 || c.Ny < 2 || c.Nz < 3 || c.Nj < 1 || c.Nj > c.Nz - 1)
        return "Nx and Ny must be at least 2, Nz at least 3, and Nj must lie in [1,Nz-1].";
    auto finitePositive = [](double v) { return std::isfinite(v) && v > 0.0; };
    if (!finitePositive(c.Lx) || !finitePositive(c.Ly) || !finitePositive(c.Lz) || !finitePositive(c.N0) ||
        !finitePositive(c.rho0) || !finitePositive(c.g) || !finitePositive(c.planetaryRadius) ||
        !finitePositive(c.rotationRate) || !std::isfinite(c.latitude) || std::abs(c.latitude) > 90.0)
        return "Physical lengths and constants must be finite and positive, and latitude must lie in [-90,90].";
    mulChecked(c.Nx, c.Ny);
    R = mulChecked(mulChecked(c.Nx, c.Ny), c.Nz);

    // Retained primary, non-Nyquist modes inside the 2/3 disk, sorted by
    // (Kh, kMode, lMode) -- WVKernelTypes.cpp:139-166.
    const auto kModes = signedDftModes(c.Nx);
    const auto lModes = signedDftModes(c.Ny);
    const double maxAbsK = 2.0 * kPi * (static_cast<double>(c.Nx / 2) / c.Lx);
    for (const int64_t lm : lModes) {
        for (const int64_t km : kModes) {
            const bool primary = lm > 0 || (selfConjugateMode(lm, c.Ny) && (km > 0 || selfConjugateMode(km, c.Nx)));
            const bool nyquist = (c.Nx % 2 == 0 && km == -static_cast<int64_t>(c.Nx / 2)) ||
                                 (c.Ny % 2 == 0 && lm == -static_cast<int64_t>(c.Ny / 2));
            if (!primary || nyquist) continue;
            const double k = 2.0 * kPi * (static_cast<double>(km) / c.Lx);
            const double l = 2.0 * kPi * (static_cast<double>(lm) / c.Ly);
            if (c.shouldAntialias && shellRadius(k, l) > 2.0 * maxAbsK / 3.0) continue;
            FourierMode fm;
            fm.kMode = km;
            fm.lMode = lm;
            fm.k = k;
            fm.l = l;
            fm.Kh = shellRadius(k, l);
            fm.cosAlpha = fm.Kh == 0.0 ? 0.0 : k / fm.Kh;
            fm.sinAlpha = fm.Kh == 0.0 ? 0.0 : l / fm.Kh;
            modes.push_back(fm);
        }
    }
    std::stable_sort(modes.begin(), modes.end(), [](const FourierMode& a, const FourierMode& b) {
        if (a.Kh != b.Kh) return a.Kh < b.Kh;
        if (a.kMode != b.kMode) return a.kMode < b.kMode;
        return a.lMode < b.lMode;
    });
    Nkl = modes.size();
    const size_t Nj = c.Nj, Nz = c.Nz;
    M = mulChecked(Nj, Nkl);
    NxHalf = c.Nx / 2 + 1;

    // Vertical structure -- WVKernelTypes.cpp:169-212.
    f = 2.0 * c.rotationRate * std::sin(c.latitude * kPi / 180.0);
    N02 = c.N0 * c.N0;
    f2 = f * f;
    waveNorm = std::sqrt(2.0 * kModalGravity / (c.Lz * (N02 - f2)));
    projPrefactor = std::sqrt(c.g * c.Lz / (2.0 * (N02 - f2)));
    if (!c.isHydrostatic && N02 <= f2)
        return "Nonhydrostatic constant stratification requires N0 squared greater than f squared.";
    if (stratified) {
        if (!c.isHydrostatic) return "Arbitrary stratification is supported for the hydrostatic transform only.";
        const size_t nz = c.Nz, nj = c.Nj;
        if (strat->z.size() != nz || strat->N2.size() != nz || strat->dLnN2.size() != nz || strat->h.size() != nj ||
            strat->P.size() != nj || strat->Q.size() != nj || strat->PFinv.size() != nz * nj ||
            strat->QGinv.size() != nz * nj || strat->PF.size() != nj * nz || strat->QG.size() != nj * nz)
            return "Vertical operators do not match Nz and Nj.";
        for (const auto* v : {&strat->z, &strat->N2, &strat->dLnN2, &strat->h, &strat->P, &strat->Q, &strat->PFinv,
                              &strat->QGinv, &strat->PF, &strat->QG})
            for (double x : *v)
                if (!std::isfinite(x)) return "Vertical operators must be finite.";
        for (size_t j = 1; j < nj; ++j)
            if (!(strat->h[j] > 0.0)) return "Equivalent depths must be positive.";
        for (size_t j = 0; j < nj; ++j)
            if (strat->P[j] == 0.0 || strat->Q[j] == 0.0) return "Preconditioners must be nonzero.";
    }
    z.resize(Nz);
    h0.resize(Nj);
    m.resize(Nj);
    Fg.resize(Nj);
    Gg.resize(Nj);
    inertialScale.assign(Nj, 0.0);
    gWaveScale.resize(Nj);
    apmW.assign(Nj, 0.0);
    const double dz = c.Lz / static_cast<double>(Nz - 1);
    for (size_t i = 0; i < Nz; ++i) z[i] = stratified ? strat->z[i] : dz * static_cast<double>(i) - c.Lz;
    for (size_t j = 0; j < Nj; ++j) {
        // arbitrary N^2: h_pm = h_0 from the eigenproblem; the barotropic entry
        // takes Lz as in the constant-N tables (it enters the energy factors only)
        m[j] = stratified ? 0.0 : static_cast<double>(j) * kPi / c.Lz;
        h0[j] = j == 0 ? c.Lz : stratified ? strat->h[j] : N02 / (c.g * m[j] * m[j]);
    }
    for (size_t j = 0; j < Nj; ++j) {
        if (stratified) {  // F and G are the normalized modes themselves
            Fg[j] = Gg[j] = gWaveScale[j] = 1.0;
            continue;
        }
        const double sign = j % 2 == 0 ? 1.0 : -1.0;
        Fg[j] = j == 0 ? 2.0 : sign * h0[j] * m[j] * std::sqrt(2.0 * kModalGravity / (c.Lz * N02));
        Gg[j] = j == 0 ? 1.0 : sign * std::sqrt(2.0 * kModalGravity / (c.Lz * N02));
        const double Gw = c.isHydrostatic || j == 0 ? Gg[j] : sign * std::sqrt(2.0 * kModalGravity / (c.Lz * (N02 - f2)));
        gWaveScale[j] = Gg[j] / (Gg[j] / Gw);
    }

    // Per-coefficient tables -- WVKernelTypes.cpp:214-284; total-energy factors
    // from FlowComponents/*Component.m (see table() below).
    for (auto* v : {&omega, &fWaveScale, &NAp, &NA0, &A0Vort, &A0Buoy, &ApmN, &ApmDScaled, &ApmTE, &A0TE}) v->assign(M, 0.0);
    for (auto* v : {&UAp, &VAp, &WAp, &UA0, &VA0, &ApmD}) v->assign(2 * M, 0.0);
    for (size_t im = 0; im < Nkl; ++im) {
        const FourierMode& hm = modes[im];
        const double Kh = hm.Kh, Kh2 = Kh * Kh;
        for (size_t j = 0; j < Nj; ++j) {
            const size_t idx = j + Nj * im;
            const double Mj = m[j];
            const double sign = j % 2 == 0 ? 1.0 : -1.0;
            const bool wave = Kh > 0.0 && j > 0, inertial = Kh == 0.0, geostrophic = Kh > 0.0, mda = Kh == 0.0 && j > 0;
            double hpm = 1.0;
            if (j > 0) hpm = stratified ? h0[j] : c.isHydrostatic ? N02 / (c.g * Mj * Mj) : (N02 - f2) / (c.g * (Mj * Mj + Kh2));
            double Fw = Fg[j], Gw = Gg[j];
            if (!c.isHydrostatic) {
                Fw = j == 0 ? 2.0 : sign * hpm * Mj * std::sqrt(2.0 * kModalGravity / (c.Lz * (N02 - f2)));
                Gw = j == 0 ? 1.0 : sign * std::sqrt(2.0 * kModalGravity / (c.Lz * (N02 - f2)));
            }
            const double om = std::sqrt(c.g * hpm * Kh2 + f2);
            const double Fwg = Fg[j] / Fw, Gwg = Gg[j] / Gw;
            omega[idx] = om;
            fWaveScale[idx] = Fg[j] / Fwg;
            if (im == 0) inertialScale[j] = 0.5 * Fwg / Fg[j];
            const double prefactor = stratified ? 0.0 : sign * std::sqrt(c.g * c.Lz / (2.0 * (N02 - f2)));
            ApmDScaled[idx] = (Mj / 2.0) * prefactor;
            if (im == 0) apmW[j] = prefactor;
            if (wave) {
                const double fs = fWaveScale[idx], gs = gWaveScale[j];
                UAp[2 * idx] = hm.cosAlpha * fs;
                UAp[2 * idx + 1] = (-(f / om) * hm.sinAlpha) * fs;
                VAp[2 * idx] = hm.sinAlpha * fs;
                VAp[2 * idx + 1] = ((f / om) * hm.cosAlpha) * fs;
                WAp[2 * idx] = 0.0 * gs;
                WAp[2 * idx + 1] = (-Kh * hpm) * gs;
                NAp[idx] = (-Kh * hpm / om) * gs;
                const double deltaScale = h0[j] * Gwg / Fg[j];
                ApmD[2 * idx] = 0.0 * deltaScale;
                ApmD[2 * idx + 1] = (-1.0 / (2.0 * Kh * hpm)) * deltaScale;
                ApmN[idx] = (-om / (2.0 * Kh * hpm)) * Gwg / Gg[j];
                ApmTE[idx] = 2.0 * hpm;
            } else if (inertial) {
                UAp[2 * idx] = fWaveScale[idx];
                VAp[2 * idx + 1] = fWaveScale[idx];
                ApmTE[idx] = j == 0 ? c.Lz : hpm;
            }
            if (geostrophic) {
                const double lr2inv = j == 0 ? 0.0 : f2 / (c.g * h0[j]);
                const double den = Kh2 + lr2inv;
                UA0[2 * idx] = 0.0 * Fg[j];
                UA0[2 * idx + 1] = (hm.l / den) * Fg[j];
                VA0[2 * idx] = 0.0 * Fg[j];
                VA0[2 * idx + 1] = (-hm.k / den) * Fg[j];
                NA0[idx] = (j == 0 ? 0.0 : -(f / c.g) / den) * Gg[j];
                A0Vort[idx] = 1.0 / Fg[j];
                A0Buoy[idx] = j == 0 ? 0.0 : (-f / h0[j]) / Gg[j];
                A0TE[idx] = h0[j] / den;
            } else if (mda) {
                NA0[idx] = Gg[j];
                A0Buoy[idx] = 1.0 / Gg[j];
                A0Vort[idx] = (f2 / (2.0 * h0[j])) / Fg[j];
                A0TE[idx] = c.g / 2.0;
            }
        }
    }

    // Half-spectrum storage rows and Hermitian completion -- WVKernelTypes.cpp:286-321.
    const size_t rows = NxHalf * c.Ny;
    std::vector<long> represented(rows, -1);
    for (size_t iw = 0; iw < Nkl; ++iw) {
        FourierMode& fm = modes[iw];
        if (fm.kMode >= 0) {
            fm.storageRow = static_cast<size_t>(fm.kMode) + NxHalf * dftIndex(fm.lMode, c.Ny);
            fm.conjugated = false;
        } else {
            fm.storageRow = static_cast<size_t>(-fm.kMode) + NxHalf * dftIndex(-fm.lMode, c.Ny);
            fm.conjugated = true;
        }
        represented[fm.storageRow] = static_cast<long>(iw);
    }
    for (size_t iL = 0; iL < c.Ny; ++iL) {
        const int64_t lm = iL <= c.Ny / 2 ? static_cast<int64_t>(iL) : static_cast<int64_t>(iL) - static_cast<int64_t>(c.Ny);
        const size_t dst = NxHalf * dftIndex(lm, c.Ny), src = NxHalf * dftIndex(-lm, c.Ny);
        if (represented[dst] < 0 && represented[src] >= 0) {
            completionRows.push_back(dst);
            completionSources.push_back(src);
        }
        if (lm == 0 || (c.Ny % 2 == 0 && std::llabs(lm) == static_cast<long long>(c.Ny / 2)))
            if (represented[dst] >= 0) selfConjugateRows.push_back(dst);
    }

    // GPU column list: every represented row plus every completion row,
    // ordered (iy, kx) so that columns adjacent in the list are adjacent kx.
    std::vector<uint8_t> isSelf(rows, 0);
    for (auto r : selfConjugateRows) isSelf[r] = 1;
    for (size_t iw = 0; iw < Nkl; ++iw) {
        const FourierMode& fm = modes[iw];
        Column col;
        col.kx = static_cast<int32_t>(fm.storageRow % NxHalf);
        col.iy = static_cast<int32_t>(fm.storageRow / NxHalf);
        col.wv = static_cast<int32_t>(iw);
        col.flags = (fm.conjugated ? kColConj : 0u) | (isSelf[fm.storageRow] ? kColSelfConj : 0u);
        columns.push_back(col);
    }
    for (size_t i = 0; i < completionRows.size(); ++i) {
        const size_t srcWv = static_cast<size_t>(represented[completionSources[i]]);
        Column col;
        col.kx = static_cast<int32_t>(completionRows[i] % NxHalf);
        col.iy = static_cast<int32_t>(completionRows[i] / NxHalf);
        col.wv = static_cast<int32_t>(srcWv);
        col.flags = (modes[srcWv].conjugated ? 0u : kColConj) | kColCompletion | (isSelf[completionRows[i]] ? kColSelfConj : 0u);
        columns.push_back(col);
    }
    std::sort(columns.begin(), columns.end(), [](const Column& a, const Column& b) {
        if (a.iy != b.iy) return a.iy < b.iy;
        return a.kx < b.kx;
    });
    wvColumn.assign(Nkl, -1);
    nkx = 1;
    for (size_t ic = 0; ic < columns.size(); ++ic) {
        nkx = std::max(nkx, static_cast<size_t>(columns[ic].kx) + 1);
        if (!(columns[ic].flags & kColCompletion)) wvColumn[static_cast<size_t>(columns[ic].wv)] = static_cast<int32_t>(ic);
    }
    nkxP = roundUp(nkx, kxTile);
    kxv.resize(nkx);
    for (size_t k = 0; k < nkx; ++k) kxv[k] = 2.0 * kPi * (static_cast<double>(k) / c.Lx);  // as the modes' k

    // Dense vertical operators.  Synthesis applies the unnormalized DCT-I / DST-I
    // of FFTW REDFT00 / RODFT00 (WVReferenceFFTEngine.cpp:96-101) to a spectrum
    // that is nonzero only for j < Nj <= Nz-2:
    //   DCT-I: X_z = x_0 + 2 sum_{j>=1} x_j cos(pi j z/(Nz-1))
    //   DST-I (on z = 1..Nz-2): X_z = 2 sum_{j>=1} x_j sin(pi j z/(Nz-1)); X_0 = X_{Nz-1} = 0
    // Analysis applies the same transforms followed by normalizeForwardDCT / DST
    // (...Kernel.cpp:176-193): scale 1/(Nz-1), DST endpoints zeroed.
    const size_t L = Nz - 1;
    Zs = roundUp(Nz, 16);
    // contraction lengths of the MFMA synthesis and analysis: whole batches of
    // 4 k-steps, also after splitting j by parity
    Js = roundUp(Nj, 32);
    Ja = roundUp(Nj, 32);
    Za = roundUp(Nz, 16);  // analysis contraction: whole batches of 4 MFMA k-steps
    synC.assign(Zs * Js, 0.0);
    synS.assign(Zs * Js, 0.0);
    anaC.assign(Ja * Za, 0.0);
    anaS.assign(Ja * Za, 0.0);
    if (stratified) {
        synGz.assign(Zs * Js, 0.0);
        synFh.assign(Zs * Js, 0.0);
        synEz.assign(Zs * Js, 0.0);
        for (size_t iz = 0; iz < Nz; ++iz) {
            for (size_t j = 0; j < Nj; ++j) {
                const double F = strat->PFinv[iz + Nz * j] * strat->P[j];
                const double G = strat->QGinv[iz + Nz * j] * strat->Q[j];
                const double Fh = j == 0 ? 0.0 : F / strat->h[j];
                synC[iz * Js + j] = F;
                synS[iz * Js + j] = G;
                synGz[iz * Js + j] = -(strat->N2[iz] / c.g) * G;
                synFh[iz * Js + j] = Fh;
                synEz[iz * Js + j] = Fh + strat->dLnN2[iz] * G;
                anaC[j * Za + iz] = strat->PF[j + Nj * iz] / strat->P[j];
                anaS[j * Za + iz] = strat->QG[j + Nj * iz] / strat->Q[j];
            }
        }
    }
    for (size_t iz = 0; iz < Nz && !stratified; ++iz) {
        for (size_t j = 0; j < Nj; ++j) {
            synC[iz * Js + j] = (j == 0 ? 1.0 : 2.0) * cosPiRatio(j * iz, L);
            const bool interior = iz >= 1 && iz + 1 <= Nz - 1 && j >= 1;
            synS[iz * Js + j] = interior && iz + 1 < Nz ? 2.0 * sinPiRatio(j * iz, L) : 0.0;
            const double wEnd = (iz == 0 || iz + 1 == Nz) ? 1.0 : 2.0;
            anaC[j * Za + iz] = (wEnd * cosPiRatio(j * iz, L)) / static_cast<double>(L);
            anaS[j * Za + iz] = (j >= 1 && iz >= 1 && iz + 1 < Nz) ? (2.0 * sinPiRatio(j * iz, L)) / static_cast<double>(L) : 0.0;
        }
    }
    // Power-of-two n: the Stockham pass twiddles of the LDS FFT, pass (Ns, R)
    // entry (r, k) at r*Ns + k - 1 holding exp(-2 pi i r k / (Ns R)), so that
    // consecutive k (consecutive threads) read consecutive entries; the pass
    // order is the FFT's (radix 2^(log2 n mod lr) first, then radix 2^lr).
    // Other n: the plain table exp(-2 pi i q / n) of the direct DFT.
    auto twiddles = [](size_t n, std::vector<double>& out, int lr) {
        out.assign(2 * n, 0.0);
        auto put = [&](size_t idx, size_t q) {
            out[2 * idx] = cosPiRatio(2 * q, n);
            out[2 * idx + 1] = -sinPiRatio(2 * q, n);
        };
        if (n & (n - 1)) {
            for (size_t q = 0; q < n; ++q) put(q, q);
            return;
        }
        int l2 = 0;
        while ((size_t{1} << l2) < n) ++l2;
        size_t Ns = 1;
        auto pass = [&](size_t R) {
            for (size_t r = 1; r < R; ++r)
                for (size_t k = 0; k < Ns; ++k)
                    if (r * Ns + k >= 1) put(r * Ns + k - 1, (r * k * (n / (Ns * R))) % n);
            Ns *= R;
        };
        if (l2 % lr) pass(size_t{1} << (l2 % lr));
        while (Ns < n) pass(size_t{1} << lr);
    };
    twiddles(c.Nx, twX, 4);
    twiddles(c.Ny, twY, 4);
    return {};
}

long Descriptor::table(const std::string& n, double* out) const {
    auto put = [&](const std::vector<double>& v) {
        if (out) std::copy(v.begin(), v.end(), out);
        return static_cast<long>(v.size());
    };
    if (n == "z") return put(z);
    if (n == "h0") return put(h0);
    if (n == "verticalWavenumber") return put(m);
    if (n == "Fg") return put(Fg);
    if (n == "Gg") return put(Gg);
    if (n == "inertialScale") return put(inertialScale);
    if (n == "gWaveScale") return put(gWaveScale);
    if (n == "apmWProjectionPrefactor") return put(apmW);
    if (n == "omega") return put(omega);
    if (n == "fWaveScale") return put(fWaveScale);
    if (n == "NApField") return put(NAp);
    if (n == "NA0Field") return put(NA0);
    if (n == "A0FromVorticity") return put(A0Vort);
    if (n == "A0FromBuoyancy") return put(A0Buoy);
    if (n == "ApmNProjection") return put(ApmN);
    if (n == "ApmDScaled") return put(ApmDScaled);
    // Total-energy factors: waves 2*h_pm (FlowComponents/WVInternalGravityWaveComponent.m:66),
    // inertial h_pm or Lz at j = 0 (WVInertialOscillationComponent.m:86-87), geostrophic
    // hke + pe = h0/(K^2 + f^2/(g h0)) (WVGeostrophicComponent.m:160-174), MDA g/2
    // (WVMeanDensityAnomalyComponent.m:97-104).
    if (n == "ApmTE") return put(ApmTE);
    if (n == "A0TE") return put(A0TE);
    if (n == "UApField") return put(UAp);
    if (n == "VApField") return put(VAp);
    if (n == "WApField") return put(WAp);
    if (n == "UA0Field") return put(UA0);
    if (n == "VA0Field") return put(VA0);
    if (n == "ApmDProjection") return put(ApmD);
    return -1;
}

}  // namespace wvamd
