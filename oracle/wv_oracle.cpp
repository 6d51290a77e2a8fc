// CPU oracle for the wave-vortex nonlinear-flux / RK4 path.
//
// TEST INFRASTRUCTURE ONLY: a plain FP64 restatement of the reference
// algorithm, used by tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline as the checker.  The product never links or calls it.
//
// It follows, stage by stage and in the same arithmetic order:
//   descriptor ............ CompiledKernel/src/WVKernelTypes.cpp:119-330
//   coefficient formulas .. CompiledKernel/src/WVCoefficientFormulas.hpp:10-116
//   transforms / flux ..... CompiledKernel/src/WVTransformConstantStratificationKernel.cpp
//                           :176-209 (normalization), :218-291 (assembly),
//                           :299-431 (projection), :433-438 (Hermitian completion),
//                           :583-684 (forward), :686-805 (inverse), :819-863 (F/G),
//                           :907-990 (derivatives + projection), :1173-1244 (flux)
//   FFT semantics ......... CompiledKernel/adapters/reference/WVReferenceFFTEngine.cpp:47-106
//   RK4 ................... PortableRuntime/src/WVRungeKutta.cpp:15-19,140-176,335-462
//   energy ................ @WVTransformConstantStratification/WVTransformConstantStratification.m:364-376
//                           with the factors of FlowComponents/*Component.m
// The FFTs are computed with this file's own radix-2 / direct DFT code
// (FFTW, the reference's production provider, is not available here).
#include "wv_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kModalGravity = 9.81;  // WVKernelTypes.cpp:13-16 (fixed 9.81 in F/G norms)

struct cplx {
    double re = 0.0, im = 0.0;
};
inline cplx add(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
inline cplx sub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
inline cplx mul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cplx mul(cplx a, double b) { return {a.re * b, a.im * b}; }
inline cplx cj(cplx a) { return {a.re, -a.im}; }
inline cplx phase(double angle) { return {std::cos(angle), std::sin(angle)}; }

// ---------------------------------------------------------------- FFT -----
// Complex DFT of length n, X_k = sum_j x_j exp(sign*2*pi*i*j*k/n).
struct DFT {
    size_t n = 0;
    bool pow2 = false;
    std::vector<cplx> w;        // w[m] = exp(-2*pi*i*m/n)
    std::vector<size_t> rev;
    void init(size_t len) {
        n = len;
        pow2 = len >= 1 && (len & (len - 1)) == 0;
        w.resize(len);
        for (size_t m = 0; m < len; ++m) {
            // exact angle reduction keeps the table symmetric
            const double a = -2.0 * kPi * static_cast<double>(m) / static_cast<double>(len);
            w[m] = {std::cos(a), std::sin(a)};
        }
        if (pow2) {
            rev.resize(len);
            size_t bits = 0;
            while ((size_t(1) << bits) < len) ++bits;
            for (size_t i = 0; i < len; ++i) {
                size_t r = 0;
                for (size_t b = 0; b < bits; ++b) if (i & (size_t(1) << b)) r |= size_t(1) << (bits - 1 - b);
                rev[i] = r;
            }
        }
    }
    // in place; scratch must hold n entries
    void run(cplx* x, int sign, cplx* scratch) const {
        if (n <= 1) return;
        if (pow2) {
            for (size_t i = 0; i < n; ++i) if (rev[i] > i) std::swap(x[i], x[rev[i]]);
            for (size_t len = 2; len <= n; len <<= 1) {
                const size_t half = len / 2, stride = n / len;
                for (size_t s = 0; s < n; s += len) {
                    for (size_t q = 0; q < half; ++q) {
                        cplx tw = w[q * stride];
                        if (sign > 0) tw = cj(tw);
                        const cplx a = x[s + q];
                        const cplx b = mul(x[s + q + half], tw);
                        x[s + q] = add(a, b);
                        x[s + q + half] = sub(a, b);
                    }
                }
            }
        } else {
            for (size_t k = 0; k < n; ++k) {
                cplx acc{};
                for (size_t j = 0; j < n; ++j) {
                    cplx tw = w[(j * k) % n];
                    if (sign > 0) tw = cj(tw);
                    acc = add(acc, mul(x[j], tw));
                }
                scratch[k] = acc;
            }
            std::copy(scratch, scratch + n, x);
        }
    }
};

int gThreads = 0;

// ------------------------------------------------------------ Oracle ------
struct Mode {
    int64_t kMode = 0, lMode = 0;
    double k = 0, l = 0, Kh = 0, cosA = 0, sinA = 0;
    size_t row = 0;
    bool conj = false;
};

struct Oracle {
    wvo_config c{};
    size_t NxHalf = 0, rows = 0, Nkl = 0, M = 0, R = 0;
    std::vector<Mode> modes;
    std::vector<size_t> completionRows, completionSources, selfConjugateRows;
    double f = 0;
    std::vector<double> z, h0, m, Fg, Gg, inertialScale, gWaveScale, apmW;
    std::vector<double> omega, fWaveScale, NAp, NA0, A0Vort, A0Buoy, ApmN, ApmDScaled, ApmTE, A0TE;
    std::vector<cplx> UAp, VAp, WAp, UA0, VA0, ApmD;
    DFT dftX, dftY, dftV;  // dftV has length 2*(Nz-1)

    using Half = std::vector<cplx>;  // one channel, index z + Nz*row

    // --- descriptor: WVKernelTypes.cpp:119-330 -----------------------------
    static std::vector<int64_t> dftModes(size_t n) {
        std::vector<int64_t> v;
        for (size_t q = 0; q < (n + 1) / 2; ++q) v.push_back(int64_t(q));
        for (size_t q = n / 2; q > 0; --q) v.push_back(-int64_t(q));
        return v;
    }
    static size_t wrap(int64_t mode, size_t n) { return size_t(mode >= 0 ? mode : int64_t(n) + mode); }
    static bool selfConj(int64_t mode, size_t n) { return mode == 0 || (n % 2 == 0 && mode == -int64_t(n / 2)); }
    static double radial(double a, double b) {
        volatile double a2 = a * a;  // two separate roundings, as in :22-29
        volatile double b2 = b * b;
        return std::sqrt(a2 + b2);
    }

    std::string build() {
        if (c.Nx < 2 || c.Ny < 2 || c.Nz < 3 || c.Nj < 1 || c.Nj > c.Nz - 1) return "Nx and Ny must be at least 2, Nz at least 3, and Nj must lie in [1,Nz-1].";
        auto pos = [](double v) { return std::isfinite(v) && v > 0.0; };
        if (!pos(c.Lx) || !pos(c.Ly) || !pos(c.Lz) || !pos(c.N0) || !pos(c.rho0) || !pos(c.g) || !pos(c.planetaryRadius) || !pos(c.rotationRate) || !std::isfinite(c.latitude) || std::abs(c.latitude) > 90.0)
            return "Physical lengths and constants must be finite and positive, and latitude must lie in [-90,90].";
        const auto kModes = dftModes(c.Nx), lModes = dftModes(c.Ny);
        const double maxAbsK = 2.0 * kPi * (double(c.Nx / 2) / c.Lx);
        for (auto lm : lModes) for (auto km : kModes) {
            const bool primary = lm > 0 || (selfConj(lm, c.Ny) && (km > 0 || selfConj(km, c.Nx)));
            const bool nyq = (c.Nx % 2 == 0 && km == -int64_t(c.Nx / 2)) || (c.Ny % 2 == 0 && lm == -int64_t(c.Ny / 2));
            if (!primary || nyq) continue;
            const double k = 2.0 * kPi * (double(km) / c.Lx);
            const double l = 2.0 * kPi * (double(lm) / c.Ly);
            if (c.shouldAntialias && radial(k, l) > 2.0 * maxAbsK / 3.0) continue;
            Mode md;
            md.kMode = km; md.lMode = lm; md.k = k; md.l = l;
            md.Kh = radial(k, l);
            md.cosA = md.Kh == 0.0 ? 0.0 : k / md.Kh;
            md.sinA = md.Kh == 0.0 ? 0.0 : l / md.Kh;
            modes.push_back(md);
        }
        std::stable_sort(modes.begin(), modes.end(), [](const Mode& a, const Mode& b) {
            if (a.Kh != b.Kh) return a.Kh < b.Kh;
            if (a.kMode != b.kMode) return a.kMode < b.kMode;
            return a.lMode < b.lMode;
        });
        Nkl = modes.size();
        M = Nkl * c.Nj;
        NxHalf = c.Nx / 2 + 1;
        rows = NxHalf * c.Ny;
        R = c.Nx * c.Ny * c.Nz;

        f = 2.0 * c.rotationRate * std::sin(c.latitude * kPi / 180.0);
        const size_t Nz = c.Nz, Nj = c.Nj;
        z.resize(Nz); h0.resize(Nj); m.resize(Nj); Fg.resize(Nj); Gg.resize(Nj);
        inertialScale.resize(Nj); gWaveScale.resize(Nj); apmW.resize(Nj);
        const double dz = c.Lz / double(Nz - 1);
        for (size_t i = 0; i < Nz; ++i) z[i] = dz * double(i) - c.Lz;
        for (size_t j = 0; j < Nj; ++j) {
            m[j] = double(j) * kPi / c.Lz;
            h0[j] = j == 0 ? c.Lz : c.N0 * c.N0 / (c.g * m[j] * m[j]);
        }
        const double N02 = c.N0 * c.N0, f2 = f * f;
        if (!c.isHydrostatic && N02 <= f2) return "Nonhydrostatic constant stratification requires N0 squared greater than f squared.";
        for (size_t j = 0; j < Nj; ++j) {  // :203-212
            const double s = j % 2 == 0 ? 1.0 : -1.0;
            const double fg = j == 0 ? 2.0 : s * h0[j] * m[j] * std::sqrt(2.0 * kModalGravity / (c.Lz * N02));
            const double gg = j == 0 ? 1.0 : s * std::sqrt(2.0 * kModalGravity / (c.Lz * N02));
            const double gw = c.isHydrostatic || j == 0 ? gg : s * std::sqrt(2.0 * kModalGravity / (c.Lz * (N02 - f2)));
            Fg[j] = fg; Gg[j] = gg; gWaveScale[j] = gg / (gg / gw);
        }
        omega.assign(M, 0); fWaveScale.assign(M, 0); NAp.assign(M, 0); NA0.assign(M, 0);
        A0Vort.assign(M, 0); A0Buoy.assign(M, 0); ApmN.assign(M, 0); ApmDScaled.assign(M, 0);
        UAp.assign(M, {}); VAp.assign(M, {}); WAp.assign(M, {}); UA0.assign(M, {}); VA0.assign(M, {}); ApmD.assign(M, {});
        ApmTE.assign(M, 0); A0TE.assign(M, 0);
        for (size_t im = 0; im < Nkl; ++im) {  // :214-284
            const Mode& hm = modes[im];
            const double Kh = hm.Kh, Kh2 = Kh * Kh;
            for (size_t j = 0; j < Nj; ++j) {
                const size_t idx = j + Nj * im;
                const double Mj = m[j];
                const double s = j % 2 == 0 ? 1.0 : -1.0;
                const bool isWave = Kh > 0.0 && j > 0, isInertial = Kh == 0.0, isGeo = Kh > 0.0, isMDA = Kh == 0.0 && j > 0;
                double hpm = 1.0;
                if (j > 0) hpm = c.isHydrostatic ? N02 / (c.g * Mj * Mj) : (N02 - f2) / (c.g * (Mj * Mj + Kh2));
                double Fw = Fg[j], Gw = Gg[j];
                if (!c.isHydrostatic) {
                    Fw = j == 0 ? 2.0 : s * hpm * Mj * std::sqrt(2.0 * kModalGravity / (c.Lz * (N02 - f2)));
                    Gw = j == 0 ? 1.0 : s * std::sqrt(2.0 * kModalGravity / (c.Lz * (N02 - f2)));
                }
                const double om = std::sqrt(c.g * hpm * Kh2 + f2);
                const double Fwg = Fg[j] / Fw, Gwg = Gg[j] / Gw;
                omega[idx] = om;
                fWaveScale[idx] = Fg[j] / Fwg;
                if (im == 0) inertialScale[j] = 0.5 * Fwg / Fg[j];
                const double pre = s * std::sqrt(c.g * c.Lz / (2.0 * (N02 - f2)));
                ApmDScaled[idx] = (Mj / 2.0) * pre;
                if (im == 0) apmW[j] = pre;
                if (isWave) {
                    const cplx uap{hm.cosA, -(f / om) * hm.sinA}, vap{hm.sinA, (f / om) * hm.cosA}, wap{0.0, -Kh * hpm};
                    const double nap = -Kh * hpm / om;
                    const cplx apmd{0.0, -1.0 / (2.0 * Kh * hpm)};
                    const double apmn = -om / (2.0 * Kh * hpm);
                    UAp[idx] = {uap.re * fWaveScale[idx], uap.im * fWaveScale[idx]};
                    VAp[idx] = {vap.re * fWaveScale[idx], vap.im * fWaveScale[idx]};
                    WAp[idx] = {wap.re * gWaveScale[j], wap.im * gWaveScale[j]};
                    NAp[idx] = nap * gWaveScale[j];
                    const double ds = h0[j] * Gwg / Fg[j];
                    ApmD[idx] = {apmd.re * ds, apmd.im * ds};
                    ApmN[idx] = apmn * Gwg / Gg[j];
                } else if (isInertial) {
                    UAp[idx] = {fWaveScale[idx], 0.0};
                    VAp[idx] = {0.0, fWaveScale[idx]};
                }
                if (isGeo) {
                    const double lr2inv = j == 0 ? 0.0 : f2 / (c.g * h0[j]);
                    const double den = Kh2 + lr2inv;
                    const cplx ua0{0.0, hm.l / den}, va0{0.0, -hm.k / den};
                    const double na0 = j == 0 ? 0.0 : -(f / c.g) / den;
                    UA0[idx] = {ua0.re * Fg[j], ua0.im * Fg[j]};
                    VA0[idx] = {va0.re * Fg[j], va0.im * Fg[j]};
                    NA0[idx] = na0 * Gg[j];
                    A0Vort[idx] = 1.0 / Fg[j];
                    A0Buoy[idx] = j == 0 ? 0.0 : (-f / h0[j]) / Gg[j];
                } else if (isMDA) {
                    NA0[idx] = Gg[j];
                    A0Buoy[idx] = 1.0 / Gg[j];
                    A0Vort[idx] = (f2 / (2.0 * h0[j])) / Fg[j];
                }
                // total-energy factors (FlowComponents/WVInternalGravityWaveComponent.m:66,
                // WVInertialOscillationComponent.m:86-87, WVGeostrophicComponent.m:160-174,
                // WVMeanDensityAnomalyComponent.m:97-104)
                if (isWave) ApmTE[idx] = 2.0 * hpm;
                else if (isInertial) ApmTE[idx] = j == 0 ? c.Lz : hpm;
                if (isGeo) {
                    const double lr2inv = j == 0 ? 0.0 : f2 / (c.g * h0[j]);
                    A0TE[idx] = h0[j] / (Kh2 + lr2inv);
                } else if (isMDA) {
                    A0TE[idx] = c.g / 2.0;
                }
            }
        }
        // half-spectrum maps :286-321
        std::vector<long> represented(rows, -1);
        for (size_t iw = 0; iw < Nkl; ++iw) {
            Mode& md = modes[iw];
            if (md.kMode >= 0) {
                md.row = size_t(md.kMode) + NxHalf * wrap(md.lMode, c.Ny);
                md.conj = false;
            } else {
                md.row = size_t(-md.kMode) + NxHalf * wrap(-md.lMode, c.Ny);
                md.conj = true;
            }
            represented[md.row] = long(iw);
        }
        for (size_t iL = 0; iL < c.Ny; ++iL) {
            const int64_t lm = iL <= c.Ny / 2 ? int64_t(iL) : int64_t(iL) - int64_t(c.Ny);
            const size_t dst = NxHalf * wrap(lm, c.Ny), src = NxHalf * wrap(-lm, c.Ny);
            if (represented[dst] < 0 && represented[src] >= 0) {
                completionRows.push_back(dst);
                completionSources.push_back(src);
            }
            if (lm == 0 || (c.Ny % 2 == 0 && std::abs(lm) == int64_t(c.Ny / 2)))
                if (represented[dst] >= 0) selfConjugateRows.push_back(dst);
        }
        dftX.init(c.Nx);
        dftY.init(c.Ny);
        dftV.init(2 * (c.Nz - 1));
        return {};
    }

    // --- primitive transforms ---------------------------------------------
    // Unnormalized 2-D c2r of one channel (Hermitian-implied), as
    // WVReferenceFFTEngine.cpp:65-86; output plane z is out[x + Nx*(y + Ny*z)].
    void c2r(const Half& h, double* out) const {
        const size_t Nx = c.Nx, Ny = c.Ny, Nz = c.Nz;
#pragma omp parallel for schedule(static)
        for (long zz = 0; zz < long(Nz); ++zz) {
            std::vector<cplx> plane(NxHalf * Ny), col(Ny), row(Nx), scratch(std::max(Nx, Ny));
            for (size_t kx = 0; kx < NxHalf; ++kx) {
                for (size_t iy = 0; iy < Ny; ++iy) col[iy] = h[size_t(zz) + Nz * (kx + NxHalf * iy)];
                dftY.run(col.data(), +1, scratch.data());
                for (size_t y = 0; y < Ny; ++y) plane[kx + NxHalf * y] = col[y];
            }
            for (size_t y = 0; y < Ny; ++y) {
                for (size_t kx = 0; kx < Nx; ++kx) {
                    if (kx < NxHalf) row[kx] = plane[kx + NxHalf * y];
                    else row[kx] = cj(plane[(Nx - kx) + NxHalf * y]);
                }
                dftX.run(row.data(), +1, scratch.data());
                for (size_t x = 0; x < Nx; ++x) out[x + Nx * (y + Ny * size_t(zz))] = row[x].re;
            }
        }
    }
    // Unnormalized 2-D r2c (e^{-i}), as WVReferenceFFTEngine.cpp:47-63.
    void r2c(const double* in, Half& h) const {
        const size_t Nx = c.Nx, Ny = c.Ny, Nz = c.Nz;
#pragma omp parallel for schedule(static)
        for (long zz = 0; zz < long(Nz); ++zz) {
            std::vector<cplx> plane(NxHalf * Ny), col(Ny), row(Nx), scratch(std::max(Nx, Ny));
            for (size_t y = 0; y < Ny; ++y) {
                for (size_t x = 0; x < Nx; ++x) row[x] = {in[x + Nx * (y + Ny * size_t(zz))], 0.0};
                dftX.run(row.data(), -1, scratch.data());
                for (size_t kx = 0; kx < NxHalf; ++kx) plane[kx + NxHalf * y] = row[kx];
            }
            for (size_t kx = 0; kx < NxHalf; ++kx) {
                for (size_t y = 0; y < Ny; ++y) col[y] = plane[kx + NxHalf * y];
                dftY.run(col.data(), -1, scratch.data());
                for (size_t iy = 0; iy < Ny; ++iy) h[size_t(zz) + Nz * (kx + NxHalf * iy)] = col[iy];
            }
        }
    }
    // In-place unnormalized DCT-I along z on every row (FFTW REDFT00 /
    // WVReferenceFFTEngine.cpp:98-101) via an even extension of length 2(Nz-1).
    void dct(Half& h) const {
        const size_t Nz = c.Nz, L = 2 * (Nz - 1);
#pragma omp parallel for schedule(static)
        for (long r = 0; r < long(rows); ++r) {
            std::vector<cplx> y(L), scratch(L);
            cplx* x = &h[Nz * size_t(r)];
            for (size_t n = 0; n < Nz; ++n) y[n] = x[n];
            for (size_t n = 1; n + 1 < Nz; ++n) y[L - n] = x[n];
            dftV.run(y.data(), -1, scratch.data());
            for (size_t q = 0; q < Nz; ++q) x[q] = y[q];
        }
    }
    // In-place unnormalized DST-I of length Nz-2 on z = 1..Nz-2 (FFTW RODFT00 /
    // WVReferenceFFTEngine.cpp:96-97); entries z = 0 and Nz-1 are left as they are.
    void dst(Half& h) const {
        const size_t Nz = c.Nz, L = 2 * (Nz - 1);
#pragma omp parallel for schedule(static)
        for (long r = 0; r < long(rows); ++r) {
            std::vector<cplx> y(L), scratch(L);
            cplx* x = &h[Nz * size_t(r)];
            const size_t N = Nz - 2;
            y[0] = {};
            for (size_t n = 0; n < N; ++n) { y[n + 1] = x[n + 1]; y[L - 1 - n] = mul(x[n + 1], -1.0); }
            y[N + 1] = {};
            dftV.run(y.data(), -1, scratch.data());
            for (size_t q = 0; q < N; ++q) x[q + 1] = mul(y[q + 1], cplx{0.0, 1.0});
        }
    }
    void normForwardDCT(Half& h) const {  // :176-183
        const double s = 1.0 / double(c.Nz - 1);
        for (size_t r = 0; r < rows; ++r) {
            cplx* x = &h[c.Nz * r];
            for (size_t q = 0; q < c.Nz; ++q) x[q] = mul(x[q], s);
            x[c.Nz - 1] = mul(x[c.Nz - 1], 0.5);
        }
    }
    void normForwardDST(Half& h) const {  // :185-193
        const double s = 1.0 / double(c.Nz - 1);
        for (size_t r = 0; r < rows; ++r) {
            cplx* x = &h[c.Nz * r];
            x[0] = {};
            for (size_t q = 1; q + 1 < c.Nz; ++q) x[q] = mul(x[q], s);
            x[c.Nz - 1] = {};
        }
    }
    void normInverseDCT(Half& h) const {  // :195-200
        for (size_t r = 0; r < rows; ++r) for (size_t q = 0; q < c.Nz; ++q) h[q + c.Nz * r] = mul(h[q + c.Nz * r], 0.5);
    }
    void normInverseDST(Half& h) const {  // :202-209
        for (size_t r = 0; r < rows; ++r) {
            cplx* x = &h[c.Nz * r];
            x[0] = {};
            for (size_t q = 1; q + 1 < c.Nz; ++q) x[q] = mul(x[q], 0.5);
            x[c.Nz - 1] = {};
        }
    }
    void completeHermitian(Half& h) const {  // :433-438
        for (size_t i = 0; i < completionRows.size(); ++i)
            for (size_t q = 0; q < c.Nz; ++q) h[q + c.Nz * completionRows[i]] = cj(h[q + c.Nz * completionSources[i]]);
        for (auto r : selfConjugateRows) for (size_t q = 0; q < c.Nz; ++q) h[q + c.Nz * r].im = 0.0;
    }

    // --- coefficient formulas (WVCoefficientFormulas.hpp) ------------------
    struct Evolved { cplx Ap, Am, A0; };
    Evolved evolved(const cplx* Ap, const cplx* Am, const cplx* A0, size_t idx, cplx p) const {
        return {mul(Ap[idx], p), mul(Am[idx], cj(p)), A0[idx]};
    }
    cplx fieldValue(int T, size_t idx, const Evolved& e) const {  // :27-33
        if (T == 0) return add(add(mul(UAp[idx], e.Ap), mul(cj(UAp[idx]), e.Am)), mul(UA0[idx], e.A0));
        if (T == 1) return add(add(mul(VAp[idx], e.Ap), mul(cj(VAp[idx]), e.Am)), mul(VA0[idx], e.A0));
        if (T == 2) return add(mul(WAp[idx], e.Ap), mul(WAp[idx], e.Am));
        return add(add(mul(e.Ap, NAp[idx]), mul(e.Am, -NAp[idx])), mul(e.A0, NA0[idx]));
    }

    // --- synthesis ----------------------------------------------------------
    // Velocity spectra (u,v,w)*1/2 with w endpoint zeroing, :269-291, then
    // completion, DCT(u,v), DST(w), c2r: :778-805.
    void synthesizeUVW(const std::vector<Evolved>& ev, double* out) const {
        std::vector<Half> h(3, Half(c.Nz * rows));
        const size_t Nj = c.Nj, Nz = c.Nz;
        for (size_t im = 0; im < Nkl; ++im) {
            const Mode& md = modes[im];
            for (size_t j = 0; j < Nj; ++j) {
                const size_t idx = j + Nj * im;
                cplx v[3] = {fieldValue(0, idx, ev[idx]), fieldValue(1, idx, ev[idx]), fieldValue(2, idx, ev[idx])};
                v[0] = mul(v[0], 0.5);
                v[1] = mul(v[1], 0.5);
                v[2] = (j == 0 || j + 1 == Nz) ? cplx{} : mul(v[2], 0.5);
                for (int t = 0; t < 3; ++t) h[t][j + Nz * md.row] = md.conj ? cj(v[t]) : v[t];
            }
        }
        for (auto& x : h) completeHermitian(x);
        dct(h[0]); dct(h[1]); dst(h[2]);
        for (int t = 0; t < 3; ++t) c2r(h[t], out + size_t(t) * R);
    }
    // Full (u,v,w,eta) synthesis :218-244, :745-776.
    void synthesizeUVWEta(const std::vector<Evolved>& ev, double* out) const {
        std::vector<Half> h(4, Half(c.Nz * rows));
        const size_t Nj = c.Nj, Nz = c.Nz;
        for (size_t im = 0; im < Nkl; ++im) {
            const Mode& md = modes[im];
            for (size_t j = 0; j < Nj; ++j) {
                const size_t idx = j + Nj * im;
                for (int t = 0; t < 4; ++t) {
                    cplx v = fieldValue(t, idx, ev[idx]);
                    v = (t >= 2 && (j == 0 || j + 1 == Nz)) ? cplx{} : mul(v, 0.5);
                    h[t][j + Nz * md.row] = md.conj ? cj(v) : v;
                }
            }
        }
        for (auto& x : h) completeHermitian(x);
        dct(h[0]); dct(h[1]); dst(h[2]); dst(h[3]);
        for (int t = 0; t < 4; ++t) c2r(h[t], out + size_t(t) * R);
    }
    // Derivative spectra of field T: :246-267 + normalizedDerivativeSpectrum
    // (WVCoefficientFormulas.hpp:42-58), then :932-943.
    void synthesizeDerivatives(int T, const std::vector<Evolved>& ev, double* out) const {
        std::vector<Half> h(3, Half(c.Nz * rows));
        const size_t Nj = c.Nj, Nz = c.Nz;
        const bool cosine = T < 2;
        for (size_t im = 0; im < Nkl; ++im) {
            const Mode& md = modes[im];
            for (size_t j = 0; j < Nj; ++j) {
                const size_t idx = j + Nj * im;
                cplx v = fieldValue(T, idx, ev[idx]);
                const bool endpoint = j == 0 || j + 1 == Nz;
                if (md.conj) v = cj(v);
                cplx dx = mul(v, cplx{0.0, md.conj ? -md.k : md.k});
                cplx dy = mul(v, cplx{0.0, md.conj ? -md.l : md.l});
                cplx dz = mul(v, cosine ? -m[j] : m[j]);
                if (cosine) {
                    dx = mul(dx, 0.5); dy = mul(dy, 0.5); dz = endpoint ? cplx{} : mul(dz, 0.5);
                } else {
                    dx = endpoint ? cplx{} : mul(dx, 0.5); dy = endpoint ? cplx{} : mul(dy, 0.5); dz = mul(dz, 0.5);
                }
                h[0][j + Nz * md.row] = dx;
                h[1][j + Nz * md.row] = dy;
                h[2][j + Nz * md.row] = dz;
            }
        }
        for (auto& x : h) completeHermitian(x);
        if (cosine) { dct(h[0]); dct(h[1]); dst(h[2]); }
        else { dst(h[0]); dst(h[1]); dct(h[2]); }
        for (int t = 0; t < 3; ++t) c2r(h[t], out + size_t(t) * R);
    }

    // --- analysis ------------------------------------------------------------
    // projectFluxTargetMode :396-431 for every retained coefficient.
    void projectTarget(int T, const Half& h, const std::vector<cplx>& ph, cplx* Fp, cplx* Fm, cplx* F0) const {
        const double scale = 1.0 / double(c.Nx * c.Ny);
        const size_t Nj = c.Nj, Nz = c.Nz;
        const bool hydro = c.isHydrostatic != 0;
        for (size_t im = 0; im < Nkl; ++im) {
            const Mode& md = modes[im];
            for (size_t j = 0; j < Nj; ++j) {
                const size_t idx = j + Nj * im;
                cplx value = h[j + Nz * md.row];
                if (md.conj) value = cj(value);
                value = mul(value, scale);
                cplx a0{};
                if (T == 0) a0 = mul(mul(value, cplx{0.0, -md.l}), A0Vort[idx]);
                if (T == 1) a0 = mul(mul(value, cplx{0.0, md.k}), A0Vort[idx]);
                if (T == 3) a0 = mul(value, A0Buoy[idx]);
                const cplx buoyValue = sub(T == 3 ? value : cplx{}, mul(a0, NA0[idx]));
                const cplx buoy = mul(buoyValue, ApmN[idx]);
                cplx wave{};
                if (hydro) {
                    if (T == 0) wave = mul(mul(value, cplx{0.0, md.k}), ApmD[idx]);
                    if (T == 1) wave = mul(mul(value, cplx{0.0, md.l}), ApmD[idx]);
                } else {
                    if (T == 0) wave = mul(value, md.cosA * ApmDScaled[idx]);
                    if (T == 1) wave = mul(value, md.sinA * ApmDScaled[idx]);
                    if (T == 2) wave = mul(value, cplx{0.0, (md.Kh / 2.0) * apmW[j]});
                }
                cplx cAp = add(wave, buoy), cAm = sub(wave, buoy);
                if (im == 0) {  // inertialCoefficientPair<Target> :115-121
                    cplx a{};
                    if (T == 0) a = mul(value, inertialScale[j]);
                    if (T == 1) a = mul(mul(value, cplx{0.0, -1.0}), inertialScale[j]);
                    cAp = a; cAm = cj(a);
                }
                Fp[idx] = add(Fp[idx], mul(cAp, cj(ph[idx])));
                Fm[idx] = add(Fm[idx], mul(cAm, ph[idx]));
                F0[idx] = add(F0[idx], a0);
            }
        }
    }

    // nonlinearFluxImpl :1173-1244; uvwGiven (nonlinearFluxUsingAdvectionFields,
    // :1160-1171): the advecting fields of the products instead of the
    // synthesized ones.
    void flux(double t, double t0, const cplx* Ap, const cplx* Am, const cplx* A0, cplx* Fp, cplx* Fm, cplx* F0,
              const double* uvwGiven = nullptr) const {
        const double elapsed = t - t0;
        std::vector<cplx> ph(M);
        std::vector<Evolved> ev(M);
        for (size_t i = 0; i < M; ++i) {
            ph[i] = phase(omega[i] * elapsed);
            Fp[i] = {}; Fm[i] = {}; F0[i] = {};
            ev[i] = evolved(Ap, Am, A0, i, ph[i]);
        }
        std::vector<double> uvw(3 * R), d(3 * R);
        if (uvwGiven) std::copy(uvwGiven, uvwGiven + 3 * R, uvw.begin());
        else synthesizeUVW(ev, uvw.data());
        const int hydroTargets[] = {0, 1, 3}, nonhydroTargets[] = {0, 1, 2, 3};
        const int nt = c.isHydrostatic ? 3 : 4;
        const int* targets = c.isHydrostatic ? hydroTargets : nonhydroTargets;
        for (int it = 0; it < nt; ++it) {
            const int T = targets[it];
            synthesizeDerivatives(T, ev, d.data());
            const double *U = uvw.data(), *V = U + R, *W = V + R;
            double* dx = d.data();
            const double *dy = dx + R, *dz = dy + R;
            for (size_t i = 0; i < R; ++i) dx[i] = -(U[i] * dx[i] + V[i] * dy[i] + W[i] * dz[i]);
            Half h(c.Nz * rows);
            r2c(dx, h);
            if (T < 2) { dct(h); normForwardDCT(h); }
            else { dst(h); normForwardDST(h); }
            projectTarget(T, h, ph, Fp, Fm, F0);
        }
    }

    // Forward transforms :583-684 with the project*Mode formulas :299-370.
    void forward(const double* fields, double t, double t0, bool withW, cplx* Ap, cplx* Am, cplx* A0) const {
        const size_t nch = withW ? 4 : 3;
        std::vector<Half> h(nch, Half(c.Nz * rows));
        for (size_t ch = 0; ch < nch; ++ch) r2c(fields + ch * R, h[ch]);
        dct(h[0]); dct(h[1]); normForwardDCT(h[0]); normForwardDCT(h[1]);
        for (size_t ch = 2; ch < nch; ++ch) { dst(h[ch]); normForwardDST(h[ch]); }
        const double scale = 1.0 / double(c.Nx * c.Ny);
        const double elapsed = t - t0;
        const size_t Nj = c.Nj, Nz = c.Nz;
        for (size_t im = 0; im < Nkl; ++im) {
            const Mode& md = modes[im];
            for (size_t j = 0; j < Nj; ++j) {
                const size_t idx = j + Nj * im;
                auto stored = [&](size_t ch) { cplx v = h[ch][j + Nz * md.row]; if (md.conj) v = cj(v); return mul(v, scale); };
                const cplx U = stored(0), V = stored(1);
                const cplx W = withW ? stored(2) : cplx{};
                const cplx N = stored(withW ? 3 : 2);
                // geostrophicCoefficient :43-45
                const cplx vort = sub(mul(V, cplx{0.0, md.k}), mul(U, cplx{0.0, md.l}));
                const cplx a0 = add(mul(vort, A0Vort[idx]), mul(N, A0Buoy[idx]));
                cplx cAp, cAm;
                if (im == 0) {  // inertialCoefficientPair :107-110
                    const cplx a = mul(sub(U, mul(V, cplx{0.0, 1.0})), inertialScale[j]);
                    cAp = a; cAm = cj(a);
                } else {
                    const cplx buoy = mul(sub(N, mul(a0, NA0[idx])), ApmN[idx]);
                    cplx wave;
                    if (!withW) {
                        const cplx div = add(mul(U, cplx{0.0, md.k}), mul(V, cplx{0.0, md.l}));
                        wave = mul(ApmD[idx], div);
                    } else {
                        const cplx delta = mul(add(mul(U, md.cosA), mul(V, md.sinA)), ApmDScaled[idx]);
                        const cplx wbar = mul(cplx{0.0, (md.Kh / 2.0) * apmW[j]}, W);
                        wave = add(delta, wbar);
                    }
                    cAp = add(wave, buoy); cAm = sub(wave, buoy);
                }
                const cplx p = phase(omega[idx] * elapsed);
                Ap[idx] = mul(cAp, cj(p));
                Am[idx] = mul(cAm, p);
                A0[idx] = a0;
            }
        }
    }

    // F/G all-derivatives :372-394, :819-863, fieldFamilyDerivativeSpectrum :60-73.
    void fgAll(bool fFamily, const cplx* Apm, const cplx* A0, double* out) const {
        std::vector<Half> h(4, Half(c.Nz * rows));
        const size_t Nj = c.Nj, Nz = c.Nz;
        for (size_t im = 0; im < Nkl; ++im) {
            const Mode& md = modes[im];
            for (size_t j = 0; j < Nj; ++j) {
                const size_t idx = j + Nj * im;
                const double ws = fFamily ? fWaveScale[idx] : gWaveScale[j];
                const double gs = fFamily ? Fg[j] : Gg[j];
                cplx v = add(mul(Apm[idx], ws), mul(A0[idx], gs));
                cplx x = mul(v, cplx{0.0, md.k}), y = mul(v, cplx{0.0, md.l}), zz = mul(v, fFamily ? -m[j] : m[j]);
                if (md.conj) { v = cj(v); x = cj(x); y = cj(y); zz = cj(zz); }
                h[0][j + Nz * md.row] = v;
                h[1][j + Nz * md.row] = x;
                h[2][j + Nz * md.row] = y;
                h[3][j + Nz * md.row] = zz;
            }
        }
        for (auto& x : h) completeHermitian(x);
        if (fFamily) {
            for (int q = 0; q < 3; ++q) { dct(h[q]); normInverseDCT(h[q]); }
            dst(h[3]); normInverseDST(h[3]);
        } else {
            for (int q = 0; q < 3; ++q) { dst(h[q]); normInverseDST(h[q]); }
            dct(h[3]); normInverseDCT(h[3]);
        }
        for (int q = 0; q < 4; ++q) c2r(h[q], out + size_t(q) * R);
    }
};

void setMsg(char* msg, size_t len, const std::string& s) {
    if (msg && len) { std::snprintf(msg, len, "%s", s.c_str()); }
}

Oracle* O(void* h) { return static_cast<Oracle*>(h); }
const cplx* C(const double* p) { return reinterpret_cast<const cplx*>(p); }
cplx* C(double* p) { return reinterpret_cast<cplx*>(p); }

}  // namespace

extern "C" {

void wvo_set_threads(int threads) {
    gThreads = threads;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
}

void* wvo_create(const wvo_config* config, char* msg, size_t msglen) {
    auto* o = new Oracle();
    o->c = *config;
    const std::string err = o->build();
    if (!err.empty()) { setMsg(msg, msglen, err); delete o; return nullptr; }
    return o;
}
void wvo_destroy(void* h) { delete O(h); }
size_t wvo_nkl(void* h) { return O(h)->Nkl; }
double wvo_coriolis(void* h) { return O(h)->f; }

void wvo_modes(void* h, int64_t* kMode, int64_t* lMode, double* k, double* l, double* Kh, double* cosA, double* sinA, int64_t* row, uint8_t* conj) {
    const auto& md = O(h)->modes;
    for (size_t i = 0; i < md.size(); ++i) {
        kMode[i] = md[i].kMode; lMode[i] = md[i].lMode; k[i] = md[i].k; l[i] = md[i].l; Kh[i] = md[i].Kh;
        cosA[i] = md[i].cosA; sinA[i] = md[i].sinA; row[i] = int64_t(md[i].row); conj[i] = md[i].conj ? 1 : 0;
    }
}

long wvo_table(void* h, const char* name, double* out) {
    Oracle* o = O(h);
    const std::string n(name);
    auto real = [&](const std::vector<double>& v) { std::copy(v.begin(), v.end(), out); return long(v.size()); };
    auto cpx = [&](const std::vector<cplx>& v) {
        for (size_t i = 0; i < v.size(); ++i) { out[2 * i] = v[i].re; out[2 * i + 1] = v[i].im; }
        return long(2 * v.size());
    };
    if (n == "z") return real(o->z);
    if (n == "h0") return real(o->h0);
    if (n == "verticalWavenumber") return real(o->m);
    if (n == "Fg") return real(o->Fg);
    if (n == "Gg") return real(o->Gg);
    if (n == "inertialScale") return real(o->inertialScale);
    if (n == "gWaveScale") return real(o->gWaveScale);
    if (n == "apmWProjectionPrefactor") return real(o->apmW);
    if (n == "omega") return real(o->omega);
    if (n == "fWaveScale") return real(o->fWaveScale);
    if (n == "NApField") return real(o->NAp);
    if (n == "NA0Field") return real(o->NA0);
    if (n == "A0FromVorticity") return real(o->A0Vort);
    if (n == "A0FromBuoyancy") return real(o->A0Buoy);
    if (n == "ApmNProjection") return real(o->ApmN);
    if (n == "ApmDScaled") return real(o->ApmDScaled);
    if (n == "ApmTE") return real(o->ApmTE);
    if (n == "A0TE") return real(o->A0TE);
    if (n == "UApField") return cpx(o->UAp);
    if (n == "VApField") return cpx(o->VAp);
    if (n == "WApField") return cpx(o->WAp);
    if (n == "UA0Field") return cpx(o->UA0);
    if (n == "VA0Field") return cpx(o->VA0);
    if (n == "ApmDProjection") return cpx(o->ApmD);
    return -1;
}

int wvo_nonlinear_flux(void* h, double t, double t0, const double* Ap, const double* Am, const double* A0, double* Fp, double* Fm, double* F0) {
    O(h)->flux(t, t0, C(Ap), C(Am), C(A0), C(Fp), C(Fm), C(F0));
    return 0;
}

int wvo_nonlinear_flux_using_advection_fields(void* h, double t, double t0, const double* Ap, const double* Am,
                                              const double* A0, const double* uvw, double* Fp, double* Fm,
                                              double* F0) {
    O(h)->flux(t, t0, C(Ap), C(Am), C(A0), C(Fp), C(Fm), C(F0), uvw);
    return 0;
}

// transformStateFieldDerivatives :719-742 (phase, then the derivative
// synthesis of :907-945 for one target).
int wvo_state_field_derivatives(void* h, double t, double t0, const double* Ap, const double* Am, const double* A0,
                                int field, double* out) {
    Oracle* o = O(h);
    if (field < 0 || field > 3) return 1;
    std::vector<Oracle::Evolved> ev(o->M);
    for (size_t i = 0; i < o->M; ++i) ev[i] = o->evolved(C(Ap), C(Am), C(A0), i, phase(o->omega[i] * (t - t0)));
    o->synthesizeDerivatives(field, ev, out);
    return 0;
}

int wvo_inverse_uvweta(void* h, double t, double t0, const double* Ap, const double* Am, const double* A0, double* fields) {
    Oracle* o = O(h);
    std::vector<Oracle::Evolved> ev(o->M);
    for (size_t i = 0; i < o->M; ++i) ev[i] = o->evolved(C(Ap), C(Am), C(A0), i, phase(o->omega[i] * (t - t0)));
    o->synthesizeUVWEta(ev, fields);
    return 0;
}

int wvo_inverse_uvw(void* h, double t, double t0, const double* Ap, const double* Am, const double* A0, double* fields) {
    Oracle* o = O(h);
    std::vector<Oracle::Evolved> ev(o->M);
    for (size_t i = 0; i < o->M; ++i) ev[i] = o->evolved(C(Ap), C(Am), C(A0), i, phase(o->omega[i] * (t - t0)));
    o->synthesizeUVW(ev, fields);
    return 0;
}

int wvo_forward_uvweta(void* h, const double* fields, double t, double t0, double* Ap, double* Am, double* A0) {
    if (O(h)->c.isHydrostatic) return 1;
    O(h)->forward(fields, t, t0, true, C(Ap), C(Am), C(A0));
    return 0;
}

int wvo_forward_uveta(void* h, const double* fields, double t, double t0, double* Ap, double* Am, double* A0) {
    if (!O(h)->c.isHydrostatic) return 1;
    O(h)->forward(fields, t, t0, false, C(Ap), C(Am), C(A0));
    return 0;
}

int wvo_fg_all_derivatives(void* h, int fFamily, const double* Apm, const double* A0, double* fields) {
    O(h)->fgAll(fFamily != 0, C(Apm), C(A0), fields);
    return 0;
}

// WVFixedStepRK4::step (WVRungeKutta.cpp:335-462) for a system with no state
// constraints, repeated nsteps times: k1 at t; stage = y + (h/2) k1;
// weighted = k1; k2 at t+h/2; weighted += 2 k2; stage = y + (h/2) k2; k3 at
// t+h/2; weighted += 2 k3; stage = y + h k3; k4 at t+h; weighted += k4;
// y = y + (h/6) weighted.  scaledSum (:15-19) is value + scale*increment.
int wvo_rk4_advance(void* h, double* Ap, double* Am, double* A0, double* t, double t0, double dt, int nsteps) {
    Oracle* o = O(h);
    const size_t M = o->M;
    std::vector<cplx> y(3 * M), stage(3 * M), k(3 * M), weighted(3 * M);
    for (size_t i = 0; i < M; ++i) { y[i] = C(Ap)[i]; y[M + i] = C(Am)[i]; y[2 * M + i] = C(A0)[i]; }
    auto eval = [&](const std::vector<cplx>& s, double time) {
        o->flux(time, t0, &s[0], &s[M], &s[2 * M], &k[0], &k[M], &k[2 * M]);
    };
    auto scaled = [](cplx v, cplx inc, double s) { return cplx{v.re + s * inc.re, v.im + s * inc.im}; };
    for (int step = 0; step < nsteps; ++step) {
        const double ti = *t;
        eval(y, ti);
        weighted = k;
        for (size_t i = 0; i < 3 * M; ++i) stage[i] = scaled(y[i], k[i], 0.5 * dt);
        eval(stage, ti + 0.5 * dt);
        for (size_t i = 0; i < 3 * M; ++i) weighted[i] = scaled(weighted[i], k[i], 2.0);
        for (size_t i = 0; i < 3 * M; ++i) stage[i] = scaled(y[i], k[i], 0.5 * dt);
        eval(stage, ti + 0.5 * dt);
        for (size_t i = 0; i < 3 * M; ++i) weighted[i] = scaled(weighted[i], k[i], 2.0);
        for (size_t i = 0; i < 3 * M; ++i) stage[i] = scaled(y[i], k[i], dt);
        eval(stage, ti + dt);
        for (size_t i = 0; i < 3 * M; ++i) weighted[i] = scaled(weighted[i], k[i], 1.0);
        for (size_t i = 0; i < 3 * M; ++i) y[i] = scaled(y[i], weighted[i], dt / 6.0);
        *t = ti + dt;
    }
    for (size_t i = 0; i < M; ++i) { C(Ap)[i] = y[i]; C(Am)[i] = y[M + i]; C(A0)[i] = y[2 * M + i]; }
    return 0;
}

void wvo_rk4_linear_test(double lre, double lim, double* yv, double dt, int nsteps) {
    const cplx lambda{lre, lim};
    cplx y{yv[0], yv[1]};
    auto scaled = [](cplx v, cplx inc, double s) { return cplx{v.re + s * inc.re, v.im + s * inc.im}; };
    for (int s = 0; s < nsteps; ++s) {
        cplx k = mul(lambda, y), w = k;
        cplx st = scaled(y, k, 0.5 * dt);
        k = mul(lambda, st); w = scaled(w, k, 2.0);
        st = scaled(y, k, 0.5 * dt);
        k = mul(lambda, st); w = scaled(w, k, 2.0);
        st = scaled(y, k, dt);
        k = mul(lambda, st); w = scaled(w, k, 1.0);
        y = scaled(y, w, dt / 6.0);
    }
    yv[0] = y.re; yv[1] = y.im;
}

double wvo_energy(void* h, const double* Ap, const double* Am, const double* A0) {
    Oracle* o = O(h);
    double e = 0.0;
    for (size_t i = 0; i < o->M; ++i) {
        const cplx p = C(Ap)[i], mm = C(Am)[i], z0 = C(A0)[i];
        e += o->ApmTE[i] * ((p.re * p.re + p.im * p.im) + (mm.re * mm.re + mm.im * mm.im)) + o->A0TE[i] * (z0.re * z0.re + z0.im * z0.im);
    }
    return e;
}

double wvo_energy_spatial(void* h, const double* fields) {
    Oracle* o = O(h);
    const size_t R = o->R, plane = o->c.Nx * o->c.Ny, Nz = o->c.Nz;
    const double dz = o->c.Lz / double(Nz - 1), N2 = o->c.N0 * o->c.N0;
    double e = 0.0;
    for (size_t zz = 0; zz < Nz; ++zz) {
        const double w = (zz == 0 || zz + 1 == Nz) ? 0.5 * dz : dz;
        double s = 0.0;
        for (size_t q = 0; q < plane; ++q) {
            const size_t i = q + plane * zz;
            const double u = fields[i], v = fields[R + i], ww = fields[2 * R + i], eta = fields[3 * R + i];
            s += u * u + v * v + (o->c.isHydrostatic ? 0.0 : ww * ww) + N2 * eta * eta;
        }
        e += w * (s / double(plane)) / 2.0;
    }
    return e;
}

}  // extern "C"
