// A WVFFTEngine (CompiledKernel/include/WaveVortexKernel/WVFFTEngine.hpp:10-50)
// with fast transforms, written for this build, so that the REFERENCE kernel
// (CompiledKernel/src, compiled unchanged) can be timed on the host cores at
// production sizes.  The reference ships only a direct-DFT engine
// (adapters/reference, O(N^2) per dimension) and an FFTW adapter; FFTW is
// not installed here.
//
// TEST INFRASTRUCTURE ONLY (the CPU baseline of the benchmark and the
// engine comparison test).  Semantics are those of
// adapters/reference/WVReferenceFFTEngine.cpp, checked against it by
// tests/test_reference_fast_engine.py:
//   horizontalRealToComplex2D  unnormalized 2-D r2c, exp(-i), kx <= Nx/2
//   horizontalComplexToReal2D  unnormalized 2-D c2r, exp(+i), the other half
//                              implied by conjugation (real parts at kx = 0
//                              and kx = Nx/2)
//   verticalDCTI  X_k = x_0 + (-1)^k x_{N-1} + 2 sum_{0<n<N-1} x_n cos(pi n k/(N-1))
//   verticalDSTI  X_k = 2 sum_n x_n sin(pi (n+1)(k+1)/(N+1))
// Method: iterative radix-2 complex FFTs (twiddles from std::cos/std::sin of
// the exact angles) for power-of-two lengths, a direct DFT otherwise; two
// real signals per complex FFT (rows of the 2-D transforms, batch pairs of
// the DCT-I / DST-I, which are the real-even / real-odd extensions of length
// 2(N-1) / 2(N+1)); batches distributed over OpenMP threads.
#include "fast_fft_engine.hpp"

#include <cmath>
#include <cstdint>
#include <functional>
#include <vector>

#include <omp.h>

namespace wvfast {
namespace {

using wavevortex::WVComplex64;
using wavevortex::WVFFTPlanKind;
using wavevortex::WVFFTPlanSpecification;
using wavevortex::WVKernelStatus;
using wavevortex::WVKernelStatusCode;

constexpr double kPi = 3.141592653589793238462643383279502884;

struct cd {
    double re, im;
};
inline cd add(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
inline cd sub(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
inline cd mul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cd conj(cd a) { return {a.re, -a.im}; }

// In-place complex DFT of length n: forward exp(-2 pi i jk/n), inverse
// exp(+2 pi i jk/n), both unnormalized.
class ComplexFft {
public:
    explicit ComplexFft(std::size_t n) : n_(n), pow2_(n > 0 && (n & (n - 1)) == 0) {
        const std::size_t m = pow2_ ? n / 2 : n;
        tw_.resize(m > 0 ? m : 1);
        for (std::size_t k = 0; k < m; ++k) {
            const double a = -2.0 * kPi * static_cast<double>(k) / static_cast<double>(n);
            tw_[k] = {std::cos(a), std::sin(a)};
        }
        if (pow2_) {
            rev_.resize(n);
            int bits = 0;
            while ((std::size_t(1) << bits) < n) ++bits;
            for (std::size_t i = 0; i < n; ++i) {
                std::size_t r = 0;
                for (int b = 0; b < bits; ++b)
                    if (i & (std::size_t(1) << b)) r |= std::size_t(1) << (bits - 1 - b);
                rev_[i] = static_cast<std::uint32_t>(r);
            }
        }
    }
    std::size_t size() const { return n_; }

    void run(cd* a, bool inverse, std::vector<cd>& scratch) const {
        if (n_ <= 1) return;
        if (!pow2_) {
            direct(a, inverse, scratch);
            return;
        }
        for (std::size_t i = 0; i < n_; ++i)
            if (rev_[i] > i) std::swap(a[i], a[rev_[i]]);
        for (std::size_t len = 2; len <= n_; len <<= 1) {
            const std::size_t half = len >> 1, step = n_ / len;
            for (std::size_t i = 0; i < n_; i += len) {
                for (std::size_t j = 0; j < half; ++j) {
                    cd w = tw_[j * step];
                    if (inverse) w.im = -w.im;
                    const cd u = a[i + j], v = mul(a[i + j + half], w);
                    a[i + j] = add(u, v);
                    a[i + j + half] = sub(u, v);
                }
            }
        }
    }

private:
    void direct(cd* a, bool inverse, std::vector<cd>& scratch) const {
        scratch.assign(a, a + n_);
        for (std::size_t k = 0; k < n_; ++k) {
            cd s{0.0, 0.0};
            for (std::size_t j = 0; j < n_; ++j) {
                cd w = tw_[(j * k) % n_];
                if (inverse) w.im = -w.im;
                s = add(s, mul(scratch[j], w));
            }
            a[k] = s;
        }
    }

    std::size_t n_;
    bool pow2_;
    std::vector<cd> tw_;
    std::vector<std::uint32_t> rev_;
};

std::vector<std::pair<std::ptrdiff_t, std::ptrdiff_t>> flattenBatches(const WVFFTPlanSpecification& s) {
    std::vector<std::pair<std::ptrdiff_t, std::ptrdiff_t>> out;
    std::function<void(std::size_t, std::ptrdiff_t, std::ptrdiff_t)> visit = [&](std::size_t d, std::ptrdiff_t i,
                                                                               std::ptrdiff_t o) {
        if (d == s.batchDimensions.size()) {
            out.emplace_back(i, o);
            return;
        }
        const auto& b = s.batchDimensions[d];
        for (std::size_t n = 0; n < b.count; ++n)
            visit(d + 1, i + static_cast<std::ptrdiff_t>(n) * b.inputStride, o + static_cast<std::ptrdiff_t>(n) * b.outputStride);
    };
    visit(0, 0, 0);
    return out;
}

class FastPlan final : public wavevortex::WVFFTPlan {
public:
    FastPlan(WVFFTPlanSpecification s, int threads)
        : spec_(std::move(s)), batches_(flattenBatches(spec_)), threads_(threads) {
        switch (spec_.kind) {
            case WVFFTPlanKind::horizontalRealToComplex2D:
            case WVFFTPlanKind::horizontalComplexToReal2D:
                fy_ = std::make_unique<ComplexFft>(spec_.transformDimensions[0].count);
                fx_ = std::make_unique<ComplexFft>(spec_.transformDimensions[1].count);
                break;
            case WVFFTPlanKind::verticalDCTI: {
                const std::size_t N = spec_.transformDimensions[0].count;
                fz_ = std::make_unique<ComplexFft>(N >= 2 ? 2 * (N - 1) : 1);
                break;
            }
            case WVFFTPlanKind::verticalDSTI:
                fz_ = std::make_unique<ComplexFft>(2 * (spec_.transformDimensions[0].count + 1));
                break;
        }
    }

    WVKernelStatus execute(const void* input, void* output) override {
        if (!input || !output) return {WVKernelStatusCode::invalidPointer, "FFT received a null pointer."};
        switch (spec_.kind) {
            case WVFFTPlanKind::horizontalRealToComplex2D:
                if (spec_.transformDimensions.size() != 2) break;
                r2c(static_cast<const double*>(input), static_cast<WVComplex64*>(output));
                return WVKernelStatus::ok();
            case WVFFTPlanKind::horizontalComplexToReal2D:
                if (spec_.transformDimensions.size() != 2) break;
                c2r(static_cast<const WVComplex64*>(input), static_cast<double*>(output));
                return WVKernelStatus::ok();
            case WVFFTPlanKind::verticalDCTI:
            case WVFFTPlanKind::verticalDSTI:
                if (spec_.transformDimensions.size() != 1) break;
                r2r(static_cast<const double*>(input), static_cast<double*>(output),
                    spec_.kind == WVFFTPlanKind::verticalDSTI);
                return WVKernelStatus::ok();
        }
        return {WVKernelStatusCode::invalidConfiguration, "Unsupported FFT plan rank."};
    }

    std::size_t persistentBytes() const noexcept override {
        return sizeof(*this) + batches_.capacity() * sizeof(batches_[0]);
    }

private:
    void r2c(const double* in, WVComplex64* out) const {
        const auto& dy = spec_.transformDimensions[0];
        const auto& dx = spec_.transformDimensions[1];
        const std::size_t Ny = dy.count, Nx = dx.count, H = Nx / 2 + 1;
        const long nb = static_cast<long>(batches_.size());
#pragma omp parallel num_threads(threads_)
        {
            std::vector<cd> tmp(Ny * H), z(Nx), col(Ny), scratch;
#pragma omp for schedule(static)
            for (long b = 0; b < nb; ++b) {
                const double* src = in + batches_[b].first;
                WVComplex64* dst = out + batches_[b].second;
                for (std::size_t y = 0; y < Ny; y += 2) {
                    const bool two = y + 1 < Ny;
                    for (std::size_t x = 0; x < Nx; ++x) {
                        const std::ptrdiff_t o = static_cast<std::ptrdiff_t>(x) * dx.inputStride;
                        z[x] = {src[static_cast<std::ptrdiff_t>(y) * dy.inputStride + o],
                                two ? src[static_cast<std::ptrdiff_t>(y + 1) * dy.inputStride + o] : 0.0};
                    }
                    fx_->run(z.data(), false, scratch);
                    for (std::size_t k = 0; k < H; ++k) {
                        const cd zk = z[k], zc = conj(z[(Nx - k) % Nx]);
                        tmp[y * H + k] = {0.5 * (zk.re + zc.re), 0.5 * (zk.im + zc.im)};
                        // (zk - zc) / (2 i)
                        if (two) tmp[(y + 1) * H + k] = {0.5 * (zk.im - zc.im), -0.5 * (zk.re - zc.re)};
                    }
                }
                for (std::size_t k = 0; k < H; ++k) {
                    for (std::size_t y = 0; y < Ny; ++y) col[y] = tmp[y * H + k];
                    fy_->run(col.data(), false, scratch);
                    for (std::size_t y = 0; y < Ny; ++y) {
                        WVComplex64& v = dst[static_cast<std::ptrdiff_t>(y) * dy.outputStride +
                                             static_cast<std::ptrdiff_t>(k) * dx.outputStride];
                        v.real = col[y].re;
                        v.imag = col[y].im;
                    }
                }
            }
        }
    }

    void c2r(const WVComplex64* in, double* out) const {
        const auto& dy = spec_.transformDimensions[0];
        const auto& dx = spec_.transformDimensions[1];
        const std::size_t Ny = dy.count, Nx = dx.count, H = Nx / 2 + 1;
        const long nb = static_cast<long>(batches_.size());
#pragma omp parallel num_threads(threads_)
        {
            std::vector<cd> tmp(Ny * H), z(Nx), col(Ny), scratch;
#pragma omp for schedule(static)
            for (long b = 0; b < nb; ++b) {
                const WVComplex64* src = in + batches_[b].first;
                double* dst = out + batches_[b].second;
                for (std::size_t k = 0; k < H; ++k) {
                    for (std::size_t y = 0; y < Ny; ++y) {
                        const WVComplex64& v = src[static_cast<std::ptrdiff_t>(y) * dy.inputStride +
                                                   static_cast<std::ptrdiff_t>(k) * dx.inputStride];
                        col[y] = {v.real, v.imag};
                    }
                    fy_->run(col.data(), true, scratch);
                    for (std::size_t y = 0; y < Ny; ++y) tmp[y * H + k] = col[y];
                }
                for (std::size_t y = 0; y < Ny; y += 2) {
                    const bool two = y + 1 < Ny;
                    for (std::size_t k = 0; k < H; ++k) {
                        cd a = tmp[y * H + k], c = two ? tmp[(y + 1) * H + k] : cd{0.0, 0.0};
                        const bool edge = k == 0 || 2 * k == Nx;  // real parts only
                        if (edge) a.im = c.im = 0.0;
                        z[k] = {a.re - c.im, a.im + c.re};  // a + i c
                        if (!edge) z[Nx - k] = {a.re + c.im, c.re - a.im};  // conj(a) + i conj(c)
                    }
                    fx_->run(z.data(), true, scratch);
                    for (std::size_t x = 0; x < Nx; ++x) {
                        const std::ptrdiff_t o = static_cast<std::ptrdiff_t>(x) * dx.outputStride;
                        dst[static_cast<std::ptrdiff_t>(y) * dy.outputStride + o] = z[x].re;
                        if (two) dst[static_cast<std::ptrdiff_t>(y + 1) * dy.outputStride + o] = z[x].im;
                    }
                }
            }
        }
    }

    void r2r(const double* in, double* out, bool sine) const {
        const auto& d = spec_.transformDimensions[0];
        const std::size_t N = d.count;
        const long npairs = static_cast<long>((batches_.size() + 1) / 2);
        if (!sine && N < 3) {  // degenerate lengths: the defining sum itself
            for (const auto& b : batches_) {
                std::vector<double> x(N);
                for (std::size_t n = 0; n < N; ++n) x[n] = in[b.first + static_cast<std::ptrdiff_t>(n) * d.inputStride];
                for (std::size_t k = 0; k < N; ++k)
                    out[b.second + static_cast<std::ptrdiff_t>(k) * d.outputStride] =
                        x.front() + (k % 2 == 0 ? x.back() : -x.back());
            }
            return;
        }
        const std::size_t M = fz_->size();
#pragma omp parallel num_threads(threads_)
        {
            std::vector<cd> z(M), scratch;
#pragma omp for schedule(static)
            for (long p = 0; p < npairs; ++p) {
                const auto& ba = batches_[2 * p];
                const bool two = static_cast<std::size_t>(2 * p + 1) < batches_.size();
                const auto& bb = two ? batches_[2 * p + 1] : ba;
                auto at = [&](const std::pair<std::ptrdiff_t, std::ptrdiff_t>& b, std::size_t n) {
                    return in[b.first + static_cast<std::ptrdiff_t>(n) * d.inputStride];
                };
                for (std::size_t i = 0; i < M; ++i) z[i] = {0.0, 0.0};
                if (!sine) {  // real-even extension of length 2 (N - 1)
                    for (std::size_t n = 0; n < N; ++n) {
                        const cd v{at(ba, n), two ? at(bb, n) : 0.0};
                        z[n] = v;
                        if (n > 0 && n + 1 < N) z[M - n] = v;
                    }
                } else {  // real-odd extension of length 2 (N + 1)
                    for (std::size_t n = 0; n < N; ++n) {
                        const cd v{at(ba, n), two ? at(bb, n) : 0.0};
                        z[n + 1] = v;
                        z[M - n - 1] = {-v.re, -v.im};
                    }
                }
                fz_->run(z.data(), false, scratch);
                for (std::size_t k = 0; k < N; ++k) {
                    // even: the transform of each real part is real; odd: it is
                    // -2 i times the sine sum
                    const double xa = sine ? -z[k + 1].im : z[k].re;
                    const double xb = sine ? z[k + 1].re : z[k].im;
                    out[ba.second + static_cast<std::ptrdiff_t>(k) * d.outputStride] = xa;
                    if (two) out[bb.second + static_cast<std::ptrdiff_t>(k) * d.outputStride] = xb;
                }
            }
        }
    }

    WVFFTPlanSpecification spec_;
    std::vector<std::pair<std::ptrdiff_t, std::ptrdiff_t>> batches_;
    int threads_;
    std::unique_ptr<ComplexFft> fx_, fy_, fz_;
};

}  // namespace

FastFFTEngine::FastFFTEngine(int threads) : threads_(threads > 0 ? threads : omp_get_max_threads()) {}

wavevortex::WVKernelStatus FastFFTEngine::createPlan(const wavevortex::WVFFTPlanSpecification& specification,
                                                     std::unique_ptr<wavevortex::WVFFTPlan>& plan) {
    if (specification.transformDimensions.empty())
        return {WVKernelStatusCode::invalidConfiguration, "FFT plan without transform dimensions."};
    plan = std::make_unique<FastPlan>(specification, threads_);
    return WVKernelStatus::ok();
}

}  // namespace wvfast
