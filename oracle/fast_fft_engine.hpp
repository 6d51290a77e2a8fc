// WVFFTEngine with radix-2 FFTs and OpenMP batches, for timing the reference
// kernel on the host (test infrastructure; see fast_fft_engine.cpp).
#pragma once

#include "WaveVortexKernel/WVFFTEngine.hpp"

namespace wvfast {

class FastFFTEngine final : public wavevortex::WVFFTEngine {
public:
    explicit FastFFTEngine(int threads = 0);  // 0: OpenMP's default thread count
    std::string identifier() const override { return "build-radix2-openmp"; }
    std::size_t persistentBytes() const noexcept override { return sizeof(*this); }
    wavevortex::WVKernelStatus createPlan(const wavevortex::WVFFTPlanSpecification& specification,
                                          std::unique_ptr<wavevortex::WVFFTPlan>& plan) override;
    int threads() const { return threads_; }

private:
    int threads_;
};

}  // namespace wvfast
