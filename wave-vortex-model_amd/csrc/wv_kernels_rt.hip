// Explicit instantiations of the runtime-length FFT kernels (the list and
// its reason are in wv_kernels_rt.hpp); wv_kernels.hip launches them.
#include "wv_kernels_device.hpp"
#include "wv_kernels_rt.hpp"

namespace wvamd {

#define WV_RT_KERNEL(...) template __VA_ARGS__;
WV_RT_KERNELS
#undef WV_RT_KERNEL

}  // namespace wvamd
