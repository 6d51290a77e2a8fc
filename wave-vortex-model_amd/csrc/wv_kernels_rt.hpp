// The runtime-length FFT kernels (any horizontal length: mixed-radix or
// direct in-LDS transforms), the heaviest instantiations to compile.  Their
// explicit instantiation definitions live in wv_kernels_rt.hip, compiled in
// parallel with wv_kernels.hip, which declares them extern and launches
// them.  WV_RT_KERNEL(declaration) expands to the one or the other.
#pragma once

#define WV_RT_YFFT(T, L2N, DERIV)                                                                                    \
    WV_RT_KERNEL(__global__ void k_yfft<T, L2N, 0, DERIV>(const cx<T>*, cx<T>*, int, int, int, int, int, int,       \
                                                          const cx<T>*, int, const int*, int, int, YDerivPlan,       \
                                                          const T*, int))
#define WV_RT_XFUSED(T, NT, L2N, EXT)                                                                                \
    WV_RT_KERNEL(__global__ void k_xfused<T, NT, 1, L2N, false, EXT>(const cx<T>*, cx<T>*, int, int, int, int, int,  \
                                                                     int, const cx<T>*, const T*, const cx<T>*, T,   \
                                                                     const double*))
#define WV_RT_XINV(T, OP)                                                                                            \
    WV_RT_KERNEL(__global__ void k_xinv<T, OP, 0>(const cx<T>*, double*, int, int, int, int, int, int, int, int,      \
                                                  const cx<T>*, const double*, double))
#define WV_RT_XFWD(T)                                                                                                \
    WV_RT_KERNEL(__global__ void k_xfwd<T, 0>(const double*, cx<T>*, int, int, int, int, int, int, int, int,         \
                                              const cx<T>*, int))
#define WV_RT_YMASK(L2N)                                                                                             \
    WV_RT_KERNEL(__global__ void k_ymask<L2N, 0>(cx<double>*, int, int, int, int, int, const cx<double>*,           \
                                                 const double*))
#define WV_RT_XADJ(T, NT)                                                                                            \
    WV_RT_KERNEL(__global__ void k_xadj<T, NT>(const cx<T>*, const cx<T>*, cx<T>*, int, int, int, int, int, int,     \
                                               const cx<T>*, const T*))

#define WV_RT_KERNELS_FOR(T)                                                                                         \
    WV_RT_YFFT(T, 0, true)                                                                                           \
    WV_RT_YFFT(T, 0, false)                                                                                          \
    WV_RT_YFFT(T, kSmooth7Rounds, true)                                                                              \
    WV_RT_YFFT(T, kSmooth7Rounds, false)                                                                             \
    WV_RT_XFUSED(T, 3, kSmooth7, false)                                                                              \
    WV_RT_XFUSED(T, 4, kSmooth7, false)                                                                              \
    WV_RT_XFUSED(T, 5, kSmooth7, false)                                                                              \
    WV_RT_XFUSED(T, 3, 0, false)                                                                                     \
    WV_RT_XFUSED(T, 4, 0, false)                                                                                     \
    WV_RT_XFUSED(T, 5, 0, false)                                                                                     \
    WV_RT_XFUSED(T, 3, 0, true)                                                                                      \
    WV_RT_XFUSED(T, 4, 0, true)                                                                                      \
    WV_RT_XFUSED(T, 5, 0, true)                                                                                      \
    WV_RT_XINV(T, kXinvFields)                                                                                       \
    WV_RT_XINV(T, kXinvApv)                                                                                          \
    WV_RT_XFWD(T)                                                                                                    \
    WV_RT_XADJ(T, 3)                                                                                                 \
    WV_RT_XADJ(T, 4)

#define WV_RT_KERNELS         \
    WV_RT_KERNELS_FOR(float)  \
    WV_RT_KERNELS_FOR(double) \
    WV_RT_XINV(double, kXinvAdvect) \
    WV_RT_YMASK(0)            \
    WV_RT_YMASK(kSmooth7Rounds)
