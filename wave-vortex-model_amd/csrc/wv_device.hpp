// Device-side helpers for the wave-vortex HIP kernels (gfx950): complex
// arithmetic, MFMA wrappers and the in-LDS batched FFT used by the y- and
// x-direction transforms.
#pragma once

#include <hip/hip_runtime.h>

namespace wvamd {

template <typename T>
struct cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }
// a * (0 + i s)
template <typename T>
__device__ __forceinline__ cx<T> cmuli(cx<T> a, T s) { return {-a.im * s, a.re * s}; }

// ---------------------------------------------------------------- MFMA ----
// 16x16x4 MFMA, one A and one B element per lane, four C elements per lane.
// Lane l holds A[l%16][l/16], B[l/16][l%16] and C[4*(l/16)+i][l%16], i=0..3.
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    using acc = floatx4;
    __device__ static __forceinline__ acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
    __device__ static __forceinline__ acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<double> {
    using acc = doublex4;
    __device__ static __forceinline__ acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
    __device__ static __forceinline__ acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// ----------------------------------------------------------------- FFT ----
// In-place batched complex FFT in LDS, all threads of the block participate.
// Element n of transform b lives at buf[n*sn + b*sb] (nb transforms of length
// N).  X_k = sum_n x_n exp(sign*2*pi*i*n*k/N); tw[q] = exp(-2*pi*i*q/N).
// Power-of-two N uses Stockham autosort passes (one radix-2 pass when log2 N
// is odd, then radix-4); other N use a direct DFT.  Each pass holds at most 16
// complex values per thread, so the caller must launch at least
// nb*N/16 threads.  The caller synchronizes before the call; the function
// returns after a final barrier.
constexpr int kFftRegs = 16;

template <typename T>
__device__ __forceinline__ void butterfly4(cx<T>* v, int sign) {
    const cx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const cx<T> a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
    // (+/- i) * a3
    const cx<T> ia3 = sign < 0 ? cx<T>{a3.im, -a3.re} : cx<T>{-a3.im, a3.re};
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, ia3);
    v[3] = csub(a1, ia3);
}

template <typename T, int R>
__device__ __forceinline__ void stockhamPass(cx<T>* buf, int N, int Ns, int sn, int sb, int nb, bool bFast,
                                             const cx<T>* __restrict__ tw, int sign) {
    constexpr int IPT = kFftRegs / R;
    const int NR = N / R;
    const int items = nb * NR;
    const int tstep = N / (Ns * R);
    cx<T> v[IPT][R];
#pragma unroll
    for (int it = 0; it < IPT; ++it) {
        const int item = threadIdx.x + it * blockDim.x;
        if (item < items) {
            const int b = bFast ? item % nb : item / NR;
            const int jj = bFast ? item / nb : item % NR;
            const int k = jj % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cx<T> x = buf[(jj + r * NR) * sn + b * sb];
                if (r > 0 && k > 0) {
                    cx<T> w = tw[r * k * tstep];
                    if (sign > 0) w.im = -w.im;
                    x = cmul(x, w);
                }
                v[it][r] = x;
            }
            if constexpr (R == 4) {
                butterfly4(v[it], sign);
            } else {
                const cx<T> a = v[it][0], c = v[it][1];
                v[it][0] = cadd(a, c);
                v[it][1] = csub(a, c);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IPT; ++it) {
        const int item = threadIdx.x + it * blockDim.x;
        if (item < items) {
            const int b = bFast ? item % nb : item / NR;
            const int jj = bFast ? item / nb : item % NR;
            const int d = (jj / Ns) * Ns * R + (jj % Ns);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[(d + r * Ns) * sn + b * sb] = v[it][r];
        }
    }
    __syncthreads();
}

template <typename T>
__device__ void directDft(cx<T>* buf, int N, int sn, int sb, int nb, bool bFast, const cx<T>* __restrict__ tw,
                          int sign) {
    cx<T> out[kFftRegs];
    const int items = nb * N;
#pragma unroll
    for (int it = 0; it < kFftRegs; ++it) {
        const int item = threadIdx.x + it * blockDim.x;
        if (item < items) {
            const int b = bFast ? item % nb : item / N;
            const int k = bFast ? item / nb : item % N;
            cx<T> acc{T(0), T(0)};
            for (int n = 0; n < N; ++n) {
                cx<T> w = tw[(n * k) % N];
                if (sign > 0) w.im = -w.im;
                acc = cadd(acc, cmul(buf[n * sn + b * sb], w));
            }
            out[it] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kFftRegs; ++it) {
        const int item = threadIdx.x + it * blockDim.x;
        if (item < items) {
            const int b = bFast ? item % nb : item / N;
            const int k = bFast ? item / nb : item % N;
            buf[k * sn + b * sb] = out[it];
        }
    }
    __syncthreads();
}

template <typename T>
__device__ void fftLds(cx<T>* buf, int N, int log2N, int sn, int sb, int nb, bool bFast, const cx<T>* __restrict__ tw,
                       int sign) {
    if (log2N < 0) {
        directDft(buf, N, sn, sb, nb, bFast, tw, sign);
        return;
    }
    int Ns = 1;
    if (log2N & 1) {
        stockhamPass<T, 2>(buf, N, Ns, sn, sb, nb, bFast, tw, sign);
        Ns = 2;
    }
    for (; Ns < N; Ns *= 4) stockhamPass<T, 4>(buf, N, Ns, sn, sb, nb, bFast, tw, sign);
}

}  // namespace wvamd
