// Device-side helpers for the wave-vortex HIP kernels (gfx950): complex
// arithmetic, MFMA wrappers and the in-LDS batched FFT used by the y- and
// x-direction transforms.
#pragma once

#include <hip/hip_runtime.h>

namespace wvamd {

template <typename T>
struct alignas(2 * sizeof(T)) cx {  // aligned as a pair: one 8- or 16-byte access
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

// FP64 products, sums and quotients rounded on their own: never contracted
// into fused multiply-adds, whichever kernel they are inlined into (device
// code is otherwise compiled with contraction, and the HIP headers' __dmul_rn
// family is a plain operator in this toolchain).  For arithmetic that must
// give the same bits in every kernel, or the bits of an uncontracted CPU
// computation.
__device__ __forceinline__ double mulRn(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double addRn(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double subRn(double a, double b) {
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ double divRn(double a, double b) {
#pragma clang fp contract(off)
    return a / b;
}

// ---------------------------------------------------------------- MFMA ----
// 16x16x4 MFMA, one A and one B element per lane, four C elements per lane.
// Lane l holds A[l%16][l/16] and B[l/16][l%16]; C element i of lane l is
// C[row(l/16, i)][l%16] with row = 4*(l/16) + i for f32 and (l/16) + 4*i for
// f64 (both measured on gfx950).
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
    __device__ static __forceinline__ int row(int kq, int i) { return 4 * kq + i; }
};
template <>
struct Mfma<double> {
    using acc = doublex4;
    __device__ static __forceinline__ acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
    __device__ static __forceinline__ acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int row(int kq, int i) { return kq + 4 * i; }
};

// ----------------------------------------------------------------- FFT ----
// In-place batched complex FFT in LDS, all threads of the block participate.
// Two batch layouts:
//   BatchRows<NB1, PAD>: transform b is contiguous (unit stride; PAD: one pad
//     slot per 16 entries so that radix-16 passes are free of bank conflicts)
//     and starts at (b % NB1)*sb + (b / NB1)*sb2; consecutive threads walk one
//     transform.
//   BatchCols: nb = 2^log2nb interleaved transforms, element n of transform b
//     at n*nb + b; consecutive threads walk the transforms.
// X_k = sum_n x_n exp(SIGN*2*pi*i*n*k/N).  Twiddles: for power-of-two N the
// pass table, entry r*Ns + k - 1 = exp(-2*pi*i*r*k/(Ns*R)) (k = 0..Ns-1) for
// the pass that combines sub-transforms of length Ns with radix R; otherwise
// tw[q] = exp(-2*pi*i*q/N).
// Power-of-two N runs Stockham autosort passes of radix 16 (then 8, 4 or 2
// for the remaining factor), each butterfly an in-register DFT, with all index
// arithmetic in shifts and masks; other N with prime factors up to 13 run
// mixed-radix passes (odd and composite radices first, then the powers of
// two) with runtime index arithmetic; the rest a direct DFT.  A pass holds
// about 16 complex values per thread: kFftThreads threads take at most
// kFftRegs * kFftThreads points per round (rows of more points run in rounds
// of whole transforms where the caller asks for them, ROUNDS).  Every
// workgroup that runs these transforms has kFftThreads threads.  The caller
// synchronizes before the call; the function returns after a final barrier.
constexpr int kFftRegs = 16;
constexpr int kFftThreads = 256;

// threadIdx.x of a kFftThreads-thread workgroup, with its range known to the
// compiler (items of a pass that every thread holds need no guard)
__device__ __forceinline__ int fftThread() {
    const int t = static_cast<int>(threadIdx.x);
    __builtin_assume(t >= 0 && t < kFftThreads);
    return t;
}

__device__ __forceinline__ int ldsIndex(int n, int sn) { return sn == 1 ? n + (n >> 4) : n * sn; }
// padded length of a unit-stride transform
__host__ __device__ constexpr int ldsLength(int N) { return N + (N >> 4); }

// Quotient and remainder by a runtime divisor 1 <= d <= 4096 for 0 <= a with
// a + d < 2^15 (FFT item and position indices), through an FP32 reciprocal
// raised by 2^-16: that covers the reciprocal's 1-ulp error and the
// product's rounding, and floor(a (1/d)(1 + 2^-16 +- 2^-22)) = floor(a / d)
// while (a / d + 1) d (2^-16 + 2^-21) < 1.  Checked exhaustively over that
// range with the reciprocal rounded down, to nearest and up.  The FFT
// passes divide item indices below (transforms per workgroup) x N <= 8 N and
// positions below N by N or a divisor of it; setup keeps 9 Nx < 2^15.
struct FastDiv {
    int d;
    float inv;
    __device__ __forceinline__ explicit FastDiv(int dv)
        : d(dv), inv(__builtin_amdgcn_rcpf(static_cast<float>(dv)) * (1.0f + 0x1p-16f)) {}
    __device__ __forceinline__ int quot(int a) const { return static_cast<int>(static_cast<float>(a) * inv); }
    __device__ __forceinline__ void divmod(int a, int& q, int& r) const {
        q = quot(a);
        r = a - __mul24(q, d);
    }
};

template <int NB1, bool PAD = true>
struct BatchRows {
    int nb, sb, sb2;
    __device__ __forceinline__ int base(int b) const {
        const unsigned u = static_cast<unsigned>(b);
        return static_cast<int>(u % NB1) * sb + static_cast<int>(u / NB1) * sb2;
    }
    __device__ __forceinline__ static int at(int n) { return PAD ? n + (n >> 4) : n; }
    // item -> (transform, position) when each transform holds 2^log2n items
    __device__ __forceinline__ void split(int item, int log2n, int& b, int& j) const {
        b = item >> log2n;
        j = item & ((1 << log2n) - 1);
    }
    __device__ __forceinline__ void splitN(int item, const FastDiv& n, int& b, int& j) const { n.divmod(item, b, j); }
};

struct BatchCols {
    int log2nb;
    __device__ __forceinline__ int nbt() const { return 1 << log2nb; }
    __device__ __forceinline__ int base(int b) const { return b; }
    __device__ __forceinline__ int at(int n) const { return n << log2nb; }
    __device__ __forceinline__ void split(int item, int, int& b, int& j) const {
        b = item & ((1 << log2nb) - 1);
        j = item >> log2nb;
    }
    __device__ __forceinline__ void splitN(int item, const FastDiv&, int& b, int& j) const { split(item, 0, b, j); }
};

template <int NB1, bool PAD>
__device__ __forceinline__ int batchCount(const BatchRows<NB1, PAD>& bt) { return bt.nb; }
__device__ __forceinline__ int batchCount(const BatchCols& bt) { return bt.nbt(); }

// exp(-2 pi i q / 16), q = 0..15
__device__ __forceinline__ double w16cos(int q) {
    constexpr double c[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                              0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                              -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                              0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
    return c[q & 15];
}
__device__ __forceinline__ double w16sin(int q) { return -w16cos((q + 12) & 15); }  // -sin(2 pi q/16)

// In-register DFT of R = 2, 4, 8 or 16 points, natural order in and out
// (radix-2 decimation in frequency, then a compile-time bit-reversal).
// Bit r of ZM marks input r as known zero: the first stage then copies or
// negates its partner instead of adding.
template <typename T, int R, int SIGN, unsigned ZM = 0>
__device__ __forceinline__ void dftReg(cx<T>* v) {
#pragma unroll
    for (int half = R / 2; half >= 1; half /= 2) {
#pragma unroll
        for (int s = 0; s < R; s += 2 * half) {
#pragma unroll
            for (int q = 0; q < half; ++q) {
                const cx<T> a = v[s + q], b = v[s + q + half];
                const bool za = half == R / 2 && ((ZM >> (s + q)) & 1u);
                const bool zb = half == R / 2 && ((ZM >> (s + q + half)) & 1u);
                v[s + q] = za ? b : zb ? a : cadd(a, b);
                const cx<T> d = za ? cx<T>{-b.re, -b.im} : zb ? a : csub(a, b);
                const int qi = q * (16 / (2 * half));  // exp(sign 2 pi i q/(2 half)) = W16^(qi)
                if (qi == 0) {
                    v[s + q + half] = d;
                } else if (qi == 4) {  // -/+ i
                    v[s + q + half] = SIGN < 0 ? cx<T>{d.im, -d.re} : cx<T>{-d.im, d.re};
                } else {
                    const T wr = T(w16cos(qi)), wi = T(SIGN < 0 ? w16sin(qi) : -w16sin(qi));
                    v[s + q + half] = {d.re * wr - d.im * wi, d.re * wi + d.im * wr};
                }
            }
        }
    }
    cx<T> tmp[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int r = 0;
#pragma unroll
        for (int b = 1, m = R / 2; b < R; b <<= 1, m >>= 1)
            if (k & b) r |= m;
        tmp[k] = v[r];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = tmp[k];
}

template <int R>
constexpr int log2Radix() { return R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4; }

// Items i0 + thread + it kFftThreads below itemEnd of one Stockham pass of
// radix R (stockhamPass).
template <typename T, int R, int SIGN, typename Bt, unsigned ZM>
__device__ __forceinline__ void stockhamRound(cx<T>* buf, int log2N, int log2Ns, const Bt& bt,
                                              const cx<T>* __restrict__ tw, int i0, int itemEnd) {
    constexpr int IPT = kFftRegs / R;
    constexpr int LR = log2Radix<R>();
    const int log2NR = log2N - LR;
    const int NR = 1 << log2NR;
    const int kmask = (1 << log2Ns) - 1;
    cx<T> v[IPT][R];
#pragma unroll
    for (int it = 0; it < IPT; ++it) {
        const int item = i0 + fftThread() + it * kFftThreads;
        if (item < itemEnd) {
            int b, jj;
            bt.split(item, log2NR, b, jj);
            const int k = jj & kmask;
            const int boff = bt.base(b);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if ((ZM >> r) & 1u) {
                    v[it][r] = cx<T>{T(0), T(0)};
                    continue;
                }
                cx<T> x = buf[bt.at(jj + r * NR) + boff];
                if (r > 0 && log2Ns > 0) {
                    // the table holds the unit twiddle at k = 0 too
                    cx<T> w = tw[(r << log2Ns) + k - 1];
                    if (SIGN > 0) w.im = -w.im;
                    x = cmul(x, w);
                }
                v[it][r] = x;
            }
            dftReg<T, R, SIGN, ZM>(v[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IPT; ++it) {
        const int item = i0 + fftThread() + it * kFftThreads;
        if (item < itemEnd) {
            int b, jj;
            bt.split(item, log2NR, b, jj);
            const int d = ((jj >> log2Ns) << (log2Ns + LR)) + (jj & kmask);
            const int boff = bt.base(b);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[bt.at(d + (r << log2Ns)) + boff] = v[it][r];
        }
    }
    __syncthreads();
}

// One Stockham pass of radix R over transforms of length 2^log2N whose
// sub-transforms so far have length 2^log2Ns.  Bit r of ZM marks input r of
// every butterfly as known zero (not read).  ROUNDS: transforms of more than
// kFftRegs * kFftThreads points together run in rounds of whole transforms
// (kCap is a multiple of NR; rows only).  Without, the caller holds at most
// that many points.
template <typename T, int R, int SIGN, typename Bt, unsigned ZM = 0, bool ROUNDS = true>
__device__ __forceinline__ void stockhamPass(cx<T>* buf, int log2N, int log2Ns, const Bt& bt,
                                             const cx<T>* __restrict__ tw) {
    const int items = batchCount(bt) << (log2N - log2Radix<R>());
    if constexpr (ROUNDS) {
        constexpr int kCap = kFftThreads * (kFftRegs / R);
        for (int i0 = 0; i0 < items; i0 += kCap)
            stockhamRound<T, R, SIGN, Bt, ZM>(buf, log2N, log2Ns, bt, tw, i0, items);
    } else {
        stockhamRound<T, R, SIGN, Bt, ZM>(buf, log2N, log2Ns, bt, tw, 0, items);
    }
}

template <typename T, int SIGN, bool ROUNDS, typename Bt>
__device__ void directDft(cx<T>* buf, int N, const Bt& bt, const cx<T>* __restrict__ tw) {
    const int nb = batchCount(bt);
    const FastDiv fN(N);
    // transforms b0 ... below bEnd
    auto round = [&](int b0, int bEnd) __attribute__((always_inline)) {
        const int i0 = b0 * N, items = bEnd * N;
        cx<T> out[kFftRegs];
#pragma unroll
        for (int it = 0; it < kFftRegs; ++it) {
            const int item = i0 + fftThread() + it * kFftThreads;
            if (item < items) {
                int b, k;
                bt.splitN(item, fN, b, k);
                cx<T> acc{T(0), T(0)};
                // twiddle index n k mod N, advanced by k per term
                for (int n = 0, q = 0; n < N; ++n) {
                    cx<T> w = tw[q];
                    if (SIGN > 0) w.im = -w.im;
                    acc = cadd(acc, cmul(buf[bt.at(n) + bt.base(b)], w));
                    q += k;
                    if (q >= N) q -= N;
                }
                out[it] = acc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < kFftRegs; ++it) {
            const int item = i0 + fftThread() + it * kFftThreads;
            if (item < items) {
                int b, k;
                bt.splitN(item, fN, b, k);
                buf[bt.at(k) + bt.base(b)] = out[it];
            }
        }
        __syncthreads();
    };
    if constexpr (ROUNDS) {
        // rounds of whole transforms of at most kFftRegs * kFftThreads points
        const int per = max(1, (kFftRegs * kFftThreads) / N);
        for (int b0 = 0; b0 < nb; b0 += per) round(b0, min(nb, b0 + per));
    } else {
        round(0, nb);
    }
}

// cos(2 pi m / P) and sin(2 pi m / P), m = 0 ... (P - 1)/2, for the odd
// primes of the mixed-radix passes (longer tables follow by symmetry).
template <int P>
struct PrimeRoots;
template <>
struct PrimeRoots<3> {
    static constexpr double c[2] = {1.0, -0.5};
    static constexpr double s[2] = {0.0, 0.8660254037844386};
};
template <>
struct PrimeRoots<5> {
    static constexpr double c[3] = {1.0, 0.30901699437494745, -0.8090169943749475};
    static constexpr double s[3] = {0.0, 0.9510565162951535, 0.5877852522924731};
};
template <>
struct PrimeRoots<7> {
    static constexpr double c[4] = {1.0, 0.6234898018587335, -0.2225209339563144, -0.9009688679024191};
    static constexpr double s[4] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.4338837391175581};
};
template <>
struct PrimeRoots<11> {
    static constexpr double c[6] = {1.0, 0.8412535328311812, 0.41541501300188644, -0.14231483827328514,
                                    -0.6548607339452851, -0.9594929736144974};
    static constexpr double s[6] = {0.0, 0.5406408174555976, 0.9096319953545183, 0.9898214418809327,
                                    0.7557495743542583, 0.28173255684142967};
};
template <>
struct PrimeRoots<13> {
    static constexpr double c[7] = {1.0, 0.8854560256532099, 0.5680647467311558, 0.12053668025532305,
                                    -0.3546048870425356, -0.7485107481711011, -0.970941817426052};
    static constexpr double s[7] = {0.0, 0.46472317204376856, 0.8229838658936564, 0.992708874098054,
                                    0.9350162426854148, 0.6631226582407952, 0.23931566428755777};
};

// In-register DFT of an odd prime P, natural order in and out: with
// a_n = x_n + x_{P-n} and b_n = x_n - x_{P-n} (n = 1 ... (P-1)/2),
//   X_0 = x_0 + sum a_n,
//   X_k, X_{P-k} = x_0 + sum a_n cos(2 pi n k / P) +/- i SIGN sum b_n sin(2 pi n k / P).
template <typename T, int P, int SIGN>
__device__ __forceinline__ void dftOddPrime(cx<T>* v) {
    constexpr int H = (P - 1) / 2;
    using PR = PrimeRoots<P>;
    cx<T> a[H + 1], b[H + 1];
#pragma unroll
    for (int n = 1; n <= H; ++n) {
        a[n] = cadd(v[n], v[P - n]);
        b[n] = csub(v[n], v[P - n]);
    }
    cx<T> out[P];
    cx<T> x0 = v[0];
    cx<T> sum = x0;
#pragma unroll
    for (int n = 1; n <= H; ++n) sum = cadd(sum, a[n]);
    out[0] = sum;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        cx<T> re = x0, im{T(0), T(0)};
#pragma unroll
        for (int n = 1; n <= H; ++n) {
            const int m = (n * k) % P;
            const T cm = T(m <= H ? PR::c[m] : PR::c[P - m]);
            const T sm = T(m <= H ? PR::s[m] : -PR::s[P - m]);
            re = {re.re + a[n].re * cm, re.im + a[n].im * cm};
            im = {im.re + b[n].re * sm, im.im + b[n].im * sm};
        }
        // i SIGN im
        const cx<T> rot = SIGN < 0 ? cx<T>{im.im, -im.re} : cx<T>{-im.im, im.re};
        out[k] = cadd(re, rot);
        out[P - k] = csub(re, rot);
    }
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = out[k];
}

// Composite radices of the mixed-radix passes: R = P1 P2 with coprime
// factors (6 = 3 2, 10 = 5 2, 12 = 3 4, 14 = 7 2, 15 = 3 5, 18 = 9 2), P1
// the odd one (3 for 15); 9 = 3 3 by Cooley-Tukey (dftCt9).  (20 = 5 4 and
// 24 = 3 8 took 320 and 384 to two passes, 4 % faster, but the larger
// kernels ran every other runtime length 2-4 % slower: not kept.)
template <int R>
constexpr int pfaFirst() { return R == 15 ? 3 : (R & 1) ? R : pfaFirst<R / 2>(); }
template <int R>
constexpr bool isPrimeRadix() { return R == 3 || R == 5 || R == 7 || R == 11 || R == 13; }
constexpr int inverseMod(int a, int m) {
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 0;
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void dftSmall(cx<T>* v);

// In-register DFT of R = P1 P2 points (gcd(P1, P2) = 1), natural order in
// and out, by the prime-factor algorithm: with the input map n = (P2 n1 +
// P1 n2) mod R and the output map k = (e1 k1 + e2 k2) mod R, e1 = 1 mod P1
// and 0 mod P2, e2 = 0 mod P1 and 1 mod P2, the DFT is a DFT of P1 points
// along n1 followed by one of P2 points along n2, with no twiddles between.
template <typename T, int P1, int P2, int SIGN>
__device__ __forceinline__ void dftPfa(cx<T>* v) {
    constexpr int R = P1 * P2;
    constexpr int e1 = P2 * inverseMod(P2 % P1, P1), e2 = P1 * inverseMod(P1 % P2, P2);
    static_assert(e1 % P1 == 1 && e1 % P2 == 0 && e2 % P1 == 0 && e2 % P2 == 1, "coprime factors");
    cx<T> a[P1][P2];
#pragma unroll
    for (int n1 = 0; n1 < P1; ++n1)
#pragma unroll
        for (int n2 = 0; n2 < P2; ++n2) a[n1][n2] = v[(P2 * n1 + P1 * n2) % R];
#pragma unroll
    for (int n2 = 0; n2 < P2; ++n2) {
        cx<T> t[P1];
#pragma unroll
        for (int n1 = 0; n1 < P1; ++n1) t[n1] = a[n1][n2];
        dftSmall<T, P1, SIGN>(t);
#pragma unroll
        for (int k1 = 0; k1 < P1; ++k1) a[k1][n2] = t[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < P1; ++k1) {
        dftSmall<T, P2, SIGN>(a[k1]);
#pragma unroll
        for (int k2 = 0; k2 < P2; ++k2) v[(e1 * k1 + e2 * k2) % R] = a[k1][k2];
    }
}

// In-register DFT of 9 points, natural order in and out, as 3 x 3
// Cooley-Tukey: with n = 3 n1 + n2 and k = k1 + 3 k2, the DFTs of 3 points
// along n1, the twiddles W9^(n2 k1) (W9 = exp(SIGN 2 pi i / 9)), then the DFTs
// of 3 points along n2.
template <typename T, int SIGN>
__device__ __forceinline__ void dftCt9(cx<T>* v) {
    // cos and sin of 2 pi m / 9 for m = n2 k1 = 1, 2, 4
    constexpr double c9[5] = {1.0, 0.766044443118978, 0.17364817766693041, 0.0, -0.9396926207859083};
    constexpr double s9[5] = {0.0, 0.6427876096865393, 0.984807753012208, 0.0, 0.3420201433256689};
    cx<T> y[3][3];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
        cx<T> t[3] = {v[n2], v[3 + n2], v[6 + n2]};
        dftOddPrime<T, 3, SIGN>(t);
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) {
            const int m = n2 * k1;
            if (m == 0) {
                y[k1][n2] = t[k1];
            } else {
                const cx<T> w{T(c9[m]), T(SIGN * s9[m])};
                y[k1][n2] = cmul(t[k1], w);
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
        dftOddPrime<T, 3, SIGN>(y[k1]);
#pragma unroll
        for (int k2 = 0; k2 < 3; ++k2) v[k1 + 3 * k2] = y[k1][k2];
    }
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void dftSmall(cx<T>* v) {
    if constexpr ((R & (R - 1)) == 0)
        dftReg<T, R, SIGN>(v);
    else if constexpr (isPrimeRadix<R>())
        dftOddPrime<T, R, SIGN>(v);
    else if constexpr (R == 9)
        dftCt9<T, SIGN>(v);
    else
        dftPfa<T, pfaFirst<R>(), R / pfaFirst<R>(), SIGN>(v);
}

// One Stockham pass of radix R (2, 4, 8, 16 or an odd prime up to 13) over
// transforms of a runtime length N whose sub-transforms so far have length
// Ns, with the full twiddle table tw[q] = exp(-2 pi i q / N): input r of the
// butterfly of item (transform, j), k = j mod Ns, is multiplied by
// exp(-2 pi i r k / (Ns R)) = tw[r k N / (Ns R)] (index < N), and its
// outputs go to (j / Ns) Ns R + k + r Ns.  A pass holds at most about
// kFftRegs complex values per thread: ceil(16 / R) butterflies (IPT; the
// composite radices run one butterfly per thread).
template <typename T, int R, int SIGN, typename Bt, int IPT = (kFftRegs + R - 1) / R, bool ROUNDS = true>
__device__ __forceinline__ void stockhamPassAny(cx<T>* buf, int N, int Ns, const Bt& bt,
                                                const cx<T>* __restrict__ tw) {
    const int NR = N / R;
    const int nb = batchCount(bt);
    const int step = N / (Ns * R);
    const FastDiv fNR(NR), fNs(Ns);
    // transforms b0 ... below bEnd
    auto round = [&](int b0, int bEnd) __attribute__((always_inline)) {
        const int i0 = b0 * NR, items = bEnd * NR;
        cx<T> v[IPT][R];
        int dst[IPT];  // output position (j / Ns) Ns R + k of each item
#pragma unroll
        for (int it = 0; it < IPT; ++it) {
            const int item = i0 + fftThread() + it * kFftThreads;
            if (item < items) {
                int b, jj, q, k;
                bt.splitN(item, fNR, b, jj);
                fNs.divmod(jj, q, k);
                const int boff = bt.base(b);
                dst[it] = __mul24(q, Ns * R) + k;
                const int ks = __mul24(k, step);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    cx<T> x = buf[bt.at(jj + r * NR) + boff];
                    if (r > 0 && Ns > 1) {
                        cx<T> w = tw[r * ks];
                        if (SIGN > 0) w.im = -w.im;
                        x = cmul(x, w);
                    }
                    v[it][r] = x;
                }
                dftSmall<T, R, SIGN>(v[it]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IPT; ++it) {
            const int item = i0 + fftThread() + it * kFftThreads;
            if (item < items) {
                int b, jj;
                bt.splitN(item, fNR, b, jj);
                const int boff = bt.base(b);
#pragma unroll
                for (int r = 0; r < R; ++r) buf[bt.at(dst[it] + r * Ns) + boff] = v[it][r];
            }
        }
        __syncthreads();
    };
    if constexpr (ROUNDS) {
        // rounds of whole transforms, at most kFftThreads IPT items each
        // (rows only)
        const int per = max(1, (kFftThreads * IPT) / NR);
        for (int b0 = 0; b0 < nb; b0 += per) round(b0, min(nb, b0 + per));
    } else {
        round(0, nb);
    }
}

template <typename T, int R, int SIGN, bool ROUNDS, typename Bt>
__device__ __forceinline__ void passAny(cx<T>* buf, int N, int Ns, const Bt& bt, const cx<T>* __restrict__ tw) {
    stockhamPassAny<T, R, SIGN, Bt, (kFftRegs + R - 1) / R, ROUNDS>(buf, N, Ns, bt, tw);
}

// Mixed-radix FFT of a runtime length N (passes of 15, 18, 6, 12, 10, 14 or
// 9 where they apply, of 3, 5, 7, 11, 13, then radix-16 passes for the power of two
// and 8/4/2 for its remainder) with the full twiddle table; false (nothing
// done) when N has a prime factor above 13.
template <typename T, int SIGN, int MAXP = 13, bool ROUNDS = true, typename Bt>
__device__ __forceinline__ bool mixedRadixFft(cx<T>* buf, int N, const Bt& bt, const cx<T>* __restrict__ tw) {
    int rest = N, p2 = 0, e3 = 0, e5 = 0, e7 = 0, e11 = 0, e13 = 0;
    while ((rest & 1) == 0) { rest >>= 1; ++p2; }
    while (rest % 3 == 0) { rest /= 3; ++e3; }
    while (rest % 5 == 0) { rest /= 5; ++e5; }
    while (rest % 7 == 0) { rest /= 7; ++e7; }
    while (rest % 11 == 0) { rest /= 11; ++e11; }
    while (rest % 13 == 0) { rest /= 13; ++e13; }
    if (rest != 1 || (MAXP < 11 && e11 + e13 > 0)) return false;
    // odd radices first: the power-of-two passes then scatter with an odd
    // stride Ns (no LDS bank conflicts in an unpadded layout).  Where every
    // butterfly of a pass has its own thread, pairs of factors run as one
    // composite pass (DFTs in registers, one pass fewer): 3 x 5, then the
    // power of two's remainder beyond its radix-16 passes with odd factors
    // (18 = 9 x 2, 6, 12, 10, 14), then 9 = 3 x 3.
    int Ns = 1;
    const int nb = batchCount(bt);
    auto single = [&](int R) { return nb * (N / R) <= kFftThreads; };
    for (; e3 > 0 && e5 > 0 && single(15); --e3, --e5, Ns *= 15)
        stockhamPassAny<T, 15, SIGN, Bt, 1, ROUNDS>(buf, N, Ns, bt, tw);
    if ((p2 & 3) == 1 && e3 > 1 && single(18)) {
        stockhamPassAny<T, 18, SIGN, Bt, 1, ROUNDS>(buf, N, Ns, bt, tw);
        e3 -= 2, --p2, Ns *= 18;
    } else if ((p2 & 3) == 1 && e3 > 0 && single(6)) {
        stockhamPassAny<T, 6, SIGN, Bt, 1, ROUNDS>(buf, N, Ns, bt, tw);
        --e3, --p2, Ns *= 6;
    } else if ((p2 & 3) == 2 && e3 > 0 && single(12)) {
        stockhamPassAny<T, 12, SIGN, Bt, 1, ROUNDS>(buf, N, Ns, bt, tw);
        --e3, p2 -= 2, Ns *= 12;
    } else if ((p2 & 3) == 1 && e5 > 0 && single(10)) {
        stockhamPassAny<T, 10, SIGN, Bt, 1, ROUNDS>(buf, N, Ns, bt, tw);
        --e5, --p2, Ns *= 10;
    } else if ((p2 & 3) == 1 && e7 > 0 && single(14)) {
        stockhamPassAny<T, 14, SIGN, Bt, 1, ROUNDS>(buf, N, Ns, bt, tw);
        --e7, --p2, Ns *= 14;
    }
    for (; e3 > 1 && single(9); e3 -= 2, Ns *= 9) stockhamPassAny<T, 9, SIGN, Bt, 1, ROUNDS>(buf, N, Ns, bt, tw);
    for (; e3 > 0; --e3, Ns *= 3) passAny<T, 3, SIGN, ROUNDS>(buf, N, Ns, bt, tw);
    for (; e5 > 0; --e5, Ns *= 5) passAny<T, 5, SIGN, ROUNDS>(buf, N, Ns, bt, tw);
    for (; e7 > 0; --e7, Ns *= 7) passAny<T, 7, SIGN, ROUNDS>(buf, N, Ns, bt, tw);
    if constexpr (MAXP >= 11) {
        for (; e11 > 0; --e11, Ns *= 11) passAny<T, 11, SIGN, ROUNDS>(buf, N, Ns, bt, tw);
        for (; e13 > 0; --e13, Ns *= 13) passAny<T, 13, SIGN, ROUNDS>(buf, N, Ns, bt, tw);
    }
    for (; p2 >= 4; p2 -= 4, Ns *= 16) passAny<T, 16, SIGN, ROUNDS>(buf, N, Ns, bt, tw);
    if (p2 == 3) passAny<T, 8, SIGN, ROUNDS>(buf, N, Ns, bt, tw);
    if (p2 == 2) passAny<T, 4, SIGN, ROUNDS>(buf, N, Ns, bt, tw);
    if (p2 == 1) passAny<T, 2, SIGN, ROUNDS>(buf, N, Ns, bt, tw);
    return true;
}

// Radix-2^LR passes from sub-transform length 2^L2S up to 2^L2N.
template <typename T, int SIGN, int L2N, int L2S, int LR, bool ROUNDS, typename Bt>
__device__ __forceinline__ void radixPasses(cx<T>* buf, const Bt& bt, const cx<T>* __restrict__ tw) {
    if constexpr (L2S < L2N) {
        stockhamPass<T, 1 << LR, SIGN, Bt, 0, ROUNDS>(buf, L2N, L2S, bt, tw);
        radixPasses<T, SIGN, L2N, L2S + LR, LR, ROUNDS>(buf, bt, tw);
    }
}

// FFT of a runtime power-of-two length 2^log2N: the leading pass takes the
// factor 2^(log2N mod 4); the rest are radix 16.
template <typename T, int SIGN, bool ROUNDS = true, typename Bt>
__device__ __forceinline__ void fftLdsPow2(cx<T>* buf, int log2N, const Bt& bt, const cx<T>* __restrict__ tw) {
    int l2s = 0;
    switch (log2N & 3) {
        case 1: stockhamPass<T, 2, SIGN, Bt, 0, ROUNDS>(buf, log2N, l2s, bt, tw); l2s = 1; break;
        case 2: stockhamPass<T, 4, SIGN, Bt, 0, ROUNDS>(buf, log2N, l2s, bt, tw); l2s = 2; break;
        case 3: stockhamPass<T, 8, SIGN, Bt, 0, ROUNDS>(buf, log2N, l2s, bt, tw); l2s = 3; break;
        default: break;
    }
    for (; l2s < log2N; l2s += 4) stockhamPass<T, 16, SIGN, Bt, 0, ROUNDS>(buf, log2N, l2s, bt, tw);
}

// FFT of a length that is not a power of two: mixed-radix passes when its
// prime factors are at most 13, else the direct DFT.  SMOOTH7: the caller
// has checked that they are at most 7 (only the passes of 2 ... 16, 3, 5
// and 7 are compiled: fewer registers).
template <typename T, int SIGN, bool SMOOTH7 = false, bool ROUNDS = true, typename Bt>
__device__ __forceinline__ void fftLdsAny(cx<T>* buf, int N, const Bt& bt, const cx<T>* __restrict__ tw) {
    if constexpr (SMOOTH7)
        mixedRadixFft<T, SIGN, 7, ROUNDS>(buf, N, bt, tw);
    else if (!mixedRadixFft<T, SIGN, 13, ROUNDS>(buf, N, bt, tw))
        directDft<T, SIGN, ROUNDS>(buf, N, bt, tw);
}

// Kernels take the length class as a template argument L2N: L2N > 0 fixes
// the length 2^L2N at compile time (N and log2N are then ignored): every
// pass and every index shift is a constant, and the passes are of radix
// 2^LR (LR = 4 or 2; the leading pass takes the remaining factor) with the
// twiddle table of that pass sequence.  Runtime lengths: L2N = 0 any
// length, kSmooth7Rounds a length whose prime factors are at most 7, both
// in passes written as rounds of whole transforms (any number of points;
// measured to take fewer registers in the y kernels); kSmooth7 the same in
// passes of one round (at most kFftRegs * kFftThreads points; fewer
// registers in the x kernels).  log2N < 0 marks a length that is not a
// power of two.
constexpr int kSmooth7 = -1;
constexpr int kSmooth7Rounds = -3;
template <int L2N>
constexpr bool fftRounds() { return L2N >= 0 || L2N == kSmooth7Rounds; }
template <int L2N>
constexpr bool fftSmooth7() { return L2N == kSmooth7 || L2N == kSmooth7Rounds; }
template <typename T, int SIGN, int L2N = 0, int LR = 4, typename Bt>
__device__ __forceinline__ void fftLds(cx<T>* buf, int N, int log2N, const Bt& bt, const cx<T>* __restrict__ tw) {
    if constexpr (L2N > 0) {
        constexpr int lead = L2N % LR;
        // rounds only where a caller holds more than kFftRegs * kFftThreads
        // points: 8 row transforms of 1024 (at most 8 transforms of 512,
        // or column tiles of at most 4096 points, elsewhere)
        constexpr bool kRounds = L2N >= 10;
        if constexpr (lead != 0) stockhamPass<T, 1 << lead, SIGN, Bt, 0, kRounds>(buf, L2N, 0, bt, tw);
        radixPasses<T, SIGN, L2N, lead, LR, kRounds>(buf, bt, tw);
        return;
    }
    if (log2N < 0)
        fftLdsAny<T, SIGN, fftSmooth7<L2N>(), fftRounds<L2N>()>(buf, N, bt, tw);
    else
        fftLdsPow2<T, SIGN, fftRounds<L2N>()>(buf, log2N, bt, tw);
}

// Row transforms of the x kernels (transform b at row b / NB1, slot b %
// NB1): power-of-two lengths in the padded layout (n + n/16, for their
// radix-16 passes, which scatter with a stride of 16), other lengths in the
// plain one (n; their odd-radix passes come first, and the addresses of the
// generic passes stay linear).  rowsPadded(log2N) says which.
__device__ __forceinline__ bool rowsPadded(int log2N) { return log2N >= 0; }
// (More than kFftRegs * kFftThreads points in all: every pass runs in rounds
// of whole transforms.)
template <typename T, int SIGN, int L2N, int NB1>
__device__ __forceinline__ void fftRows(cx<T>* buf, int N, int log2N, int nb, int sb, int sb2,
                                        const cx<T>* __restrict__ tw) {
    if constexpr (L2N > 0) {
        fftLds<T, SIGN, L2N>(buf, N, log2N, BatchRows<NB1, true>{nb, sb, sb2}, tw);
    } else {
        if (rowsPadded(log2N))
            fftLdsPow2<T, SIGN, fftRounds<L2N>()>(buf, log2N, BatchRows<NB1, true>{nb, sb, sb2}, tw);
        else
            fftLdsAny<T, SIGN, fftSmooth7<L2N>(), fftRounds<L2N>()>(buf, N, BatchRows<NB1, false>{nb, sb, sb2},
                                                                 tw);
    }
}

// nb unit-stride padded transforms at buf + b stride (any number of points:
// the passes run in rounds of whole transforms).
template <typename T, int SIGN, int L2N = 0>
__device__ __forceinline__ void fftLdsBatched(cx<T>* buf, int N, int log2N, int nb, int stride,
                                              const cx<T>* __restrict__ tw) {
    fftLds<T, SIGN, L2N>(buf, N, log2N, BatchRows<1>{nb, 0, stride}, tw);
}

// Forward (SIGN = -1) FFT of NB 256-point transforms (unit stride, padded,
// transform b at bt.base(b)) in the two radix-16 Stockham passes of fftLds
// (same twiddle table tw), on the first 32 NB threads, two per butterfly.
// With x_r (r < 16) the twiddled inputs of a butterfly and W =
// exp(-2 pi i / 16), its outputs 2p + s (s = 0, 1) are the 8-point DFT over
// r < 8 of
//   u_r = (x_r + (-1)^s x_{r+8}) W^{r s},
// so lane s of a pair forms u from all 16 inputs and runs an 8-point DFT:
// every lane busy, and fewer instructions per lane than one radix-16
// butterfly.  The caller synchronizes before the call; the function returns
// after a final barrier.
template <typename T, int NB, int NB1>
__device__ __forceinline__ void fftFwd256(cx<T>* buf, const BatchRows<NB1>& bt, const cx<T>* __restrict__ tw) {
    static_assert(32 * NB <= kFftThreads, "two threads per butterfly");
    const int t = fftThread();
    const bool active = t < 32 * NB;  // whole waves when NB is even
    // lane set-up inside the active branch: the idle waves skip it
    int s = 0, jj = 0;
    T sg = T(1);
    cx<T>* z = buf;
    cx<T> c[8];  // W^{r s}, r = 1..7
    if (active) {
        s = t & 1;
        jj = (t >> 1) & 15;
        sg = s ? T(-1) : T(1);
        z = buf + bt.base(t >> 5);
#pragma unroll
        for (int r = 1; r < 8; ++r) c[r] = s ? cx<T>{T(w16cos(r)), T(w16sin(r))} : cx<T>{T(1), T(0)};
    }
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        cx<T> u[8];
        if (active) {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                cx<T> x0 = z[BatchRows<1>::at(jj + 16 * r)];
                cx<T> x1 = z[BatchRows<1>::at(jj + 16 * (r + 8))];
                if (pass == 1) {
                    // pass twiddles exp(-2 pi i r k / 256), k = jj (the table
                    // holds the unit twiddle at k = 0 too)
                    if (r > 0) x0 = cmul(x0, tw[(r << 4) + jj - 1]);
                    x1 = cmul(x1, tw[((r + 8) << 4) + jj - 1]);
                }
                const cx<T> v{x0.re + sg * x1.re, x0.im + sg * x1.im};
                u[r] = r == 0 ? v : cmul(v, c[r]);
            }
            dftReg<T, 8, -1>(u);
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int p = 0; p < 8; ++p) {
                const int q = 2 * p + s;
                z[BatchRows<1>::at(pass == 0 ? 16 * jj + q : jj + 16 * q)] = u[p];
            }
        }
        __syncthreads();
    }
}

// Copies the N-entry twiddle table into LDS (block-wide; caller synchronizes).
template <typename T>
__device__ __forceinline__ void loadTwiddles(cx<T>* dst, const cx<T>* __restrict__ src, int N) {
    for (int i = fftThread(); i < N; i += kFftThreads) dst[i] = src[i];
}

}  // namespace wvamd
