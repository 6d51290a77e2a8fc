// MI355X (gfx950) implementation of the wave-vortex nonlinear flux and the
// fixed-step RK4 that drives it.  Host entry points are the C-ABI declared in
// include/wv_hip.h.
//
// One right-hand side (reference: WVTransformConstantStratificationKernel::
// nonlinearFluxImpl, CompiledKernel/src/WVTransformConstantStratificationKernel.cpp:1173-1244)
// runs as six kernels, all resident in HBM:
//
//   k_fields   per coefficient: phase exp(i w (t-t0)) and the evolved field
//              spectra u, v, w, eta (WVCoefficientFormulas.hpp:15-33)
//   k_synth    per retained half-spectrum column: spectral assembly of the 15
//              (12 hydrostatic) channels -- U, V, W and d/dx, d/dy, d/dz of
//              every target field (...Kernel.cpp:246-291) -- and the vertical
//              DCT-I / DST-I as dense MFMA products over the Nj retained modes
//   k_yfft     inverse complex FFT along y on [Ny, kx-tile] LDS tiles
//   k_xfused   per (y, z) row: c2r of all channels along x, the advection
//              products F_T = -(U dT/dx + V dT/dy + W dT/dz) (...Kernel.cpp:1230)
//              and r2c of the products, all in LDS; spatial fields never
//              reach HBM
//   k_yfft     forward complex FFT along y
//   k_proj     per coefficient tile: vertical forward DCT-I / DST-I with its
//              normalization as MFMA products over z, the wave-vortex
//              projection of every target (...Kernel.cpp:396-431) and either
//              the flux store or the fused RK4 stage update
//              (PortableRuntime/src/WVRungeKutta.cpp:335-462)
//
// Half-spectrum data are kept only for kx < nkx (the largest retained kx)
// in [channel][z][iy][kx] order, so y-FFT tiles read and write contiguous kx
// runs and the x transforms read contiguous rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "wv_descriptor.hpp"
#include "wv_device.hpp"
#include "wv_hip.h"

namespace wvamd {

// ------------------------------------------------------------ device data --
struct ColumnDev {
    int iy, kx, wv;
    unsigned flags;
    double k, l, Kh, cosA, sinA;
};

// Vertical synthesis plan.  A channel is the evolved spectrum of field
// `field` (0=u, 1=v, 2=w, 3=eta) times valueScale or, with dz set, its z
// derivative in spectral form (factor -m_j/2 into the sine family, +m_j/2
// into the cosine family: the constant-N assembly, ...Kernel.cpp:240,286-288,
// WVCoefficientFormulas.hpp:42-58, whose value channels carry valueScale 1/2),
// synthesized with the group's matrix `mat`:
//   constant N:  0 DCT-I, 1 DST-I
//   arbitrary N: 0 F, 1 G, 2 -(N^2/g) G, 3 F/h, 4 F/h + dLnN2 G (z derivatives
//                are dense operators, valueScale 1)
// The result goes to up to three half-spectrum channels: outV (the value),
// outX (i k times it) and outY (i l times it), k and l the wavenumbers of the
// storage row; -1 = not written.  Horizontal derivatives commute with the
// vertical transform, so they cost a scaling instead of a synthesis.  The
// sine-family endpoint zeroing is carried by the DST-I / G matrices, whose
// j = 0 column and end rows are zero.
struct SynthChannel {
    int field, dz, outV, outX, outY;
};
struct SynthGroup {
    int mat, nch;
    SynthChannel ch[4];
};
constexpr int kMaxGroups = 8;
constexpr int kSynthMats = 5;      // distinct synthesis operators
constexpr int kSynthMatSlots = 8;  // operands one synthesis launch can take
struct SynthPlan {
    int ngroups;
    double valueScale;
    int skipCompletion;  // leave the Hermitian-completion columns untouched
    SynthGroup g[kMaxGroups];
};
template <typename T>
struct SynthMats {
    const T* m[kSynthMatSlots];
};

enum EpilogueMode : int { kFluxOut = 0, kRk1 = 1, kRk2 = 2, kRk3 = 3, kRk4 = 4 };

struct ProjTables {
    const double* omega;
    const double* A0Vort;
    const double* A0Buoy;
    const double* NA0;
    const double* ApmN;
    const double* ApmDScaled;
    const cx<double>* ApmD;
    const double* apmW;
    const double* inertialScale;
};

struct Epilogue {
    int mode;
    double scale;  // 0.5h (stages 1,2), h (stage 3), h/6 (stage 4)
    // RK modes: also store the field spectra of the state the stage produces
    // (the next stage's input; after stage 4 the new y) at elapsed time
    // nextElapsed, the input of the next flux evaluation.  Stages 1-3 then
    // do not store the stage state itself (st*): the next evaluation starts
    // from these spectra and nothing else reads it.
    int emitFields;
    double nextElapsed;
    cx<double>* out0;
    cx<double>* out1;
    cx<double>* out2;
    cx<double>* y0;
    cx<double>* y1;
    cx<double>* y2;
    cx<double>* st0;
    cx<double>* st1;
    cx<double>* st2;
    cx<double>* w0;
    cx<double>* w1;
    cx<double>* w2;
};

// ----------------------------------------------------------------- kernels -

// Index of the calling wave in its workgroup, as a wave-uniform (scalar) value.
__device__ __forceinline__ int waveIndex() { return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6)); }

// Workgroups go to the 8 XCDs in turn (block b to XCD b % 8), each XCD with
// its own L2.  This renumbers them so that every XCD runs one contiguous range
// of logical blocks, and blocks that share cache lines share an L2.
__device__ __forceinline__ int xcdBlock(int b, int nblocks) {
    constexpr int kXcds = 8;
    const int x = b % kXcds, q = nblocks / kXcds, r = nblocks % kXcds;
    return x * q + min(x, r) + b / kXcds;
}



// Phase p = exp(i omega elapsed) (phase(), WVCoefficientFormulas.hpp:15).
__device__ __forceinline__ cx<double> phaseAt(double omega, double elapsed) {
    double sn, cs;
    sincos(mulRn(omega, elapsed), &sn, &cs);  // the angle rounded on its own in every kernel
    return cx<double>{cs, sn};
}

// Evolved field spectra (u, v, w, eta) of coefficient idx with phase p
// (WVCoefficientFormulas.hpp:15-33), stored to FV[f*M + idx].
// Every operation is rounded on its own (no contraction into fused
// multiply-adds), so the spectra are the same bits in every kernel that
// forms them (k_fields from memory, k_stage from the stage it produced).
template <typename T>
__device__ __forceinline__ void storeFieldSpectra(cx<double> UAp, cx<double> VAp, cx<double> WAp, double NAp,
                                                  cx<double> UA0, cx<double> VA0, double NA0, int idx, int M,
                                                  cx<double> Ap, cx<double> Am, cx<double> A0, cx<double> p,
                                                  cx<T>* __restrict__ FV) {
    auto mul = [](cx<double> a, cx<double> b) {
        return cx<double>{subRn(mulRn(a.re, b.re), mulRn(a.im, b.im)),
                          addRn(mulRn(a.re, b.im), mulRn(a.im, b.re))};
    };
    auto add = [](cx<double> a, cx<double> b) { return cx<double>{addRn(a.re, b.re), addRn(a.im, b.im)}; };
    auto scale = [](cx<double> a, double s) { return cx<double>{mulRn(a.re, s), mulRn(a.im, s)}; };
    const cx<double> ap = mul(Ap, p), am = mul(Am, cconj(p)), a0 = A0;
    const cx<double> u = add(add(mul(UAp, ap), mul(cconj(UAp), am)), mul(UA0, a0));
    const cx<double> v = add(add(mul(VAp, ap), mul(cconj(VAp), am)), mul(VA0, a0));
    const cx<double> w = add(mul(WAp, ap), mul(WAp, am));
    const cx<double> e = add(add(scale(ap, NAp), scale(am, -NAp)), scale(a0, NA0));
    FV[idx] = {T(u.re), T(u.im)};
    FV[M + idx] = {T(v.re), T(v.im)};
    FV[2 * M + idx] = {T(w.re), T(w.im)};
    FV[3 * M + idx] = {T(e.re), T(e.im)};
}


// F- or G-family field spectrum Apm*waveScale + A0*geostrophicScale per
// coefficient (fieldFamilyDerivativeSpectrum, WVCoefficientFormulas.hpp:60-73).
template <typename T>
__global__ __launch_bounds__(256) void k_fgfields(const cx<double>* __restrict__ Apm, const cx<double>* __restrict__ A0,
                                                  const double* __restrict__ fWaveScale,
                                                  const double* __restrict__ gWaveScale, const double* __restrict__ Fg,
                                                  const double* __restrict__ Gg, int Nj, int M, int fFamily,
                                                  cx<T>* __restrict__ FV) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M) return;
    const int j = idx % Nj;
    const double ws = fFamily ? fWaveScale[idx] : gWaveScale[j];
    const double gs = fFamily ? Fg[j] : Gg[j];
    const cx<double> v = cadd(cscale(Apm[idx], ws), cscale(A0[idx], gs));
    FV[idx] = {T(v.re), T(v.im)};
}

// Column stride (complex entries) of the synthesis staging buffer: at least
// Js, and 8 words modulo 64 banks apart (f32: Jp = 4 mod 32; f64: Jp = 2 mod 16).
template <typename T>
__host__ __device__ constexpr int synthJStride(int Js) {
    return sizeof(T) == 4 ? Js + ((4 - Js % 32) % 32 + 32) % 32 : Js + ((2 - Js % 16) % 16 + 16) % 16;
}

// Value of lane l ^ 1 (DPP quad permutation 1,0,3,2).
__device__ __forceinline__ float pairSwap(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double pairSwap(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Stores of one folded synthesis task (rows z = 16 zt + row(kq, i) of the
// tile and their mirrors L - z; z <= L/2 only) for the NQ channel slots of
// group g from slot q0: the value (E + O at z, E - O or O - E at L - z, times
// scale[q]) and i k / i l times it, whose other part the partner lane n ^ 1
// holds.
template <typename T, int NQ>
__device__ __forceinline__ void storeFoldedTask(const SynthGroup& g, int q0, int zt, int kq, int L, int zStride,
                                                size_t chStride, int colOff, bool store, const T (&scale)[NQ], T ks,
                                                T ls, int part, cx<T>* __restrict__ H,
                                                const typename Mfma<T>::acc (&accE)[NQ],
                                                const typename Mfma<T>::acc (&accO)[NQ]) {
    using MF = Mfma<T>;
    const int half = L / 2, mat = g.mat, nch = g.nch;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        if (q0 + q >= nch) continue;
        const SynthChannel& sc = g.ch[q0 + q];
        const int oV = sc.outV, oX = sc.outX, oY = sc.outY;
        // rows z = zr0 + i dzr and L - z of this lane, as byte offsets
        // within a channel (uniform channel bases, 32-bit lane offsets)
        const int zr0 = 16 * zt + MF::row(kq, 0);
        constexpr int dzr = sizeof(T) == 4 ? 1 : 4;  // MF::row(kq, i + 1) - MF::row(kq, i)
        const unsigned offZ = (static_cast<unsigned>(zr0) * zStride + colOff) * sizeof(T);
        const unsigned offL = (static_cast<unsigned>(L - zr0) * zStride + colOff) * sizeof(T);
        const unsigned stepB = static_cast<unsigned>(dzr * zStride) * sizeof(T);
        char* const hb = reinterpret_cast<char*>(H);
        const size_t chBytes = chStride * sizeof(T);
        char* const hV = hb + (oV >= 0 ? oV : 0) * chBytes;
        const bool xy = oX >= 0 || oY >= 0;  // the i k and i l outputs (uniform)
        const T sq = scale[q];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int z = zr0 + i * dzr;
            const T e = accE[q][i], o = accO[q][i];
            const T v1 = (e + o) * sq, v2 = (mat == 0 ? e - o : o - e) * sq;  // at z and at L - z
            const unsigned o1 = offZ + i * stepB, o2 = offL - i * stepB;
            const bool lo = z <= half && store, hi = lo && L - z != z;
            if (oV >= 0) {
                if (lo) *reinterpret_cast<T*>(hV + o1) = v1;
                if (hi) *reinterpret_cast<T*>(hV + o2) = v2;
            }
            if (xy) {
                // partner values with every lane active (a DPP read from an
                // inactive lane returns nothing)
                const T w1 = pairSwap(v1), w2 = pairSwap(v2);
                const T r1 = part ? w1 : -w1;  // part of i * value
                const T r2 = part ? w2 : -w2;
                char* const hX = hb + (oX >= 0 ? oX : 0) * chBytes;
                char* const hY = hb + (oY >= 0 ? oY : 0) * chBytes;
                if (oX >= 0) {
                    if (lo) *reinterpret_cast<T*>(hX + o1) = ks * r1;
                    if (hi) *reinterpret_cast<T*>(hX + o2) = ks * r2;
                }
                if (oY >= 0) {
                    if (lo) *reinterpret_cast<T*>(hY + o1) = ls * r1;
                    if (hi) *reinterpret_cast<T*>(hY + o2) = ls * r2;
                }
            }
        }
    }
}

// Spectral assembly + vertical synthesis (MFMA).  One workgroup owns 8
// consecutive columns: it stages their evolved field spectra (all j < Nj) and
// m_j in LDS once, then its waves loop over (group, z-chunk) tasks, each
// computing ZT tiles of 16 z levels for NQ channel slots of the group:
//   H[out][z][iy][kx] = sum_j syn_fam[z][j] * value_ch(j, column).
// Complex values are two real GEMM columns; lane l owns GEMM column l%16
// (column (l%16)/2, real or imaginary part) and k-row l/16.  The contraction
// length Js is a multiple of 16, so the step loop runs in whole batches with
// no per-step conditions; channel slots beyond the group's channels and z
// tiles beyond the last carry zero coefficients or are not stored.
//
// FOLD (constant N, the DCT-I / DST-I of Nz - 1 = L intervals): row L - z of
// those matrices is row z times (-1)^j (DCT-I) or -(-1)^j (DST-I), so with
// E_z and O_z the sums over even and odd j the outputs are E + O at z and
// E - O (DCT-I) or O - E (DST-I) at L - z, for z <= L/2 only: half the MFMA
// work.  mats then holds the even-j parts of the two matrices (rows z <=
// L/2, Zs of them) in m[0], m[1], the odd-j parts in m[2], m[3], and the
// same with column j times +m_j / 2 (DCT-I) or -m_j / 2 (DST-I), the
// operators of the z derivatives, in m[4] ... m[7].  A task is one z tile of
// NQ channel slots of one group (ZT is 1), all values or all z derivatives,
// so that the MFMA B operand is the staged spectrum itself.
template <typename T, int ZT, int NQ, int KS, bool FOLD>
__global__ __launch_bounds__(512) void k_synth(const cx<T>* __restrict__ FV, const ColumnDev* __restrict__ cols,
                                               int ncol, int M, SynthMats<T> mats, int Zs, int zmid, int Js,
                                               int Nz, int Nj, int Ny, int nkxP, const T* __restrict__ mj,
                                               SynthPlan plan, cx<T>* __restrict__ H) {
    using MF = Mfma<T>;
    extern __shared__ __align__(16) unsigned char smem[];
    // [4 fields][8 columns][Jp]: the staging stores run along j, and the
    // column stride puts the 8 columns x 4 rows x 2 parts that a wave reads
    // in distinct banks
    const int Jp = synthJStride<T>(Js);
    cx<T>* fv = reinterpret_cast<cx<T>*>(smem);
    T* mjs = reinterpret_cast<T*>(fv + 4 * 8 * Jp);  // [Js]
    __shared__ ColumnDev scol[8];
    const int lane = threadIdx.x & 63;
    const int waveInBlock = waveIndex();
    const int nwaves = blockDim.x >> 6;
    const int cg = xcdBlock(blockIdx.x, gridDim.x);
    const int ztiles = Zs / 16;
    const int nzc = (ztiles + ZT - 1) / ZT;
    const int nsteps = FOLD ? Js / 8 : Js / 4;  // folded: Js/2 even (odd) j per sum
    if (threadIdx.x < 8) {
        const int col = cg * 8 + threadIdx.x;
        if (col < ncol) scol[threadIdx.x] = cols[col];
        else scol[threadIdx.x] = ColumnDev{0, 0, -1, 0, 0.0, 0.0, 0.0, 0.0, 0.0};
    }
    // m_j for the z-derivative channels (Js <= 4 blockDim.x; the folded
    // operators carry m_j themselves)
    T mv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = threadIdx.x + k * blockDim.x;
        mv[k] = FOLD ? T(0) : mj[j < Nj ? j : Nj - 1];
    }
    __syncthreads();
    // stage spectra: wave w copies rows (field, column) r = w, w + nwaves, ...
    // (kR rows at a time, all loads issued before the stores), its lanes
    // walking j; rows of absent columns and entries j >= Nj are zeros.  The
    // loads are unconditional (entries j >= Nj read entry Nj - 1, one line
    // for the whole wave): a load under a branch would make every load wait
    // for the ones before it.
    {
        constexpr int kR = 4;    // rows per batch
        constexpr int kH = 4;    // 64-entry chunks of a row (Js <= 256)
        for (int r0 = waveInBlock; r0 < 32; r0 += kR * nwaves) {
            cx<T> v[kR][kH];
#pragma unroll
            for (int i = 0; i < kR; ++i) {
                const int r = r0 + i * nwaves;
                const int rc = r < 32 ? r : 31;
                const int wv = scol[rc & 7].wv;
                const cx<T>* src = FV + static_cast<size_t>(rc >> 3) * M + static_cast<size_t>(Nj) * (wv >= 0 ? wv : 0);
#pragma unroll
                for (int h = 0; h < kH; ++h) {
                    const int j = lane + 64 * h;
                    v[i][h] = src[j < Nj ? j : Nj - 1];
                }
            }
            __builtin_amdgcn_sched_barrier(0);  // all loads issued before the first store
#pragma unroll
            for (int i = 0; i < kR; ++i) {
                const int r = r0 + i * nwaves;
                if (r >= 32) continue;
                const bool live = scol[r & 7].wv >= 0;
#pragma unroll
                for (int h = 0; h < kH; ++h) {
                    const int j = lane + 64 * h;
                    if (j < Js) fv[r * Jp + j] = (live && j < Nj) ? v[i][h] : cx<T>{T(0), T(0)};
                }
            }
        }
    }
    if constexpr (!FOLD) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int j = threadIdx.x + k * blockDim.x;
            if (j < Js) mjs[j] = j < Nj ? mv[k] : T(0);
        }
    }
    __syncthreads();
    const int n = lane & 15, kq = lane >> 4;
    const int c = n >> 1, part = n & 1;
    const ColumnDev cd = scol[c];
    const bool colValid = cd.wv >= 0;
    // Stored value = conj(mode value) on conjugated rows, real on
    // self-conjugate rows; the storage row's wavenumbers are the mode's times
    // the same sign.
    const T rowSign = (cd.flags & kColConj) ? T(-1) : T(1);
    const T partWeight = part ? ((cd.flags & kColSelfConj) ? T(0) : rowSign) : T(1);
    const T ks = rowSign * T(cd.k), ls = rowSign * T(cd.l);
    const bool store = colValid && !(plan.skipCompletion && (cd.flags & kColCompletion));
    const int colOff = 2 * (cd.iy * nkxP + cd.kx) + part;  // word offset in a z plane
    const int zStride = 2 * Ny * nkxP;                      // words per z plane
    const size_t chStride = static_cast<size_t>(Nz) * zStride;
    // lane reads the real (part 0) or imaginary (part 1) word of fv[.][c][j]
    const T* fvw = reinterpret_cast<const T*>(fv) + 2 * c * Jp + part;
    if constexpr (FOLD) {
        const int L = Nz - 1;
        // channel slots of a group in parts of NQ
        int maxch = 0;
        for (int g = 0; g < plan.ngroups; ++g) maxch = max(maxch, plan.g[g].nch);
        const int parts = (maxch + NQ - 1) / NQ;
        const int ntasks = plan.ngroups * parts * ztiles;
        // operand rows of task tk's even-j and odd-j matrices (lane offset
        // included); a part's channels are all values or all z derivatives
        auto matRows = [&](int tk, const T*& pE, const T*& pO) {
            const int g = tk % plan.ngroups;
            const int zt = tk / (plan.ngroups * parts);
            const int m = plan.g[g].mat + (plan.g[g].ch[((tk / plan.ngroups) % parts) * NQ].dz ? 4 : 0);
            pE = mats.m[m] + static_cast<size_t>(zt) * nsteps * 64 + lane;
            pO = mats.m[m + 2] + static_cast<size_t>(zt) * nsteps * 64 + lane;
        };
        constexpr int kS = KS;
        // A operands in register batches of kS steps, two named batches used
        // in turn; the first batch of a wave's next task is loaded before the
        // current task's stores, so that its wait does not include them (one
        // counter covers a wave's loads and stores in issue order).
        T e0[kS], o0[kS], e1[kS], o1[kS];
        {
            const T *pE, *pO;
            matRows(waveInBlock < ntasks ? waveInBlock : ntasks - 1, pE, pO);
#pragma unroll
            for (int u = 0; u < kS; ++u) {
                e0[u] = pE[u * 64];
                o0[u] = pO[u * 64];
            }
        }
        for (int task = waveInBlock; task < ntasks; task += nwaves) {
            const int gi = task % plan.ngroups;
            const int pt = (task / plan.ngroups) % parts;
            const int zt = task / (plan.ngroups * parts);
            const int nch = plan.g[gi].nch;
            const int q0 = pt * NQ;  // first channel of this part
            // B is the staged spectrum itself; the channel's factor (the value
            // scale, or 1 for z derivatives, whose -/+ m_j / 2 is in the
            // operator) and the column's part weight scale the outputs
            const bool dzPart = plan.g[gi].ch[q0].dz != 0;
            const T oscale = (dzPart ? T(1) : T(plan.valueScale)) * partWeight;
            int fieldOff[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const bool live = q0 + q < nch;
                fieldOff[q] = live ? plan.g[gi].ch[q0 + q].field * 16 * Jp : 0;
            }
            const T *AE, *AO;
            matRows(task, AE, AO);
            typename MF::acc accE[NQ], accO[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) accE[q] = accO[q] = MF::zero();
            auto load = [&](T (&dE)[kS], T (&dO)[kS], int s0) {
#pragma unroll
                for (int u = 0; u < kS; ++u) {
                    dE[u] = AE[(s0 + u) * 64];
                    dO[u] = AO[(s0 + u) * 64];
                }
            };
            auto run = [&](const T (&aE)[kS], const T (&aO)[kS], int s0) {
                T bE[kS][NQ], bO[kS][NQ];
#pragma unroll
                for (int u = 0; u < kS; ++u) {
                    const int je = 2 * (4 * (s0 + u) + kq);  // even j; the odd one is je + 1
                    const T* fj = fvw + 2 * je;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        bE[u][q] = fj[fieldOff[q]];
                        bO[u][q] = fj[fieldOff[q] + 2];
                    }
                }
#pragma unroll
                for (int u = 0; u < kS; ++u)
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        accE[q] = MF::run(aE[u], bE[u][q], accE[q]);
                        accO[q] = MF::run(aO[u], bO[u][q], accO[q]);
                    }
            };
            // (the scheduling barriers keep each batch's loads a whole batch
            // of MFMA work ahead of their use; left alone, the compiler moves
            // them down to just before it)
            int s0 = 0;
            for (; s0 + 2 * kS < nsteps; s0 += 2 * kS) {
                load(e1, o1, s0 + kS);
                __builtin_amdgcn_sched_barrier(0);
                run(e0, o0, s0);
                load(e0, o0, s0 + 2 * kS);
                __builtin_amdgcn_sched_barrier(0);
                run(e1, o1, s0 + kS);
            }
            load(e1, o1, s0 + kS);
            __builtin_amdgcn_sched_barrier(0);
            run(e0, o0, s0);
            {
                // the next task's first batch (the last task reloads its own)
                const int next = task + nwaves < ntasks ? task + nwaves : task;
                const T *pE, *pO;
                matRows(next, pE, pO);
#pragma unroll
                for (int u = 0; u < kS; ++u) {
                    e0[u] = pE[u * 64];
                    o0[u] = pO[u * 64];
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            run(e1, o1, s0 + kS);
            if (zmid >= 0 && zt == 0) {
                // The midpoint row z = zmid = L/2, alone in the tile after the
                // last one contracted on MFMA (Nz = 16 m + 1): a dot product
                // over j per lane (its k-row group kq), summed over the four
                // groups, with that tile's row 0 as coefficients (lane 16 kq
                // of each step).  At the midpoint z = L - z: one value, E + O.
                const T* cE = AE + static_cast<size_t>(ztiles - zt) * nsteps * 64 + 16 * kq - lane;
                const T* cO = AO + static_cast<size_t>(ztiles - zt) * nsteps * 64 + 16 * kq - lane;
                T mid[NQ];
#pragma unroll
                for (int q = 0; q < NQ; ++q) mid[q] = T(0);
                for (int s = 0; s < nsteps; s += 4) {  // nsteps is a multiple of 4
                    T ce[4], co[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        ce[u] = cE[(s + u) * 64];
                        co[u] = cO[(s + u) * 64];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const T* fj = fvw + 4 * (4 * (s + u) + kq);
#pragma unroll
                        for (int q = 0; q < NQ; ++q) mid[q] += ce[u] * fj[fieldOff[q]] + co[u] * fj[fieldOff[q] + 2];
                    }
                }
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    mid[q] += __shfl_xor(mid[q], 16);
                    mid[q] += __shfl_xor(mid[q], 32);
                    // i k and i l times the value: the partner lane n ^ 1
                    // holds the other part (read with every lane active)
                    const T v = mid[q] * oscale;
                    const T other = pairSwap(v);
                    if (q0 + q >= nch) continue;
                    const SynthChannel& sc = plan.g[gi].ch[q0 + q];
                    const size_t o = static_cast<size_t>(zmid) * zStride + colOff;
                    const T rot = part ? other : -other;
                    if (kq == 0 && store) {
                        if (sc.outV >= 0) reinterpret_cast<T*>(H)[sc.outV * chStride + o] = v;
                        if (sc.outX >= 0) reinterpret_cast<T*>(H)[sc.outX * chStride + o] = ks * rot;
                        if (sc.outY >= 0) reinterpret_cast<T*>(H)[sc.outY * chStride + o] = ls * rot;
                    }
                }
            }
            T scale[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) scale[q] = oscale;
            storeFoldedTask<T, NQ>(plan.g[gi], q0, zt, kq, L, zStride, chStride, colOff, store, scale, ks, ls, part, H,
                                   accE, accO);
        }
        return;
    }
    for (int task = waveInBlock; task < plan.ngroups * nzc; task += nwaves) {
        const int gi = task % plan.ngroups;
        const int zc = task / plan.ngroups;
        const int mat = plan.g[gi].mat, nch = plan.g[gi].nch;
        int fieldOff[NQ];
        T coefV[NQ], coefZ[NQ];
        // value channels: valueScale; spectral z-derivative channels: -/+ m_j / 2
        const T zsign = mat == 1 ? T(-0.5) : T(0.5);
        const T vscale = T(plan.valueScale);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const bool live = q < nch;
            const bool dz = live && plan.g[gi].ch[q].dz;
            fieldOff[q] = live ? plan.g[gi].ch[q].field * 16 * Jp : 0;
            coefV[q] = live && !dz ? vscale * partWeight : T(0);
            coefZ[q] = dz ? zsign * partWeight : T(0);
        }
        const T* A = mats.m[mat];
        const int zt0 = zc * ZT;
        typename MF::acc acc[NQ][ZT];
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int r = 0; r < ZT; ++r) acc[q][r] = MF::zero();
        // A operands in register batches of kS steps: two named batches used in
        // turn (no register copies, so each wait covers only the batch in use);
        // the step count is a multiple of 2 kS
        constexpr int kS = KS;
        const T* Ar[ZT];
#pragma unroll
        for (int r = 0; r < ZT; ++r) {
            const int zt = zt0 + r < ztiles ? zt0 + r : ztiles - 1;
            Ar[r] = A + static_cast<size_t>(zt) * nsteps * 64 + lane;
        }
        auto load = [&](T (&dst)[ZT][kS], int s0) {
#pragma unroll
            for (int r = 0; r < ZT; ++r)
#pragma unroll
                for (int u = 0; u < kS; ++u) dst[r][u] = Ar[r][(s0 + u) * 64];
        };
        auto run = [&](const T (&a)[ZT][kS], int s0) {
            T b[kS][NQ];
#pragma unroll
            for (int u = 0; u < kS; ++u) {
                const int j = 4 * (s0 + u) + kq;
                const T mjj = mjs[j];
                const T* fj = fvw + 2 * j;
#pragma unroll
                for (int q = 0; q < NQ; ++q) b[u][q] = fj[fieldOff[q]] * (coefV[q] + coefZ[q] * mjj);
            }
#pragma unroll
            for (int u = 0; u < kS; ++u)
#pragma unroll
                for (int q = 0; q < NQ; ++q)
#pragma unroll
                    for (int r = 0; r < ZT; ++r) acc[q][r] = MF::run(a[r][u], b[u][q], acc[q][r]);
        };
        // (scheduling barriers keep each load a whole batch ahead of its use)
        T a0[ZT][kS], a1[ZT][kS];
        load(a0, 0);
        int s0 = 0;
        for (; s0 + 2 * kS < nsteps; s0 += 2 * kS) {
            load(a1, s0 + kS);
            __builtin_amdgcn_sched_barrier(0);
            run(a0, s0);
            load(a0, s0 + 2 * kS);
            __builtin_amdgcn_sched_barrier(0);
            run(a1, s0 + kS);
        }
        load(a1, s0 + kS);
        __builtin_amdgcn_sched_barrier(0);
        run(a0, s0);
        run(a1, s0 + kS);
        // outputs: the value and i k / i l times it (the partner lane n^1
        // holds the other part of the same complex entry)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (q >= nch) continue;
            const int oV = plan.g[gi].ch[q].outV, oX = plan.g[gi].ch[q].outX, oY = plan.g[gi].ch[q].outY;
            T* hV = reinterpret_cast<T*>(H) + (oV >= 0 ? oV : 0) * chStride + colOff;
            T* hX = reinterpret_cast<T*>(H) + (oX >= 0 ? oX : 0) * chStride + colOff;
            T* hY = reinterpret_cast<T*>(H) + (oY >= 0 ? oY : 0) * chStride + colOff;
#pragma unroll
            for (int r = 0; r < ZT; ++r) {
                const int zt = zt0 + r;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const T v = acc[q][r][i];
                    const T other = pairSwap(v);
                    const T rot = part ? other : -other;  // part of i * (re + i im)
                    const int z = 16 * zt + MF::row(kq, i);
                    if (zt < ztiles && z < Nz && store) {
                        const int o = z * zStride;
                        if (oV >= 0) hV[o] = v;
                        if (oX >= 0) hX[o] = ks * rot;
                        if (oY >= 0) hY[o] = ls * rot;
                    }
                }
            }
        }
    }
}

// The folded synthesis with split-FP16 MFMA products (the mixed-precision
// form with compensated half-precision vertical GEMMs): every operand x is
// scaled by a power of two s and held as hi = fp16(x s) and lo = fp16(x s -
// hi), so x s = hi + lo to 22 bits, and each product is summed in FP32 as
// lo_a hi_b + hi_a lo_b + hi_a hi_b (the lo_a lo_b term, 2^-22 of the
// product, is dropped).  The operators (SynthMats16, prepared on the host)
// have one scale per matrix pair; the staged spectra have one per (field,
// column), chosen from the column's largest entry so that it lies in
// [2^14, 2^15): every hi part is a normal FP16 number and the columns of
// small modes keep their relative precision.  The outputs are scaled back
// exactly.  Same tasks, tiles and outputs as the folded k_synth, with the
// midpoint row z = L/2 in the last MFMA tile (no separate dot product).
// K is the folded contraction length (Jh = Js / 2 even- or odd-j entries)
// in N32 steps of 32 (v_mfma_f32_16x16x32_f16: lane l holds k = 8 (l/16) +
// t, t < 8) and TAIL steps of 16 (v_mfma_f32_16x16x16_f16: k = 4 (l/16) + t,
// t < 4).
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
struct SynthMats16 {
    const _Float16* m[kSynthMatSlots];  // per matrix: [tile][N32 steps x (hi, lo) x 64 x 8 | TAIL x (hi, lo) x 64 x 4]
    float inv[kSynthMatSlots];          // 1 / (the matrix's scale)
};
// Power-of-two scale that puts a nonzero magnitude m in [2^14, 2^15), as an
// exponent clamped so that the scale stays a normal float.
__host__ __device__ inline int split16Exponent(float m) {
    if (!(m > 0.0f)) return 0;
    int ex;
    frexpf(m, &ex);  // m = f 2^ex, f in [0.5, 1)
    const int e = 15 - ex;
    return e > 100 ? 100 : e < -100 ? -100 : e;
}
__device__ __forceinline__ void splitHalf(float v, _Float16& hi, _Float16& lo) {
    hi = static_cast<_Float16>(v);
    lo = static_cast<_Float16>(v - static_cast<float>(hi));
}

template <int NQ, int N32, int TAIL>
__global__ __launch_bounds__(256) void k_synth16(const cx<float>* __restrict__ FV, const ColumnDev* __restrict__ cols,
                                                 int ncol, int M, SynthMats16 mats, int Js, int Nz, int Nj, int Ny,
                                                 int nkxP, int ztiles, SynthPlan plan, cx<float>* __restrict__ H) {
    using MF = Mfma<float>;
    constexpr int NC8 = 4 * N32 + 2 * TAIL;          // 8-entry chunks of the folded contraction
    constexpr int kTileHalves = N32 * 2 * 64 * 8 + TAIL * 2 * 64 * 4;
    // staged spectra: half8 [field 4][parity 2][hl 2][chunk NC8][column c 8 x part 2]
    extern __shared__ __align__(16) unsigned char smem[];
    half8* sb = reinterpret_cast<half8*>(smem);
    __shared__ ColumnDev scol[8];
    __shared__ unsigned rowMax[32];  // per (field, column): largest |entry| as float bits
    __shared__ float rowInv[32];     // 1 / (its scale)
    const int lane = threadIdx.x & 63;
    const int waveInBlock = waveIndex();
    const int nwaves = blockDim.x >> 6;
    const int cg = xcdBlock(blockIdx.x, gridDim.x);
    if (threadIdx.x < 8) {
        const int col = cg * 8 + threadIdx.x;
        if (col < ncol) scol[threadIdx.x] = cols[col];
        else scol[threadIdx.x] = ColumnDev{0, 0, -1, 0, 0.0, 0.0, 0.0, 0.0, 0.0};
    }
    if (threadIdx.x < 32) rowMax[threadIdx.x] = 0u;
    __syncthreads();
    // staging: item (field, chunk g, column c, part) takes one part of the
    // entries j = 16 g ... 16 g + 15 of the column (even and odd j of chunk
    // g); chunks past Js are zeros.  Sixteen consecutive items write the 16
    // column-part slots of one 256-byte row of the staged layout (no bank
    // conflicts).  The values stay in registers across the barrier that
    // completes the row maxima.
    constexpr int kItems = 4 * NC8 * 16;
    constexpr int kItemsPerThread = (kItems + 255) / 256;
    float v[kItemsPerThread][16];
#pragma unroll
    for (int it = 0; it < kItemsPerThread; ++it) {
        const int item = threadIdx.x + 256 * it;
        const int part = item & 1, c = (item >> 1) & 7, g = (item >> 4) % NC8, field = (item >> 4) / NC8;
        if (field < 4) {
            const int wv = scol[c].wv;
            const cx<float>* src = FV + static_cast<size_t>(field) * M + static_cast<size_t>(Nj) * (wv >= 0 ? wv : 0);
            float mx = 0.0f;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int j = 16 * g + t;
                const cx<float> x = src[j < Nj ? j : Nj - 1];
                v[it][t] = (wv >= 0 && j < Nj && j < Js) ? (part ? x.im : x.re) : 0.0f;
                mx = fmaxf(mx, fabsf(v[it][t]));
            }
            atomicMax(&rowMax[field * 8 + c], __float_as_uint(mx));
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kItemsPerThread; ++it) {
        const int item = threadIdx.x + 256 * it;
        const int part = item & 1, c = (item >> 1) & 7, g = (item >> 4) % NC8, field = (item >> 4) / NC8;
        if (field >= 4) continue;
        const int r = field * 8 + c;
        const int e = split16Exponent(__uint_as_float(rowMax[r]));
        const float sc = ldexpf(1.0f, e);
        if (g == 0 && part == 0) rowInv[r] = ldexpf(1.0f, -e);
#pragma unroll
        for (int par = 0; par < 2; ++par) {
            half8 hi, lo;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                _Float16 h, l;
                splitHalf(v[it][2 * t + par] * sc, h, l);
                hi[t] = h;
                lo[t] = l;
            }
            const int base = ((field * 2 + par) * 2) * NC8 + g;
            sb[base * 16 + 2 * c + part] = hi;
            sb[(base + NC8) * 16 + 2 * c + part] = lo;
        }
    }
    __syncthreads();
    const int n = lane & 15, kq = lane >> 4;
    const int c = n >> 1, part = n & 1;
    const ColumnDev cd = scol[c];
    const bool colValid = cd.wv >= 0;
    const float rowSign = (cd.flags & kColConj) ? -1.0f : 1.0f;
    const float partWeight = part ? ((cd.flags & kColSelfConj) ? 0.0f : rowSign) : 1.0f;
    const float ks = rowSign * static_cast<float>(cd.k), ls = rowSign * static_cast<float>(cd.l);
    const bool store = colValid && !(plan.skipCompletion && (cd.flags & kColCompletion));
    const int colOff = 2 * (cd.iy * nkxP + cd.kx) + part;
    const int zStride = 2 * Ny * nkxP;
    const size_t chStride = static_cast<size_t>(Nz) * zStride;
    const int L = Nz - 1;
    int maxch = 0;
    for (int gg = 0; gg < plan.ngroups; ++gg) maxch = max(maxch, plan.g[gg].nch);
    const int parts = (maxch + NQ - 1) / NQ;
    const int ntasks = plan.ngroups * parts * ztiles;
    for (int task = waveInBlock; task < ntasks; task += nwaves) {
        const int gi = task % plan.ngroups;
        const int pt = (task / plan.ngroups) % parts;
        const int zt = task / (plan.ngroups * parts);
        const SynthGroup& grp = plan.g[gi];
        const int q0 = pt * NQ;
        const bool dzPart = grp.ch[q0].dz != 0;
        const int m = grp.mat + (dzPart ? 4 : 0);
        const _Float16* AE = mats.m[m] + static_cast<size_t>(zt) * kTileHalves;
        const _Float16* AO = mats.m[m + 2] + static_cast<size_t>(zt) * kTileHalves;
        int field[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) field[q] = q0 + q < grp.nch ? grp.ch[q0 + q].field : 0;
        typename MF::acc accE[NQ], accO[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) accE[q] = accO[q] = MF::zero();
#pragma unroll
        for (int s = 0; s < N32; ++s) {
            const half8 aEh = reinterpret_cast<const half8*>(AE + (2 * s) * 512)[lane];
            const half8 aEl = reinterpret_cast<const half8*>(AE + (2 * s + 1) * 512)[lane];
            const half8 aOh = reinterpret_cast<const half8*>(AO + (2 * s) * 512)[lane];
            const half8 aOl = reinterpret_cast<const half8*>(AO + (2 * s + 1) * 512)[lane];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int bE = ((field[q] * 2 + 0) * 2) * NC8 + 4 * s + kq, bO = ((field[q] * 2 + 1) * 2) * NC8 + 4 * s + kq;
                const half8 bEh = sb[bE * 16 + n], bEl = sb[(bE + NC8) * 16 + n];
                const half8 bOh = sb[bO * 16 + n], bOl = sb[(bO + NC8) * 16 + n];
                accE[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aEl, bEh, accE[q], 0, 0, 0);
                accE[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aEh, bEl, accE[q], 0, 0, 0);
                accE[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aEh, bEh, accE[q], 0, 0, 0);
                accO[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aOl, bOh, accO[q], 0, 0, 0);
                accO[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aOh, bOl, accO[q], 0, 0, 0);
                accO[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aOh, bOh, accO[q], 0, 0, 0);
            }
        }
        if constexpr (TAIL) {
            // k = 32 N32 + 4 (l/16) + t: chunk 4 N32 + kq/2, entries 4 (kq & 1) ...
            const _Float16* tE = AE + N32 * 2 * 512;
            const _Float16* tO = AO + N32 * 2 * 512;
            const half4 aEh = reinterpret_cast<const half4*>(tE)[lane];
            const half4 aEl = reinterpret_cast<const half4*>(tE + 256)[lane];
            const half4 aOh = reinterpret_cast<const half4*>(tO)[lane];
            const half4 aOl = reinterpret_cast<const half4*>(tO + 256)[lane];
            const int ch8 = 4 * N32 + (kq >> 1), sub = 4 * (kq & 1);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int bE = ((field[q] * 2 + 0) * 2) * NC8 + ch8, bO = ((field[q] * 2 + 1) * 2) * NC8 + ch8;
                const _Float16* pEh = reinterpret_cast<const _Float16*>(sb + bE * 16 + n) + sub;
                const _Float16* pEl = reinterpret_cast<const _Float16*>(sb + (bE + NC8) * 16 + n) + sub;
                const _Float16* pOh = reinterpret_cast<const _Float16*>(sb + bO * 16 + n) + sub;
                const _Float16* pOl = reinterpret_cast<const _Float16*>(sb + (bO + NC8) * 16 + n) + sub;
                const half4 bEh = *reinterpret_cast<const half4*>(pEh), bEl = *reinterpret_cast<const half4*>(pEl);
                const half4 bOh = *reinterpret_cast<const half4*>(pOh), bOl = *reinterpret_cast<const half4*>(pOl);
                accE[q] = __builtin_amdgcn_mfma_f32_16x16x16f16(aEl, bEh, accE[q], 0, 0, 0);
                accE[q] = __builtin_amdgcn_mfma_f32_16x16x16f16(aEh, bEl, accE[q], 0, 0, 0);
                accE[q] = __builtin_amdgcn_mfma_f32_16x16x16f16(aEh, bEh, accE[q], 0, 0, 0);
                accO[q] = __builtin_amdgcn_mfma_f32_16x16x16f16(aOl, bOh, accO[q], 0, 0, 0);
                accO[q] = __builtin_amdgcn_mfma_f32_16x16x16f16(aOh, bOl, accO[q], 0, 0, 0);
                accO[q] = __builtin_amdgcn_mfma_f32_16x16x16f16(aOh, bOh, accO[q], 0, 0, 0);
            }
        }
        // the channel's factor (value scale, or 1 for z derivatives), the
        // column's part weight, and the operand scales undone
        const float oscale = (dzPart ? 1.0f : static_cast<float>(plan.valueScale)) * partWeight * mats.inv[m];
        float scale[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) scale[q] = oscale * rowInv[field[q] * 8 + c];
        storeFoldedTask<float, NQ>(grp, q0, zt, kq, L, zStride, chStride, colOff, store, scale, ks, ls, part, H, accE,
                                   accO);
    }
}

// Complex FFT along y of [Ny, KT] tiles: in/out layout [ch][z][iy][nkxP].
// rowLimit[tile] (if given) is the largest |ky| retained in the kx tile: input
// rows r < iy < Ny - r are known zeros (skipIn) or output rows there are not
// needed (skipOut).  L2N > 0 fixes Ny = 2^L2N and the tile width 2^L2KT at
// compile time (the runtime Ny, log2Ny and log2KT are then ignored);
// L2N <= 0 takes runtime lengths (kSmooth7, kSmooth7Rounds: prime factors
// at most 7).
// Channel plan of the inverse y transform with y derivatives: block z takes
// input channel src[z] to output channel src[z] and, when der[z] >= 0, also
// i l times it to output channel der[z] (l = lrow[iy], the y wavenumber of
// storage row iy), from the same loaded tile.
struct YDerivPlan {
    int src[16];
    int der[16];
};

template <typename T, int L2N, int L2KT, bool DERIV>
__global__ __launch_bounds__(kFftThreads) void k_yfft(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                                                      int NyArg, int log2NyArg, int nkx, int nkxP, int log2KTArg,
                                                      int Nz, const cx<T>* __restrict__ twY, int sign,
                                                      const int* __restrict__ rowLimit, int skipIn, int skipOut,
                                                      YDerivPlan yp, const T* __restrict__ lrow, int lineRemap) {
    const int Ny = L2N > 0 ? 1 << (L2N > 0 ? L2N : 0) : NyArg;
    const int log2Ny = L2N > 0 ? L2N : log2NyArg;
    const int log2KT = L2N > 0 ? L2KT : log2KTArg;
    const int KT = 1 << log2KT;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem);
    cx<T>* tile = tw + Ny;
    // lineRemap (tile rows under 64 bytes, e.g. FP32 at Ny = 1024): neighbouring
    // tiles, which share lines, are numbered onto one XCD and its L2
    // (measured at 1024 x 1024: inverse 5.7 -> 2.2 ms; at 64-byte rows it
    // cost 10 %)
    int bx = static_cast<int>(blockIdx.x), by = static_cast<int>(blockIdx.y), bz = static_cast<int>(blockIdx.z);
    if (lineRemap) {
        const int X = static_cast<int>(gridDim.x), Y = static_cast<int>(gridDim.y);
        const int l = xcdBlock(bx + X * (by + Y * bz), X * Y * static_cast<int>(gridDim.z));
        bx = l % X;
        by = (l / X) % Y;
        bz = l / (X * Y);
    }
    const int kx0 = bx * KT;
    const int lim = rowLimit ? rowLimit[bx] : Ny;
    const int ch = DERIV ? yp.src[bz] : bz;
    const size_t planeStride = static_cast<size_t>(Ny) * nkxP;
    const size_t plane = (static_cast<size_t>(ch) * Nz + by) * planeStride;
    const int nel = Ny * KT;
    loadTwiddles(tw, twY, Ny);
    // all of a thread's loads are issued before its LDS stores; entries per
    // thread: nel / blockDim for a compiled length, else at most 16 (nel <=
    // 16 blockDim, checked at setup)
    constexpr int kMaxPerThread = L2N > 0 ? ((1 << (L2N > 0 ? L2N + L2KT : 0)) + kFftThreads - 1) / kFftThreads : 16;
    cx<T> v[kMaxPerThread];
#pragma unroll
    for (int i = 0; i < kMaxPerThread; ++i) {
        const int e = fftThread() + i * kFftThreads;
        const int ec = e < nel ? e : nel - 1;  // clamped: loads are unconditional
        const int kk = ec & (KT - 1), iy = ec >> log2KT;
        // rows known to be zero read row 0 instead (loaded by other lanes)
        const bool zero = skipIn && iy > lim && iy < Ny - lim;
        v[i] = in[plane + static_cast<size_t>(zero ? 0 : iy) * nkxP + kx0 + kk];
    }
    __builtin_amdgcn_sched_barrier(0);  // all loads issued before the first store
    // (both inlined: a call would move the staged values v to scratch)
    auto fill = [&](bool deriv) __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < kMaxPerThread; ++i) {
            const int e = fftThread() + i * kFftThreads;
            const int iy = e >> log2KT;
            const bool zero = skipIn && iy > lim && iy < Ny - lim;
            cx<T> x = v[i];
            if (deriv) {
                const T l = lrow[iy < Ny ? iy : Ny - 1];
                x = cx<T>{l * -x.im, l * x.re};  // i l x
            }
            if (e < nel) tile[e] = zero ? cx<T>{T(0), T(0)} : x;
        }
    };
    auto transform = [&](size_t outPlane) __attribute__((always_inline)) {
        __syncthreads();
        if (sign > 0)
            fftLds<T, +1, L2N>(tile, Ny, log2Ny, BatchCols{log2KT}, tw);
        else
            fftLds<T, -1, L2N>(tile, Ny, log2Ny, BatchCols{log2KT}, tw);
        // columns kx >= nkx (the tile's padding) are not written
        for (int e = fftThread(); e < nel; e += kFftThreads) {
            const int kk = e & (KT - 1), iy = e >> log2KT;
            if ((skipOut && iy > lim && iy < Ny - lim) || kx0 + kk >= nkx) continue;
            out[outPlane + static_cast<size_t>(iy) * nkxP + kx0 + kk] = tile[e];
        }
    };
    fill(false);
    transform(plane);
    if (DERIV && yp.der[bz] >= 0) {
        __syncthreads();  // the tile is refilled
        fill(true);
        transform((static_cast<size_t>(yp.der[bz]) * Nz + by) * planeStride);
    }
}

// q = e / N, r = e % N; shifts when N = 2^log2N (log2N >= 0).
__device__ __forceinline__ void divmodN(int e, int N, int log2N, int& q, int& r) {
    if (log2N >= 0) {
        q = e >> log2N;
        r = e & (N - 1);
    } else {
        q = e / N;
        r = e - q * N;
    }
}

// Hermitian-extended x spectrum of channel `ch` at row (z, y): the c2r input
// of WVReferenceFFTEngine.cpp:65-86 (imaginary part of kx = 0 ignored).
template <typename T>
__device__ __forceinline__ cx<T> hermitianX(const cx<T>* __restrict__ G, size_t rowBase, size_t chStride, int ch,
                                            int k, int Nx, int nkx) {
    const cx<T>* row = G + ch * chStride + rowBase;
    if (k == 0) return {row[0].re, T(0)};
    if (k < nkx) return row[k];
    if (Nx - k < nkx) return cconj(row[Nx - k]);
    return {T(0), T(0)};
}

// LDS layout of the x transforms: transform stride NxP = the padded length
// (N + N/16, conflict-free radix-16 passes; for f32 it is 16 entries modulo a
// 256-byte bank sweep, so the two transforms a half-wave reads fill disjoint
// banks), row stride PAIRS*NxP.
template <typename T>
__host__ __device__ constexpr int xfusedStride(int Nx) { return ldsLength(Nx); }
template <typename T>
__host__ __device__ constexpr int xfusedRowStride(int Nx, int pairs) { return pairs * xfusedStride<T>(Nx); }
// LDS bytes of a k_xfused block of RB rows (twiddle table + transforms)
template <typename T>
__host__ __device__ constexpr size_t xfusedLdsBytes(int Nx, int pairs, int RB) {
    return (static_cast<size_t>(Nx) + static_cast<size_t>(RB) * xfusedRowStride<T>(Nx, pairs)) * sizeof(cx<T>);
}
// rows per k_xfused block: one.  At Nx = 256 in FP32 a one-row block takes
// 19 KB of LDS, so 8 blocks (32 waves) share a CU; two-row blocks (37 KB, 4
// blocks, all 256 threads busy in the inverse passes) measured 4 % slower.
template <typename T>
__host__ __device__ constexpr int xfusedRows(int, int) {
    return 1;
}

// Spatial channels of the flux products, packed two per complex x transform.
// Pair p takes half-spectrum channel srcA[p] (real part of the transform),
// multiplied by i k when ikA[p] is set, and srcB[p] (imaginary part, -1 =
// none).  A complex FFT's rounding error scales with the larger of its two
// signals, so each pair holds quantities of one kind and size (the reference
// transforms every channel on its own).  Products: F_t = -(U d/dx T +
// V d/dy T + W d/dz T) with the spatial channel numbers (2p + part) of
// (dx, dy, dz) T in prod[t] and of U, V, W in uvw.
// Nonhydrostatic (half-spectrum channels u v w eta, dy u..eta, dz u..eta):
//   (U, V) (dxu, dyu) (dxv, dyv) (dxw, dyw) (dxeta, dyeta) (dzu, dzv) (W, dzw) (dzeta, -)
// Hydrostatic (u v w eta, dy u v eta, dz u v eta):
//   (U, V) (dxu, dyu) (dxv, dyv) (dxeta, dyeta) (dzu, dzv) (W, -) (dzeta, -)
template <int NT>
struct XMap;
template <>
struct XMap<4> {
    static constexpr int targets = 4;
    static constexpr int pairs = 8;
    static constexpr int srcA[8] = {0, 0, 1, 2, 3, 8, 2, 11};
    static constexpr int ikA[8] = {0, 1, 1, 1, 1, 0, 0, 0};
    static constexpr int srcB[8] = {1, 4, 5, 6, 7, 9, 10, -1};
    static constexpr int uvw[3] = {0, 1, 12};
    static constexpr int prod[4][3] = {{2, 3, 10}, {4, 5, 11}, {6, 7, 13}, {8, 9, 14}};
};
template <>
struct XMap<3> {
    static constexpr int targets = 3;
    static constexpr int pairs = 7;
    static constexpr int srcA[7] = {0, 0, 1, 3, 7, 2, 9};
    static constexpr int ikA[7] = {0, 1, 1, 1, 0, 0, 0};
    static constexpr int srcB[7] = {1, 4, 5, 6, 8, -1, -1};
    static constexpr int uvw[3] = {0, 1, 10};
    static constexpr int prod[3][3] = {{2, 3, 8}, {4, 5, 9}, {6, 7, 12}};
};
// Nonhydrostatic with d/dz w from continuity (the mixed-precision flux of the
// constant-N transform): the wave-vortex modes are divergence-free one by one
// (i (k UAp + l VAp) + m WAp = 0, i (k UA0 + l VA0) = 0), so d/dz w =
// -(d/dx u + d/dy v) and the d/dz w channel is neither synthesized nor
// transformed.  Half-spectrum channels u v w eta, dy u..eta, dz u, dz v,
// dz eta:
//   (U, V) (dxu, dyu) (dxv, dyv) (dxw, dyw) (dxeta, dyeta) (dzu, dzv) (W, dzeta)
// prod entry -1: -(dxu + dyv) (spatial channels 2 and 5).
template <>
struct XMap<5> {
    static constexpr int targets = 4;
    static constexpr int pairs = 7;
    static constexpr int srcA[7] = {0, 0, 1, 2, 3, 8, 2};
    static constexpr int ikA[7] = {0, 1, 1, 1, 1, 0, 0};
    static constexpr int srcB[7] = {1, 4, 5, 6, 7, 9, 10};
    static constexpr int uvw[3] = {0, 1, 12};
    static constexpr int prod[4][3] = {{2, 3, 10}, {4, 5, 11}, {6, 7, -1}, {8, 9, 13}};
};

// Per (y, z) row: c2r of the spatial channels along x (two real transforms
// per complex FFT), the NT advection products and r2c of the products.
// kxv[k] = the x wavenumber of half-spectrum column k.  L2N > 0 fixes
// Nx = 2^L2N at compile time (the runtime Nx and log2Nx are then ignored);
// L2N <= 0 takes runtime lengths (kSmooth7, kSmooth7Rounds: prime factors
// at most 7).
// The forward transforms run the same radix-16 passes as the inverse ones:
// at Nx = 256 with two threads per butterfly (fftFwd256, one full wave for
// the two transforms of a row), at other lengths on the first 16 threads per
// transform.
// PRUNE (Nx = 256, nkx <= 96): entries 96 ... 159 of every inverse
// transform lie in the zero band, so inputs r = 6 ... 9 of every butterfly
// of the first radix-16 pass (entries jj + 16 r) are known zeros: neither
// written nor read.
// EXT (nonlinearFluxUsingAdvectionFields, ...Kernel.cpp:1160-1171): the
// advecting U, V, W of the products are the caller's FP64 fields adv
// ([3][Nz][Ny][Nx]) instead of the synthesized ones.
template <typename T, int NT, int RB, int L2N, bool PRUNE, bool EXT = false>
__global__ __launch_bounds__(kFftThreads) void k_xfused(const cx<T>* __restrict__ G, cx<T>* __restrict__ P,
                                                        int NxArg, int log2NxArg, int Ny, int Nz, int nkx, int nkxP,
                                                        const cx<T>* __restrict__ twX, const T* __restrict__ kxv, const cx<T>* __restrict__ zeros,
                                                        T etaScale, const double* __restrict__ adv) {
    using XM = XMap<NT>;
    constexpr int PAIRS = XM::pairs;
    const int Nx = L2N > 0 ? 1 << (L2N > 0 ? L2N : 0) : NxArg;
    const int log2Nx = L2N > 0 ? L2N : log2NxArg;
    extern __shared__ __align__(16) unsigned char smem[];
    const int NxP = xfusedStride<T>(Nx);
    // LDS position of entry n of a transform (the layout of fftRows)
    const bool pad = L2N > 0 || rowsPadded(log2Nx);
    auto xi = [=](int n) { return pad ? ldsIndex(n, 1) : n; };
    const int RS = xfusedRowStride<T>(Nx, PAIRS);
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem);
    cx<T>* buf = tw + Nx;  // [RB][RS], row rr holds PAIRS transforms of stride NxP
    const int nrows = Ny * Nz;
    const int row0 = blockIdx.x * RB;  // row = z*Ny + y
    const size_t chStride = static_cast<size_t>(Nz) * Ny * nkxP;
    loadTwiddles(tw, twX, Nx);
    // Transform t = rr*PAIRS + p packs spatial channels 2p (real part) and
    // 2p+1 (imaginary part) of row rr.  Wave w fills transforms w, w + 4, ...:
    // each retained entry k < nkx is read once and written at k and,
    // conjugated, at Nx - k; the band between is zeroed.  A wave issues the
    // loads of all its transforms (128 entries each) before storing them.
    static_assert(RB * PAIRS <= 16, "at most four transforms per wave");
    constexpr int kTPW = (RB * PAIRS + 3) / 4;  // transforms per wave
    const int lane = threadIdx.x & 63, wave = waveIndex();
    for (int kc = 0; kc < nkx; kc += 128) {
        cx<T> va[kTPW][2], vb[kTPW][2];
#pragma unroll
        for (int i = 0; i < kTPW; ++i) {
            const int t = wave + 4 * i;
            const int rr = t / PAIRS, p = t % PAIRS;
            const int row = row0 + rr;
            const int sa = XM::srcA[p], sb = XM::srcB[p];
            // a missing channel or row reads the zero row, so the values need
            // no select; addresses are clamped into the arrays so that every
            // load is issued unconditionally (no branch, no wait between them)
            const bool live = t < RB * PAIRS && row < nrows;
            const size_t rowOff = static_cast<size_t>(row) * nkxP;
            const cx<T>* A = live ? G + sa * chStride + rowOff : zeros;
            const cx<T>* B = live && sb >= 0 ? G + sb * chStride + rowOff : zeros;
            // (32-bit byte offsets from the uniform bases: no 64-bit
            // address arithmetic per lane)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = kc + lane + 64 * h;
                const unsigned off = static_cast<unsigned>(k < nkx ? k : nkx - 1) * sizeof(cx<T>);
                va[i][h] = *reinterpret_cast<const cx<T>*>(reinterpret_cast<const char*>(A) + off);
                vb[i][h] = *reinterpret_cast<const cx<T>*>(reinterpret_cast<const char*>(B) + off);
            }
        }
        T kk[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = kc + lane + 64 * h;
            const unsigned off = static_cast<unsigned>(k < nkx ? k : nkx - 1) * sizeof(T);
            kk[h] = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(kxv) + off);
        }
        __builtin_amdgcn_sched_barrier(0);  // all loads issued before the first store
#pragma unroll
        for (int i = 0; i < kTPW; ++i) {
            const int t = wave + 4 * i;
            if (t >= RB * PAIRS) continue;
            const int p = t % PAIRS;
            const bool ik = XM::ikA[p] != 0;
            cx<T>* z = buf + (t / PAIRS) * RS + p * NxP;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = kc + lane + 64 * h;
                if (k >= nkx) continue;
                cx<T> a = va[i][h];
                const cx<T> b = vb[i][h];
                if (ik) a = cmuli(a, kk[h]);
                // k = 0: real parts only (the imaginary part of the kx = 0
                // column is ignored by the c2r transform)
                const T aim = (h == 0 && k == 0) ? T(0) : a.im, bim = (h == 0 && k == 0) ? T(0) : b.im;
                z[xi(k)] = {a.re - bim, aim + b.re};
                if (k > 0) z[xi(Nx - k)] = {a.re + bim, b.re - aim};
            }
        }
    }
    static_assert(!PRUNE || L2N == 8, "pruned first pass for Nx = 256");
    for (int t = wave; t < RB * PAIRS; t += 4) {
        cx<T>* z = buf + (t / PAIRS) * RS + (t % PAIRS) * NxP;
        if constexpr (PRUNE) {
            // the band entries that the first pass reads: nkx ... 95, 160 ... Nx - nkx
            const int lo = 96 - nkx;
            for (int e = lane; e < lo + (Nx - nkx - 159); e += 64) z[ldsIndex(e < lo ? nkx + e : 160 + e - lo, 1)] = cx<T>{T(0), T(0)};
        } else {
            for (int k = nkx + lane; k <= Nx - nkx; k += 64) z[xi(k)] = cx<T>{T(0), T(0)};
        }
    }
    __syncthreads();
    if constexpr (PRUNE) {
        const BatchRows<PAIRS> bt{RB * PAIRS, NxP, RS};
        stockhamPass<T, 16, +1, BatchRows<PAIRS>, 0x3C0u>(buf, 8, 0, bt, tw);
        stockhamPass<T, 16, +1>(buf, 8, 4, bt, tw);
    } else {
        fftRows<T, +1, L2N, PAIRS>(buf, Nx, log2Nx, RB * PAIRS, NxP, RS, tw);
    }
    for (int e = fftThread(); e < RB * Nx; e += kFftThreads) {
        int rr, x;
        if constexpr (RB == 1) {
            rr = 0;
            x = e;
        } else {
            divmodN(e, Nx, log2Nx, rr, x);
        }
        const int xl = xi(x);
        cx<T>* rb = buf + rr * RS;
        T ch[2 * PAIRS];
#pragma unroll
        for (int p = 0; p < PAIRS; ++p) {
            const cx<T> v = rb[p * NxP + xl];
            ch[2 * p] = v.re;
            ch[2 * p + 1] = v.im;
        }
        T U = ch[XM::uvw[0]], V = ch[XM::uvw[1]], W = ch[XM::uvw[2]];
        if constexpr (EXT) {
            const int row = row0 + rr;
            const size_t R = static_cast<size_t>(nrows) * Nx;
            const size_t i = static_cast<size_t>(row < nrows ? row : nrows - 1) * Nx + x;
            U = static_cast<T>(adv[i]);
            V = static_cast<T>(adv[R + i]);
            W = static_cast<T>(adv[2 * R + i]);
        }
        T F[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int t = 0; t < XM::targets; ++t) {
            const T dz = XM::prod[t][2] >= 0 ? ch[XM::prod[t][2] >= 0 ? XM::prod[t][2] : 0] : -(ch[2] + ch[5]);
            F[t] = -(U * ch[XM::prod[t][0]] + V * ch[XM::prod[t][1]] + W * dz);
        }
        // F_eta shares a transform with F_w: scaled by etaScale (a power of two
        // near N0) to the size of an acceleration, removed exactly below
        rb[xl] = {F[0], F[1]};
        rb[NxP + xl] = {F[2], F[3] * etaScale};
    }
    __syncthreads();
    // forward transforms: pairs 0 and 1 of every row (Nx = 256: two threads
    // per radix-16 butterfly)
    if constexpr (L2N == 8)
        fftFwd256<T, 2 * RB>(buf, BatchRows<2>{2 * RB, NxP, RS}, tw);
    else
        fftRows<T, -1, L2N, 2>(buf, Nx, log2Nx, 2 * RB, NxP, RS, tw);
    // wave w unpacks (row, pair) rq = w, w + nwaves, ...; its lanes walk k
    constexpr int nwaves = kFftThreads / 64;
    for (int rq = waveIndex(); rq < RB * 2; rq += nwaves) {
        const int q = rq & 1, rr = rq >> 1;
        const int row = row0 + rr;
        if (row >= nrows) continue;
        // uniform bases of the row's two target rows
        cx<T>* const Pa = P + (2 * q) * chStride + static_cast<size_t>(row) * nkxP;
        cx<T>* const Pb = P + (2 * q + 1) * chStride + static_cast<size_t>(row) * nkxP;
        const cx<T>* rb = buf + rr * RS + q * NxP;
        const T h = T(0.5), hb = q == 1 ? h / etaScale : h;
        for (int k = threadIdx.x & 63; k < nkx; k += 64) {
            const cx<T> Za = rb[xi(k)];
            const cx<T> Zb = rb[xi(k == 0 ? 0 : Nx - k)];
            const unsigned off = static_cast<unsigned>(k) * sizeof(cx<T>);
            if (2 * q < XM::targets)
                *reinterpret_cast<cx<T>*>(reinterpret_cast<char*>(Pa) + off) = {(Za.re + Zb.re) * h, (Za.im - Zb.im) * h};
            if (2 * q + 1 < XM::targets)
                *reinterpret_cast<cx<T>*>(reinterpret_cast<char*>(Pb) + off) = {(Za.im + Zb.im) * hb, (Zb.re - Za.re) * hb};
        }
    }
}

// c2r along x of nch channels to FP64 real fields [Nx,Ny,Nz,nch] (inverse
// transforms).  A workgroup takes rpb consecutive (y, z) rows: its real
// signals s = (row, channel) = (s / nch, s % nch) are paired two per complex
// transform (2q real part, 2q + 1 imaginary part), so that small channel
// counts still give the radix-16 passes several transforms.  OP selects what
// is written instead of the fields themselves:
//   kXinvAdvect (nch = 3, the scalar advection's d/dx, d/dy, d/dz):
//     out = -(u, v, w) . (the three fields) of each row, adv = [u, v, w] ([3][R]);
//   kXinvApv (nch = 9, the derivative channels of kApvChannels): the
//     available potential vorticity of each row, coriolis = f.
enum XinvOp : int { kXinvFields = 0, kXinvAdvect = 1, kXinvApv = 2 };
// Channels of the APV x transform: d/dy u, d/dz u, d/dx v, d/dz v, d/dx w,
// d/dy w, d/dx eta, d/dy eta, d/dz eta.
constexpr int kApvChannels = 9;
template <typename T, int OP = kXinvFields, int L2N = 0>
__global__ __launch_bounds__(256) void k_xinv(const cx<T>* __restrict__ G, double* __restrict__ out, int nch, int rpb,
                                              int NxArg, int log2NxArg, int Ny, int Nz, int nkx, int nkxP,
                                              const cx<T>* __restrict__ twX, const double* __restrict__ adv = nullptr,
                                              double coriolis = 0.0) {
    const int Nx = L2N > 0 ? 1 << (L2N > 0 ? L2N : 0) : NxArg;  // L2N > 0: the length fixed at compile time
    const int log2Nx = L2N > 0 ? L2N : log2NxArg;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nrows = Ny * Nz, row0 = static_cast<int>(blockIdx.x) * rpb;
    const int nsig = rpb * nch, ntr = (nsig + 1) / 2;
    const int NxP = ldsLength(Nx);
    const size_t chStride = static_cast<size_t>(Nz) * Ny * nkxP;
    for (int e = threadIdx.x; e < ntr * Nx; e += blockDim.x) {
        const int q = e / Nx, k = e % Nx;
        cx<T> A{T(0), T(0)}, B{T(0), T(0)};
        const int sa = 2 * q, sb = 2 * q + 1;
        if (row0 + sa / nch < nrows)
            A = hermitianX(G, static_cast<size_t>(row0 + sa / nch) * nkxP, chStride, sa % nch, k, Nx, nkx);
        if (sb < nsig && row0 + sb / nch < nrows)
            B = hermitianX(G, static_cast<size_t>(row0 + sb / nch) * nkxP, chStride, sb % nch, k, Nx, nkx);
        buf[q * NxP + ldsIndex(k, 1)] = {A.re - B.im, A.im + B.re};
    }
    __syncthreads();
    fftLdsBatched<T, +1, L2N>(buf, Nx, log2Nx, ntr, NxP, twX);
    const size_t R = static_cast<size_t>(Nx) * Ny * Nz;
    auto field = [&](int sg, int x) {
        const cx<T> v = buf[(sg / 2) * NxP + ldsIndex(x, 1)];
        return static_cast<double>((sg & 1) ? v.im : v.re);
    };
    if constexpr (OP == kXinvApv) {
        // APVOperation.m:15-26 with eta_true = eta: zeta_z - zeta_x eta_x -
        // zeta_y eta_y - (zeta_z + f) eta_z, zeta_x = w_y - v_z, zeta_y =
        // u_z - w_x, zeta_z = v_x - u_y (classDefinedOperationForKnownVariable.m:126-137)
        for (int e = threadIdx.x; e < rpb * Nx; e += blockDim.x) {
            const int r = e / Nx, x = e % Nx, row = row0 + r;
            if (row >= nrows) continue;
            const int c0 = kApvChannels * r;
            const double uy = field(c0, x), uz = field(c0 + 1, x), vx = field(c0 + 2, x), vz = field(c0 + 3, x);
            const double wx = field(c0 + 4, x), wy = field(c0 + 5, x);
            const double ex = field(c0 + 6, x), ey = field(c0 + 7, x), ez = field(c0 + 8, x);
            const double zx = wy - vz, zy = uz - wx, zz = vx - uy;
            out[static_cast<size_t>(row) * Nx + x] = zz - zx * ex - zy * ey - (zz + coriolis) * ez;
        }
        return;
    }
    if constexpr (OP == kXinvAdvect) {
        for (int e = threadIdx.x; e < rpb * Nx; e += blockDim.x) {
            const int r = e / Nx, x = e % Nx, row = row0 + r;
            if (row >= nrows) continue;
            const size_t i = static_cast<size_t>(row) * Nx + x;
            out[i] = -(adv[i] * field(3 * r, x) + adv[R + i] * field(3 * r + 1, x) + adv[2 * R + i] * field(3 * r + 2, x));
        }
        return;
    }
    for (int e = threadIdx.x; e < nsig * Nx; e += blockDim.x) {
        const int sg = e / Nx, x = e % Nx, row = row0 + sg / nch;
        if (row >= nrows) continue;
        const cx<T> v = buf[(sg / 2) * NxP + ldsIndex(x, 1)];
        out[(sg % nch) * R + static_cast<size_t>(row) * Nx + x] = static_cast<double>((sg & 1) ? v.im : v.re);
    }
}

// advectFGridScalar (...Kernel.cpp:1032-1138), FP64: the derivative assembly
// of the scalar's half spectrum C[0][z][l][k] (k < nxh = Nx/2 + 1, row
// stride nxhP), for a workgroup's 8 columns k0 ... k0 + 7 of row l.  With
// v = scale * C[0] (scale = 1/(Nx Ny)): C[0] = i k v (k = 0 at the x Nyquist
// column), C[1] = i l v, C[2] = sum_w DT[w][z] v_w, DT the transposed column operator of
// the reference's DCT-I, -m_j, DST-I sequence (:1078-1094; rows z = 0 and
// Nz - 1 zero); imaginary parts cleared on the self-conjugate rows
// (:1095-1100).  LDS: the 8 columns over z.
constexpr double kScalarPi = 3.14159265358979323846;
__global__ __launch_bounds__(1024) void k_scalarDeriv(cx<double>* __restrict__ C, int Nx, int Ny, int Nz, int nxh,
                                                     int nxhP, double Lx, double Ly, const double* __restrict__ D,
                                                     double scale) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<double>* col = reinterpret_cast<cx<double>*>(smem);  // [Nz][8]
    const int k0 = static_cast<int>(blockIdx.x) * 8, l = static_cast<int>(blockIdx.y);
    const size_t plane = static_cast<size_t>(Ny) * nxhP;
    const size_t chStride = static_cast<size_t>(Nz) * plane;
    const size_t base = static_cast<size_t>(l) * nxhP + k0;
    for (int e = threadIdx.x; e < Nz * 8; e += blockDim.x) {
        const int z = e >> 3, c = e & 7;
        const cx<double> v = k0 + c < nxhP ? C[z * plane + base + c] : cx<double>{0.0, 0.0};
        col[e] = cx<double>{v.re * scale, v.im * scale};
    }
    __syncthreads();
    const int lMode = l < (Ny + 1) / 2 ? l : l - Ny;
    const double lw = 2.0 * kScalarPi * static_cast<double>(lMode) / Ly;
    const bool lSelf = l == 0 || (Ny % 2 == 0 && l == Ny / 2);
    for (int e = threadIdx.x; e < Nz * 8; e += blockDim.x) {
        const int z = e >> 3, c = e & 7, k = k0 + c;
        if (k >= nxh) continue;
        const bool nyq = Nx % 2 == 0 && k == Nx / 2;
        const double kw = nyq ? 0.0 : 2.0 * kScalarPi * static_cast<double>(k) / Lx;
        const cx<double> v = col[e];
        cx<double> dx{-v.im * kw, v.re * kw}, dy{-v.im * lw, v.re * lw};
        if (lSelf && (k == 0 || nyq)) dx.im = dy.im = 0.0;
        const size_t o = z * plane + base + c;
        C[o] = dx;
        C[chStride + o] = dy;
    }
    // d/dz on FP64 MFMA 16x16x4: the tile's 8 complex columns are the 16 real
    // GEMM columns n (column n / 2, real or imaginary part n % 2); wave w
    // takes 16-row z tiles w, w + nwaves, ... and contracts over w in steps of
    // 4 (entries beyond Nz are zeros).  Lane l: A = DT[4 s + l/16][16 zt + l%16],
    // B = the staged value (w = 4 s + l/16, n = l%16); C element i is
    // z = 16 zt + l/16 + 4 i, n = l%16.
    using MF = Mfma<double>;
    const int lane = threadIdx.x & 63, nwaves = static_cast<int>(blockDim.x >> 6);
    const int n = lane & 15, kq = lane >> 4, c = n >> 1, part = n & 1;
    const double* colw = reinterpret_cast<const double*>(col) + 2 * c + part;  // word of (w, c, part) at w * 16
    const int k = k0 + c;
    const bool nyqc = Nx % 2 == 0 && k == Nx / 2;
    const bool zeroIm = part == 1 && lSelf && (k == 0 || nyqc);
    const int nsteps = (Nz + 3) / 4, ztiles = (Nz + 15) / 16;
    for (int zt = static_cast<int>(threadIdx.x >> 6); zt < ztiles; zt += nwaves) {
        const int za = 16 * zt + n;  // this lane's A row
        typename MF::acc acc = MF::zero();
        for (int st = 0; st < nsteps; ++st) {
            const int w = 4 * st + kq;
            const double a = (w < Nz && za < Nz) ? D[static_cast<size_t>(w) * Nz + za] : 0.0;
            const double b = w < Nz ? colw[w * 16] : 0.0;
            acc = MF::run(a, b, acc);
        }
        if (k >= nxh) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int z = 16 * zt + MF::row(kq, i);
            if (z >= Nz) continue;
            reinterpret_cast<double*>(C + 2 * chStride + z * plane + base + c)[part] = zeroIm ? 0.0 : acc[i];
        }
    }
}

// The scalar antialias's y round trip in one pass (:1014-1029): a workgroup
// loads its [Ny x KT] tile of the x half spectrum C[z][y][k], transforms it
// along y, weights each (ky, kx) entry by keep (1/(Nx Ny) on the kept rows,
// 0 elsewhere), transforms back and stores the columns k < nkx.
template <int L2N, int L2KT>
__global__ __launch_bounds__(kFftThreads) void k_ymask(cx<double>* __restrict__ C, int NyArg, int log2NyArg, int nkx,
                                                       int nkxP, int log2KTArg, const cx<double>* __restrict__ twY,
                                                       const double* __restrict__ keep) {
    const int Ny = L2N > 0 ? 1 << (L2N > 0 ? L2N : 0) : NyArg;
    const int log2Ny = L2N > 0 ? L2N : log2NyArg;
    const int log2KT = L2N > 0 ? L2KT : log2KTArg;
    const int KT = 1 << log2KT;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<double>* tw = reinterpret_cast<cx<double>*>(smem);
    cx<double>* tile = tw + Ny;
    const int kx0 = static_cast<int>(blockIdx.x) * KT;
    const size_t plane = static_cast<size_t>(blockIdx.y) * Ny * nkxP;
    const int nel = Ny * KT;
    loadTwiddles(tw, twY, Ny);
    for (int e = fftThread(); e < nel; e += kFftThreads)
        tile[e] = C[plane + static_cast<size_t>(e >> log2KT) * nkxP + kx0 + (e & (KT - 1))];
    __syncthreads();
    fftLds<double, -1, L2N>(tile, Ny, log2Ny, BatchCols{log2KT}, tw);
    for (int e = fftThread(); e < nel; e += kFftThreads) {
        const double w = keep[static_cast<size_t>(e >> log2KT) * nkxP + kx0 + (e & (KT - 1))];
        tile[e] = cx<double>{tile[e].re * w, tile[e].im * w};
    }
    __syncthreads();
    fftLds<double, +1, L2N>(tile, Ny, log2Ny, BatchCols{log2KT}, tw);
    for (int e = fftThread(); e < nel; e += kFftThreads) {
        const int kk = e & (KT - 1);
        if (kx0 + kk < nkx) C[plane + static_cast<size_t>(e >> log2KT) * nkxP + kx0 + kk] = tile[e];
    }
}

// Projection targets of the NT transformed products: u, v, w, eta
// (nonhydrostatic) or u, v, eta (hydrostatic).
template <int NT>
__host__ __device__ constexpr int projTarget(int t) { return NT == 4 ? t : (t == 2 ? 3 : t); }

// k_proj stages one family's targets at a time in FP64 (its LDS then bounds
// the workgroups per CU), all targets at once in FP32.
template <typename T>
constexpr bool projSplitStaging() { return sizeof(T) == 8; }
// LDS bytes of k_proj: the staged columns (Zq rows per target; one family's
// targets at most two, or all), later the transposed results.
template <typename T>
constexpr size_t projLdsBytes(int NT, int Zq, int Ja) {
    const int nst = projSplitStaging<T>() ? (NT < 2 ? NT : 2) : NT;
    const size_t staged = static_cast<size_t>(nst) * Zq * 8 * sizeof(cx<T>);
    const size_t results = static_cast<size_t>(Ja) * 4 * 17 * sizeof(T);
    return staged > results ? staged : results;
}

// Vertical analysis (MFMA) of the NT transformed products: the spectral
// value of every target at (j, mode), stored as Tg[t][j + Nj*mode] (the
// storage row's conjugation undone; the 1/(Nx Ny) scaling and the
// wave-vortex projection follow in k_stage).
// One workgroup owns 8 consecutive columns and has one wave per 16-row tile
// of j; each target's Q columns (all z) are staged in LDS once and shared by
// the waves.  ana.m[0] and m[1] are the cosine- and sine-family analysis
// matrices ([Ja x Za]).
// FOLD (constant N): column L - z of the DCT-I / DST-I analysis (L = Nz - 1)
// is column z times (-1)^j or -(-1)^j, so rows of even j contract the sums
// s_z = q_z + q_{L-z} (DCT-I) or the differences d_z = q_z - q_{L-z}
// (DST-I) over z <= L/2, rows of odd j the other: half the MFMA work.  Za is
// then the folded row count, ana.m[0], m[1] hold the even-j rows and m[2],
// m[3] the odd-j rows, and the first Ja/32 waves take even j.
template <typename T, int NT, int KS, bool FOLD>
__global__ __launch_bounds__(1024) void k_proj(const cx<T>* __restrict__ Q, const ColumnDev* __restrict__ cols,
                                               int ncol, SynthMats<T> ana, int Ja, int Za, int Zk, int zmid, int Nz,
                                               int Nj, int Ny, int nkxP, int M, cx<T>* __restrict__ Tg) {
    using MF = Mfma<T>;
    extern __shared__ __align__(16) unsigned char smem[];
    // staged targets, slot s: [s][Za][8], folded [s][2][Za][8] (sums,
    // differences); slot = target, or with split staging the target's index
    // within its family (t & 1)
    constexpr bool kSplit = projSplitStaging<T>();
    cx<T>* qs = reinterpret_cast<cx<T>*>(smem);
    __shared__ ColumnDev scol[8];
    const int lane = threadIdx.x & 63;
    const int jt = waveIndex();
    const int cg = xcdBlock(blockIdx.x, gridDim.x);
    if (threadIdx.x < 8) {
        const int col = cg * 8 + threadIdx.x;
        if (col < ncol) scol[threadIdx.x] = cols[col];
        else scol[threadIdx.x] = ColumnDev{0, 0, -1, kColCompletion, 0.0, 0.0, 0.0, 0.0, 0.0};
    }
    const int n = lane & 15, kq = lane >> 4;
    const int part = n & 1;
    const size_t tStride = static_cast<size_t>(Nz) * Ny * nkxP;
    const int tsteps = Za / 4;  // operator steps per row tile (the table's layout)
    const int nsteps = Zk / 4;  // steps contracted on MFMA: rows z < Zk
    const int L = Nz - 1;
    __syncthreads();  // scol ready
    // stage the columns of family fam's targets (0, 1 | 2, 3; NT = 3: 0, 1 |
    // 2, whose projection target is 3), or of all targets (fam < 0): lane
    // (zr, c) of wave w takes column c of levels z = 8 g + zr, g = w,
    // w + nwaves, ...; all loads of a batch (two levels each when folded)
    // precede its stores.
    auto stage = [&](int fam) {
        constexpr int NS = kSplit ? 2 : NT;  // slots staged per call
        const int t0 = fam < 0 ? 0 : 2 * fam;
        const int nwaves = blockDim.x >> 6;
        const int c = lane & 7, zr = lane >> 3;
        const ColumnDev& cc = scol[c];
        const size_t plane = static_cast<size_t>(Ny) * nkxP;
        const size_t colBase = static_cast<size_t>(cc.iy) * nkxP + cc.kx;
        const bool colLive = cc.wv >= 0;
        for (int g = jt; 8 * g < Za; g += nwaves) {
            const int z = 8 * g + zr;
            const int zc = z < Nz ? z : Nz - 1;  // clamped: loads are unconditional
            const int zm = L - zc >= 0 ? L - zc : 0;
            cx<T> v[NS], w[NS];
#pragma unroll
            for (int tf = 0; tf < NS; ++tf) {
                const int tt = t0 + tf < NT ? t0 + tf : NT - 1;
                v[tf] = Q[tt * tStride + zc * plane + colBase];
                if constexpr (FOLD) w[tf] = Q[tt * tStride + zm * plane + colBase];
            }
            __builtin_amdgcn_sched_barrier(0);  // all loads issued before the first store
            if (z >= Za) continue;
            const cx<T> zero{T(0), T(0)};
#pragma unroll
            for (int tf = 0; tf < NS; ++tf) {
                if (t0 + tf >= NT) continue;
                if constexpr (FOLD) {
                    const bool ok = colLive && 2 * z <= L, pair = z < L - z;
                    const cx<T> sv = !ok ? zero : pair ? cadd(v[tf], w[tf]) : v[tf];
                    const cx<T> dv = !ok || !pair ? zero : csub(v[tf], w[tf]);
                    qs[((2 * tf) * Za + z) * 8 + c] = sv;
                    qs[((2 * tf + 1) * Za + z) * 8 + c] = dv;
                } else {
                    qs[(tf * Za + z) * 8 + c] = colLive && z < Nz ? v[tf] : zero;
                }
            }
        }
    };
    // analysis: each family's matrix is streamed once for all its targets,
    // in register batches of kS steps loaded one batch ahead
    typename MF::acc acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = MF::zero();
    const T* qw = reinterpret_cast<const T*>(qs) + 2 * (n >> 1) + part;  // [t][z][c] words
    const int nEven = Ja / 32;
    const int P = FOLD && jt >= nEven ? 1 : 0;        // folded: parity of this wave's j
    const int jh = FOLD ? jt - P * nEven : jt;        // row tile within the parity
    constexpr int kS = KS;  // the step count Za / 4 is a multiple of 2 kS
#pragma unroll
    for (int fam = 0; fam < 2; ++fam) {
        if (kSplit || fam == 0) {
            if (fam > 0) __syncthreads();  // every wave has read the previous family's block
            stage(kSplit ? fam : -1);
            __syncthreads();
        }
        const T* At = ana.m[fam + 2 * P] + static_cast<size_t>(jh) * tsteps * 64;  // this row tile
        const T* Aj = At + lane;
        // staged block read by this family: the columns themselves, or
        // (folded) the sums or differences, by family and parity of j
        const int par = FOLD ? (fam ^ P) : 0;
        // two named batches used in turn (no register copies, so the waits
        // before the MFMAs cover only the batch in use)
        auto load = [&](T (&dst)[kS], int s0) {
#pragma unroll
            for (int u = 0; u < kS; ++u) dst[u] = Aj[(s0 + u) * 64];
        };
        auto run = [&](const T (&a)[kS], int s0) {
            T b[kS][NT];
#pragma unroll
            for (int u = 0; u < kS; ++u) {
                const int zrow = 4 * (s0 + u) + kq;
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    if ((projTarget<NT>(t) >= 2) == (fam == 1))
                        b[u][t] = FOLD ? qw[2 * (((2 * (kSplit ? t & 1 : t) + par) * Za + zrow) * 8)]
                                       : qw[2 * (((kSplit ? t & 1 : t) * Za + zrow) * 8)];
            }
#pragma unroll
            for (int u = 0; u < kS; ++u)
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    if ((projTarget<NT>(t) >= 2) == (fam == 1)) acc[t] = MF::run(a[u], b[u][t], acc[t]);
        };
        // (scheduling barriers keep each load a whole batch ahead of its use)
        T a0[kS], a1[kS];
        load(a0, 0);
        int s0 = 0;
        for (; s0 + 2 * kS < nsteps; s0 += 2 * kS) {
            load(a1, s0 + kS);
            __builtin_amdgcn_sched_barrier(0);
            run(a0, s0);
            load(a0, s0 + 2 * kS);
            __builtin_amdgcn_sched_barrier(0);
            run(a1, s0 + kS);
        }
        load(a1, s0 + kS);
        __builtin_amdgcn_sched_barrier(0);
        run(a0, s0);
        run(a1, s0 + kS);
        if (zmid >= 0) {
            // the one staged row z = zmid beyond the MFMA tiles (the folded
            // midpoint L/2): a rank-1 term, operator column zmid of this
            // tile's rows times the staged value of the lane's column
            T am[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) am[i] = At[(zmid >> 2) * 64 + 16 * (zmid & 3) + MF::row(kq, i)];
            const int par = FOLD ? (fam ^ P) : 0;
#pragma unroll
            for (int t = 0; t < NT; ++t)
                if ((projTarget<NT>(t) >= 2) == (fam == 1)) {
                    const T sm = FOLD ? qw[2 * (((2 * (kSplit ? t & 1 : t) + par) * Za + zmid) * 8)]
                                      : qw[2 * (((kSplit ? t & 1 : t) * Za + zmid) * 8)];
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[t][i] += am[i] * sm;
                }
        }
    }
    // The analysis results go through LDS as res[t][j][n] (the staged columns
    // are no longer needed) so that the stores' lanes walk j: lane (c, jj)
    // takes rows j = 16 jt + 8 s + jj of column c.
    constexpr int kResStride = 17;  // words per row j
    __syncthreads();  // every wave has finished reading qs
    T* res = reinterpret_cast<T*>(smem);
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 16 * jh + MF::row(kq, i);
            const int j = FOLD ? 2 * r + P : r;
            res[(t * Ja + j) * kResStride + n] = acc[t][i];
        }
    __syncthreads();
    const int ec = lane >> 3, jj = lane & 7;
    const ColumnDev cd = scol[ec];
    if (cd.wv < 0 || (cd.flags & kColCompletion)) return;
    const bool conj = (cd.flags & kColConj) != 0;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
        const int j = 16 * jt + 8 * ii + jj;
        if (j >= Nj) continue;
        const size_t idx = static_cast<size_t>(j) + static_cast<size_t>(Nj) * static_cast<size_t>(cd.wv);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const T* rt = res + (t * Ja + j) * kResStride + 2 * ec;
            Tg[t * static_cast<size_t>(M) + idx] = cx<T>{rt[0], conj ? -rt[1] : rt[1]};
        }
    }
}

// The folded vertical analysis with split-FP16 MFMA products (the
// mixed-precision form of k_proj<float, NT, ., true>; the scheme of
// k_synth16): the analysis operators (ana.m[fam + 2 P], prepared on the
// host: per row tile [N32 x (hi, lo) x 64 x 8 | TAIL x (hi, lo) x 64 x 4],
// lane l holding rows j = 16 tile + l%16 and folded columns z = 32 s + 8
// (l/16) + t, resp. 32 N32 + 4 (l/16) + t) with one scale per matrix, the
// staged sums and differences with one scale per (target, sums or
// differences, column) from their largest entry.  The contraction covers
// every folded row z <= L/2 (the midpoint row included).  Same waves (one
// per 16-row tile of j, even j first) and the same transposed stores of Tg.
// Targets T0 ... T0 + NTB - 1 of the NT (fixed at compile time, so that
// every MFMA of the body is unconditional) for column group cg.
template <int NT, int NTB, int T0, int N32, int TAIL>
__device__ __forceinline__ void proj16Body(const cx<float>* __restrict__ Qall, const ColumnDev* __restrict__ cols,
                                           int ncol, const SynthMats16& ana, int Ja, int Nz, int Nj, int Ny, int nkxP,
                                           int M, cx<float>* __restrict__ TgAll, int cg, unsigned char* smem,
                                           ColumnDev* scol, unsigned* colMax, float* colInv) {
    using MF = Mfma<float>;
    constexpr int NC8 = 4 * N32 + 2 * TAIL;  // 8-row chunks of the contraction
    constexpr int kTileHalves = N32 * 1024 + TAIL * 512;
    constexpr int kItems = NTB * 16 * NC8;   // staging items (target, chunk, column, part)
    constexpr int t0 = T0;
    const size_t tStrideAll = static_cast<size_t>(Nz) * Ny * nkxP;
    const cx<float>* __restrict__ Q = Qall + t0 * tStrideAll;
    cx<float>* __restrict__ Tg = TgAll + t0 * static_cast<size_t>(M);
    half8* sb = reinterpret_cast<half8*>(smem);  // [target][sums, diffs][hl][chunk][column c x part]
    const int lane = threadIdx.x & 63;
    const int jt = waveIndex();
    if (threadIdx.x < 8) {
        const int col = cg * 8 + threadIdx.x;
        if (col < ncol) scol[threadIdx.x] = cols[col];
        else scol[threadIdx.x] = ColumnDev{0, 0, -1, kColCompletion, 0.0, 0.0, 0.0, 0.0, 0.0};
    }
    if (threadIdx.x < NTB * 16) colMax[threadIdx.x] = 0u;
    __syncthreads();
    const size_t tStride = static_cast<size_t>(Nz) * Ny * nkxP, plane = static_cast<size_t>(Ny) * nkxP;
    const int L = Nz - 1;
    // staging: item (target t, chunk g, column c, part) forms one part of the
    // folded sums and differences of rows z = 8 g ... 8 g + 7; sixteen
    // consecutive items write the 16 column-part slots of one 256-byte row
    // (launched with at least kItems / 2 threads: two items per thread at
    // most; a wave whose items are all past the end skips the second)
    float sv[2][8], dv[2][8];
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        if (static_cast<int>(threadIdx.x - lane) + it * static_cast<int>(blockDim.x) >= kItems) break;
        const int item = threadIdx.x + it * blockDim.x;
        const int part = item & 1, c = (item >> 1) & 7, g = (item >> 4) % NC8, t = (item >> 4) / NC8;
        const bool live = item < kItems && t0 + t < NT;
        const ColumnDev& cc = scol[c];
        const size_t colBase = static_cast<size_t>(cc.iy) * nkxP + cc.kx;
        const bool colLive = live && cc.wv >= 0;
        const int tt = t < NTB && t0 + t < NT ? t : 0;
        float mxs = 0.0f, mxd = 0.0f;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int z = 8 * g + i;
            const int zc = z < Nz ? z : Nz - 1;  // clamped: loads are unconditional
            const int zm = L - zc >= 0 ? L - zc : 0;
            const cx<float> v = Q[tt * tStride + zc * plane + colBase];
            const cx<float> w = Q[tt * tStride + zm * plane + colBase];
            const float a = part ? v.im : v.re, b = part ? w.im : w.re;
            const bool ok = colLive && 2 * z <= L, pair = z < L - z;
            sv[it][i] = !ok ? 0.0f : pair ? a + b : a;
            dv[it][i] = !ok || !pair ? 0.0f : a - b;
            mxs = fmaxf(mxs, fabsf(sv[it][i]));
            mxd = fmaxf(mxd, fabsf(dv[it][i]));
        }
        if (live) {
            atomicMax(&colMax[(t * 2 + 0) * 8 + c], __float_as_uint(mxs));
            atomicMax(&colMax[(t * 2 + 1) * 8 + c], __float_as_uint(mxd));
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int item = threadIdx.x + it * blockDim.x;
        if (item >= kItems) continue;
        const int part = item & 1, c = (item >> 1) & 7, g = (item >> 4) % NC8, t = (item >> 4) / NC8;
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {
            const int slot = (t * 2 + sd) * 8 + c;
            const int e = split16Exponent(__uint_as_float(colMax[slot]));
            const float sc = ldexpf(1.0f, e);
            if (g == 0 && part == 0) colInv[slot] = ldexpf(1.0f, -e);
            half8 hi, lo;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                _Float16 h, l;
                splitHalf((sd ? dv[it][i] : sv[it][i]) * sc, h, l);
                hi[i] = h;
                lo[i] = l;
            }
            const int base = ((t * 2 + sd) * 2) * NC8 + g;
            sb[base * 16 + 2 * c + part] = hi;
            sb[(base + NC8) * 16 + 2 * c + part] = lo;
        }
    }
    __syncthreads();
    const int n = lane & 15, kq = lane >> 4;
    const int nEven = Ja / 32;
    const int P = jt >= nEven ? 1 : 0;  // parity of this wave's j
    const int jh = jt - P * nEven;      // row tile within the parity
    // family of each of this workgroup's targets (u, v: cosine; w, eta:
    // sine); -1 past the last target
    auto tfamOf = [](int t) { return T0 + t < NT ? (projTarget<NT>(T0 + t) >= 2 ? 1 : 0) : -1; };
    int tfam[NTB];
#pragma unroll
    for (int t = 0; t < NTB; ++t) tfam[t] = tfamOf(t);
    typename MF::acc acc[NTB];
#pragma unroll
    for (int t = 0; t < NTB; ++t) acc[t] = MF::zero();
    float scale[NTB];
#pragma unroll
    for (int t = 0; t < NTB; ++t) scale[t] = 0.0f;
#pragma unroll
    for (int fam = 0; fam < 2; ++fam) {
        const _Float16* A = ana.m[fam + 2 * P] + static_cast<size_t>(jh) * kTileHalves;
        const int sd = fam ^ P;  // even j of the cosine family: sums, ...
        bool any = false;
#pragma unroll
        for (int t = 0; t < NTB; ++t) any = any || tfam[t] == fam;
        if (!any) continue;  // uniform: no target of this family here
#pragma unroll
        for (int t = 0; t < NTB; ++t)
            if (tfam[t] == fam) scale[t] = ana.inv[fam + 2 * P] * colInv[(t * 2 + sd) * 8 + (n >> 1)];
#pragma unroll
        for (int s = 0; s < N32; ++s) {
            const half8 aH = reinterpret_cast<const half8*>(A + (2 * s) * 512)[lane];
            const half8 aL = reinterpret_cast<const half8*>(A + (2 * s + 1) * 512)[lane];
#pragma unroll
            for (int t = 0; t < NTB; ++t)
                if (tfam[t] == fam) {
                    const int b = ((t * 2 + sd) * 2) * NC8 + 4 * s + kq;
                    const half8 bH = sb[b * 16 + n], bL = sb[(b + NC8) * 16 + n];
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aL, bH, acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aH, bL, acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aH, bH, acc[t], 0, 0, 0);
                }
        }
        if constexpr (TAIL) {
            const _Float16* tA = A + N32 * 1024;
            const half4 aH = reinterpret_cast<const half4*>(tA)[lane];
            const half4 aL = reinterpret_cast<const half4*>(tA + 256)[lane];
            const int ch8 = 4 * N32 + (kq >> 1), sub = 4 * (kq & 1);
#pragma unroll
            for (int t = 0; t < NTB; ++t)
                if (tfam[t] == fam) {
                    const int b = ((t * 2 + sd) * 2) * NC8 + ch8;
                    const half4 bH = *reinterpret_cast<const half4*>(reinterpret_cast<const _Float16*>(sb + b * 16 + n) + sub);
                    const half4 bL =
                        *reinterpret_cast<const half4*>(reinterpret_cast<const _Float16*>(sb + (b + NC8) * 16 + n) + sub);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(aL, bH, acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(aH, bL, acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(aH, bH, acc[t], 0, 0, 0);
                }
        }
    }
    // the results through LDS as res[t][j][n], stores walking j (as k_proj)
    constexpr int kResStride = 17;
    __syncthreads();  // every wave has finished reading the staged operands
    float* res = reinterpret_cast<float*>(smem);
#pragma unroll
    for (int t = 0; t < NTB; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = 2 * (16 * jh + MF::row(kq, i)) + P;
            res[(t * Ja + j) * kResStride + n] = acc[t][i] * scale[t];
        }
    __syncthreads();
    const int ec = lane >> 3, jj = lane & 7;
    const ColumnDev cd = scol[ec];
    if (cd.wv < 0 || (cd.flags & kColCompletion)) return;
    const bool conj = (cd.flags & kColConj) != 0;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
        const int j = 16 * jt + 8 * ii + jj;
        if (j >= Nj) continue;
        const size_t idx = static_cast<size_t>(j) + static_cast<size_t>(Nj) * static_cast<size_t>(cd.wv);
#pragma unroll
        for (int t = 0; t < NTB; ++t) {
            if (t0 + t >= NT) continue;
            const float* rt = res + (t * Ja + j) * kResStride + 2 * ec;
            Tg[t * static_cast<size_t>(M) + idx] = cx<float>{rt[0], conj ? -rt[1] : rt[1]};
        }
    }
}

// Split-FP16 analysis: workgroup -> (column group cg, targets t0 ... t0 +
// NTB - 1 of the NT); one XCD runs consecutive column groups of one target
// set.
template <int NT, int NTB, int N32, int TAIL>
__global__ __launch_bounds__(1024) void k_proj16(const cx<float>* __restrict__ Qall, const ColumnDev* __restrict__ cols,
                                                 int ncol, SynthMats16 ana, int Ja, int Nz, int Nj, int Ny, int nkxP,
                                                 int M, cx<float>* __restrict__ TgAll) {
    static_assert(NT <= 2 * NTB, "at most two target sets");
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ ColumnDev scol[8];
    __shared__ unsigned colMax[NTB * 16];  // per (target, sums or diffs, column): largest |entry| as float bits
    __shared__ float colInv[NTB * 16];
    const int ncg = (ncol + 7) / 8;
    const int logical = xcdBlock(blockIdx.x, gridDim.x);
    const int cg = logical % ncg;
    if (logical < ncg)
        proj16Body<NT, NTB, 0, N32, TAIL>(Qall, cols, ncol, ana, Ja, Nz, Nj, Ny, nkxP, M, TgAll, cg, smem, scol,
                                          colMax, colInv);
    else if constexpr (NT > NTB)
        proj16Body<NT, NTB, NTB, N32, TAIL>(Qall, cols, ncol, ana, Ja, Nz, Nj, Ny, nkxP, M, TgAll, cg, smem, scol,
                                            colMax, colInv);
}

// Per-mode wavenumbers of the projection.
struct ModeDev {
    double k, l, Kh, cosA, sinA;
};

// Inputs of the per-coefficient tables: constants and per-j arrays.
struct CoefConsts {
    double f, f2, N02, g, Lz;
    double S;   // sqrt(2 g_modal / (Lz (N0^2 - f^2))): nonhydrostatic wave normalization
    double P0;  // sqrt(g Lz / (2 (N0^2 - f^2))): wave-projection prefactor (constant N)
    int hydro, stratified;
    const double* m;    // m_j
    const double* h0;   // equivalent depths
    const double* Fg;   // F, G normalizations
    const double* Gg;
    const double* gws;  // G wave scale
    const double* apmW;
    const double* inertialScale;
    // per-j FP64 factors of the mixed-precision projection tables
    // (coefficientsF): [kMixedJ][NjMixed]
    const double* mixedJ;
    int NjMixed;
};
// rows of CoefConsts::mixedJ
enum MixedJ : int {
    kMjGwgOverGg = 0,  // (Gg / Gw) / Gg
    kMjDeltaScale,     // (h0 Gg / Gw) / Fg
    kMjLr2inv,         // f^2 / (g h0), 0 at j = 0
    kMjA0VortGeo,      // 1 / Fg
    kMjA0BuoyGeo,      // (-f / h0) / Gg, 0 at j = 0
    kMjApmDScaled,     // (m_j / 2) prefactor
    kMjA0VortMda,      // (f^2 / (2 h0)) / Fg
    kMjA0BuoyMda,      // 1 / Gg
    kMjNA0Num,         // -(f / g) Gg, 0 at j = 0
    kMixedJ
};

// The per-coefficient tables of coefficient (j, mode) -- the descriptor's
// formulas (WVKernelTypes.cpp:214-284) -- evaluated with the same operations
// in the same order, each rounded on its own (no contraction), so the values
// equal the host tables bit for bit.
struct CoefRow {
    double om, NAp, NA0, A0Vort, A0Buoy, ApmN, ApmDScaled;
    cx<double> UAp, VAp, WAp, UA0, VA0, ApmD;
};
__device__ __forceinline__ CoefRow coefficients(const CoefConsts& c, int j, const ModeDev& md) {
    auto mul = [](double a, double b) { return mulRn(a, b); };
    auto dv = [](double a, double b) { return divRn(a, b); };
    auto add = [](double a, double b) { return addRn(a, b); };
    CoefRow r{};
    const double Kh = md.Kh, Kh2 = mul(Kh, Kh);
    const double Mj = c.m[j], h0 = c.h0[j], Fg = c.Fg[j], Gg = c.Gg[j];
    const double sign = j % 2 == 0 ? 1.0 : -1.0;
    const bool wave = Kh > 0.0 && j > 0, inertial = Kh == 0.0, geostrophic = Kh > 0.0, mda = Kh == 0.0 && j > 0;
    const double N02mf2 = subRn(c.N02, c.f2);
    double hpm = 1.0;
    if (j > 0)
        hpm = c.stratified ? h0 : c.hydro ? dv(c.N02, mul(mul(c.g, Mj), Mj)) : dv(N02mf2, mul(c.g, add(mul(Mj, Mj), Kh2)));
    double Fw = Fg, Gw = Gg;
    if (!c.hydro) {
        Fw = j == 0 ? 2.0 : mul(mul(mul(sign, hpm), Mj), c.S);
        Gw = j == 0 ? 1.0 : mul(sign, c.S);
    }
    const double om = __dsqrt_rn(add(mul(mul(c.g, hpm), Kh2), c.f2));
    const double Fwg = dv(Fg, Fw), Gwg = dv(Gg, Gw);
    const double fs = dv(Fg, Fwg);
    r.om = om;
    const double prefactor = c.stratified ? 0.0 : mul(sign, c.P0);
    r.ApmDScaled = mul(dv(Mj, 2.0), prefactor);
    if (wave) {
        const double gs = c.gws[j];
        const double fo = dv(c.f, om);
        r.UAp = {mul(md.cosA, fs), mul(mul(-fo, md.sinA), fs)};
        r.VAp = {mul(md.sinA, fs), mul(mul(fo, md.cosA), fs)};
        r.WAp = {mul(0.0, gs), mul(mul(-Kh, hpm), gs)};
        r.NAp = mul(dv(mul(-Kh, hpm), om), gs);
        const double deltaScale = dv(mul(h0, Gwg), Fg);
        const double twoKh = mul(mul(2.0, Kh), hpm);
        r.ApmD = {mul(0.0, deltaScale), mul(dv(-1.0, twoKh), deltaScale)};
        r.ApmN = dv(mul(dv(-om, twoKh), Gwg), Gg);
    } else if (inertial) {
        r.UAp = {fs, 0.0};
        r.VAp = {0.0, fs};
    }
    if (geostrophic) {
        const double lr2inv = j == 0 ? 0.0 : dv(c.f2, mul(c.g, h0));
        const double den = add(Kh2, lr2inv);
        r.UA0 = {mul(0.0, Fg), mul(dv(md.l, den), Fg)};
        r.VA0 = {mul(0.0, Fg), mul(dv(-md.k, den), Fg)};
        r.NA0 = mul(j == 0 ? 0.0 : dv(-dv(c.f, c.g), den), Gg);
        r.A0Vort = dv(1.0, Fg);
        r.A0Buoy = j == 0 ? 0.0 : dv(dv(-c.f, h0), Gg);
    } else if (mda) {
        r.NA0 = Gg;
        r.A0Buoy = dv(1.0, Gg);
        r.A0Vort = dv(dv(c.f2, mul(2.0, h0)), Fg);
    }
    return r;
}

// The mixed-precision form of the same tables.  The projection tables
// (A0Vort, A0Buoy, NA0, ApmN, ApmD, ApmDScaled) stay FP64: the wave-vortex
// projection of a nearly balanced flow sums terms that cancel to a small
// remainder, and FP32 tables would bias that remainder the same way on every
// step.  Their per-j factors come from host tables (CoefConsts::mixedJ), so
// a coefficient needs four FP64 quotients and one square root.  The
// synthesis tables of the field spectra (UAp, VAp, WAp, NAp, UA0, VA0, NA0)
// are FP32 (fast quotients): they only perturb the flux's input at FP32
// rounding, below the FP32 transforms' own error.  The frequency is FP64:
// the phase needs it to FP64 accuracy (the angle omega t reaches hundreds of
// radians over a run).
struct CoefRowF {
    double om, NA0, A0Vort, A0Buoy, ApmN, ApmDScaled;
    cx<double> ApmD;
    float NApF, NA0F;
    cx<float> UAp, VAp, WAp, UA0, VA0;
};
__device__ __forceinline__ CoefRowF coefficientsF(const CoefConsts& c, int j, const ModeDev& md) {
    // a / b as a times the hardware reciprocal of b (about 2 ulp)
    auto dv = [](float a, float b) { return a * __builtin_amdgcn_rcpf(b); };
    const int Nj = c.NjMixed;
    auto pj = [&](int row) { return c.mixedJ[row * Nj + j]; };
    CoefRowF r{};
    const double Kh = md.Kh, Kh2 = Kh * Kh, Mj = c.m[j];
    const bool wave = Kh > 0.0 && j > 0, inertial = Kh == 0.0, geostrophic = Kh > 0.0, mda = Kh == 0.0 && j > 0;
    double hpm = 1.0;
    if (j > 0)
        hpm = c.stratified ? c.h0[j] : c.hydro ? c.N02 / (c.g * Mj * Mj) : (c.N02 - c.f2) / (c.g * (Mj * Mj + Kh2));
    const double om = sqrt(c.g * hpm * Kh2 + c.f2);
    r.om = om;
    r.ApmDScaled = pj(kMjApmDScaled);
    // FP32 synthesis tables: fs = Fg / (Fg / Fw) = Fw
    const float Khf = static_cast<float>(Kh), hpmf = static_cast<float>(hpm), omf = static_cast<float>(om);
    const float Fg = static_cast<float>(c.Fg[j]), f = static_cast<float>(c.f);
    const float sign = j % 2 == 0 ? 1.0f : -1.0f;
    const float fs = c.hydro ? Fg : j == 0 ? 2.0f : sign * hpmf * static_cast<float>(Mj) * static_cast<float>(c.S);
    const float cosA = static_cast<float>(md.cosA), sinA = static_cast<float>(md.sinA);
    if (wave) {
        const float gs = static_cast<float>(c.gws[j]);
        const float fo = dv(f, omf);
        r.UAp = {cosA * fs, -fo * sinA * fs};
        r.VAp = {sinA * fs, fo * cosA * fs};
        r.WAp = {0.0f, -Khf * hpmf * gs};
        r.NApF = dv(-Khf * hpmf, omf) * gs;
        const double twoKh = 2.0 * Kh * hpm;
        r.ApmD = {0.0, -pj(kMjDeltaScale) / twoKh};
        r.ApmN = (-om / twoKh) * pj(kMjGwgOverGg);
    } else if (inertial) {
        r.UAp = {fs, 0.0f};
        r.VAp = {0.0f, fs};
    }
    if (geostrophic) {
        const double den = Kh2 + pj(kMjLr2inv);
        const float rden = __builtin_amdgcn_rcpf(static_cast<float>(den));
        r.UA0 = {0.0f, static_cast<float>(md.l) * rden * Fg};
        r.VA0 = {0.0f, -static_cast<float>(md.k) * rden * Fg};
        r.NA0 = pj(kMjNA0Num) / den;
        r.A0Vort = pj(kMjA0VortGeo);
        r.A0Buoy = pj(kMjA0BuoyGeo);
    } else if (mda) {
        r.NA0 = c.Gg[j];
        r.A0Buoy = pj(kMjA0BuoyMda);
        r.A0Vort = pj(kMjA0VortMda);
    }
    r.NA0F = static_cast<float>(r.NA0);
    return r;
}

// exp(i omega elapsed) in FP32: the angle formed and reduced modulo 2 pi in
// FP64 (two-part 2 pi, Cody-Waite), then the FP32 sine and cosine of the
// reduced angle.
__device__ __forceinline__ cx<float> phaseF(double omega, double elapsed) {
    constexpr double kTwoPiHi = 6.283185307179586232, kTwoPiLo = 2.4492935982947064e-16;
    const double a = omega * elapsed;
    const double q = rint(a * (1.0 / kTwoPiHi));
    const double r = fma(-q, kTwoPiLo, fma(-q, kTwoPiHi, a));
    float sn, cs;
    sincosf(static_cast<float>(r), &sn, &cs);
    return cx<float>{cs, sn};
}

// storeFieldSpectra with the FP32 tables.
__device__ __forceinline__ void storeFieldSpectraF(const CoefRowF& cr, int idx, int M, cx<double> Apd, cx<double> Amd,
                                                   cx<double> A0d, cx<float> p, cx<float>* __restrict__ FV) {
    auto f = [](cx<double> a) { return cx<float>{static_cast<float>(a.re), static_cast<float>(a.im)}; };
    const cx<float> ap = cmul(f(Apd), p), am = cmul(f(Amd), cconj(p)), a0 = f(A0d);
    const cx<float> u = cadd(cadd(cmul(cr.UAp, ap), cmul(cconj(cr.UAp), am)), cmul(cr.UA0, a0));
    const cx<float> v = cadd(cadd(cmul(cr.VAp, ap), cmul(cconj(cr.VAp), am)), cmul(cr.VA0, a0));
    const cx<float> w = cadd(cmul(cr.WAp, ap), cmul(cr.WAp, am));
    const cx<float> e = cadd(cadd(cscale(ap, cr.NApF), cscale(am, -cr.NApF)), cscale(a0, cr.NA0F));
    FV[idx] = u;
    FV[M + idx] = v;
    FV[2 * M + idx] = w;
    FV[3 * M + idx] = e;
}

// Evolved field spectra of (Ap, Am, A0) at the given elapsed time, with the
// coefficient tables formed by coefficients() -- the same values k_stage
// forms when it emits the next evaluation's spectra, so that an RK4 run gives
// the same bits however its steps are split over calls.
template <typename T>
__global__ __launch_bounds__(256) void k_fields(const cx<double>* __restrict__ Ap, const cx<double>* __restrict__ Am,
                                                const cx<double>* __restrict__ A0,
                                                const ModeDev* __restrict__ modes, int Nj, CoefConsts cc,
                                                double elapsed, int M, cx<T>* __restrict__ FV) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M) return;
    const int wv = idx / Nj, j = idx - wv * Nj;
    if constexpr (sizeof(T) == 4) {
        const CoefRowF cr = coefficientsF(cc, j, modes[wv]);
        storeFieldSpectraF(cr, idx, M, Ap[idx], Am[idx], A0[idx], phaseF(cr.om, elapsed), FV);
    } else {
        const CoefRow cr = coefficients(cc, j, modes[wv]);
        storeFieldSpectra(cr.UAp, cr.VAp, cr.WAp, cr.NAp, cr.UA0, cr.VA0, cr.NA0, idx, M, Ap[idx], Am[idx], A0[idx],
                          phaseAt(cr.om, elapsed), FV);
    }
}

// The flux store or RK4 stage of WVFixedStepRK4::step for one coefficient,
// given its flux k3 = (Fp, Fm, F0); next = the state it produces (RK modes).
// The weighted flux sum k1 + 2 k2 + 2 k3 + k4 is stored in the flux's own
// precision T (the fluxes of the mixed path carry FP32 accuracy; the state
// update y + h/6 w is FP64 in both).  Stage arithmetic as scaledSum
// (value + scale*increment, WVRungeKutta.cpp:15-19), rounded per operation as
// on the CPU.
template <typename T>
__device__ __forceinline__ void rkStageUpdate(const Epilogue& ep, int idx, const cx<double> (&k3)[3],
                                              cx<double> (&next)[3]) {
    cx<double>* outs[3] = {ep.out0, ep.out1, ep.out2};
    cx<double>* ys[3] = {ep.y0, ep.y1, ep.y2};
    cx<double>* sts[3] = {ep.st0, ep.st1, ep.st2};
    cx<T>* ws[3] = {reinterpret_cast<cx<T>*>(ep.w0), reinterpret_cast<cx<T>*>(ep.w1), reinterpret_cast<cx<T>*>(ep.w2)};
    auto wload = [&](int a) {
        const cx<T> v = ws[a][idx];
        return cx<double>{static_cast<double>(v.re), static_cast<double>(v.im)};
    };
    auto wstore = [&](int a, cx<double> v) { ws[a][idx] = {static_cast<T>(v.re), static_cast<T>(v.im)}; };
    auto axpy = [](cx<double> y, cx<double> x, double s) {
        return cx<double>{addRn(y.re, mulRn(s, x.re)), addRn(y.im, mulRn(s, x.im))};
    };
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const cx<double> kv = k3[a];
        switch (ep.mode) {
            case kFluxOut: outs[a][idx] = kv; break;
            case kRk1:
                wstore(a, kv);
                next[a] = axpy(ys[a][idx], kv, ep.scale);
                if (!ep.emitFields) sts[a][idx] = next[a];
                break;
            case kRk2:
            case kRk3:
                wstore(a, axpy(wload(a), kv, 2.0));
                next[a] = axpy(ys[a][idx], kv, ep.scale);
                if (!ep.emitFields) sts[a][idx] = next[a];
                break;
            default: {
                const cx<double> w = axpy(wload(a), kv, 1.0);
                next[a] = axpy(ys[a][idx], w, ep.scale);
                ys[a][idx] = next[a];
                break;
            }
        }
    }
}

// Wave-vortex projection of one target's analysed value (projectFluxTargetMode,
// ...Kernel.cpp) added to (Fp, Fm, F0) with phase p, in arithmetic R; the
// tables in cr (CoefRow or CoefRowF).
template <typename R, int NT, typename Row>
__device__ __forceinline__ void projectTarget(int target, cx<R> value, const Row& cr, const ModeDev& md, R apmW,
                                              R inertialScale, bool inertial, cx<R> p, cx<R>& Fp, cx<R>& Fm,
                                              cx<R>& F0) {
    constexpr bool hydroT = NT == 3;  // hydrostatic: ApmD, else ApmDScaled
    const R k = static_cast<R>(md.k), l = static_cast<R>(md.l);
    cx<R> a0{R(0), R(0)};
    if (target == 0) a0 = cscale(cmuli(value, -l), static_cast<R>(cr.A0Vort));
    if (target == 1) a0 = cscale(cmuli(value, k), static_cast<R>(cr.A0Vort));
    if (target == 3) a0 = cscale(value, static_cast<R>(cr.A0Buoy));
    const cx<R> buoyValue = csub(target == 3 ? value : cx<R>{R(0), R(0)}, cscale(a0, static_cast<R>(cr.NA0)));
    const cx<R> buoy = cscale(buoyValue, static_cast<R>(cr.ApmN));
    cx<R> wave{R(0), R(0)};
    if constexpr (hydroT) {
        const cx<R> apmd{static_cast<R>(cr.ApmD.re), static_cast<R>(cr.ApmD.im)};
        if (target == 0) wave = cmul(cmuli(value, k), apmd);
        if (target == 1) wave = cmul(cmuli(value, l), apmd);
    } else {
        const R dsc = static_cast<R>(cr.ApmDScaled);
        if (target == 0) wave = cscale(value, static_cast<R>(md.cosA) * dsc);
        if (target == 1) wave = cscale(value, static_cast<R>(md.sinA) * dsc);
        if (target == 2) wave = cmuli(value, (static_cast<R>(md.Kh) / R(2)) * apmW);
    }
    cx<R> cAp = cadd(wave, buoy), cAm = csub(wave, buoy);
    if (inertial) {
        cx<R> a{R(0), R(0)};
        if (target == 0) a = cscale(value, inertialScale);
        if (target == 1) a = cscale(cmuli(value, R(-1)), inertialScale);
        cAp = a;
        cAm = cconj(a);
    }
    Fp = cadd(Fp, cmul(cAp, cconj(p)));
    Fm = cadd(Fm, cmul(cAm, p));
    F0 = cadd(F0, a0);
}

// Wave-vortex projection of the analysed products tv (projectFluxTargetMode,
// ...Kernel.cpp) of coefficient idx = j + Nj wv and the flux store or RK4
// stage of WVFixedStepRK4::step; in RK modes with emitFields set it also
// forms the field spectra of the state it produces (the input of the next
// flux evaluation).  FP64 (T = double): tables and phase in FP64, the
// tables bit for bit the host's; mixed (T = float): the tables of
// coefficientsF and the FP32 phase of phaseF.  The projection and the stage
// arithmetic on the state are FP64 in both.
template <typename T, int NT>
__device__ __forceinline__ void stageCoefficient(const cx<T> (&tv)[NT], int idx, int j, int wv,
                                                 const ModeDev* __restrict__ modes, int M, const CoefConsts& cc,
                                                 double elapsed, double hscale, const Epilogue& ep,
                                                 cx<T>* __restrict__ FV) {
    const ModeDev md = modes[wv];
    const bool inertial = wv == 0;
    // the coefficient tables, formed here instead of read from memory
    using Row = typename std::conditional<sizeof(T) == 4, CoefRowF, CoefRow>::type;
    Row cr;
    cx<T> p;
    if constexpr (sizeof(T) == 4) {
        cr = coefficientsF(cc, j, md);
        p = phaseF(cr.om, elapsed);
    } else {
        cr = coefficients(cc, j, md);
        p = phaseAt(cr.om, elapsed);
    }
    // the projection in FP64 in both precisions (its terms cancel to a
    // small remainder for a nearly balanced flow)
    auto d = [](cx<T> a) { return cx<double>{static_cast<double>(a.re), static_cast<double>(a.im)}; };
    const cx<double> pd = d(p);
    cx<double> Fp{0.0, 0.0}, Fm{0.0, 0.0}, F0{0.0, 0.0};
    const double apmW = cc.apmW[j], iscale = cc.inertialScale[j];
#pragma unroll
    for (int t = 0; t < NT; ++t)
        projectTarget<double, NT>(projTarget<NT>(t), cscale(d(tv[t]), hscale), cr, md, apmW, iscale, inertial, pd, Fp,
                                  Fm, F0);
    const cx<double> k3[3] = {Fp, Fm, F0};
    cx<double> next[3];
    rkStageUpdate<T>(ep, idx, k3, next);
    if (ep.mode != kFluxOut && ep.emitFields) {
        // stages 2 and 4 produce the state of their own time: the same phase
        if constexpr (sizeof(T) == 4) {
            const cx<float> pn = ep.nextElapsed != elapsed ? phaseF(cr.om, ep.nextElapsed) : p;
            storeFieldSpectraF(cr, idx, M, next[0], next[1], next[2], pn, FV);
        } else {
            const cx<double> pn = ep.nextElapsed != elapsed ? phaseAt(cr.om, ep.nextElapsed) : p;
            storeFieldSpectra(cr.UAp, cr.VAp, cr.WAp, cr.NAp, cr.UA0, cr.VA0, cr.NA0, idx, M, next[0], next[1],
                              next[2], pn, FV);
        }
    }
}

// stageCoefficient over the analysed products Tg[t][idx], one coefficient
// (j, mode) per thread.
template <typename T, int NT>
__global__ __launch_bounds__(256) void k_stage(const cx<T>* __restrict__ Tg, const ModeDev* __restrict__ modes,
                                               int Nj, int M, CoefConsts cc, double elapsed, double hscale,
                                               Epilogue ep, cx<T>* __restrict__ FV) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M) return;
    const int wv = idx / Nj, j = idx - wv * Nj;
    cx<T> tv[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) tv[t] = Tg[t * static_cast<size_t>(M) + idx];
    stageCoefficient<T, NT>(tv, idx, j, wv, modes, M, cc, elapsed, hscale, ep, FV);
}

// r2c along x of nf real FP64 fields [Nx,Ny,Nz,nf] into P[c][z][y][kx < nkx]
// (the horizontal forward plans of the forward transforms).  A workgroup
// takes rpb consecutive (y, z) rows, their signals paired as in k_xinv.
template <typename T, int L2N = 0>
__global__ __launch_bounds__(256) void k_xfwd(const double* __restrict__ in, cx<T>* __restrict__ P, int nf, int rpb,
                                              int NxArg, int log2NxArg, int Ny, int Nz, int nkx, int nkxP,
                                              const cx<T>* __restrict__ twX, int adjointWeights) {
    const int Nx = L2N > 0 ? 1 << (L2N > 0 ? L2N : 0) : NxArg;  // L2N > 0: the length fixed at compile time
    const int log2Nx = L2N > 0 ? L2N : log2NxArg;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nrows = Ny * Nz, row0 = static_cast<int>(blockIdx.x) * rpb;
    const int nsig = rpb * nf, ntr = (nsig + 1) / 2;
    const int NxP = ldsLength(Nx);
    const size_t R = static_cast<size_t>(Nx) * Ny * Nz;
    auto value = [&](int sg, int x) -> T {
        const int row = row0 + sg / nf;
        if (sg >= nsig || row >= nrows) return T(0);
        return T(in[(sg % nf) * R + static_cast<size_t>(row) * Nx + x]);
    };
    for (int e = threadIdx.x; e < ntr * Nx; e += blockDim.x) {
        const int q = e / Nx, x = e % Nx;
        buf[q * NxP + ldsIndex(x, 1)] = {value(2 * q, x), value(2 * q + 1, x)};
    }
    __syncthreads();
    fftLdsBatched<T, -1, L2N>(buf, Nx, log2Nx, ntr, NxP, twX);
    const size_t chStride = static_cast<size_t>(Nz) * Ny * nkxP;
    for (int e = threadIdx.x; e < ntr * nkx; e += blockDim.x) {
        const int q = e / nkx, k = e % nkx;
        const cx<T> Za = buf[q * NxP + ldsIndex(k, 1)];
        const cx<T> Zb = buf[q * NxP + ldsIndex(k == 0 ? 0 : Nx - k, 1)];
        // adjointWeights: the transpose of the Hermitian c2r (twice the r2c
        // for kx > 0; kx = 0 is real either way)
        const T h = adjointWeights && k > 0 ? T(1) : T(0.5);
        const int sa = 2 * q, sb = 2 * q + 1;
        if (row0 + sa / nf < nrows)
            P[(sa % nf) * chStride + static_cast<size_t>(row0 + sa / nf) * nkxP + k] =
                {(Za.re + Zb.re) * h, (Za.im - Zb.im) * h};
        if (sb < nsig && row0 + sb / nf < nrows)
            P[(sb % nf) * chStride + static_cast<size_t>(row0 + sb / nf) * nkxP + k] =
                {(Za.im + Zb.im) * h, (Zb.re - Za.re) * h};
    }
}

// Vertical analysis of the forward transforms (MFMA, as k_proj) followed by
// projectNonhydrostaticMode / projectHydrostaticMode (...Kernel.cpp:299-370):
// channels (U, V, W, N) nonhydrostatic or (U, V, N) hydrostatic.
template <typename T, int NF>
__global__ __launch_bounds__(256) void k_projfwd(const cx<T>* __restrict__ Q, const ColumnDev* __restrict__ cols,
                                                 int ncol, const T* __restrict__ anaC, const T* __restrict__ anaS,
                                                 int Ja, int Za, int Nz, int Nj, int Ny, int nkxP, ProjTables tb,
                                                 double elapsed, double hscale, cx<double>* __restrict__ outAp,
                                                 cx<double>* __restrict__ outAm, cx<double>* __restrict__ outA0) {
    using MF = Mfma<T>;
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + waveIndex();
    const int ncg = (ncol + 7) / 8;
    const int njt = Ja / 16;
    if (wave >= ncg * njt) return;
    // consecutive waves share a column group, so its Q columns stay in L2
    const int jt = wave % njt;
    const int cg = wave / njt;
    const int n = lane & 15, kq = lane >> 4;
    const int col = cg * 8 + (n >> 1);
    const int part = n & 1;
    const bool colValid = col < ncol;
    ColumnDev cd{};
    if (colValid) cd = cols[col];
    const size_t tStride = static_cast<size_t>(Nz) * Ny * nkxP;
    const size_t colOff = static_cast<size_t>(cd.iy) * nkxP + cd.kx;
    typename MF::acc acc[NF];
    const int nsteps = Za / 4;
#pragma unroll
    for (int c = 0; c < NF; ++c) {
        const T* A = c < 2 ? anaC : anaS;  // U, V cosine family; W, N sine family
        typename MF::acc a4 = MF::zero();
        for (int s = 0; s < nsteps; ++s) {
            const int z = 4 * s + kq;
            T b = T(0);
            if (colValid && z < Nz) {
                const cx<T> q = Q[c * tStride + static_cast<size_t>(z) * Ny * nkxP + colOff];
                b = part ? q.im : q.re;
            }
            a4 = MF::run(A[(static_cast<size_t>(jt) * nsteps + s) * 64 + lane], b, a4);
        }
        acc[c] = a4;
    }
    double cre[NF][4], cim[NF][4];
#pragma unroll
    for (int c = 0; c < NF; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double mine = static_cast<double>(acc[c][i]);
            const double other = __shfl_xor(mine, 1);
            cre[c][i] = part ? other : mine;
            cim[c][i] = part ? mine : other;
        }
    if (!colValid || (cd.flags & kColCompletion)) return;
    const bool conj = (cd.flags & kColConj) != 0;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
        const int i = 2 * part + ii;
        const int j = 16 * jt + MF::row(kq, i);
        if (j >= Nj) continue;
        const size_t idx = static_cast<size_t>(j) + static_cast<size_t>(Nj) * static_cast<size_t>(cd.wv);
        cx<double> v[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) v[c] = cscale(cx<double>{cre[c][i], conj ? -cim[c][i] : cim[c][i]}, hscale);
        const cx<double> U = v[0], V = v[1], N = v[NF - 1];
        const cx<double> vort = csub(cmuli(V, cd.k), cmuli(U, cd.l));
        const cx<double> A0 = cadd(cscale(vort, tb.A0Vort[idx]), cscale(N, tb.A0Buoy[idx]));
        cx<double> cAp, cAm;
        if (cd.wv == 0) {
            const cx<double> a = cscale(csub(U, cmuli(V, 1.0)), tb.inertialScale[j]);
            cAp = a;
            cAm = cconj(a);
        } else {
            cx<double> wave;
            if (NF == 3) {
                wave = cmul(tb.ApmD[idx], cadd(cmuli(U, cd.k), cmuli(V, cd.l)));
            } else {
                const cx<double> delta = cscale(cadd(cscale(U, cd.cosA), cscale(V, cd.sinA)), tb.ApmDScaled[idx]);
                wave = cadd(delta, cmuli(v[2], (cd.Kh / 2.0) * tb.apmW[j]));
            }
            const cx<double> buoy = cscale(csub(N, cscale(A0, tb.NA0[idx])), tb.ApmN[idx]);
            cAp = cadd(wave, buoy);
            cAm = csub(wave, buoy);
        }
        double sn, cs;
        sincos(tb.omega[idx] * elapsed, &sn, &cs);
        const cx<double> p{cs, sn};
        outAp[idx] = cmul(cAp, cconj(p));
        outAm[idx] = cmul(cAm, p);
        outA0[idx] = A0;
    }
}

// Deterministic two-pass FP64 energy reduction.
__global__ __launch_bounds__(256) void k_energy_partial(const cx<double>* __restrict__ Ap,
                                                        const cx<double>* __restrict__ Am,
                                                        const cx<double>* __restrict__ A0,
                                                        const double* __restrict__ ApmTE,
                                                        const double* __restrict__ A0TE, int M,
                                                        double* __restrict__ partial) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        const cx<double> p = Ap[i], m = Am[i], z = A0[i];
        acc += ApmTE[i] * ((p.re * p.re + p.im * p.im) + (m.re * m.re + m.im * m.im)) + A0TE[i] * (z.re * z.re + z.im * z.im);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_energy_final(const double* __restrict__ partial, int n,
                                                      double* __restrict__ out) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// Scratch of the allocation-free energy entry point (wv_total_energy_no_alloc_device):
// the partial sums of its reduction, one per workgroup of k_energy_partial.
// A file-scope device array, so each GPU holds one copy for the life of the
// process; host code reaches it only through the per-device address taken
// with hipGetSymbolAddress at create.
constexpr int kEnergyPartials = 1024;
static __device__ __align__(256) double scratch[kEnergyPartials];

#include "wv_adjoint_kernels.hpp"

// ------------------------------------------------------------------- host --
namespace {

thread_local std::string gLastError;

struct Failure {
    wv_status code;
    std::string message;
};

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Failure{WV_FFT_EXECUTION_FAILURE, std::string(what) + ": " + hipGetErrorString(e)};
}

int log2Exact(size_t n) {
    if (n == 0 || (n & (n - 1)) != 0) return -1;
    int b = 0;
    while ((size_t(1) << b) < n) ++b;
    return b;
}

// prime factors of n all at most 7 (the lengths of the kSmooth7 kernels)
bool smooth7(size_t n) {
    if (n == 0) return false;
    for (size_t p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

template <typename T>
size_t csize() { return sizeof(cx<T>); }

struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t n) {
        bytes = n;
        if (n == 0) n = 16;
        if (hipMalloc(&p, n) != hipSuccess) {
            p = nullptr;
            throw Failure{WV_ALLOCATION_FAILURE, "Unable to allocate device workspace."};
        }
    }
    template <typename U>
    U* as() const { return static_cast<U*>(p); }
};

struct StageRecord {
    std::string name;
    double totalMs = 0.0;
    uint64_t launches = 0;
    double totalBytes = 0.0;  // algorithmic bytes summed over the timed launches
};

}  // namespace

struct KernelImpl {
    Descriptor d;
    wv_precision precision = WV_PRECISION_FP64;
    int device = 0;
    int NT = 4, NCH = 15, NH = 12, targetCode = 0, KT = 16;  // targets, spatial and half-spectrum channels
    size_t ncol = 0;
    size_t tsize = 16;  // bytes of one field complex value
    std::atomic<bool> executing{false};

    DeviceBuffer cols, omega, UAp, VAp, WAp, NAp, UA0, VA0, NA0, A0Vort, A0Buoy, ApmN, ApmDScaled, ApmD, ApmTE, A0TE,
        apmW, inertialScale, fWaveScale, gWaveScale, Fg, Gg;
    DeviceBuffer synC, synS, anaC, anaS, synGz, synFh, synEz, twX, twY, mj, kxv, zeroRow;
    // even / odd j parts of the DCT-I, DST-I synthesis (rows z <= (Nz-1)/2) and
    // of the analysis (columns z <= (Nz-1)/2): constant N
    DeviceBuffer synCE, synSE, synCO, synSO, anaCE, anaCO, anaSE, anaSO;
    DeviceBuffer synCmE, synSmE, synCmO, synSmO;  // folded, columns times +/- m_j / 2
    // split-FP16 (hi, lo) forms of the eight folded synthesis operators, in
    // the order of the folded launch's slots, and their scales
    DeviceBuffer syn16[kSynthMatSlots];
    float syn16Inv[kSynthMatSlots] = {};
    int syn16N32 = 0, syn16Tail = 0;  // contraction steps of 32 and 16 (0, 0: not available)
    // the same for the four folded analysis operators (anaCE, anaSE, anaCO, anaSO)
    DeviceBuffer ana16[4];
    float ana16Inv[4] = {};
    int ana16N32 = 0, ana16Tail = 0;
    bool fold = false;
    size_t Zh = 0;  // folded z rows (multiple of 16)
    DeviceBuffer FV, H, G, P, Q, Tg, modeTab, jTab, mixedJTab, stage, weighted, partial, scalar;
    CoefConsts coefConsts{};
    DeviceBuffer rowLimit;  // per y-FFT kx tile: largest retained |ky|
    DeviceBuffer lrow;      // per storage row iy: its y wavenumber (rowSign * l of its columns)
    bool lrowValid = false;
    double retainedRowFraction = 1.0;
    size_t maxLdsPerBlock = 65536;

    // counters of the reference's WVKernelMetrics (...Kernel.hpp:12-45)
    wv_kernel_metrics metrics{};
    uint64_t rkStepCount = 0;
    double lastStepSize = 0.0;
    bool scalarPrepared = false;

    // Calls on one handle share its workspaces.  Each call records lastCall
    // on its stream once its work is enqueued; a call on another stream
    // first makes that stream wait for it, so calls stay ordered on the
    // device whatever streams they come from.
    hipEvent_t lastCall = nullptr;
    hipStream_t lastStream = nullptr;
    bool lastCallValid = false;
    void orderAfterLastCall(hipStream_t s) {
        if (lastCallValid && s != lastStream) hipCheck(hipStreamWaitEvent(s, lastCall, 0), "hipStreamWaitEvent");
    }
    void recordCall(hipStream_t s) {
        if (!lastCall) hipCheck(hipEventCreateWithFlags(&lastCall, hipEventDisableTiming), "hipEventCreate");
        hipCheck(hipEventRecord(lastCall, s), "hipEventRecord");
        lastStream = s;
        lastCallValid = true;
    }

    // Test hook (the counterpart of the reference tests' failing FFT engine,
    // PortableRuntime/tests/TestWVForcingAndRK4.cpp:313-320): when set to
    // n > 0, the n-th next timed launch is not made and throws
    // WV_FFT_EXECUTION_FAILURE on the host.
    long failAfterLaunches = 0;

    bool instrument = false;
    std::vector<StageRecord> records;
    struct Pending {
        int record;
        double bytes;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;

    static size_t sumBytes(std::initializer_list<const DeviceBuffer*> bs) {
        size_t s = 0;
        for (const DeviceBuffer* b : bs) s += b->bytes;
        return s;
    }
    // descriptor tables and column / mode maps on the device
    size_t descriptorBytes() const {
        return sumBytes({&rowLimit, &lrow, &cols, &omega, &UAp, &VAp, &WAp, &NAp, &UA0, &VA0, &NA0, &A0Vort, &A0Buoy,
                         &ApmN, &ApmDScaled, &ApmD, &ApmTE, &A0TE, &apmW, &inertialScale, &fWaveScale, &gWaveScale,
                         &Fg, &Gg, &modeTab, &jTab, &mixedJTab, &mainColumn, &compColumn, &scKeep});
    }
    // transform operators and twiddle tables (the "plans")
    size_t planBytes() const {
        size_t s = sumBytes({&synC, &synS, &anaC, &anaS, &synGz, &synFh, &synEz, &twX, &twY, &mj, &zeroRow, &kxv,
                             &synCE, &synSE, &synCO, &synSO, &synCmE, &synSmE, &synCmO, &synSmO, &anaCE, &anaCO,
                             &anaSE, &anaSO, &anaTC, &anaTS, &scD, &scTwX, &scTwY});
        for (const DeviceBuffer& b : synT) s += b.bytes;
        return s;
    }
    // workspaces: half spectra, coefficient work arrays, reductions
    size_t scratchBytes() const {
        return sumBytes({&FV, &H, &G, &P, &Q, &Tg, &stage, &weighted, &partial, &scalar, &scC, &Qs, &Gs, &colbuf,
                         &rkWork});
    }
    size_t persistentBytes() const { return descriptorBytes() + planBytes() + scratchBytes(); }

    int recordIndex(const char* name) {
        for (size_t i = 0; i < records.size(); ++i)
            if (records[i].name == name) return static_cast<int>(i);
        records.push_back({name, 0.0, 0, 0.0});
        return static_cast<int>(records.size() - 1);
    }

    template <typename Launch>
    void timed(const char* name, double bytes, hipStream_t s, Launch&& launch) {
        if (failAfterLaunches > 0 && --failAfterLaunches == 0)
            throw Failure{WV_FFT_EXECUTION_FAILURE, std::string("Injected failure before launch: ") + name};
        if (!instrument) {
            launch();
            hipCheck(hipGetLastError(), name);
            return;
        }
        Pending pd{recordIndex(name), bytes, nullptr, nullptr};
        hipCheck(hipEventCreate(&pd.a), "hipEventCreate");
        hipCheck(hipEventCreate(&pd.b), "hipEventCreate");
        hipCheck(hipEventRecord(pd.a, s), "hipEventRecord");
        launch();
        hipCheck(hipGetLastError(), name);
        hipCheck(hipEventRecord(pd.b, s), "hipEventRecord");
        pending.push_back(pd);
    }

    void drainTimings() {
        for (auto& pd : pending) {
            hipCheck(hipEventSynchronize(pd.b), "hipEventSynchronize");
            float ms = 0.f;
            hipCheck(hipEventElapsedTime(&ms, pd.a, pd.b), "hipEventElapsedTime");
            records[pd.record].totalMs += ms;
            records[pd.record].launches += 1;
            records[pd.record].totalBytes += pd.bytes;
            (void)hipEventDestroy(pd.a);
            (void)hipEventDestroy(pd.b);
        }
        pending.clear();
    }

    // ---- setup ----------------------------------------------------------
    template <typename U>
    void upload(DeviceBuffer& b, const std::vector<U>& v) {
        b.alloc(v.size() * sizeof(U));
        if (!v.empty()) hipCheck(hipMemcpy(b.p, v.data(), v.size() * sizeof(U), hipMemcpyHostToDevice), "upload");
    }
    template <typename T>
    void uploadAs(DeviceBuffer& b, const std::vector<double>& v) {
        std::vector<T> tmp(v.begin(), v.end());
        upload(b, tmp);
    }

    // MFMA A operands in [tile][step][lane] order: lane l of step s in tile r
    // holds M[16 r + l%16][4 s + l/16] of a row-major [rows x cols] matrix.
    static std::vector<double> mfmaOperand(const std::vector<double>& m, size_t rows, size_t cols) {
        const size_t tiles = rows / 16, steps = cols / 4;
        std::vector<double> out(tiles * steps * 64);
        for (size_t r = 0; r < tiles; ++r)
            for (size_t st = 0; st < steps; ++st)
                for (size_t l = 0; l < 64; ++l) out[(r * steps + st) * 64 + l] = m[(16 * r + l % 16) * cols + 4 * st + l / 16];
        return out;
    }

    // The folded synthesis operators in split-FP16 form for k_synth16: per
    // matrix (rows z <= half of tile 16 zt + l%16, folded columns jh), per
    // tile: N32 steps of [hi | lo][64 lanes][8] (jh = 32 s + 8 (l/16) + t)
    // then, with a tail, [hi | lo][64][4] (jh = 32 N32 + 4 (l/16) + t);
    // columns past Jh are zero.  The even and odd matrices of a pair share
    // one power-of-two scale, so that their sums share it.
    void setupSplit16Synthesis(size_t half, size_t Jh) {
        const size_t rem = Jh % 32;
        const int n32 = static_cast<int>(Jh / 32 + (rem > 16 ? 1 : 0)), tail = rem > 0 && rem <= 16 ? 1 : 0;
        if (n32 > 2) return;  // longer contractions stay on the FP32 operators
        syn16N32 = n32;
        syn16Tail = tail;
        struct Src {
            const std::vector<double>* m;
            size_t par;
            double scale;
        };
        const Src src[kSynthMatSlots] = {{&d.synC, 0, 0.0}, {&d.synS, 0, 0.0}, {&d.synC, 1, 0.0}, {&d.synS, 1, 0.0},
                                         {&d.synC, 0, 0.5}, {&d.synS, 0, -0.5}, {&d.synC, 1, 0.5}, {&d.synS, 1, -0.5}};
        auto entry = [&](int slot, size_t z, size_t jh) {
            if (z > half || jh >= Jh) return 0.0;
            const size_t j = 2 * jh + src[slot].par;
            const double a = (*src[slot].m)[z * d.Js + j];
            return src[slot].scale == 0.0 ? a : j < d.c.Nj ? a * (src[slot].scale * d.m[j]) : 0.0;
        };
        const size_t tiles = Zh / 16, tileHalves = static_cast<size_t>(n32) * 1024 + static_cast<size_t>(tail) * 512;
        for (int pair = 0; pair < kSynthMatSlots; ++pair) {
            if (pair % 4 >= 2) continue;  // slots m and m + 2 (even, odd) are prepared together
            float mx = 0.0f;
            for (int slot : {pair, pair + 2})
                for (size_t z = 0; z <= half; ++z)
                    for (size_t jh = 0; jh < Jh; ++jh) mx = std::max(mx, static_cast<float>(std::fabs(entry(slot, z, jh))));
            const int e = split16Exponent(mx);
            const float sc = std::ldexp(1.0f, e);
            for (int slot : {pair, pair + 2}) {
                std::vector<_Float16> out(tiles * tileHalves);
                auto put = [&](size_t base, size_t z, size_t jh) {
                    const float v = static_cast<float>(entry(slot, z, jh)) * sc;
                    const _Float16 hi = static_cast<_Float16>(v);
                    out[base] = hi;
                    return static_cast<_Float16>(v - static_cast<float>(hi));
                };
                for (size_t zt = 0; zt < tiles; ++zt) {
                    _Float16* tile = out.data() + zt * tileHalves;
                    for (int st = 0; st < n32; ++st)
                        for (size_t l = 0; l < 64; ++l)
                            for (size_t t = 0; t < 8; ++t) {
                                const size_t o = static_cast<size_t>(st) * 1024 + l * 8 + t;
                                tile[o + 512] = put(zt * tileHalves + o, 16 * zt + l % 16, 32 * st + 8 * (l / 16) + t);
                            }
                    if (tail)
                        for (size_t l = 0; l < 64; ++l)
                            for (size_t t = 0; t < 4; ++t) {
                                const size_t o = static_cast<size_t>(n32) * 1024 + l * 4 + t;
                                tile[o + 256] = put(zt * tileHalves + o, 16 * zt + l % 16, 32 * n32 + 4 * (l / 16) + t);
                            }
                }
                upload(syn16[slot], out);
                syn16Inv[slot] = std::ldexp(1.0f, -e);
            }
        }
    }

    // The folded analysis operators in split-FP16 form for k_proj16: per row
    // tile (rows jh = 16 tile + l%16 of the even- or odd-j part), N32 steps
    // of [hi | lo][64][8] (z = 32 s + 8 (l/16) + t) then, with a tail,
    // [hi | lo][64][4] (z = 32 N32 + 4 (l/16) + t); one scale per matrix.
    void setupSplit16Analysis(size_t half, size_t Jah) {
        const int n32 = static_cast<int>(Zh / 32), tail = Zh % 32 ? 1 : 0;
        if (n32 < 1 || n32 > 2 || Jah % 16) return;  // other lengths stay on the FP32 operators
        ana16N32 = n32;
        ana16Tail = tail;
        const std::vector<double>* src[4] = {&d.anaC, &d.anaS, &d.anaC, &d.anaS};
        const size_t tiles = Jah / 16, tileHalves = static_cast<size_t>(n32) * 1024 + static_cast<size_t>(tail) * 512;
        for (int slot = 0; slot < 4; ++slot) {
            const size_t par = slot / 2;
            auto entry = [&](size_t jh, size_t z) { return z > half ? 0.0 : (*src[slot])[(2 * jh + par) * d.Za + z]; };
            float mx = 0.0f;
            for (size_t jh = 0; jh < Jah; ++jh)
                for (size_t z = 0; z <= half; ++z) mx = std::max(mx, static_cast<float>(std::fabs(entry(jh, z))));
            const int e = split16Exponent(mx);
            const float sc = std::ldexp(1.0f, e);
            std::vector<_Float16> out(tiles * tileHalves);
            auto put = [&](size_t o, size_t jh, size_t z, size_t loOff) {
                const float v = static_cast<float>(entry(jh, z)) * sc;
                const _Float16 hi = static_cast<_Float16>(v);
                out[o] = hi;
                out[o + loOff] = static_cast<_Float16>(v - static_cast<float>(hi));
            };
            for (size_t jt = 0; jt < tiles; ++jt) {
                const size_t base = jt * tileHalves;
                for (int st = 0; st < n32; ++st)
                    for (size_t l = 0; l < 64; ++l)
                        for (size_t t = 0; t < 8; ++t)
                            put(base + static_cast<size_t>(st) * 1024 + l * 8 + t, 16 * jt + l % 16,
                                32 * st + 8 * (l / 16) + t, 512);
                if (tail)
                    for (size_t l = 0; l < 64; ++l)
                        for (size_t t = 0; t < 4; ++t)
                            put(base + static_cast<size_t>(n32) * 1024 + l * 4 + t, 16 * jt + l % 16,
                                32 * n32 + 4 * (l / 16) + t, 256);
            }
            upload(ana16[slot], out);
            ana16Inv[slot] = std::ldexp(1.0f, -e);
        }
    }

    template <typename T>
    void setupTyped() {
        uploadAs<T>(synC, mfmaOperand(d.synC, d.Zs, d.Js));
        uploadAs<T>(synS, mfmaOperand(d.synS, d.Zs, d.Js));
        uploadAs<T>(anaC, mfmaOperand(d.anaC, d.Ja, d.Za));
        uploadAs<T>(anaS, mfmaOperand(d.anaS, d.Ja, d.Za));
        if (d.stratified) {
            uploadAs<T>(synGz, mfmaOperand(d.synGz, d.Zs, d.Js));
            uploadAs<T>(synFh, mfmaOperand(d.synFh, d.Zs, d.Js));
            uploadAs<T>(synEz, mfmaOperand(d.synEz, d.Zs, d.Js));
        }
        if (fold) {
            // rows z <= half of the synthesis, split by the parity of j
            const size_t half = (d.c.Nz - 1) / 2, Jh = d.Js / 2, Jah = d.Ja / 2;
            Zh = (half + 1 + 15) / 16 * 16;
            // (scale != 0: column j times scale m_j, the z-derivative operator)
            auto synPart = [&](const std::vector<double>& m, size_t par, double scale = 0.0) {
                std::vector<double> out(Zh * Jh, 0.0);
                for (size_t z = 0; z <= half; ++z)
                    for (size_t jh = 0; jh < Jh; ++jh) {
                        const size_t j = 2 * jh + par;
                        const double a = m[z * d.Js + j];
                        out[z * Jh + jh] = scale == 0.0 ? a : j < d.c.Nj ? a * (scale * d.m[j]) : 0.0;
                    }
                return mfmaOperand(out, Zh, Jh);
            };
            auto anaPart = [&](const std::vector<double>& m, size_t par) {
                std::vector<double> out(Jah * Zh, 0.0);
                for (size_t jh = 0; jh < Jah; ++jh)
                    for (size_t z = 0; z <= half; ++z) out[jh * Zh + z] = m[(2 * jh + par) * d.Za + z];
                return mfmaOperand(out, Jah, Zh);
            };
            uploadAs<T>(synCE, synPart(d.synC, 0));
            uploadAs<T>(synSE, synPart(d.synS, 0));
            uploadAs<T>(synCO, synPart(d.synC, 1));
            uploadAs<T>(synSO, synPart(d.synS, 1));
            uploadAs<T>(synCmE, synPart(d.synC, 0, 0.5));
            uploadAs<T>(synSmE, synPart(d.synS, 0, -0.5));
            uploadAs<T>(synCmO, synPart(d.synC, 1, 0.5));
            uploadAs<T>(synSmO, synPart(d.synS, 1, -0.5));
            if (precision == WV_PRECISION_MIXED_SPLIT16) {
                setupSplit16Synthesis(half, Jh);
                setupSplit16Analysis(half, Jah);
            }
            uploadAs<T>(anaCE, anaPart(d.anaC, 0));
            uploadAs<T>(anaCO, anaPart(d.anaC, 1));
            uploadAs<T>(anaSE, anaPart(d.anaS, 0));
            uploadAs<T>(anaSO, anaPart(d.anaS, 1));
        }
        uploadAs<T>(twX, d.twX);
        uploadAs<T>(zeroRow, std::vector<double>(2 * d.nkxP, 0.0));  // a row of zero spectra
        uploadAs<T>(twY, d.twY);
        {
            // y wavenumber of each storage row, as the synthesis forms it
            // (rowSign * l of every column of the row); valid when all
            // columns of a row agree
            std::vector<double> l(d.c.Ny, 0.0);
            std::vector<char> seen(d.c.Ny, 0);
            lrowValid = true;
            for (const Column& c : d.columns) {
                const double sgn = (c.flags & kColConj) ? -1.0 : 1.0;
                const double v = sgn * d.modes[static_cast<size_t>(c.wv)].l;
                const size_t iy = static_cast<size_t>(c.iy);
                if (seen[iy] && l[iy] != v) lrowValid = false;
                l[iy] = v;
                seen[iy] = 1;
            }
            uploadAs<T>(lrow, l);
        }
        uploadAs<T>(mj, d.m);
        uploadAs<T>(kxv, d.kxv);
        tsize = csize<T>();
    }

    void setup(const wv_config& cfg, wv_precision prec, int dev, const Stratification* strat = nullptr) {
        precision = prec;
        device = dev;
        hipCheck(hipSetDevice(dev), "hipSetDevice");
        // y-transform tile: KT columns of Ny, at most 4096 points; for FP64
        // at most 2048 while that keeps 8 columns (128-byte rows): Ny = 256
        // then runs 8 columns in a 32 KB tile, 8 workgroups per CU instead
        // of 4 (FP32 keeps 16 columns, which measured faster than 8; FP64 at
        // Ny = 512 keeps 8, 4 measured slower)
        // (runtime-length kernels stage 16 entries per thread whatever the
        // tile: they keep the widest tile)
        KT = static_cast<int>(std::max<size_t>(1, std::min<size_t>(16, 4096 / std::max<size_t>(cfg.Ny, 1))));
        const bool ny2 = cfg.Ny > 0 && (cfg.Ny & (cfg.Ny - 1)) == 0;
        if (prec == WV_PRECISION_FP64 && ny2 && KT > 8 && static_cast<size_t>(KT) * cfg.Ny > 2048)
            KT = static_cast<int>(std::max<size_t>(8, 2048 / cfg.Ny));
        while (KT & (KT - 1)) --KT;  // power of two
        std::string err;
        try {
            err = d.build(cfg, static_cast<size_t>(KT), strat);
        } catch (const std::overflow_error& e) {
            throw Failure{WV_SIZE_OVERFLOW, e.what()};
        } catch (const std::bad_alloc&) {
            throw Failure{WV_ALLOCATION_FAILURE, "Unable to allocate immutable kernel descriptor storage."};
        }
        if (!err.empty()) throw Failure{WV_INVALID_CONFIGURATION, err};
        const bool hydro = cfg.isHydrostatic != 0;
        NT = hydro ? 3 : 4;
        NCH = 3 + 3 * NT;
        NH = hydro ? 10 : 12;
        targetCode = hydro ? (0 | (1 << 4) | (3 << 8)) : (0 | (1 << 4) | (2 << 8) | (3 << 12));
        const size_t pairs = static_cast<size_t>(NT == 4 ? XMap<4>::pairs : XMap<3>::pairs);
        // In-LDS FFTs hold at most kFftRegs * kFftThreads = 4096 points per
        // call: one transform of Nx or KT columns of Ny (the x kernels take a
        // row's transforms in several calls when they hold more); the fused x
        // kernel keeps a row's `pairs` transforms and the twiddles in LDS.
        const size_t fftPoints = static_cast<size_t>(kFftRegs) * kFftThreads;
        if (cfg.Nx > fftPoints || static_cast<size_t>(KT) * cfg.Ny > fftPoints)
            throw Failure{WV_UNSUPPORTED_OPERATION, "Horizontal grid too large for the in-LDS FFTs (Nx, Ny <= 4096)."};
        int maxLds = 0;
        hipCheck(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "hipDeviceGetAttribute");
        maxLdsPerBlock = static_cast<size_t>(maxLds);
        const size_t elem = prec == WV_PRECISION_FP64 ? 16 : 8;
        const size_t xLds = (cfg.Nx + pairs * static_cast<size_t>(ldsLength(static_cast<int>(cfg.Nx)))) * elem;
        const size_t yLds = (static_cast<size_t>(KT) + 1) * cfg.Ny * elem;
        if (xLds > static_cast<size_t>(maxLds) || yLds > static_cast<size_t>(maxLds))
            throw Failure{WV_UNSUPPORTED_OPERATION, "Horizontal grid needs more LDS per workgroup than the device provides."};
        // staged rows: Za, or 2 x ((Nz-1)/2 + 1) rounded up to 16 when folded
        const int Zq = static_cast<int>(std::max(d.Za, 2 * (((d.c.Nz - 1) / 2 + 1 + 15) / 16 * 16)));
        const size_t pLds = prec == WV_PRECISION_FP64 ? projLdsBytes<double>(NT, Zq, static_cast<int>(d.Ja))
                                                      : projLdsBytes<float>(NT, Zq, static_cast<int>(d.Ja));
        if (d.Ja > 256 || pLds > static_cast<size_t>(maxLds))
            throw Failure{WV_UNSUPPORTED_OPERATION, "Vertical resolution too large for the in-LDS analysis (Nj <= 256)."};
        // columns
        ncol = d.columns.size();
        std::vector<ColumnDev> hc(ncol);
        for (size_t i = 0; i < ncol; ++i) {
            const Column& c = d.columns[i];
            const FourierMode& fm = d.modes[static_cast<size_t>(c.wv)];
            hc[i] = {c.iy, c.kx, c.wv, c.flags, fm.k, fm.l, fm.Kh, fm.cosAlpha, fm.sinAlpha};
        }
        upload(cols, hc);
        {
            const size_t tiles = d.nkxP / static_cast<size_t>(KT);
            std::vector<int> lim(tiles, 0);
            for (const Column& c : d.columns) {
                const long ky = c.iy <= static_cast<int>(cfg.Ny / 2) ? c.iy : c.iy - static_cast<long>(cfg.Ny);
                int& L = lim[static_cast<size_t>(c.kx) / static_cast<size_t>(KT)];
                L = std::max(L, static_cast<int>(std::labs(ky)));
            }
            size_t kept = 0;
            for (int L : lim) kept += std::min(cfg.Ny, 2 * static_cast<size_t>(L) + 1);
            retainedRowFraction = static_cast<double>(kept) / static_cast<double>(tiles * cfg.Ny);
            upload(rowLimit, lim);
        }
        upload(omega, d.omega);
        upload(UAp, d.UAp);
        upload(VAp, d.VAp);
        upload(WAp, d.WAp);
        upload(NAp, d.NAp);
        upload(UA0, d.UA0);
        upload(VA0, d.VA0);
        upload(NA0, d.NA0);
        upload(A0Vort, d.A0Vort);
        upload(A0Buoy, d.A0Buoy);
        upload(ApmN, d.ApmN);
        upload(ApmDScaled, d.ApmDScaled);
        upload(ApmD, d.ApmD);
        upload(ApmTE, d.ApmTE);
        upload(A0TE, d.A0TE);
        upload(apmW, d.apmW);
        upload(inertialScale, d.inertialScale);
        upload(fWaveScale, d.fWaveScale);
        upload(gWaveScale, d.gWaveScale);
        upload(Fg, d.Fg);
        upload(Gg, d.Gg);
        fold = !d.stratified;
        if (precision == WV_PRECISION_FP64) setupTyped<double>();
        else setupTyped<float>();
        const size_t plane = cfg.Ny * d.nkxP;
        const size_t chField = cfg.Nz * plane;
        // the synthesis addresses a channel with 32-bit byte offsets
        if (chField * tsize >= (size_t(1) << 32))
            throw Failure{WV_UNSUPPORTED_OPERATION, "Grid too large: one spatial channel exceeds 4 GiB."};
        FV.alloc(4 * d.M * tsize);
        H.alloc(static_cast<size_t>(NH) * chField * tsize);
        G.alloc(static_cast<size_t>(NH) * chField * tsize);
        P.alloc(static_cast<size_t>(NT) * chField * tsize);
        Q.alloc(static_cast<size_t>(NT) * chField * tsize);
        Tg.alloc(static_cast<size_t>(NT) * d.M * tsize);
        {
            std::vector<ModeDev> hm(d.Nkl);
            for (size_t i = 0; i < d.Nkl; ++i) {
                const FourierMode& fm = d.modes[i];
                hm[i] = {fm.k, fm.l, fm.Kh, fm.cosAlpha, fm.sinAlpha};
            }
            upload(modeTab, hm);
            // per-j inputs of the coefficient tables: m, h0, Fg, Gg, G wave
            // scale, apmW, inertialScale
            const size_t Nj = d.c.Nj;
            std::vector<double> jt(7 * Nj);
            for (size_t j = 0; j < Nj; ++j) {
                jt[j] = d.m[j];
                jt[Nj + j] = d.h0[j];
                jt[2 * Nj + j] = d.Fg[j];
                jt[3 * Nj + j] = d.Gg[j];
                jt[4 * Nj + j] = d.gWaveScale[j];
                jt[5 * Nj + j] = d.apmW[j];
                jt[6 * Nj + j] = d.inertialScale[j];
            }
            upload(jTab, jt);
            const double* jp = jTab.as<double>();
            // per-j factors of the mixed-precision projection tables
            // (the descriptor's formulas, WVKernelTypes.cpp:214-284, with
            // everything that depends on j alone formed here)
            std::vector<double> mj(static_cast<size_t>(kMixedJ) * Nj, 0.0);
            for (size_t j = 0; j < Nj; ++j) {
                const double h0 = d.h0[j], Fg = d.Fg[j], Gg = d.Gg[j], sign = j % 2 == 0 ? 1.0 : -1.0;
                const double Gw = cfg.isHydrostatic ? Gg : j == 0 ? 1.0 : sign * d.waveNorm;
                const double Gwg = Gg / Gw;
                const double prefactor = d.stratified ? 0.0 : sign * d.projPrefactor;
                auto at = [&](int row) -> double& { return mj[static_cast<size_t>(row) * Nj + j]; };
                at(kMjGwgOverGg) = Gwg / Gg;
                at(kMjDeltaScale) = h0 * Gwg / Fg;
                at(kMjLr2inv) = j == 0 ? 0.0 : d.f2 / (cfg.g * h0);
                at(kMjA0VortGeo) = 1.0 / Fg;
                at(kMjA0BuoyGeo) = j == 0 ? 0.0 : (-d.f / h0) / Gg;
                at(kMjApmDScaled) = d.m[j] / 2.0 * prefactor;
                at(kMjA0VortMda) = j == 0 ? 0.0 : d.f2 / (2.0 * h0) / Fg;
                at(kMjA0BuoyMda) = 1.0 / Gg;
                at(kMjNA0Num) = j == 0 ? 0.0 : -(d.f / cfg.g) * Gg;
            }
            upload(mixedJTab, mj);
            coefConsts = CoefConsts{d.f, d.f2, d.N02, cfg.g, cfg.Lz, d.waveNorm, d.projPrefactor,
                                    cfg.isHydrostatic ? 1 : 0, d.stratified ? 1 : 0,
                                    jp, jp + Nj, jp + 2 * Nj, jp + 3 * Nj, jp + 4 * Nj, jp + 5 * Nj, jp + 6 * Nj,
                                    mixedJTab.as<double>(), static_cast<int>(Nj)};
        }
        hipCheck(hipMemset(H.p, 0, H.bytes), "hipMemset");
        hipCheck(hipMemset(P.p, 0, P.bytes), "hipMemset");
        stage.alloc(3 * d.M * sizeof(cx<double>));
        weighted.alloc(3 * d.M * sizeof(cx<double>));
        partial.alloc(kEnergyPartials * sizeof(double));
        scalar.alloc(16 * sizeof(double));
        hipCheck(hipDeviceSynchronize(), "setup");
    }

    // ---- pipeline -----------------------------------------------------------
    // Half-spectrum channels of the flux (nonhydrostatic / hydrostatic):
    //   u v w eta, d/dy of each target, d/dz of each target
    // (the d/dx channels are formed in the x transform).  For an arbitrary
    // stratification the eta target's z channel is d/dz eta + eta dLnN2.
    // derived: d/dz w from continuity (XMap<5>; u v w eta, dy, dz u, dz v,
    // dz eta: 11 channels).
    SynthPlan fluxPlan(bool derived = false) const {
        SynthPlan sp{};
        sp.valueScale = 0.5;
        sp.ngroups = 2;
        if (derived) {
            sp.g[0] = {0, 3, {{0, 0, 0, -1, 4}, {1, 0, 1, -1, 5}, {3, 1, 10, -1, -1}, {}}};
            sp.g[1] = {1, 4, {{2, 0, 2, -1, 6}, {3, 0, 3, -1, 7}, {0, 1, 8, -1, -1}, {1, 1, 9, -1, -1}}};
        } else if (d.stratified) {
            sp.valueScale = 1.0;
            sp.ngroups = 4;
            sp.g[0] = {0, 2, {{0, 0, 0, -1, 4}, {1, 0, 1, -1, 5}, {}, {}}};
            sp.g[1] = {1, 2, {{2, 0, 2, -1, -1}, {3, 0, 3, -1, 6}, {}, {}}};
            sp.g[2] = {2, 2, {{0, 0, 7, -1, -1}, {1, 0, 8, -1, -1}, {}, {}}};
            sp.g[3] = {4, 1, {{3, 0, 9, -1, -1}, {}, {}, {}}};
        } else if (NT == 4) {
            sp.g[0] = {0, 4, {{0, 0, 0, -1, 4}, {1, 0, 1, -1, 5}, {2, 1, 10, -1, -1}, {3, 1, 11, -1, -1}}};
            sp.g[1] = {1, 4, {{2, 0, 2, -1, 6}, {3, 0, 3, -1, 7}, {0, 1, 8, -1, -1}, {1, 1, 9, -1, -1}}};
        } else {
            sp.g[0] = {0, 3, {{0, 0, 0, -1, 4}, {1, 0, 1, -1, 5}, {3, 1, 9, -1, -1}, {}}};
            sp.g[1] = {1, 4, {{2, 0, 2, -1, -1}, {3, 0, 3, -1, 6}, {0, 1, 7, -1, -1}, {1, 1, 8, -1, -1}}};
        }
        return sp;
    }

    // The flux plan without the y-derivative outputs, which the inverse y
    // transform forms instead (plan in *yp; false when the storage rows'
    // y wavenumbers are not one per row).
    bool fluxPlanLean(SynthPlan& sp, YDerivPlan& yp, bool derived = false) const {
        if (!lrowValid) return false;
        sp = fluxPlan(derived);
        yp = YDerivPlan{};
        int n = 0;
        for (int g = 0; g < sp.ngroups; ++g)
            for (int q = 0; q < sp.g[g].nch; ++q) {
                SynthChannel& c = sp.g[g].ch[q];
                if (c.outX >= 0 || c.outV < 0) return false;
                yp.src[n] = c.outV;
                yp.der[n] = c.outY;
                ++n;
                c.outY = -1;
            }
        return n <= 16;
    }

    template <typename T>
    SynthMats<T> synthMats() const {
        return SynthMats<T>{{synC.as<T>(), synS.as<T>(), d.stratified ? synGz.as<T>() : nullptr,
                             d.stratified ? synFh.as<T>() : nullptr, d.stratified ? synEz.as<T>() : nullptr}};
    }

    ProjTables projTables() const {
        return ProjTables{omega.as<double>(), A0Vort.as<double>(),     A0Buoy.as<double>(),
                          NA0.as<double>(),   ApmN.as<double>(),       ApmDScaled.as<double>(),
                          ApmD.as<cx<double>>(), apmW.as<double>(),    inertialScale.as<double>()};
    }

    template <typename T>
    void launchFields(const double* Ap, const double* Am, const double* A0, double elapsed, hipStream_t s) {
        const int M = static_cast<int>(d.M);
        // reads Ap, Am, A0 (48 B) per coefficient (the tables are formed from
        // per-mode and per-j values), writes four field spectra
        timed("fields", 48.0 * M + 4.0 * M * sizeof(cx<T>), s, [&] {
            hipLaunchKernelGGL(k_fields<T>, dim3((M + 255) / 256), dim3(256), 0, s,
                               reinterpret_cast<const cx<double>*>(Ap), reinterpret_cast<const cx<double>*>(Am),
                               reinterpret_cast<const cx<double>*>(A0), modeTab.as<ModeDev>(),
                               static_cast<int>(d.c.Nj), coefConsts, elapsed, M, FV.as<cx<T>>());
        });
    }

    // FV -> H with the synthesis matrices unless given other buffers/matrices.
    template <typename T>
    void launchSynth(const SynthPlan& sp, hipStream_t s, const DeviceBuffer* in = nullptr,
                     const DeviceBuffer* out = nullptr, const SynthMats<T>* mats = nullptr,
                     const char* name = "synth") {
        constexpr int ZT = sizeof(T) == 8 ? 2 : 3;  // z tiles per task (FP64 accumulators are twice the size)
        int maxch = 0;
        for (int g = 0; g < sp.ngroups; ++g) maxch = std::max(maxch, sp.g[g].nch);
        const int ncg = static_cast<int>((ncol + 7) / 8);
        int nout = 0;
        for (int g = 0; g < sp.ngroups; ++g)
            for (int q = 0; q < sp.g[g].nch; ++q)
                nout += (sp.g[g].ch[q].outV >= 0) + (sp.g[g].ch[q].outX >= 0) + (sp.g[g].ch[q].outY >= 0);
        const double bytes = static_cast<double>(nout) * d.c.Nz * ncol * sizeof(cx<T>) + 4.0 * d.M * sizeof(cx<T>);
        const size_t lds = 4 * 8 * static_cast<size_t>(synthJStride<T>(static_cast<int>(d.Js))) * sizeof(cx<T>) + d.Js * sizeof(T);
        constexpr int FQ = 2;
        // the folded DCT-I / DST-I path serves the constant-N matrices, for
        // plans whose parts of FQ channels are each all values or all z
        // derivatives (one operator per task)
        bool homogeneous = true;
        for (int g = 0; g < sp.ngroups; ++g)
            for (int q = 0; q < sp.g[g].nch; ++q)
                homogeneous = homogeneous && sp.g[g].ch[q].dz == sp.g[g].ch[q - q % FQ].dz;
        const bool folded = fold && mats == nullptr && homogeneous;
        // folded, when the last 16-row tile holds only the midpoint row
        // z = L/2 (Nz = 16 m + 1): that row is a dot product beside the MFMA
        // tiles
        const int half = static_cast<int>((d.c.Nz - 1) / 2);
        const bool midRow = folded && (d.c.Nz - 1) % 2 == 0 && (half + 1) % 16 == 1;
        const int ztiles = static_cast<int>((folded ? Zh : d.Zs) / 16) - (midRow ? 1 : 0);
        const int zmid = midRow ? half : -1;
        // Folded tasks take FQ = 2 channel slots (64 VGPRs in FP32: 8 waves
        // per SIMD, so 6 workgroups per CU, as many as the LDS holds) and run
        // on four waves, one per SIMD, which deal the tasks round-robin.  (A
        // five-wave workgroup puts two of its waves on one SIMD, which then
        // has twice the work of the others.)  Unfolded: the largest of 4 ... 1
        // waves that divides the tasks.
        const int tasks = folded ? sp.ngroups * ((maxch + FQ - 1) / FQ) * ztiles : sp.ngroups * ((ztiles + ZT - 1) / ZT);
        int waves = folded ? std::min(4, tasks) : 4;
        while (!folded && tasks % waves) --waves;
        const int nsteps = static_cast<int>(folded ? d.Js / 8 : d.Js / 4);
        const bool ks4 = nsteps % 8 == 0;
        using Fn = void (*)(const cx<T>*, const ColumnDev*, int, int, SynthMats<T>, int, int, int, int, int, int,
                            int, const T*, SynthPlan, cx<T>*);
        Fn fn;
        // folded FP32: operand batches of up to 6 steps (a wave keeps a
        // whole batch of loads in flight; the MFMA work of one small batch is
        // much shorter than a load's latency).  FP64 keeps short batches (its
        // registers bound the resident waves).
        if (folded)
            fn = sizeof(T) == 4 && nsteps % 12 == 0 ? k_synth<T, 1, FQ, 6, true>
                                  : ks4 ? k_synth<T, 1, FQ, 4, true> : k_synth<T, 1, FQ, 2, true>;
        else
            fn = maxch <= 2 ? (ks4 ? k_synth<T, ZT, 2, 4, false> : k_synth<T, ZT, 2, 2, false>)
                            : (ks4 ? k_synth<T, ZT, 4, 4, false> : k_synth<T, ZT, 4, 2, false>);
        const SynthMats<T> m = folded ? SynthMats<T>{{synCE.as<T>(), synSE.as<T>(), synCO.as<T>(), synSO.as<T>(),
                                                      synCmE.as<T>(), synSmE.as<T>(), synCmO.as<T>(), synSmO.as<T>()}}
                                      : (mats ? *mats : synthMats<T>());
        if constexpr (sizeof(T) == 4) {
            if (folded && syn16N32 + syn16Tail > 0 && maxch <= 4) {
                // split-FP16 products: every tile on MFMA (the midpoint row
                // included), FQ channel slots per task, four waves
                using Fn16 = void (*)(const cx<float>*, const ColumnDev*, int, int, SynthMats16, int, int, int, int,
                                      int, int, SynthPlan, cx<float>*);
                const int key = syn16N32 * 2 + syn16Tail;
                const Fn16 fn16 = key == 1 ? k_synth16<FQ, 0, 1> : key == 2 ? k_synth16<FQ, 1, 0>
                                : key == 3 ? k_synth16<FQ, 1, 1> : key == 4 ? k_synth16<FQ, 2, 0> : k_synth16<FQ, 2, 1>;
                SynthMats16 m16{};
                for (int i = 0; i < kSynthMatSlots; ++i) {
                    m16.m[i] = syn16[i].as<_Float16>();
                    m16.inv[i] = syn16Inv[i];
                }
                const int nc8 = 4 * syn16N32 + 2 * syn16Tail;
                const size_t lds16 = static_cast<size_t>(4096) * nc8;
                timed(name, bytes, s, [&] {
                    hipLaunchKernelGGL(fn16, dim3(static_cast<unsigned>(ncg)), dim3(256), lds16, s,
                                       in ? in->as<cx<float>>() : FV.as<cx<float>>(), cols.as<ColumnDev>(),
                                       static_cast<int>(ncol), static_cast<int>(d.M), m16, static_cast<int>(d.Js),
                                       static_cast<int>(d.c.Nz), static_cast<int>(d.c.Nj), static_cast<int>(d.c.Ny),
                                       static_cast<int>(d.nkxP), static_cast<int>(Zh / 16), sp,
                                       out ? out->as<cx<float>>() : H.as<cx<float>>());
                });
                return;
            }
        }
        timed(name, bytes, s, [&] {
            hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(ncg)), dim3(64 * waves), lds, s,
                               in ? in->as<cx<T>>() : FV.as<cx<T>>(), cols.as<ColumnDev>(), static_cast<int>(ncol),
                               static_cast<int>(d.M), m, ztiles * 16, zmid, static_cast<int>(d.Js),
                               static_cast<int>(d.c.Nz), static_cast<int>(d.c.Nj), static_cast<int>(d.c.Ny),
                               static_cast<int>(d.nkxP), mj.as<T>(), sp, out ? out->as<cx<T>>() : H.as<cx<T>>());
        });
    }

    // skipIn: input rows outside the tile's retained ky band are zeros;
    // skipOut: output rows outside it are not needed.  yp (inverse flux
    // transform): the channel plan with y derivatives; nch is then its
    // number of input channels.
    template <typename T>
    void launchYfft(const DeviceBuffer& in, DeviceBuffer& out, int nch, int sign, hipStream_t s, const char* name,
                    bool skipIn, bool skipOut, const YDerivPlan* yp = nullptr) {
        const size_t lds = (static_cast<size_t>(KT) + 1) * d.c.Ny * sizeof(cx<T>);
        const double frac = (skipIn || skipOut) ? 0.5 * (1.0 + retainedRowFraction) : 1.0;
        int nout = nch;
        if (yp)
            for (int c = 0; c < nch; ++c) nout += yp->der[c] >= 0;
        const double bytes = frac * (nch + nout) * static_cast<double>(d.c.Nz) * d.c.Ny * d.nkxP * sizeof(cx<T>);
        // Ny = 64 ... 1024 with their default tile widths run kernels
        // compiled for that length; others the runtime-length kernel.
        using Fn = void (*)(const cx<T>*, cx<T>*, int, int, int, int, int, int, const cx<T>*, int, const int*, int,
                            int, YDerivPlan, const T*, int);
        const int l2y = log2Exact(d.c.Ny), l2t = log2Exact(KT);
        const bool dv = yp != nullptr;
        Fn fn = dv ? k_yfft<T, 0, 0, true> : k_yfft<T, 0, 0, false>;
        if (smooth7(d.c.Ny)) fn = dv ? k_yfft<T, kSmooth7Rounds, 0, true> : k_yfft<T, kSmooth7Rounds, 0, false>;
        if (l2y == 6 && l2t == 4) fn = dv ? k_yfft<T, 6, 4, true> : k_yfft<T, 6, 4, false>;
        if (l2y == 7 && l2t == 4) fn = dv ? k_yfft<T, 7, 4, true> : k_yfft<T, 7, 4, false>;
        if (l2y == 8 && l2t == 4) fn = dv ? k_yfft<T, 8, 4, true> : k_yfft<T, 8, 4, false>;
        if (l2y == 8 && l2t == 3) fn = dv ? k_yfft<T, 8, 3, true> : k_yfft<T, 8, 3, false>;
        if (l2y == 9 && l2t == 3) fn = dv ? k_yfft<T, 9, 3, true> : k_yfft<T, 9, 3, false>;
        if (l2y == 10 && l2t == 2) fn = dv ? k_yfft<T, 10, 2, true> : k_yfft<T, 10, 2, false>;
        const YDerivPlan plan = yp ? *yp : YDerivPlan{};
        timed(name, bytes, s, [&] {
            hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(d.nkxP / KT), static_cast<unsigned>(d.c.Nz), nch),
                               dim3(kFftThreads), lds, s, in.as<cx<T>>(), out.as<cx<T>>(), static_cast<int>(d.c.Ny),
                               l2y, static_cast<int>(d.nkx), static_cast<int>(d.nkxP), l2t, static_cast<int>(d.c.Nz),
                               twY.as<cx<T>>(), sign, rowLimit.as<int>(), skipIn ? 1 : 0, skipOut ? 1 : 0, plan,
                               lrow.as<T>(), KT * sizeof(cx<T>) < 64 ? 1 : 0);
        });
    }

    template <typename T, int NTc, int RB, int L2N>
    void launchXfusedKernel(hipStream_t s) {
        constexpr int pairs = XMap<NTc>::pairs;
        const int Nx = static_cast<int>(d.c.Nx), Ny = static_cast<int>(d.c.Ny), Nz = static_cast<int>(d.c.Nz);
        const int nrows = Ny * Nz;
        const size_t lds = xfusedLdsBytes<T>(Nx, pairs, RB);
        if (lds > maxLdsPerBlock) throw Failure{WV_UNSUPPORTED_OPERATION, "x transforms exceed the LDS of a workgroup."};
        // power of two nearest N0 (exact scaling of the eta product)
        const T es = static_cast<T>(std::ldexp(1.0, static_cast<int>(std::lround(std::log2(d.c.N0 > 0 ? d.c.N0 : 1.0)))));
        auto fn = k_xfused<T, NTc, RB, L2N, false>;
        if constexpr (L2N == 8)
            if (d.nkx <= 96) fn = k_xfused<T, NTc, RB, L2N, true>;
        hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>((nrows + RB - 1) / RB)),
                           dim3(kFftThreads), lds, s, G.as<cx<T>>(), P.as<cx<T>>(), Nx, log2Exact(d.c.Nx), Ny, Nz,
                           static_cast<int>(d.nkx), static_cast<int>(d.nkxP), twX.as<cx<T>>(), kxv.as<T>(),
                           zeroRow.as<cx<T>>(), es, static_cast<const double*>(nullptr));
    }

    // Nx = 64 ... 1024 run kernels compiled for that length; others the
    // runtime-length kernels.
    template <typename T, int NTc>
    void launchXfusedFor(hipStream_t s) {
        constexpr int pairs = XMap<NTc>::pairs;
        switch (log2Exact(d.c.Nx)) {
            case 6: return launchXfusedKernel<T, NTc, xfusedRows<T>(64, pairs), 6>(s);
            case 7: return launchXfusedKernel<T, NTc, xfusedRows<T>(128, pairs), 7>(s);
            case 8: return launchXfusedKernel<T, NTc, xfusedRows<T>(256, pairs), 8>(s);
            case 9: return launchXfusedKernel<T, NTc, xfusedRows<T>(512, pairs), 9>(s);
            case 10: return launchXfusedKernel<T, NTc, 1, 10>(s);
            default: break;
        }
        // a 7-smooth length whose row transforms fit one round of the in-LDS
        // FFTs: the single-round kernel; else passes in rounds
        static_assert(xfusedRows<T>(64, pairs) == 1, "one-row blocks at every length");
        if (static_cast<size_t>(pairs) * d.c.Nx <= static_cast<size_t>(kFftRegs) * kFftThreads && smooth7(d.c.Nx))
            launchXfusedKernel<T, NTc, 1, kSmooth7>(s);
        else
            launchXfusedKernel<T, NTc, 1, 0>(s);
    }

    // The products with the caller's advecting fields (runtime-length
    // kernel, any Nx).
    template <typename T, int NTc>
    void launchXfusedExt(const double* adv, hipStream_t s) {
        constexpr int pairs = XMap<NTc>::pairs;
        const int Nx = static_cast<int>(d.c.Nx), Ny = static_cast<int>(d.c.Ny), Nz = static_cast<int>(d.c.Nz);
        const size_t lds = xfusedLdsBytes<T>(Nx, pairs, 1);
        if (lds > maxLdsPerBlock) throw Failure{WV_UNSUPPORTED_OPERATION, "x transforms exceed the LDS of a workgroup."};
        const T es = static_cast<T>(std::ldexp(1.0, static_cast<int>(std::lround(std::log2(d.c.N0 > 0 ? d.c.N0 : 1.0)))));
        hipLaunchKernelGGL((k_xfused<T, NTc, 1, 0, false, true>), dim3(static_cast<unsigned>(Ny * Nz)),
                           dim3(kFftThreads), lds, s, G.as<cx<T>>(), P.as<cx<T>>(), Nx, log2Exact(d.c.Nx), Ny, Nz,
                           static_cast<int>(d.nkx), static_cast<int>(d.nkxP), twX.as<cx<T>>(), kxv.as<T>(),
                           zeroRow.as<cx<T>>(), es, adv);
    }

    // derived: the d/dz w channel is not in G (XMap<5>).
    template <typename T>
    void launchXfused(hipStream_t s, const double* adv = nullptr, bool derived = false) {
        const double nh = derived ? NH - 1 : NH;
        const double bytes = (nh + NT) * d.c.Nz * d.c.Ny * d.nkx * sizeof(cx<T>) + (adv ? 3.0 * d.R * sizeof(double) : 0.0);
        timed("xfused", bytes, s, [&] {
            if (adv)
                derived ? launchXfusedExt<T, 5>(adv, s) : NT == 4 ? launchXfusedExt<T, 4>(adv, s) : launchXfusedExt<T, 3>(adv, s);
            else if (derived)
                launchXfusedFor<T, 5>(s);
            else if (NT == 4)
                launchXfusedFor<T, 4>(s);
            else
                launchXfusedFor<T, 3>(s);
        });
    }

    // The mixed-precision flux of the constant-N nonhydrostatic transform
    // forms d/dz w from continuity (XMap<5>); the FP64 flux synthesizes every
    // derivative, as the reference's dataflow does.
    bool deriveDzW() const { return precision != WV_PRECISION_FP64 && NT == 4 && !d.stratified && lrowValid; }

    // Vertical analysis of the transformed products into Tg.
    template <typename T>
    void launchProj(hipStream_t s) {
        const int ncg = static_cast<int>((ncol + 7) / 8);
        const double bytes = static_cast<double>(NT) * (static_cast<double>(d.c.Nz) * ncol + d.M) * sizeof(cx<T>);
        if constexpr (sizeof(T) == 4) {
            const int njt = static_cast<int>(d.Ja / 16);
            const int nc8 = 4 * ana16N32 + 2 * ana16Tail;
            // two targets per workgroup (u, v of the cosine family; w, eta
            // or eta alone of the sine family): half the staged operands,
            // twice the workgroups
            constexpr int kNTB = 2;
            if (fold && ana16N32 > 0 && 2 * 64 * njt >= kNTB * 16 * nc8) {
                // split-FP16 analysis: the staged operands (kNTB targets x
                // sums, differences x hi, lo x chunks x 16 columns of 16 B),
                // later the transposed results
                const size_t staged = static_cast<size_t>(kNTB) * 4 * nc8 * 16 * 16;
                const size_t results = static_cast<size_t>(d.Ja) * kNTB * 17 * sizeof(float);
                const size_t lds = std::max(staged, results);
                using Fn16 = void (*)(const cx<float>*, const ColumnDev*, int, SynthMats16, int, int, int, int, int,
                                      int, cx<float>*);
                const int key = (ana16N32 - 1) * 2 + ana16Tail;
                const Fn16 fn16 = NT == 4 ? (key == 0 ? k_proj16<4, kNTB, 1, 0> : key == 1 ? k_proj16<4, kNTB, 1, 1>
                                             : key == 2 ? k_proj16<4, kNTB, 2, 0> : k_proj16<4, kNTB, 2, 1>)
                                          : (key == 0 ? k_proj16<3, kNTB, 1, 0> : key == 1 ? k_proj16<3, kNTB, 1, 1>
                                             : key == 2 ? k_proj16<3, kNTB, 2, 0> : k_proj16<3, kNTB, 2, 1>);
                SynthMats16 a16{};
                for (int i = 0; i < 4; ++i) {
                    a16.m[i] = ana16[i].as<_Float16>();
                    a16.inv[i] = ana16Inv[i];
                }
                timed("proj", bytes, s, [&] {
                    hipLaunchKernelGGL(fn16, dim3(static_cast<unsigned>(ncg * ((NT + kNTB - 1) / kNTB))),
                                       dim3(static_cast<unsigned>(64 * njt)), lds,
                                       s, Q.as<cx<float>>(), cols.as<ColumnDev>(), static_cast<int>(ncol), a16,
                                       static_cast<int>(d.Ja), static_cast<int>(d.c.Nz), static_cast<int>(d.c.Nj),
                                       static_cast<int>(d.c.Ny), static_cast<int>(d.nkxP), static_cast<int>(d.M),
                                       Tg.as<cx<float>>());
                });
                return;
            }
        }
        timed("proj", bytes, s, [&] {
            const int njt = static_cast<int>(d.Ja / 16);
            const dim3 grid(static_cast<unsigned>(ncg)), block(static_cast<unsigned>(64 * njt));
            const int Zq = static_cast<int>(fold ? Zh : d.Za);  // contraction rows (staged)
            const size_t lds = projLdsBytes<T>(NT, fold ? 2 * Zq : Zq, static_cast<int>(d.Ja));
            // folded, when the last 16-row tile holds only the midpoint row
            // z = L/2 (Nz = 16 m + 1): MFMA over the full tiles, that row as
            // a rank-1 term
            const int half = static_cast<int>((d.c.Nz - 1) / 2);
            const bool mid = fold && (d.c.Nz - 1) % 2 == 0 && (half + 1) % 16 == 1;
            const int Zk = mid ? Zq - 16 : Zq, zmid = mid ? half : -1;
            const int ncolI = static_cast<int>(ncol), Ja = static_cast<int>(d.Ja), Nz = static_cast<int>(d.c.Nz),
                      Nj = static_cast<int>(d.c.Nj), Ny = static_cast<int>(d.c.Ny), nkxP = static_cast<int>(d.nkxP),
                      M = static_cast<int>(d.M);
            // batch length: 4, 3 or 2 steps, whichever divides the step count in pairs
            const int nsteps = Zk / 4;
            const int ks = nsteps % 8 == 0 ? 4 : nsteps % 6 == 0 ? 3 : 2;
            using Fn = void (*)(const cx<T>*, const ColumnDev*, int, SynthMats<T>, int, int, int, int, int, int, int,
                                int, int, cx<T>*);
            Fn fn;
            if (fold)
                fn = NT == 4 ? (ks == 4 ? k_proj<T, 4, 4, true> : ks == 3 ? k_proj<T, 4, 3, true> : k_proj<T, 4, 2, true>)
                             : (ks == 4 ? k_proj<T, 3, 4, true> : ks == 3 ? k_proj<T, 3, 3, true> : k_proj<T, 3, 2, true>);
            else
                fn = NT == 4 ? (ks == 4 ? k_proj<T, 4, 4, false> : ks == 3 ? k_proj<T, 4, 3, false> : k_proj<T, 4, 2, false>)
                             : (ks == 4 ? k_proj<T, 3, 4, false> : ks == 3 ? k_proj<T, 3, 3, false> : k_proj<T, 3, 2, false>);
            const SynthMats<T> ana = fold ? SynthMats<T>{{anaCE.as<T>(), anaSE.as<T>(), anaCO.as<T>(), anaSO.as<T>(), nullptr}}
                                          : SynthMats<T>{{anaC.as<T>(), anaS.as<T>(), nullptr, nullptr, nullptr}};
            hipLaunchKernelGGL(fn, grid, block, lds, s, Q.as<cx<T>>(), cols.as<ColumnDev>(), ncolI, ana, Ja, Zq, Zk, zmid,
                               Nz, Nj, Ny, nkxP, M, Tg.as<cx<T>>());
        });
    }

    // HBM bytes of the stage epilogue's coefficient arrays: flux out 48 B per
    // coefficient; RK stage 1 reads y, writes w (and the stage unless the
    // next input's spectra are formed: 96 or 144 B); stages 2, 3 read w, y
    // and write w (and the stage: 144 or 192 B); stage 4 reads w, y and
    // writes y (144 B); with the next input's spectra four spectra (y and
    // the stage state are FP64 triples, 48 B; the weighted sum is a triple
    // in T, 3 sizeof(cx<T>))
    template <typename T>
    double stageEpilogueBytes(const Epilogue& ep) const {
        const double stBytes = ep.emitFields ? 0.0 : 48.0, wB = 3.0 * sizeof(cx<T>);
        const double epBytes = ep.mode == kFluxOut ? 48.0
                               : ep.mode == kRk1   ? 48.0 + wB + stBytes
                               : ep.mode == kRk4   ? 96.0 + wB
                                                   : 48.0 + 2.0 * wB + stBytes;
        const double fieldBytes = (ep.mode != kFluxOut && ep.emitFields) ? 4.0 * sizeof(cx<T>) : 0.0;
        return static_cast<double>(d.M) * (epBytes + fieldBytes);
    }

    // Projection + flux store or RK stage (+ the next flux input's spectra).
    template <typename T>
    void launchStage(double elapsed, const Epilogue& ep, hipStream_t s) {
        const int M = static_cast<int>(d.M);
        const double hscale = 1.0 / static_cast<double>(d.c.Nx * d.c.Ny);
        // per coefficient: the NT analysed values and the epilogue's arrays
        const double bytes = static_cast<double>(M) * NT * sizeof(cx<T>) + stageEpilogueBytes<T>(ep);
        // the kernel reads y and w in every mode
        Epilogue e = ep;
        if (e.mode == kFluxOut) {
            e.y0 = e.w0 = e.out0;
            e.y1 = e.w1 = e.out1;
            e.y2 = e.w2 = e.out2;
        }
        timed("stage", bytes, s, [&] {
            auto fn = NT == 4 ? k_stage<T, 4> : k_stage<T, 3>;
            hipLaunchKernelGGL(fn, dim3((M + 255) / 256), dim3(256), 0, s, Tg.as<cx<T>>(), modeTab.as<ModeDev>(),
                               static_cast<int>(d.c.Nj), M, coefConsts, elapsed, hscale, e, FV.as<cx<T>>());
        });
    }

    template <typename T>
    // fieldsReady: FV already holds the spectra of (Ap, Am, A0) at t (the
    // previous RK stage formed them).  adv: the caller's advecting fields
    // [3][Nz][Ny][Nx] (nonlinearFluxUsingAdvectionFields), else synthesized.
    void rhs(const double* Ap, const double* Am, const double* A0, double t, double t0, const Epilogue& ep,
             hipStream_t s, bool fieldsReady = false, const double* adv = nullptr) {
        const double elapsed = t - t0;
        if (!fieldsReady) launchFields<T>(Ap, Am, A0, elapsed, s);
        SynthPlan lean;
        YDerivPlan yp;
        const bool derived = deriveDzW();
        if (fluxPlanLean(lean, yp, derived)) {
            int nsrc = 0;
            for (int g = 0; g < lean.ngroups; ++g) nsrc += lean.g[g].nch;
            launchSynth<T>(lean, s);
            launchYfft<T>(H, G, nsrc, +1, s, "yfft_inv", true, false, &yp);
        } else {
            launchSynth<T>(fluxPlan(), s);
            launchYfft<T>(H, G, NH, +1, s, "yfft_inv", true, false);
        }
        launchXfused<T>(s, adv, derived);
        launchYfft<T>(P, Q, NT, -1, s, "yfft_fwd", false, true);
        launchProj<T>(s);
        launchStage<T>(elapsed, ep, s);
        // the logical transforms of the reference's schedule (one velocity
        // c2r, per target one derivative c2r and one r2c; vertical: three
        // velocity channels, three derivative channels and one analysis per
        // target), counted as its plan executions
        ++metrics.nonlinearFluxCallCount;
        metrics.nonlinearFluxPhaseEvaluationCount += d.M;
        if (!adv) ++metrics.advectionVelocityReconstructionCount;
        metrics.horizontalExecutionCount += (adv ? 0 : 1) + 2 * static_cast<size_t>(NT);
        metrics.verticalExecutionCount += (adv ? 0 : 3) + 4 * static_cast<size_t>(NT);
        metrics.executionCount = metrics.horizontalExecutionCount + metrics.verticalExecutionCount;
    }

    // The generic x transforms take rows in groups of about 16 real signals
    // (8 complex transforms per workgroup), fewer above Nx = 512 (the
    // workgroup's transforms hold about 4096 points; a row with more signals
    // than that takes its transforms in several calls).
    int xRowsPerBlock(int signalsPerRow) const {
        const int Nx = static_cast<int>(d.c.Nx);
        const int cap = std::max(2, std::min(16, 2 * (kFftRegs * kFftThreads) / std::max(Nx, 1)));
        return std::max(1, cap / std::max(1, signalsPerRow));
    }
    template <typename T>
    void launchXfwdRows(const double* in, cx<T>* P, int nf, int nkx, int nkxP, const cx<T>* tw, int adjointWeights,
                        hipStream_t s) {
        const int rpb = xRowsPerBlock(nf), ntr = (rpb * nf + 1) / 2;
        const int nrows = static_cast<int>(d.c.Ny * d.c.Nz);
        const size_t lds = static_cast<size_t>(ntr) * ldsLength(static_cast<int>(d.c.Nx)) * sizeof(cx<T>);
        if (lds > maxLdsPerBlock) throw Failure{WV_UNSUPPORTED_OPERATION, "x transforms exceed the LDS of a workgroup."};
        auto fn = k_xfwd<T, 0>;
        switch (log2Exact(d.c.Nx)) {  // lengths 64 ... 512 compiled as such
            case 6: fn = k_xfwd<T, 6>; break;
            case 7: fn = k_xfwd<T, 7>; break;
            case 8: fn = k_xfwd<T, 8>; break;
            case 9: fn = k_xfwd<T, 9>; break;
            default: break;
        }
        hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>((nrows + rpb - 1) / rpb)), dim3(256), lds, s, in, P,
                           nf, rpb, static_cast<int>(d.c.Nx), log2Exact(d.c.Nx), static_cast<int>(d.c.Ny),
                           static_cast<int>(d.c.Nz), nkx, nkxP, tw, adjointWeights);
    }
    template <typename T>
    void launchXinvRows(const cx<T>* G, double* out, int nch, int nkx, int nkxP, const cx<T>* tw, hipStream_t s) {
        launchXinvOp<T, kXinvFields>(G, out, nch, nkx, nkxP, tw, s);
    }
    template <typename T, int OP>
    void launchXinvOp(const cx<T>* G, double* out, int nch, int nkx, int nkxP, const cx<T>* tw, hipStream_t s,
                      double coriolis = 0.0) {
        const int rpb = xRowsPerBlock(nch), ntr = (rpb * nch + 1) / 2;
        const int nrows = static_cast<int>(d.c.Ny * d.c.Nz);
        const size_t lds = static_cast<size_t>(ntr) * ldsLength(static_cast<int>(d.c.Nx)) * sizeof(cx<T>);
        if (lds > maxLdsPerBlock) throw Failure{WV_UNSUPPORTED_OPERATION, "x transforms exceed the LDS of a workgroup."};
        auto fn = k_xinv<T, OP, 0>;
        switch (log2Exact(d.c.Nx)) {  // lengths 64 ... 512 compiled as such
            case 6: fn = k_xinv<T, OP, 6>; break;
            case 7: fn = k_xinv<T, OP, 7>; break;
            case 8: fn = k_xinv<T, OP, 8>; break;
            case 9: fn = k_xinv<T, OP, 9>; break;
            default: break;
        }
        hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>((nrows + rpb - 1) / rpb)), dim3(256), lds, s, G, out,
                           nch, rpb, static_cast<int>(d.c.Nx), log2Exact(d.c.Nx), static_cast<int>(d.c.Ny),
                           static_cast<int>(d.c.Nz), nkx, nkxP, tw, static_cast<const double*>(nullptr), coriolis);
    }

    // ---- advectFGridScalar (FP64 in either precision, as the reference's)
    // Scratch, allocated on the first call and kept: three half spectra
    // [3][Nz][Ny][nxhP] (nxh = Nx/2 + 1 columns padded to the y tile), the
    // d/dz column operator, the antialias row weights and FP64 twiddles.
    DeviceBuffer scC, scD, scKeep, scTwX, scTwY;
    int scNxh = 0, scNxhP = 0;

    // d/dz of an F-grid column (...Kernel.cpp:1078-1094) as one matrix: the
    // DCT-I (REDFT00) over z divided by L = Nz - 1 (last coefficient halved),
    // coefficients j = 1 ... L-1 times -pi j / Lz (j = 0 and L dropped), then
    // halved and synthesized by the DST-I (RODFT00) of z = 1 ... L-1:
    //   D[z][w] = sum_j -m_j sin(pi j z / L) c_w cos(pi j w / L) / L,
    // c_w = 1 at w = 0 and L, 2 inside; rows z = 0 and L are zero.
    static std::vector<double> scalarDzOperator(size_t Nz, double Lz) {
        const size_t L = Nz - 1;
        std::vector<double> D(Nz * Nz, 0.0);
        if (L < 2) return D;
        auto trig = [L](size_t p, bool sine) {
            const double a = kScalarPi * static_cast<double>(p % (2 * L)) / static_cast<double>(L);
            return sine ? std::sin(a) : std::cos(a);
        };
        for (size_t z = 1; z < L; ++z)
            for (size_t w = 0; w <= L; ++w) {
                const double cw = (w == 0 || w == L) ? 1.0 : 2.0;
                double acc = 0.0;
                for (size_t j = 1; j < L; ++j) {
                    const double mj = kScalarPi * static_cast<double>(j) / Lz;
                    acc += -mj * trig(j * z, true) * cw * trig(j * w, false);
                }
                D[z * Nz + w] = acc / static_cast<double>(L);
            }
        return D;
    }

    void ensureScalarScratch() {
        if (scD.p) return;
        const size_t Nx = d.c.Nx, Ny = d.c.Ny, Nz = d.c.Nz;
        scNxh = static_cast<int>(Nx / 2 + 1);
        scNxhP = (scNxh + KT - 1) / KT * KT;
        const size_t yLds = (static_cast<size_t>(KT) + 1) * Ny * sizeof(cx<double>);
        // the generic x transforms hold (rows per block x signals + 1) / 2
        // complex transforms: 8 for one signal (r2c) and for three (c2r)
        const size_t xTransforms = static_cast<size_t>((xRowsPerBlock(1) + 1) / 2);
        const size_t xLds = xTransforms * static_cast<size_t>(ldsLength(static_cast<int>(Nx))) * sizeof(cx<double>);
        const size_t zLds = Nz * 8 * sizeof(cx<double>);
        if (yLds > maxLdsPerBlock || xLds > maxLdsPerBlock || zLds > maxLdsPerBlock)
            throw Failure{WV_UNSUPPORTED_OPERATION, "Scalar advection needs more LDS per workgroup than the device provides."};
        const size_t half = Nz * Ny * static_cast<size_t>(scNxhP);
        scC.alloc(3 * half * sizeof(cx<double>));
        {
            const std::vector<double> D = scalarDzOperator(Nz, d.c.Lz);
            std::vector<double> DT(Nz * Nz);
            for (size_t z = 0; z < Nz; ++z)
                for (size_t w = 0; w < Nz; ++w) DT[w * Nz + z] = D[z * Nz + w];
            upload(scD, DT);  // [w][z]
        }
        // the rows that store or complete a wave-vortex mode (:467-471)
        std::vector<double> keep(Ny * static_cast<size_t>(scNxhP), 0.0);
        const double scale = 1.0 / static_cast<double>(Nx * Ny);
        for (const Column& c : d.columns) keep[static_cast<size_t>(c.iy) * scNxhP + static_cast<size_t>(c.kx)] = scale;
        upload(scKeep, keep);
        upload(scTwX, d.twX);
        upload(scTwY, d.twY);
        scalarPrepared = true;
    }

    void advectScalar(const double* scalar, const double* adv, bool antialias, double* rhs, hipStream_t s) {
        ensureScalarScratch();
        const int Nx = static_cast<int>(d.c.Nx), Ny = static_cast<int>(d.c.Ny), Nz = static_cast<int>(d.c.Nz);
        const int l2y = log2Exact(d.c.Ny), l2t = log2Exact(KT);
        cx<double>* C = scC.as<cx<double>>();
        const cx<double>* twx = scTwX.as<cx<double>>();
        auto xfwd = [&](const double* in) {
            launchXfwdRows<double>(in, C, 1, scNxh, scNxhP, twx, 0, s);
        };
        using YFn = void (*)(const cx<double>*, cx<double>*, int, int, int, int, int, int, const cx<double>*, int,
                             const int*, int, int, YDerivPlan, const double*, int);
        YFn yfn = k_yfft<double, 0, 0, false>;
        if (smooth7(d.c.Ny)) yfn = k_yfft<double, kSmooth7Rounds, 0, false>;
        if (l2y == 6 && l2t == 4) yfn = k_yfft<double, 6, 4, false>;
        if (l2y == 7 && l2t == 4) yfn = k_yfft<double, 7, 4, false>;
        if (l2y == 8 && l2t == 4) yfn = k_yfft<double, 8, 4, false>;
        if (l2y == 8 && l2t == 3) yfn = k_yfft<double, 8, 3, false>;
        if (l2y == 9 && l2t == 3) yfn = k_yfft<double, 9, 3, false>;
        if (l2y == 10 && l2t == 2) yfn = k_yfft<double, 10, 2, false>;
        auto yfft = [&](int nch, int sign) {  // in place: a workgroup loads its whole tile before it stores
            hipLaunchKernelGGL(yfn, dim3(static_cast<unsigned>(scNxhP / KT), static_cast<unsigned>(Nz), nch),
                               dim3(kFftThreads), (static_cast<size_t>(KT) + 1) * Ny * sizeof(cx<double>), s, C, C, Ny,
                               l2y, scNxh, scNxhP, l2t, Nz, scTwY.as<cx<double>>(), sign, nullptr, 0, 0, YDerivPlan{},
                               nullptr, KT * sizeof(cx<double>) < 64 ? 1 : 0);
        };
        auto xinv = [&](double* out, int nch) {
            launchXinvRows<double>(C, out, nch, scNxh, scNxhP, twx, s);
        };
        xfwd(scalar);
        yfft(1, -1);
        // one wave per 16-row z tile of d/dz, at most 16 waves
        const unsigned derivThreads = static_cast<unsigned>(64 * std::min(16, (Nz + 15) / 16));
        hipLaunchKernelGGL(k_scalarDeriv, dim3(static_cast<unsigned>((scNxhP + 7) / 8), static_cast<unsigned>(Ny)),
                           dim3(derivThreads),
                           static_cast<size_t>(Nz) * 8 * sizeof(cx<double>), s, C, Nx, Ny, Nz, scNxh, scNxhP, d.c.Lx,
                           d.c.Ly, scD.as<double>(), 1.0 / static_cast<double>(d.c.Nx * d.c.Ny));
        yfft(3, +1);
        {
            // c2r of d/dx, d/dy, d/dz with the product -(u, v, w) . grad in the epilogue
            const int rpb = xRowsPerBlock(3), ntr = (rpb * 3 + 1) / 2;
            const int nrows = Ny * Nz;
            auto pfn = k_xinv<double, kXinvAdvect, 0>;
            switch (log2Exact(d.c.Nx)) {
                case 6: pfn = k_xinv<double, kXinvAdvect, 6>; break;
                case 7: pfn = k_xinv<double, kXinvAdvect, 7>; break;
                case 8: pfn = k_xinv<double, kXinvAdvect, 8>; break;
                case 9: pfn = k_xinv<double, kXinvAdvect, 9>; break;
                default: break;
            }
            hipLaunchKernelGGL(pfn, dim3(static_cast<unsigned>((nrows + rpb - 1) / rpb)), dim3(256),
                               static_cast<size_t>(ntr) * ldsLength(Nx) * sizeof(cx<double>), s, C, rhs, 3, rpb, Nx,
                               log2Exact(d.c.Nx), Ny, Nz, scNxh, scNxhP, twx, adv, 0.0);
        }
        if (antialias) {
            xfwd(rhs);
            using MFn = void (*)(cx<double>*, int, int, int, int, int, const cx<double>*, const double*);
            MFn mfn = smooth7(d.c.Ny) ? k_ymask<kSmooth7Rounds, 0> : k_ymask<0, 0>;
            if (l2y == 6 && l2t == 4) mfn = k_ymask<6, 4>;
            if (l2y == 7 && l2t == 4) mfn = k_ymask<7, 4>;
            if (l2y == 8 && l2t == 4) mfn = k_ymask<8, 4>;
            if (l2y == 8 && l2t == 3) mfn = k_ymask<8, 3>;
            if (l2y == 9 && l2t == 3) mfn = k_ymask<9, 3>;
            if (l2y == 10 && l2t == 2) mfn = k_ymask<10, 2>;
            hipLaunchKernelGGL(mfn, dim3(static_cast<unsigned>(scNxhP / KT), static_cast<unsigned>(Nz)),
                               dim3(kFftThreads), (static_cast<size_t>(KT) + 1) * Ny * sizeof(cx<double>), s, C, Ny,
                               l2y, scNxh, scNxhP, l2t, scTwY.as<cx<double>>(), scKeep.as<double>());
            xinv(rhs, 1);
        }
        hipCheck(hipGetLastError(), "advectFGridScalar");
        // the reference's plan executions (...Kernel.cpp:1059-1129): r2c,
        // DCT-I and DST-I of the z derivative, c2r of the three gradients;
        // the antialias adds an r2c and a c2r
        ++metrics.scalarAdvectionCount;
        metrics.horizontalExecutionCount += antialias ? 4 : 2;
        metrics.verticalExecutionCount += 2;
        if (antialias) ++metrics.scalarAntialiasCount;
        metrics.executionCount = metrics.horizontalExecutionCount + metrics.verticalExecutionCount;
    }

    void rhsAny(const double* Ap, const double* Am, const double* A0, double t, double t0, const Epilogue& ep,
                hipStream_t s, bool fieldsReady = false, const double* adv = nullptr) {
        if (precision == WV_PRECISION_FP64) rhs<double>(Ap, Am, A0, t, t0, ep, s, fieldsReady, adv);
        else rhs<float>(Ap, Am, A0, t, t0, ep, s, fieldsReady, adv);
    }

    // transformStateFieldDerivatives (...Kernel.cpp:719-742): d/dx, d/dy and
    // d/dz of field f (0 u, 1 v, 2 w, 3 eta) at time t, out [3][Nz][Ny][Nx].
    // u and v are of the cosine (F) family, w and eta of the sine (G) family;
    // d/dz turns one family into the other (the DST-I / DCT-I with -/+ m_j / 2,
    // or for an arbitrary stratification -(N^2/g) G and F/h).
    template <typename T>
    void stateDerivatives(const double* Ap, const double* Am, const double* A0, double t, double t0, int f,
                          double* out, hipStream_t s) {
        launchFields<T>(Ap, Am, A0, t - t0, s);
        const bool cosFamily = f < 2;
        SynthPlan sp{};
        sp.valueScale = d.stratified ? 1.0 : 0.5;
        sp.ngroups = 2;
        sp.g[0] = {cosFamily ? 0 : 1, 1, {{f, 0, -1, 0, 1}, {}, {}, {}}};
        if (d.stratified)
            sp.g[1] = {cosFamily ? 2 : 3, 1, {{f, 0, 2, -1, -1}, {}, {}, {}}};
        else
            sp.g[1] = {cosFamily ? 1 : 0, 1, {{f, 1, 2, -1, -1}, {}, {}, {}}};
        launchSynth<T>(sp, s);
        launchYfft<T>(H, G, 3, +1, s, "yfft_inv", true, false);
        timed("xinv", 3.0 * d.R * (8.0 + 0.67 * sizeof(cx<T>)), s, [&] {
            launchXinvRows<T>(G.as<cx<T>>(), out, 3, static_cast<int>(d.nkx), static_cast<int>(d.nkxP), twX.as<cx<T>>(), s);
        });
        metrics.horizontalExecutionCount += 1;
        metrics.verticalExecutionCount += 3;
        metrics.executionCount = metrics.horizontalExecutionCount + metrics.verticalExecutionCount;
    }

    // Available potential vorticity (Operations/APVOperation.m:15-26) of the
    // state at time t, out [Nz][Ny][Nx].  eta_true = eta here: with the
    // constant-N background as the no-motion profile (the default,
    // shouldUseTrueNoMotionProfile = false) the density is linear in z, so
    // the inversion of EtaTrueOperation.m returns z - eta exactly.  One
    // synthesis of the nine derivative channels (kApvChannels order), the y
    // inverse, and the x inverse forming the APV products in LDS.
    template <typename T>
    void apv(const double* Ap, const double* Am, const double* A0, double t, double t0, double* out,
             hipStream_t s) {
        launchFields<T>(Ap, Am, A0, t - t0, s);
        SynthPlan sp{};
        sp.valueScale = 0.5;
        sp.ngroups = 4;
        // cosine values: u -> d/dy (0), v -> d/dx (2); sine values: w -> d/dx,
        // d/dy (4, 5), eta -> d/dx, d/dy (6, 7); z derivatives of the cosine
        // fields on the sine operator (1, 3) and of eta on the cosine one (8)
        sp.g[0] = {0, 2, {{0, 0, -1, -1, 0}, {1, 0, -1, 2, -1}, {}, {}}};
        sp.g[1] = {1, 2, {{2, 0, -1, 4, 5}, {3, 0, -1, 6, 7}, {}, {}}};
        sp.g[2] = {1, 2, {{0, 1, 1, -1, -1}, {1, 1, 3, -1, -1}, {}, {}}};
        sp.g[3] = {0, 1, {{3, 1, 8, -1, -1}, {}, {}, {}}};
        launchSynth<T>(sp, s);
        launchYfft<T>(H, G, kApvChannels, +1, s, "yfft_inv", true, false);
        timed("xapv", d.R * (8.0 + kApvChannels * 0.67 * sizeof(cx<T>)), s, [&] {
            launchXinvOp<T, kXinvApv>(G.as<cx<T>>(), out, kApvChannels, static_cast<int>(d.nkx),
                                      static_cast<int>(d.nkxP), twX.as<cx<T>>(), s, d.f);
        });
        metrics.horizontalExecutionCount += 1;
        metrics.verticalExecutionCount += 4;
        metrics.executionCount = metrics.horizontalExecutionCount + metrics.verticalExecutionCount;
    }

    template <typename T>
    void inverse(const double* Ap, const double* Am, const double* A0, double t, double t0, int nf, double* out,
                 hipStream_t s) {
        launchFields<T>(Ap, Am, A0, t - t0, s);
        SynthPlan sp{};
        sp.valueScale = d.stratified ? 1.0 : 0.5;
        sp.ngroups = 2;
        sp.g[0] = {0, 2, {{0, 0, 0, -1, -1}, {1, 0, 1, -1, -1}, {}, {}}};
        if (nf == 4) sp.g[1] = {1, 2, {{2, 0, 2, -1, -1}, {3, 0, 3, -1, -1}, {}, {}}};
        else sp.g[1] = {1, 1, {{2, 0, 2, -1, -1}, {}, {}, {}}};
        launchSynth<T>(sp, s);
        launchYfft<T>(H, G, nf, +1, s, "yfft_inv", true, false);
        timed("xinv", static_cast<double>(nf) * d.R * (8.0 + 0.67 * sizeof(cx<T>)), s, [&] {
            launchXinvRows<T>(G.as<cx<T>>(), out, nf, static_cast<int>(d.nkx), static_cast<int>(d.nkxP), twX.as<cx<T>>(), s);
        });
    }

    // Forward transform (...Kernel.cpp:583-684): r2c along x, FFT along y,
    // vertical analysis and the wave-vortex projection formulas.
    template <typename T>
    void forward(const double* fields, double t, double t0, int nf, double* Ap, double* Am, double* A0,
                 hipStream_t s) {
        timed("xfwd", static_cast<double>(nf) * d.R * (8.0 + 0.67 * sizeof(cx<T>)), s, [&] {
            launchXfwdRows<T>(fields, P.as<cx<T>>(), nf, static_cast<int>(d.nkx), static_cast<int>(d.nkxP), twX.as<cx<T>>(), 0, s);
        });
        launchYfft<T>(P, Q, nf, -1, s, "yfft_fwd", false, true);
        const int ncg = static_cast<int>((ncol + 7) / 8);
        const int waves = ncg * static_cast<int>(d.Ja / 16);
        ProjTables tb{omega.as<double>(), A0Vort.as<double>(), A0Buoy.as<double>(), NA0.as<double>(),
                      ApmN.as<double>(), ApmDScaled.as<double>(), ApmD.as<cx<double>>(), apmW.as<double>(),
                      inertialScale.as<double>()};
        const double hscale = 1.0 / static_cast<double>(d.c.Nx * d.c.Ny);
        const dim3 grid((waves + 3) / 4), block(256);
        const int ncolI = static_cast<int>(ncol), Ja = static_cast<int>(d.Ja), Za = static_cast<int>(d.Za),
                  Nz = static_cast<int>(d.c.Nz), Nj = static_cast<int>(d.c.Nj), Ny = static_cast<int>(d.c.Ny),
                  nkxP = static_cast<int>(d.nkxP);
        auto* oAp = reinterpret_cast<cx<double>*>(Ap);
        auto* oAm = reinterpret_cast<cx<double>*>(Am);
        auto* oA0 = reinterpret_cast<cx<double>*>(A0);
        timed("projfwd", static_cast<double>(nf) * d.c.Nz * ncol * sizeof(cx<T>) + 48.0 * d.M, s, [&] {
            if (nf == 4)
                hipLaunchKernelGGL((k_projfwd<T, 4>), grid, block, 0, s, Q.as<cx<T>>(), cols.as<ColumnDev>(), ncolI,
                                   anaC.as<T>(), anaS.as<T>(), Ja, Za, Nz, Nj, Ny, nkxP, tb, t - t0, hscale, oAp, oAm,
                                   oA0);
            else
                hipLaunchKernelGGL((k_projfwd<T, 3>), grid, block, 0, s, Q.as<cx<T>>(), cols.as<ColumnDev>(), ncolI,
                                   anaC.as<T>(), anaS.as<T>(), Ja, Za, Nz, Nj, Ny, nkxP, tb, t - t0, hscale, oAp, oAm,
                                   oA0);
        });
    }

    // transformToSpatialDomainWith{F,G}AllDerivatives (...Kernel.cpp:819-863):
    // value, d/dx, d/dy, d/dz of an F- (cosine) or G- (sine) family field.
    template <typename T>
    void fgAll(const double* Apm, const double* A0, bool fFamily, double* out, hipStream_t s) {
        const int M = static_cast<int>(d.M);
        timed("fgfields", 48.0 * M + M * sizeof(cx<T>), s, [&] {
            hipLaunchKernelGGL(k_fgfields<T>, dim3((M + 255) / 256), dim3(256), 0, s,
                               reinterpret_cast<const cx<double>*>(Apm), reinterpret_cast<const cx<double>*>(A0),
                               fWaveScale.as<double>(), gWaveScale.as<double>(), Fg.as<double>(), Gg.as<double>(),
                               static_cast<int>(d.c.Nj), M, fFamily ? 1 : 0, FV.as<cx<T>>());
        });
        SynthPlan sp{};
        sp.ngroups = 2;
        const int a = fFamily ? 0 : 1, b = 1 - a;
        sp.valueScale = d.stratified ? 1.0 : 0.5;
        sp.g[0] = {a, 1, {{0, 0, 0, 1, 2}, {}, {}, {}}};
        if (d.stratified)  // d/dz of F is -(N^2/g) G, of G is F/h
            sp.g[1] = {fFamily ? 2 : 3, 1, {{0, 0, 3, -1, -1}, {}, {}, {}}};
        else
            sp.g[1] = {b, 1, {{0, 1, 3, -1, -1}, {}, {}, {}}};
        launchSynth<T>(sp, s);
        launchYfft<T>(H, G, 4, +1, s, "yfft_inv", true, false);
        timed("xinv", 4.0 * d.R * (8.0 + 0.67 * sizeof(cx<T>)), s, [&] {
            launchXinvRows<T>(G.as<cx<T>>(), out, 4, static_cast<int>(d.nkx), static_cast<int>(d.nkxP), twX.as<cx<T>>(), s);
        });
    }

    // Workgroups of the energy reduction: one partial sum each.
    int energyBlocks() const {
        return std::max(1, std::min(kEnergyPartials, static_cast<int>((d.M + 255) / 256)));
    }
    // The total-energy reduction with its partial sums in `partials`
    // (energyBlocks() of them) and the result in `out`, both device memory.
    void energyInto(const double* Ap, const double* Am, const double* A0, double* partials, double* out,
                    hipStream_t s) const {
        const int M = static_cast<int>(d.M);
        const int blocks = energyBlocks();
        hipLaunchKernelGGL(k_energy_partial, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const cx<double>*>(Ap),
                           reinterpret_cast<const cx<double>*>(Am), reinterpret_cast<const cx<double>*>(A0),
                           ApmTE.as<double>(), A0TE.as<double>(), M, partials);
        hipCheck(hipGetLastError(), "energy");
        hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, s, partials, blocks, out);
        hipCheck(hipGetLastError(), "energy");
    }

    double energy(const double* Ap, const double* Am, const double* A0, hipStream_t s) {
        energyInto(Ap, Am, A0, partial.as<double>(), scalar.as<double>(), s);
        double e = 0.0;
        hipCheck(hipMemcpyAsync(&e, scalar.p, sizeof(double), hipMemcpyDeviceToHost, s), "energy copy");
        hipCheck(hipStreamSynchronize(s), "energy sync");
        return e;
    }

    // ---- adjoint ---------------------------------------------------------------
    // Buffers and transposed operators of the vector-Jacobian products,
    // allocated on the first adjoint call (none of the forward entry points
    // touch them).
    bool adjointReady = false;
    DeviceBuffer anaTC, anaTS, synT[kSynthMats], Qs, Gs, colbuf, mainColumn, compColumn;
    DeviceBuffer rkWork;  // 7 coefficient triples: s2, s3, s4, kbar, acc, tmp, ywork

    static std::vector<double> transposeOp(const std::vector<double>& m, size_t rows, size_t cols, size_t orow,
                                           size_t ocol, size_t nr, size_t nc) {
        // out[orow x ocol] with out[c][r] = m[r][c] for r < nr, c < nc
        std::vector<double> out(orow * ocol, 0.0);
        for (size_t r = 0; r < nr && r < rows; ++r)
            for (size_t c = 0; c < nc && c < cols; ++c) out[c * ocol + r] = m[r * cols + c];
        return out;
    }

    template <typename T>
    void setupAdjointTyped() {
        const size_t Nz = d.c.Nz, Nj = d.c.Nj;
        // analysis [Ja x Za] -> synthesis layout [Zs x Js]
        uploadAs<T>(anaTC, mfmaOperand(transposeOp(d.anaC, d.Ja, d.Za, d.Zs, d.Js, Nj, Nz), d.Zs, d.Js));
        uploadAs<T>(anaTS, mfmaOperand(transposeOp(d.anaS, d.Ja, d.Za, d.Zs, d.Js, Nj, Nz), d.Zs, d.Js));
        // synthesis [Zs x Js] -> analysis layout [Ja x Za]
        const std::vector<double>* syn[kSynthMats] = {&d.synC, &d.synS, &d.synGz, &d.synFh, &d.synEz};
        for (int m = 0; m < kSynthMats; ++m)
            if (!syn[m]->empty())
                uploadAs<T>(synT[m], mfmaOperand(transposeOp(*syn[m], d.Zs, d.Js, d.Ja, d.Za, Nz, Nj), d.Ja, d.Za));
    }

    void ensureAdjoint() {
        if (adjointReady) return;
        if (precision == WV_PRECISION_FP64) setupAdjointTyped<double>();
        else setupAdjointTyped<float>();
        const size_t plane = static_cast<size_t>(d.c.Ny) * d.nkxP, chField = d.c.Nz * plane;
        Qs.alloc(static_cast<size_t>(NT) * chField * tsize);
        Gs.alloc(static_cast<size_t>(NH) * chField * tsize);
        hipCheck(hipMemset(Qs.p, 0, Qs.bytes), "hipMemset");
        hipCheck(hipMemset(Gs.p, 0, Gs.bytes), "hipMemset");
        colbuf.alloc(4 * ncol * d.c.Nj * tsize);
        std::vector<int> comp(d.Nkl, -1);
        for (size_t ic = 0; ic < d.columns.size(); ++ic)
            if (d.columns[ic].flags & kColCompletion) comp[static_cast<size_t>(d.columns[ic].wv)] = static_cast<int>(ic);
        upload(mainColumn, std::vector<int>(d.wvColumn.begin(), d.wvColumn.end()));
        upload(compColumn, comp);
        rkWork.alloc(7 * 3 * d.M * sizeof(cx<double>));
        hipCheck(hipDeviceSynchronize(), "adjoint setup");
        adjointReady = true;
    }

    template <typename T>
    SynthMats<T> anaTMats() const {
        return SynthMats<T>{{anaTC.as<T>(), anaTS.as<T>(), nullptr, nullptr, nullptr}};
    }
    template <typename T>
    SynthMats<T> synTMats() const {
        return SynthMats<T>{{synT[0].as<T>(), synT[1].as<T>(), synT[2].as<T>(), synT[3].as<T>(), synT[4].as<T>()}};
    }

    // P^T as a synthesis: per-target cotangents of the analysed spectra (in
    // FV) through the transposed analysis matrices into the main columns of Qs.
    SynthPlan projAdjPlan() const {
        SynthPlan sp{};
        sp.valueScale = 1.0;
        sp.skipCompletion = 1;
        sp.ngroups = 2;
        sp.g[0] = {0, 2, {{0, 0, 0, -1, -1}, {1, 0, 1, -1, -1}, {}, {}}};
        if (NT == 4) sp.g[1] = {1, 2, {{2, 0, 2, -1, -1}, {3, 0, 3, -1, -1}, {}, {}}};
        else sp.g[1] = {1, 1, {{2, 0, 2, -1, -1}, {}, {}, {}}};
        return sp;
    }

    template <typename T>
    void launchSynthAdj(const SynthPlan& sp, const DeviceBuffer& Hstar, hipStream_t s) {
        const int ncg = static_cast<int>((ncol + 7) / 8);
        const int njt = static_cast<int>(d.Ja / 16);
        const size_t lds = 4 * d.Za * 8 * sizeof(cx<T>);
        timed("synthadj", 0.0, s, [&] {
            hipLaunchKernelGGL(k_synthadj<T>, dim3(static_cast<unsigned>(ncg)), dim3(static_cast<unsigned>(64 * njt)),
                               lds, s, Hstar.as<cx<T>>(), cols.as<ColumnDev>(), static_cast<int>(ncol),
                               synTMats<T>(), static_cast<int>(d.Ja), static_cast<int>(d.Za),
                               static_cast<int>(d.c.Nz), static_cast<int>(d.c.Nj), static_cast<int>(d.c.Ny),
                               static_cast<int>(d.nkxP), mj.as<T>(), sp, colbuf.as<cx<T>>());
        });
    }

    template <typename T>
    void launchFieldsAdj(double elapsed, int mode, double* Gp, double* Gm, double* G0, hipStream_t s) {
        const int M = static_cast<int>(d.M);
        timed("fieldsadj", 0.0, s, [&] {
            hipLaunchKernelGGL(k_fieldsadj<T>, dim3((M + 255) / 256), dim3(256), 0, s, colbuf.as<cx<T>>(),
                               static_cast<int>(ncol), mainColumn.as<int>(), compColumn.as<int>(), omega.as<double>(),
                               UAp.as<cx<double>>(), VAp.as<cx<double>>(), WAp.as<cx<double>>(), NAp.as<double>(),
                               UA0.as<cx<double>>(), VA0.as<cx<double>>(), NA0.as<double>(), elapsed,
                               static_cast<int>(d.c.Nj), M, mode, reinterpret_cast<cx<double>*>(Gp),
                               reinterpret_cast<cx<double>*>(Gm), reinterpret_cast<cx<double>*>(G0));
        });
    }

    // (Gp, Gm, G0) = J(y)^T (Lp, Lm, L0) for the flux at time t.
    template <typename T>
    void fluxVjp(const double* Ap, const double* Am, const double* A0, double t, double t0, const double* Lp,
                 const double* Lm, const double* L0, double* Gp, double* Gm, double* G0, hipStream_t s) {
        const double elapsed = t - t0;
        // forward spectra at y (as rhs)
        launchFields<T>(Ap, Am, A0, elapsed, s);
        launchSynth<T>(fluxPlan(), s);
        launchYfft<T>(H, G, NH, +1, s, "yfft_inv", true, false);
        // P^T
        const int M = static_cast<int>(d.M);
        ProjTables tb{omega.as<double>(), A0Vort.as<double>(), A0Buoy.as<double>(), NA0.as<double>(),
                      ApmN.as<double>(), ApmDScaled.as<double>(), ApmD.as<cx<double>>(), apmW.as<double>(),
                      inertialScale.as<double>()};
        const double hscale = 1.0 / static_cast<double>(d.c.Nx * d.c.Ny);
        const int hydro = d.c.isHydrostatic ? 1 : 0;
        timed("projadj", 0.0, s, [&] {
            auto* lp = reinterpret_cast<const cx<double>*>(Lp);
            auto* lm = reinterpret_cast<const cx<double>*>(Lm);
            auto* l0 = reinterpret_cast<const cx<double>*>(L0);
            if (NT == 4)
                hipLaunchKernelGGL((k_projadj<T, 4>), dim3((M + 255) / 256), dim3(256), 0, s, lp, lm, l0,
                                   cols.as<ColumnDev>(), mainColumn.as<int>(), static_cast<int>(d.c.Nj), M,
                                   targetCode, hydro, tb, elapsed, hscale, FV.as<cx<T>>());
            else
                hipLaunchKernelGGL((k_projadj<T, 3>), dim3((M + 255) / 256), dim3(256), 0, s, lp, lm, l0,
                                   cols.as<ColumnDev>(), mainColumn.as<int>(), static_cast<int>(d.c.Nj), M,
                                   targetCode, hydro, tb, elapsed, hscale, FV.as<cx<T>>());
        });
        const SynthMats<T> at = anaTMats<T>();
        launchSynth<T>(projAdjPlan(), s, &FV, &Qs, &at, "synth_projadj");
        launchYfft<T>(Qs, P, NT, +1, s, "yfft_projadj", true, false);
        // products transposed along x
        const int Nx = static_cast<int>(d.c.Nx), l2 = log2Exact(d.c.Nx), Ny = static_cast<int>(d.c.Ny),
                  Nz = static_cast<int>(d.c.Nz), nkx = static_cast<int>(d.nkx), nkxP = static_cast<int>(d.nkxP);
        const int pairs = NT == 4 ? XMap<4>::pairs : XMap<3>::pairs;
        const size_t lds = (static_cast<size_t>(Nx) + static_cast<size_t>(pairs + 2) * xfusedStride<T>(Nx)) *
                           sizeof(cx<T>);
        if (lds > maxLdsPerBlock)
            throw Failure{WV_UNSUPPORTED_OPERATION, "Adjoint x transforms exceed the LDS of a workgroup."};
        timed("xadj", 0.0, s, [&] {
            const dim3 grid(static_cast<unsigned>(Ny * Nz));
            if (NT == 4)
                hipLaunchKernelGGL((k_xadj<T, 4>), grid, dim3(256), lds, s, G.as<cx<T>>(), P.as<cx<T>>(),
                                   Gs.as<cx<T>>(), Nx, l2, Ny, Nz, nkx, nkxP, twX.as<cx<T>>(), kxv.as<T>());
            else
                hipLaunchKernelGGL((k_xadj<T, 3>), grid, dim3(256), lds, s, G.as<cx<T>>(), P.as<cx<T>>(),
                                   Gs.as<cx<T>>(), Nx, l2, Ny, Nz, nkx, nkxP, twX.as<cx<T>>(), kxv.as<T>());
        });
        launchYfft<T>(Gs, G, NH, -1, s, "yfft_adj", false, true);
        launchSynthAdj<T>(fluxPlan(), G, s);
        launchFieldsAdj<T>(elapsed, 0, Gp, Gm, G0, s);
    }

    // Transpose of transformWaveVortexToUVWEta: cotangent fields [4][Nz][Ny][Nx]
    // (u, v, w, eta) -> coefficient cotangents.
    template <typename T>
    void inverseVjp(const double* fieldsBar, double t, double t0, double* Gp, double* Gm, double* G0, hipStream_t s) {
        const int nf = 4;
        timed("xfwd_adj", 0.0, s, [&] {
            launchXfwdRows<T>(fieldsBar, Gs.as<cx<T>>(), nf, static_cast<int>(d.nkx), static_cast<int>(d.nkxP), twX.as<cx<T>>(), 1, s);
        });
        launchYfft<T>(Gs, G, nf, -1, s, "yfft_adj", false, true);
        SynthPlan sp{};
        sp.valueScale = d.stratified ? 1.0 : 0.5;
        sp.ngroups = 2;
        sp.g[0] = {0, 2, {{0, 0, 0, -1, -1}, {1, 0, 1, -1, -1}, {}, {}}};
        sp.g[1] = {1, 2, {{2, 0, 2, -1, -1}, {3, 0, 3, -1, -1}, {}, {}}};
        launchSynthAdj<T>(sp, G, s);
        launchFieldsAdj<T>(t - t0, 0, Gp, Gm, G0, s);
    }

    void fluxVjpAny(const double* Ap, const double* Am, const double* A0, double t, double t0, const double* Lp,
                    const double* Lm, const double* L0, double* Gp, double* Gm, double* G0, hipStream_t s) {
        ensureAdjoint();
        if (precision == WV_PRECISION_FP64) fluxVjp<double>(Ap, Am, A0, t, t0, Lp, Lm, L0, Gp, Gm, G0, s);
        else fluxVjp<float>(Ap, Am, A0, t, t0, Lp, Lm, L0, Gp, Gm, G0, s);
    }

    // out_a = alpha x_a + beta y_a over the coefficient triples (pointers to
    // the first of three consecutive arrays of M complex doubles each).
    void axpby(const cx<double>* x, const cx<double>* y, double alpha, double beta, cx<double>* out, hipStream_t s) {
        const int M = static_cast<int>(d.M);
        hipLaunchKernelGGL(k_axpby3, dim3((M + 255) / 256), dim3(256), 0, s, x, x + M, x + 2 * M, y, y + M, y + 2 * M,
                           alpha, beta, M, out, out + M, out + 2 * M);
        hipCheck(hipGetLastError(), "axpby");
    }

    void fluxInto(const cx<double>* y, double t, double t0, cx<double>* out, hipStream_t s) {
        const size_t M = d.M;
        Epilogue ep{};
        ep.mode = kFluxOut;
        ep.out0 = out;
        ep.out1 = out + M;
        ep.out2 = out + 2 * M;
        const double* yd = reinterpret_cast<const double*>(y);
        rhsAny(yd, yd + 2 * M, yd + 4 * M, t, t0, ep, s);
    }

    void vjpInto(const cx<double>* y, double t, double t0, const cx<double>* lam, cx<double>* out, hipStream_t s) {
        const size_t M = d.M;
        const double* yd = reinterpret_cast<const double*>(y);
        const double* ld = reinterpret_cast<const double*>(lam);
        double* od = reinterpret_cast<double*>(out);
        fluxVjpAny(yd, yd + 2 * M, yd + 4 * M, t, t0, ld, ld + 2 * M, ld + 4 * M, od, od + 2 * M, od + 4 * M, s);
    }

    // One classic RK4 step's transpose: ybar <- (d y_{n+1} / d y_n)^T ybar, the
    // stages recomputed from y_n exactly as the forward step forms them.
    void rk4StepVjp(const cx<double>* y, double t, double t0, double h, cx<double>* ybar, hipStream_t s) {
        ensureAdjoint();
        const size_t M3 = 3 * d.M;
        cx<double>* w = rkWork.as<cx<double>>();
        cx<double>* s2 = w;
        cx<double>* s3 = w + M3;
        cx<double>* s4 = w + 2 * M3;
        cx<double>* kb = w + 3 * M3;
        cx<double>* acc = w + 4 * M3;
        cx<double>* tmp = w + 5 * M3;
        fluxInto(y, t, t0, tmp, s);
        axpby(y, tmp, 1.0, 0.5 * h, s2, s);
        fluxInto(s2, t + 0.5 * h, t0, tmp, s);
        axpby(y, tmp, 1.0, 0.5 * h, s3, s);
        fluxInto(s3, t + 0.5 * h, t0, tmp, s);
        axpby(y, tmp, 1.0, h, s4, s);
        axpby(ybar, ybar, h / 6.0, 0.0, kb, s);  // kbar4
        vjpInto(s4, t + h, t0, kb, tmp, s);
        axpby(ybar, tmp, 1.0, 1.0, acc, s);
        axpby(ybar, tmp, h / 3.0, h, kb, s);  // kbar3
        vjpInto(s3, t + 0.5 * h, t0, kb, tmp, s);
        axpby(acc, tmp, 1.0, 1.0, acc, s);
        axpby(ybar, tmp, h / 3.0, 0.5 * h, kb, s);  // kbar2
        vjpInto(s2, t + 0.5 * h, t0, kb, tmp, s);
        axpby(acc, tmp, 1.0, 1.0, acc, s);
        axpby(ybar, tmp, h / 6.0, 0.5 * h, kb, s);  // kbar1
        vjpInto(y, t, t0, kb, tmp, s);
        axpby(acc, tmp, 1.0, 1.0, ybar, s);
    }

    // One forward RK4 step in place (the arithmetic of wv_rk4_advance_device).
    // nsteps classic RK4 steps of the state y[0..2] in place, *t advanced.
    // Each stage's epilogue forms the next flux evaluation's field spectra
    // (with the same times as WVFixedStepRK4::step: t + h/2, t + h/2, t + h),
    // so only the first evaluation of the call starts from the coefficients.
    void rk4Advance(cx<double>* const y[3], double* t, double t0, double h, int nsteps, hipStream_t s) {
        rk4AdvanceSteps(y, t, t0, std::vector<double>(static_cast<size_t>(std::max(nsteps, 0)), h), s);
    }

    // Classic RK4 steps of the given sizes.  A step's stages 1-3 write only
    // the workspace; its stage 4 forms y + h/6 weighted in place, so every
    // step is committed by its last launch: when a launch fails, y holds the
    // state after the last whole step and *t that step's end time.
    void rk4AdvanceSteps(cx<double>* const y[3], double* t, double t0, const std::vector<double>& steps,
                         hipStream_t s) {
        const int nsteps = static_cast<int>(steps.size());
        const size_t M = d.M;
        cx<double>* st = stage.as<cx<double>>();
        cx<double>* w = weighted.as<cx<double>>();
        Epilogue ep{};
        ep.y0 = y[0];
        ep.y1 = y[1];
        ep.y2 = y[2];
        ep.st0 = st;
        ep.st1 = st + M;
        ep.st2 = st + 2 * M;
        ep.w0 = w;
        ep.w1 = w + M;
        ep.w2 = w + 2 * M;
        ep.emitFields = 1;
        const double* sd = reinterpret_cast<const double*>(st);
        const double* y0 = reinterpret_cast<const double*>(y[0]);
        const double* y1 = reinterpret_cast<const double*>(y[1]);
        const double* y2 = reinterpret_cast<const double*>(y[2]);
        for (int n = 0; n < nsteps; ++n) {
            const double ti = *t;
            const double h = steps[static_cast<size_t>(n)];
            ep.mode = kRk1;
            ep.scale = 0.5 * h;
            ep.nextElapsed = (ti + 0.5 * h) - t0;
            rhsAny(y0, y1, y2, ti, t0, ep, s, n > 0);
            ep.mode = kRk2;
            rhsAny(sd, sd + 2 * M, sd + 4 * M, ti + 0.5 * h, t0, ep, s, true);
            ep.mode = kRk3;
            ep.scale = h;
            ep.nextElapsed = (ti + h) - t0;
            rhsAny(sd, sd + 2 * M, sd + 4 * M, ti + 0.5 * h, t0, ep, s, true);
            ep.mode = kRk4;
            ep.scale = h / 6.0;
            ep.emitFields = n + 1 < nsteps ? 1 : 0;  // the next step's first evaluation
            rhsAny(sd, sd + 2 * M, sd + 4 * M, ti + h, t0, ep, s, true);
            ep.emitFields = 1;
            *t = ti + h;
            ++rkStepCount;
            lastStepSize = h;
        }
    }

    void rk4Step(cx<double>* y, double t, double t0, double h, hipStream_t s) {
        cx<double>* const yy[3] = {y, y + d.M, y + 2 * d.M};
        rk4Advance(yy, &t, t0, h, 1, s);
    }

    // Transpose of nsteps RK4 steps from y0 at time t: ybar (cotangent of the
    // final state) <- cotangent of y0.  The forward states are recomputed
    // from checkpoints stored every `interval` steps (held on the device only
    // for the duration of the call).
    void rk4Adjoint(const cx<double>* y0, double t, double t0, double h, int nsteps, int interval, cx<double>* ybar,
                    hipStream_t s) {
        ensureAdjoint();
        const size_t M3 = 3 * d.M, bytes = M3 * sizeof(cx<double>);
        const int ncp = (nsteps + interval - 1) / interval;
        DeviceBuffer cp, seg;
        cp.alloc(static_cast<size_t>(std::max(ncp, 1)) * bytes);
        seg.alloc(static_cast<size_t>(interval) * bytes);
        cx<double>* yw = rkWork.as<cx<double>>() + 6 * M3;
        // step start times accumulated as wv_rk4_advance_device accumulates them
        std::vector<double> tn(static_cast<size_t>(nsteps) + 1, t);
        for (int n = 0; n < nsteps; ++n) tn[n + 1] = tn[n] + h;
        hipCheck(hipMemcpyAsync(yw, y0, bytes, hipMemcpyDeviceToDevice, s), "copy");
        for (int n = 0; n < nsteps; ++n) {
            if (n % interval == 0)
                hipCheck(hipMemcpyAsync(cp.as<cx<double>>() + static_cast<size_t>(n / interval) * M3, yw, bytes,
                                        hipMemcpyDeviceToDevice, s), "checkpoint");
            if (n + 1 < nsteps) rk4Step(yw, tn[n], t0, h, s);  // the final state is not needed
        }
        for (int c = ncp - 1; c >= 0; --c) {
            const int n0 = c * interval, n1 = std::min(nsteps, n0 + interval);
            hipCheck(hipMemcpyAsync(yw, cp.as<cx<double>>() + static_cast<size_t>(c) * M3, bytes,
                                    hipMemcpyDeviceToDevice, s), "restore");
            for (int n = n0; n < n1; ++n) {
                hipCheck(hipMemcpyAsync(seg.as<cx<double>>() + static_cast<size_t>(n - n0) * M3, yw, bytes,
                                        hipMemcpyDeviceToDevice, s), "segment");
                if (n + 1 < n1) rk4Step(yw, tn[n], t0, h, s);
            }
            for (int n = n1 - 1; n >= n0; --n)
                rk4StepVjp(seg.as<cx<double>>() + static_cast<size_t>(n - n0) * M3, tn[n], t0, h, ybar, s);
        }
        hipCheck(hipStreamSynchronize(s), "adjoint");
    }
};

}  // namespace wvamd

// ==================================================================== C-ABI ==
using wvamd::Failure;
using wvamd::KernelImpl;
using wvamd::hipCheck;

struct wv_kernel {
    KernelImpl impl;
};

namespace {

wv_status fail(wv_status code, const std::string& msg) {
    wvamd::gLastError = msg;
    return code;
}

template <typename F>
wv_status guarded(F&& f) {
    try {
        return f();
    } catch (const Failure& e) {
        return fail(e.code, e.message);
    } catch (const std::bad_alloc&) {
        return fail(WV_ALLOCATION_FAILURE, "Host allocation failed.");
    } catch (const std::overflow_error& e) {
        return fail(WV_SIZE_OVERFLOW, e.what());
    } catch (const std::exception& e) {
        return fail(WV_NUMERICAL_FAILURE, e.what());
    }
}

struct ExecGuard {
    std::atomic<bool>& flag;
    bool entered;
    explicit ExecGuard(std::atomic<bool>& f) : flag(f) {
        bool expected = false;
        entered = flag.compare_exchange_strong(expected, true);
    }
    ~ExecGuard() {
        if (entered) flag.store(false);
    }
};

hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

// One call on a handle: the reentrancy guard (ExecutionGuard, ...Kernel.cpp:76-84),
// the handle's device, ordering after the handle's previous call, the work,
// and the event that orders the next call.  body(s) throws Failure on error.
template <typename F>
wv_status execute(KernelImpl& k, void* stream, const char* busy, F&& body) {
    ExecGuard g(k.executing);
    if (!g.entered) return fail(WV_REENTRANT_EXECUTION, busy);
    return guarded([&] {
        hipCheck(hipSetDevice(k.device), "hipSetDevice");
        const hipStream_t s = S(stream);
        k.orderAfterLastCall(s);
        body(s);
        k.recordCall(s);
        return WV_SUCCESS;
    });
}
constexpr const char* kNotReentrant = "Kernel operations are not reentrant.";

// Device address of `scratch` on each GPU, taken once per device at create
// with that device current (the array's name in host code is a host-side
// stand-in, never a device address).
constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
double* gScratch[kMaxDevices] = {};

void registerScratch(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw Failure{WV_UNSUPPORTED_OPERATION, "Device index beyond the library's per-device scratch table."};
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (gScratch[device]) return;
    void* ptr = nullptr;
    hipCheck(hipSetDevice(device), "hipSetDevice");
    hipCheck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(wvamd::scratch)), "hipGetSymbolAddress(scratch)");
    gScratch[device] = static_cast<double*>(ptr);
}

double* scratchOf(int device) {
    std::lock_guard<std::mutex> lock(gScratchMutex);
    return device >= 0 && device < kMaxDevices ? gScratch[device] : nullptr;
}

bool overlaps(const void* a, size_t an, const void* b, size_t bn) {
    if (!a || !b || an == 0 || bn == 0) return false;
    const auto x = reinterpret_cast<uintptr_t>(a), y = reinterpret_cast<uintptr_t>(b);
    return x < y + bn && y < x + an;
}

// validateStateAndFlux (WVKernelTypes.cpp:343-362) for device pointers.
wv_status validateFlux(const KernelImpl& k, double t, double t0, const void* Ap, const void* Am, const void* A0,
                       const void* Fp, const void* Fm, const void* F0) {
    if (!std::isfinite(t) || !std::isfinite(t0)) return fail(WV_INVALID_CONFIGURATION, "State times must be finite.");
    const char* names[6] = {"Ap", "Am", "A0", "Fp", "Fm", "F0"};
    const void* ptrs[6] = {Ap, Am, A0, Fp, Fm, F0};
    for (int i = 0; i < 6; ++i)
        if (!ptrs[i] && k.d.M) return fail(WV_INVALID_POINTER, std::string(names[i]) + " has a null data pointer.");
    const size_t bytes = k.d.M * 16;
    for (int i = 3; i < 6; ++i) {
        for (int j = i + 1; j < 6; ++j)
            if (overlaps(ptrs[i], bytes, ptrs[j], bytes)) return fail(WV_OVERLAPPING_ARRAYS, "Flux outputs must not overlap.");
        for (int j = 0; j < 3; ++j)
            if (overlaps(ptrs[i], bytes, ptrs[j], bytes))
                return fail(WV_OVERLAPPING_ARRAYS, "Flux outputs must not overlap state inputs.");
    }
    return WV_SUCCESS;
}

// validateInverseOwnership: a spatial bundle of nf fields may not overlap the
// state's coefficient arrays.
wv_status validateBundleOwnership(const KernelImpl& k, const void* fields, int nf, const void* Ap, const void* Am,
                                  const void* A0, const char* what) {
    const size_t cb = k.d.M * 16, fb = k.d.R * static_cast<size_t>(nf) * sizeof(double);
    for (const void* in : {Ap, Am, A0})
        if (overlaps(fields, fb, in, cb)) return fail(WV_OVERLAPPING_ARRAYS, std::string(what) + " must not overlap coefficient inputs.");
    return WV_SUCCESS;
}

// The three arrays of a state or cotangent must not overlap one another.
bool overlapping3(const void* a, const void* b, const void* c, size_t bytes) {
    return overlaps(a, bytes, b, bytes) || overlaps(a, bytes, c, bytes) || overlaps(b, bytes, c, bytes);
}

}  // namespace

extern "C" {

const char* wv_last_error(void) { return wvamd::gLastError.c_str(); }
uint32_t wv_contract_version(void) { return WV_CONTRACT_VERSION; }

wv_status wv_create(const wv_config* config, wv_precision precision, int device, wv_kernel** kernel) {
    if (!config || !kernel) return fail(WV_INVALID_POINTER, "wv_create requires a configuration and an output handle.");
    *kernel = nullptr;
    if (precision != WV_PRECISION_FP64 && precision != WV_PRECISION_MIXED && precision != WV_PRECISION_MIXED_SPLIT16)
        return fail(WV_INVALID_CONFIGURATION, "Unknown precision.");
    return guarded([&] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            return fail(WV_UNSUPPORTED_OPERATION, "No HIP device is available.");
        if (device < 0 || device >= count) return fail(WV_INVALID_CONFIGURATION, "Invalid HIP device index.");
        std::unique_ptr<wv_kernel> k(new wv_kernel());
        k->impl.setup(*config, precision, device);
        registerScratch(device);
        *kernel = k.release();
        return WV_SUCCESS;
    });
}

wv_status wv_create_hydrostatic(const wv_config* config, const wv_vertical_operators* ops, wv_precision precision,
                                int device, wv_kernel** kernel) {
    if (!config || !ops || !kernel)
        return fail(WV_INVALID_POINTER, "wv_create_hydrostatic requires a configuration, operators and an output handle.");
    *kernel = nullptr;
    if (precision != WV_PRECISION_FP64 && precision != WV_PRECISION_MIXED && precision != WV_PRECISION_MIXED_SPLIT16)
        return fail(WV_INVALID_CONFIGURATION, "Unknown precision.");
    for (const double* p : {ops->z, ops->N2, ops->dLnN2, ops->h, ops->P, ops->Q, ops->PFinv, ops->PF, ops->QGinv, ops->QG})
        if (!p) return fail(WV_INVALID_POINTER, "Vertical operator pointers must be non-null.");
    if (!config->isHydrostatic)
        return fail(WV_INVALID_CONFIGURATION, "Arbitrary stratification is supported for the hydrostatic transform only.");
    if (config->Nz < 3 || config->Nj < 1 || config->Nj > config->Nz - 1)
        return fail(WV_INVALID_CONFIGURATION, "Nz must be at least 3 and Nj must lie in [1,Nz-1].");
    return guarded([&] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            return fail(WV_UNSUPPORTED_OPERATION, "No HIP device is available.");
        if (device < 0 || device >= count) return fail(WV_INVALID_CONFIGURATION, "Invalid HIP device index.");
        const size_t nz = config->Nz, nj = config->Nj;
        wvamd::Stratification st;
        st.z.assign(ops->z, ops->z + nz);
        st.N2.assign(ops->N2, ops->N2 + nz);
        st.dLnN2.assign(ops->dLnN2, ops->dLnN2 + nz);
        st.h.assign(ops->h, ops->h + nj);
        st.P.assign(ops->P, ops->P + nj);
        st.Q.assign(ops->Q, ops->Q + nj);
        st.PFinv.assign(ops->PFinv, ops->PFinv + nz * nj);
        st.QGinv.assign(ops->QGinv, ops->QGinv + nz * nj);
        st.PF.assign(ops->PF, ops->PF + nj * nz);
        st.QG.assign(ops->QG, ops->QG + nj * nz);
        std::unique_ptr<wv_kernel> k(new wv_kernel());
        k->impl.setup(*config, precision, device, &st);
        registerScratch(device);
        *kernel = k.release();
        return WV_SUCCESS;
    });
}

void wv_destroy(wv_kernel* kernel) {
    if (!kernel) return;
    (void)hipSetDevice(kernel->impl.device);
    (void)hipDeviceSynchronize();
    for (auto& pd : kernel->impl.pending) {
        (void)hipEventDestroy(pd.a);
        (void)hipEventDestroy(pd.b);
    }
    if (kernel->impl.lastCall) (void)hipEventDestroy(kernel->impl.lastCall);
    delete kernel;
}

size_t wv_nkl(const wv_kernel* kernel) { return kernel ? kernel->impl.d.Nkl : 0; }
double wv_coriolis_frequency(const wv_kernel* kernel) { return kernel ? kernel->impl.d.f : 0.0; }
size_t wv_persistent_bytes(const wv_kernel* kernel) { return kernel ? kernel->impl.persistentBytes() : 0; }
size_t wv_scratch_bytes(const wv_kernel* kernel) { return kernel ? kernel->impl.scratchBytes() : 0; }

wv_status wv_metrics(const wv_kernel* kernel, wv_kernel_metrics* out) {
    if (!kernel || !out) return fail(WV_INVALID_POINTER, "wv_metrics requires a kernel and an output.");
    const KernelImpl& k = kernel->impl;
    wv_kernel_metrics m = k.metrics;
    m.descriptorBytes = k.descriptorBytes();
    m.planBytes = k.planBytes();
    m.engineBytes = 0;
    m.kernelManagementBytes = sizeof(wv_kernel);
    m.halfSpectrumScratchCapacityBytes = k.scratchBytes();
    m.realScratchCapacityBytes = 0;
    m.scratchCapacityBytes = m.halfSpectrumScratchCapacityBytes + m.realScratchCapacityBytes;
    m.scratchHighWaterBytes = m.scratchCapacityBytes;
    m.planCount = 17 + (k.scalarPrepared ? 1 : 0);
    auto mean = [&](std::initializer_list<const char*> names) {
        double sec = 0.0;
        for (const char* n : names)
            for (const auto& r : k.records)
                if (r.name == n && r.launches) sec += 1e-3 * r.totalMs / static_cast<double>(r.launches);
        return sec;
    };
    m.phaseSeconds = mean({"fields"});
    m.reconstructionSeconds = mean({"synth", "yfft_inv"});
    m.productSeconds = mean({"xfused"});
    m.projectionSeconds = mean({"yfft_fwd", "proj", "stage"});
    *out = m;
    return WV_SUCCESS;
}

const char* wv_engine_identifier(const wv_kernel* kernel) { return kernel ? "wvamd-hip-gfx950" : nullptr; }
const char* wv_engine_library_identity(const wv_kernel* kernel) { return kernel ? "libwvamd.so (HIP, gfx950)" : nullptr; }
const char* wv_nonlinear_flux_schedule_identifier(const wv_kernel* kernel) {
    return kernel ? "mfma-synthesis/y-fft/fused-x-products/y-fft/mfma-analysis/stage" : nullptr;
}
const char* wv_phase_implementation_identifier(const wv_kernel* kernel) { return kernel ? "device-sincos-fp64" : nullptr; }
const char* wv_coefficient_arithmetic_mode_identifier(const wv_kernel* kernel) {
    return kernel ? "natural-dimensional-prescaled" : nullptr;
}
const char* wv_inverse_normalization_placement_identifier(const wv_kernel* kernel) {
    return kernel ? "coefficient-production" : nullptr;
}
const char* wv_optimization_implementation_identifier(const wv_kernel* kernel) {
    if (!kernel) return nullptr;
    return kernel->impl.precision == WV_PRECISION_FP64   ? "hip-gfx950-fp64"
           : kernel->impl.precision == WV_PRECISION_MIXED ? "hip-gfx950-mixed-fp32"
                                                          : "hip-gfx950-mixed-split-fp16";
}
size_t wv_coefficient_worker_count(const wv_kernel* kernel) { (void)kernel; return 0; }
size_t wv_phase_reservation_bytes(const wv_kernel* kernel) {
    if (!kernel) return 0;
    const auto& c = kernel->impl.d.c;
    return kernel->impl.d.NxHalf * c.Ny * c.Nz * 16;
}

wv_status wv_descriptor_modes(const wv_kernel* kernel, int64_t* kMode, int64_t* lMode, double* k, double* l,
                              double* Kh, double* cosAlpha, double* sinAlpha, int64_t* storageRow,
                              uint8_t* conjugated) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    const auto& md = kernel->impl.d.modes;
    for (size_t i = 0; i < md.size(); ++i) {
        if (kMode) kMode[i] = md[i].kMode;
        if (lMode) lMode[i] = md[i].lMode;
        if (k) k[i] = md[i].k;
        if (l) l[i] = md[i].l;
        if (Kh) Kh[i] = md[i].Kh;
        if (cosAlpha) cosAlpha[i] = md[i].cosAlpha;
        if (sinAlpha) sinAlpha[i] = md[i].sinAlpha;
        if (storageRow) storageRow[i] = static_cast<int64_t>(md[i].storageRow);
        if (conjugated) conjugated[i] = md[i].conjugated ? 1 : 0;
    }
    return WV_SUCCESS;
}

long wv_descriptor_table(const wv_kernel* kernel, const char* name, double* out) {
    if (!kernel || !name) return -1;
    return kernel->impl.d.table(name, out);
}

wv_status wv_nonlinear_flux_device(wv_kernel* kernel, double t, double t0, const double* Ap, const double* Am,
                                   const double* A0, double* Fp, double* Fm, double* F0, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (auto st = validateFlux(k, t, t0, Ap, Am, A0, Fp, Fm, F0); st != WV_SUCCESS) return st;
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        wvamd::Epilogue ep{};
        ep.mode = wvamd::kFluxOut;
        ep.out0 = reinterpret_cast<wvamd::cx<double>*>(Fp);
        ep.out1 = reinterpret_cast<wvamd::cx<double>*>(Fm);
        ep.out2 = reinterpret_cast<wvamd::cx<double>*>(F0);
        k.rhsAny(Ap, Am, A0, t, t0, ep, s);
    });
}

wv_status wv_nonlinear_flux(wv_kernel* kernel, double t, double t0, const double* Ap, const double* Am,
                            const double* A0, double* Fp, double* Fm, double* F0) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (auto st = validateFlux(k, t, t0, Ap, Am, A0, Fp, Fm, F0); st != WV_SUCCESS) return st;
    return guarded([&] {
        hipCheck(hipSetDevice(k.device), "hipSetDevice");
        const size_t bytes = k.d.M * 16;
        wvamd::DeviceBuffer buf;
        buf.alloc(6 * bytes);
        char* b = buf.as<char>();
        hipCheck(hipMemcpy(b, Ap, bytes, hipMemcpyHostToDevice), "copy");
        hipCheck(hipMemcpy(b + bytes, Am, bytes, hipMemcpyHostToDevice), "copy");
        hipCheck(hipMemcpy(b + 2 * bytes, A0, bytes, hipMemcpyHostToDevice), "copy");
        const double* d = reinterpret_cast<const double*>(b);
        double* o = reinterpret_cast<double*>(b + 3 * bytes);
        const size_t n = k.d.M * 2;
        const wv_status st = wv_nonlinear_flux_device(kernel, t, t0, d, d + n, d + 2 * n, o, o + n, o + 2 * n, nullptr);
        if (st != WV_SUCCESS) return st;
        hipCheck(hipMemcpy(Fp, o, bytes, hipMemcpyDeviceToHost), "copy");
        hipCheck(hipMemcpy(Fm, o + n, bytes, hipMemcpyDeviceToHost), "copy");
        hipCheck(hipMemcpy(F0, o + 2 * n, bytes, hipMemcpyDeviceToHost), "copy");
        k.metrics.bytesCopied += 6 * bytes;
        return WV_SUCCESS;
    });
}

wv_status wv_nonlinear_flux_with_advection_fields_device(wv_kernel* kernel, double t, double t0, const double* Ap,
                                                         const double* Am, const double* A0, double* Fp,
                                                         double* Fm, double* F0, double* uvw, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (auto st = validateFlux(k, t, t0, Ap, Am, A0, Fp, Fm, F0); st != WV_SUCCESS) return st;
    if (!uvw) return fail(WV_INVALID_POINTER, "Advection fields has a null data pointer.");
    if (auto st = validateBundleOwnership(k, uvw, 3, Ap, Am, A0, "Advection fields"); st != WV_SUCCESS) return st;
    const size_t cb = k.d.M * 16, fb = 3 * k.d.R * sizeof(double);
    for (const void* o : {static_cast<const void*>(Fp), static_cast<const void*>(Fm), static_cast<const void*>(F0)})
        if (overlaps(uvw, fb, o, cb)) return fail(WV_OVERLAPPING_ARRAYS, "Advection fields must not overlap flux outputs.");
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        if (k.precision == WV_PRECISION_FP64) k.inverse<double>(Ap, Am, A0, t, t0, 3, uvw, s);
        else k.inverse<float>(Ap, Am, A0, t, t0, 3, uvw, s);
        wvamd::Epilogue ep{};
        ep.mode = wvamd::kFluxOut;
        ep.out0 = reinterpret_cast<wvamd::cx<double>*>(Fp);
        ep.out1 = reinterpret_cast<wvamd::cx<double>*>(Fm);
        ep.out2 = reinterpret_cast<wvamd::cx<double>*>(F0);
        k.rhsAny(Ap, Am, A0, t, t0, ep, s);
    });
}

wv_status wv_nonlinear_flux_using_advection_fields_device(wv_kernel* kernel, double t, double t0, const double* Ap,
                                                          const double* Am, const double* A0, const double* uvw,
                                                          double* Fp, double* Fm, double* F0, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (auto st = validateFlux(k, t, t0, Ap, Am, A0, Fp, Fm, F0); st != WV_SUCCESS) return st;
    if (!uvw) return fail(WV_INVALID_POINTER, "Prepared advection fields has a null data pointer.");
    if (auto st = validateBundleOwnership(k, uvw, 3, Ap, Am, A0, "Prepared advection fields"); st != WV_SUCCESS)
        return st;
    const size_t cb = k.d.M * 16, fb = 3 * k.d.R * sizeof(double);
    for (const void* o : {static_cast<const void*>(Fp), static_cast<const void*>(Fm), static_cast<const void*>(F0)})
        if (overlaps(uvw, fb, o, cb))
            return fail(WV_OVERLAPPING_ARRAYS, "Prepared advection fields must not overlap flux outputs.");
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        wvamd::Epilogue ep{};
        ep.mode = wvamd::kFluxOut;
        ep.out0 = reinterpret_cast<wvamd::cx<double>*>(Fp);
        ep.out1 = reinterpret_cast<wvamd::cx<double>*>(Fm);
        ep.out2 = reinterpret_cast<wvamd::cx<double>*>(F0);
        k.rhsAny(Ap, Am, A0, t, t0, ep, s, false, uvw);
    });
}

wv_status wv_prepare_scalar_advection(wv_kernel* kernel) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    return execute(k, nullptr, kNotReentrant, [&](hipStream_t) { k.ensureScalarScratch(); });
}

wv_status wv_advect_f_grid_scalar_device(wv_kernel* kernel, const double* scalar, const double* advectionFields,
                                         int shouldAntialias, double* rightHandSide, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (!scalar || !rightHandSide)
        return fail(WV_INVALID_POINTER, "F-grid scalar input and output pointers must be nonnull.");
    if (!advectionFields) return fail(WV_INVALID_POINTER, "Advection fields has a null data pointer.");
    const size_t bytes = k.d.R * sizeof(double);
    if (overlaps(scalar, bytes, rightHandSide, bytes) || overlaps(advectionFields, 3 * bytes, scalar, bytes) ||
        overlaps(advectionFields, 3 * bytes, rightHandSide, bytes))
        return fail(WV_OVERLAPPING_ARRAYS, "Tracer scalar, velocity, and RHS arrays must not overlap.");
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        k.advectScalar(scalar, advectionFields, shouldAntialias != 0, rightHandSide, s);
    });
}

static wv_status inverseCommon(wv_kernel* kernel, double t, double t0, const double* Ap, const double* Am,
                               const double* A0, double* fields, int nf, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (!Ap || !Am || !A0 || !fields) return fail(WV_INVALID_POINTER, "Inverse transform has a null data pointer.");
    if (!std::isfinite(t) || !std::isfinite(t0)) return fail(WV_INVALID_CONFIGURATION, "State times must be finite.");
    if (auto st = validateBundleOwnership(k, fields, nf, Ap, Am, A0, "Inverse spatial output"); st != WV_SUCCESS)
        return st;
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        if (k.precision == WV_PRECISION_FP64) k.inverse<double>(Ap, Am, A0, t, t0, nf, fields, s);
        else k.inverse<float>(Ap, Am, A0, t, t0, nf, fields, s);
    });
}

wv_status wv_transform_wave_vortex_to_uvweta_device(wv_kernel* kernel, double t, double t0, const double* Ap,
                                                    const double* Am, const double* A0, double* fields,
                                                    void* stream) {
    return inverseCommon(kernel, t, t0, Ap, Am, A0, fields, 4, stream);
}

wv_status wv_transform_wave_vortex_to_uvw_device(wv_kernel* kernel, double t, double t0, const double* Ap,
                                                 const double* Am, const double* A0, double* fields, void* stream) {
    return inverseCommon(kernel, t, t0, Ap, Am, A0, fields, 3, stream);
}

wv_status wv_transform_state_field_derivatives_device(wv_kernel* kernel, double t, double t0, const double* Ap,
                                                      const double* Am, const double* A0, int field,
                                                      double* derivatives, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (field < 0 || field > 3) return fail(WV_INVALID_CONFIGURATION, "Unknown dynamical field derivative target.");
    if (!derivatives) return fail(WV_INVALID_POINTER, "Field derivatives has a null data pointer.");
    if (!Ap) return fail(WV_INVALID_POINTER, "Ap has a null data pointer.");
    if (!Am) return fail(WV_INVALID_POINTER, "Am has a null data pointer.");
    if (!A0) return fail(WV_INVALID_POINTER, "A0 has a null data pointer.");
    if (!std::isfinite(t) || !std::isfinite(t0)) return fail(WV_INVALID_CONFIGURATION, "State times must be finite.");
    if (auto st = validateBundleOwnership(k, derivatives, 3, Ap, Am, A0, "Field derivatives"); st != WV_SUCCESS)
        return st;
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        if (k.precision == WV_PRECISION_FP64) k.stateDerivatives<double>(Ap, Am, A0, t, t0, field, derivatives, s);
        else k.stateDerivatives<float>(Ap, Am, A0, t, t0, field, derivatives, s);
    });
}

wv_status wv_apv_device(wv_kernel* kernel, double t, double t0, const double* Ap, const double* Am,
                        const double* A0, double* apv, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (k.d.stratified)
        return fail(WV_UNSUPPORTED_OPERATION,
                    "APV needs eta_true, which equals eta only for the constant-N background profile.");
    if (!apv) return fail(WV_INVALID_POINTER, "APV has a null data pointer.");
    if (!Ap) return fail(WV_INVALID_POINTER, "Ap has a null data pointer.");
    if (!Am) return fail(WV_INVALID_POINTER, "Am has a null data pointer.");
    if (!A0) return fail(WV_INVALID_POINTER, "A0 has a null data pointer.");
    if (!std::isfinite(t) || !std::isfinite(t0)) return fail(WV_INVALID_CONFIGURATION, "State times must be finite.");
    if (auto st = validateBundleOwnership(k, apv, 1, Ap, Am, A0, "APV"); st != WV_SUCCESS) return st;
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        if (k.precision == WV_PRECISION_FP64) k.apv<double>(Ap, Am, A0, t, t0, apv, s);
        else k.apv<float>(Ap, Am, A0, t, t0, apv, s);
    });
}

static wv_status forwardCommon(wv_kernel* kernel, const double* fields, double t, double t0, double* Ap,
                               double* Am, double* A0, int nf, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    const bool hydro = k.d.c.isHydrostatic != 0;
    if (nf == 4 && hydro) return fail(WV_INVALID_CONFIGURATION, "transformUVWEtaToWaveVortex requires a nonhydrostatic kernel.");
    if (nf == 3 && !hydro) return fail(WV_INVALID_CONFIGURATION, "transformUVEtaToWaveVortex requires a hydrostatic kernel.");
    if (!fields) return fail(WV_INVALID_POINTER, "Fields has a null data pointer.");
    if (!Ap || !Am || !A0) return fail(WV_INVALID_POINTER, "Coefficient output has a null data pointer.");
    if (!std::isfinite(t) || !std::isfinite(t0)) return fail(WV_INVALID_CONFIGURATION, "State times must be finite.");
    const size_t cb = k.d.M * 16, fb = k.d.R * nf * sizeof(double);
    void* outs[3] = {Ap, Am, A0};
    for (int i = 0; i < 3; ++i) {
        if (overlaps(outs[i], cb, fields, fb)) return fail(WV_OVERLAPPING_ARRAYS, "Forward coefficient outputs must not overlap the spatial input.");
        for (int j = i + 1; j < 3; ++j)
            if (overlaps(outs[i], cb, outs[j], cb)) return fail(WV_OVERLAPPING_ARRAYS, "Forward coefficient outputs must not overlap each other.");
    }
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        if (k.precision == WV_PRECISION_FP64) k.forward<double>(fields, t, t0, nf, Ap, Am, A0, s);
        else k.forward<float>(fields, t, t0, nf, Ap, Am, A0, s);
    });
}

wv_status wv_transform_uvweta_to_wave_vortex_device(wv_kernel* kernel, const double* fields, double t, double t0,
                                                    double* Ap, double* Am, double* A0, void* stream) {
    return forwardCommon(kernel, fields, t, t0, Ap, Am, A0, 4, stream);
}

wv_status wv_transform_uveta_to_wave_vortex_device(wv_kernel* kernel, const double* fields, double t, double t0,
                                                   double* Ap, double* Am, double* A0, void* stream) {
    return forwardCommon(kernel, fields, t, t0, Ap, Am, A0, 3, stream);
}

static wv_status fgCommon(wv_kernel* kernel, const double* Apm, const double* A0, double* fields, bool fFamily,
                          void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (!Apm) return fail(WV_INVALID_POINTER, "Apm has a null data pointer.");
    if (!A0) return fail(WV_INVALID_POINTER, "A0 has a null data pointer.");
    if (!fields) return fail(WV_INVALID_POINTER, "Field derivatives has a null data pointer.");
    const size_t cb = k.d.M * 16, fb = k.d.R * 4 * sizeof(double);
    if (overlaps(fields, fb, Apm, cb) || overlaps(fields, fb, A0, cb))
        return fail(WV_OVERLAPPING_ARRAYS, "Derivative output must not overlap coefficient inputs.");
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        if (k.precision == WV_PRECISION_FP64) k.fgAll<double>(Apm, A0, fFamily, fields, s);
        else k.fgAll<float>(Apm, A0, fFamily, fields, s);
    });
}

wv_status wv_transform_to_spatial_domain_with_f_all_derivatives_device(wv_kernel* kernel, const double* Apm,
                                                                       const double* A0, double* fields,
                                                                       void* stream) {
    return fgCommon(kernel, Apm, A0, fields, true, stream);
}

wv_status wv_transform_to_spatial_domain_with_g_all_derivatives_device(wv_kernel* kernel, const double* Apm,
                                                                       const double* A0, double* fields,
                                                                       void* stream) {
    return fgCommon(kernel, Apm, A0, fields, false, stream);
}

namespace {
// State checks shared by the RK4 entry points (WVFixedStepRK4::step,
// WVRungeKutta.cpp:335-344): non-null, pairwise non-overlapping arrays.
wv_status validateRk4State(const KernelImpl& k, const double* Ap, const double* Am, const double* A0, const double* t) {
    if (!Ap || !Am || !A0 || !t) return fail(WV_INVALID_POINTER, "RK4 state has a null data pointer.");
    if (overlapping3(Ap, Am, A0, k.d.M * 16)) return fail(WV_OVERLAPPING_ARRAYS, "RK4 state arrays must not overlap.");
    if (!std::isfinite(*t)) return fail(WV_INVALID_CONFIGURATION, "State times must be finite.");
    return WV_SUCCESS;
}
}  // namespace

wv_status wv_rk4_advance_device(wv_kernel* kernel, double* Ap, double* Am, double* A0, double* t, double t0,
                                double h, int nsteps, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (auto st = validateRk4State(k, Ap, Am, A0, t); st != WV_SUCCESS) return st;
    if (!std::isfinite(t0)) return fail(WV_INVALID_CONFIGURATION, "State times must be finite.");
    if (!(h > 0.0) || !std::isfinite(h)) return fail(WV_INVALID_CONFIGURATION, "RK4 step size must be finite and positive.");
    if (nsteps < 0) return fail(WV_INVALID_CONFIGURATION, "RK4 step count must be nonnegative.");
    return execute(k, stream, "RK4 stepping is not reentrant.", [&](hipStream_t s) {
        using wvamd::cx;
        cx<double>* const y[3] = {reinterpret_cast<cx<double>*>(Ap), reinterpret_cast<cx<double>*>(Am),
                                  reinterpret_cast<cx<double>*>(A0)};
        k.rk4Advance(y, t, t0, h, nsteps, s);
    });
}

wv_status wv_rk4_advance_to_time_device(wv_kernel* kernel, double* Ap, double* Am, double* A0, double* t, double t0,
                                        double finalTime, double h, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    KernelImpl& k = kernel->impl;
    if (auto st = validateRk4State(k, Ap, Am, A0, t); st != WV_SUCCESS) return st;
    if (!std::isfinite(t0)) return fail(WV_INVALID_CONFIGURATION, "State times must be finite.");
    if (finalTime < *t || !std::isfinite(finalTime))
        return fail(WV_INVALID_CONFIGURATION, "RK4 cannot advance backward or to a nonfinite time.");
    if (!(h > 0.0) || !std::isfinite(h)) return fail(WV_INVALID_CONFIGURATION, "RK4 step size must be finite and positive.");
    // the step sizes advanceToTime takes, with its time arithmetic
    std::vector<double> steps;
    for (double tt = *t; tt < finalTime;) {
        const double stepSize = std::min(h, finalTime - tt);
        if (!(stepSize > 0.0)) break;
        steps.push_back(stepSize);
        tt = tt + stepSize;
    }
    const wv_status st = execute(k, stream, "RK4 stepping is not reentrant.", [&](hipStream_t s) {
        using wvamd::cx;
        cx<double>* const y[3] = {reinterpret_cast<cx<double>*>(Ap), reinterpret_cast<cx<double>*>(Am),
                                  reinterpret_cast<cx<double>*>(A0)};
        k.rk4AdvanceSteps(y, t, t0, steps, s);
    });
    if (st == WV_SUCCESS) *t = finalTime;
    return st;
}

wv_status wv_rk4_counters(const wv_kernel* kernel, uint64_t* stepCount, double* lastStepSize) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    if (stepCount) *stepCount = kernel->impl.rkStepCount;
    if (lastStepSize) *lastStepSize = kernel->impl.lastStepSize;
    return WV_SUCCESS;
}

namespace {
// The three coefficient arrays of a state or cotangent must be distinct.
bool distinct3(const void* a, const void* b, const void* c) { return a != b && a != c && b != c; }
}  // namespace

wv_status wv_nonlinear_flux_vjp_device(wv_kernel* kernel, double t, double t0, const double* Ap, const double* Am,
                                       const double* A0, const double* Lp, const double* Lm, const double* L0,
                                       double* Gp, double* Gm, double* G0, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    if (!Ap || !Am || !A0 || !Lp || !Lm || !L0 || !Gp || !Gm || !G0)
        return fail(WV_INVALID_POINTER, "Flux adjoint has a null data pointer.");
    if (!std::isfinite(t) || !std::isfinite(t0)) return fail(WV_INVALID_CONFIGURATION, "Times must be finite.");
    KernelImpl& k = kernel->impl;
    const size_t cb = k.d.M * 16;
    for (const double* o : {Gp, Gm, G0})
        for (const double* i : {Ap, Am, A0, Lp, Lm, L0})
            if (overlaps(o, cb, i, cb)) return fail(WV_OVERLAPPING_ARRAYS, "Adjoint outputs may not alias the inputs.");
    if (!distinct3(Gp, Gm, G0) || overlapping3(Gp, Gm, G0, cb))
        return fail(WV_OVERLAPPING_ARRAYS, "Adjoint outputs must be distinct.");
    return execute(k, stream, kNotReentrant,
                   [&](hipStream_t s) { k.fluxVjpAny(Ap, Am, A0, t, t0, Lp, Lm, L0, Gp, Gm, G0, s); });
}

wv_status wv_transform_wave_vortex_to_uvweta_vjp_device(wv_kernel* kernel, const double* fieldsBar, double t,
                                                        double t0, double* Gp, double* Gm, double* G0, void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    if (!fieldsBar || !Gp || !Gm || !G0) return fail(WV_INVALID_POINTER, "Transform adjoint has a null data pointer.");
    if (!std::isfinite(t) || !std::isfinite(t0)) return fail(WV_INVALID_CONFIGURATION, "Times must be finite.");
    KernelImpl& k = kernel->impl;
    const size_t cb = k.d.M * 16;
    if (!distinct3(Gp, Gm, G0) || overlapping3(Gp, Gm, G0, cb))
        return fail(WV_OVERLAPPING_ARRAYS, "Adjoint outputs must be distinct.");
    if (auto st = validateBundleOwnership(k, fieldsBar, 4, Gp, Gm, G0, "Field cotangents"); st != WV_SUCCESS) return st;
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        k.ensureAdjoint();
        if (k.precision == WV_PRECISION_FP64) k.inverseVjp<double>(fieldsBar, t, t0, Gp, Gm, G0, s);
        else k.inverseVjp<float>(fieldsBar, t, t0, Gp, Gm, G0, s);
    });
}

namespace {
// Copies a state or cotangent given as three arrays into one packed triple.
void packTriple(const double* a, const double* b, const double* c, wvamd::cx<double>* dst, size_t M, hipStream_t s) {
    const size_t bytes = M * sizeof(wvamd::cx<double>);
    hipCheck(hipMemcpyAsync(dst, a, bytes, hipMemcpyDeviceToDevice, s), "pack");
    hipCheck(hipMemcpyAsync(dst + M, b, bytes, hipMemcpyDeviceToDevice, s), "pack");
    hipCheck(hipMemcpyAsync(dst + 2 * M, c, bytes, hipMemcpyDeviceToDevice, s), "pack");
}
void unpackTriple(const wvamd::cx<double>* src, double* a, double* b, double* c, size_t M, hipStream_t s) {
    const size_t bytes = M * sizeof(wvamd::cx<double>);
    hipCheck(hipMemcpyAsync(a, src, bytes, hipMemcpyDeviceToDevice, s), "unpack");
    hipCheck(hipMemcpyAsync(b, src + M, bytes, hipMemcpyDeviceToDevice, s), "unpack");
    hipCheck(hipMemcpyAsync(c, src + 2 * M, bytes, hipMemcpyDeviceToDevice, s), "unpack");
}
}  // namespace

wv_status wv_rk4_step_vjp_device(wv_kernel* kernel, const double* Ap, const double* Am, const double* A0, double t,
                                 double t0, double h, double* Lp, double* Lm, double* L0, void* stream) {
    return wv_rk4_adjoint_device(kernel, Ap, Am, A0, t, t0, h, 1, 1, Lp, Lm, L0, stream);
}

wv_status wv_rk4_adjoint_device(wv_kernel* kernel, const double* Ap, const double* Am, const double* A0, double t,
                                double t0, double h, int nsteps, int interval, double* Lp, double* Lm, double* L0,
                                void* stream) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    if (!Ap || !Am || !A0 || !Lp || !Lm || !L0) return fail(WV_INVALID_POINTER, "RK4 adjoint has a null data pointer.");
    if (!(h > 0.0) || !std::isfinite(h) || !std::isfinite(t) || !std::isfinite(t0))
        return fail(WV_INVALID_CONFIGURATION, "RK4 step size must be finite and positive, times finite.");
    if (nsteps < 0 || interval < 1) return fail(WV_INVALID_CONFIGURATION, "Step count must be nonnegative, interval positive.");
    KernelImpl& k = kernel->impl;
    const size_t cb = k.d.M * 16;
    if (!distinct3(Lp, Lm, L0) || overlapping3(Lp, Lm, L0, cb))
        return fail(WV_OVERLAPPING_ARRAYS, "Cotangent arrays must be distinct.");
    for (const double* o : {Lp, Lm, L0})
        for (const double* i : {Ap, Am, A0})
            if (overlaps(o, cb, i, cb)) return fail(WV_OVERLAPPING_ARRAYS, "Cotangents may not alias the state.");
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) {
        using wvamd::cx;
        k.ensureAdjoint();
        const size_t M = k.d.M;
        wvamd::DeviceBuffer y0, ybar;
        y0.alloc(3 * M * sizeof(cx<double>));
        ybar.alloc(3 * M * sizeof(cx<double>));
        packTriple(Ap, Am, A0, y0.as<cx<double>>(), M, s);
        packTriple(Lp, Lm, L0, ybar.as<cx<double>>(), M, s);
        k.rk4Adjoint(y0.as<cx<double>>(), t, t0, h, nsteps, interval, ybar.as<cx<double>>(), s);
        unpackTriple(ybar.as<cx<double>>(), Lp, Lm, L0, M, s);
        hipCheck(hipStreamSynchronize(s), "adjoint");
    });
}

wv_status wv_total_energy_device(wv_kernel* kernel, const double* Ap, const double* Am, const double* A0,
                                 double* energy, void* stream) {
    if (!kernel || !energy) return fail(WV_INVALID_POINTER, "Total energy requires a kernel and an output.");
    if (!Ap || !Am || !A0) return fail(WV_INVALID_POINTER, "Total energy state has a null data pointer.");
    KernelImpl& k = kernel->impl;
    return execute(k, stream, kNotReentrant, [&](hipStream_t s) { *energy = k.energy(Ap, Am, A0, s); });
}

wv_status wv_total_energy_no_alloc_device(wv_kernel* kernel, const double* Ap, const double* Am, const double* A0,
                                          double* energy, void* stream) {
    if (!kernel || !energy) return fail(WV_INVALID_POINTER, "Total energy requires a kernel and a device output.");
    if (!Ap || !Am || !A0) return fail(WV_INVALID_POINTER, "Total energy state has a null data pointer.");
    const KernelImpl& k = kernel->impl;
    // k_energy_partial strides an int index over M coefficients
    if (k.d.M > static_cast<size_t>(INT_MAX) - static_cast<size_t>(256 * wvamd::kEnergyPartials))
        return fail(WV_SIZE_OVERFLOW, "Coefficient count exceeds the allocation-free energy reduction's maximum.");
    if (k.energyBlocks() > wvamd::kEnergyPartials)
        return fail(WV_SIZE_OVERFLOW, "Energy reduction needs more partial sums than the static scratch holds (1024).");
    return guarded([&] {
        double* partials = scratchOf(k.device);
        if (!partials) return fail(WV_INVALID_CONFIGURATION, "No scratch address for this kernel's device.");
        hipCheck(hipSetDevice(k.device), "hipSetDevice");
        k.energyInto(Ap, Am, A0, partials, energy, S(stream));
        return WV_SUCCESS;
    });
}

wv_status wv_set_stage_instrumentation(wv_kernel* kernel, int enabled) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    kernel->impl.instrument = enabled != 0;
    return WV_SUCCESS;
}

int wv_stage_timings(wv_kernel* kernel, wv_stage_timing* out, int capacity) {
    if (!kernel) return -1;
    try {
        kernel->impl.drainTimings();
    } catch (const Failure& e) {
        fail(e.code, e.message);
        return -1;
    }
    const auto& r = kernel->impl.records;
    for (int i = 0; i < capacity && i < static_cast<int>(r.size()); ++i) {
        std::snprintf(out[i].name, sizeof(out[i].name), "%s", r[static_cast<size_t>(i)].name.c_str());
        out[i].total_ms = r[static_cast<size_t>(i)].totalMs;
        out[i].launches = r[static_cast<size_t>(i)].launches;
        out[i].bytes_per_launch = r[static_cast<size_t>(i)].launches
                                      ? r[static_cast<size_t>(i)].totalBytes / static_cast<double>(r[static_cast<size_t>(i)].launches)
                                      : 0.0;
    }
    return static_cast<int>(r.size());
}

void wv_reset_stage_timings(wv_kernel* kernel) {
    if (!kernel) return;
    try {
        kernel->impl.drainTimings();
    } catch (...) {
    }
    for (auto& r : kernel->impl.records) {
        r.totalMs = 0.0;
        r.launches = 0;
        r.totalBytes = 0.0;
    }
}

wv_status wv_debug_inject_launch_failure(wv_kernel* kernel, long n) {
    if (!kernel) return fail(WV_INVALID_POINTER, "null kernel");
    if (n < 0) return fail(WV_INVALID_CONFIGURATION, "Launch count must be nonnegative.");
    kernel->impl.failAfterLaunches = n;
    return WV_SUCCESS;
}

}  // extern "C"
