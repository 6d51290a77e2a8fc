// Device data and kernels of the wave-vortex flux (the templates and their
// device helpers), shared by the two translation units that instantiate
// them: wv_kernels.hip (the host code, the kernels compiled per
// power-of-two length) and wv_kernels_rt.hip (the runtime-length FFT
// kernels, compiled in parallel with it).  Non-template kernels have
// internal linkage (maybe-unused: the runtime-length unit launches none of
// them): each unit that launches one holds its own copy.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <type_traits>

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
        stockhamPass<T, 16, +1, BatchRows<PAIRS>, 0x3C0u, false>(buf, 8, 0, bt, tw);
        stockhamPass<T, 16, +1, BatchRows<PAIRS>, 0, false>(buf, 8, 4, bt, tw);
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
[[maybe_unused]] static __global__ __launch_bounds__(1024) void k_scalarDeriv(cx<double>* __restrict__ C, int Nx, int Ny, int Nz, int nxh,
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
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_energy_partial(const cx<double>* __restrict__ Ap,
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

[[maybe_unused]] static __global__ __launch_bounds__(256) void k_energy_final(const double* __restrict__ partial, int n,
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

#include "wv_adjoint_kernels.hpp"

}  // namespace wvamd
