// Adjoint (vector-Jacobian) kernels of the nonlinear flux.  Included by
// wv_kernels_device.hpp after the forward kernels, whose data structures they
// use.
//
// The flux is F(y) = P(products(S y, S y)) with S the synthesis (coefficients
// -> spatial u, v, w, eta and derivatives, linear) and P the projection
// (spatial flux of each target -> coefficients, linear).  Its transpose
// Jacobian applied to a cotangent lambda is
//   y* = S^T( products'(S y)^T (P^T lambda) ),
// evaluated here stage by stage in reverse order.  Inner products are the
// real ones: <a, b> = Re sum conj(a) b for complex arrays, sum a b for real
// fields, so that <J dy, lambda> = <dy, J^T lambda>.  The transpose of a
// multiplication by a complex constant c is a multiplication by conj(c), of a
// conjugation a conjugation, of taking the real part taking the real part.

// P^T, first half: the per-coefficient projection formulas of k_proj
// (projectFluxTargetMode) transposed.  For every coefficient (j, mode) and
// target T it writes hscale * value_T*, the cotangent of the target's
// vertically analysed spectrum at the mode's storage column (the column
// conjugation is applied by the synthesis that follows).
template <typename T, int NT>
__global__ __launch_bounds__(256) void k_projadj(const cx<double>* __restrict__ Lp, const cx<double>* __restrict__ Lm,
                                                 const cx<double>* __restrict__ L0,
                                                 const ColumnDev* __restrict__ cols,
                                                 const int* __restrict__ wvColumn, int Nj, int M, int targetCode,
                                                 int hydro, ProjTables tb, double elapsed, double hscale,
                                                 cx<T>* __restrict__ CV) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M) return;
    const int j = idx % Nj, wv = idx / Nj;
    const ColumnDev cd = cols[wvColumn[wv]];
    double sn, cs;
    sincos(tb.omega[idx] * elapsed, &sn, &cs);
    const cx<double> p{cs, sn};
    const cx<double> cApBar = cmul(Lp[idx], p);         // Fp = sum cAp conj(p)
    const cx<double> cAmBar = cmul(Lm[idx], cconj(p));  // Fm = sum cAm p
    const cx<double> a0Bar = L0[idx];                   // F0 = sum a0
    const double a0v = tb.A0Vort[idx], a0b = tb.A0Buoy[idx], na0 = tb.NA0[idx], apmn = tb.ApmN[idx];
    const bool inertial = wv == 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int target = (targetCode >> (4 * t)) & 0xF;
        // forward: a0 = alpha value, cAp = (beta + gamma) value, cAm = (beta - gamma) value,
        // gamma = apmn (delta_T3 - na0 alpha); inertial: cAp = iota value, cAm = conj(iota value)
        cx<double> alpha{0.0, 0.0};
        if (target == 0) alpha = {0.0, -cd.l * a0v};
        if (target == 1) alpha = {0.0, cd.k * a0v};
        if (target == 3) alpha = {a0b, 0.0};
        const cx<double> gamma = cscale(csub(target == 3 ? cx<double>{1.0, 0.0} : cx<double>{0.0, 0.0},
                                             cscale(alpha, na0)),
                                        apmn);
        cx<double> beta{0.0, 0.0};
        if (hydro) {
            if (target == 0) beta = cmul(cx<double>{0.0, cd.k}, tb.ApmD[idx]);
            if (target == 1) beta = cmul(cx<double>{0.0, cd.l}, tb.ApmD[idx]);
        } else {
            if (target == 0) beta = {cd.cosA * tb.ApmDScaled[idx], 0.0};
            if (target == 1) beta = {cd.sinA * tb.ApmDScaled[idx], 0.0};
            if (target == 2) beta = {0.0, (cd.Kh / 2.0) * tb.apmW[j]};
        }
        cx<double> v = cmul(cconj(alpha), a0Bar);
        if (inertial) {
            cx<double> iota{0.0, 0.0};
            if (target == 0) iota = {tb.inertialScale[j], 0.0};
            if (target == 1) iota = {0.0, -tb.inertialScale[j]};
            v = cadd(v, cmul(cconj(iota), cApBar));
            v = cadd(v, cconj(cmul(iota, cAmBar)));
        } else {
            v = cadd(v, cmul(cconj(cadd(beta, gamma)), cApBar));
            v = cadd(v, cmul(cconj(csub(beta, gamma)), cAmBar));
        }
        v = cscale(v, hscale);
        CV[static_cast<size_t>(t) * M + idx] = {T(v.re), T(v.im)};
    }
}

// Transpose of the x part of the flux for one (y, z) row:
//   forward fields   c2r of the spectra G (as k_xfused)
//   mu_T             transpose of the r2c of the products: the cotangents P*
//                    of the targets (kx < nkx), mu(x) = Re sum_k P*_k e^{+ikx}
//   product adjoint  U* = -sum_T mu_T dxT, V* = -sum mu_T dyT, W* = -sum mu_T dzT,
//                    dxT* = -mu_T U, dyT* = -mu_T V, dzT* = -mu_T W
//   c2r transpose    twice the r2c for kx > 0, the real part at kx = 0, and
//                    -i k for the channels formed as i k times a spectrum,
//                    summed into the cotangent Gs of each spectrum channel.
// One row per workgroup of 256 threads.
template <int NT>
struct XAdjMap;
template <>
struct XAdjMap<4> {  // spectrum channel <- spatial channel (plain), spatial channel (times i k)
    static constexpr int nh = 12;
    static constexpr int plain[12] = {0, 1, 12, -1, 3, 5, 7, 9, 10, 11, 13, 14};
    static constexpr int ik[12] = {2, 4, 6, 8, -1, -1, -1, -1, -1, -1, -1, -1};
};
template <>
struct XAdjMap<3> {
    static constexpr int nh = 10;
    static constexpr int plain[10] = {0, 1, 10, -1, 3, 5, 7, 8, 9, 12};
    static constexpr int ik[10] = {2, 4, -1, 6, -1, -1, -1, -1, -1, -1};
};

template <typename T, int NT>
__global__ __launch_bounds__(256) void k_xadj(const cx<T>* __restrict__ G, const cx<T>* __restrict__ Ps,
                                              cx<T>* __restrict__ Gs, int Nx, int log2Nx, int Ny, int Nz, int nkx,
                                              int nkxP, const cx<T>* __restrict__ twX, const T* __restrict__ kxv) {
    using XM = XMap<NT>;
    using XA = XAdjMap<NT>;
    constexpr int PAIRS = XM::pairs;
    constexpr int TP = (NT + 1) / 2;  // target pairs
    extern __shared__ __align__(16) unsigned char smem[];
    const int NxP = xfusedStride<T>(Nx);
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem);
    cx<T>* buf = tw + Nx;            // [PAIRS][NxP] fields, then their cotangents
    cx<T>* mub = buf + PAIRS * NxP;  // [TP][NxP] mu of the targets
    const int row = blockIdx.x;      // z*Ny + y
    const size_t chStride = static_cast<size_t>(Nz) * Ny * nkxP;
    const size_t rowOff = static_cast<size_t>(row) * nkxP;
    loadTwiddles(tw, twX, Nx);
    const int lane = threadIdx.x & 63, wave = waveIndex();
    // spectra: the forward pairs (as k_xfused) and the target cotangents
    // (halved for kx > 0, real part at kx = 0), Hermitian-extended
    for (int t = wave; t < PAIRS + TP; t += 4) {
        cx<T>* zb = t < PAIRS ? buf + t * NxP : mub + (t - PAIRS) * NxP;
        for (int k = lane; k < nkx; k += 64) {
            cx<T> a, b;
            if (t < PAIRS) {
                const int sa = XM::srcA[t], sb = XM::srcB[t];
                a = G[sa * chStride + rowOff + k];
                b = sb >= 0 ? G[sb * chStride + rowOff + k] : cx<T>{T(0), T(0)};
                const T kk = kxv[k];
                if (XM::ikA[t]) a = cmuli(a, kk);
                if (k == 0) a.im = b.im = T(0);
            } else {
                const int ta = 2 * (t - PAIRS), tb = ta + 1;
                a = Ps[ta * chStride + rowOff + k];
                b = tb < NT ? Ps[tb * chStride + rowOff + k] : cx<T>{T(0), T(0)};
                if (k == 0) {
                    a.im = b.im = T(0);
                } else {
                    a = cscale(a, T(0.5));
                    b = cscale(b, T(0.5));
                }
            }
            zb[ldsIndex(k, 1)] = {a.re - b.im, a.im + b.re};
            if (k > 0) zb[ldsIndex(Nx - k, 1)] = {a.re + b.im, b.re - a.im};
        }
        for (int k = nkx + lane; k <= Nx - nkx; k += 64) zb[ldsIndex(k, 1)] = cx<T>{T(0), T(0)};
    }
    __syncthreads();
    fftLdsBatched<T, +1>(buf, Nx, log2Nx, PAIRS, NxP, tw);
    fftLdsBatched<T, +1>(mub, Nx, log2Nx, TP, NxP, tw);
    // products transposed, in place per x
    for (int x = threadIdx.x; x < Nx; x += blockDim.x) {
        const int xl = ldsIndex(x, 1);
        T ch[2 * PAIRS], mu[2 * TP];
#pragma unroll
        for (int p = 0; p < PAIRS; ++p) {
            const cx<T> v = buf[p * NxP + xl];
            ch[2 * p] = v.re;
            ch[2 * p + 1] = v.im;
        }
#pragma unroll
        for (int p = 0; p < TP; ++p) {
            const cx<T> v = mub[p * NxP + xl];
            mu[2 * p] = v.re;
            mu[2 * p + 1] = v.im;
        }
        const T U = ch[XM::uvw[0]], V = ch[XM::uvw[1]], W = ch[XM::uvw[2]];
        T out[2 * PAIRS];
#pragma unroll
        for (int c = 0; c < 2 * PAIRS; ++c) out[c] = T(0);
        T Ub = T(0), Vb = T(0), Wb = T(0);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            Ub -= mu[t] * ch[XM::prod[t][0]];
            Vb -= mu[t] * ch[XM::prod[t][1]];
            Wb -= mu[t] * ch[XM::prod[t][2]];
            out[XM::prod[t][0]] = -mu[t] * U;
            out[XM::prod[t][1]] = -mu[t] * V;
            out[XM::prod[t][2]] = -mu[t] * W;
        }
        out[XM::uvw[0]] = Ub;
        out[XM::uvw[1]] = Vb;
        out[XM::uvw[2]] = Wb;
#pragma unroll
        for (int p = 0; p < PAIRS; ++p) buf[p * NxP + xl] = {out[2 * p], out[2 * p + 1]};
    }
    __syncthreads();
    fftLdsBatched<T, -1>(buf, Nx, log2Nx, PAIRS, NxP, tw);
    // unpack the real transforms, weight, fold into the spectrum channels
    for (int c = wave; c < XA::nh; c += 4) {
        for (int k = lane; k < nkx; k += 64) {
            cx<T> acc{T(0), T(0)};
            const int sp = XA::plain[c], si = XA::ik[c];
#pragma unroll
            for (int which = 0; which < 2; ++which) {
                const int s = which ? si : sp;
                if (s < 0) continue;
                const cx<T>* zb = buf + (s >> 1) * NxP;
                const cx<T> Za = zb[ldsIndex(k, 1)];
                const cx<T> Zb = zb[ldsIndex(k == 0 ? 0 : Nx - k, 1)];
                const T h = T(0.5);
                cx<T> X = (s & 1) ? cx<T>{(Za.im + Zb.im) * h, (Zb.re - Za.re) * h}
                                  : cx<T>{(Za.re + Zb.re) * h, (Za.im - Zb.im) * h};
                X = k == 0 ? cx<T>{X.re, T(0)} : cscale(X, T(2));
                if (which) X = cmuli(X, -kxv[k]);  // transpose of i k
                acc = cadd(acc, X);
            }
            Gs[c * chStride + rowOff + k] = acc;
        }
    }
}

// S^T for the vertical part: for one group of 8 columns, every synthesis
// channel's cotangent acc*(z) = H*[outV] - i k H*[outX] - i l H*[outY] is
// contracted over z with the transposed synthesis matrix (MFMA, one wave per
// 16 rows of j), the channel coefficient (valueScale or -/+ m_j/2) and the
// column's conjugation are transposed, and the result is summed per field
// into colbuf[f][column][j].
template <typename T>
__global__ __launch_bounds__(1024) void k_synthadj(const cx<T>* __restrict__ Hs, const ColumnDev* __restrict__ cols,
                                                   int ncol, SynthMats<T> matsT, int Ja, int Za, int Nz, int Nj,
                                                   int Ny, int nkxP, const T* __restrict__ mj, SynthPlan plan,
                                                   cx<T>* __restrict__ colbuf) {
    using MF = Mfma<T>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* qs = reinterpret_cast<cx<T>*>(smem);  // [4][Za][8]
    __shared__ ColumnDev scol[8];
    const int lane = threadIdx.x & 63;
    const int jt = waveIndex();
    const int cg = blockIdx.x;
    if (threadIdx.x < 8) {
        const int col = cg * 8 + threadIdx.x;
        if (col < ncol) scol[threadIdx.x] = cols[col];
        else scol[threadIdx.x] = ColumnDev{0, 0, -1, 0, 0.0, 0.0, 0.0, 0.0, 0.0};
    }
    const int n = lane & 15, kq = lane >> 4;
    const int part = n & 1;
    const size_t chs = static_cast<size_t>(Nz) * Ny * nkxP;
    const int nsteps = Za / 4;
    const T* qw = reinterpret_cast<const T*>(qs) + 2 * (n >> 1) + part;
    double fre[4][2], fim[4][2];
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int ii = 0; ii < 2; ++ii) fre[f][ii] = fim[f][ii] = 0.0;
    for (int gi = 0; gi < plan.ngroups; ++gi) {
        const int nch = plan.g[gi].nch;
        __syncthreads();  // scol ready / previous group consumed
        for (int e = threadIdx.x; e < nch * Za * 8; e += blockDim.x) {
            const int c = e & 7, zq = e >> 3;
            const int q = zq / Za, z = zq - q * Za;
            const ColumnDev& cc = scol[c];
            cx<T> v{T(0), T(0)};
            if (cc.wv >= 0 && z < Nz) {
                const size_t o = (static_cast<size_t>(z) * Ny + cc.iy) * nkxP + cc.kx;
                const T rs = (cc.flags & kColConj) ? T(-1) : T(1);
                const T ks = rs * T(cc.k), ls = rs * T(cc.l);
                const int oV = plan.g[gi].ch[q].outV, oX = plan.g[gi].ch[q].outX, oY = plan.g[gi].ch[q].outY;
                if (oV >= 0) v = cadd(v, Hs[oV * chs + o]);
                if (oX >= 0) v = cadd(v, cmuli(Hs[oX * chs + o], -ks));
                if (oY >= 0) v = cadd(v, cmuli(Hs[oY * chs + o], -ls));
            }
            qs[e] = v;
        }
        __syncthreads();
        const T* Aj = matsT.m[plan.g[gi].mat] + static_cast<size_t>(jt) * nsteps * 64 + lane;
        for (int q = 0; q < nch; ++q) {
            typename MF::acc a4 = MF::zero();
            for (int s = 0; s < nsteps; ++s)
                a4 = MF::run(Aj[static_cast<size_t>(s) * 64], qw[2 * ((q * Za + 4 * s + kq) * 8)], a4);
            // complex B*(j) for this lane's items, transposed channel
            // coefficient and column conjugation
            const ColumnDev cc = scol[n >> 1];
            const int field = plan.g[gi].ch[q].field, dz = plan.g[gi].ch[q].dz;
            const double zsign = plan.g[gi].mat == 1 ? -0.5 : 0.5;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double mine = static_cast<double>(a4[i]);
                const double other = __shfl_xor(mine, 1);
                const double re = part ? other : mine, im = part ? mine : other;
                if ((i >> 1) == part) {
                    const int j = 16 * jt + MF::row(kq, i);
                    const double c = dz ? zsign * static_cast<double>(j < Nj ? mj[j] : T(0)) : plan.valueScale;
                    double vim = im;
                    if (cc.flags & kColConj) vim = -vim;
                    if (cc.flags & kColSelfConj) vim = 0.0;
#pragma unroll
                    for (int f = 0; f < 4; ++f)
                        if (f == field) {
                            fre[f][i & 1] += c * re;
                            fim[f][i & 1] += c * vim;
                        }
                }
            }
        }
    }
    const int col = cg * 8 + (n >> 1);
    if (col >= ncol) return;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
        const int j = 16 * jt + MF::row(kq, 2 * part + ii);
        if (j >= Nj) continue;
#pragma unroll
        for (int f = 0; f < 4; ++f)
            colbuf[(static_cast<size_t>(f) * ncol + col) * Nj + j] = {T(fre[f][ii]), T(fim[f][ii])};
    }
}

// S^T for the spectra of the fields (k_fields transposed): gathers the
// cotangent of each field's value at the mode's storage column and at its
// Hermitian-completion column, and maps it back to (Ap, Am, A0) at t0.
// mode 0: out = result; mode 1: out += result.
template <typename T>
__global__ __launch_bounds__(256) void k_fieldsadj(const cx<T>* __restrict__ colbuf, int ncol,
                                                   const int* __restrict__ mainColumn,
                                                   const int* __restrict__ compColumn, const double* __restrict__ omega,
                                                   const cx<double>* __restrict__ UAp, const cx<double>* __restrict__ VAp,
                                                   const cx<double>* __restrict__ WAp, const double* __restrict__ NAp,
                                                   const cx<double>* __restrict__ UA0, const cx<double>* __restrict__ VA0,
                                                   const double* __restrict__ NA0, double elapsed, int Nj, int M,
                                                   int mode, cx<double>* __restrict__ Gp, cx<double>* __restrict__ Gm,
                                                   cx<double>* __restrict__ G0) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M) return;
    const int j = idx % Nj, wv = idx / Nj;
    const int c0 = mainColumn[wv], c1 = compColumn[wv];
    cx<double> fb[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const cx<T> a = colbuf[(static_cast<size_t>(f) * ncol + c0) * Nj + j];
        fb[f] = {static_cast<double>(a.re), static_cast<double>(a.im)};
        if (c1 >= 0) {
            const cx<T> b = colbuf[(static_cast<size_t>(f) * ncol + c1) * Nj + j];
            fb[f] = cadd(fb[f], cx<double>{static_cast<double>(b.re), static_cast<double>(b.im)});
        }
    }
    // u = UAp ap + conj(UAp) am + UA0 a0, v likewise, w = WAp (ap + am),
    // eta = NAp (ap - am) + NA0 a0 (WVCoefficientFormulas.hpp:23-33)
    const cx<double> apB = cadd(cadd(cmul(cconj(UAp[idx]), fb[0]), cmul(cconj(VAp[idx]), fb[1])),
                                cadd(cmul(cconj(WAp[idx]), fb[2]), cscale(fb[3], NAp[idx])));
    const cx<double> amB = cadd(cadd(cmul(UAp[idx], fb[0]), cmul(VAp[idx], fb[1])),
                                cadd(cmul(cconj(WAp[idx]), fb[2]), cscale(fb[3], -NAp[idx])));
    const cx<double> a0B = cadd(cadd(cmul(cconj(UA0[idx]), fb[0]), cmul(cconj(VA0[idx]), fb[1])),
                                cscale(fb[3], NA0[idx]));
    double sn, cs;
    sincos(omega[idx] * elapsed, &sn, &cs);
    const cx<double> p{cs, sn};
    const cx<double> rp = cmul(apB, cconj(p)), rm = cmul(amB, p);
    if (mode == 0) {
        Gp[idx] = rp;
        Gm[idx] = rm;
        G0[idx] = a0B;
    } else {
        Gp[idx] = cadd(Gp[idx], rp);
        Gm[idx] = cadd(Gm[idx], rm);
        G0[idx] = cadd(G0[idx], a0B);
    }
}

// out_a = alpha x_a + beta y_a for the three coefficient arrays
// (x, y, out may alias elementwise).
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_axpby3(const cx<double>* x0, const cx<double>* x1, const cx<double>* x2,
                                                const cx<double>* y0, const cx<double>* y1, const cx<double>* y2,
                                                double alpha, double beta, int M, cx<double>* o0, cx<double>* o1,
                                                cx<double>* o2) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M) return;
    auto f = [&](const cx<double>* x, const cx<double>* y, cx<double>* o) {
        const cx<double> a = x[idx], b = y[idx];
        o[idx] = {addRn(mulRn(alpha, a.re), mulRn(beta, b.re)),
                  addRn(mulRn(alpha, a.im), mulRn(beta, b.im))};
    };
    f(x0, y0, o0);
    f(x1, y1, o1);
    f(x2, y2, o2);
}
