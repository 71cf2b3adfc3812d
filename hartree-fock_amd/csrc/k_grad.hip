// k_grad.hip - RHF nuclear-gradient kernels: the adjoint of the J/K build.
//
// No reference counterpart (the reference computes energies only, src/SCF.cpp:23-136).
// These kernels differentiate the energy the engine's integrals define:
//   dE/dX = sum P dh/dX - sum W dS/dX + 1/2 sum Gamma d(mn|ls)/dX      (+ dVnn/dX on the host)
//   Gamma_mnls = P_mn P_ls - 1/4 (P_ml P_ns + P_ms P_nl).
// A Cartesian Gaussian's centre derivative is d/dA_x g_i = 2a g_{i+1} - i g_{i-1}; on the 1-D
// Hermite coefficients of a product it is 2a E^{i+1,j}_t - i E^{i-1,j}_t (first centre) and
// 2b E^{i,j+1}_t - j E^{i,j-1}_t (second centre).  The Hermite Coulomb integrals R_tuv are
// those of the undifferentiated quartet, one order higher.  The operator derivative of the
// nuclear attraction is dR_tuv(P - C)/dC_x = -R_{t+1,u,v}.  Integrals are recomputed, never
// stored.
//
// Work layout: one block per work item.  The two-electron kernel (four wavefronts) takes one
// screened shell quartet of the J/K work list at a time (grid-stride); the block walks the
// primitive quartets together, builds the Boys values, R (in LDS, in place, one angular shell
// per step) and the E tables (in LDS), and the threads split the Cartesian component quartets.
// The one-electron kernel (one wavefront) takes one (shell pair, operator) item at a time the same way: the
// overlap and kinetic terms of the pair, or the attraction of one nucleus.  Each block adds its
// results into its own row of a partial-gradient array (no atomics); k_grad_reduce sums the
// rows.
#include "engine.hpp"
#include "md.hpp"

namespace hfx {
namespace {

constexpr int RD = 10;      // R cube edge: t, u, v <= 9 (L + 1 for (dd|dd))
constexpr int EI = 4;       // E tables of the two-electron kernel: i, j <= 3, t < 8
constexpr int ET = 8;
constexpr int OI = 4;       // E tables of the one-electron kernel: i <= 3, j <= 5, t < 10
constexpr int OJ = 6;
constexpr int OT = 10;

__device__ __forceinline__ int cart_exp(int l, int c, int d) {
    return d == 0 ? md::cart_i(l, c) : (d == 1 ? md::cart_j(l, c) : md::cart_k(l, c));
}

// 1-D Hermite coefficients E[i][j][t] (i <= li, j <= lj) of one direction, written by one
// lane: i-steps with PA at j = 0, then j-steps with PB for every i (the recursion of
// src/IntegralEngine.cpp:1025-1079); entries with t > i + j come out zero.
template <class R, int NI, int NJ, int NT>
__device__ void build_e(R* E, int li, int lj, R pa, R pb, R oo2p) {
    auto at = [&](int i, int j, int t) -> R& { return E[(i * NJ + j) * NT + t]; };
    for (int t = 0; t < NT; ++t) at(0, 0, t) = (t == 0) ? R(1) : R(0);
    for (int i = 0; i < li; ++i)
        for (int t = 0; t < NT; ++t) {
            R v = pa * at(i, 0, t);
            if (t > 0) v += oo2p * at(i, 0, t - 1);
            if (t + 1 < NT) v += R(t + 1) * at(i, 0, t + 1);
            at(i + 1, 0, t) = v;
        }
    for (int j = 0; j < lj; ++j)
        for (int i = 0; i <= li; ++i)
            for (int t = 0; t < NT; ++t) {
                R v = pb * at(i, j, t);
                if (t > 0) v += oo2p * at(i, j, t - 1);
                if (t + 1 < NT) v += R(t + 1) * at(i, j, t + 1);
                at(i, j + 1, t) = v;
            }
}

// F_n(T) for one order n per lane: asymptotic above T = 30, else a Taylor expansion around
// the nearest point of the precision's table (global memory, BOYS_ORDERS per point).
template <class R>
__device__ R boys_one(int n, R T, const R* __restrict__ tab) {
    if (T > R(30)) {
        const R rt = md::dev_rsqrt(T);
        R f = R(0.8862269254527579) * rt;
        const R invT = rt * rt;
        for (int m = 0; m < n; ++m) f = (R(m) + R(0.5)) * f * invT;
        return f;
    }
    constexpr int NT = md::BoysTerms<R>::nt;
    using G = md::BoysGrid<R>;
    int i = (int)(T * G::inv_step + R(0.5));
    i = i < 0 ? 0 : (i > G::points - 1 ? G::points - 1 : i);
    const R x = R(i) * G::step - T;
    const R* row = tab + i * md::BOYS_ORDERS + n;
    R t = row[NT - 1];
    for (int k = NT - 2; k >= 0; --k) t = row[k] + (x * (R(1) / R(k + 1))) * t;
    return t;
}

// R^0_tuv for t + u + v <= M into X (cube of edge RD), from Fs[n] = (-2 alpha)^n F_n:
// levels n = M-1 .. 0 are built in place, shells s = M-n .. 1 top down, so that every
// source (shells s-1, s-2) still holds level n+1.  Called by the whole block.
template <class R>
__device__ void build_r(R* X, const R* Fs, int M, R px, R py, R pz) {
    const int lane = threadIdx.x;
    if (lane == 0) X[0] = Fs[M];
    __syncthreads();
    for (int n = M - 1; n >= 0; --n) {
        for (int s = M - n; s >= 1; --s) {
            int k = lane;
            if (k < (s + 1) * (s + 2) / 2) {
                int t = 0, u = 0, v = 0;
                for (v = 0; v <= s; ++v) {
                    const int cnt = s - v + 1;
                    if (k < cnt) {
                        u = k;
                        t = s - v - u;
                        break;
                    }
                    k -= cnt;
                }
                R val;
                if (t > 0) {
                    val = px * X[(t - 1) + RD * (u + RD * v)];
                    if (t > 1) val += R(t - 1) * X[(t - 2) + RD * (u + RD * v)];
                } else if (u > 0) {
                    val = py * X[t + RD * ((u - 1) + RD * v)];
                    if (u > 1) val += R(u - 1) * X[t + RD * ((u - 2) + RD * v)];
                } else {
                    val = pz * X[t + RD * (u + RD * (v - 1))];
                    if (v > 1) val += R(v - 1) * X[t + RD * (u + RD * (v - 2))];
                }
                X[t + RD * (u + RD * v)] = val;
            }
            __syncthreads();
        }
        if (lane == 0) X[0] = Fs[n];
        __syncthreads();
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <class R>
struct GradLists;
template <>
struct GradLists<double> {
    __device__ static const int* ps(const GradArgs& g) { return g.pair_ps; }
    __device__ static const int* np(const GradArgs& g) { return g.pair_np; }
    __device__ static const PrimPair64* pp(const GradArgs& g) { return g.pp64; }
    __device__ static const double* ab(const GradArgs& g) { return g.pab64; }
    __device__ static const double* boys(const GradArgs& g) { return g.boys64; }
};
template <>
struct GradLists<float> {
    __device__ static const int* ps(const GradArgs& g) { return g.pair_ps32; }
    __device__ static const int* np(const GradArgs& g) { return g.pair_np32; }
    __device__ static const PrimPair32* pp(const GradArgs& g) { return g.pp32; }
    __device__ static const float* ab(const GradArgs& g) { return g.pab32; }
    __device__ static const float* boys(const GradArgs& g) { return g.boys32; }
};

// ---------------------------------------------------------------------------------------
// Two-electron part: 1/2 deg sum_{block} Gamma d(ab|cd)/dX for every screened quartet of
// one class; centres A, B, C explicitly, D = -(A + B + C) by translational invariance.
// ---------------------------------------------------------------------------------------
constexpr int GEN_T = 256;  // threads of the generic two-electron kernel's block (four wavefronts)

// At least four wavefronts per SIMD (128 VGPRs, a 160-byte spill in FP64): the kernel waits on
// LDS, and more resident blocks hide it (dddp 13.2 ms at two per SIMD, 10.6 at three, 8.9 at
// four, 8.5 at six with a 368-byte spill).
template <class R>
__global__ __launch_bounds__(GEN_T, 4) void k_grad_eri(GradArgs g, int la, int lb, int lc, int ld) {
    using L = GradLists<R>;
    __shared__ R X[RD * RD * RD];
    __shared__ R Eb[3 * EI * EI * ET];
    __shared__ R Ek[3 * EI * EI * ET];
    __shared__ double Gm[1296];
    __shared__ R Fs[16];
    __shared__ double red[GEN_T / 64][9];
    const int lane = threadIdx.x;
    const long long lo = g.range[0], hi = g.range[1];
    const int nA = md::ncart(la), nB = md::ncart(lb), nC = md::ncart(lc), nD = md::ncart(ld);
    const int ncomp = nA * nB * nC * nD;
    // blockIdx.y: this block's slice of the component quartets (a class with few quartets and
    // many components fills the machine only with several blocks per quartet)
    const int per = (ncomp + gridDim.y - 1) / gridDim.y;
    const int c_lo = blockIdx.y * per, c_hi = min(ncomp, c_lo + per);
    const int M = la + lb + lc + ld + 1;
    double* row = g.part + (size_t)blockIdx.x * g.part_stride + 2 * g.n3;
    for (long long q = lo + blockIdx.x; q < hi; q += gridDim.x) {
        const uint2 ent = g.list[q];
        const int p1 = (int)ent.x, p2 = (int)ent.y;
        const int sa = g.pair_a[p1], sb = g.pair_b[p1], sc = g.pair_a[p2], sd = g.pair_b[p2];
        const double deg = (sa != sb ? 2.0 : 1.0) * (sc != sd ? 2.0 : 1.0) * (p1 != p2 ? 2.0 : 1.0);
        const int ia = g.sh_ao[sa], ib = g.sh_ao[sb], ic = g.sh_ao[sc], id = g.sh_ao[sd];
        const int N = g.n_ao;
        // Gamma of the block, with the Cartesian normalization ratios and 1/2 deg folded in
        for (int c = c_lo + lane; c < c_hi; c += GEN_T) {
            const int D = c % nD, C = (c / nD) % nC, B = (c / (nD * nC)) % nB, A = c / (nD * nC * nB);
            const size_t i = ia + A, j = ib + B, k = ic + C, l = id + D;
            auto P = [&](size_t r, size_t s) { return g.P[r + s * N]; };
            const double gam = P(i, j) * P(k, l) - 0.25 * (P(i, k) * P(j, l) + P(i, l) * P(j, k));
            Gm[c] = 0.5 * deg * gam * g.sh_cf[6 * sa + A] * g.sh_cf[6 * sb + B] * g.sh_cf[6 * sc + C] *
                    g.sh_cf[6 * sd + D];
        }
        double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
        const int ps1 = L::ps(g)[p1], np1 = L::np(g)[p1], ps2 = L::ps(g)[p2], np2 = L::np(g)[p2];
        for (int x1 = 0; x1 < np1; ++x1) {
            const auto pb = L::pp(g)[ps1 + x1];
            const R ea = L::ab(g)[2 * (ps1 + x1)], eb = L::ab(g)[2 * (ps1 + x1) + 1];
            __syncthreads();  // Eb is still read by the previous bra primitive's lanes
            if (lane < 3) {
                const R pa_d = lane == 0 ? pb.pax : (lane == 1 ? pb.pay : pb.paz);
                const R pb_d = lane == 0 ? pb.pbx : (lane == 1 ? pb.pby : pb.pbz);
                build_e<R, EI, EI, ET>(Eb + lane * EI * EI * ET, la + 1, lb + 1, pa_d, pb_d, pb.oo2p);
            }
            for (int x2 = 0; x2 < np2; ++x2) {
                const auto pk = L::pp(g)[ps2 + x2];
                const R ec = L::ab(g)[2 * (ps2 + x2)], ed = L::ab(g)[2 * (ps2 + x2) + 1];
                (void)ed;
                const R pq = pb.p + pk.p;
                const R alpha = pb.p * pk.p / pq;
                const R px = pb.px - pk.px, py = pb.py - pk.py, pz = pb.pz - pk.pz;
                const R T = alpha * (px * px + py * py + pz * pz);
                const R pref = R(34.98683665524972) * pb.kp * pk.kp * md::dev_rsqrt(pq);  // 2 pi^2.5
                __syncthreads();  // X, Ek and Fs are still read by the previous primitive quartet
                if (lane <= M) {
                    R f = boys_one<R>(lane, T, L::boys(g));
                    R m2a = R(1);
                    for (int n = 0; n < lane; ++n) m2a *= R(-2) * alpha;
                    Fs[lane] = m2a * f;
                }
                if (lane >= 32 && lane < 35) {
                    const int d = lane - 32;
                    const R pa_d = d == 0 ? pk.pax : (d == 1 ? pk.pay : pk.paz);
                    const R pb_d = d == 0 ? pk.pbx : (d == 1 ? pk.pby : pk.pbz);
                    build_e<R, EI, EI, ET>(Ek + d * EI * EI * ET, lc + 1, ld + 1, pa_d, pb_d, pk.oo2p);
                }
                __syncthreads();
                build_r<R>(X, Fs, M, px, py, pz);
                for (int c = c_lo + lane; c < c_hi; c += GEN_T) {
                    const R w = R(Gm[c]) * pref;
                    if (w == R(0)) continue;
                    const int D = c % nD, C = (c / nD) % nC, B = (c / (nD * nC)) % nB, A = c / (nD * nC * nB);
                    int e1[3], e2[3], e3[3], e4[3];
                    for (int d = 0; d < 3; ++d) {
                        e1[d] = cart_exp(la, A, d);
                        e2[d] = cart_exp(lb, B, d);
                        e3[d] = cart_exp(lc, C, d);
                        e4[d] = cart_exp(ld, D, d);
                    }
                    auto E1 = [&](int d, int i, int j, int t) -> R {
                        return (i < 0 || j < 0) ? R(0) : Eb[((d * EI + i) * EI + j) * ET + t];
                    };
                    auto E2 = [&](int d, int i, int j, int t) -> R {
                        return (i < 0 || j < 0) ? R(0) : Ek[((d * EI + i) * EI + j) * ET + t];
                    };
                    // bra: base, d/dA, d/dB of dimension d at Hermite index t
                    auto eb0 = [&](int d, int t) { return E1(d, e1[d], e2[d], t); };
                    auto ebA = [&](int d, int t) {
                        return R(2) * ea * E1(d, e1[d] + 1, e2[d], t) - R(e1[d]) * E1(d, e1[d] - 1, e2[d], t);
                    };
                    auto ebB = [&](int d, int t) {
                        return R(2) * eb * E1(d, e1[d], e2[d] + 1, t) - R(e2[d]) * E1(d, e1[d], e2[d] - 1, t);
                    };
                    // ket: base and d/dC with the Hermite sign (-1)^t folded in
                    auto sg = [](int t) { return (t & 1) ? R(-1) : R(1); };
                    auto ek0 = [&](int d, int t) { return sg(t) * E2(d, e3[d], e4[d], t); };
                    auto ekC = [&](int d, int t) {
                        return sg(t) * (R(2) * ec * E2(d, e3[d] + 1, e4[d], t) - R(e3[d]) * E2(d, e3[d] - 1, e4[d], t));
                    };
                    const int lx = e1[0] + e2[0], ly = e1[1] + e2[1], lz = e1[2] + e2[2];
                    const int kx = e3[0] + e4[0], ky = e3[1] + e4[1], kz = e3[2] + e4[2];
                    R sA[3] = {0, 0, 0}, sB[3] = {0, 0, 0}, sC[3] = {0, 0, 0};
                    // bra derivatives: the bra box reaches one index further in each direction
                    for (int v = 0; v <= lz + 1; ++v)
                        for (int u = 0; u <= ly + 1; ++u)
                            for (int t = 0; t <= lx + 1; ++t) {
                                const int over = (t > lx) + (u > ly) + (v > lz);
                                if (over > 1) continue;
                                R s = 0;
                                for (int vv = 0; vv <= kz; ++vv) {
                                    const R fz = ek0(2, vv);
                                    for (int uu = 0; uu <= ky; ++uu) {
                                        const R fyz = fz * ek0(1, uu);
                                        for (int tt = 0; tt <= kx; ++tt)
                                            s += fyz * ek0(0, tt) * X[(t + tt) + RD * ((u + uu) + RD * (v + vv))];
                                    }
                                }
                                const R x0 = eb0(0, t), y0 = eb0(1, u), z0 = eb0(2, v);
                                sA[0] += ebA(0, t) * y0 * z0 * s;
                                sB[0] += ebB(0, t) * y0 * z0 * s;
                                sA[1] += x0 * ebA(1, u) * z0 * s;
                                sB[1] += x0 * ebB(1, u) * z0 * s;
                                sA[2] += x0 * y0 * ebA(2, v) * s;
                                sB[2] += x0 * y0 * ebB(2, v) * s;
                            }
                    // ket derivatives on centre C
                    for (int v = 0; v <= lz; ++v)
                        for (int u = 0; u <= ly; ++u)
                            for (int t = 0; t <= lx; ++t) {
                                const R e = eb0(0, t) * eb0(1, u) * eb0(2, v);
                                if (e == R(0)) continue;
                                R cx = 0, cy = 0, cz = 0;
                                for (int vv = 0; vv <= kz + 1; ++vv)
                                    for (int uu = 0; uu <= ky + 1; ++uu)
                                        for (int tt = 0; tt <= kx + 1; ++tt) {
                                            const int over = (tt > kx) + (uu > ky) + (vv > kz);
                                            if (over > 1) continue;
                                            const R r = X[(t + tt) + RD * ((u + uu) + RD * (v + vv))];
                                            const R x0 = ek0(0, tt), y0 = ek0(1, uu), z0 = ek0(2, vv);
                                            cx += ekC(0, tt) * y0 * z0 * r;
                                            cy += x0 * ekC(1, uu) * z0 * r;
                                            cz += x0 * y0 * ekC(2, vv) * r;
                                        }
                                sC[0] += e * cx;
                                sC[1] += e * cy;
                                sC[2] += e * cz;
                            }
                    for (int d = 0; d < 3; ++d) {
                        gA[d] += (double)(w * sA[d]);
                        gB[d] += (double)(w * sB[d]);
                        gC[d] += (double)(w * sC[d]);
                    }
                }
            }
        }
        double v[9];
        for (int d = 0; d < 3; ++d) {
            v[d] = wave_sum(gA[d]);
            v[3 + d] = wave_sum(gB[d]);
            v[6 + d] = wave_sum(gC[d]);
        }
        if ((lane & 63) == 0)
            for (int x = 0; x < 9; ++x) red[lane >> 6][x] = v[x];
        __syncthreads();
        if (lane == 0) {
            for (int w = 1; w < GEN_T / 64; ++w)
                for (int x = 0; x < 9; ++x) v[x] += red[w][x];
            const int aa = g.sh_atom[sa], ab = g.sh_atom[sb], ac = g.sh_atom[sc], ad = g.sh_atom[sd];
            // the blocks of the other slices add into the same row
            for (int d = 0; d < 3; ++d) {
                atomicAdd(&row[3 * aa + d], v[d]);
                atomicAdd(&row[3 * ab + d], v[3 + d]);
                atomicAdd(&row[3 * ac + d], v[6 + d]);
                atomicAdd(&row[3 * ad + d], -(v[d] + v[3 + d] + v[6 + d]));
            }
        }
        __syncthreads();  // Gm and red are rewritten by the next quartet
    }
}

// ---------------------------------------------------------------------------------------
// One-electron part over the canonical shell pairs: -W dS and P dT on both centres, P dV
// on both centres and, for every nucleus C, the operator term on C.  A pair with a != b
// stands for both orders of its block.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_grad_one(GradArgs g) {
    __shared__ double E[3 * OI * OJ * OT];
    __shared__ double X[RD * RD * RD];
    __shared__ double Fs[16];
    const int lane = threadIdx.x;
    double* rowS = g.part + (size_t)blockIdx.x * g.part_stride;
    double* rowH = rowS + g.n3;
    const int N = g.n_ao;
    // one work item per (shell pair, operator): item % (n_atoms + 1) = 0 is the overlap and
    // kinetic terms, 1 + C the attraction of nucleus C
    const long long n_items = (long long)g.n_pairs * (g.n_atoms + 1);
    for (long long item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int pid = (int)(item / (g.n_atoms + 1));
        const int C = (int)(item % (g.n_atoms + 1)) - 1;
        const int sa = g.pair_a[pid], sb = g.pair_b[pid];
        const int la = g.sh_l[sa], lb = g.sh_l[sb];
        const int nA = md::ncart(la), nB = md::ncart(lb), ncomp = nA * nB;
        const double mult = (sa != sb) ? 2.0 : 1.0;
        const double* A = g.ctr + 3 * sa;
        const double* B = g.ctr + 3 * sb;
        const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
        const int ia = g.sh_ao[sa], ib = g.sh_ao[sb];
        double gSA[3] = {0, 0, 0}, gSB[3] = {0, 0, 0}, gHA[3] = {0, 0, 0}, gHB[3] = {0, 0, 0};
        const int M = la + lb + 1;
        {
            double gCC[3] = {0, 0, 0};
            for (int x1 = 0; x1 < g.sh_np[sa]; ++x1)
                for (int x2 = 0; x2 < g.sh_np[sb]; ++x2) {
                    const double a = g.alpha[g.sh_po[sa] + x1], b = g.alpha[g.sh_po[sb] + x2];
                    const double p = a + b, oop = 1.0 / p;
                    const double P3[3] = {(a * A[0] + b * B[0]) * oop, (a * A[1] + b * B[1]) * oop,
                                          (a * A[2] + b * B[2]) * oop};
                    const double K = exp(-a * b * oop * AB2);
                    const double cp = g.coeff[g.sh_po[sa] + x1] * g.coeff[g.sh_po[sb] + x2];
                    __syncthreads();  // E, X, Fs are still read by the previous primitive pair
                    if (lane < 3)
                        build_e<double, OI, OJ, OT>(E + lane * OI * OJ * OT, la + 1, lb + 3, P3[lane] - A[lane],
                                                    P3[lane] - B[lane], 0.5 * oop);
                    if (C >= 0) {
                        const double* Cx = g.atom_xyz + 3 * C;
                        const double pcx = P3[0] - Cx[0], pcy = P3[1] - Cx[1], pcz = P3[2] - Cx[2];
                        const double T = p * (pcx * pcx + pcy * pcy + pcz * pcz);
                        if (lane <= M) {
                            double f = boys_one<double>(lane, T, g.boys64);
                            double m2p = 1.0;
                            for (int n = 0; n < lane; ++n) m2p *= -2.0 * p;
                            Fs[lane] = m2p * f;
                        }
                        __syncthreads();
                        build_r<double>(X, Fs, M, pcx, pcy, pcz);
                    } else {
                        __syncthreads();
                    }
                    auto Et = [&](int d, int i, int j, int t) -> double {
                        return (i < 0 || j < 0) ? 0.0 : E[((d * OI + i) * OJ + j) * OT + t];
                    };
                    for (int c = lane; c < ncomp; c += 64) {
                        const int Y = c % nB, Xc = c / nB;
                        int c1[3], c2[3];
                        for (int d = 0; d < 3; ++d) {
                            c1[d] = cart_exp(la, Xc, d);
                            c2[d] = cart_exp(lb, Y, d);
                        }
                        const int i = ia + Xc, j = ib + Y;
                        const double f = g.norm[g.sh_no[sa] + x1 * nA + Xc] * g.norm[g.sh_no[sb] + x2 * nB + Y] * cp *
                                         mult;
                        const double pd = g.P[i + (size_t)j * N];
                        // derivative-or-base 1-D coefficient: der 0 base, 1 d/dA, 2 d/dB
                        auto ed = [&](int d, int jj, int t, int der) -> double {
                            const int ii = c1[d];
                            if (der == 0) return Et(d, ii, jj, t);
                            if (der == 1) return 2.0 * a * Et(d, ii + 1, jj, t) - ii * Et(d, ii - 1, jj, t);
                            return 2.0 * b * Et(d, ii, jj + 1, t) - jj * Et(d, ii, jj - 1, t);
                        };
                        if (C < 0) {
                            const double w = g.W[i + (size_t)j * N];
                            const double sq = pow(M_PI * oop, 1.5) * K * f;
                            auto kin = [&](int dx, int der) {
                                double sx[3], tx[3];
                                for (int d = 0; d < 3; ++d) {
                                    const int dd = (d == dx) ? der : 0;
                                    const int l2 = c2[d];
                                    sx[d] = ed(d, l2, 0, dd);
                                    double t = b * (2 * l2 + 1) * sx[d] - 2.0 * b * b * ed(d, l2 + 2, 0, dd);
                                    if (l2 >= 2) t -= 0.5 * l2 * (l2 - 1) * ed(d, l2 - 2, 0, dd);
                                    tx[d] = t;
                                }
                                return tx[0] * sx[1] * sx[2] + sx[0] * tx[1] * sx[2] + sx[0] * sx[1] * tx[2];
                            };
                            for (int x = 0; x < 3; ++x)
                                for (int der = 1; der <= 2; ++der) {
                                    double ds = 1.0;
                                    for (int y = 0; y < 3; ++y) ds *= ed(y, c2[y], 0, y == x ? der : 0);
                                    (der == 1 ? gSA : gSB)[x] -= w * sq * ds;
                                    (der == 1 ? gHA : gHB)[x] += pd * sq * kin(x, der);
                                }
                        } else {
                            const double zf = -g.atom_z[C] * (2.0 * M_PI * oop) * K * f * pd;
                            if (zf == 0.0) continue;
                            const int nx = c1[0] + c2[0] + 1, ny = c1[1] + c2[1] + 1, nz = c1[2] + c2[2] + 1;
                            double sa_[3] = {0, 0, 0}, sb_[3] = {0, 0, 0}, sc_[3] = {0, 0, 0};
                            for (int v = 0; v <= nz; ++v)
                                for (int u = 0; u <= ny; ++u)
                                    for (int t = 0; t <= nx; ++t) {
                                        const int over = (t == nx) + (u == ny) + (v == nz);
                                        if (over > 1) continue;
                                        const double r = X[t + RD * (u + RD * v)];
                                        const double x0 = ed(0, c2[0], t, 0), y0 = ed(1, c2[1], u, 0),
                                                     z0 = ed(2, c2[2], v, 0);
                                        sa_[0] += ed(0, c2[0], t, 1) * y0 * z0 * r;
                                        sb_[0] += ed(0, c2[0], t, 2) * y0 * z0 * r;
                                        sa_[1] += x0 * ed(1, c2[1], u, 1) * z0 * r;
                                        sb_[1] += x0 * ed(1, c2[1], u, 2) * z0 * r;
                                        sa_[2] += x0 * y0 * ed(2, c2[2], v, 1) * r;
                                        sb_[2] += x0 * y0 * ed(2, c2[2], v, 2) * r;
                                        if (over == 0) {
                                            const double e = x0 * y0 * z0;
                                            sc_[0] -= e * X[(t + 1) + RD * (u + RD * v)];
                                            sc_[1] -= e * X[t + RD * ((u + 1) + RD * v)];
                                            sc_[2] -= e * X[t + RD * (u + RD * (v + 1))];
                                        }
                                    }
                            for (int d = 0; d < 3; ++d) {
                                gHA[d] += zf * sa_[d];
                                gHB[d] += zf * sb_[d];
                                gCC[d] += zf * sc_[d];
                            }
                        }
                    }
                }
            if (C >= 0) {
                double v[3];
                for (int d = 0; d < 3; ++d) v[d] = wave_sum(gCC[d]);
                if (lane == 0)
                    for (int d = 0; d < 3; ++d) rowH[3 * C + d] += v[d];
            }
        }
        double v[12];
        for (int d = 0; d < 3; ++d) {
            v[d] = wave_sum(gSA[d]);
            v[3 + d] = wave_sum(gSB[d]);
            v[6 + d] = wave_sum(gHA[d]);
            v[9 + d] = wave_sum(gHB[d]);
        }
        if (lane == 0) {
            const int aa = g.sh_atom[sa], ab = g.sh_atom[sb];
            for (int d = 0; d < 3; ++d) {
                if (C < 0) {
                    rowS[3 * aa + d] += v[d];
                    rowS[3 * ab + d] += v[3 + d];
                }
                rowH[3 * aa + d] += v[6 + d];
                rowH[3 * ab + d] += v[9 + d];
            }
        }
        __syncthreads();  // E, X and Fs are rewritten by the next item
    }
}

// out[x] = sum over blocks of part[block][x], x < 3 * n3.
__global__ void k_grad_reduce(const double* __restrict__ part, long long stride, int n_rows, int n, double* out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    double s = 0.0;
    for (int r = 0; r < n_rows; ++r) s += part[(size_t)r * stride + x];
    out[x] = s;
}

}  // namespace

// Launchers of the compile-time-unrolled classes (k_grad_eri.hip: one per chunk of bra
// components), registered at load time, kept in chunk order.
namespace {
constexpr int GRAD_UNITS_MAX = 64;
struct GradUnits {
    int n = 0;
    bool fp32 = true;  // every unit of the class has FP32 kernels
    int first[GRAD_UNITS_MAX];
    GradClassFn fn[GRAD_UNITS_MAX];
};
GradUnits& grad_units(int combo) {
    static GradUnits table[N_COMBOS];
    return table[combo];
}
}  // namespace

bool register_grad_class(int combo, int first_chunk, GradClassFn fn, bool fp32) {
    if (combo < 0 || combo >= N_COMBOS) return false;
    GradUnits& u = grad_units(combo);
    if (u.n == GRAD_UNITS_MAX) return false;
    u.fp32 = u.fp32 && fp32;
    int i = u.n++;
    for (; i > 0 && u.first[i - 1] > first_chunk; --i) {
        u.first[i] = u.first[i - 1];
        u.fn[i] = u.fn[i - 1];
    }
    u.first[i] = first_chunk;
    u.fn[i] = fn;
    return true;
}

bool grad_class_has_fp32(int la, int lb, int lc, int ld) {
    const GradUnits& u = grad_units(combo_of(pair_class(la, lb), pair_class(lc, ld)));
    return u.n > 0 && u.fp32;
}

int grad_class_units(int la, int lb, int lc, int ld) {
    return grad_units(combo_of(pair_class(la, lb), pair_class(lc, ld))).n;
}

void launch_grad_unit(int fp32, const GradArgs& g, int la, int lb, int lc, int ld, int unit, int grid, hipStream_t s) {
    const GradUnits& u = grad_units(combo_of(pair_class(la, lb), pair_class(lc, ld)));
    if (u.n > 0) {
        u.fn[unit](fp32, g, grid, s);
        return;
    }
    const int ncomp = md::ncart(la) * md::ncart(lb) * md::ncart(lc) * md::ncart(ld);
    const dim3 blocks(grid, std::min(4, (ncomp + GEN_T - 1) / GEN_T));  // component slices
    if (fp32)
        hipLaunchKernelGGL(k_grad_eri<float>, blocks, dim3(GEN_T), 0, s, g, la, lb, lc, ld);
    else
        hipLaunchKernelGGL(k_grad_eri<double>, blocks, dim3(GEN_T), 0, s, g, la, lb, lc, ld);
}

void launch_grad_one(const GradArgs& g, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_grad_one, dim3(grid), dim3(64), 0, s, g);
}

void launch_grad_reduce(const GradArgs& g, int n_rows, double* out, hipStream_t s) {
    const int n = 3 * g.n3;
    hipLaunchKernelGGL(k_grad_reduce, dim3((n + 255) / 256), dim3(256), 0, s, g.part, g.part_stride, n_rows, n, out);
}

}  // namespace hfx
