// k_eri.hip - contracted ERIs of one quartet class, digested into J/K on the fly.
//
// Compiled once per class with -DHFX_COMBO=k (k = pc1(pc1+1)/2 + pc2, pair class
// pc = la(la+1)/2 + lb), so the 21 classes build in parallel.
//
// One thread owns one shell quartet (bra pair, ket pair) from the class's work list and a
// chunk of CB bra Cartesian components; chunks are separate launches so that every array
// index is a compile-time constant.  Per primitive quartet the thread evaluates the Boys
// function (table slice in LDS), the Hermite R^0_{tuv} (scaled by the prefactor), the bra
// contraction Y_{t'u'v'} = sum_tuv E^{ab}_tuv R_{t+t',u+u',v+v'} for each of its bra
// components, and accumulates sum_{t'u'v'} (-1)^{t'+u'+v'} E^{cd}_{t'u'v'} Y_{t'u'v'} into
// the contracted integrals.  This is the primitive integral of the reference's
// electronRepulsion (src/IntegralEngine.cpp:897-934) in factored order.
//
// Schedule: blocks claim contiguous chunks of the class's list (bra-major), so a block keeps
// one bra pair (a,b) for many consecutive quartets.  The exchange rows of that pair,
// K[a,:] and K[b,:], are collected in LDS with LDS atomics, with a mask of the ket shells
// touched, and the touched columns are added to the accumulators once when the block moves
// to another bra pair; quartets of other bra pairs in the same iteration, and the Coulomb
// terms of the ket pairs, go to the accumulators directly.
//
// Digestion: the full block of the quartet is used with the permutational degeneracy
// s = (a==b ? 1/2 : 1)(c==d ? 1/2 : 1)(ab==cd ? 1/2 : 1); the integrals below tau_S are
// dropped as the reference's AO test (src/SCF.cpp:670-671).  Contributions go to
//   AJ[i,j] += 2 s I D[k,l],  AJ[k,l] += 2 s I D[i,j],
//   AK[i,k] += s I D[j,l], AK[i,l] += s I D[j,k], AK[j,k] += s I D[i,l], AK[j,l] += s I D[i,k]
// accumulated per thread in registers and flushed with FP64 atomics (or, in the deterministic
// mode, two-word fixed-point integer atomics) into packed lower
// triangles (both orientations of (i,j) land in one element, i.e. J = AJ + AJ^T is formed
// on the fly; the diagonal is doubled when the triangles are expanded).  This equals the reference's canonical-AO digestion
// (src/SCF.cpp:633-711) summed over the 8 permutations.
#include "digest.hpp"
#include "engine.hpp"
#include "md.hpp"

#ifndef HFX_COMBO
#error "HFX_COMBO must be defined"
#endif

#define HFX_CAT2(a, b) a##b
#define HFX_CAT(a, b) HFX_CAT2(a, b)

namespace hfx {
namespace {

using namespace md;
using namespace digest;

constexpr int K_COMBO = HFX_COMBO;
constexpr int PC1 = combo_pc1(K_COMBO), PC2 = combo_pc2(K_COMBO);
constexpr int LA = pc_la(PC1), LB = pc_lb(PC1), LC = pc_la(PC2), LD = pc_lb(PC2);
constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
constexpr int NAB = NA * NB, NCD = NC * ND;
// Chunking of the bra components: a thread accumulates at most `acc` integrals (cb bra
// components x all NCD ket components), so a class is split into nch chunks.  The classes with
// L >= 4 run at one wavefront per SIMD whatever their accumulators, so their threads hold more
// (FP32: 72, 108 for L >= 5; FP64: 72): fewer chunks repeat the Boys function, R^0_tuv and the
// E tables fewer times.
constexpr int ACC32_HI = 108;
constexpr int ACC64_HI = 72;
template <class R>
struct Chunks {
    static constexpr int acc = sizeof(R) == 4 ? (LA + LB + LC + LD >= 5 ? ACC32_HI : (LA + LB + LC + LD >= 4 ? 72 : 36))
                                              : (LA + LB + LC + LD >= 4 ? ACC64_HI : 36);
    static constexpr int cb_max = cmax(1, acc / NCD);
    static constexpr int nch = (NAB + cb_max - 1) / cb_max;
    static constexpr int cb = (NAB + nch - 1) / nch;
};
constexpr int NCH = Chunks<double>::nch;  // the translation units are sliced by FP64 chunks
constexpr int LTOT = LA + LB + LC + LD;
constexpr int LCD = LC + LD;
constexpr double TWO_PI_2_5 = 34.98683665524972;  // 2 pi^(5/2)
// Light classes take two ket primitive pairs per bra load.
constexpr bool KET_PAIRS = LTOT <= 2;

template <class R>
__device__ __forceinline__ const PrimPair<R>* pp_of(const EriArgs& a);
template <>
__device__ __forceinline__ const PrimPair<double>* pp_of<double>(const EriArgs& a) { return a.pp64; }
template <>
__device__ __forceinline__ const PrimPair<float>* pp_of<float>(const EriArgs& a) { return a.pp32; }
template <class R>
__device__ __forceinline__ const R* boys_of(const EriArgs& a);
template <>
__device__ __forceinline__ const double* boys_of<double>(const EriArgs& a) { return a.boys64; }
template <>
__device__ __forceinline__ const float* boys_of<float>(const EriArgs& a) { return a.boys32; }

// Is bra component row ia (of shell a) / ib (of shell b) touched by chunk C0?
__host__ __device__ constexpr bool chunk_has_a(int c0, int cb, int ia) {
    for (int ab = c0; ab < c0 + cb && ab < NAB; ++ab)
        if (ab / NB == ia) return true;
    return false;
}
__host__ __device__ constexpr bool chunk_has_b(int c0, int cb, int ib) {
    for (int ab = c0; ab < c0 + cb && ab < NAB; ++ab)
        if (ab % NB == ib) return true;
    return false;
}

// Stored blocks (the stored-integral mode): the block of the r-th entry of a class list sits in
// a group of 64 entries, element-major: element ab*NCD + cd of entry r at
// ((r / 64) * NAB*NCD + ab*NCD + cd) * 64 + r % 64, so the 64 lanes of a wavefront, which hold 64
// consecutive entries, read every element with one contiguous 512-byte access.
__device__ __forceinline__ long long stored_at(long long r, int e) {
    return ((r >> 6) * (long long)(NAB * NCD) + e) * 64 + (r & 63);
}

// tile: the block's LDS exchange rows (NA + NB rows of a.tile_n, then the same for K_beta),
// holding bra pair tile_pid; NULL when the block keeps no rows.
// STORED: the integrals are read from a.stored (normalized, written by the store mode) instead
// of being computed; no Boys table.
template <class R, int C0, bool STORED, bool FX>
__device__ __forceinline__ void quartet(const EriArgs& a, long long qi, bool valid, const R* tab, double* tile,
                                        int tile_pid) {
    constexpr int CB = Chunks<R>::cb;
    const uint2 e = a.list[qi];
    const int pid = (int)e.x, qid = (int)e.y;
    const int* __restrict__ pps = sizeof(R) == 8 ? a.pair_ps : a.pair_ps32;
    const int* __restrict__ pnp = sizeof(R) == 8 ? a.pair_np : a.pair_np32;
    const int ps = STORED ? 0 : pps[pid], np = (valid && !STORED) ? pnp[pid] : 0;
    const int qs = STORED ? 0 : pps[qid], nq = (valid && !STORED) ? pnp[qid] : 0;
    const PrimPair<R>* __restrict__ pp = pp_of<R>(a);

    R acc[CB][NCD];
    sfor<CB>([&](auto c) { sfor<NCD>([&](auto d) { acc[c][d] = R(0); }); });

    // One primitive quartet: bra primitive pair A (with its E coefficients), ket B (with F).
    auto prim = [&](const PrimPair<R>& A, const Ecoef<LA, LB, R>& Ex, const Ecoef<LA, LB, R>& Ey,
                    const Ecoef<LA, LB, R>& Ez, const PrimPair<R>& B, const Ecoef<LC, LD, R>& Fx,
                    const Ecoef<LC, LD, R>& Fy, const Ecoef<LC, LD, R>& Fz) {
        const R pq = A.p + B.p;
        const R rs = dev_rsqrt(pq);
        const R alpha = (A.p * B.p) * (rs * rs);
        const R X = A.px - B.px, Y = A.py - B.py, Z = A.pz - B.pz;
        const R T = alpha * (X * X + Y * Y + Z * Z);
        const R pref = R(TWO_PI_2_5) * rs * A.kp * B.kp;
        R F[LTOT + 1];
        boys<LTOT>(T, tab, F);
        R base[LTOT + 1];
        {
            const R m2a = R(-2) * alpha;
            R pw = pref;
            sfor<LTOT + 1>([&](auto n) {
                base[n] = pw * F[n];
                if constexpr (decltype(n)::value < LTOT) pw *= m2a;
            });
        }
        R Rt[nherm(LTOT)];
        hermite_r<LTOT>(base, X, Y, Z, Rt);

        sfor<CB>([&](auto c_) {
            constexpr int c = decltype(c_)::value;
            constexpr int ab = C0 + c;
            if constexpr (ab < NAB) {
                constexpr int ia = ab / NB, ib = ab % NB;
                constexpr int ax = cart_i(LA, ia), ay = cart_j(LA, ia), az = cart_k(LA, ia);
                constexpr int bx = cart_i(LB, ib), by = cart_j(LB, ib), bz = cart_k(LB, ib);
                R Yv[nherm(LCD)];
                sfor<nherm(LCD)>([&](auto h_) {
                    constexpr int h = decltype(h_)::value;
                    constexpr int t2 = h_t(h), u2 = h_u(h), v2 = h_v(h);
                    Yv[h] = sum_of<az + bz + 1>([&](auto v_) {
                        constexpr int v = decltype(v_)::value;
                        return Ez.e[az][bz][v] * sum_of<ay + by + 1>([&](auto u_) {
                                   constexpr int u = decltype(u_)::value;
                                   return Ey.e[ay][by][u] * sum_of<ax + bx + 1>([&](auto t_) {
                                              constexpr int t = decltype(t_)::value;
                                              return Ex.e[ax][bx][t] * Rt[hidx(t + t2, u + u2, v + v2)];
                                          });
                               });
                    });
                });
                sfor<NCD>([&](auto cd_) {
                    constexpr int cd = decltype(cd_)::value;
                    constexpr int jc = cd / ND, jd = cd % ND;
                    constexpr int cx = cart_i(LC, jc), cy = cart_j(LC, jc), cz = cart_k(LC, jc);
                    constexpr int dx = cart_i(LD, jd), dy = cart_j(LD, jd), dz = cart_k(LD, jd);
                    acc[c][cd] += sum_of<cz + dz + 1>([&](auto v_) {
                        constexpr int v = decltype(v_)::value;
                        return Fz.e[cz][dz][v] * sum_of<cy + dy + 1>([&](auto u_) {
                                   constexpr int u = decltype(u_)::value;
                                   return Fy.e[cy][dy][u] * sum_of<cx + dx + 1>([&](auto t_) {
                                              constexpr int t = decltype(t_)::value;
                                              constexpr R sg = ((t + u + v) & 1) ? R(-1) : R(1);
                                              return (sg * Fx.e[cx][dx][t]) * Yv[hidx(t, u, v)];
                                          });
                               });
                    });
                });
            }
        });
    };

    // Outer loop over the ket's primitive pairs (private to the lane), inner loop over the
    // bra's: the work list groups quartets by bra pair, so the inner loads are shared by the
    // lanes of a wavefront.  For the light classes the outer loop takes two ket primitive
    // pairs at a time, so each bra load serves two independent primitive quartets; an odd
    // last one gets a pass of its own.
    int iq = 0;
    if constexpr (KET_PAIRS) {
        // In FP32 the bra loop is unrolled by two with the loads alternating between A0 and A1,
        // so no register copies are needed for the one-ahead prefetch; pp[ps + np] is the next
        // pair's first primitive pair or the zero element past the end: always readable.
        const PrimPair<R>* __restrict__ pa = pp + ps;
        for (; iq + 1 < nq; iq += 2) {
            const PrimPair<R> B0 = pp[qs + iq], B1 = pp[qs + iq + 1];
            Ecoef<LC, LD, R> Fx0, Fy0, Fz0, Fx1, Fy1, Fz1;
            hermite_e<LC, LD>(B0.oo2p, B0.pax, B0.pbx, Fx0);
            hermite_e<LC, LD>(B0.oo2p, B0.pay, B0.pby, Fy0);
            hermite_e<LC, LD>(B0.oo2p, B0.paz, B0.pbz, Fz0);
            hermite_e<LC, LD>(B1.oo2p, B1.pax, B1.pbx, Fx1);
            hermite_e<LC, LD>(B1.oo2p, B1.pay, B1.pby, Fy1);
            hermite_e<LC, LD>(B1.oo2p, B1.paz, B1.pbz, Fz1);
            auto both = [&](const PrimPair<R>& A) {
                Ecoef<LA, LB, R> Ex, Ey, Ez;
                hermite_e<LA, LB>(A.oo2p, A.pax, A.pbx, Ex);
                hermite_e<LA, LB>(A.oo2p, A.pay, A.pby, Ey);
                hermite_e<LA, LB>(A.oo2p, A.paz, A.pbz, Ez);
                prim(A, Ex, Ey, Ez, B0, Fx0, Fy0, Fz0);
                prim(A, Ex, Ey, Ez, B1, Fx1, Fy1, Fz1);
            };
            if constexpr (sizeof(R) == 4) {
                PrimPair<R> A0 = pa[0];
                int ip = 0;
                for (; ip + 1 < np; ip += 2) {
                    const PrimPair<R> A1 = pa[ip + 1];
                    both(A0);
                    A0 = pa[ip + 2];
                    both(A1);
                }
                if (ip < np) both(A0);
            } else {
                // FP64: the register cost of the unrolled loop outweighs its savings
                PrimPair<R> A;
                if (np > 0) A = pa[0];
                for (int ip = 0; ip < np; ++ip) {
                    const PrimPair<R> An = pa[ip + 1 < np ? ip + 1 : ip];
                    both(A);
                    A = An;
                }
            }
        }
    }
    for (; iq < nq; ++iq) {
        const PrimPair<R> B = pp[qs + iq];
        Ecoef<LC, LD, R> Fx, Fy, Fz;
        hermite_e<LC, LD>(B.oo2p, B.pax, B.pbx, Fx);
        hermite_e<LC, LD>(B.oo2p, B.pay, B.pby, Fy);
        hermite_e<LC, LD>(B.oo2p, B.paz, B.pbz, Fz);
        // The next bra primitive pair is requested one iteration ahead to hide its latency.
        PrimPair<R> A;
        if (np > 0) A = pp[ps];
        for (int ip = 0; ip < np; ++ip) {
            const PrimPair<R> An = pp[ps + (ip + 1 < np ? ip + 1 : ip)];
            Ecoef<LA, LB, R> Ex, Ey, Ez;
            hermite_e<LA, LB>(A.oo2p, A.pax, A.pbx, Ex);
            hermite_e<LA, LB>(A.oo2p, A.pay, A.pby, Ey);
            hermite_e<LA, LB>(A.oo2p, A.paz, A.pbz, Ez);
            prim(A, Ex, Ey, Ez, B, Fx, Fy, Fz);
            A = An;
        }
    }

    const int sa = a.pair_a[pid], sb = a.pair_b[pid], sc = a.pair_a[qid], sd = a.pair_b[qid];
    double cfa[NA], cfb[NB], cfc[NC], cfd[ND];
    if constexpr (STORED) {
        // the stored integrals are normalized already
        sfor<NA>([&](auto x) { cfa[x] = 1.0; });
        sfor<NB>([&](auto x) { cfb[x] = 1.0; });
        sfor<NC>([&](auto x) { cfc[x] = 1.0; });
        sfor<ND>([&](auto x) { cfd[x] = 1.0; });
        if (valid) {
            const double* __restrict__ src = a.stored + stored_at(qi - a.range[0], 0);
            sfor<CB>([&](auto c_) {
                constexpr int ab = C0 + decltype(c_)::value;
                if constexpr (ab < NAB) {
                    sfor<NCD>([&](auto cd_) {
                        constexpr int cd = decltype(cd_)::value;
                        acc[decltype(c_)::value][cd] = __builtin_nontemporal_load(&src[(ab * NCD + cd) * 64]);
                    });
                }
            });
        }
    } else {
        sfor<NA>([&](auto x) { cfa[x] = a.sh_cf[6 * sa + x]; });
        sfor<NB>([&](auto x) { cfb[x] = a.sh_cf[6 * sb + x]; });
        sfor<NC>([&](auto x) { cfc[x] = a.sh_cf[6 * sc + x]; });
        sfor<ND>([&](auto x) { cfd[x] = a.sh_cf[6 * sd + x]; });
    }

    if (!STORED && a.out) {  // store mode: normalized block, no screening
        if (!valid) return;
        // at the entry's offset, or in the stored-integral layout when no offsets are given
        double* out = a.out_off ? a.out + a.out_off[qi] : a.out + stored_at(qi - a.range[0], 0);
        const int es = a.out_off ? 1 : 64;
        sfor<CB>([&](auto c_) {
            constexpr int ab = C0 + decltype(c_)::value;
            if constexpr (ab < NAB) {
                sfor<NCD>([&](auto cd_) {
                    constexpr int cd = decltype(cd_)::value;
                    out[(ab * NCD + cd) * es] = (double)acc[decltype(c_)::value][cd] * (cfa[ab / NB] * cfb[ab % NB]) *
                                                (cfc[cd / ND] * cfd[cd % ND]);
                });
            }
        });
        return;
    } else {
        const long long base = a.sh_tbase[sa];
        const int oA = a.sh_ao[sa], oB = a.sh_ao[sb], oC = a.sh_ao[sc], oD = a.sh_ao[sd];
        double s = 1.0;
        if (sa == sb) s *= 0.5;
        if (sc == sd) s *= 0.5;
        if (pid == qid) s *= 0.5;
        const double tauS = a.tauS;
        const double* __restrict__ Dt = a.Dt + base;
        const double* __restrict__ Da = a.Da + base;
        const double* __restrict__ Db = a.Db ? a.Db + base : nullptr;

        // Screened, degeneracy-scaled integral of bra component ab (chunk slot c) and ket cd.
        auto val = [&](auto c_, auto cd_) -> double {
            constexpr int c = decltype(c_)::value, cd = decltype(cd_)::value;
            constexpr int ab = C0 + c;
            const double v = (double)acc[c][cd] * (cfa[ab / NB] * cfb[ab % NB]) * (cfc[cd / ND] * cfd[cd % ND]);
            return (fabs(v) < tauS) ? 0.0 : v * s;
        };
        const long long copy = (long long)(xcd_id() % a.n_copies) * a.acc_stride;
        constexpr int fx = FX ? 1 : 0;
        double* AJ = a.AJ + copy + (base << fx);
        {
            double Jab[CB], Jcd[NCD];
            sfor<CB>([&](auto x) { Jab[x] = 0.0; });
            sfor<NCD>([&](auto x) { Jcd[x] = 0.0; });
            sfor<NCD>([&](auto cd_) {
                constexpr int cd = decltype(cd_)::value;
                const long long k = oC + cd / ND, l = oD + cd % ND;
                const double dkl = Dt[tri_index(k, l)];
                sfor<CB>([&](auto c_) {
                    constexpr int c = decltype(c_)::value;
                    constexpr int ab = C0 + c;
                    if constexpr (ab < NAB) {
                        const long long i = oA + ab / NB, j = oB + ab % NB;
                        const double v = val(c_, cd_);
                        Jab[c] += v * dkl;
                        if constexpr (!STORED) Jcd[cd] += v * Dt[tri_index(i, j)];
                    }
                });
            });
            // Lanes of one wavefront that share the bra pair (contiguous runs of the work
            // list) sum their J[a,b] block first; the first lane of each run flushes it.
            const int lane = threadIdx.x & 63;
            const int key = valid ? pid : -1 - lane;
            const int prev = __shfl_up(key, 1);
            const bool head = (lane == 0) || (prev != key);
            const unsigned long long heads = __ballot(head);
            const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
            const int run = __popcll(heads & le);
            bool same[6];
            sfor<6>([&](auto k_) {
                constexpr int off = 1 << decltype(k_)::value;
                const int r = __shfl_down(run, off);
                same[decltype(k_)::value] = (lane + off < 64) && (r == run);
            });
            sfor<CB>([&](auto c_) {
                constexpr int c = decltype(c_)::value;
                constexpr int ab = C0 + c;
                if constexpr (ab < NAB) {
                    double x = Jab[c];
                    sfor<6>([&](auto k_) {
                        constexpr int off = 1 << decltype(k_)::value;
                        const double o = __shfl_down(x, off);
                        if (same[decltype(k_)::value]) x += o;
                    });
                    if (head && valid) acc_add<FX>(AJ, tri_index(oA + ab / NB, oB + ab % NB), 2.0 * x);
                }
            });
            // The stored digest leaves the ket-side Coulomb terms to the ket-major pass (k_jket.hip).
            if constexpr (!STORED) {
                sfor<NCD>([&](auto cd_) {
                    constexpr int cd = decltype(cd_)::value;
                    acc_add<FX>(AJ, tri_index(oC + cd / ND, oD + cd % ND), 2.0 * Jcd[cd]);
                });
            }
        }
        // Exchange with one spin density Dx into the accumulator AKx (rows of a and b into the
        // LDS rows T instead when the block holds this bra pair).
        const bool in_tile = tile != nullptr && pid == tile_pid;
        const int tn = a.tile_n;
        // The exchange terms of the rows of one bra shell (SIDE 0: rows i of a, paired with
        // density row j of b; SIDE 1: rows j of b with density row i of a), with one spin
        // density Dx into the accumulator AKx, or into the LDS rows T when the block holds
        // this bra pair.  The two sides run one after the other, each with half of the
        // accumulators and density loads live, which keeps the digest's register peak (the
        // kernel's allocation) close to that of the primitive loop.
        auto exchange_side = [&](auto side_, const double* __restrict__ Dx, double* AKx, double* T) {
            constexpr int SIDE = decltype(side_)::value;
            constexpr int NR = SIDE == 0 ? NA : NB;  // rows of this side's shell
            double Kc[NR][NC], Kd[NR][ND];
            sfor<NR>([&](auto x) {
                sfor<NC>([&](auto y) { Kc[x][y] = 0.0; });
                sfor<ND>([&](auto y) { Kd[x][y] = 0.0; });
            });
            sfor<NCD>([&](auto cd_) {
                constexpr int cd = decltype(cd_)::value;
                constexpr int jc = cd / ND, jd = cd % ND;
                const long long k = oC + jc, l = oD + jd;
                sfor<CB>([&](auto c_) {
                    constexpr int c = decltype(c_)::value;
                    constexpr int ab = C0 + c;
                    if constexpr (ab < NAB) {
                        constexpr int ia = ab / NB, ib = ab % NB;
                        constexpr int r = SIDE == 0 ? ia : ib;
                        const long long o = SIDE == 0 ? oB + ib : oA + ia;  // the other bra index
                        const double v = val(c_, cd_);
                        Kc[r][jc] += v * Dx[tri_index(o, l)];
                        Kd[r][jd] += v * Dx[tri_index(o, k)];
                    }
                });
            });
            const int orow = SIDE == 0 ? oA : oB;
            sfor<NR>([&](auto r_) {
                constexpr int r = decltype(r_)::value;
                constexpr bool has = SIDE == 0 ? chunk_has_a(C0, CB, r) : chunk_has_b(C0, CB, r);
                if constexpr (has) {
                    if (in_tile) {
                        sfor<NC>([&](auto jc_) {
                            constexpr int jc = decltype(jc_)::value;
                            lds_add(&T[(SIDE * NA + r) * tn + oC + jc], Kc[r][jc]);
                        });
                        sfor<ND>([&](auto jd_) {
                            constexpr int jd = decltype(jd_)::value;
                            lds_add(&T[(SIDE * NA + r) * tn + oD + jd], Kd[r][jd]);
                        });
                    } else {
                        sfor<NC>([&](auto jc_) {
                            constexpr int jc = decltype(jc_)::value;
                            acc_add<FX>(AKx, tri_index(orow + r, oC + jc), Kc[r][jc]);
                        });
                        sfor<ND>([&](auto jd_) {
                            constexpr int jd = decltype(jd_)::value;
                            acc_add<FX>(AKx, tri_index(orow + r, oD + jd), Kd[r][jd]);
                        });
                    }
                }
            });
        };
        auto exchange = [&](const double* __restrict__ Dx, double* AKx, double* T) {
            __builtin_amdgcn_sched_barrier(0);
            exchange_side(md::ic<0>{}, Dx, AKx, T);
            __builtin_amdgcn_sched_barrier(0);
            exchange_side(md::ic<1>{}, Dx, AKx, T);
            __builtin_amdgcn_sched_barrier(0);
        };
        exchange(Da, a.AKa + copy + (base << fx), tile);
        if (Db) exchange(Db, a.AKb + copy + (base << fx), tile ? tile + (NA + NB) * tn : nullptr);
        if (in_tile && valid) {
            unsigned int* mask = (unsigned int*)(tile + (Db ? 2 : 1) * (NA + NB) * tn);
            const int lc = a.sh_local[sc], ld = a.sh_local[sd];
            // most ket shells are already marked: test before the (contended) LDS atomic
            if (!((mask[lc >> 5] >> (lc & 31)) & 1u))
                __hip_atomic_fetch_or(&mask[lc >> 5], 1u << (lc & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (!((mask[ld >> 5] >> (ld & 31)) & 1u))
                __hip_atomic_fetch_or(&mask[ld >> 5], 1u << (ld & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

// Adds the columns of the touched ket shells of the block's LDS exchange rows of bra pair
// pid to the accumulators and clears them and the mask; called by the whole block.
__device__ void flush_rows(const EriArgs& a, double* tile, int pid, int n_tiles) {
    const int sa = a.pair_a[pid], sb = a.pair_b[pid];
    const int tn = a.tile_n;
    const int oA = a.sh_ao[sa], oB = a.sh_ao[sb];
    const int g0 = sa - a.sh_local[sa], ns = a.sh_ns[sa];
    // Never reached in the fixed-point mode (the engine keeps no LDS rows then), so the rows go
    // out as plain FP64 atomics.
    const long long copy = (long long)(xcd_id() % a.n_copies) * a.acc_stride + a.sh_tbase[sa];
    unsigned int* mask = (unsigned int*)(tile + n_tiles * (NA + NB) * tn);
    for (int sh = threadIdx.x; sh < ns; sh += blockDim.x) {
        if (!((mask[sh >> 5] >> (sh & 31)) & 1u)) continue;
        const int o = a.sh_ao[g0 + sh], nc = a.sh_nao[g0 + sh];
        for (int t = 0; t < n_tiles; ++t) {
            double* AK = (t == 0 ? a.AKa : a.AKb) + copy;
            double* T = tile + t * (NA + NB) * tn;
            for (int r = 0; r < NA + NB; ++r) {
                const int row = r < NA ? oA + r : oB + (r - NA);
                for (int c = 0; c < nc; ++c) {
                    const double v = T[r * tn + o + c];
                    T[r * tn + o + c] = 0.0;
                    if (v != 0.0)
                        __hip_atomic_fetch_add(&AK[tri_index(row, o + c)], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
    __syncthreads();
    for (int w = threadIdx.x; w < (ns + 31) / 32; w += blockDim.x) mask[w] = 0u;
}

// The work of one chunk kernel on virtual block vb of nvb (several chunks share a launch).
template <class R, int C0, bool STORED, bool FX>
__device__ __forceinline__ void run_chunk(const EriArgs& a, int vb, int nvb, R* tab, double* tile_mem,
                                          long long& s_chunk) {
    const long long lo = a.range[0], hi = a.range[1];
    const long long n = hi - lo;
    if (a.chunk_ctr == nullptr) {
        const long long first = (long long)vb * blockDim.x;
        if (first >= n) return;  // the whole block has no work: uniform exit before the barrier
        if constexpr (!STORED) load_boys<LTOT, R>(boys_of<R>(a), tab);
        __syncthreads();
        // Grid-stride over the list; the trip count is block-uniform, so every lane reaches
        // the wavefront reductions of the digest.
        const long long stride = (long long)nvb * blockDim.x;
        for (long long base = first; base < n; base += stride) {
            const long long w = base + threadIdx.x;
            const bool valid = w < n;
            quartet<R, C0, STORED, FX>(a, lo + (valid ? w : 0), valid, tab, nullptr, -1);
        }
        return;
    }
    if (n <= 0) return;
    unsigned int* ctr = a.chunk_ctr + C0 / Chunks<R>::cb;
    const int n_tiles = a.tile_n > 0 ? (a.Db ? 2 : 1) : 0;
    double* tile = n_tiles ? tile_mem : nullptr;
    if constexpr (!STORED) load_boys<LTOT, R>(boys_of<R>(a), tab);
    for (int x = threadIdx.x; x < n_tiles * (NA + NB) * a.tile_n; x += blockDim.x) tile_mem[x] = 0.0;
    if (tile)
        for (int w = threadIdx.x; w < (a.tile_shells + 31) / 32; w += blockDim.x)
            ((unsigned int*)(tile_mem + n_tiles * (NA + NB) * a.tile_n))[w] = 0u;
    int tile_pid = -1;
    // Chunks of at most chunk_len entries, shorter for short lists so that every block claims
    // about four (the lists are sorted heaviest first: several claims per block balance them).
    long long L = (n + 4LL * nvb - 1) / (4LL * nvb);
    L = ((L + blockDim.x - 1) / blockDim.x) * blockDim.x;
    if (L > a.chunk_len) L = a.chunk_len;
    for (;;) {
        __syncthreads();  // every thread has read s_chunk of the previous round
        if (threadIdx.x == 0) s_chunk = (long long)atomicAdd(ctr, 1u) * L;
        __syncthreads();
        const long long c0 = s_chunk;
        if (c0 >= n) break;
        const long long c1 = c0 + L < n ? c0 + L : n;
        const long long nit = (c1 - c0 + blockDim.x - 1) / blockDim.x;
        const long long rot = a.stagger ? (long long)((blockIdx.x * 2654435761u) >> 8) % nit : 0;
        for (long long it = 0; it < nit; ++it) {
            const long long base = c0 + (it + rot < nit ? it + rot : it + rot - nit) * blockDim.x;
            if (tile) {
                // The rows follow the bra pair of the iteration's first entry once it holds
                // at least a quarter of the iteration.  pid0 and tile_pid are the same in every
                // thread, so while the rows already hold pid0 the block skips the count and its
                // barrier (the rows only take atomic adds until the next flush).
                const int pid0 = (int)a.list[lo + base].x;
                if (pid0 != tile_pid) {
                    const long long w = base + threadIdx.x;
                    const int mine = (w < c1 && (int)a.list[lo + w].x == pid0) ? 1 : 0;
                    const int cnt = __syncthreads_count(mine);
                    if (cnt >= (int)blockDim.x / 4) {
                        if (tile_pid >= 0) flush_rows(a, tile, tile_pid, n_tiles);
                        __syncthreads();
                        tile_pid = pid0;
                    }
                }
            }
            const long long w = base + threadIdx.x;
            const bool valid = w < c1;
            quartet<R, C0, STORED, FX>(a, lo + (valid ? w : c0), valid, tab, tile, tile_pid);
        }
    }
    if (tile && tile_pid >= 0) {
        __syncthreads();
        flush_rows(a, tile, tile_pid, n_tiles);
    }
}

// Chunks [CH0, CH0 + NC) of the class in one launch: block b works on chunk b % NC as
// virtual block b / NC, so the chunks are interleaved over the machine and run side by side.
// The class index is a template argument only so that profiles name each class's kernels.
// STORED: the FP64 stored-integral digest (no Boys table in LDS).
template <int KC, class R, int CH0, int NC, bool STORED, bool FX>
__global__ __launch_bounds__(256) void k_eri(EriArgs a) {
    using S = BoysSlice<LTOT, R>;
    __shared__ __align__(16) R tab[STORED ? 2 : BoysGrid<R>::points * S::nord];
    __shared__ long long s_chunk;
    extern __shared__ double tile_mem[];
    const int ch = (int)(blockIdx.x % NC), vb = (int)(blockIdx.x / NC), nvb = (int)(gridDim.x / NC);
    sfor<NC>([&](auto c_) {
        constexpr int c = decltype(c_)::value;
        if (ch == c) run_chunk<R, (CH0 + c) * Chunks<R>::cb, STORED, FX>(a, vb, nvb, tab, tile_mem, s_chunk);
    });
}

#ifndef HFX_CHUNK_BEGIN
#define HFX_CHUNK_BEGIN 0
#endif
#ifndef HFX_CHUNK_END
#define HFX_CHUNK_END 1000
#endif
constexpr int CH_BEGIN = HFX_CHUNK_BEGIN;
// Chunks [CH_BEGIN, CH_END) of this translation unit in the chunking of precision R.
template <class R>
constexpr int ch_end() {
    return HFX_CHUNK_END < Chunks<R>::nch ? HFX_CHUNK_END : Chunks<R>::nch;
}

template <class R>
void launch_part(int fp32, const EriArgs& a, int grid, const hipStream_t* cs) {
    (void)fp32;
    constexpr int NC = ch_end<R>() > CH_BEGIN ? ch_end<R>() - CH_BEGIN : 0;
    if constexpr (NC > 0) {
        const size_t lds = a.tile_n > 0 ? (size_t)(a.Db ? 2 : 1) * (NA + NB) * a.tile_n * sizeof(double) +
                                              (size_t)((a.tile_shells + 31) / 32) * sizeof(unsigned int)
                                        : 0;
        const int blocks = (grid > NC ? grid / NC : 1) * NC;
        if constexpr (sizeof(R) == 8) {
            if (a.stored) {
                if (a.fixed)
                    hipLaunchKernelGGL((k_eri<K_COMBO, R, CH_BEGIN, NC, true, true>), dim3(blocks), dim3(256), lds,
                                       cs[CH_BEGIN], a);
                else
                    hipLaunchKernelGGL((k_eri<K_COMBO, R, CH_BEGIN, NC, true, false>), dim3(blocks), dim3(256), lds,
                                       cs[CH_BEGIN], a);
                return;
            }
        }
        if (a.fixed)
            hipLaunchKernelGGL((k_eri<K_COMBO, R, CH_BEGIN, NC, false, true>), dim3(blocks), dim3(256), lds,
                               cs[CH_BEGIN], a);
        else
            hipLaunchKernelGGL((k_eri<K_COMBO, R, CH_BEGIN, NC, false, false>), dim3(blocks), dim3(256), lds,
                               cs[CH_BEGIN], a);
    }
}

const bool registered64 =
    register_eri_part(K_COMBO, 0, CH_BEGIN, ch_end<double>(), Chunks<double>::nch, &launch_part<double>);
const bool registered32 = ch_end<float>() > CH_BEGIN
                              ? register_eri_part(K_COMBO, 1, CH_BEGIN, ch_end<float>(), Chunks<float>::nch,
                                                  &launch_part<float>)
                              : false;

}  // namespace
}  // namespace hfx
