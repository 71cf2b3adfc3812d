// k_grad_eri.hip - two-electron part of the RHF nuclear gradient for one quartet class, every
// angular loop unrolled at compile time (the pattern of the J/K kernels of k_eri.hip).
//
// Compiled per class with -DHFX_GRAD_COMBO=k (k as in k_eri.hip).  Up to la+lb+lc+ld = 3 one
// thread takes a whole quartet.  Above that the whole quartet spills to scratch, so the bra
// components are split into chunks of HFX_GRAD_CB: each thread takes one (quartet, chunk) and
// repeats the Boys values, R and the E tables for its chunk.  A translation unit holds the
// chunks HFX_GRAD_CHUNK_BEGIN .. + HFX_GRAD_CHUNKS - 1 of its class, one kernel each, and
// registers each as a launch unit of the class.  The classes not compiled here use the generic
// kernel of k_grad.hip.  The chunked classes are compiled in FP64 only
// (-DHFX_GRAD_FP64_ONLY): a mixed-precision gradient runs their whole list, the quartets below
// the cut included, through the FP64 kernels in one pass (a second, FP32 pass over the short
// tail of each chunk took longer than the FP64 work it replaced), and the compile time halves.
//
// One thread per screened quartet (and chunk) of the J/K work list built from (P, P/2).  The
// thread writes the part of the quartet's
//   Gamma_{ab,cd} = 1/2 deg [P_ij P_kl - 1/4 (P_ik P_jl + P_il P_jk)] N_a N_b N_c N_d
// that belongs to its chunk into its own LDS slots, then walks the primitive quartets: F_0..F_{L+1} (L = la+lb+lc+ld),
// R^0_tuv for t+u+v <= L+1 scaled by the prefactor, E of the bra up to (la+1, lb) and of the
// ket up to (lc+1, ld).  For every component pair (ab, cd) it forms
//   Y[t,u,v]   = sum_{t'u'v'} (-1)^{t'+u'+v'} E^{cd}_{t'u'v'} R_{t+t',u+u',v+v'}
//   Y_C[t,u,v] = the same with the ket's C_x (C_y, C_z) derivative coefficients
// over the bra's Hermite box (extended by one in one direction where needed) and contracts
//   dA_x:  sum (2a E^{i+1,j}_t - i E^{i-1,j}_t) E_u E_v Y[t,u,v]       (centre A)
//   dP_x:  sum E_t E_u E_v Y[t+1,u,v]                                   (A and B together:
//          moving both centres shifts the product's Hermite Gaussians, Lambda_t -> Lambda_{t+1})
//   dC_x:  sum E_t E_u E_v Y_C_x[t,u,v]
// weighted by Gamma.  Then dE/dB = dP - dA and, by translational invariance of the integral,
// dE/dD = -dP - dC.  No reference counterpart (the reference computes energies only).
#include "engine.hpp"
#include "md.hpp"

#ifndef HFX_GRAD_COMBO
#error "HFX_GRAD_COMBO must be defined"
#endif

namespace hfx {
namespace {

using namespace md;

constexpr int K_COMBO = HFX_GRAD_COMBO;
constexpr int PC1 = combo_pc1(K_COMBO), PC2 = combo_pc2(K_COMBO);
constexpr int LA = pc_la(PC1), LB = pc_lb(PC1), LC = pc_la(PC2), LD = pc_lb(PC2);
constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
constexpr int NAB = NA * NB, NCD = NC * ND;
constexpr int LT = LA + LB + LC + LD + 1;  // highest Boys order and R level
#ifndef HFX_GRAD_CB
#define HFX_GRAD_CB NAB
#endif
#ifndef HFX_GRAD_CHUNK_BEGIN
#define HFX_GRAD_CHUNK_BEGIN 0
#endif
#ifndef HFX_GRAD_CHUNKS
#define HFX_GRAD_CHUNKS 1
#endif
constexpr int CB = HFX_GRAD_CB;  // bra components per thread
constexpr int CH_BEGIN = HFX_GRAD_CHUNK_BEGIN, CH_COUNT = HFX_GRAD_CHUNKS;
static_assert(CB >= 1 && CB <= NAB && CH_COUNT >= 1 && (CH_BEGIN + CH_COUNT - 1) * CB < NAB,
              "chunks of bra components outside the class");
// Threads per block: the Gamma slots of every thread fit in 40 KB of LDS.
constexpr int block_threads(int bytes) { return bytes * 256 <= 40960 ? 256 : (bytes * 128 <= 40960 ? 128 : 64); }
constexpr int GRAD_LDS_MAX = 4096;  // gradient entries (3 per atom) a block keeps in LDS

template <class R>
struct Lists;
template <>
struct Lists<double> {
    __device__ static const int* ps(const GradArgs& g) { return g.pair_ps; }
    __device__ static const int* np(const GradArgs& g) { return g.pair_np; }
    __device__ static const PrimPair64* pp(const GradArgs& g) { return g.pp64; }
    __device__ static const double* ab(const GradArgs& g) { return g.pab64; }
    __device__ static const double* boys(const GradArgs& g) { return g.boys64; }
};
template <>
struct Lists<float> {
    __device__ static const int* ps(const GradArgs& g) { return g.pair_ps32; }
    __device__ static const int* np(const GradArgs& g) { return g.pair_np32; }
    __device__ static const PrimPair32* pp(const GradArgs& g) { return g.pp32; }
    __device__ static const float* ab(const GradArgs& g) { return g.pab32; }
    __device__ static const float* boys(const GradArgs& g) { return g.boys32; }
};

// E^{I,J}_T of one direction (zero outside T <= I + J).
template <int I, int J, int T, int EA, int EB, class R>
__device__ __forceinline__ R ec(const Ecoef<EA, EB, R>& E) {
    if constexpr (I >= 0 && J >= 0 && I <= EA && J <= EB && T <= I + J)
        return E.e[I][J][T];
    else
        return R(0);
}
// Derivative coefficient of the first centre: 2a E^{I+1,J}_T - I E^{I-1,J}_T.
template <int I, int J, int T, int EA, int EB, class R>
__device__ __forceinline__ R ed(const Ecoef<EA, EB, R>& E, R twoa) {
    R v = twoa * ec<I + 1, J, T>(E);
    if constexpr (I >= 1 && T <= I + J - 1) v -= R(I) * E.e[I - 1][J][T];
    return v;
}

// One chunk: bra components C0 .. C0 + CB - 1 (those below NAB) of every quartet of the list,
// one quartet per thread; results into dst (the block's LDS gradient row, or its row of g.part).
template <class R, int C0>
__device__ __forceinline__ void grad_chunk(const GradArgs& g, R* gam, double* dst) {
    constexpr int BT = block_threads(CB * NCD * (int)sizeof(R));
    using Ls = Lists<R>;
    R* gm = gam + threadIdx.x;  // this thread's slots, BT apart
    const long long lo = g.range[0], hi = g.range[1];
    const int N = g.n_ao;
    const auto* __restrict__ pp = Ls::pp(g);
    const R* __restrict__ pab = Ls::ab(g);
    const R* __restrict__ btab = Ls::boys(g) + BoysSlice<LT, R>::first;
    for (long long q = lo + (long long)blockIdx.x * BT + threadIdx.x; q < hi; q += (long long)gridDim.x * BT) {
        const uint2 ent = g.list[q];
        const int p1 = (int)ent.x, p2 = (int)ent.y;
        const int sa = g.pair_a[p1], sb = g.pair_b[p1], sc = g.pair_a[p2], sd = g.pair_b[p2];
        const double deg = (sa != sb ? 2.0 : 1.0) * (sc != sd ? 2.0 : 1.0) * (p1 != p2 ? 2.0 : 1.0);
        {
            const int oA = g.sh_ao[sa], oB = g.sh_ao[sb], oC = g.sh_ao[sc], oD = g.sh_ao[sd];
            sfor<CB>([&](auto c_) {
                constexpr int c = decltype(c_)::value;
                constexpr int ab = C0 + c;
                if constexpr (ab < NAB) {
                constexpr int A = ab / NB, B = ab % NB;
                const size_t i = oA + A, j = oB + B;
                const double cab = 0.5 * deg * g.sh_cf[6 * sa + A] * g.sh_cf[6 * sb + B];
                const double pij = g.P[i + j * N];
                sfor<NCD>([&](auto cd_) {
                    constexpr int cd = decltype(cd_)::value;
                    constexpr int C = cd / ND, D = cd % ND;
                    const size_t k = oC + C, l = oD + D;
                    const double gmm =
                        pij * g.P[k + l * N] - 0.25 * (g.P[i + k * N] * g.P[j + l * N] + g.P[i + l * N] * g.P[j + k * N]);
                    gm[(c * NCD + cd) * BT] = (R)(gmm * cab * g.sh_cf[6 * sc + C] * g.sh_cf[6 * sd + D]);
                });
                }
            });
        }
        double gA[3] = {0, 0, 0}, gP[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
        const int ps1 = Ls::ps(g)[p1], np1 = Ls::np(g)[p1], ps2 = Ls::ps(g)[p2], np2 = Ls::np(g)[p2];
        for (int x1 = 0; x1 < np1; ++x1) {
            const PrimPair<R> pb = pp[ps1 + x1];
            const R twoa = R(2) * pab[2 * (ps1 + x1)];
            Ecoef<LA + 1, LB, R> Ex, Ey, Ez;
            hermite_e<LA + 1, LB>(pb.oo2p, pb.pax, pb.pbx, Ex);
            hermite_e<LA + 1, LB>(pb.oo2p, pb.pay, pb.pby, Ey);
            hermite_e<LA + 1, LB>(pb.oo2p, pb.paz, pb.pbz, Ez);
            R sA[3] = {R(0), R(0), R(0)}, sP[3] = {R(0), R(0), R(0)}, sC[3] = {R(0), R(0), R(0)};
            for (int x2 = 0; x2 < np2; ++x2) {
                const PrimPair<R> pk = pp[ps2 + x2];
                const R twoc = R(2) * pab[2 * (ps2 + x2)];
                Ecoef<LC + 1, LD, R> Fx, Fy, Fz;
                hermite_e<LC + 1, LD>(pk.oo2p, pk.pax, pk.pbx, Fx);
                hermite_e<LC + 1, LD>(pk.oo2p, pk.pay, pk.pby, Fy);
                hermite_e<LC + 1, LD>(pk.oo2p, pk.paz, pk.pbz, Fz);
                const R pq = pb.p + pk.p;
                const R rs = dev_rsqrt(pq);
                const R alpha = (pb.p * pk.p) * (rs * rs);
                const R X = pb.px - pk.px, Y = pb.py - pk.py, Z = pb.pz - pk.pz;
                const R T = alpha * (X * X + Y * Y + Z * Z);
                const R pref = R(34.98683665524972) * rs * pb.kp * pk.kp;  // 2 pi^(5/2)
                R F[LT + 1];
                boys<LT, R, BOYS_ORDERS>(T, btab, F);
                R base[LT + 1];
                {
                    const R m2a = R(-2) * alpha;
                    R pw = pref;
                    sfor<LT + 1>([&](auto n_) {
                        constexpr int n = decltype(n_)::value;
                        base[n] = pw * F[n];
                        if constexpr (n < LT) pw *= m2a;
                    });
                }
                R Rt[nherm(LT)];
                hermite_r<LT>(base, X, Y, Z, Rt);
                sfor<CB>([&](auto c_) {
                    constexpr int c = decltype(c_)::value;
                    constexpr int ab = C0 + c;
                    if constexpr (ab < NAB) {
                    constexpr int A = ab / NB, B = ab % NB;
                    constexpr int ax = cart_i(LA, A), ay = cart_j(LA, A), az = cart_k(LA, A);
                    constexpr int bx = cart_i(LB, B), by = cart_j(LB, B), bz = cart_k(LB, B);
                    constexpr int lx = ax + bx, ly = ay + by, lz = az + bz;
                    sfor<NCD>([&](auto cd_) {
                        constexpr int cd = decltype(cd_)::value;
                        constexpr int C = cd / ND, D = cd % ND;
                        constexpr int cx = cart_i(LC, C), cy = cart_j(LC, C), cz = cart_k(LC, C);
                        constexpr int dx = cart_i(LD, D), dy = cart_j(LD, D), dz = cart_k(LD, D);
                        constexpr int kx = cx + dx, ky = cy + dy, kz = cz + dz;
                        const R w = gm[(c * NCD + cd) * BT];
                        // Y over the bra box extended by one in at most one direction
                        R Yv[lx + 2][ly + 2][lz + 2];
                        sfor<lz + 2>([&](auto v_) {
                            constexpr int v = decltype(v_)::value;
                            sfor<ly + 2>([&](auto u_) {
                                constexpr int u = decltype(u_)::value;
                                sfor<lx + 2>([&](auto t_) {
                                    constexpr int t = decltype(t_)::value;
                                    if constexpr ((t > lx) + (u > ly) + (v > lz) <= 1) {
                                        Yv[t][u][v] = sum_of<kz + 1>([&](auto vv_) {
                                            constexpr int vv = decltype(vv_)::value;
                                            return ec<cz, dz, vv>(Fz) * sum_of<ky + 1>([&](auto uu_) {
                                                       constexpr int uu = decltype(uu_)::value;
                                                       return ec<cy, dy, uu>(Fy) * sum_of<kx + 1>([&](auto tt_) {
                                                                  constexpr int tt = decltype(tt_)::value;
                                                                  constexpr R sg = ((tt + uu + vv) & 1) ? R(-1) : R(1);
                                                                  return (sg * ec<cx, dx, tt>(Fx)) *
                                                                         Rt[hidx(t + tt, u + uu, v + vv)];
                                                              });
                                                   });
                                        });
                                    }
                                });
                            });
                        });
                        // centre A and the joint translation P of the bra
                        R iA[3], iP[3];
                        iA[0] = sum_of<lz + 1>([&](auto v_) {
                            constexpr int v = decltype(v_)::value;
                            return ec<az, bz, v>(Ez) * sum_of<ly + 1>([&](auto u_) {
                                       constexpr int u = decltype(u_)::value;
                                       return ec<ay, by, u>(Ey) * sum_of<lx + 2>([&](auto t_) {
                                                  constexpr int t = decltype(t_)::value;
                                                  return ed<ax, bx, t>(Ex, twoa) * Yv[t][u][v];
                                              });
                                   });
                        });
                        iA[1] = sum_of<lz + 1>([&](auto v_) {
                            constexpr int v = decltype(v_)::value;
                            return ec<az, bz, v>(Ez) * sum_of<ly + 2>([&](auto u_) {
                                       constexpr int u = decltype(u_)::value;
                                       return ed<ay, by, u>(Ey, twoa) * sum_of<lx + 1>([&](auto t_) {
                                                  constexpr int t = decltype(t_)::value;
                                                  return ec<ax, bx, t>(Ex) * Yv[t][u][v];
                                              });
                                   });
                        });
                        iA[2] = sum_of<lz + 2>([&](auto v_) {
                            constexpr int v = decltype(v_)::value;
                            return ed<az, bz, v>(Ez, twoa) * sum_of<ly + 1>([&](auto u_) {
                                       constexpr int u = decltype(u_)::value;
                                       return ec<ay, by, u>(Ey) * sum_of<lx + 1>([&](auto t_) {
                                                  constexpr int t = decltype(t_)::value;
                                                  return ec<ax, bx, t>(Ex) * Yv[t][u][v];
                                              });
                                   });
                        });
                        sfor<3>([&](auto d_) {
                            constexpr int d = decltype(d_)::value;
                            iP[d] = sum_of<lz + 1>([&](auto v_) {
                                constexpr int v = decltype(v_)::value;
                                return ec<az, bz, v>(Ez) * sum_of<ly + 1>([&](auto u_) {
                                           constexpr int u = decltype(u_)::value;
                                           return ec<ay, by, u>(Ey) * sum_of<lx + 1>([&](auto t_) {
                                                      constexpr int t = decltype(t_)::value;
                                                      return ec<ax, bx, t>(Ex) *
                                                             Yv[t + (d == 0)][u + (d == 1)][v + (d == 2)];
                                                  });
                                       });
                            });
                        });
                        // centre C: the ket's derivative coefficients over the bra box
                        R iC[3];
                        sfor<3>([&](auto d_) {
                            constexpr int d = decltype(d_)::value;
                            iC[d] = sum_of<lz + 1>([&](auto v_) {
                                constexpr int v = decltype(v_)::value;
                                return ec<az, bz, v>(Ez) * sum_of<ly + 1>([&](auto u_) {
                                           constexpr int u = decltype(u_)::value;
                                           return ec<ay, by, u>(Ey) * sum_of<lx + 1>([&](auto t_) {
                                                      constexpr int t = decltype(t_)::value;
                                                      return ec<ax, bx, t>(Ex) * sum_of<kz + 1 + (d == 2)>([&](auto vv_) {
                                                                 constexpr int vv = decltype(vv_)::value;
                                                                 const R fz = d == 2 ? ed<cz, dz, vv>(Fz, twoc)
                                                                                     : ec<cz, dz, vv>(Fz);
                                                                 return fz * sum_of<ky + 1 + (d == 1)>([&](auto uu_) {
                                                                            constexpr int uu = decltype(uu_)::value;
                                                                            const R fy = d == 1 ? ed<cy, dy, uu>(Fy, twoc)
                                                                                                : ec<cy, dy, uu>(Fy);
                                                                            return fy * sum_of<kx + 1 + (d == 0)>([&](auto tt_) {
                                                                                       constexpr int tt = decltype(tt_)::value;
                                                                                       constexpr R sg = ((tt + uu + vv) & 1) ? R(-1) : R(1);
                                                                                       const R fx = d == 0 ? ed<cx, dx, tt>(Fx, twoc)
                                                                                                           : ec<cx, dx, tt>(Fx);
                                                                                       return (sg * fx) * Rt[hidx(t + tt, u + uu, v + vv)];
                                                                                   });
                                                                        });
                                                             });
                                                  });
                                       });
                            });
                        });
                        sfor<3>([&](auto d_) {
                            constexpr int d = decltype(d_)::value;
                            sA[d] += w * iA[d];
                            sP[d] += w * iP[d];
                            sC[d] += w * iC[d];
                        });
                        // one component pair at a time: interleaving them only raises the
                        // register peak
                        __builtin_amdgcn_sched_barrier(0);
                    });
                    }
                });
            }
            sfor<3>([&](auto d_) {
                constexpr int d = decltype(d_)::value;
                gA[d] += (double)sA[d];
                gP[d] += (double)sP[d];
                gC[d] += (double)sC[d];
            });
        }
        const int aa = g.sh_atom[sa], ab = g.sh_atom[sb], ac = g.sh_atom[sc], ad = g.sh_atom[sd];
        for (int d = 0; d < 3; ++d) {
            atomicAdd(&dst[3 * aa + d], gA[d]);
            atomicAdd(&dst[3 * ab + d], gP[d] - gA[d]);
            atomicAdd(&dst[3 * ac + d], gC[d]);
            atomicAdd(&dst[3 * ad + d], -(gP[d] + gC[d]));
        }
    }
}

// One chunk per kernel (C0: its first bra component); KC (the class) names the kernel in
// profiles.  The chunks of a class are separate launch units that the caller spreads over its
// streams, so blocks of different kernels share the rows of g.part: the LDS rows are added with
// atomics.  (One kernel for a unit's chunks, blockIdx.y picking the chunk, ran the same but
// took half as long again to compile: the compiler allocates registers over all the bodies.)
template <class R, int KC, int C0>
__global__ __launch_bounds__(block_threads(CB * NCD * (int)sizeof(R))) void k_grad_eri_t(GradArgs g) {
    constexpr int BT = block_threads(CB * NCD * (int)sizeof(R));
    __shared__ R gam[CB * NCD * BT];
    extern __shared__ double sgrad[];
    const int n3 = g.n3;
    const bool lds_rows = n3 <= GRAD_LDS_MAX;
    if (lds_rows)
        for (int x = threadIdx.x; x < n3; x += BT) sgrad[x] = 0.0;
    __syncthreads();
    double* row = g.part + (size_t)blockIdx.x * g.part_stride + 2 * n3;
    grad_chunk<R, C0>(g, gam, lds_rows ? sgrad : row);
    if (lds_rows) {
        __syncthreads();
        for (int x = threadIdx.x; x < n3; x += BT)
            if (sgrad[x] != 0.0) atomicAdd(&row[x], sgrad[x]);
    }
}

// Chunk I of this unit (chunk CH_BEGIN + I of the class).
template <int I>
void launch(int fp32, const GradArgs& g, int grid, hipStream_t s) {
    constexpr int c0 = (CH_BEGIN + I) * CB;
    const size_t lds = g.n3 <= GRAD_LDS_MAX ? (size_t)g.n3 * sizeof(double) : 0;
#ifdef HFX_GRAD_FP64_ONLY
    (void)fp32;
#else
    if (fp32)
        hipLaunchKernelGGL((k_grad_eri_t<float, K_COMBO, c0>), dim3(grid), dim3(block_threads(CB * NCD * 4)), lds, s, g);
    else
#endif
        hipLaunchKernelGGL((k_grad_eri_t<double, K_COMBO, c0>), dim3(grid), dim3(block_threads(CB * NCD * 8)), lds, s,
                           g);
}

#ifdef HFX_GRAD_FP64_ONLY
constexpr bool HAS_FP32 = false;
#else
constexpr bool HAS_FP32 = true;
#endif
template <int I>
bool register_chunks() {
    if constexpr (I < CH_COUNT)
        return register_grad_class(K_COMBO, CH_BEGIN + I, &launch<I>, HAS_FP32) && register_chunks<I + 1>();
    else
        return true;
}
const bool registered = register_chunks<0>();

}  // namespace
}  // namespace hfx
