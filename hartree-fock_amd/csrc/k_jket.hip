// k_jket.hip - ket-side Coulomb terms of the stored integrals, from their ket-major copy.
//
// Compiled once per quartet class with -DHFX_COMBO=k (k as in k_eri.hip).
//
// The stored digest (k_eri<..., STORED = true>) adds the bra-side Coulomb terms J[a,b] (reduced
// over the lanes that share the bra) and the exchange terms (rows of the bra's shells, in LDS).
// The remaining terms
//   AJ[k,l] += 2 s sum_ab I(ab|kl) D[i,j]
// (the second half of the reference's Coulomb digestion, src/SCF.cpp:677-681) would be one
// global atomic per ket component per quartet in a bra-major list.  Here the list is ket-major
// (entries {ket pair, bra pair}, sorted by ket pair; hfx_store_eris builds it with the stored
// blocks transposed: element cd*NAB + ab of entry r at ((r/64) NAB NCD + cd NAB + ab) 64 + r mod 64),
// so the 64 lanes of a wavefront read each element with one 512-byte access and mostly share
// the ket pair: they sum over their run of equal ket pairs and the run's first lane adds the
// totals.  The screening of single integrals (|I| < tau_S dropped, src/SCF.cpp:670-671) and
// the degeneracy s are those of the main digest.
#include "digest.hpp"
#include "engine.hpp"
#include "md.hpp"

#ifndef HFX_COMBO
#error "HFX_COMBO must be defined"
#endif

namespace hfx {
namespace {

using namespace md;
using namespace digest;

constexpr int K_COMBO = HFX_COMBO;
constexpr int PC1 = combo_pc1(K_COMBO), PC2 = combo_pc2(K_COMBO);
constexpr int LA = pc_la(PC1), LB = pc_lb(PC1), LC = pc_la(PC2), LD = pc_lb(PC2);
constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
constexpr int NAB = NA * NB, NCD = NC * ND;
// Bra components whose loads one thread issues together (every load of the group in flight
// before the sums that use them).
constexpr int AB_GROUP = NCD <= 3 ? 12 : (NCD <= 9 ? 4 : 1);

template <bool FX>
__global__ __launch_bounds__(256) void k_jket(EriArgs a) {
    const long long lo = a.range[0], n = a.range[1] - lo;
    const long long copy = (long long)(xcd_id() % a.n_copies) * a.acc_stride;
    constexpr int fx = FX ? 1 : 0;
    const int lane = threadIdx.x & 63;
    const double tauS = a.tauS;
    // the trip count is the same for the whole block, so every lane reaches the reductions
    for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < n; b0 += (long long)gridDim.x * blockDim.x) {
        const long long r = b0 + threadIdx.x;
        const bool valid = r < n;
        const uint2 e = a.list[lo + (valid ? r : 0)];
        const int qid = (int)e.x, pid = (int)e.y;
        const int sa = a.pair_a[pid], sb = a.pair_b[pid], sc = a.pair_a[qid], sd = a.pair_b[qid];
        double s = 1.0;
        if (sa == sb) s *= 0.5;
        if (sc == sd) s *= 0.5;
        if (pid == qid) s *= 0.5;
        const long long base = a.sh_tbase[sa];
        const double* __restrict__ Dt = a.Dt + base;
        const int oA = a.sh_ao[sa], oB = a.sh_ao[sb], oC = a.sh_ao[sc], oD = a.sh_ao[sd];
        const double* __restrict__ src =
            a.stored + ((valid ? r : 0) >> 6) * (long long)(NAB * NCD) * 64 + ((valid ? r : 0) & 63);
        double Jcd[NCD];
        sfor<NCD>([&](auto x) { Jcd[x] = 0.0; });
        if (valid) {
            for (int g = 0; g < NAB; g += AB_GROUP) {
                double v[AB_GROUP][NCD], d[AB_GROUP];
                sfor<AB_GROUP>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    const int ab = g + j;
                    if (ab < NAB) {
                        d[j] = Dt[tri_index(oA + ab / NB, oB + ab % NB)];
                        sfor<NCD>([&](auto cd_) {
                            constexpr int cd = decltype(cd_)::value;
                            v[j][cd] = __builtin_nontemporal_load(&src[(cd * NAB + ab) * 64]);
                        });
                    } else {
                        d[j] = 0.0;
                        sfor<NCD>([&](auto cd_) { v[j][decltype(cd_)::value] = 0.0; });
                    }
                });
                sfor<AB_GROUP>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    sfor<NCD>([&](auto cd_) {
                        constexpr int cd = decltype(cd_)::value;
                        const double x = v[j][cd];
                        Jcd[cd] += (fabs(x) < tauS ? 0.0 : x * s) * d[j];
                    });
                });
            }
        }
        bool head, same[6];
        run_flags(valid ? qid : -1 - lane, head, same);
        double* AJ = a.AJ + copy + (base << fx);
        sfor<NCD>([&](auto cd_) {
            constexpr int cd = decltype(cd_)::value;
            const double x = run_sum(Jcd[cd], same);
            if (head && valid) acc_add<FX>(AJ, tri_index(oC + cd / ND, oD + cd % ND), 2.0 * x);
        });
    }
}

void launch(const EriArgs& a, int grid, hipStream_t s) {
    if (a.fixed)
        hipLaunchKernelGGL(k_jket<true>, dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_jket<false>, dim3(grid), dim3(256), 0, s, a);
}

const bool registered = register_jket(K_COMBO, &launch);

}  // namespace
}  // namespace hfx
