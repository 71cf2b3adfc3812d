// engine.hip - the MI355X direct-SCF Fock-build engine behind include/hfx.h.
//
// Host side: the flat basis (IntegralEngine::IntegralEngine, src/IntegralEngine.cpp:13-75),
// shell pairs binned by angular-momentum class, primitive-pair data (computePrimitivePairData,
// src/IntegralEngine.cpp:995-1022, precomputed once per geometry), the Boys table, and the
// per-build pipeline.  Device side (this file): shell-pair density maxima, the screening
// kernel that writes the per-class work lists, the symmetrization of J/K and the
// one-electron integrals.  The ERI + J/K kernels live in k_eri.hip.
//
// One J/K build (src/SCF.cpp:523-744) on the engine's stream:
//   k_dmax        shell-pair maxima of |dD| (SCF.cpp:543-575)
//   k_screen<0>   per class: count, for every bra pair, the canonical quartets passing
//                 Q(a,b)Q(c,d) >= tau_S and the density test (SCF.cpp:602-623)
//   scan          exclusive prefix sum of the counts (hipcub)
//   k_screen<1>   write the survivors at their scanned positions (FP64 and FP32 parts)
//   k_ranges      list ranges from the scanned offsets
//   k_eri<...>    contracted ERIs + digestion into AJ/AK (k_eri.hip)
//   k_symm        J = AJ + AJ^T, K = AK + AK^T (the reference's selfadjointView, SCF.cpp:727-728)
// No host synchronization inside the pipeline: kernels read list sizes from device memory.
#include "../../include/hfx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "digest.hpp"
#include "engine.hpp"
#include "md.hpp"
#include "scf_device.hpp"

using namespace hfx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t err_ = (x);                                                             \
        if (err_ != hipSuccess)                                                            \
            return fail(HFX_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(err_)); \
    } while (0)

// ---------------------------------------------------------------------------------------
// Registry of ERI class kernels (filled by the k_eri translation units at load time).
// ---------------------------------------------------------------------------------------
struct PartTable {
    EriPartFn fn[2][N_COMBOS][64];
    int begin[2][N_COMBOS][64];
    int end[2][N_COMBOS][64];
    int nparts[2][N_COMBOS];
    int nch[2][N_COMBOS];
};
PartTable& parts() {
    static PartTable t;  // zero-initialized before any dynamic initializer runs
    return t;
}
JketFn& jket_fn(int combo) {
    static JketFn t[N_COMBOS];  // zero-initialized likewise
    return t[combo];
}

}  // namespace

namespace hfx {

bool register_jket(int combo, JketFn fn) {
    if (combo < 0 || combo >= N_COMBOS) return false;
    jket_fn(combo) = fn;
    return true;
}

bool register_eri_part(int combo, int fp32, int begin, int end, int nch, EriPartFn fn) {
    PartTable& t = parts();
    if (combo < 0 || combo >= N_COMBOS || fp32 < 0 || fp32 > 1 || t.nparts[fp32][combo] >= 64) return false;
    const int k = t.nparts[fp32][combo]++;
    t.fn[fp32][combo][k] = fn;
    t.begin[fp32][combo][k] = begin;
    t.end[fp32][combo][k] = end;
    t.nch[fp32][combo] = nch;
    return true;
}

int eri_chunks(int combo, int fp32) {
    PartTable& t = parts();
    const int n = t.nch[fp32][combo];
    std::vector<int> cover(n > 0 ? n : 0, 0);
    for (int k = 0; k < t.nparts[fp32][combo]; ++k)
        for (int c = t.begin[fp32][combo][k]; c < t.end[fp32][combo][k] && c < n; ++c) cover[c]++;
    for (int c = 0; c < n; ++c)
        if (cover[c] != 1) return -1;
    return n > 0 ? n : -1;
}

int eri_parts(int combo, int fp32, int* begin, int* end) {
    PartTable& t = parts();
    for (int k = 0; k < t.nparts[fp32][combo]; ++k) {
        begin[k] = t.begin[fp32][combo][k];
        end[k] = std::min(t.end[fp32][combo][k], t.nch[fp32][combo]);
    }
    return t.nparts[fp32][combo];
}

void launch_eri(int combo, int fp32, const EriArgs& a, int grid, const hipStream_t* chunk_streams) {
    PartTable& t = parts();
    for (int k = 0; k < t.nparts[fp32][combo]; ++k) t.fn[fp32][combo][k](fp32, a, grid, chunk_streams);
}

}  // namespace hfx

namespace {

constexpr int CMAX = 6;  // Cartesian components of a d shell

// ---------------------------------------------------------------------------------------
// Boys tables: F_m(T_i), m = 0..16, and e^{-T_i} in the last column, from the series
// F_m(T) = e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)) in long double.
// ---------------------------------------------------------------------------------------
long double boys_series_ld(int m, long double T) {
    long double term = 1.0L / (2 * m + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
        term *= 2.0L * T / (2 * m + 2 * k + 1);
        sum += term;
        if (term < 1e-22L * sum) break;
    }
    return sum * expl(-T);
}

// ---------------------------------------------------------------------------------------
// Device kernels of the build pipeline.
// ---------------------------------------------------------------------------------------
__global__ void k_fill(double* p, long long n, double v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_scale(const double* __restrict__ in, double* __restrict__ out, long long n, double f) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = f * in[i];
}

__global__ void k_fill_u64(unsigned long long* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

__global__ void k_fill_u32(unsigned int* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

struct DmaxArgs {
    const int* pair_a;
    const int* pair_b;
    const int* sh_ao;
    const int* sh_nao;
    const long long* sh_mbase;
    const int* sh_mn;
    const int* sh_local;
    const long long* sh_qbase;
    const int* sh_ns;
    const double *Dt, *Da, *Db;
    double *Mt, *Ma, *Mb;
    int n_pairs;
};

// max |dD| over the AO block (rows of the higher-index shell, columns of the lower one, as
// SCF.cpp:550-566), stored symmetrically in the molecule's nS x nS table.
__global__ void k_dmax(DmaxArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_pairs) return;
    const int s1 = max(a.pair_a[p], a.pair_b[p]), s2 = min(a.pair_a[p], a.pair_b[p]);
    const long long base = a.sh_mbase[s1];
    const long long N = a.sh_mn[s1];
    const int r0 = a.sh_ao[s1], nr = a.sh_nao[s1], c0 = a.sh_ao[s2], nc = a.sh_nao[s2];
    double mt = 0.0, ma = 0.0, mb = 0.0;
    for (int r = r0; r < r0 + nr; ++r)
        for (int c = c0; c < c0 + nc; ++c) {
            const long long x = base + r + c * N;
            mt = fmax(mt, fabs(a.Dt[x]));
            ma = fmax(ma, fabs(a.Da[x]));
            if (a.Db) mb = fmax(mb, fabs(a.Db[x]));
        }
    const long long q = a.sh_qbase[s1];
    const int ns = a.sh_ns[s1], l1 = a.sh_local[s1], l2 = a.sh_local[s2];
    a.Mt[q + l1 + (long long)l2 * ns] = a.Mt[q + l2 + (long long)l1 * ns] = mt;
    a.Ma[q + l1 + (long long)l2 * ns] = a.Ma[q + l2 + (long long)l1 * ns] = ma;
    if (a.Db) a.Mb[q + l1 + (long long)l2 * ns] = a.Mb[q + l2 + (long long)l1 * ns] = mb;
}

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// The bra and ket lists of one quartet class for the screening kernels (a device table of
// N_COMBOS entries, rebuilt whenever the class lists are re-sorted).
struct ScreenClass {
    const int* bra_ids;  // pair ids in sorted order
    int n_bra;           // 0: the class has no work
    const int4* ket_info;  // {pair, local a | local b << 16, nprim, 0} of each ket-list position
    const double* ket_q;   // Q of each ket-list position
    const int* ket_seg_start;
    const int* ket_seg_end;
    const int* ket_molseg;  // [mol] -> first segment, [mol+1] -> end (per ket class)
    int diagonal;           // bra class == ket class: only ket positions <= bra position
    long long slot64, slot32;
};

struct ScreenArgs {
    const ScreenClass* cls;  // [N_COMBOS]
    // per pair / shell data
    const int4* pair_info;  // {shell a, shell b, local a | local b << 16, nprim pairs}
    const int* pair_npx;    // primitive pairs the kernels execute: FP64 count | FP32 count << 16
    const int* pair_mol;
    const double* pair_q;
    const long long* sh_qbase;
    const int* sh_ns;
    const double *Mt, *Ma, *Mb;
    int ns_max;             // LDS row length (largest molecule's shell count)
    double tauS, tauD, cut;
    bool mixed;
    long long nS_total;
    // output
    uint2* list;
    unsigned long long* counts;        // pass 0: survivors of bra i at [slot64 + i] / [slot32 + i]
    const unsigned long long* offs;    // pass 1: exclusive scan of counts = list positions
    unsigned long long* stats;  // [6 k]: hash sum, hash xor, primitive quartets FP64 / FP32 of class k
                                // (unpruned model counts), then the executed ones FP64 / FP32
};

// One wavefront per bra pair (a, b).  Ket candidates: for every nprim segment of the bra's
// molecule in the ket class, the Schwarz prefix (segments are sorted by Q descending, so
// Q(ab)Q(cd) >= tau_S holds on a prefix), then the density test per candidate.  The rows a
// and b of the spin-density maxima are staged in LDS (dynamic shared memory, one region per
// wavefront).  PASS 0 counts each bra's survivors; after an exclusive scan of the counts,
// PASS 1 writes them at their positions (PASS 2 likewise, without the survivor-list hash of
// the statistics), so every class list is ordered like its bra list
// (primitive-pair count descending: the heaviest work first).
// All classes go out in one launch: the blocks of class k follow those of classes < k, each
// class taking ceil(n_bra / wavefronts per block) blocks.
template <int PASS>
__global__ __launch_bounds__(256) void k_screen(ScreenArgs a) {
    extern __shared__ double srow[];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int wpb = (int)(blockDim.x >> 6);
    int k = 0, b0 = 0;
    for (; k < N_COMBOS; ++k) {
        const int nb = (a.cls[k].n_bra + wpb - 1) / wpb;
        if ((int)blockIdx.x < b0 + nb) break;
        b0 += nb;
    }
    if (k == N_COMBOS) return;  // whole block beyond the last class (none launched)
    const ScreenClass K = a.cls[k];
    const int bpos = ((int)blockIdx.x - b0) * wpb + wib;
    const bool active = bpos < K.n_bra;
    const int nrow = a.Mb ? 4 : 2;
    double* rows = srow + (size_t)wib * nrow * a.ns_max;
    int pid = 0, sa = 0, sb = 0, la = 0, lb = 0, mol = 0;
    long long qb = 0, ns = 0;
    double qab = 0.0, tab2 = 0.0;
    long long npa = 0;
    int npxa = 0;
    if (active) {
        pid = K.bra_ids[bpos];
        const int4 pi = a.pair_info[pid];
        sa = pi.x;
        sb = pi.y;
        la = pi.z & 0xffff;
        lb = pi.z >> 16;
        npa = pi.w;
        npxa = a.pair_npx[pid];
        mol = a.pair_mol[pid];
        qab = a.pair_q[pid];
        qb = a.sh_qbase[sa];
        ns = a.sh_ns[sa];
        tab2 = 2.0 * a.Mt[qb + la + lb * ns];
        for (int x = lane; x < ns; x += 64) {
            rows[x] = a.Ma[qb + la + (long long)x * ns];
            rows[a.ns_max + x] = a.Ma[qb + lb + (long long)x * ns];
            if (a.Mb) {
                rows[2 * a.ns_max + x] = a.Mb[qb + la + (long long)x * ns];
                rows[3 * a.ns_max + x] = a.Mb[qb + lb + (long long)x * ns];
            }
        }
    }
    __syncthreads();
    if (!active) return;  // uniform per wavefront; no barriers below
    const double* ra = rows;
    const double* rb = rows + a.ns_max;
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    unsigned long long h_sum = 0, h_xor = 0, prims = 0, prims32 = 0, ex64 = 0, ex32 = 0;
    const int sg0 = K.ket_molseg[mol], sg1 = K.ket_molseg[mol + 1];
    unsigned long long n64 = 0, n32 = 0, b64 = 0, b32 = 0;
    if (PASS >= 1) {
        b64 = a.offs[K.slot64 + bpos];
        b32 = a.offs[K.slot32 + bpos];
    }
    {
        for (int sg = sg0; sg < sg1; ++sg) {
            const int s0 = K.ket_seg_start[sg];
            int s1 = K.ket_seg_end[sg];
            if (K.diagonal) s1 = min(s1, bpos + 1);
            if (s1 <= s0) continue;
            // Schwarz prefix: first x in [s0, s1) with qab * Q(x) < tau_S.
            int lo = s0, hi = s1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (qab * K.ket_q[mid] < a.tauS)
                    hi = mid;
                else
                    lo = mid + 1;
            }
            const int xend = lo;
            for (int x0 = s0; x0 < xend; x0 += 64) {
                const int x = x0 + lane;
                bool keep = false, f32 = false;
                int qid = 0;
                int4 pk = make_int4(0, 0, 0, 0);
                double mx = 0.0;
                if (x < xend) {
                    pk = K.ket_info[x];
                    qid = pk.x;
                    const int lc = pk.y & 0xffff, ld = pk.y >> 16;
                    mx = fmax(tab2, 2.0 * a.Mt[qb + lc + (long long)ld * ns]);
                    mx = fmax(mx, fmax(fmax(ra[lc], ra[ld]), fmax(rb[lc], rb[ld])));
                    if (a.Mb) {
                        const double* rc = rows + 2 * a.ns_max;
                        const double* rd = rows + 3 * a.ns_max;
                        mx = fmax(mx, fmax(fmax(rc[lc], rc[ld]), fmax(rd[lc], rd[ld])));
                    }
                    keep = !(mx < a.tauD);
                    f32 = keep && a.mixed && (qab * K.ket_q[x] * mx < a.cut);
                }
                const unsigned long long m64 = __ballot(keep && !f32);
                const unsigned long long m32 = __ballot(keep && f32);
                if (PASS == 0) {
                    n64 += __popcll(m64);
                    n32 += __popcll(m32);
                    continue;
                }
                if (keep) {
                    const uint2 ent = make_uint2((unsigned)pid, (unsigned)qid);
                    if (!f32)
                        a.list[(long long)(b64 + __popcll(m64 & lt_mask))] = ent;
                    else
                        a.list[(long long)(b32 + __popcll(m32 & lt_mask))] = ent;
                    if (f32) {
                        prims32 += (unsigned long long)(npa * pk.z);
                        ex32 += (unsigned long long)((npxa >> 16) * (pk.w >> 16));
                    } else {
                        prims += (unsigned long long)(npa * pk.z);
                        ex64 += (unsigned long long)((npxa & 0xffff) * (pk.w & 0xffff));
                    }
                    if constexpr (PASS == 1) {
                        const int4 pq = a.pair_info[qid];
                        const int sc = pq.x, sd = pq.y;
                        // Canonical (reference loop order) form of the quartet for the hash.
                        long long A = max(sa, sb), B = min(sa, sb), C = max(sc, sd), D = min(sc, sd);
                        if (A * (A + 1) / 2 + B < C * (C + 1) / 2 + D) {
                            long long t = A;
                            A = C;
                            C = t;
                            t = B;
                            B = D;
                            D = t;
                        }
                        const unsigned long long h = splitmix64(
                            (unsigned long long)(((A * a.nS_total + B) * a.nS_total + C) * a.nS_total + D));
                        h_sum += h;
                        h_xor ^= h;
                    }
                }
                b64 += __popcll(m64);
                b32 += __popcll(m32);
            }
        }
    }
    if (PASS == 0) {
        if (lane == 0) {
            a.counts[K.slot64 + bpos] = n64;
            a.counts[K.slot32 + bpos] = n32;
        }
        return;
    }
    for (int off = 32; off > 0; off >>= 1) {
        h_sum += __shfl_xor(h_sum, off);
        h_xor ^= __shfl_xor(h_xor, off);
        prims += __shfl_xor(prims, off);
        prims32 += __shfl_xor(prims32, off);
        ex64 += __shfl_xor(ex64, off);
        ex32 += __shfl_xor(ex32, off);
    }
    if (lane == 0) {
        if (h_sum) atomicAdd(&a.stats[6 * k + 0], h_sum);
        if (h_xor) atomicXor(&a.stats[6 * k + 1], h_xor);
        if (prims) atomicAdd(&a.stats[6 * k + 2], prims);
        if (prims32) atomicAdd(&a.stats[6 * k + 3], prims32);
        if (ex64) atomicAdd(&a.stats[6 * k + 4], ex64);
        if (ex32) atomicAdd(&a.stats[6 * k + 5], ex32);
    }
}

// ranges[2k..2k+1] = FP64 entries of class k, ranges[42+2k..] = FP32 entries; slots[3k..3k+2]
// are the class's first FP64 slot, first FP32 slot and end slot in the scanned counts.
__global__ void k_ranges(const unsigned long long* offs, const long long* slots, long long* ranges) {
    const int k = threadIdx.x;
    if (k >= N_COMBOS) return;
    ranges[2 * k] = (long long)offs[slots[3 * k]];
    ranges[2 * k + 1] = (long long)offs[slots[3 * k + 1]];
    ranges[2 * N_COMBOS + 2 * k] = (long long)offs[slots[3 * k + 1]];
    ranges[2 * N_COMBOS + 2 * k + 1] = (long long)offs[slots[3 * k + 2]];
    // both parts of the class (they are adjacent in the list)
    ranges[4 * N_COMBOS + 2 * k] = (long long)offs[slots[3 * k]];
    ranges[4 * N_COMBOS + 2 * k + 1] = (long long)offs[slots[3 * k + 2]];
}

// out = sum over the accumulator copies of A + A^T, within every molecule block
// (blockIdx.y = molecule).
// Sum of the per-XCD packed-triangle copies, expanded to the full symmetric block; the
// diagonal holds one orientation only and is doubled.
__global__ void k_symm(const double* A, long long stride, int copies, double* out, const long long* mbase,
                       const long long* tbase, const int* mn) {
    const int m = blockIdx.y;
    const long long base = mbase[m], tb = tbase[m];
    const long long N = mn[m];
    const long long n = N * N;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const long long r = e % N, c = e / N;
        const long long t = tb + tri_index(r, c);
        double v = 0.0;
        for (int k = 0; k < copies; ++k) v += A[k * stride + t];
        out[base + e] = (r == c) ? 2.0 * v : v;
    }
}

// The same for fixed-point accumulators (element t of a copy is the int64 pair [2t, 2t+1] =
// (hi, lo), value hi 2^-32 + lo 2^-72): the copies are summed as integers, so the result does
// not depend on how the additions were distributed over the copies.
__global__ void k_symm_fixed(const double* A, long long stride, int copies, double* out, const long long* mbase,
                             const long long* tbase, const int* mn) {
    const int m = blockIdx.y;
    const long long base = mbase[m], tb = tbase[m];
    const long long N = mn[m];
    const long long n = N * N;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const long long r = e % N, c = e / N;
        const long long t = tb + tri_index(r, c);
        unsigned long long hi = 0, lo = 0;
        for (int k = 0; k < copies; ++k) {
            const unsigned long long* q = reinterpret_cast<const unsigned long long*>(A + k * stride) + 2 * t;
            hi += q[0];
            lo += q[1];
        }
        const double v = ldexp((double)(long long)hi, -32) + ldexp((double)(long long)lo, -72);
        out[base + e] = (r == c) ? 2.0 * v : v;
    }
}

// Packed lower triangle of each molecule's block of a symmetric matrix.
__global__ void k_pack_tri(const double* __restrict__ D, double* __restrict__ T, const long long* mbase,
                           const long long* tbase, const int* mn) {
    const int m = blockIdx.y;
    const long long base = mbase[m], tb = tbase[m];
    const long long N = mn[m];
    const long long n = N * (N + 1) / 2;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        // e = i (i + 1) / 2 + j with j <= i
        long long i = (long long)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > e) --i;
        while ((i + 1) * (i + 2) / 2 <= e) ++i;
        const long long j = e - i * (i + 1) / 2;
        T[tb + e] = D[base + i + j * N];
    }
}

// ---------------------------------------------------------------------------------------
// One-electron integrals (IntegralEngine.cpp:78-368), one thread per shell pair of a class.
// KIND 0: overlap, 1: kinetic, 2: nuclear attraction.
// ---------------------------------------------------------------------------------------
struct OneArgs {
    const int* ids;  // pair ids of one class
    int n;
    const int* pair_a;
    const int* pair_b;
    const int* sh_po;
    const int* sh_np;
    const int* sh_no;
    const int* sh_ao;
    const long long* sh_mbase;
    const int* sh_mn;
    const int* sh_mol;
    const double* ctr;
    const double* alpha;
    const double* coeff;
    const double* norm;
    const double* atom_z;
    const double* atom_xyz;
    const int* mol_atom0;  // [mol], [mol+1]: atom range of a molecule
    const double* boys;    // full table, BOYS_ORDERS per grid point
    double* out;
};

template <int LA, int LB, int KIND>
__global__ __launch_bounds__(128) void k_one(OneArgs o) {
    using namespace md;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= o.n) return;
    const int pid = o.ids[t];
    const int sa = o.pair_a[pid], sb = o.pair_b[pid];
    constexpr int NA = ncart(LA), NB = ncart(LB);
    constexpr int LBE = (KIND == 1) ? LB + 2 : LB;
    constexpr int L = LA + LB;
    const double Ax = o.ctr[3 * sa], Ay = o.ctr[3 * sa + 1], Az = o.ctr[3 * sa + 2];
    const double Bx = o.ctr[3 * sb], By = o.ctr[3 * sb + 1], Bz = o.ctr[3 * sb + 2];
    const double AB2 = (Ax - Bx) * (Ax - Bx) + (Ay - By) * (Ay - By) + (Az - Bz) * (Az - Bz);
    double acc[NA][NB];
    sfor<NA>([&](auto i) { sfor<NB>([&](auto j) { acc[i][j] = 0.0; }); });
    const int npa = o.sh_np[sa], npb = o.sh_np[sb];
    for (int pa = 0; pa < npa; ++pa) {
        const double al = o.alpha[o.sh_po[sa] + pa];
        for (int pb = 0; pb < npb; ++pb) {
            const double be = o.alpha[o.sh_po[sb] + pb];
            const double p = al + be, oop = 1.0 / p;
            const double Px = (al * Ax + be * Bx) * oop, Py = (al * Ay + be * By) * oop, Pz = (al * Az + be * Bz) * oop;
            const double K = exp(-al * be * oop * AB2);
            const double cp = o.coeff[o.sh_po[sa] + pa] * o.coeff[o.sh_po[sb] + pb];
            Ecoef<LA, LBE, double> Ex, Ey, Ez;
            hermite_e<LA, LBE>(0.5 * oop, Px - Ax, Px - Bx, Ex);
            hermite_e<LA, LBE>(0.5 * oop, Py - Ay, Py - By, Ey);
            hermite_e<LA, LBE>(0.5 * oop, Pz - Az, Pz - Bz, Ez);
            const double* na = o.norm + o.sh_no[sa] + pa * NA;
            const double* nb = o.norm + o.sh_no[sb] + pb * NB;
            if constexpr (KIND == 0 || KIND == 1) {
                const double pref = pow(M_PI * oop, 1.5) * K * cp;
                sfor<NA>([&](auto ia_) {
                    constexpr int ia = decltype(ia_)::value;
                    constexpr int ax = cart_i(LA, ia), ay = cart_j(LA, ia), az = cart_k(LA, ia);
                    sfor<NB>([&](auto ib_) {
                        constexpr int ib = decltype(ib_)::value;
                        constexpr int bx = cart_i(LB, ib), by = cart_j(LB, ib), bz = cart_k(LB, ib);
                        const double sx = Ex.e[ax][bx][0], sy = Ey.e[ay][by][0], sz = Ez.e[az][bz][0];
                        double v;
                        if constexpr (KIND == 0) {
                            v = sx * sy * sz;
                        } else {
                            // Laplacian on b (IntegralEngine.cpp:217-233)
                            const double t1 = be * (2 * (bx + by + bz) + 3) * sx * sy * sz;
                            const double t2 = -2.0 * be * be *
                                              (Ex.e[ax][bx + 2][0] * sy * sz + sx * Ey.e[ay][by + 2][0] * sz +
                                               sx * sy * Ez.e[az][bz + 2][0]);
                            double t3 = 0.0;
                            if constexpr (bx >= 2) t3 -= bx * (bx - 1) * Ex.e[ax][bx - 2][0] * sy * sz;
                            if constexpr (by >= 2) t3 -= by * (by - 1) * sx * Ey.e[ay][by - 2][0] * sz;
                            if constexpr (bz >= 2) t3 -= bz * (bz - 1) * sx * sy * Ez.e[az][bz - 2][0];
                            v = t1 + t2 + 0.5 * t3;
                        }
                        acc[ia][ib] += na[ia] * nb[ib] * pref * v;
                    });
                });
            } else {
                const double pref = (2.0 * M_PI * oop) * K * cp;
                const int mol = o.sh_mol[sa];
                for (int at = o.mol_atom0[mol]; at < o.mol_atom0[mol + 1]; ++at) {
                    const double X = Px - o.atom_xyz[3 * at], Y = Py - o.atom_xyz[3 * at + 1],
                                 Z = Pz - o.atom_xyz[3 * at + 2];
                    const double T = p * (X * X + Y * Y + Z * Z);
                    double F[L + 1];
                    boys<L, double, BOYS_ORDERS>(T, o.boys + BoysSlice<L, double>::first, F);
                    double base[L + 1];
                    double pw = -o.atom_z[at] * pref;
                    sfor<L + 1>([&](auto n) {
                        base[n] = pw * F[n];
                        pw *= -2.0 * p;
                    });
                    double Rt[nherm(L)];
                    hermite_r<L>(base, X, Y, Z, Rt);
                    sfor<NA>([&](auto ia_) {
                        constexpr int ia = decltype(ia_)::value;
                        constexpr int ax = cart_i(LA, ia), ay = cart_j(LA, ia), az = cart_k(LA, ia);
                        sfor<NB>([&](auto ib_) {
                            constexpr int ib = decltype(ib_)::value;
                            constexpr int bx = cart_i(LB, ib), by = cart_j(LB, ib), bz = cart_k(LB, ib);
                            const double v = sum_of<az + bz + 1>([&](auto v_) {
                                constexpr int w = decltype(v_)::value;
                                return Ez.e[az][bz][w] * sum_of<ay + by + 1>([&](auto u_) {
                                           constexpr int u = decltype(u_)::value;
                                           return Ey.e[ay][by][u] * sum_of<ax + bx + 1>([&](auto t_) {
                                                      constexpr int tt = decltype(t_)::value;
                                                      return Ex.e[ax][bx][tt] * Rt[hidx(tt, u, w)];
                                                  });
                                       });
                            });
                            acc[ia][ib] += na[ia] * nb[ib] * v;
                        });
                    });
                }
            }
        }
    }
    const long long base = o.sh_mbase[sa];
    const long long N = o.sh_mn[sa];
    const int oA = o.sh_ao[sa], oB = o.sh_ao[sb];
    sfor<NA>([&](auto ia) {
        sfor<NB>([&](auto ib) {
            o.out[base + (oA + ia) + (oB + ib) * N] = acc[ia][ib];
            o.out[base + (oB + ib) + (oA + ia) * N] = acc[ia][ib];
        });
    });
}

template <int KIND>
void launch_one(int pc, const OneArgs& o, hipStream_t s) {
    const int grid = (o.n + 127) / 128;
    if (o.n == 0) return;
    switch (pc) {
        case 0: hipLaunchKernelGGL((k_one<0, 0, KIND>), dim3(grid), dim3(128), 0, s, o); break;
        case 1: hipLaunchKernelGGL((k_one<1, 0, KIND>), dim3(grid), dim3(128), 0, s, o); break;
        case 2: hipLaunchKernelGGL((k_one<1, 1, KIND>), dim3(grid), dim3(128), 0, s, o); break;
        case 3: hipLaunchKernelGGL((k_one<2, 0, KIND>), dim3(grid), dim3(128), 0, s, o); break;
        case 4: hipLaunchKernelGGL((k_one<2, 1, KIND>), dim3(grid), dim3(128), 0, s, o); break;
        case 5: hipLaunchKernelGGL((k_one<2, 2, KIND>), dim3(grid), dim3(128), 0, s, o); break;
    }
}

// Build parameters that kernels of the pipeline take by value (the graph cache key).
struct BuildParams {
    double tauS, tauD, cut;
    bool mixed, uhf, timing;
    bool hash = true;     // the survivor-list hash (statistics only)
    bool stored = false;  // digest the stored integrals (no screening pass)
    bool det = false;     // fixed-point accumulators (bitwise-reproducible J/K)
    bool operator==(const BuildParams& o) const {
        return tauS == o.tauS && tauD == o.tauD && cut == o.cut && mixed == o.mixed && uhf == o.uhf &&
               timing == o.timing && hash == o.hash && stored == o.stored && det == o.det;
    }
};

// ---------------------------------------------------------------------------------------
// Device memory owned by an engine.
// ---------------------------------------------------------------------------------------
struct DevArena {
    std::vector<void*> ptrs;
    template <class T>
    T* alloc(size_t n) {
        void* p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
    ~DevArena() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

}  // namespace

// Staging memory of hfx_one_electron_noalloc: the output blocks (SCRATCH_OUT doubles), the atoms'
// charges and positions (SCRATCH_ATOMS of each) and the per-molecule atom ranges.  Every GPU holds
// its own copy for the life of the process; host code uses the address hipGetSymbolAddress gives
// for the current device (scratch_address), never the array's name.
namespace {
constexpr long long SCRATCH_OUT = 1LL << 21;  // 16 MiB of output staging
constexpr int SCRATCH_ATOMS = 4096;
constexpr int SCRATCH_MOLS = 4096;
constexpr long long SCRATCH_DOUBLES = SCRATCH_OUT + 4LL * SCRATCH_ATOMS + (SCRATCH_MOLS + 2) / 2 + 1;
}  // namespace
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// The device address of `scratch` on `device` (which must be current), looked up once per device.
int scratch_address(int device, double** out) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return fail(HFX_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (r != hipSuccess || !p)
            return fail(HFX_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(r));
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return HFX_OK;
}
}  // namespace

struct hfx_engine {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    // Streams the class kernels of a build are spread over (side[0] is `stream`); the
    // captured graph has one branch per stream.
    static constexpr int N_SIDE = 4;
    int n_side = N_SIDE;  // streams in use (HFX_BUILD_STREAMS, 1..N_SIDE, or hfx_set_launch)
    int blocks_per_cu = 2;  // persistent blocks per CU of a class launch (HFX_BLOCKS_PER_CU, hfx_set_launch)
    int chunk_len = 0;      // list entries a block claims at a time (0: the defaults below; HFX_CHUNK_LEN)
    int stagger = 0;        // staggered walk of the claimed runs (HFX_STAGGER)
    hipStream_t side[N_SIDE] = {};
    hipEvent_t ev_fork = nullptr, ev_join[N_SIDE] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_async = nullptr;  // orders hfx_jk_async builds against the engine's stream
    double* scratch = nullptr;      // this device's address of the allocation-free staging array
    DevArena mem;

    // molecules
    int n_mol = 0;
    std::vector<int> mol_N, mol_nS, mol_shell0, mol_atom0;
    std::vector<long long> mol_mbase, mol_qbase, mol_tbase;
    long long n_elems = 0, n_qelems = 0, n_telems = 0;
    int nS = 0, N_tot = 0;

    // shells (global index), host
    std::vector<int> sh_l, sh_np, sh_po, sh_no, sh_ao, sh_nao, sh_mol, sh_local, sh_ns, sh_mn;
    std::vector<long long> sh_mbase, sh_qbase, sh_tbase;
    std::vector<double> sh_ctr, sh_cf, alpha, coeff, norm;

    // pairs
    int n_pairs = 0;
    std::vector<int> pair_a, pair_b, pair_ps, pair_np, pair_pc, pair_mol;
    std::vector<int> pair_nprim;  // unpruned primitive-pair count (the model's work)
    std::map<long long, int> pair_of;
    long long n_pp = 0;

    // class lists (sorted) and segments
    std::vector<int> cls_ids[N_PAIR_CLASSES];
    std::vector<double> cls_q[N_PAIR_CLASSES];
    std::vector<int> seg_start[N_PAIR_CLASSES], seg_end[N_PAIR_CLASSES], molseg[N_PAIR_CLASSES];
    std::vector<double> pair_q;
    std::vector<double> sort_q;  // the Q table the current order was built from
    std::vector<double> own_q;   // the engine's Schwarz table
    std::vector<long long> reg;  // class regions in the work list, [N_COMBOS + 1]

    // device copies
    int *d_pair_a = nullptr, *d_pair_b = nullptr, *d_pair_ps = nullptr, *d_pair_np = nullptr, *d_pair_mol = nullptr;
    int *d_pair_ps32 = nullptr, *d_pair_np32 = nullptr;
    std::vector<int> pair_ps32, pair_np32;
    long long n_pp32 = 0;
    int4* d_pair_info = nullptr;
    int* d_pair_npx = nullptr;  // pair_np | pair_np32 << 16
    int ns_max = 0;
    double* d_pair_q = nullptr;
    PrimPair64* d_pp64 = nullptr;
    PrimPair32* d_pp32 = nullptr;
    double* d_pab64 = nullptr;  // exponents (a, b) per primitive pair (gradient kernels)
    float* d_pab32 = nullptr;
    int* d_sh_l = nullptr;
    // gradient buffers, grown on demand
    int* d_sh_atom = nullptr;
    double *d_atom_z = nullptr, *d_atom_xyz = nullptr, *d_gpart = nullptr, *d_gout = nullptr;
    int atom_cap = 0;
    long long gpart_cap = 0;
    int *d_sh_ao = nullptr, *d_sh_nao = nullptr, *d_sh_mn = nullptr, *d_sh_local = nullptr, *d_sh_ns = nullptr;
    int *d_sh_po = nullptr, *d_sh_np = nullptr, *d_sh_no = nullptr, *d_sh_mol = nullptr;
    long long *d_sh_mbase = nullptr, *d_sh_qbase = nullptr, *d_sh_tbase = nullptr;
    double *d_sh_cf = nullptr, *d_ctr = nullptr, *d_alpha = nullptr, *d_coeff = nullptr, *d_norm = nullptr;
    int* d_cls_ids[N_PAIR_CLASSES] = {};
    int4* d_cls_info[N_PAIR_CLASSES] = {};
    double* d_cls_q[N_PAIR_CLASSES] = {};
    int *d_seg_start[N_PAIR_CLASSES] = {}, *d_seg_end[N_PAIR_CLASSES] = {}, *d_molseg[N_PAIR_CLASSES] = {};
    double* d_boys64 = nullptr;
    float* d_boys32 = nullptr;
    long long* d_mbase = nullptr;
    long long* d_tbase = nullptr;
    int* d_mn = nullptr;
    long long* d_reg = nullptr;

    // per-build buffers
    double *d_Dt = nullptr, *d_Da = nullptr, *d_Db = nullptr;
    double *d_Tt = nullptr, *d_Ta = nullptr, *d_Tb = nullptr;  // packed triangles of d_Dt, d_Da, d_Db
    double *d_Mt = nullptr, *d_Ma = nullptr, *d_Mb = nullptr;
    double *d_AJ = nullptr, *d_AKa = nullptr, *d_AKb = nullptr;  // ACC_COPIES packed-triangle copies each
    static constexpr int ACC_COPIES = 8;  // one per XCD of an MI355X
    double *d_J = nullptr, *d_Ka = nullptr, *d_Kb = nullptr;
    uint2* d_list = nullptr;
    long long list_cap = 0;
    unsigned long long* d_counts = nullptr;  // [n_slots + 1] survivors per (class, precision, bra)
    unsigned long long* d_offs = nullptr;    // exclusive scan of d_counts
    long long n_slots = 0;
    std::vector<long long> slots;            // [3 * N_COMBOS]: FP64 slot, FP32 slot, end slot
    long long* d_slots = nullptr;
    void* d_scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    unsigned long long* d_stats = nullptr;  // [6 * N_COMBOS]
    hipEvent_t cev[2 * N_COMBOS + 1] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    BuildParams graph_key{-1.0, -1.0, -1.0, false, false, false};
    long long* d_ranges = nullptr;          // [6*N_COMBOS]: FP64 parts, FP32 parts, both
    ScreenClass* d_screen_cls = nullptr;     // [N_COMBOS]: class lists for the screening kernels
    int screen_nbra[N_COMBOS] = {};          // their bra counts (0: class without work)
    unsigned int* d_chunk_ctr = nullptr;     // [2*N_COMBOS*CHUNK_SLOTS]: chunk counters per class kernel
    static constexpr int CHUNK_SLOTS = 64;   // chunk kernels per class, at most
    static constexpr int CHUNK_LEN = 2048;   // list entries a block claims at a time (HFX_CHUNK_LEN to change)
    static constexpr int STORED_CHUNK_LEN = 8192;  // the same for stored builds
    int max_N = 0;

    // stored integrals (hfx_store_eris): the Schwarz-screened list of every class, its ranges and
    // the normalized blocks in the stored-integral layout, one region per class
    uint2* d_slist = nullptr;
    long long* d_sranges = nullptr;  // [2 * N_COMBOS]
    double* d_store = nullptr;
    // the ket-major copy for the Coulomb pass: per class, entries {ket pair, bra pair} sorted by
    // ket pair (same ranges as d_slist) and their blocks transposed, in the same layout
    uint2* d_tlist = nullptr;
    double* d_storeT = nullptr;
    long long store_n[N_COMBOS] = {}, store_base[N_COMBOS] = {};
    long long store_elems = 0;
    double store_tauS = -1.0;
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_key = BuildParams{-1.0, -1.0, -1.0, false, false, false};
    }
    void release_store() {
        if (d_store) drop_graph();  // a captured stored build reads these buffers
        if (d_slist) (void)hipFree(d_slist);
        if (d_sranges) (void)hipFree(d_sranges);
        if (d_store) (void)hipFree(d_store);
        if (d_tlist) (void)hipFree(d_tlist);
        if (d_storeT) (void)hipFree(d_storeT);
        d_slist = nullptr;
        d_sranges = nullptr;
        d_store = nullptr;
        d_tlist = nullptr;
        d_storeT = nullptr;
        for (int k = 0; k < N_COMBOS; ++k) store_n[k] = store_base[k] = 0;
        store_elems = 0;
        store_tauS = -1.0;
    }

    ~hfx_engine() {
        release_store();
        drop_graph();
        for (hipEvent_t& x : cev)
            if (x) (void)hipEventDestroy(x);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ev_async) (void)hipEventDestroy(ev_async);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        for (int i = 1; i < N_SIDE; ++i) {
            if (ev_join[i]) (void)hipEventDestroy(ev_join[i]);
            if (side[i]) (void)hipStreamDestroy(side[i]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <class T>
int upload(hfx_engine* e, T*& dst, const std::vector<T>& src) {
    dst = e->mem.alloc<T>(src.size());
    if (!dst) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
    if (!src.empty()) HIPCHK(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return HFX_OK;
}

int ncart_h(int l) { return (l + 1) * (l + 2) / 2; }

double model_flop_h(int la, int lb, int lc, int ld) {  // per primitive quartet, bra = larger pair class
    auto nherm = [](int L) { return (L + 1) * (L + 2) * (L + 3) / 6; };
    auto hsum = [](int l1, int l2) {
        double s = 0;
        for (int i1 = l1; i1 >= 0; --i1)
            for (int j1 = l1 - i1; j1 >= 0; --j1)
                for (int i2 = l2; i2 >= 0; --i2)
                    for (int j2 = l2 - i2; j2 >= 0; --j2)
                        s += (i1 + i2 + 1) * (j1 + j2 + 1) * ((l1 - i1 - j1) + (l2 - i2 - j2) + 1);
        return s;
    };
    const int L = la + lb + lc + ld;
    const double c4 = (double)(L + 4) * (L + 3) * (L + 2) * (L + 1) / 24.0;
    return (20 + 3 * (L + 1)) + 3 * c4 + 2.0 * nherm(la + lb) * hsum(lc, ld) +
           2.0 * ncart_h(lc) * ncart_h(ld) * hsum(la, lb) + 2.0 * ncart_h(la) * ncart_h(lb) * ncart_h(lc) * ncart_h(ld);
}

// Sort every class list by (molecule, nprim desc, Q desc, id) using the Q table qt
// (packed per molecule), build the nprim segments, and upload.
int sort_pairs(hfx_engine* e, const std::vector<double>& qt) {
    e->pair_q.resize(e->n_pairs);
    for (int p = 0; p < e->n_pairs; ++p) {
        const int a = e->pair_a[p], b = e->pair_b[p];
        e->pair_q[p] = qt[e->sh_qbase[a] + e->sh_local[a] + (long long)e->sh_local[b] * e->sh_ns[a]];
    }
    for (int pc = 0; pc < N_PAIR_CLASSES; ++pc) {
        std::vector<int>& ids = e->cls_ids[pc];
        std::sort(ids.begin(), ids.end(), [&](int x, int y) {
            if (e->pair_mol[x] != e->pair_mol[y]) return e->pair_mol[x] < e->pair_mol[y];
            if (e->pair_np[x] != e->pair_np[y]) return e->pair_np[x] > e->pair_np[y];
            if (e->pair_q[x] != e->pair_q[y]) return e->pair_q[x] > e->pair_q[y];
            return x < y;
        });
        e->cls_q[pc].resize(ids.size());
        for (size_t i = 0; i < ids.size(); ++i) e->cls_q[pc][i] = e->pair_q[ids[i]];
        e->seg_start[pc].clear();
        e->seg_end[pc].clear();
        e->molseg[pc].assign(e->n_mol + 1, 0);
        size_t i = 0;
        for (int m = 0; m < e->n_mol; ++m) {
            e->molseg[pc][m] = (int)e->seg_start[pc].size();
            while (i < ids.size() && e->pair_mol[ids[i]] == m) {
                size_t j = i;
                while (j < ids.size() && e->pair_mol[ids[j]] == m && e->pair_np[ids[j]] == e->pair_np[ids[i]]) ++j;
                e->seg_start[pc].push_back((int)i);
                e->seg_end[pc].push_back((int)j);
                i = j;
            }
        }
        e->molseg[pc][e->n_mol] = (int)e->seg_start[pc].size();
        auto put = [&](auto*& dst, const auto& src) -> int {
            using T = typename std::decay_t<decltype(src)>::value_type;
            if (!dst) {
                dst = e->mem.alloc<T>(std::max<size_t>(src.size(), 1) + 1);
                if (!dst) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
            }
            if (!src.empty()) HIPCHK(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
            return HFX_OK;
        };
        int rc;
        if ((rc = put(e->d_cls_ids[pc], ids)) || (rc = put(e->d_cls_q[pc], e->cls_q[pc]))) return rc;
        {
            // what the screening needs of each position, in one load:
            // {pair, local a | local b << 16, unpruned nprim, executed FP64 | FP32 nprim << 16}
            std::vector<int4> info(ids.size());
            for (size_t i = 0; i < ids.size(); ++i) {
                const int p = ids[i];
                info[i] = make_int4(p, e->sh_local[e->pair_a[p]] | (e->sh_local[e->pair_b[p]] << 16), e->pair_nprim[p],
                                    e->pair_np[p] | (e->pair_np32[p] << 16));
            }
            if ((rc = put(e->d_cls_info[pc], info))) return rc;
        }
        // segment arrays can change length between sorts only through nprim runs, which are fixed
        // by the pair set; their capacity is the number of pairs in the class.
        if (!e->d_seg_start[pc]) {
            e->d_seg_start[pc] = e->mem.alloc<int>(ids.size() + 1);
            e->d_seg_end[pc] = e->mem.alloc<int>(ids.size() + 1);
            if (!e->d_seg_start[pc] || !e->d_seg_end[pc]) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
        }
        if (!e->seg_start[pc].empty()) {
            HIPCHK(hipMemcpy(e->d_seg_start[pc], e->seg_start[pc].data(), e->seg_start[pc].size() * sizeof(int),
                             hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e->d_seg_end[pc], e->seg_end[pc].data(), e->seg_end[pc].size() * sizeof(int),
                             hipMemcpyHostToDevice));
        }
        if ((rc = put(e->d_molseg[pc], e->molseg[pc]))) return rc;
    }
    if (!e->d_pair_q) {
        e->d_pair_q = e->mem.alloc<double>(e->n_pairs);
        if (!e->d_pair_q) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
    }
    HIPCHK(hipMemcpy(e->d_pair_q, e->pair_q.data(), e->n_pairs * sizeof(double), hipMemcpyHostToDevice));
    e->sort_q = qt;
    // The screening kernels' class table.
    {
        std::vector<ScreenClass> t(N_COMBOS);
        for (int k = 0; k < N_COMBOS; ++k) {
            const int p1 = combo_pc1(k), p2 = combo_pc2(k);
            ScreenClass c{};
            const bool live = !e->cls_ids[p1].empty() && !e->cls_ids[p2].empty();
            c.bra_ids = e->d_cls_ids[p1];
            c.n_bra = live ? (int)e->cls_ids[p1].size() : 0;
            c.ket_info = e->d_cls_info[p2];
            c.ket_q = e->d_cls_q[p2];
            c.ket_seg_start = e->d_seg_start[p2];
            c.ket_seg_end = e->d_seg_end[p2];
            c.ket_molseg = e->d_molseg[p2];
            c.diagonal = p1 == p2;
            c.slot64 = e->slots[3 * k];
            c.slot32 = e->slots[3 * k + 1];
            t[k] = c;
            e->screen_nbra[k] = c.n_bra;
        }
        if (!e->d_screen_cls) {
            e->d_screen_cls = e->mem.alloc<ScreenClass>(N_COMBOS);
            if (!e->d_screen_cls) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
        }
        HIPCHK(hipMemcpy(e->d_screen_cls, t.data(), N_COMBOS * sizeof(ScreenClass), hipMemcpyHostToDevice));
    }
    // A captured build baked in the previous order's launch shapes: capture again.
    e->graph_key.tauS = -1.0;
    return HFX_OK;
}

EriArgs eri_args(hfx_engine* e) {
    EriArgs a{};
    a.pair_a = e->d_pair_a;
    a.pair_b = e->d_pair_b;
    a.pair_ps = e->d_pair_ps;
    a.pair_np = e->d_pair_np;
    a.pair_ps32 = e->d_pair_ps32;
    a.pair_np32 = e->d_pair_np32;
    a.pp64 = e->d_pp64;
    a.pp32 = e->d_pp32;
    a.sh_ao = e->d_sh_ao;
    a.sh_mbase = e->d_sh_mbase;
    a.sh_tbase = e->d_sh_tbase;
    a.sh_local = e->d_sh_local;
    a.sh_nao = e->d_sh_nao;
    a.sh_ns = e->d_sh_ns;
    a.tile_shells = e->ns_max;
    a.sh_mn = e->d_sh_mn;
    a.sh_cf = e->d_sh_cf;
    a.boys64 = e->d_boys64;
    a.boys32 = e->d_boys32;
    a.n_copies = 1;
    a.acc_stride = 0;
    return a;
}

// Blocks per class launch: two per CU (HFX_BLOCKS_PER_CU to change).  The launches of a build
// run four at a time on the streams, and small persistent grids let them share the machine.
int eri_grid(hfx_engine* e) { return e->n_cu * e->blocks_per_cu; }

// Store-mode ERI blocks for a list of (bra pair, ket pair) entries of one class.
int store_blocks(hfx_engine* e, int combo, const std::vector<uint2>& ents, std::vector<double>& out,
                 std::vector<long long>& offs) {
    const int pc1 = combo_pc1(combo), pc2 = combo_pc2(combo);
    const long long blk = (long long)ncart_h(pc_la(pc1)) * ncart_h(pc_lb(pc1)) * ncart_h(pc_la(pc2)) * ncart_h(pc_lb(pc2));
    const long long n = (long long)ents.size();
    offs.resize(n);
    for (long long i = 0; i < n; ++i) offs[i] = i * blk;
    out.assign(n * blk, 0.0);
    if (n == 0) return HFX_OK;
    if (eri_chunks(combo, 0) < 0) return fail(HFX_ERR_UNSUPPORTED, "ERI kernels missing for class " + std::to_string(combo));
    DevArena tmp;
    uint2* dl = tmp.alloc<uint2>(n);
    long long* doff = tmp.alloc<long long>(n);
    double* dout = tmp.alloc<double>(n * blk);
    long long* drange = tmp.alloc<long long>(2);
    if (!dl || !doff || !dout || !drange) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
    const long long rng[2] = {0, n};
    HIPCHK(hipMemcpy(dl, ents.data(), n * sizeof(uint2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(doff, offs.data(), n * sizeof(long long), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(drange, rng, sizeof(rng), hipMemcpyHostToDevice));
    EriArgs a = eri_args(e);
    a.list = dl;
    a.range = drange;
    a.out = dout;
    a.out_off = doff;
    std::vector<hipStream_t> cs(std::max(1, eri_chunks(combo, 0)), e->stream);
    launch_eri(combo, 0, a, eri_grid(e), cs.data());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(out.data(), dout, n * blk * sizeof(double), hipMemcpyDeviceToHost));
    return HFX_OK;
}

// Schwarz table from the diagonal quartets (ab|ab) of every pair (IntegralEngine.cpp:695-728).
int compute_schwarz(hfx_engine* e) {
    e->own_q.assign(e->n_qelems, 0.0);
    for (int pc = 0; pc < N_PAIR_CLASSES; ++pc) {
        const std::vector<int>& ids = e->cls_ids[pc];
        if (ids.empty()) continue;
        std::vector<uint2> ents(ids.size());
        for (size_t i = 0; i < ids.size(); ++i) ents[i] = make_uint2(ids[i], ids[i]);
        std::vector<double> out;
        std::vector<long long> offs;
        const int combo = combo_of(pc, pc);
        int rc = store_blocks(e, combo, ents, out, offs);
        if (rc) return rc;
        const int nab = ncart_h(pc_la(pc)) * ncart_h(pc_lb(pc));
        for (size_t i = 0; i < ids.size(); ++i) {
            double mx = 0.0;
            for (int ab = 0; ab < nab; ++ab) mx = std::max(mx, out[offs[i] + (long long)ab * nab + ab]);
            const int a = e->pair_a[ids[i]], b = e->pair_b[ids[i]];
            const long long q = e->sh_qbase[a];
            const int ns = e->sh_ns[a];
            const double v = std::sqrt(mx);
            e->own_q[q + e->sh_local[a] + (long long)e->sh_local[b] * ns] = v;
            e->own_q[q + e->sh_local[b] + (long long)e->sh_local[a] * ns] = v;
        }
    }
    return HFX_OK;
}

bool finite_all(const double* x, long long n) {
    for (long long i = 0; i < n; ++i)
        if (!std::isfinite(x[i])) return false;
    return true;
}

int create_impl(int n_mol, const hfx_basis* bases, int device, hfx_engine** out) {
    if (!out || !bases || n_mol <= 0) return fail(HFX_ERR_INVALID, "hfx_create: bad arguments");
    *out = nullptr;
    for (int c = 0; c < N_COMBOS; ++c)
        if (eri_chunks(c, 0) < 0 || eri_chunks(c, 1) < 0)
            return fail(HFX_ERR_UNSUPPORTED, "library built without the kernels of class " + std::to_string(c));
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(HFX_ERR_INVALID, "hfx_create: no such device");
    HIPCHK(hipSetDevice(device));
    std::unique_ptr<hfx_engine> e(new hfx_engine());
    e->device = device;
    {
        const int rc = scratch_address(device, &e->scratch);
        if (rc) return rc;
    }
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    e->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    e->side[0] = e->stream;
    for (int i = 1; i < hfx_engine::N_SIDE; ++i) {
        HIPCHK(hipStreamCreateWithFlags(&e->side[i], hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&e->ev_join[i], hipEventDisableTiming));
    }
    HIPCHK(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
    if (const char* v = std::getenv("HFX_BUILD_STREAMS")) e->n_side = std::max(1, std::min(hfx_engine::N_SIDE, std::atoi(v)));
    if (const char* v = std::getenv("HFX_BLOCKS_PER_CU")) e->blocks_per_cu = std::max(1, std::min(16, std::atoi(v)));
    if (const char* v = std::getenv("HFX_CHUNK_LEN")) e->chunk_len = std::max(256, std::min(65536, std::atoi(v)));
    if (const char* v = std::getenv("HFX_STAGGER")) e->stagger = std::atoi(v) ? 1 : 0;
    HIPCHK(hipEventCreate(&e->ev0));
    HIPCHK(hipEventCreate(&e->ev1));
    HIPCHK(hipEventCreateWithFlags(&e->ev_async, hipEventDisableTiming));
    for (hipEvent_t& x : e->cev) HIPCHK(hipEventCreate(&x));

    // ---- flatten the molecules' bases ----
    e->n_mol = n_mol;
    int prim_base = 0, norm_base = 0;
    for (int m = 0; m < n_mol; ++m) {
        const hfx_basis* b = &bases[m];
        if (b->n_shells <= 0 || !b->l || !b->n_prim || !b->prim_offset || !b->ao_offset || !b->norm_offset ||
            !b->centers || !b->alpha || !b->coeff || !b->norm || !b->cart)
            return fail(HFX_ERR_INVALID, "hfx_basis: missing field");
        e->mol_shell0.push_back(e->nS);
        e->mol_N.push_back(b->n_aos);
        e->mol_nS.push_back(b->n_shells);
        e->mol_mbase.push_back(e->n_elems);
        e->mol_qbase.push_back(e->n_qelems);
        e->mol_tbase.push_back(e->n_telems);
        int ao = 0;
        for (int s = 0; s < b->n_shells; ++s) {
            const int l = b->l[s];
            if (l < 0 || l > HFX_MAX_L) return fail(HFX_ERR_UNSUPPORTED, "angular momentum above d");
            const int nao = ncart_h(l);
            if (b->ao_offset[s] != ao) return fail(HFX_ERR_INVALID, "hfx_basis: ao_offset not contiguous");
            if (b->n_prim[s] <= 0) return fail(HFX_ERR_INVALID, "hfx_basis: shell without primitives");
            for (int c = 0; c < nao; ++c) {  // the kernels assume the reference's Cartesian order
                const int* cc = &b->cart[3 * (ao + c)];
                if (cc[0] != md::cart_i(l, c) || cc[1] != md::cart_j(l, c) || cc[2] != md::cart_k(l, c))
                    return fail(HFX_ERR_UNSUPPORTED, "hfx_basis: unexpected Cartesian order");
            }
            e->sh_l.push_back(l);
            e->sh_np.push_back(b->n_prim[s]);
            e->sh_po.push_back(prim_base + b->prim_offset[s]);
            e->sh_no.push_back(norm_base + b->norm_offset[s]);
            e->sh_ao.push_back(ao);
            e->sh_nao.push_back(nao);
            e->sh_mol.push_back(m);
            e->sh_local.push_back(s);
            e->sh_ns.push_back(b->n_shells);
            e->sh_mn.push_back(b->n_aos);
            e->sh_mbase.push_back(e->n_elems);
            e->sh_qbase.push_back(e->n_qelems);
            e->sh_tbase.push_back(e->n_telems);
            for (int x = 0; x < 3; ++x) e->sh_ctr.push_back(b->centers[3 * s + x]);
            // normalization ratio of each component (independent of the primitive)
            const double* nrm = &b->norm[b->norm_offset[s]];
            for (int c = 0; c < CMAX; ++c) e->sh_cf.push_back(c < nao ? nrm[c] / nrm[0] : 0.0);
            ao += nao;
        }
        if (ao != b->n_aos) return fail(HFX_ERR_INVALID, "hfx_basis: n_aos mismatch");
        for (int p = 0; p < b->n_prims; ++p) {
            e->alpha.push_back(b->alpha[p]);
            e->coeff.push_back(b->coeff[p]);
        }
        for (int x = 0; x < b->n_norm; ++x) e->norm.push_back(b->norm[x]);
        prim_base += b->n_prims;
        norm_base += b->n_norm;
        e->nS += b->n_shells;
        e->N_tot += b->n_aos;
        e->n_elems += (long long)b->n_aos * b->n_aos;
        e->n_qelems += (long long)b->n_shells * b->n_shells;
        e->n_telems += (long long)b->n_aos * (b->n_aos + 1) / 2;
    }

    // ---- shell pairs and primitive-pair data ----
    constexpr double PRUNE64 = 1e-20, PRUNE32 = 1e-11;
    std::vector<PrimPair64> pp64;
    std::vector<PrimPair32> pp32;
    std::vector<double> pab64;
    std::vector<float> pab32;
    std::vector<int> cls_count(N_PAIR_CLASSES, 0);
    for (int m = 0; m < n_mol; ++m) {
        const int s0 = e->mol_shell0[m], s1 = s0 + e->mol_nS[m];
        for (int i = s0; i < s1; ++i)
            for (int j = s0; j <= i; ++j) {
                int a = i, b = j;
                if (e->sh_l[b] > e->sh_l[a]) std::swap(a, b);
                const int pc = pair_class(e->sh_l[a], e->sh_l[b]);
                const int pid = e->n_pairs++;
                e->pair_a.push_back(a);
                e->pair_b.push_back(b);
                e->pair_pc.push_back(pc);
                e->pair_mol.push_back(m);
                e->pair_nprim.push_back(e->sh_np[a] * e->sh_np[b]);
                e->pair_of[(long long)j * e->nS + i] = pid;
                e->cls_ids[pc].push_back(pid);
                cls_count[pc]++;
                const double* A = &e->sh_ctr[3 * a];
                const double* B = &e->sh_ctr[3 * b];
                const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
                e->pair_ps.push_back((int)pp64.size());
                e->pair_ps32.push_back((int)pp32.size());
                for (int pa = 0; pa < e->sh_np[a]; ++pa)
                    for (int pb = 0; pb < e->sh_np[b]; ++pb) {
                        const double al = e->alpha[e->sh_po[a] + pa], be = e->alpha[e->sh_po[b] + pb];
                        const double p = al + be, oop = 1.0 / p;
                        PrimPair64 d;
                        d.p = p;
                        d.px = (al * A[0] + be * B[0]) * oop;
                        d.py = (al * A[1] + be * B[1]) * oop;
                        d.pz = (al * A[2] + be * B[2]) * oop;
                        d.pax = d.px - A[0];
                        d.pay = d.py - A[1];
                        d.paz = d.pz - A[2];
                        d.pbx = d.px - B[0];
                        d.pby = d.py - B[1];
                        d.pbz = d.pz - B[2];
                        d.oo2p = 0.5 * oop;
                        const double K = std::exp(-al * be * oop * AB2);
                        const double ga = e->norm[e->sh_no[a] + pa * e->sh_nao[a]];
                        const double gb = e->norm[e->sh_no[b] + pb * e->sh_nao[b]];
                        const double kc = K * (e->coeff[e->sh_po[a] + pa] * ga) * (e->coeff[e->sh_po[b] + pb] * gb);
                        d.kp = kc * oop;
                        if (std::fabs(kc) >= PRUNE64) {
                            pp64.push_back(d);
                            pab64.push_back(al);
                            pab64.push_back(be);
                        }
                        if (std::fabs(kc) >= PRUNE32) {
                            pab32.push_back((float)al);
                            pab32.push_back((float)be);
                        }
                        if (std::fabs(kc) >= PRUNE32)
                            pp32.push_back(PrimPair32{(float)d.p, (float)d.px, (float)d.py, (float)d.pz, (float)d.kp,
                                                      (float)d.pax, (float)d.pay, (float)d.paz, (float)d.pbx,
                                                      (float)d.pby, (float)d.pbz, (float)d.oo2p});
                    }
                e->pair_np.push_back((int)pp64.size() - e->pair_ps.back());
                e->pair_np32.push_back((int)pp32.size() - e->pair_ps32.back());
            }
    }
    e->n_pp = (long long)pp64.size();
    e->n_pp32 = (long long)pp32.size();
    // One zero element past the end: the class kernels load the primitive pair after the one
    // they use one iteration ahead, without a bounds clamp (its kp is 0).
    pp64.push_back(PrimPair64{});
    pp32.push_back(PrimPair32{});
    std::vector<int4> pinfo(e->n_pairs);
    for (int p = 0; p < e->n_pairs; ++p) {
        const int a = e->pair_a[p], b = e->pair_b[p];
        pinfo[p] = make_int4(a, b, e->sh_local[a] | (e->sh_local[b] << 16), e->pair_nprim[p]);
    }
    std::vector<int> npx(e->n_pairs);
    for (int p = 0; p < e->n_pairs; ++p) {
        if (e->pair_np[p] > 65535 || e->pair_np32[p] > 32767)
            return fail(HFX_ERR_UNSUPPORTED, "more than 32767 primitive pairs in one shell pair");
        npx[p] = e->pair_np[p] | (e->pair_np32[p] << 16);
    }
    for (int m = 0; m < n_mol; ++m) e->ns_max = std::max(e->ns_max, e->mol_nS[m]);
    if (e->ns_max > 65535) return fail(HFX_ERR_UNSUPPORTED, "more than 65535 shells in one molecule");
    // ---- Boys tables ----
    // FP64 at T_i = 0.1 i (i <= 300), FP32 at T_i = 0.05 i (i <= 600), each grid point as the
    // kernels form it in their precision.
    std::vector<double> b64((size_t)md::BoysGrid<double>::points * md::BOYS_ORDERS);
    std::vector<float> b32((size_t)md::BoysGrid<float>::points * md::BOYS_ORDERS);
    for (int i = 0; i < md::BoysGrid<double>::points; ++i) {
        const double T64 = (double)i * md::BoysGrid<double>::step;
        for (int m = 0; m < md::BOYS_EXP_COL; ++m)
            b64[(size_t)i * md::BOYS_ORDERS + m] = (double)boys_series_ld(m, (long double)T64);
        b64[(size_t)i * md::BOYS_ORDERS + md::BOYS_EXP_COL] = (double)std::exp(-(long double)T64);
    }
    for (int i = 0; i < md::BoysGrid<float>::points; ++i) {
        const float T32 = (float)i * md::BoysGrid<float>::step;
        for (int m = 0; m < md::BOYS_EXP_COL; ++m)
            b32[(size_t)i * md::BOYS_ORDERS + m] = (float)boys_series_ld(m, (long double)T32);
        b32[(size_t)i * md::BOYS_ORDERS + md::BOYS_EXP_COL] = (float)std::exp(-(long double)T32);
    }

    // ---- work-list regions: unscreened canonical quartets per class ----
    e->reg.assign(N_COMBOS + 1, 0);
    {
        std::vector<std::vector<long long>> cnt(n_mol, std::vector<long long>(N_PAIR_CLASSES, 0));
        for (int p = 0; p < e->n_pairs; ++p) cnt[e->pair_mol[p]][e->pair_pc[p]]++;
        long long acc = 0;
        for (int k = 0; k < N_COMBOS; ++k) {
            e->reg[k] = acc;
            const int p1 = combo_pc1(k), p2 = combo_pc2(k);
            for (int m = 0; m < n_mol; ++m)
                acc += (p1 == p2) ? cnt[m][p1] * (cnt[m][p1] + 1) / 2 : cnt[m][p1] * cnt[m][p2];
        }
        e->reg[N_COMBOS] = acc;
        e->list_cap = acc;
    }

    // ---- device upload ----
    int rc;
    if ((rc = upload(e.get(), e->d_pair_a, e->pair_a)) || (rc = upload(e.get(), e->d_pair_b, e->pair_b)) ||
        (rc = upload(e.get(), e->d_pair_ps, e->pair_ps)) || (rc = upload(e.get(), e->d_pair_np, e->pair_np)) ||
        (rc = upload(e.get(), e->d_pair_mol, e->pair_mol)) || (rc = upload(e.get(), e->d_pp64, pp64)) ||
        (rc = upload(e.get(), e->d_pair_info, pinfo)) || (rc = upload(e.get(), e->d_pair_ps32, e->pair_ps32)) ||
        (rc = upload(e.get(), e->d_pair_np32, e->pair_np32)) ||
        (rc = upload(e.get(), e->d_pp32, pp32)) || (rc = upload(e.get(), e->d_sh_ao, e->sh_ao)) ||
        (rc = upload(e.get(), e->d_sh_nao, e->sh_nao)) || (rc = upload(e.get(), e->d_sh_mn, e->sh_mn)) ||
        (rc = upload(e.get(), e->d_sh_local, e->sh_local)) || (rc = upload(e.get(), e->d_sh_ns, e->sh_ns)) ||
        (rc = upload(e.get(), e->d_sh_po, e->sh_po)) || (rc = upload(e.get(), e->d_sh_np, e->sh_np)) ||
        (rc = upload(e.get(), e->d_sh_no, e->sh_no)) || (rc = upload(e.get(), e->d_sh_mol, e->sh_mol)) ||
        (rc = upload(e.get(), e->d_sh_mbase, e->sh_mbase)) || (rc = upload(e.get(), e->d_sh_qbase, e->sh_qbase)) ||
        (rc = upload(e.get(), e->d_sh_tbase, e->sh_tbase)) || (rc = upload(e.get(), e->d_tbase, e->mol_tbase)) ||
        (rc = upload(e.get(), e->d_sh_cf, e->sh_cf)) || (rc = upload(e.get(), e->d_ctr, e->sh_ctr)) ||
        (rc = upload(e.get(), e->d_alpha, e->alpha)) || (rc = upload(e.get(), e->d_coeff, e->coeff)) ||
        (rc = upload(e.get(), e->d_norm, e->norm)) || (rc = upload(e.get(), e->d_boys64, b64)) ||
        (rc = upload(e.get(), e->d_boys32, b32)) || (rc = upload(e.get(), e->d_mbase, e->mol_mbase)) ||
        (rc = upload(e.get(), e->d_mn, e->mol_N)) || (rc = upload(e.get(), e->d_reg, e->reg)) ||
        (rc = upload(e.get(), e->d_pab64, pab64)) || (rc = upload(e.get(), e->d_pab32, pab32)) ||
        (rc = upload(e.get(), e->d_sh_l, e->sh_l)) || (rc = upload(e.get(), e->d_pair_npx, npx)))
        return rc;
    const long long ne = e->n_elems, nq = e->n_qelems;
    e->d_Dt = e->mem.alloc<double>(ne);
    e->d_Da = e->mem.alloc<double>(ne);
    e->d_Db = e->mem.alloc<double>(ne);
    e->d_Mt = e->mem.alloc<double>(nq);
    e->d_Ma = e->mem.alloc<double>(nq);
    e->d_Mb = e->mem.alloc<double>(nq);
    const long long nt = e->n_telems;
    e->d_Tt = e->mem.alloc<double>(nt);
    e->d_Ta = e->mem.alloc<double>(nt);
    e->d_Tb = e->mem.alloc<double>(nt);
    // two words per element, for the fixed-point pairs of deterministic builds
    e->d_AJ = e->mem.alloc<double>(2 * nt * hfx_engine::ACC_COPIES);
    e->d_AKa = e->mem.alloc<double>(2 * nt * hfx_engine::ACC_COPIES);
    e->d_AKb = e->mem.alloc<double>(2 * nt * hfx_engine::ACC_COPIES);
    e->d_J = e->mem.alloc<double>(ne);
    e->d_Ka = e->mem.alloc<double>(ne);
    e->d_Kb = e->mem.alloc<double>(ne);
    e->d_list = e->mem.alloc<uint2>(e->list_cap);
    e->slots.assign(3 * N_COMBOS, 0);
    for (int k = 0; k < N_COMBOS; ++k) {
        const long long nb = (long long)e->cls_ids[combo_pc1(k)].size();
        const bool live = !e->cls_ids[combo_pc1(k)].empty() && !e->cls_ids[combo_pc2(k)].empty();
        e->slots[3 * k] = e->n_slots;
        e->slots[3 * k + 1] = e->n_slots + (live ? nb : 0);
        e->slots[3 * k + 2] = e->n_slots + (live ? 2 * nb : 0);
        e->n_slots = e->slots[3 * k + 2];
    }
    e->d_counts = e->mem.alloc<unsigned long long>(e->n_slots + 1);
    e->d_offs = e->mem.alloc<unsigned long long>(e->n_slots + 1);
    if ((rc = upload(e.get(), e->d_slots, e->slots))) return rc;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, e->scan_tmp_bytes, e->d_counts, e->d_offs,
                                            (int)(e->n_slots + 1), e->stream));
    e->d_scan_tmp = e->mem.alloc<char>(e->scan_tmp_bytes);
    e->d_stats = e->mem.alloc<unsigned long long>(6 * N_COMBOS);
    e->d_ranges = e->mem.alloc<long long>(6 * N_COMBOS);
    e->d_chunk_ctr = e->mem.alloc<unsigned int>(2 * N_COMBOS * hfx_engine::CHUNK_SLOTS);
    for (int m = 0; m < n_mol; ++m) e->max_N = std::max(e->max_N, e->mol_N[m]);
    if (!e->d_Tt || !e->d_Ta || !e->d_Tb || !e->d_Dt || !e->d_Da || !e->d_Db || !e->d_Mt || !e->d_Ma || !e->d_Mb || !e->d_AJ || !e->d_AKa || !e->d_AKb ||
        !e->d_J || !e->d_Ka || !e->d_Kb || !e->d_list || !e->d_counts || !e->d_offs ||
        !e->d_scan_tmp || !e->d_stats || !e->d_ranges || !e->d_chunk_ctr)
        return fail(HFX_ERR_NOMEM, "hipMalloc failed for the build buffers");

    // ---- Schwarz table, then the class order it implies ----
    {
        // Class lists in creation order are enough for the diagonal quartets.
        if ((rc = compute_schwarz(e.get()))) return rc;
        if ((rc = sort_pairs(e.get(), e->own_q))) return rc;
    }
    *out = e.release();
    return HFX_OK;
}

// The device pipeline of one build, from the staged densities (d_Dt, d_Da, d_Db) to the
// symmetrized outputs (d_J, d_Ka, d_Kb).  No host synchronization, no memset: it is
// captured into a hipGraph as is.  With P.timing, events are recorded after every class
// kernel group (ev_slot[2k + precision] = event index, -1 when absent).
// Screening of a build: shell-pair density maxima of the staged densities, the two
// counting/writing passes over every class and the per-class list ranges (d_ranges).
int enqueue_screen(hfx_engine* e, const BuildParams& P, hipStream_t s) {
    const bool uhf = P.uhf, mixed = P.mixed;

    // shell-pair density maxima
    {
        DmaxArgs d{};
        d.pair_a = e->d_pair_a;
        d.pair_b = e->d_pair_b;
        d.sh_ao = e->d_sh_ao;
        d.sh_nao = e->d_sh_nao;
        d.sh_mbase = e->d_sh_mbase;
        d.sh_mn = e->d_sh_mn;
        d.sh_local = e->d_sh_local;
        d.sh_qbase = e->d_sh_qbase;
        d.sh_ns = e->d_sh_ns;
        d.Dt = e->d_Dt;
        d.Da = e->d_Da;
        d.Db = uhf ? e->d_Db : nullptr;
        d.Mt = e->d_Mt;
        d.Ma = e->d_Ma;
        d.Mb = e->d_Mb;
        d.n_pairs = e->n_pairs;
        hipLaunchKernelGGL(k_dmax, dim3((e->n_pairs + 255) / 256), dim3(256), 0, s, d);
    }
    hipLaunchKernelGGL(k_fill_u64, dim3(std::max<long long>(1, std::min<long long>(1024, (e->n_slots + 256) / 256))),
                       dim3(256), 0, s, e->d_counts, (int)(e->n_slots + 1));
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(128), 0, s, e->d_stats, 6 * N_COMBOS);
    ScreenArgs a{};
    a.cls = e->d_screen_cls;
    a.pair_info = e->d_pair_info;
    a.pair_npx = e->d_pair_npx;
    a.pair_mol = e->d_pair_mol;
    a.pair_q = e->d_pair_q;
    a.sh_qbase = e->d_sh_qbase;
    a.sh_ns = e->d_sh_ns;
    a.Mt = e->d_Mt;
    a.Ma = e->d_Ma;
    a.Mb = uhf ? e->d_Mb : nullptr;
    a.ns_max = e->ns_max;
    a.tauS = P.tauS;
    a.tauD = P.tauD;
    a.cut = P.cut;
    a.mixed = mixed;
    a.nS_total = e->nS;
    a.list = e->d_list;
    a.counts = e->d_counts;
    a.offs = e->d_offs;
    a.stats = e->d_stats;
    const size_t row_bytes = (size_t)(uhf ? 4 : 2) * e->ns_max * sizeof(double);
    const int wpb = (int)std::min<size_t>(4, (60 * 1024) / row_bytes);
    if (wpb < 1) return fail(HFX_ERR_UNSUPPORTED, "molecule too large for the screening kernel's LDS rows");
    long long blocks = 0;
    for (int k = 0; k < N_COMBOS; ++k) blocks += (e->screen_nbra[k] + wpb - 1) / wpb;
    if (blocks > 0) {
        const dim3 sg((unsigned)blocks), sb(64 * wpb);
        hipLaunchKernelGGL(k_screen<0>, sg, sb, wpb * row_bytes, s, a);
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(e->d_scan_tmp, e->scan_tmp_bytes, e->d_counts, e->d_offs,
                                                (int)(e->n_slots + 1), s));
        if (P.hash)
            hipLaunchKernelGGL(k_screen<1>, sg, sb, wpb * row_bytes, s, a);
        else
            hipLaunchKernelGGL(k_screen<2>, sg, sb, wpb * row_bytes, s, a);
    } else {
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(e->d_scan_tmp, e->scan_tmp_bytes, e->d_counts, e->d_offs,
                                                (int)(e->n_slots + 1), s));
    }
    hipLaunchKernelGGL(k_ranges, dim3(1), dim3(64), 0, s, e->d_offs, e->d_slots, e->d_ranges);
    return HFX_OK;
}

// Static cost (arbitrary units) of one chunk kernel of class k in precision fp32 (1) or
// FP64 (0): the unscreened model FLOPs of the class at its mean unpruned primitive counts,
// split over the class's chunk kernels, plus a fixed launch cost.  In mixed mode most
// quartets run in FP32; the FP64 share is small but runs on an under-filled machine.
double chunk_cost(hfx_engine* e, int k, int fp32, bool mixed) {
    const int p1 = combo_pc1(k), p2 = combo_pc2(k);
    if (e->d_store && fp32 < 0) {  // stored-integral digest: the bytes it reads
        const int nch = std::max(1, eri_chunks(k, 0));
        const double blk = (double)ncart_h(pc_la(p1)) * ncart_h(pc_lb(p1)) * ncart_h(pc_la(p2)) * ncart_h(pc_lb(p2));
        return 1e-6 * (double)e->store_n[k] * (blk + 1.0) / nch + 1.0;
    }
    const std::vector<int>& b = e->cls_ids[p1];
    const std::vector<int>& c = e->cls_ids[p2];
    if (b.empty() || c.empty()) return 0.0;
    auto mean_np = [&](const std::vector<int>& ids) {
        double t = 0.0;
        for (int id : ids) t += e->pair_nprim[id];
        return t / (double)ids.size();
    };
    const double nb = (double)b.size(), nc = (double)c.size();
    const double quartets = p1 == p2 ? 0.5 * nb * (nb + 1.0) : nb * nc;
    const int la = pc_la(p1), lb = pc_lb(p1), lc = pc_la(p2), ld = pc_lb(p2);
    const double per = mean_np(b) * mean_np(c) * model_flop_h(la, lb, lc, ld) +
                       12.0 * ncart_h(la) * ncart_h(lb) * ncart_h(lc) * ncart_h(ld);
    const double share = !mixed ? 1.0 : (fp32 ? 0.9 : 0.25);
    const int nch = std::max(1, eri_chunks(k, fp32));
    return 1e-9 * quartets * per * share / nch + 1.0;
}

int enqueue_build(hfx_engine* e, const BuildParams& P, hipStream_t s, int* ev_slot, int& n_ev) {
    const bool uhf = P.uhf, mixed = P.mixed, timing = P.timing;
    const long long ne = e->n_elems;
    n_ev = 0;
    if (!P.stored) {
        const int rc = enqueue_screen(e, P, s);
        if (rc) return rc;
    }
    const long long nt = e->n_telems;
    {
        const int g = e->n_cu * 4;
        const long long nc = nt * hfx_engine::ACC_COPIES * (P.det ? 2 : 1);
        hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, s, e->d_AJ, nc, 0.0);
        hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, s, e->d_AKa, nc, 0.0);
        if (uhf) hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, s, e->d_AKb, nc, 0.0);
        const dim3 gp(std::max(1, std::min(e->n_cu * 4, (int)((nt / e->n_mol + 255) / 256))), e->n_mol);
        hipLaunchKernelGGL(k_pack_tri, gp, dim3(256), 0, s, e->d_Dt, e->d_Tt, e->d_mbase, e->d_tbase, e->d_mn);
        hipLaunchKernelGGL(k_pack_tri, gp, dim3(256), 0, s, e->d_Da, e->d_Ta, e->d_mbase, e->d_tbase, e->d_mn);
        if (uhf) hipLaunchKernelGGL(k_pack_tri, gp, dim3(256), 0, s, e->d_Db, e->d_Tb, e->d_mbase, e->d_tbase, e->d_mn);
    }
    EriArgs ea = eri_args(e);
    ea.list = e->d_list;
    ea.Dt = e->d_Tt;
    ea.Da = e->d_Ta;
    ea.Db = uhf ? e->d_Tb : nullptr;
    ea.AJ = e->d_AJ;
    ea.AKa = e->d_AKa;
    ea.AKb = e->d_AKb;
    ea.acc_stride = P.det ? 2 * nt : nt;
    ea.n_copies = hfx_engine::ACC_COPIES;
    ea.fixed = P.det ? 1 : 0;
    ea.tauS = P.tauS;
    ea.out = nullptr;
    ea.out_off = nullptr;
    if (P.stored) ea.list = e->d_slist;
    const int grid = eri_grid(e);
    hipLaunchKernelGGL(k_fill_u32, dim3(4), dim3(256), 0, s, e->d_chunk_ctr, 2 * N_COMBOS * hfx_engine::CHUNK_SLOTS);
    // Longer runs flush a stored build's exchange rows less often; in the direct build they leave
    // a long tail on the heaviest-first lists (ssss FP32 2.9 -> 7.5 ms alone at 8192).
    ea.chunk_len = e->chunk_len ? e->chunk_len : (P.stored ? hfx_engine::STORED_CHUNK_LEN : hfx_engine::CHUNK_LEN);
    ea.stagger = e->stagger;
    // LDS exchange rows of a class kernel: NA + NB rows of max_N doubles (twice for UHF), next
    // to the kernel's Boys slice, within 64 KB per block; classes that do not fit run without.
    auto tile_cols = [&](int k, int fp32) {
        if (P.det) return 0;  // LDS atomics would add the exchange rows in a run-dependent order
        const int p1 = combo_pc1(k), p2 = combo_pc2(k);
        const int la = pc_la(p1), lb = pc_lb(p1), L = la + lb + pc_la(p2) + pc_lb(p2);
        const size_t boys = P.stored ? 16 : (size_t)md::boys_points(!fp32) * md::boys_slice_nord(L, !fp32) * (fp32 ? 4 : 8);
        const size_t rows = (size_t)(uhf ? 2 : 1) * (ncart_h(la) + ncart_h(lb)) * e->max_N * sizeof(double);
        const size_t mask = (size_t)((e->ns_max + 31) / 32) * sizeof(unsigned int);
        return boys + rows + mask + 1024 <= 64 * 1024 ? e->max_N : 0;
    };
    // Chunk kernels go to the streams by longest-processing-time-first on a static cost
    // estimate; with per-class timing everything stays on the build stream, in class order.
    const int n_str = timing ? 1 : e->n_side;
    std::vector<hipStream_t> cs[2 * N_COMBOS];
    {
        // one item per launch: the chunks [b, e) of one translation unit of a class
        struct Item {
            double cost;
            int slot, b, e;
        };
        std::vector<Item> items;
        int pb[64], pe[64];
        for (int k = 0; k < N_COMBOS; ++k) {
            for (int f = 0; f < (mixed ? 2 : 1); ++f) {
                const int nch = eri_chunks(k, f);
                const int np = eri_parts(k, f, pb, pe);
                cs[2 * k + f].assign(std::max(1, nch), s);
                const double c = chunk_cost(e, k, P.stored ? -1 : f, mixed);
                for (int x = 0; x < np; ++x)
                    if (pe[x] > pb[x]) items.push_back({c * (pe[x] - pb[x]), 2 * k + f, pb[x], pe[x]});
            }
        }
        std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
        double load[hfx_engine::N_SIDE] = {};
        for (const Item& it : items) {
            int best = 0;
            for (int i = 1; i < n_str; ++i)
                if (load[i] < load[best]) best = i;
            load[best] += it.cost;
            // slot 0 is the stream the build is enqueued on (the engine's own for hfx_jk, the
            // caller's for hfx_jk_async); slots 1.. are the engine's side streams, forked from it
            for (int ch = it.b; ch < it.e; ++ch) cs[it.slot][ch] = best == 0 ? s : e->side[best];
        }
    }
    if (n_str > 1) {
        HIPCHK(hipEventRecord(e->ev_fork, s));
        for (int i = 1; i < n_str; ++i) HIPCHK(hipStreamWaitEvent(e->side[i], e->ev_fork, 0));
    }
    if (timing) HIPCHK(hipEventRecord(e->cev[n_ev++], s));
    for (int k = 0; k < N_COMBOS; ++k) {
        ev_slot[2 * k] = ev_slot[2 * k + 1] = -1;
        const int p1 = combo_pc1(k), p2 = combo_pc2(k);
        if (e->cls_ids[p1].empty() || e->cls_ids[p2].empty()) continue;
        if (P.stored && e->store_n[k] == 0) continue;
        ea.range = (P.stored ? e->d_sranges : e->d_ranges) + 2 * k;
        ea.stored = P.stored ? e->d_store + e->store_base[k] : nullptr;
        ea.chunk_ctr = e->d_chunk_ctr + (2 * k) * hfx_engine::CHUNK_SLOTS;
        ea.tile_n = tile_cols(k, 0);
        launch_eri(k, 0, ea, grid, cs[2 * k].data());
        if (P.stored) {
            // the ket-side Coulomb terms from the ket-major copy, on the stream of chunk 0
            EriArgs ta = ea;
            ta.list = e->d_tlist;
            ta.stored = e->d_storeT + e->store_base[k];
            jket_fn(k)(ta, grid, cs[2 * k][0]);
        }
        if (timing) {
            HIPCHK(hipEventRecord(e->cev[n_ev], s));
            ev_slot[2 * k] = n_ev++;
        }
        if (mixed) {
            ea.range = e->d_ranges + 2 * N_COMBOS + 2 * k;
            ea.chunk_ctr = e->d_chunk_ctr + (2 * k + 1) * hfx_engine::CHUNK_SLOTS;
            ea.tile_n = tile_cols(k, 1);
            launch_eri(k, 1, ea, grid, cs[2 * k + 1].data());
            if (timing) {
                HIPCHK(hipEventRecord(e->cev[n_ev], s));
                ev_slot[2 * k + 1] = n_ev++;
            }
        }
    }
    if (n_str > 1) {
        for (int i = 1; i < n_str; ++i) {
            HIPCHK(hipEventRecord(e->ev_join[i], e->side[i]));
            HIPCHK(hipStreamWaitEvent(s, e->ev_join[i], 0));
        }
    }
    {
        const dim3 g(std::max(1, std::min(e->n_cu * 4, (int)((ne + 255) / 256))), e->n_mol);
        const int cp = hfx_engine::ACC_COPIES;
        auto symm = P.det ? k_symm_fixed : k_symm;
        const long long st = ea.acc_stride;
        hipLaunchKernelGGL(symm, g, dim3(256), 0, s, e->d_AJ, st, cp, e->d_J, e->d_mbase, e->d_tbase, e->d_mn);
        hipLaunchKernelGGL(symm, g, dim3(256), 0, s, e->d_AKa, st, cp, e->d_Ka, e->d_mbase, e->d_tbase, e->d_mn);
        if (uhf)
            hipLaunchKernelGGL(symm, g, dim3(256), 0, s, e->d_AKb, st, cp, e->d_Kb, e->d_mbase, e->d_tbase, e->d_mn);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipGetLastError());
    return HFX_OK;
}

// Class lists sorted by the Schwarz table Q (the caller's, or the engine's own when NULL).
int use_q(hfx_engine* e, const double* Q, const char* who) {
    if (Q) {
        if (!finite_all(Q, e->n_qelems)) return fail(HFX_ERR_NONFINITE, std::string(who) + ": non-finite Q");
        if (e->sort_q.size() != (size_t)e->n_qelems ||
            std::memcmp(e->sort_q.data(), Q, e->n_qelems * sizeof(double)) != 0) {
            std::vector<double> q(Q, Q + e->n_qelems);
            return sort_pairs(e, q);
        }
    } else if (e->sort_q != e->own_q) {
        return sort_pairs(e, e->own_q);
    }
    return HFX_OK;
}

// The engine's captured build for parameters P (captured again when they change).
int ensure_graph(hfx_engine* e, const BuildParams& P) {
    if (e->graph_exec && e->graph_key == P) return HFX_OK;
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
    int ev_slot[2 * N_COMBOS];
    int n_ev = 0;
    hipStream_t s = e->stream;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_build(e, P, s, ev_slot, n_ev);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess) return fail(HFX_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
    e->graph = g;
    HIPCHK(hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0));
    e->graph_key = P;
    return HFX_OK;
}

// Checks shared by hfx_jk and hfx_jk_async; fills the build parameters.
int jk_check(hfx_engine* e, const hfx_jk_args* args, double* J, double* Ka, double* Kb, bool want_hash,
             const char* who, BuildParams& P) {
    if (!e || !args || !args->dD_tot || !args->dD_alpha || !J || !Ka)
        return fail(HFX_ERR_INVALID, std::string(who) + ": null argument");
    const bool uhf = args->dD_beta != nullptr;
    if (uhf && !Kb) return fail(HFX_ERR_INVALID, std::string(who) + ": dD_beta given without Kb");
    if (args->precision != HFX_FP64 && args->precision != HFX_MIXED)
        return fail(HFX_ERR_INVALID, std::string(who) + ": unknown precision");
    const bool stored = (args->flags & HFX_JK_STORED) != 0;
    if (stored && !e->d_store)
        return fail(HFX_ERR_INVALID, std::string(who) + ": HFX_JK_STORED without stored integrals (call hfx_store_eris first)");
    if (stored && args->precision != HFX_FP64)
        return fail(HFX_ERR_UNSUPPORTED, std::string(who) + ": the stored integrals are FP64; HFX_JK_STORED needs HFX_FP64");
    if (!(args->schwarz_thresh >= 0.0) || !(args->density_thresh >= 0.0) || !(args->fp32_cut >= 0.0))
        return fail(HFX_ERR_INVALID, std::string(who) + ": negative or NaN threshold");
    P = BuildParams{args->schwarz_thresh, args->density_thresh, args->fp32_cut,
                    args->precision == HFX_MIXED && args->fp32_cut > 0.0, uhf,
                    (args->flags & HFX_JK_CLASS_TIMING) != 0};
    P.hash = want_hash;
    P.det = (args->flags & HFX_JK_DETERMINISTIC) != 0;
    if (stored) {  // no density test, no FP32 split: the stored list as it is
        P.tauD = 0.0;
        P.cut = 0.0;
        P.mixed = false;
        P.hash = false;
        P.stored = true;
    }
    return HFX_OK;
}

int jk_impl(hfx_engine* e, const hfx_jk_args* args, double* J, double* Ka, double* Kb, hfx_jk_stats* st) {
    BuildParams P;
    {
        const int rc = jk_check(e, args, J, Ka, Kb, st != nullptr && !(args && (args->flags & HFX_JK_NO_STATS)),
                                "hfx_jk", P);
        if (rc) return rc;
    }
    const bool uhf = P.uhf, stored = P.stored;
    const bool devp = (args->flags & HFX_JK_DEVICE_PTRS) != 0;
    const long long ne = e->n_elems;
    if (!devp) {
        if (!finite_all(args->dD_tot, ne) || !finite_all(args->dD_alpha, ne) || (uhf && !finite_all(args->dD_beta, ne)))
            return fail(HFX_ERR_NONFINITE, "hfx_jk: non-finite density");
    }
    if (!stored) {  // the stored list does not depend on the class order
        const int rc = use_q(e, args->Q, "hfx_jk");
        if (rc) return rc;
    }
    HIPCHK(hipSetDevice(e->device));
    hipStream_t s = e->stream;
    const hipMemcpyKind h2d = devp ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind d2h = devp ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIPCHK(hipMemcpyAsync(e->d_Dt, args->dD_tot, ne * sizeof(double), h2d, s));
    HIPCHK(hipMemcpyAsync(e->d_Da, args->dD_alpha, ne * sizeof(double), h2d, s));
    if (uhf) HIPCHK(hipMemcpyAsync(e->d_Db, args->dD_beta, ne * sizeof(double), h2d, s));
    HIPCHK(hipEventRecord(e->ev0, s));
    int ev_slot[2 * N_COMBOS];
    int n_ev = 0;
    const bool timing = P.timing;
    if ((args->flags & HFX_JK_USE_GRAPH) && !timing) {
        const int rc = ensure_graph(e, P);
        if (rc) return rc;
        HIPCHK(hipGraphLaunch(e->graph_exec, s));
    } else {
        const int rc = enqueue_build(e, P, s, ev_slot, n_ev);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(e->ev1, s));
    HIPCHK(hipMemcpyAsync(J, e->d_J, ne * sizeof(double), d2h, s));
    HIPCHK(hipMemcpyAsync(Ka, e->d_Ka, ne * sizeof(double), d2h, s));
    if (uhf) HIPCHK(hipMemcpyAsync(Kb, e->d_Kb, ne * sizeof(double), d2h, s));
    long long rng[4 * N_COMBOS];
    unsigned long long cnt[2 * N_COMBOS], stv[6 * N_COMBOS];
    const bool want_stats = st && !(args->flags & HFX_JK_NO_STATS) && !stored;
    if (want_stats) {
        HIPCHK(hipMemcpyAsync(rng, e->d_ranges, sizeof(rng), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(stv, e->d_stats, sizeof(stv), hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    if (st) {
        std::memset(st, 0, sizeof(*st));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        st->device_ms = ms;
        if (want_stats) {
            for (int k = 0; k < N_COMBOS; ++k) {
                cnt[2 * k] = (unsigned long long)(rng[2 * k + 1] - rng[2 * k]);
                cnt[2 * k + 1] = (unsigned long long)(rng[2 * N_COMBOS + 2 * k + 1] - rng[2 * N_COMBOS + 2 * k]);
            }
            double flop = 0.0, xflop = 0.0;
            for (int k = 0; k < N_COMBOS; ++k) {
                const long long c = (long long)(cnt[2 * k] + cnt[2 * k + 1]);
                st->class_count[k] = c;
                st->n_quartets += c;
                st->n_fp32 += (long long)cnt[2 * k + 1];
                st->hash_sum += stv[6 * k];
                st->hash_xor ^= stv[6 * k + 1];
                const int p1 = combo_pc1(k), p2 = combo_pc2(k);
                const int la = pc_la(p1), lb = pc_lb(p1), lc = pc_la(p2), ld = pc_lb(p2);
                const double f = model_flop_h(la, lb, lc, ld);
                const double dig = 12.0 * ncart_h(la) * ncart_h(lb) * ncart_h(lc) * ncart_h(ld);
                st->class_gflop64[k] = ((double)stv[6 * k + 2] * f + dig * (double)cnt[2 * k]) * 1e-9;
                st->class_gflop32[k] = ((double)stv[6 * k + 3] * f + dig * (double)cnt[2 * k + 1]) * 1e-9;
                st->class_exec_gflop64[k] = ((double)stv[6 * k + 4] * f + dig * (double)cnt[2 * k]) * 1e-9;
                st->class_exec_gflop32[k] = ((double)stv[6 * k + 5] * f + dig * (double)cnt[2 * k + 1]) * 1e-9;
                flop += st->class_gflop64[k] + st->class_gflop32[k];
                xflop += st->class_exec_gflop64[k] + st->class_exec_gflop32[k];
            }
            st->model_gflop = flop;
            st->exec_gflop = xflop;
        }
        if (stored && !(args->flags & HFX_JK_NO_STATS)) {  // the stored list; digest FLOPs only
            double flop = 0.0;
            for (int k = 0; k < N_COMBOS; ++k) {
                const int p1 = combo_pc1(k), p2 = combo_pc2(k);
                const double dig = 12.0 * ncart_h(pc_la(p1)) * ncart_h(pc_lb(p1)) * ncart_h(pc_la(p2)) * ncart_h(pc_lb(p2));
                st->class_count[k] = e->store_n[k];
                st->n_quartets += e->store_n[k];
                st->class_gflop64[k] = st->class_exec_gflop64[k] = dig * (double)e->store_n[k] * 1e-9;
                flop += st->class_gflop64[k];
            }
            st->model_gflop = st->exec_gflop = flop;
        }
        if (timing) {
            int prev = 0;
            for (int x = 0; x < 2 * N_COMBOS; ++x) {
                if (ev_slot[x] < 0) continue;
                float cms = 0.f;
                HIPCHK(hipEventElapsedTime(&cms, e->cev[prev], e->cev[ev_slot[x]]));
                (x % 2 == 0 ? st->class_ms64 : st->class_ms32)[x / 2] = cms;
                prev = ev_slot[x];
            }
        }
    }
    return HFX_OK;
}

// hfx_jk_async: the build of jk_impl enqueued on the caller's stream, nothing waited for.
int jk_async_impl(hfx_engine* e, const hfx_jk_args* args, double* J, double* Ka, double* Kb, hipStream_t s) {
    BuildParams P;
    {
        const int rc = jk_check(e, args, J, Ka, Kb, false, "hfx_jk_async", P);
        if (rc) return rc;
    }
    if (!(args->flags & HFX_JK_DEVICE_PTRS))
        return fail(HFX_ERR_INVALID, "hfx_jk_async: needs HFX_JK_DEVICE_PTRS (device-resident densities and outputs)");
    if (P.timing) return fail(HFX_ERR_INVALID, "hfx_jk_async: HFX_JK_CLASS_TIMING needs the blocking hfx_jk");
    HIPCHK(hipSetDevice(e->device));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(s, &cap));
    const bool capturing = cap != hipStreamCaptureStatusNone;
    if (!P.stored) {
        const bool resort = args->Q ? (e->sort_q.size() != (size_t)e->n_qelems ||
                                       std::memcmp(e->sort_q.data(), args->Q, e->n_qelems * sizeof(double)) != 0)
                                    : (e->sort_q != e->own_q);
        if (resort && capturing)
            return fail(HFX_ERR_INVALID, "hfx_jk_async: a new Schwarz table re-sorts the pair lists synchronously; "
                                         "make one call with it before capturing");
        const int rc = use_q(e, args->Q, "hfx_jk_async");
        if (rc) return rc;
    }
    const bool graph = (args->flags & HFX_JK_USE_GRAPH) != 0;
    if (graph && capturing)
        return fail(HFX_ERR_INVALID, "hfx_jk_async: HFX_JK_USE_GRAPH on a capturing stream (capture the call without it)");
    if (graph) {
        const int rc = ensure_graph(e, P);
        if (rc) return rc;
    }
    if (!capturing) {  // after the engine's earlier work
        HIPCHK(hipEventRecord(e->ev_async, e->stream));
        HIPCHK(hipStreamWaitEvent(s, e->ev_async, 0));
    }
    const long long ne = e->n_elems;
    const hipMemcpyKind d2d = hipMemcpyDeviceToDevice;
    HIPCHK(hipMemcpyAsync(e->d_Dt, args->dD_tot, ne * sizeof(double), d2d, s));
    HIPCHK(hipMemcpyAsync(e->d_Da, args->dD_alpha, ne * sizeof(double), d2d, s));
    if (P.uhf) HIPCHK(hipMemcpyAsync(e->d_Db, args->dD_beta, ne * sizeof(double), d2d, s));
    if (graph) {
        HIPCHK(hipGraphLaunch(e->graph_exec, s));
    } else {
        int ev_slot[2 * N_COMBOS];
        int n_ev = 0;
        const int rc = enqueue_build(e, P, s, ev_slot, n_ev);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(J, e->d_J, ne * sizeof(double), d2d, s));
    HIPCHK(hipMemcpyAsync(Ka, e->d_Ka, ne * sizeof(double), d2d, s));
    if (P.uhf) HIPCHK(hipMemcpyAsync(Kb, e->d_Kb, ne * sizeof(double), d2d, s));
    if (!capturing) {  // the engine's later calls wait for this build
        HIPCHK(hipEventRecord(e->ev_async, s));
        HIPCHK(hipStreamWaitEvent(e->stream, e->ev_async, 0));
    }
    return HFX_OK;
}

// Stored integrals (IntegralEngine::electronRepulsionTensor, src/IntegralEngine.cpp:635-692): the
// canonical quartets with Q(a,b)Q(c,d) >= tau_S, found by the screening pass with every density
// test passing (zero density, tau_D = 0), copied into a list of their own, and their normalized
// blocks computed once by the store mode of the FP64 class kernels, in the layout the stored
// digest reads (element-major groups of 64 entries).  Replaces any earlier store.
// Device memory owned by one scope (temporary buffers of a set-up step).
struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes) {
        if (hipMalloc(&p, bytes) != hipSuccess) p = nullptr;
    }
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// Ket-major copy of the stored integrals: per class, the entries sorted by ket pair (stable, so
// the bras of one ket stay in list order), the list {ket pair, bra pair} and the blocks
// transposed (element cd*NAB + ab of entry r' at ((r'/64) NAB NCD + cd NAB + ab) 64 + r' mod 64).
__global__ void k_ket_keys(const uint2* __restrict__ list, long long n, unsigned int* keys, unsigned int* idx) {
    for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < n; r += (long long)gridDim.x * blockDim.x) {
        keys[r] = list[r].y;
        idx[r] = (unsigned int)r;
    }
}
__global__ void k_ket_transpose(const uint2* __restrict__ list, const unsigned int* __restrict__ perm, long long n,
                                int nab, int ncd, const double* __restrict__ src, uint2* tlist, double* dst) {
    const long long blk = (long long)nab * ncd;
    const long long total = n * blk;
    for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < total;
         x += (long long)gridDim.x * blockDim.x) {
        // consecutive threads: consecutive destination entries of one transposed element
        const long long et = x / n, rt = x - et * n;
        const long long r = perm[rt];
        const long long cd = et / nab, ab = et - cd * nab;
        const long long es = ab * ncd + cd;
        dst[((rt >> 6) * blk + et) * 64 + (rt & 63)] = src[((r >> 6) * blk + es) * 64 + (r & 63)];
        if (et == 0) tlist[rt] = make_uint2(list[r].y, list[r].x);
    }
}

int store_ket_major(hfx_engine* e, const long long* sr, hipStream_t s) {
    long long nmax = 0;
    for (int k = 0; k < N_COMBOS; ++k) nmax = std::max(nmax, e->store_n[k]);
    if (nmax == 0) return HFX_OK;
    if (nmax >= (1LL << 32)) return fail(HFX_ERR_UNSUPPORTED, "hfx_store_eris: more than 2^32 quartets in one class");
    DeviceBuffer keys(4 * nmax), keys2(4 * nmax), idx(4 * nmax), idx2(4 * nmax);
    size_t tmp_bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (unsigned int*)keys.p, (unsigned int*)keys2.p,
                                              (unsigned int*)idx.p, (unsigned int*)idx2.p, (int)nmax, 0, 32, s));
    DeviceBuffer tmp(std::max<size_t>(tmp_bytes, 16));
    if (!keys.p || !keys2.p || !idx.p || !idx2.p || !tmp.p) return fail(HFX_ERR_NOMEM, "hfx_store_eris: hipMalloc failed");
    const int g = e->n_cu * 8;
    for (int k = 0; k < N_COMBOS; ++k) {
        const long long n = e->store_n[k];
        if (n == 0) continue;
        const int p1 = combo_pc1(k), p2 = combo_pc2(k);
        const int nab = ncart_h(pc_la(p1)) * ncart_h(pc_lb(p1)), ncd = ncart_h(pc_la(p2)) * ncart_h(pc_lb(p2));
        const uint2* list = e->d_slist + sr[2 * k];
        hipLaunchKernelGGL(k_ket_keys, dim3(g), dim3(256), 0, s, list, n, (unsigned int*)keys.p, (unsigned int*)idx.p);
        size_t tb = tmp_bytes;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, (unsigned int*)keys.p, (unsigned int*)keys2.p,
                                                  (unsigned int*)idx.p, (unsigned int*)idx2.p, (int)n, 0, 32, s));
        hipLaunchKernelGGL(k_ket_transpose, dim3(g), dim3(256), 0, s, list, (const unsigned int*)idx2.p, n, nab, ncd,
                           (const double*)(e->d_store + e->store_base[k]), e->d_tlist + sr[2 * k],
                           e->d_storeT + e->store_base[k]);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));  // the temporary buffers are freed on return
    return HFX_OK;
}

int store_impl(hfx_engine* e, const double* Q, double tauS, int64_t* n_out, int64_t* bytes_out) {
    if (!e) return fail(HFX_ERR_INVALID, "hfx_store_eris: null engine");
    if (!(tauS >= 0.0) || !std::isfinite(tauS)) return fail(HFX_ERR_INVALID, "hfx_store_eris: bad schwarz_thresh");
    {
        const int rc = use_q(e, Q, "hfx_store_eris");
        if (rc) return rc;
    }
    HIPCHK(hipSetDevice(e->device));
    e->release_store();
    hipStream_t s = e->stream;
    const long long ne = e->n_elems;
    const int g = e->n_cu * 4;
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, s, e->d_Dt, ne, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, s, e->d_Da, ne, 0.0);
    BuildParams P{tauS, 0.0, 0.0, false, false, false};
    P.hash = false;
    {
        const int rc = enqueue_screen(e, P, s);
        if (rc) return rc;
    }
    long long rng[4 * N_COMBOS];
    HIPCHK(hipMemcpyAsync(rng, e->d_ranges, sizeof(rng), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    long long n_tot = 0, elems = 0, sr[2 * N_COMBOS];
    for (int k = 0; k < N_COMBOS; ++k) {
        const int p1 = combo_pc1(k), p2 = combo_pc2(k);
        const long long blk = (long long)ncart_h(pc_la(p1)) * ncart_h(pc_lb(p1)) * ncart_h(pc_la(p2)) * ncart_h(pc_lb(p2));
        const long long n = rng[2 * k + 1] - rng[2 * k];
        e->store_n[k] = n;
        e->store_base[k] = elems;
        sr[2 * k] = n_tot;
        sr[2 * k + 1] = n_tot + n;
        n_tot += n;
        elems += (n + 63) / 64 * 64 * blk;
    }
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    // the blocks twice (bra-major, and the ket-major copy of the Coulomb pass) and the lists
    const double need = 2.0 * (double)elems * sizeof(double) + 2.0 * (double)n_tot * sizeof(uint2);
    if (need + (1ull << 30) > (double)free_b) {
        e->release_store();
        return fail(HFX_ERR_NOMEM, "hfx_store_eris: the stored integrals need " + std::to_string(need / 1e9) +
                                       " GB, " + std::to_string(free_b / 1e9) + " GB free");
    }
    if (hipMalloc(&e->d_slist, std::max<long long>(1, n_tot) * sizeof(uint2)) != hipSuccess ||
        hipMalloc(&e->d_sranges, sizeof(sr)) != hipSuccess ||
        hipMalloc(&e->d_store, std::max<long long>(1, elems) * sizeof(double)) != hipSuccess ||
        hipMalloc(&e->d_tlist, std::max<long long>(1, n_tot) * sizeof(uint2)) != hipSuccess ||
        hipMalloc(&e->d_storeT, std::max<long long>(1, elems) * sizeof(double)) != hipSuccess) {
        e->release_store();
        return fail(HFX_ERR_NOMEM, "hfx_store_eris: hipMalloc failed");
    }
    e->store_elems = elems;
    HIPCHK(hipMemcpyAsync(e->d_sranges, sr, sizeof(sr), hipMemcpyHostToDevice, s));
    for (int k = 0; k < N_COMBOS; ++k)
        if (e->store_n[k] > 0)
            HIPCHK(hipMemcpyAsync(e->d_slist + sr[2 * k], e->d_list + rng[2 * k], e->store_n[k] * sizeof(uint2),
                                  hipMemcpyDeviceToDevice, s));
    EriArgs a = eri_args(e);
    a.list = e->d_slist;
    a.out_off = nullptr;
    a.tauS = tauS;
    for (int k = 0; k < N_COMBOS; ++k) {
        if (e->store_n[k] == 0) continue;
        if (eri_chunks(k, 0) < 0 || !jket_fn(k)) {
            e->release_store();
            return fail(HFX_ERR_UNSUPPORTED, "ERI kernels missing for class " + std::to_string(k));
        }
        a.range = e->d_sranges + 2 * k;
        a.out = e->d_store + e->store_base[k];
        std::vector<hipStream_t> cs(std::max(1, eri_chunks(k, 0)), s);
        launch_eri(k, 0, a, eri_grid(e), cs.data());
    }
    HIPCHK(hipGetLastError());
    {
        const int rc = store_ket_major(e, sr, s);
        if (rc) {
            e->release_store();
            return rc;
        }
    }
    HIPCHK(hipStreamSynchronize(s));
    e->store_tauS = tauS;
    if (n_out) *n_out = n_tot;
    if (bytes_out) *bytes_out = (int64_t)need;
    return HFX_OK;
}

// The RHF gradient: screening of the J/K build on (P, P/2), the two-electron kernels on
// the surviving quartets, the one-electron kernel on every shell pair, the reduction of
// the per-block rows; the nuclear repulsion term is added on the host.
// Static cost (arbitrary units) of the two-electron gradient of class k: its unscreened
// quartets at the mean primitive counts of its pair classes, times the model FLOPs of the three
// raised-shell ERI classes a derivative needs.
double grad_cost(hfx_engine* e, int k) {
    const int p1 = combo_pc1(k), p2 = combo_pc2(k);
    const std::vector<int>& b = e->cls_ids[p1];
    const std::vector<int>& c = e->cls_ids[p2];
    if (b.empty() || c.empty()) return 0.0;
    auto mean_np = [&](const std::vector<int>& ids) {
        double t = 0.0;
        for (int id : ids) t += e->pair_nprim[id];
        return t / (double)ids.size();
    };
    const double nb = (double)b.size(), nc = (double)c.size();
    const double quartets = p1 == p2 ? 0.5 * nb * (nb + 1.0) : nb * nc;
    const int la = pc_la(p1), lb = pc_lb(p1), lc = pc_la(p2), ld = pc_lb(p2);
    const double f = model_flop_h(la + 1, lb, lc, ld) + model_flop_h(la, lb + 1, lc, ld) + model_flop_h(la, lb, lc + 1, ld);
    return 1e-9 * quartets * mean_np(b) * mean_np(c) * f + 1.0;
}

int grad_impl(hfx_engine* e, const hfx_grad_args* a, double* grad, double* parts, hfx_grad_stats* st) {
    if (!e || !a || !grad || !a->P || !a->W || a->n_atoms <= 0 || !a->Z || !a->xyz)
        return fail(HFX_ERR_INVALID, "hfx_rhf_gradient: null argument");
    if (e->n_mol != 1) return fail(HFX_ERR_INVALID, "hfx_rhf_gradient: single-molecule engines only");
    if (a->precision != HFX_FP64 && a->precision != HFX_MIXED)
        return fail(HFX_ERR_INVALID, "hfx_rhf_gradient: unknown precision");
    const bool devp = (a->flags & HFX_JK_DEVICE_PTRS) != 0;
    const long long ne = e->n_elems;
    const int nat = a->n_atoms, n3 = 3 * nat;
    if (!devp && (!finite_all(a->P, ne) || !finite_all(a->W, ne)))
        return fail(HFX_ERR_NONFINITE, "hfx_rhf_gradient: non-finite density");
    if (!finite_all(a->Z, nat) || !finite_all(a->xyz, n3)) return fail(HFX_ERR_NONFINITE, "hfx_rhf_gradient: atoms");
    std::vector<int> sat(e->nS);
    for (int sh = 0; sh < e->nS; ++sh) {
        if (a->shell_atom) {
            sat[sh] = a->shell_atom[sh];
        } else {
            sat[sh] = -1;
            for (int A = 0; A < nat && sat[sh] < 0; ++A)
                if (e->sh_ctr[3 * sh] == a->xyz[3 * A] && e->sh_ctr[3 * sh + 1] == a->xyz[3 * A + 1] &&
                    e->sh_ctr[3 * sh + 2] == a->xyz[3 * A + 2])
                    sat[sh] = A;
        }
        if (sat[sh] < 0 || sat[sh] >= nat) return fail(HFX_ERR_INVALID, "hfx_rhf_gradient: shell not on an atom");
    }
    // Engine-owned gradient buffers, grown on demand.
    const int grid = e->n_cu * 8;
    const long long stride = ((3LL * n3 + 31) / 32) * 32;
    if (!e->d_sh_atom) {
        e->d_sh_atom = e->mem.alloc<int>(e->nS);
        if (!e->d_sh_atom) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
    }
    if (nat > e->atom_cap) {
        e->d_atom_z = e->mem.alloc<double>(nat);
        e->d_atom_xyz = e->mem.alloc<double>(3LL * nat);
        e->d_gout = e->mem.alloc<double>(3LL * n3);
        if (!e->d_atom_z || !e->d_atom_xyz || !e->d_gout) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
        e->atom_cap = nat;
    }
    if (stride * grid > e->gpart_cap) {
        e->d_gpart = e->mem.alloc<double>(stride * grid);
        if (!e->d_gpart) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
        e->gpart_cap = stride * grid;
    }
    HIPCHK(hipSetDevice(e->device));
    hipStream_t s = e->stream;
    const hipMemcpyKind h2d = devp ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIPCHK(hipMemcpyAsync(e->d_sh_atom, sat.data(), e->nS * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(e->d_atom_z, a->Z, nat * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(e->d_atom_xyz, a->xyz, n3 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(e->d_Dt, a->P, ne * sizeof(double), h2d, s));
    HIPCHK(hipMemcpyAsync(e->d_Db, a->W, ne * sizeof(double), h2d, s));
    hipLaunchKernelGGL(k_scale, dim3(std::max(1, std::min(e->n_cu * 4, (int)((ne + 255) / 256)))), dim3(256), 0, s,
                       e->d_Dt, e->d_Da, ne, 0.5);
    hipLaunchKernelGGL(k_fill, dim3(std::max(1, std::min(e->n_cu * 4, (int)((stride * grid + 255) / 256)))), dim3(256),
                       0, s, e->d_gpart, stride * grid, 0.0);
    // Q order of the engine's own table
    if (e->sort_q != e->own_q) {
        const int rc = sort_pairs(e, e->own_q);
        if (rc) return rc;
    }
    const bool mixed = a->precision == HFX_MIXED && a->fp32_cut > 0.0;
    const BuildParams bp{a->schwarz_thresh, a->density_thresh, a->fp32_cut, mixed, false, false};
    {
        const int rc = enqueue_screen(e, bp, s);
        if (rc) return rc;
    }
    GradArgs g{};
    g.pair_a = e->d_pair_a;
    g.pair_b = e->d_pair_b;
    g.pair_ps = e->d_pair_ps;
    g.pair_np = e->d_pair_np;
    g.pp64 = e->d_pp64;
    g.pab64 = e->d_pab64;
    g.pair_ps32 = e->d_pair_ps32;
    g.pair_np32 = e->d_pair_np32;
    g.pp32 = e->d_pp32;
    g.pab32 = e->d_pab32;
    g.boys64 = e->d_boys64;
    g.boys32 = e->d_boys32;
    g.sh_l = e->d_sh_l;
    g.sh_ao = e->d_sh_ao;
    g.sh_np = e->d_sh_np;
    g.sh_po = e->d_sh_po;
    g.sh_no = e->d_sh_no;
    g.sh_cf = e->d_sh_cf;
    g.sh_atom = e->d_sh_atom;
    g.ctr = e->d_ctr;
    g.alpha = e->d_alpha;
    g.coeff = e->d_coeff;
    g.norm = e->d_norm;
    g.atom_z = e->d_atom_z;
    g.atom_xyz = e->d_atom_xyz;
    g.n_atoms = nat;
    g.n_pairs = e->n_pairs;
    g.n_ao = e->N_tot;
    g.P = e->d_Dt;
    g.W = e->d_Db;
    g.list = e->d_list;
    g.part = e->d_gpart;
    g.part_stride = stride;
    g.n3 = n3;
    // per-class timing: an event after every class (created for this call only)
    const bool timing = (a->flags & HFX_JK_CLASS_TIMING) != 0 && st != nullptr;
    std::vector<hipEvent_t> cev;
    if (timing) {
        cev.resize(N_COMBOS);
        for (auto& ev : cev) HIPCHK(hipEventCreate(&ev));
    }
    HIPCHK(hipEventRecord(e->ev0, s));
    // The class kernels go to the engine's streams (forked from s, joined back before the
    // one-electron part) by longest-processing-time-first on a static cost estimate: the model
    // FLOPs of the class's unscreened quartets at its mean primitive counts, ~30x for the
    // generic kernel (measured: 0.06 against 0.3-14 model TFLOP/s for the unrolled ones), split
    // over its launch units.  With per-class timing everything stays on s, in class order.
    const int n_str = timing ? 1 : e->n_side;
    hipStream_t ss[hfx_engine::N_SIDE];
    ss[0] = s;
    for (int i = 1; i < n_str; ++i) ss[i] = e->side[i];
    struct Item {
        double cost;
        int k, fp32, unit;
    };
    std::vector<Item> items;
    for (int k = 0; k < N_COMBOS; ++k) {
        const int p1 = combo_pc1(k), p2 = combo_pc2(k);
        if (e->cls_ids[p1].empty() || e->cls_ids[p2].empty()) continue;
        const int la = pc_la(p1), lb = pc_lb(p1), lc = pc_la(p2), ld = pc_lb(p2);
        const int units = grad_class_units(la, lb, lc, ld);
        const double c = grad_cost(e, k) * (units == 0 ? 30.0 : 1.0) / std::max(1, units);
        const bool split = mixed && grad_class_has_fp32(la, lb, lc, ld);
        for (int f = 0; f < (split ? 2 : 1); ++f)
            for (int u = 0; u < std::max(1, units); ++u) items.push_back({c, k, f, u});
    }
    auto launch = [&](const Item& it, hipStream_t st) {
        const int p1 = combo_pc1(it.k), p2 = combo_pc2(it.k);
        const int la = pc_la(p1), lb = pc_lb(p1), lc = pc_la(p2), ld = pc_lb(p2);
        const bool split = mixed && grad_class_has_fp32(la, lb, lc, ld);
        g.range = e->d_ranges + (it.fp32 ? 2 * N_COMBOS : (mixed && !split ? 4 * N_COMBOS : 0)) + 2 * it.k;
        launch_grad_unit(it.fp32, g, la, lb, lc, ld, it.unit, grid, st);
    };
    if (timing) {
        for (int k = 0; k < N_COMBOS; ++k) {
            for (const Item& it : items)
                if (it.k == k) launch(it, s);
            HIPCHK(hipEventRecord(cev[k], s));
        }
    } else {
        std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
        if (n_str > 1) {
            HIPCHK(hipEventRecord(e->ev_fork, s));
            for (int i = 1; i < n_str; ++i) HIPCHK(hipStreamWaitEvent(e->side[i], e->ev_fork, 0));
        }
        double load[hfx_engine::N_SIDE] = {};
        for (const Item& it : items) {
            int best = 0;
            for (int i = 1; i < n_str; ++i)
                if (load[i] < load[best]) best = i;
            load[best] += it.cost;
            launch(it, ss[best]);
        }
        for (int i = 1; i < n_str; ++i) {
            HIPCHK(hipEventRecord(e->ev_join[i], e->side[i]));
            HIPCHK(hipStreamWaitEvent(s, e->ev_join[i], 0));
        }
    }
    HIPCHK(hipEventRecord(e->ev1, s));
    launch_grad_one(g, grid, s);
    launch_grad_reduce(g, grid, e->d_gout, s);
    HIPCHK(hipGetLastError());
    std::vector<double> out(3 * (size_t)n3);
    long long rng[4 * N_COMBOS];
    unsigned long long stv[6 * N_COMBOS];
    HIPCHK(hipMemcpyAsync(out.data(), e->d_gout, out.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(rng, e->d_ranges, sizeof(rng), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(stv, e->d_stats, sizeof(stv), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // nuclear repulsion: d/dX_A sum_{A<B} Z_A Z_B / R_AB
    std::vector<double> gn(n3, 0.0);
    for (int A = 0; A < nat; ++A)
        for (int B = 0; B < nat; ++B) {
            if (A == B) continue;
            const double dx = a->xyz[3 * A] - a->xyz[3 * B], dy = a->xyz[3 * A + 1] - a->xyz[3 * B + 1],
                         dz = a->xyz[3 * A + 2] - a->xyz[3 * B + 2];
            const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            const double f = a->Z[A] * a->Z[B] / (r * r * r);
            gn[3 * A] -= f * dx;
            gn[3 * A + 1] -= f * dy;
            gn[3 * A + 2] -= f * dz;
        }
    for (int x = 0; x < n3; ++x) grad[x] = gn[x] + out[x] + out[n3 + x] + out[2 * n3 + x];
    if (parts) {
        std::copy(gn.begin(), gn.end(), parts);
        std::copy(out.begin(), out.end(), parts + n3);
    }
    if (st) {
        std::memset(st, 0, sizeof(*st));
        for (int k = 0; k < N_COMBOS; ++k) {
            st->n_quartets += rng[2 * k + 1] - rng[2 * k];
            const long long n32 = rng[2 * N_COMBOS + 2 * k + 1] - rng[2 * N_COMBOS + 2 * k];
            st->n_quartets += n32;
            const int p1 = combo_pc1(k), p2 = combo_pc2(k);
            if (grad_class_has_fp32(pc_la(p1), pc_lb(p1), pc_la(p2), pc_lb(p2))) st->n_fp32 += n32;
            st->class_prims[k] = (int64_t)(stv[6 * k + 2] + stv[6 * k + 3]);  // both parts, unpruned pair lists
        }
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        st->device_ms = ms;
        if (timing) {
            for (int k = 0; k < N_COMBOS; ++k) {
                HIPCHK(hipEventElapsedTime(&ms, k ? cev[k - 1] : e->ev0, cev[k]));
                st->class_ms[k] = ms;
            }
            for (auto& ev : cev) (void)hipEventDestroy(ev);
        }
    }
    return HFX_OK;
}

}  // namespace

namespace hfx_internal {

int set_error(int code, const std::string& msg) { return fail(code, msg); }

int engine_device(const hfx_engine* e, int* device) {
    if (!e || !device) return fail(HFX_ERR_INVALID, "null engine");
    *device = e->device;
    return HFX_OK;
}

}  // namespace hfx_internal

extern "C" {

const char* hfx_last_error(void) { return g_err.c_str(); }

int hfx_create(const hfx_basis* basis, int device, hfx_engine** out) { return create_impl(1, basis, device, out); }

int hfx_create_batched(int32_t n_mol, const hfx_basis* bases, int device, hfx_engine** out) {
    return create_impl(n_mol, bases, device, out);
}

int hfx_destroy(hfx_engine* e) {
    if (!e) return HFX_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    delete e;
    return HFX_OK;
}

int hfx_sizes(const hfx_engine* e, int64_t* n_aos, int64_t* n_shells, int64_t* n_mol, int64_t* elems) {
    if (!e) return fail(HFX_ERR_INVALID, "null engine");
    if (n_aos) *n_aos = e->N_tot;
    if (n_shells) *n_shells = e->nS;
    if (n_mol) *n_mol = e->n_mol;
    if (elems) *elems = e->n_elems;
    return HFX_OK;
}

// The one-electron kernels of every pair class, writing the packed N_m x N_m blocks to dout; the
// atoms of molecule m are dz/dxyz[da0[m] .. da0[m+1]) (device pointers).
static int launch_one_electron(hfx_engine* e, int kind, double* dout, const double* dz, const double* dxyz,
                               const int* da0) {
    OneArgs o{};
    o.pair_a = e->d_pair_a;
    o.pair_b = e->d_pair_b;
    o.sh_po = e->d_sh_po;
    o.sh_np = e->d_sh_np;
    o.sh_no = e->d_sh_no;
    o.sh_ao = e->d_sh_ao;
    o.sh_mbase = e->d_sh_mbase;
    o.sh_mn = e->d_sh_mn;
    o.sh_mol = e->d_sh_mol;
    o.ctr = e->d_ctr;
    o.alpha = e->d_alpha;
    o.coeff = e->d_coeff;
    o.norm = e->d_norm;
    o.atom_z = dz;
    o.atom_xyz = dxyz;
    o.mol_atom0 = da0;
    o.boys = e->d_boys64;
    o.out = dout;
    for (int pc = 0; pc < N_PAIR_CLASSES; ++pc) {
        o.ids = e->d_cls_ids[pc];
        o.n = (int)e->cls_ids[pc].size();
        if (kind == 0) launch_one<0>(pc, o, e->stream);
        else if (kind == 1) launch_one<1>(pc, o, e->stream);
        else launch_one<2>(pc, o, e->stream);
    }
    HIPCHK(hipGetLastError());
    return HFX_OK;
}

static int one_electron(hfx_engine* e, int kind, int32_t nat, const double* Z, const double* xyz,
                        const int32_t* atom_mol, double* M) {
    if (!e || !M) return fail(HFX_ERR_INVALID, "null argument");
    if (kind == 2 && (nat < 0 || (nat > 0 && (!Z || !xyz)))) return fail(HFX_ERR_INVALID, "hfx_nuclear: bad atoms");
    HIPCHK(hipSetDevice(e->device));
    DevArena tmp;
    double* dout = tmp.alloc<double>(e->n_elems);
    if (!dout) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
    std::vector<double> az, axyz;
    std::vector<int> atom0(e->n_mol + 1, 0);
    if (kind == 2) {
        if (e->n_mol > 1 && !atom_mol) return fail(HFX_ERR_INVALID, "hfx_nuclear: atom_mol required for batched engines");
        for (int m = 0; m < e->n_mol; ++m) {
            atom0[m] = (int)az.size();
            for (int a = 0; a < nat; ++a)
                if (!atom_mol || atom_mol[a] == m) {
                    az.push_back(Z[a]);
                    axyz.insert(axyz.end(), xyz + 3 * a, xyz + 3 * a + 3);
                }
        }
        atom0[e->n_mol] = (int)az.size();
        if (!finite_all(az.data(), (long long)az.size()) || !finite_all(axyz.data(), (long long)axyz.size()))
            return fail(HFX_ERR_NONFINITE, "hfx_nuclear: non-finite atoms");
    }
    double* dz = tmp.alloc<double>(az.size());
    double* dxyz = tmp.alloc<double>(axyz.size());
    int* da0 = tmp.alloc<int>(atom0.size());
    if (!dz || !dxyz || !da0) return fail(HFX_ERR_NOMEM, "hipMalloc failed");
    if (!az.empty()) {
        HIPCHK(hipMemcpy(dz, az.data(), az.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dxyz, axyz.data(), axyz.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(da0, atom0.data(), atom0.size() * sizeof(int), hipMemcpyHostToDevice));
    {
        const int rc = launch_one_electron(e, kind, dout, dz, dxyz, da0);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(M, dout, e->n_elems * sizeof(double), hipMemcpyDeviceToHost));
    return HFX_OK;
}

// hfx_one_electron_noalloc: the same kernels with the staging of the atoms and of the output in
// the device's `scratch` array (hfx_create looked its address up once per device).
static int one_electron_noalloc(hfx_engine* e, int kind, int32_t nat, const double* Z, const double* xyz,
                                const int32_t* atom_mol, double* M) {
    if (!e || !M) return fail(HFX_ERR_INVALID, "hfx_one_electron_noalloc: null argument");
    if (kind < 0 || kind > 2) return fail(HFX_ERR_INVALID, "hfx_one_electron_noalloc: kind must be 0, 1 or 2");
    if (!e->scratch) return fail(HFX_ERR_DEVICE, "hfx_one_electron_noalloc: no scratch address for this device");
    if (e->n_elems > SCRATCH_OUT)
        return fail(HFX_ERR_INVALID, "hfx_one_electron_noalloc: " + std::to_string(e->n_elems) +
                                         " matrix elements exceed the allocation-free limit of " +
                                         std::to_string(SCRATCH_OUT) + " (use hfx_overlap/hfx_kinetic/hfx_nuclear)");
    if (e->n_mol > SCRATCH_MOLS)
        return fail(HFX_ERR_INVALID, "hfx_one_electron_noalloc: more than " + std::to_string(SCRATCH_MOLS) + " molecules");
    const int na = kind == 2 ? nat : 0;
    if (kind == 2) {
        if (nat < 0 || (nat > 0 && (!Z || !xyz))) return fail(HFX_ERR_INVALID, "hfx_one_electron_noalloc: bad atoms");
        if (nat > SCRATCH_ATOMS)
            return fail(HFX_ERR_INVALID, "hfx_one_electron_noalloc: " + std::to_string(nat) +
                                             " atoms exceed the allocation-free limit of " + std::to_string(SCRATCH_ATOMS));
        if (e->n_mol > 1 && !atom_mol)
            return fail(HFX_ERR_INVALID, "hfx_one_electron_noalloc: atom_mol required for batched engines");
        if (!finite_all(Z, nat) || !finite_all(xyz, 3LL * nat))
            return fail(HFX_ERR_NONFINITE, "hfx_one_electron_noalloc: non-finite atoms");
    }
    // atom ranges per molecule; the atoms must come grouped by molecule, in molecule order
    static thread_local int atom0[SCRATCH_MOLS + 1];
    {
        int a = 0;
        for (int m = 0; m < e->n_mol; ++m) {
            atom0[m] = a;
            while (a < na && (!atom_mol || atom_mol[a] == m)) ++a;
        }
        atom0[e->n_mol] = a;
        if (a != na)
            return fail(HFX_ERR_INVALID, "hfx_one_electron_noalloc: atoms must be grouped by molecule in molecule order");
    }
    double* dout = e->scratch;
    double* dz = e->scratch + SCRATCH_OUT;
    double* dxyz = dz + SCRATCH_ATOMS;
    int* da0 = reinterpret_cast<int*>(dxyz + 3 * SCRATCH_ATOMS);
    hipStream_t s = e->stream;
    HIPCHK(hipSetDevice(e->device));
    if (na > 0) {
        HIPCHK(hipMemcpyAsync(dz, Z, na * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dxyz, xyz, 3 * (size_t)na * sizeof(double), hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipMemcpyAsync(da0, atom0, (e->n_mol + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    {
        const int rc = launch_one_electron(e, kind, dout, dz, dxyz, da0);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(M, dout, e->n_elems * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return HFX_OK;
}

int hfx_overlap(hfx_engine* e, double* S) { return one_electron(e, 0, 0, nullptr, nullptr, nullptr, S); }
int hfx_kinetic(hfx_engine* e, double* T) { return one_electron(e, 1, 0, nullptr, nullptr, nullptr, T); }
int hfx_nuclear(hfx_engine* e, int32_t nat, const double* Z, const double* xyz, const int32_t* atom_mol, double* V) {
    return one_electron(e, 2, nat, Z, xyz, atom_mol, V);
}

int hfx_one_electron_noalloc(hfx_engine* e, int32_t kind, int32_t n_atoms, const double* Z, const double* xyz,
                             const int32_t* atom_mol, double* M) {
    return one_electron_noalloc(e, kind, n_atoms, Z, xyz, atom_mol, M);
}

int hfx_schwarz(hfx_engine* e, double* Q) {
    if (!e || !Q) return fail(HFX_ERR_INVALID, "null argument");
    std::memcpy(Q, e->own_q.data(), e->n_qelems * sizeof(double));
    return HFX_OK;
}

int hfx_eri_quartets(hfx_engine* e, int64_t n, const int32_t* abcd, const int64_t* offsets, double* out) {
    if (!e || e->n_mol != 1 || (n > 0 && (!abcd || !offsets || !out)))
        return fail(HFX_ERR_INVALID, "hfx_eri_quartets: bad arguments");
    HIPCHK(hipSetDevice(e->device));
    struct Req {
        long long q;
        bool swap_bk, swap_ab, swap_cd;
    };
    std::vector<std::vector<uint2>> ents(N_COMBOS);
    std::vector<std::vector<Req>> reqs(N_COMBOS);
    for (long long q = 0; q < n; ++q) {
        const int* s = &abcd[4 * q];
        for (int k = 0; k < 4; ++k)
            if (s[k] < 0 || s[k] >= e->nS) return fail(HFX_ERR_INVALID, "hfx_eri_quartets: shell index out of range");
        const int p1 = e->pair_of.at((long long)std::min(s[0], s[1]) * e->nS + std::max(s[0], s[1]));
        const int p2 = e->pair_of.at((long long)std::min(s[2], s[3]) * e->nS + std::max(s[2], s[3]));
        int bra = p1, ket = p2;
        bool swap_bk = false;
        if (e->pair_pc[p1] < e->pair_pc[p2]) {
            std::swap(bra, ket);
            swap_bk = true;
        }
        const int k = combo_of(e->pair_pc[bra], e->pair_pc[ket]);
        Req r;
        r.q = q;
        r.swap_bk = swap_bk;
        // Does the requested (a,b) [(c,d)] order differ from the pair's stored orientation?
        r.swap_ab = e->pair_a[p1] != s[0];
        r.swap_cd = e->pair_a[p2] != s[2];
        ents[k].push_back(make_uint2(bra, ket));
        reqs[k].push_back(r);
    }
    for (int k = 0; k < N_COMBOS; ++k) {
        if (ents[k].empty()) continue;
        std::vector<double> blk;
        std::vector<long long> offs;
        int rc = store_blocks(e, k, ents[k], blk, offs);
        if (rc) return rc;
        for (size_t x = 0; x < ents[k].size(); ++x) {
            const Req& r = reqs[k][x];
            const int* s = &abcd[4 * r.q];
            const int na = ncart_h(e->sh_l[s[0]]), nb = ncart_h(e->sh_l[s[1]]), nc = ncart_h(e->sh_l[s[2]]),
                      nd = ncart_h(e->sh_l[s[3]]);
            // stored block: [bra pair comps][ket pair comps] in the pairs' own orientation
            const int bp = ents[k][x].x, kp = ents[k][x].y;
            const int nbk = ncart_h(e->sh_l[e->pair_a[kp]]) * ncart_h(e->sh_l[e->pair_b[kp]]);
            const int nkb = ncart_h(e->sh_l[e->pair_b[bp]]);
            const int nkd = ncart_h(e->sh_l[e->pair_b[kp]]);
            const bool same = (s[0] == s[2] && s[1] == s[3]);
            double* dst = out + offsets[r.q];
            for (int A = 0; A < na; ++A)
                for (int B = 0; B < nb; ++B)
                    for (int Cc = 0; Cc < nc; ++Cc)
                        for (int D = 0; D < nd; ++D) {
                            const long long i = e->sh_ao[s[0]] + A, j = e->sh_ao[s[1]] + B;
                            const long long kk = e->sh_ao[s[2]] + Cc, l = e->sh_ao[s[3]] + D;
                            double v = 0.0;
                            const bool zero = (j > i) || (l > kk) || (same && i * (i + 1) / 2 + j < kk * (kk + 1) / 2 + l);
                            if (!zero) {
                                // components in the pair orientations
                                const int p1a = r.swap_ab ? B : A, p1b = r.swap_ab ? A : B;
                                const int p2a = r.swap_cd ? D : Cc, p2b = r.swap_cd ? Cc : D;
                                int ba, bb, ka, kb;
                                if (!r.swap_bk) {
                                    ba = p1a; bb = p1b; ka = p2a; kb = p2b;
                                } else {
                                    ba = p2a; bb = p2b; ka = p1a; kb = p1b;
                                }
                                v = blk[offs[x] + (long long)(ba * nkb + bb) * nbk + (ka * nkd + kb)];
                            }
                            dst[((long long)A * nb + B) * nc * nd + (long long)Cc * nd + D] = v;
                        }
        }
    }
    return HFX_OK;
}

int hfx_jk(hfx_engine* e, const hfx_jk_args* args, double* J, double* Ka, double* Kb, hfx_jk_stats* st) {
    return jk_impl(e, args, J, Ka, Kb, st);
}

int hfx_set_launch(hfx_engine* e, int32_t blocks_per_cu, int32_t chunk_len, int32_t streams) {
    if (!e) return fail(HFX_ERR_INVALID, "hfx_set_launch: null engine");
    if (blocks_per_cu < 0 || blocks_per_cu > 16 || (chunk_len != 0 && (chunk_len < 256 || chunk_len > 65536)) ||
        streams < 0 || streams > hfx_engine::N_SIDE)
        return fail(HFX_ERR_INVALID, "hfx_set_launch: blocks_per_cu 1..16, chunk_len 256..65536, streams 1..4 (0 keeps)");
    if (blocks_per_cu) e->blocks_per_cu = blocks_per_cu;
    if (chunk_len) e->chunk_len = chunk_len;
    if (streams) e->n_side = streams;
    e->drop_graph();  // the captured build has the old launch shapes
    return HFX_OK;
}

int hfx_jk_async(hfx_engine* e, const hfx_jk_args* args, double* J, double* Ka, double* Kb, void* stream) {
    return jk_async_impl(e, args, J, Ka, Kb, static_cast<hipStream_t>(stream));
}

int hfx_store_eris(hfx_engine* e, const double* Q, double schwarz_thresh, int64_t* n_quartets, int64_t* n_bytes) {
    try {
        return store_impl(e, Q, schwarz_thresh, n_quartets, n_bytes);
    } catch (const std::exception& x) {
        if (e) e->release_store();
        return fail(HFX_ERR_NOMEM, std::string("hfx_store_eris: ") + x.what());
    }
}

int hfx_release_eris(hfx_engine* e) {
    if (!e) return fail(HFX_ERR_INVALID, "hfx_release_eris: null engine");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->release_store();
    return HFX_OK;
}

int hfx_rhf_gradient(hfx_engine* e, const hfx_grad_args* args, double* grad, double* parts, hfx_grad_stats* stats) {
    return grad_impl(e, args, grad, parts, stats);
}

}  // extern "C"
