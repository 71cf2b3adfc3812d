// engine.hpp - device-side data layout shared by the engine and its kernel translation units.
//
// HBM layout (all arrays allocated once per engine):
//   shells       sh_ao (AO offset inside its molecule), sh_mbase (offset of the molecule's
//                N_m x N_m block in every packed matrix), sh_mn (N_m), sh_cf[6*s + c]
//                (Cartesian normalization ratio N(p,c)/N(p,0), independent of p)
//   shell pairs  pair_a, pair_b (la >= lb; a >= b when la == lb), pair_ps/pair_np (range of
//                the pair's primitive pairs)
//   prim pairs   PrimPair64 / PrimPair32, 96 / 48 bytes each, pair-major; a primitive pair
//                whose |K c_a c_b g_a g_b| is below 1e-20 (FP64 list) or 1e-11 (FP32 list)
//                contributes nothing at that precision and is left out
//   work lists   uint2 {bra pair, ket pair} per surviving shell quartet, one region per
//                quartet class: FP64 entries fill the region from its start, FP32 entries
//                from its end
//   matrices     N x N column-major blocks, one per molecule, concatenated; the kernels'
//                density reads and J/K accumulators use packed lower triangles of the same
//                blocks (N(N+1)/2 each, concatenated), which halves their cache footprint
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hfx {

// Everything a primitive pair contributes to an ERI.
//   kp = K_ab c_a c_b g_a g_b / p with K_ab = exp(-ab/p |A-B|^2) (src/IntegralEngine.cpp:1006-1007),
//   c the contraction coefficients and g the primitive part of the normalization.
template <class R>
struct PrimPair {
    R p, px, py, pz, kp, pax, pay, paz, pbx, pby, pbz, oo2p;
};
using PrimPair64 = PrimPair<double>;
using PrimPair32 = PrimPair<float>;

// Pair class of (la >= lb) and quartet class of (pair class pc1 >= pc2).
__host__ __device__ constexpr int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }
__host__ __device__ constexpr int combo_of(int pc1, int pc2) { return pc1 * (pc1 + 1) / 2 + pc2; }
__host__ __device__ constexpr int combo_pc1(int k) {
    int p = 0;
    while ((p + 1) * (p + 2) / 2 <= k) ++p;
    return p;
}
__host__ __device__ constexpr int combo_pc2(int k) { return k - combo_pc1(k) * (combo_pc1(k) + 1) / 2; }
__host__ __device__ constexpr int pc_la(int pc) { return combo_pc1(pc); }
__host__ __device__ constexpr int pc_lb(int pc) { return combo_pc2(pc); }

// Packed lower triangle of a symmetric block: element (i, j) of either orientation.
__host__ __device__ __forceinline__ long long tri_index(long long i, long long j) {
    return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

constexpr int N_PAIR_CLASSES = 6;
constexpr int N_COMBOS = 21;

struct EriArgs {
    const int* pair_a;
    const int* pair_b;
    const int* pair_ps;    // primitive pairs of each shell pair kept for FP64 work
    const int* pair_np;
    const int* pair_ps32;  // the (more strongly pruned) primitive pairs for FP32 work
    const int* pair_np32;
    const PrimPair64* pp64;
    const PrimPair32* pp32;
    const int* sh_ao;
    const long long* sh_mbase;
    const long long* sh_tbase;  // offset of the shell's molecule in the packed triangles
    const int* sh_mn;
    const double* sh_cf;
    const uint2* list;
    const long long* range;  // [lo, hi) of this class's entries in list
    const double* Dt;  // packed triangles of the density increments
    const double* Da;
    const double* Db;  // NULL: no K_beta
    double* AJ;    // packed-triangle accumulators: n_copies copies, acc_stride elements apart;
    double* AKa;   // each XCD adds into its own copy (selected by the XCC_ID hardware
    double* AKb;   // register); element (i,j) holds the sum of both orientations
    long long acc_stride;
    int n_copies;
    int fixed;     // 1: accumulators hold fixed-point pairs (element t at [2t, 2t+1], int64 each)
    double tauS;
    double* out;               // store mode (non-NULL): blocks [ab][cd], canonical class order
    const long long* out_off;  // store mode: offset of each entry's block (NULL: the stored-integral
                               // layout of the class list, element-major in groups of 64 entries)
    const double* stored;      // non-NULL: digest the class's stored integrals (FP64) instead of
                               // computing them; same layout as the store mode without offsets
    const double* boys64;
    const float* boys32;
    // Chunked schedule (NULL: grid-stride over the list): blocks claim chunk_len consecutive
    // entries at a time from chunk_ctr[chunk index of the launch], so a block stays on one bra
    // pair for many quartets and collects that pair's exchange rows in LDS (tile_n columns
    // per row; 0: no LDS rows).
    unsigned int* chunk_ctr;
    int chunk_len;     // largest chunk; smaller lists use shorter chunks
    int stagger;       // 1: each block walks its claimed run from a block-dependent starting
                       // iteration (wrapping around), so that concurrent blocks of one class
                       // reach the same ket pairs at different times
    int tile_n;
    int tile_shells;   // bits of the touched-ket-shell mask kept with the rows
    const int* sh_local;  // index of a shell inside its molecule
    const int* sh_nao;
    const int* sh_ns;     // shells of the shell's molecule
};

// Arguments of the RHF-gradient kernels (k_grad.hip).
struct GradArgs {
    const int* pair_a;
    const int* pair_b;
    const int* pair_ps;
    const int* pair_np;
    const PrimPair64* pp64;
    const double* pab64;  // exponents (a, b) of every FP64 primitive pair
    const int* pair_ps32;
    const int* pair_np32;
    const PrimPair32* pp32;
    const float* pab32;
    const double* boys64;
    const float* boys32;
    const int* sh_l;
    const int* sh_ao;
    const int* sh_np;
    const int* sh_po;
    const int* sh_no;
    const double* sh_cf;
    const int* sh_atom;
    const double* ctr;
    const double* alpha;
    const double* coeff;
    const double* norm;
    const double* atom_z;
    const double* atom_xyz;
    int n_atoms;
    int n_pairs;
    int n_ao;
    const double* P;  // N x N total density
    const double* W;  // N x N energy-weighted density
    const uint2* list;
    const long long* range;
    double* part;            // per-block rows: [-W dS | P dh | two-electron], n3 each
    long long part_stride;
    int n3;                  // 3 * n_atoms
};
// Launch units of a class's two-electron gradient: grad_class_units of them (its unrolled
// kernels, one per chunk of bra components), or one (the generic kernel) when that is 0.  Every
// kernel adds into g.part with atomics, so launches on different streams may overlap.
int grad_class_units(int la, int lb, int lc, int ld);
void launch_grad_unit(int fp32, const GradArgs& g, int la, int lb, int lc, int ld, int unit, int grid, hipStream_t s);
// The two-electron gradient kernels of one quartet class, compiled with every angular loop
// unrolled (k_grad_eri.hip): one launcher per chunk of bra components (its index first_chunk;
// fp32: it has an FP32 kernel too); classes without any use the generic kernel of k_grad.hip.
using GradClassFn = void (*)(int fp32, const GradArgs& g, int grid, hipStream_t s);
bool register_grad_class(int combo, int first_chunk, GradClassFn fn, bool fp32);
// A mixed-precision gradient runs an FP32 pass for the class only where this holds (unrolled
// FP32 kernels); elsewhere the class's whole list goes through the FP64 kernels in one pass.
bool grad_class_has_fp32(int la, int lb, int lc, int ld);
void launch_grad_one(const GradArgs& g, int grid, hipStream_t s);
void launch_grad_reduce(const GradArgs& g, int n_rows, double* out, hipStream_t s);

// Each kernel translation unit compiles chunks [begin, end) of one class and registers a
// launcher for them at load time; launch_eri runs every registered part of a class, each as
// one launch (on the stream of its first chunk, chunk_streams[begin]).
// A class has nch chunks in each precision (fewer in FP32 for the heavy classes).
using EriPartFn = void (*)(int fp32, const EriArgs& a, int grid, const hipStream_t* chunk_streams);
// The ket-major Coulomb pass of a class's stored integrals (k_jket.hip, one translation unit per
// class), registered at load time.
using JketFn = void (*)(const EriArgs& a, int grid, hipStream_t s);
bool register_jket(int combo, JketFn fn);
bool register_eri_part(int combo, int fp32, int begin, int end, int nch, EriPartFn fn);
// Number of chunks of a class in a precision, or -1 when its parts do not cover [0, nch).
int eri_chunks(int combo, int fp32);
// Chunk ranges [begin[k], end[k]) of the class's launches in a precision; returns their number (<= 64).
int eri_parts(int combo, int fp32, int* begin, int* end);
void launch_eri(int combo, int fp32, const EriArgs& a, int grid, const hipStream_t* chunk_streams);

}  // namespace hfx
