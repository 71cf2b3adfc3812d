/*
 * hfx.h - C ABI of the MI355X direct-SCF Fock-build engine.
 *
 * The engine replaces the reference's C++ class IntegralEngine (include/IntegralEngine.hpp:8-35)
 * and the shell-quartet loop of SCF::buildFockMatrixDirect (src/SCF.cpp:523-744) with one call
 * per Fock build.  Everything crossing this boundary is a plain pointer plus sizes; no C++ or
 * torch types, no exceptions.  Every function returns an int status (HFX_OK = 0); on failure
 * hfx_last_error() returns a thread-local message.
 *
 * Matrices are N x N, column-major (element (r, c) at r + c*N), full and symmetric, as the
 * reference's Eigen::MatrixXd after selfadjointView<Upper> (src/SCF.cpp:727-728).
 * The caller owns every host array for the duration of a call; the engine keeps no pointer.
 * Calls on one engine are serialized on that engine's HIP stream; different engines are
 * independent.
 */
#ifndef HFX_H
#define HFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HFX_OK = 0,
    HFX_ERR_INVALID = 1,     /* bad handle, NULL pointer, inconsistent sizes            */
    HFX_ERR_UNSUPPORTED = 2, /* angular momentum above d, unexpected Cartesian order    */
    HFX_ERR_NOMEM = 3,       /* device or host allocation failed                        */
    HFX_ERR_DEVICE = 4,      /* a HIP call failed                                       */
    HFX_ERR_NONFINITE = 5,   /* NaN or Inf in an input matrix                           */
    HFX_ERR_NOT_CONVERGED = 6
};

enum { HFX_FP64 = 0, HFX_MIXED = 1 };

/* hfx_jk_args.flags */
enum {
    HFX_JK_DEVICE_PTRS = 1, /* density inputs and J/K outputs are device pointers           */
    HFX_JK_USE_GRAPH = 2,   /* replay the engine's captured hipGraph of the build pipeline   */
    HFX_JK_NO_STATS = 4,    /* skip the survivor statistics read-back                         */
    HFX_JK_CLASS_TIMING = 8, /* time every class kernel group with HIP events (stats.class_ms*) */
    HFX_JK_STORED = 16,      /* digest the integrals kept by hfx_store_eris: no density test, FP64 only,
                              * args->Q and density_thresh unused (the stored list is fixed);
                              * the conventional build SCF::buildFockMatrix (src/SCF.cpp:233-521) */
    HFX_JK_DETERMINISTIC = 32 /* bitwise-reproducible J/K: every contribution is added as a 2 x 64-bit
                              * fixed-point integer (resolution 2^-72, range +-2^31), so the sum does not
                              * depend on the order in which the device performs the additions; slower
                              * (no LDS exchange rows, two integer atomics per term).  The reference's
                              * critical-section reduction (src/SCF.cpp:717-724) is order-dependent too;
                              * this mode has no counterpart there. */
};

#define HFX_MAX_L 2
#define HFX_N_CLASSES 21 /* quartet classes (la>=lb, lc>=ld, (la,lb)>=(lc,ld)) for l <= 2 */

/* Flat basis layout: the struct-of-arrays of IntegralEngine::IntegralEngine
 * (src/IntegralEngine.cpp:13-75, fields include/IntegralEngine.hpp:145-168).
 * Shell order, Cartesian order and per-primitive per-component normalization are those of
 * BasisSet::BasisSet (src/Basis.cpp:14-86). */
typedef struct hfx_basis {
    int32_t n_shells;
    int32_t n_prims;            /* sum of n_prim                                       */
    int32_t n_aos;              /* sum of (l+1)(l+2)/2                                 */
    int32_t n_norm;             /* sum of n_prim * n_ao                                */
    const int32_t* l;           /* [n_shells]  lShell                                  */
    const int32_t* n_prim;      /* [n_shells]  nPrimitives                             */
    const int32_t* prim_offset; /* [n_shells]  primitiveOffsets                        */
    const int32_t* ao_offset;   /* [n_shells]  aoOffsets                               */
    const int32_t* norm_offset; /* [n_shells]  normFactorOffsets                       */
    const double* centers;      /* [3*n_shells] x,y,z of each shell (3 x nS, col-major) */
    const double* alpha;        /* [n_prims]   exponents                               */
    const double* coeff;        /* [n_prims]   contraction coefficients                */
    const double* norm;         /* [n_norm]    norm[norm_offset[s] + p*n_ao + comp]    */
    const int32_t* cart;        /* [3*n_aos]   (i,j,k) of each AO                      */
} hfx_basis;

typedef struct hfx_engine hfx_engine;

/* Arguments of one J/K build.  Replaces the members SCF::buildFockMatrixDirect reads:
 * D_tot - D_tot_prev, D_alpha - D_alpha_prev, D_beta - D_beta_prev (src/SCF.cpp:536-541),
 * Q (src/SCF.cpp:605), options.schwartzThreshold / densityThreshold (src/SCF.cpp:605,622). */
typedef struct hfx_jk_args {
    const double* dD_tot;   /* N x N                                                       */
    const double* dD_alpha; /* N x N                                                       */
    const double* dD_beta;  /* N x N, or NULL for RHF (no K_beta)                          */
    const double* Q;        /* nS x nS Schwarz table, or NULL for the engine's own         */
    double schwarz_thresh;  /* tau_S: shell test Q(a,b)Q(c,d) < tau_S and AO test |eri|<tau_S */
    double density_thresh;  /* tau_D                                                       */
    int32_t precision;      /* HFX_FP64 or HFX_MIXED                                       */
    double fp32_cut;        /* MIXED: quartets with Q(ab)Q(cd)max|dD| < fp32_cut run in FP32 */
    int32_t flags;          /* HFX_JK_*                                                    */
} hfx_jk_args;

/* Survivor statistics of one build.  The list hash is an order-independent multiset hash of
 * the surviving shell quartets written in the reference's canonical form (a>=b, c>=d,
 * a(a+1)/2+b >= c(c+1)/2+d; src/SCF.cpp:602-603), so two builds have the same work list
 * exactly when counts and both hashes agree. */
typedef struct hfx_jk_stats {
    int64_t n_quartets;                   /* surviving shell quartets                     */
    int64_t n_fp32;                       /* of which computed in FP32 (MIXED)            */
    int64_t class_count[HFX_N_CLASSES];   /* survivors per angular-momentum class         */
    uint64_t hash_sum;                    /* sum of splitmix64(((a*nS+b)*nS+c)*nS+d)      */
    uint64_t hash_xor;                    /* xor of the same                              */
    double model_gflop;                   /* algorithmic GFLOP of the surviving work      */
    double device_ms;                     /* device time of the build (HIP events)        */
    double class_gflop64[HFX_N_CLASSES];  /* model GFLOP of each class's FP64 / FP32 part */
    double class_gflop32[HFX_N_CLASSES];
    double class_ms64[HFX_N_CLASSES];     /* HFX_JK_CLASS_TIMING: device ms of each part  */
    double class_ms32[HFX_N_CLASSES];
    /* The same model FLOP count over the primitive quartets the kernels actually execute: the
     * engine leaves out primitive pairs whose |K c_a c_b N_a N_b| is below 1e-20 (FP64 work) or
     * 1e-11 (FP32 work), which model_gflop above still counts. */
    double exec_gflop;
    double class_exec_gflop64[HFX_N_CLASSES];
    double class_exec_gflop32[HFX_N_CLASSES];
} hfx_jk_stats;

/* IntegralEngine::IntegralEngine (src/IntegralEngine.cpp:13-75).  Builds the device basis,
 * primitive-pair data, the Boys table and the engine's Schwarz table. */
int hfx_create(const hfx_basis* basis, int device, hfx_engine** out);

/* Many independent molecules behind one engine: quartets never mix molecules, and every
 * N x N argument or result becomes the concatenation of the molecules' N_m x N_m
 * column-major blocks, in molecule order.  No reference counterpart (the reference runs one
 * molecule per process). */
int hfx_create_batched(int32_t n_mol, const hfx_basis* bases, int device, hfx_engine** out);

int hfx_destroy(hfx_engine* e);

/* Sizes of the engine's layout: AOs, shells, molecules, and the packed N x N element count. */
int hfx_sizes(const hfx_engine* e, int64_t* n_aos, int64_t* n_shells, int64_t* n_mol,
              int64_t* n_matrix_elems);

/* IntegralEngine::overlapMatrix (src/IntegralEngine.cpp:145-159). */
int hfx_overlap(hfx_engine* e, double* S);
/* IntegralEngine::kineticMatrix (src/IntegralEngine.cpp:241-256). */
int hfx_kinetic(hfx_engine* e, double* T);
/* IntegralEngine::nuclearAttractionMatrix (src/IntegralEngine.cpp:355-368); atoms are the
 * point charges Z at xyz[3*i..3*i+2] (Atom in include/Utils.hpp).  Batched engines take the
 * atoms of every molecule concatenated, with atom_mol[i] naming the molecule of atom i
 * (NULL for single-molecule engines). */
int hfx_nuclear(hfx_engine* e, int32_t n_atoms, const double* Z, const double* xyz,
                const int32_t* atom_mol, double* V);

/* The one-electron matrices without any device allocation: kind 0 = hfx_overlap, 1 = hfx_kinetic,
 * 2 = hfx_nuclear (same arguments and results as those calls; for kinds 0 and 1 the atoms are
 * ignored).  No counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The atoms
 * and the output are staged in a static device array of 16.2 MiB that every GPU the library runs on
 * holds until the process ends, so the engine's packed matrices may hold at most 2^21 elements
 * (sum of N_m^2: one molecule of up to 1448 functions), at most 4096 atoms and 4096 molecules, with
 * the atoms grouped by molecule in molecule order; larger inputs fail with HFX_ERR_INVALID.  Calls
 * on one GPU share that array: calls that overlap in time on one GPU race (serialize them). */
int hfx_one_electron_noalloc(hfx_engine* e, int32_t kind, int32_t n_atoms, const double* Z, const double* xyz,
                             const int32_t* atom_mol, double* M);

/* IntegralEngine::schwartzScreeningMatrix (src/IntegralEngine.cpp:695-728): nS x nS. */
int hfx_schwarz(hfx_engine* e, double* Q);

/* IntegralEngine::electronRepulsion(a,b,c,d, std::vector<double>&)
 * (src/IntegralEngine.cpp:731-992), for n quartets at once: quartet q is shells
 * abcd[4q..4q+3]; its naoA*naoB*naoC*naoD values go to out[offsets[q]...] in row-major
 * (a,b,c,d) order, with the entries the reference leaves at zero set to zero
 * (aoB > aoA, aoD > aoC, and (A==C && B==D && I < K); src/IntegralEngine.cpp:857-882). */
int hfx_eri_quartets(hfx_engine* e, int64_t n, const int32_t* abcd, const int64_t* offsets,
                     double* out);

/* One direct J/K build: the work of src/SCF.cpp:536-728 (density maxima, screened quartet
 * loop, digestion, reduction, symmetrization).  Outputs the increments J[dD_tot],
 * K_alpha[dD_alpha] and, when dD_beta is given, K_beta[dD_beta]; the caller forms
 * F = F_prev + J - K (src/SCF.cpp:731).  Kb may be NULL for RHF.  stats may be NULL. */
int hfx_jk(hfx_engine* e, const hfx_jk_args* args, double* J, double* Ka, double* Kb,
           hfx_jk_stats* stats);

/* hfx_jk enqueued on the caller's HIP stream without waiting for it (the graph-capturable
 * variant of the build): `stream` is a hipStream_t passed as void* (NULL: the null stream).
 * For a caller that overlaps the build with its own work, or captures a whole SCF iteration
 * (SCF::run calls the build at src/SCF.cpp:57-58) into its own hipGraph.
 *   - args->flags must include HFX_JK_DEVICE_PTRS: the densities and J/K are device buffers that
 *     stay valid until the stream has passed the end of the build; HFX_JK_CLASS_TIMING is refused
 *     and no statistics are returned (they would need a host read-back).
 *   - The call records no host synchronization and no memset; the build runs on `stream` and on
 *     the engine's side streams, forked from and joined back into `stream`.
 *   - HFX_JK_USE_GRAPH launches the engine's own captured build on `stream`; refused while
 *     `stream` is capturing (capture the call without it instead).
 *   - A Schwarz table that differs from the previous call's re-sorts the engine's pair lists
 *     synchronously, which a capturing stream refuses (HFX_ERR_INVALID): make one call with the
 *     table before capturing.
 *   - Calls on one engine must not overlap: outside capture the engine orders this build after
 *     its earlier work and its later calls after this build; a captured build must not be
 *     replayed while another call on the same engine runs. */
int hfx_jk_async(hfx_engine* e, const hfx_jk_args* args, double* J, double* Ka, double* Kb, void* stream);

/* Launch parameters of this engine's class kernels, the autotuner's knobs (0 keeps a value):
 * blocks_per_cu = persistent blocks of 256 threads per CU in each class launch (1..16, default 2),
 * chunk_len = longest run of list entries a block claims at a time (256..65536; default 2048, 8192 in
 * stored builds), streams = streams a build's class kernels are spread over (1..4, default 4).  From
 * the next build on (a captured build is captured again).  No reference counterpart: the reference's
 * quartet loop has the fixed OpenMP schedule(dynamic,32) (src/SCF.cpp:592). */
int hfx_set_launch(hfx_engine* e, int32_t blocks_per_cu, int32_t chunk_len, int32_t streams);

/* Conventional (stored-integral) mode.  Computes once, and keeps in HBM, the normalized
 * integral blocks of every canonical shell quartet with Q(a,b)Q(c,d) >= schwarz_thresh (Q as
 * in hfx_jk_args: the caller's table or, when NULL, the engine's own).  Replaces
 * IntegralEngine::electronRepulsionTensor (include/IntegralEngine.hpp:30,
 * src/IntegralEngine.cpp:635-692), which fills the reference's ElectronRepulsionTensor
 * (include/Molecule.hpp:10-80); here only the surviving blocks are kept, not the dense packed
 * tensor.  Afterwards hfx_jk with HFX_JK_STORED digests them against any density (the
 * conventional build of src/SCF.cpp:233-521: no density test, the AO test |eri| < tau_S with the
 * build's schwarz_thresh).  Fails with HFX_ERR_NOMEM when the blocks do not fit in the free
 * device memory.  Outputs the number of stored quartets and the bytes held (either may be NULL).
 * A later call replaces the store; hfx_release_eris (or hfx_destroy) frees it. */
int hfx_store_eris(hfx_engine* e, const double* Q, double schwarz_thresh, int64_t* n_quartets,
                   int64_t* n_bytes);
int hfx_release_eris(hfx_engine* e);

/* RHF nuclear gradient inputs.  No reference counterpart: the reference computes energies
 * only (src/SCF.cpp:23-136); this is the analytic derivative of that energy,
 *   dE/dX = sum P dh/dX - sum W dS/dX + 1/2 sum Gamma d(mn|ls)/dX + dVnn/dX,
 *   Gamma_mnls = P_mn P_ls - 1/4 (P_ml P_ns + P_ms P_nl),
 * with P the total density and W = 2 sum_occ eps_i C_i C_i^T. */
typedef struct hfx_grad_args {
    const double* P;           /* N x N total density                                      */
    const double* W;           /* N x N energy-weighted density                            */
    int32_t n_atoms;
    const double* Z;           /* nuclear charges                                          */
    const double* xyz;         /* 3 x n_atoms positions (Bohr); shells must sit on atoms    */
    const int32_t* shell_atom; /* atom of every shell, or NULL to match shell centres       */
    double schwarz_thresh;     /* quartets with Q(ab)Q(cd) < tau_S are skipped               */
    double density_thresh;     /* and those failing the hfx_jk density test on (P, P/2)      */
    int32_t precision;         /* HFX_FP64, or HFX_MIXED with fp32_cut as in hfx_jk; the
                                  FP32 part covers the classes with la+lb+lc+ld <= 3, the
                                  heavier ones run in FP64 throughout                       */
    double fp32_cut;
    int32_t flags;             /* HFX_JK_DEVICE_PTRS: P and W are device pointers;
                                  HFX_JK_CLASS_TIMING: per-class times in the stats        */
} hfx_grad_args;

typedef struct hfx_grad_stats {
    int64_t n_quartets;        /* shell quartets differentiated                            */
    int64_t n_fp32;            /* of which differentiated in FP32 (MIXED)                  */
    double device_ms;          /* device time of the two-electron part                     */
    double class_ms[21];       /* with HFX_JK_CLASS_TIMING in the flags: device time of each
                                  quartet class (both precisions), in hfx_jk_stats' order   */
    int64_t class_prims[21];   /* primitive quartets differentiated in each class          */
} hfx_grad_stats;

/* grad[3*n_atoms] (x,y,z per atom).  parts, when not NULL, receives 4 x 3*n_atoms:
 * nuclear repulsion, -W.dS, P.d(T+V) (basis and operator derivatives), two-electron.
 * Single-molecule engines only.  stats may be NULL. */
int hfx_rhf_gradient(hfx_engine* e, const hfx_grad_args* args, double* grad, double* parts,
                     hfx_grad_stats* stats);

/* Thread-local description of the last error on this thread. */
const char* hfx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HFX_H */
