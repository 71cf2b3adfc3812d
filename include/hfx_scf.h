/*
 * hfx_scf.h - host SCF driver over the hfx.h engine ABI.
 *
 * The reference's SCF class (include/SCF.hpp, src/SCF.cpp) owns the iteration loop, the
 * core-Hamiltonian guess, DIIS, damping and level shifting, and calls the integral engine
 * for the one-electron matrices, the Schwarz table and, every iteration, the direct Fock
 * build.  This driver is the host counterpart of that class, written against hfx.h only: it
 * runs unchanged over the HIP engine or over any other library exporting hfx.h.
 */
#ifndef HFX_SCF_H
#define HFX_SCF_H

#include <stdint.h>

#include "hfx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SCFOptions (include/SCF.hpp:9-32) plus the engine's precision controls. */
typedef struct hfx_scf_options {
    int32_t max_iter;          /* maxIter            30      */
    double energy_tol;         /* energyTol          1e-6    */
    double density_tol;        /* densityTol         1e-6    */
    int32_t use_diis;          /* useDIIS            1       */
    int32_t diis_size;         /* DIISmaxSize        6       */
    double diis_error_tol;     /* DIISErrorTol       1e-6    */
    double schwarz_thresh;     /* schwartzThreshold  1e-10   */
    double density_thresh;     /* densityThreshold   1e-10   */
    int32_t unrestricted;      /* -1: multiplicity != 1 (Input.cpp:127-130), 0 RHF, 1 UHF */
    int32_t guess_mix;         /* guessMix (0..10)   0       */
    int32_t damp;              /* damp, percent      0       */
    int32_t max_damp_iter;     /* maxDampIter        30      */
    double stop_damp_thresh;   /* stopDampThresh     0       */
    double level_shift;        /* levelShift         0       */
    double lshift_gap_tol;     /* lshiftGapTol       +inf    */
    int32_t max_lshift_iter;   /* maxLshiftIter      30      */
    double stop_lshift_thresh; /* stopLshiftThresh   0       */
    int32_t precision;         /* HFX_FP64 or HFX_MIXED for the incremental builds          */
    double fp32_cut;           /* MIXED: FP32 below this quartet bound                      */
    int32_t final_fp64;        /* MIXED: after convergence, restart from a full FP64 build  */
                               /* and iterate to convergence again                         */
    int32_t jk_flags;          /* extra HFX_JK_* flags for every build                      */
    int32_t full_builds;       /* 1: every build from the full density (no increments)      */
    int32_t direct;            /* DIRECT             1: direct builds; 0: conventional, the  */
                               /* integrals stored once (hfx_store_eris) and every build from */
                               /* the full density through them (FP64; src/SCF.cpp:57-60)     */
    double switch_tol;         /* MIXED + final_fp64: leave the mixed builds once |dE| falls  */
                               /* below this, then continue with FP64 builds (the first one   */
                               /* full) keeping the DIIS history; 0: converge in mixed        */
                               /* precision first, then restart FP64 with a fresh DIIS        */
    int32_t device_linalg;     /* 1: the iteration's linear algebra (DIIS, X^T F X, the       */
                               /* eigensolver, densities, energy) on the engine's device      */
                               /* (rocBLAS/rocSOLVER, loaded on first use) with the builds    */
                               /* stream-ordered through hfx_jk_async; the host reads a few   */
                               /* scalars per iteration.  Single-molecule runs; diis_size     */
                               /* 1..15.  0 (default): the host driver's own linear algebra   */
    const double* orbital_space; /* NULL (default): the orbitals span all N AOs.  Otherwise an */
                                 /* N x M column-major matrix whose columns span the orbital  */
                                 /* space, e.g. the Cartesian-to-spherical transform of the    */
                                 /* shells (hfx.Basis.spherical_transform): X = T (T^T S T)^-1/2, */
                                 /* every matrix stays in the Cartesian AO basis, the DIIS     */
                                 /* error is X^T (FDS - SDF) X.  Host linear algebra, single-  */
                                 /* molecule runs only                                        */
    int32_t n_orbital_space;     /* M, 1 <= M <= N                                            */
} hfx_scf_options;

typedef struct hfx_scf_result {
    double e_total, e_electronic, e_nuclear;
    int32_t iterations, converged;
    double last_dE, last_dD, last_diis_err;
    double e_total_mixed;        /* MIXED + final_fp64: energy at the mixed-precision stop */
    int32_t iterations_mixed;
    double time_total_s, time_init_s, time_jk_s, time_diag_s;
    int64_t n_jk_builds;
    int64_t n_quartets_last;     /* survivors of the last build                             */
    /* Optional outputs, caller-allocated (N x N or N), may be NULL. */
    double* F_alpha;
    double* D_alpha;
    double* C_alpha;
    double* eps_alpha;
    double* F_beta;
    double* D_beta;
    double* C_beta;
    double* eps_beta;
    double* h_core;
    double* S;
    /* Optional per-iteration trace: trace[4*it + {0,1,2,3}] = E_elec, dE, dD, DIIS error. */
    double* trace;
    int32_t trace_cap;           /* number of iterations the trace can hold                 */
    double time_linalg_load_s;   /* device_linalg: seconds spent loading rocBLAS/rocSOLVER   */
                                 /* and creating the device's handle (first run in a process */
                                 /* only; inside time_init_s and time_total_s)               */
    int32_t n_mo;                /* orbitals: N, or M with an orbital_space; C and eps keep  */
                                 /* N columns, the last N - M with zero coefficients and an  */
                                 /* orbital energy of 1e6                                     */
} hfx_scf_result;

void hfx_scf_default_options(hfx_scf_options* opt);

/* SCF::run (src/SCF.cpp:23-136) on the molecule (Z, xyz in Bohr) in the given basis. */
int hfx_scf_run(const hfx_basis* basis, int32_t n_atoms, const double* Z, const double* xyz,
                int32_t charge, int32_t multiplicity, const hfx_scf_options* opt, int device,
                hfx_scf_result* res);

/* The same loop over an engine the caller created (the engine is not destroyed).  With
 * opt->direct = 0 the run stores its own integrals on that engine (hfx_store_eris with the
 * engine's Schwarz table and opt->schwarz_thresh), replacing any the caller stored before,
 * and releases them when it returns: a caller that keeps stored integrals on an engine should
 * not pass it to a conventional run. */
int hfx_scf_run_engine(hfx_engine* e, int32_t n_atoms, const double* Z, const double* xyz,
                       int32_t charge, int32_t multiplicity, const hfx_scf_options* opt,
                       hfx_scf_result* res);

/* Many independent molecules through one batched engine (hfx_create_batched): each molecule
 * runs SCF::run (src/SCF.cpp:23-136) with its own DIIS/damping/level-shift state, and every
 * iteration's Fock builds of all molecules still iterating are ONE hfx_jk call.  Atoms of all
 * molecules are concatenated in molecule order, atom_mol[i] names the molecule of atom i;
 * charge and multiplicity are per molecule (NULL: neutral singlets); res holds n_mol results.
 * time_jk_s and n_jk_builds count the shared batched builds.  No reference counterpart (the
 * reference runs one molecule per process). */
int hfx_scf_run_batched(int32_t n_mol, const hfx_basis* bases, int32_t n_atoms, const double* Z,
                        const double* xyz, const int32_t* atom_mol, const int32_t* charge,
                        const int32_t* multiplicity, const hfx_scf_options* opt, int device,
                        hfx_scf_result* res);

/* The same over a batched engine the caller created; n_aos_mol[m] = AOs of molecule m.
 * opt->direct = 0 replaces and then releases the engine's stored integrals, as above. */
int hfx_scf_run_batched_engine(hfx_engine* e, int32_t n_mol, const int32_t* n_aos_mol, int32_t n_atoms,
                               const double* Z, const double* xyz, const int32_t* atom_mol,
                               const int32_t* charge, const int32_t* multiplicity,
                               const hfx_scf_options* opt, hfx_scf_result* res);

#ifdef __cplusplus
}
#endif

#endif /* HFX_SCF_H */
