// scf_device.hpp - the seam between the host SCF driver (hfx_scf.cpp) and the libraries it is
// linked into.  The HIP library defines these functions in engine.hip (set_error) and
// hfx_scf_gpu.hip (device_cycles); a library without device linear algebra defines them to
// refuse.
#pragma once

#include <string>

#include "../../include/hfx_scf.h"

namespace hfx_internal {

// Records msg as the thread's hfx_last_error() text and returns code.
int set_error(int code, const std::string& msg);

// The device an engine runs on (HIP library only).
int engine_device(const hfx_engine* e, int* device);

// What the host driver has prepared before the SCF cycles start (SCF::initialize,
// src/SCF.cpp:138-162): sizes, electron counts, the nuclear repulsion, and the host copies of
// S, h = T + V (n x n, column-major) and the Schwarz table (nS x nS).
struct ScfSetup {
    int n, nS, na, nb;
    bool uhf;
    int precision;
    double e_nuc;
    const double* S;
    const double* h;
    const double* Q;
};

// The SCF cycles of SCF::run (src/SCF.cpp:51-114) with the iteration's linear algebra on the
// engine's device: fills every field of res except time_total_s; time_init_s covers S^-1/2 and
// the core-Hamiltonian guess.
int device_cycles(hfx_engine* e, const hfx_scf_options& o, const ScfSetup& su, hfx_scf_result* res);

}  // namespace hfx_internal
