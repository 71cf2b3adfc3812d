// hfx_scf.cpp - host SCF driver over the engine ABI (include/hfx_scf.h).
//
// Host counterpart of the reference's SCF class, restated without Eigen:
//   SCF::run ............................ src/SCF.cpp:23-136
//   SCF::initialize ..................... src/SCF.cpp:138-162
//   SCF::computeInitialGuessDensity ..... src/SCF.cpp:165-231
//   incremental Fock update ............. src/SCF.cpp:536-541, 731-743 (J/K from the engine)
//   SCF::diagonalizeAndUpdate ........... src/SCF.cpp:747-835
//   DIIS update/extrapolate/B matrix .... src/DIIS.cpp:11-81
//   inverseSqrtMatrix ................... src/Utils.cpp:117-138
// Linear algebra (eigensolver, QR with column pivoting) comes from linalg.hpp.

#include "../../include/hfx_scf.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <string>
#include <vector>

#include "linalg.hpp"
#include "scf_device.hpp"

namespace {

using Mat = std::vector<double>;
using clk = std::chrono::steady_clock;

double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

double trace_prod(int n, const double* A, const double* B) {  // tr(A B)
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) s += A[i + (size_t)k * n] * B[k + (size_t)i * n];
    return s;
}

// Orbital energy given to the directions outside an orbital subspace.
constexpr double REMOVED_EPS = 1.0e6;

double frob(const Mat& A) {
    double s = 0.0;
    for (double x : A) s += x * x;
    return std::sqrt(s);
}

// inverseSqrtMatrix (Utils.cpp:117-138): V diag(1/sqrt(w), 0 for w <= 0) V^T.
Mat inv_sqrt(int n, const Mat& S) {
    Mat V((size_t)n * n), w(n), T((size_t)n * n), X((size_t)n * n);
    hfx_la::eigh(n, S.data(), w.data(), V.data());
    for (int c = 0; c < n; ++c) {
        const double f = w[c] > 0 ? 1.0 / std::sqrt(w[c]) : 0.0;
        for (int r = 0; r < n; ++r) T[r + (size_t)c * n] = V[r + (size_t)c * n] * f;
    }
    hfx_la::mm(n, T.data(), V.data(), X.data(), false, true);
    return X;
}

// Orthogonalizer of the orbital space spanned by the M columns of Tm (n x M, column-major):
// X = Tm (Tm^T S Tm)^-1/2 in the first M columns, zero columns after them (n x n).  With
// Tm = 1 this is inverseSqrtMatrix (Utils.cpp:117-138); with the Cartesian-to-spherical
// transform of the shells, the SCF runs in the spherical-harmonic orbital space while every
// matrix stays in the engine's Cartesian AO basis.
Mat subspace_inv_sqrt(int n, int M, const double* Tm, const Mat& S) {
    Mat ST((size_t)n * M), Ss((size_t)M * M), X((size_t)n * n, 0.0);
    hfx_la::gemm(false, false, n, M, n, 1.0, S.data(), n, Tm, n, 0.0, ST.data(), n);
    hfx_la::gemm(true, false, M, M, n, 1.0, Tm, n, ST.data(), n, 0.0, Ss.data(), M);
    const Mat Xs = inv_sqrt(M, Ss);
    hfx_la::gemm(false, false, n, M, M, 1.0, Tm, n, Xs.data(), M, 0.0, X.data(), n);
    return X;
}

// DIIS (src/DIIS.cpp): FIFO of at most maxSize error vectors and Fock matrices.
struct Diis {
    size_t max_size;
    std::deque<Mat> err, fa, fb;
    explicit Diis(size_t m) : max_size(m) {}
    void update(const Mat& Fa, const Mat& Fb, const Mat& e) {
        if (err.size() >= max_size) {
            err.pop_front();
            fa.pop_front();
            fb.pop_front();
        }
        err.push_back(e);
        fa.push_back(Fa);
        fb.push_back(Fb);
    }
    void extrapolate(Mat& Fa, Mat& Fb) const {
        if (err.empty()) return;
        const int m = (int)err.size() + 1;
        Mat B((size_t)m * m, 0.0);
        for (int i = 0; i + 1 < m; ++i)
            for (int j = 0; j <= i; ++j) {
                double s = 0.0;
                const Mat &ei = err[i], &ej = err[j];
                for (size_t x = 0; x < ei.size(); ++x) s += ei[x] * ej[x];
                B[i + (size_t)j * m] = B[j + (size_t)i * m] = s;
            }
        for (int i = 0; i < m; ++i) {
            B[(m - 1) + (size_t)i * m] = -1.0;
            B[i + (size_t)(m - 1) * m] = -1.0;
        }
        B[(m - 1) + (size_t)(m - 1) * m] = 0.0;
        std::vector<double> b(m, 0.0);
        b[m - 1] = -1.0;
        const std::vector<double> c = hfx_la::lstsq_qrcp(m, m, B, b);
        std::fill(Fa.begin(), Fa.end(), 0.0);
        std::fill(Fb.begin(), Fb.end(), 0.0);
        for (int i = 0; i + 1 < m; ++i) {
            const Mat &ai = fa[i], &bi = fb[i];
            for (size_t x = 0; x < Fa.size(); ++x) {
                Fa[x] += c[i] * ai[x];
                Fb[x] += c[i] * bi[x];
            }
        }
    }
    double err_norm() const { return err.empty() ? 0.0 : frob(err.back()); }
};

struct Scf {
    hfx_engine* e = nullptr;
    hfx_scf_options o{};
    int n = 0, nS = 0, na = 0, nb = 0;
    int m = 0;  // orbitals: n, or the M of an orbital subspace (opt->orbital_space)
    bool uhf = false;
    double e_nuc = 0, e_elec = 0, dE = 0, dD = 0, diis_err = 0, damp_c = 0;
    bool ls_a = false, ls_b = false;
    int iteration = 0;
    Mat S, X, h, Q;
    Mat Fa, Fb, Fa_prev, Fb_prev, Ca, Cb, Da, Db, Da_prev, Db_prev, Dt, Dt_prev;
    std::vector<double> eps_a, eps_b;
    double t_jk = 0, t_diag = 0;
    int64_t n_builds = 0, n_quart_last = 0;
    int precision = HFX_FP64;

    Mat mat() const { return Mat((size_t)n * n, 0.0); }

    // X^T F X (+ level shift) -> eigenpairs -> C = X V (SCF.cpp:775-790).
    void diag(const Mat& F, bool shift, const Mat& Cold, int nocc, Mat& C, std::vector<double>& w) {
        const auto t0 = clk::now();
        Mat T = mat(), Fp = mat();
        hfx_la::mm(n, F.data(), X.data(), T.data());
        hfx_la::mm(n, X.data(), T.data(), Fp.data(), true, false);
        if (shift) {
            // F' += (X S C) diag(shift on virtuals) (X S C)^T  (SCF.cpp:781-786)
            Mat SC = mat(), XSC = mat(), XSCs = mat(), add = mat();
            hfx_la::mm(n, S.data(), Cold.data(), SC.data());
            hfx_la::mm(n, X.data(), SC.data(), XSC.data());
            XSCs = XSC;
            for (int c = 0; c < n; ++c)
                for (int r = 0; r < n; ++r) XSCs[r + (size_t)c * n] *= (c >= nocc) ? o.level_shift : 0.0;
            hfx_la::mm(n, XSCs.data(), XSC.data(), add.data(), false, true);
            for (size_t x = 0; x < Fp.size(); ++x) Fp[x] += add[x];
        }
        Mat V = mat();
        w.assign(n, 0.0);
        if (m < n) {
            // an orbital subspace: X has zero columns after the first m, so the eigenproblem is
            // the leading m x m block; the other directions keep zero coefficient columns and a
            // marker orbital energy above the spectrum
            Mat Fs((size_t)m * m), Vs((size_t)m * m);
            std::vector<double> ws(m);
            for (int c = 0; c < m; ++c)
                for (int r = 0; r < m; ++r) Fs[r + (size_t)c * m] = Fp[r + (size_t)c * n];
            hfx_la::eigh(m, Fs.data(), ws.data(), Vs.data());
            for (int c = 0; c < m; ++c) {
                w[c] = ws[c];
                for (int r = 0; r < m; ++r) V[r + (size_t)c * n] = Vs[r + (size_t)c * m];
            }
            for (int c = m; c < n; ++c) {
                w[c] = REMOVED_EPS;
                V[c + (size_t)c * n] = 1.0;
            }
        } else {
            hfx_la::eigh(n, Fp.data(), w.data(), V.data());
        }
        C = mat();
        hfx_la::mm(n, X.data(), V.data(), C.data());
        t_diag += secs(t0, clk::now());
    }

    // C_occ C_occ^T
    Mat density(const Mat& C, int nocc) {
        Mat D = mat();
        if (nocc > 0) hfx_la::gemm(false, true, n, n, nocc, 1.0, C.data(), n, C.data(), n, 0.0, D.data(), n);
        return D;
    }

    double energy() const {  // E_elec = 1/2 tr(Da (h + Fa)) + 1/2 tr(Db (h + Fb))  (SCF.cpp:833-834)
        Mat ha((size_t)n * n), hb((size_t)n * n);
        for (size_t x = 0; x < ha.size(); ++x) {
            ha[x] = h[x] + Fa[x];
            hb[x] = h[x] + Fb[x];
        }
        return 0.5 * trace_prod(n, Da.data(), ha.data()) + 0.5 * trace_prod(n, Db.data(), hb.data());
    }

    // The density increments of this iteration (SCF.cpp:536-541); with full, the increments
    // restart from zero (F_prev = h, D_prev = 0), i.e. a full rebuild.
    void deltas(bool full, double* dDt, double* dDa, double* dDb) {
        if (full) {
            std::fill(Dt_prev.begin(), Dt_prev.end(), 0.0);
            std::fill(Da_prev.begin(), Da_prev.end(), 0.0);
            std::fill(Db_prev.begin(), Db_prev.end(), 0.0);
            Fa_prev = h;
            Fb_prev = h;
        }
        for (size_t x = 0; x < Dt.size(); ++x) {
            dDt[x] = Dt[x] - Dt_prev[x];
            dDa[x] = Da[x] - Da_prev[x];
            if (uhf) dDb[x] = Db[x] - Db_prev[x];
        }
    }

    // F = F_prev + J - K; F_prev = F (SCF.cpp:731-743).
    void apply_jk(const double* J, const double* Ka, const double* Kb) {
        for (size_t x = 0; x < Fa.size(); ++x) Fa[x] = Fa_prev[x] + J[x] - Ka[x];
        Fa_prev = Fa;
        if (uhf) {
            for (size_t x = 0; x < Fb.size(); ++x) Fb[x] = Fb_prev[x] + J[x] - Kb[x];
            Fb_prev = Fb;
        } else {
            Fb = Fa;
            Fb_prev = Fa_prev;
        }
    }

    int build_fock(bool full) {  // SCF::buildFockMatrixDirect (SCF.cpp:523-744) via hfx_jk
        Mat dDt = mat(), dDa = mat(), dDb;
        if (uhf) dDb = mat();
        deltas(full, dDt.data(), dDa.data(), uhf ? dDb.data() : nullptr);
        Mat J = mat(), Ka = mat(), Kb;
        if (uhf) Kb = mat();
        hfx_jk_args a{};
        a.dD_tot = dDt.data();
        a.dD_alpha = dDa.data();
        a.dD_beta = uhf ? dDb.data() : nullptr;
        a.Q = Q.data();
        a.schwarz_thresh = o.schwarz_thresh;
        a.density_thresh = o.density_thresh;
        a.precision = precision;
        a.fp32_cut = o.fp32_cut;
        a.flags = o.jk_flags;
        hfx_jk_stats st{};
        const auto t0 = clk::now();
        const int rc = hfx_jk(e, &a, J.data(), Ka.data(), uhf ? Kb.data() : nullptr, &st);
        t_jk += secs(t0, clk::now());
        if (rc) return rc;
        n_builds++;
        n_quart_last = st.n_quartets;
        apply_jk(J.data(), Ka.data(), uhf ? Kb.data() : nullptr);
        return HFX_OK;
    }

    // Everything after the Fock build in one SCF cycle (SCF.cpp:60-91): DIIS update and
    // extrapolation on FDS - SDF, diagonalization, energy, convergence test.
    bool after_build(Diis* diis, double last, hfx_scf_result* res, int slot) {
        if (diis) {
            Mat t1 = mat(), t2 = mat(), ea = mat(), eb = mat();
            hfx_la::mm(n, Fa.data(), Da.data(), t1.data());
            hfx_la::mm(n, t1.data(), S.data(), t2.data());
            hfx_la::mm(n, S.data(), Da.data(), t1.data());
            hfx_la::mm(n, t1.data(), Fa.data(), ea.data());
            for (size_t x = 0; x < ea.size(); ++x) ea[x] = t2[x] - ea[x];
            hfx_la::mm(n, Fb.data(), Db.data(), t1.data());
            hfx_la::mm(n, t1.data(), S.data(), t2.data());
            hfx_la::mm(n, S.data(), Db.data(), t1.data());
            hfx_la::mm(n, t1.data(), Fb.data(), eb.data());
            for (size_t x = 0; x < eb.size(); ++x) eb[x] = t2[x] - eb[x] + ea[x];
            if (m < n) {
                // in an orbital subspace the AO commutator need not vanish at convergence; its
                // S-projection P^T (FDS - SDF) P onto the subspace (P = X X^T S) does.  For the
                // full space P = 1 and this is the AO commutator itself, so DIIS keeps the
                // reference's metric.
                hfx_la::mm(n, eb.data(), X.data(), t1.data());
                hfx_la::mm(n, X.data(), t1.data(), t2.data(), true, false);  // X^T e X
                hfx_la::mm(n, S.data(), X.data(), ea.data());                // S X
                hfx_la::mm(n, ea.data(), t2.data(), t1.data());
                hfx_la::mm(n, t1.data(), ea.data(), eb.data(), false, true);
            }
            diis->update(Fa, Fb, eb);
            diis->extrapolate(Fa, Fb);
        }
        diagonalize_and_update();
        dE = std::abs(e_elec - last);
        double sum = 0.0;
        for (size_t x = 0; x < Dt.size(); ++x) {
            const double d = Dt[x] - Dt_prev[x];
            sum += d * d;
        }
        dD = std::sqrt(sum);
        diis_err = diis ? diis->err_norm() : 0.0;
        if (res && res->trace && slot < res->trace_cap) {
            res->trace[4 * slot + 0] = e_elec;
            res->trace[4 * slot + 1] = dE;
            res->trace[4 * slot + 2] = dD;
            res->trace[4 * slot + 3] = diis_err;
        }
        return dE < o.energy_tol && dD < o.density_tol && (!diis || diis_err < o.diis_error_tol);
    }

    void diagonalize_and_update() {  // SCF.cpp:747-835
        if (o.damp > 0 && o.max_damp_iter != 0 && o.max_damp_iter > iteration &&
            (dE > o.stop_damp_thresh || iteration == 0))
            damp_c = o.damp / 100.0;
        else
            damp_c = 0.0;
        if (o.level_shift > 0 && iteration < o.max_lshift_iter && (dE > o.stop_lshift_thresh || iteration == 0)) {
            const double ga = (na < n && na > 0) ? eps_a[na] - eps_a[na - 1] : std::numeric_limits<double>::infinity();
            const double gb = (nb < n && nb > 0) ? eps_b[nb] - eps_b[nb - 1] : std::numeric_limits<double>::infinity();
            ls_a = ga < o.lshift_gap_tol;
            ls_b = gb < o.lshift_gap_tol;
        } else {
            ls_a = ls_b = false;
        }
        Mat Cnew;
        diag(Fa, ls_a, Ca, na, Cnew, eps_a);
        Ca = Cnew;
        Da_prev = Da;
        {
            const Mat P = density(Ca, na);
            for (size_t x = 0; x < Da.size(); ++x) Da[x] = (1 - damp_c) * P[x] + damp_c * Da[x];
        }
        Db_prev = Db;
        if (uhf) {
            diag(Fb, ls_b, Cb, nb, Cnew, eps_b);
            Cb = Cnew;
            const Mat P = density(Cb, nb);
            for (size_t x = 0; x < Db.size(); ++x) Db[x] = (1 - damp_c) * P[x] + damp_c * Db[x];
        } else {
            Cb = Ca;
            eps_b = eps_a;
            Db = Da;
        }
        Dt_prev = Dt;
        for (size_t x = 0; x < Dt.size(); ++x) Dt[x] = Da[x] + Db[x];
        e_elec = energy();
    }

    void initial_guess() {  // SCF.cpp:165-231
        Fa = h;
        Fa_prev = Fa;
        diag(Fa, false, Mat(), na, Ca, eps_a);
        Da_prev = mat();
        Da = density(Ca, na);
        Db_prev = mat();
        if (uhf) {
            if (o.guess_mix > 0 && na >= 1 && na < m) {
                const double k = o.guess_mix / 10.0;
                const double f = 1.0 / std::sqrt(1 + k * k);
                for (int r = 0; r < n; ++r) {
                    const double homo = Ca[r + (size_t)(na - 1) * n], lumo = Ca[r + (size_t)na * n];
                    Ca[r + (size_t)(na - 1) * n] = f * (homo + k * lumo);
                    Ca[r + (size_t)na * n] = f * (lumo - k * homo);
                }
                Da = density(Ca, na);
            }
            Fb = h;
            Fb_prev = Fb;
            diag(Fb, false, Mat(), nb, Cb, eps_b);
            Db = density(Cb, nb);
        } else {
            Fb = Fa;
            Fb_prev = Fa_prev;
            Cb = Ca;
            eps_b = eps_a;
            Db = Da;
        }
        Dt_prev = mat();
        Dt = mat();
        for (size_t x = 0; x < Dt.size(); ++x) Dt[x] = Da[x] + Db[x];
        e_elec = energy();
    }

    // The SCF cycles of SCF::run (SCF.cpp:51-114).
    // One stage of SCF cycles; with switch_dE > 0 the stage also ends (not converged) once
    // |dE| < switch_dE after its first iteration.
    int cycles(Diis* diis, bool first_full, int& it_done, bool& conv, hfx_scf_result* res, int trace_base,
               double switch_dE = 0.0) {
        conv = false;
        for (iteration = 0; iteration < o.max_iter; ++iteration) {
            const double last = e_elec;
            const int rc = build_fock(o.full_builds || (first_full && iteration == 0));
            if (rc) return rc;
            if (after_build(diis, last, res, trace_base + iteration)) {
                conv = true;
                it_done = iteration + 1;
                return HFX_OK;
            }
            if (switch_dE > 0.0 && iteration > 0 && dE < switch_dE) {
                it_done = iteration + 1;
                return HFX_OK;
            }
        }
        it_done = o.max_iter;
        return HFX_OK;
    }

    void remove_level_shift() {  // SCF.cpp:93-107
        if (ls_a)
            for (int i = na; i < m; ++i) eps_a[i] -= o.level_shift;
        if (ls_b)
            for (int i = nb; i < m; ++i) eps_b[i] -= o.level_shift;
    }

    void write_result(hfx_scf_result* res) const {
        res->e_electronic = e_elec;
        res->e_nuclear = e_nuc;
        res->e_total = e_elec + e_nuc;
        res->last_dE = dE;
        res->last_dD = dD;
        res->last_diis_err = diis_err;
        res->time_jk_s = t_jk;
        res->time_diag_s = t_diag;
        res->n_jk_builds = n_builds;
        res->n_quartets_last = n_quart_last;
        res->time_linalg_load_s = 0.0;
        res->n_mo = m;
        auto put = [&](double* dst, const Mat& src) {
            if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(double));
        };
        put(res->F_alpha, Fa);
        put(res->D_alpha, Da);
        put(res->C_alpha, Ca);
        put(res->F_beta, Fb);
        put(res->D_beta, Db);
        put(res->C_beta, Cb);
        put(res->h_core, h);
        put(res->S, S);
        if (res->eps_alpha) std::memcpy(res->eps_alpha, eps_a.data(), n * sizeof(double));
        if (res->eps_beta) std::memcpy(res->eps_beta, eps_b.data(), n * sizeof(double));
    }

    // Electron counts (SCF.cpp:140-142, Molecule.cpp:67-72) and the nuclear repulsion
    // (Molecule.cpp:75-89).
    int setup_molecule(int n_atoms, const double* Z, const double* xyz, int charge, int mult) {
        long nel = -charge;
        for (int i = 0; i < n_atoms; ++i) nel += (long)std::lround(Z[i]);
        if (mult < 1 || nel < 0 || ((nel + mult - 1) % 2) != 0 || nel + 1 < mult) return HFX_ERR_INVALID;
        na = (int)((nel + mult - 1) / 2);
        nb = (int)((nel - mult + 1) / 2);
        m = n;
        if (o.orbital_space) {
            if (o.n_orbital_space <= 0 || o.n_orbital_space > n) return HFX_ERR_INVALID;
            m = o.n_orbital_space;
        }
        if (na > m) return HFX_ERR_INVALID;
        uhf = o.unrestricted < 0 ? (mult != 1) : (o.unrestricted != 0);
        precision = o.precision;
        double en = 0.0;
        for (int i = 0; i < n_atoms; ++i)
            for (int j = i + 1; j < n_atoms; ++j) {
                const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                             dz = xyz[3 * i + 2] - xyz[3 * j + 2];
                en += (Z[i] * Z[j]) / std::sqrt(dx * dx + dy * dy + dz * dz);
            }
        e_nuc = en;
        return HFX_OK;
    }
};

// Conventional mode (DIRECT false, src/SCF.cpp:57-60, 159): every build from the full density
// through the integrals stored once per run; they are FP64, so the builds are too.
void conventional_options(hfx_scf_options& o) {
    if (o.direct) return;
    o.precision = HFX_FP64;
    o.final_fp64 = 0;
    o.full_builds = 1;
    o.jk_flags |= HFX_JK_STORED;
}

// Frees the engine's stored integrals when the run that stored them ends.
struct StoreGuard {
    hfx_engine* e = nullptr;
    ~StoreGuard() {
        if (e) (void)hfx_release_eris(e);
    }
};

}  // namespace

extern "C" {

void hfx_scf_default_options(hfx_scf_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->max_iter = 30;
    o->energy_tol = 1.0e-6;
    o->density_tol = 1.0e-6;
    o->use_diis = 1;
    o->diis_size = 6;
    o->diis_error_tol = 1.0e-6;
    o->schwarz_thresh = 1.0e-10;
    o->density_thresh = 1.0e-10;
    o->unrestricted = -1;
    o->guess_mix = 0;
    o->damp = 0;
    o->max_damp_iter = 30;
    o->stop_damp_thresh = 0.0;
    o->level_shift = 0.0;
    o->lshift_gap_tol = std::numeric_limits<double>::infinity();
    o->max_lshift_iter = 30;
    o->stop_lshift_thresh = 0.0;
    o->precision = HFX_FP64;
    o->fp32_cut = 0.0;
    o->final_fp64 = 1;
    o->jk_flags = 0;
    o->full_builds = 0;
    o->direct = 1;
    o->switch_tol = 1.0e-4;
}

int hfx_scf_run_engine(hfx_engine* e, int32_t n_atoms, const double* Z, const double* xyz, int32_t charge,
                       int32_t mult, const hfx_scf_options* opt, hfx_scf_result* res) {
    if (!e || !res || n_atoms <= 0 || !Z || !xyz) return HFX_ERR_INVALID;
    const auto t0 = clk::now();
    Scf s;
    s.e = e;
    if (opt)
        s.o = *opt;
    else
        hfx_scf_default_options(&s.o);
    conventional_options(s.o);
    int64_t N = 0, nS = 0, nmol = 0, elems = 0;
    int rc = hfx_sizes(e, &N, &nS, &nmol, &elems);
    if (rc) return rc;
    if (nmol != 1) return HFX_ERR_INVALID;
    s.n = (int)N;
    s.nS = (int)nS;
    if ((rc = s.setup_molecule(n_atoms, Z, xyz, charge, mult))) return rc;

    // SCF::initialize (SCF.cpp:138-162).
    Mat T = s.mat(), V = s.mat();
    s.S = s.mat();
    s.Q.assign((size_t)nS * nS, 0.0);
    if ((rc = hfx_kinetic(e, T.data())) || (rc = hfx_nuclear(e, n_atoms, Z, xyz, nullptr, V.data())) ||
        (rc = hfx_overlap(e, s.S.data())) || (rc = hfx_schwarz(e, s.Q.data())))
        return rc;
    StoreGuard store;
    if (!s.o.direct) {  // IntegralEngine::electronRepulsionTensor (SCF.cpp:159)
        if ((rc = hfx_store_eris(e, s.Q.data(), s.o.schwarz_thresh, nullptr, nullptr))) return rc;
        store.e = e;
    }
    s.h = s.mat();
    for (size_t x = 0; x < s.h.size(); ++x) s.h[x] = T[x] + V[x];
    if (s.o.device_linalg && s.o.orbital_space)
        return hfx_internal::set_error(HFX_ERR_UNSUPPORTED, "hfx_scf_run: orbital_space needs the host linear algebra "
                                                           "(device_linalg = 0)");
    if (s.o.device_linalg) {  // the cycles with the iteration's linear algebra on the device
        const hfx_internal::ScfSetup su{s.n, s.nS, s.na, s.nb, s.uhf, s.precision, s.e_nuc, s.S.data(), s.h.data(),
                                        s.Q.data()};
        const double t_setup = secs(t0, clk::now());
        if ((rc = hfx_internal::device_cycles(e, s.o, su, res))) return rc;
        res->time_init_s += t_setup;
        res->time_total_s = secs(t0, clk::now());
        return HFX_OK;
    }
    s.X = s.o.orbital_space ? subspace_inv_sqrt(s.n, s.m, s.o.orbital_space, s.S) : inv_sqrt(s.n, s.S);
    s.initial_guess();
    const auto t_init = clk::now();

    Diis diis_mixed(s.o.diis_size);
    int its = 0;
    bool conv = false;
    const bool final64 = s.precision == HFX_MIXED && s.o.final_fp64;
    const bool early = final64 && s.o.switch_tol > 0.0;
    rc = s.cycles(s.o.use_diis ? &diis_mixed : nullptr, false, its, conv, res, 0, early ? s.o.switch_tol : 0.0);
    if (rc) return rc;
    int total_its = its;
    res->e_total_mixed = s.e_elec + s.e_nuc;
    res->iterations_mixed = its;
    if (final64) {
        // Restart the incremental sequence from a full FP64 build of the current density and
        // iterate to convergence in FP64: after an early switch the DIIS history of the mixed
        // builds is kept (their Fock matrices differ from FP64 ones far below the DIIS errors
        // at that point), otherwise a fresh DIIS starts.
        s.precision = HFX_FP64;
        Diis diis2(s.o.diis_size);
        int its2 = 0;
        rc = s.cycles(s.o.use_diis ? (early ? &diis_mixed : &diis2) : nullptr, true, its2, conv, res, its);
        if (rc) return rc;
        total_its += its2;
    }
    s.remove_level_shift();
    s.write_result(res);
    res->iterations = total_its;
    res->converged = conv ? 1 : 0;
    res->time_init_s = secs(t0, t_init);
    res->time_total_s = secs(t0, clk::now());
    return HFX_OK;
}

int hfx_scf_run(const hfx_basis* basis, int32_t n_atoms, const double* Z, const double* xyz, int32_t charge,
                int32_t mult, const hfx_scf_options* opt, int device, hfx_scf_result* res) {
    hfx_engine* e = nullptr;
    const auto t0 = clk::now();
    int rc = hfx_create(basis, device, &e);
    if (rc) return rc;
    const double t_create = secs(t0, clk::now());
    rc = hfx_scf_run_engine(e, n_atoms, Z, xyz, charge, mult, opt, res);
    hfx_destroy(e);
    if (rc == HFX_OK) {
        res->time_init_s += t_create;
        res->time_total_s += t_create;
    }
    return rc;
}

// Many molecules through one batched engine.  Each molecule runs the cycles of SCF::run
// (SCF.cpp:51-114) with its own DIIS, damping and level-shift state, but every iteration's
// Fock builds go out as ONE hfx_jk over the packed density increments of all molecules.  A
// molecule that has converged contributes a zero increment, which the density screening
// removes before any integral work, and its Fock matrix is left alone.
int hfx_scf_run_batched_engine(hfx_engine* e, int32_t n_mol, const int32_t* n_aos_mol, int32_t n_atoms,
                               const double* Z, const double* xyz, const int32_t* atom_mol,
                               const int32_t* charge, const int32_t* mult, const hfx_scf_options* opt,
                               hfx_scf_result* res) {
    if (!e || !res || n_mol <= 0 || !n_aos_mol || n_atoms <= 0 || !Z || !xyz || !atom_mol)
        return HFX_ERR_INVALID;
    const auto t0 = clk::now();
    int64_t N = 0, nS = 0, nm = 0, elems = 0;
    int rc = hfx_sizes(e, &N, &nS, &nm, &elems);
    if (rc) return rc;
    if (nm != n_mol) return HFX_ERR_INVALID;
    if (opt && opt->orbital_space)
        return hfx_internal::set_error(HFX_ERR_UNSUPPORTED, "hfx_scf_run_batched: orbital_space is per molecule; "
                                                           "run such molecules one at a time");
    hfx_scf_options o;
    if (opt)
        o = *opt;
    else
        hfx_scf_default_options(&o);
    if (o.device_linalg)
        return hfx_internal::set_error(HFX_ERR_UNSUPPORTED, "hfx_scf_run_batched: device_linalg runs one molecule");
    conventional_options(o);

    // Per-molecule atoms (contiguous runs of atom_mol) and packed block offsets.
    std::vector<int> atom_begin(n_mol + 1, 0);
    for (int i = 0; i < n_atoms; ++i) {
        const int m = atom_mol[i];
        if (m < 0 || m >= n_mol || (i > 0 && m < atom_mol[i - 1])) return HFX_ERR_INVALID;
        atom_begin[m + 1] = i + 1;
    }
    for (int m = 0; m < n_mol; ++m)
        if (atom_begin[m + 1] < atom_begin[m]) atom_begin[m + 1] = atom_begin[m];
    std::vector<size_t> off(n_mol + 1, 0);
    for (int m = 0; m < n_mol; ++m) off[m + 1] = off[m] + (size_t)n_aos_mol[m] * n_aos_mol[m];
    if ((int64_t)off[n_mol] != elems) return HFX_ERR_INVALID;

    std::vector<Scf> mol(n_mol);
    for (int m = 0; m < n_mol; ++m) {
        Scf& s = mol[m];
        s.e = e;
        s.o = o;
        s.n = n_aos_mol[m];
        const int a0 = atom_begin[m], na = atom_begin[m + 1] - a0;
        if (na <= 0) return HFX_ERR_INVALID;
        if ((rc = s.setup_molecule(na, Z + a0, xyz + 3 * a0, charge ? charge[m] : 0, mult ? mult[m] : 1)))
            return rc;
    }
    bool any_uhf = false;
    for (const Scf& s : mol) any_uhf |= s.uhf;

    // SCF::initialize (SCF.cpp:138-162) for all molecules from the packed one-electron blocks.
    {
        Mat T(elems), V(elems), S(elems);
        if ((rc = hfx_kinetic(e, T.data())) || (rc = hfx_nuclear(e, n_atoms, Z, xyz, atom_mol, V.data())) ||
            (rc = hfx_overlap(e, S.data())))
            return rc;
#pragma omp parallel for schedule(dynamic, 1)
        for (int m = 0; m < n_mol; ++m) {
            Scf& s = mol[m];
            s.S.assign(S.begin() + off[m], S.begin() + off[m + 1]);
            s.h.resize(off[m + 1] - off[m]);
            for (size_t x = 0; x < s.h.size(); ++x) s.h[x] = T[off[m] + x] + V[off[m] + x];
            s.X = inv_sqrt(s.n, s.S);
            s.initial_guess();
        }
    }
    StoreGuard store;
    if (!o.direct) {  // the engine's own Schwarz table, as the batched builds use
        if ((rc = hfx_store_eris(e, nullptr, o.schwarz_thresh, nullptr, nullptr))) return rc;
        store.e = e;
    }
    const auto t_init = clk::now();

    Mat dDt(elems), dDa(elems), dDb(any_uhf ? elems : 0), J(elems), Ka(elems), Kb(any_uhf ? elems : 0);
    std::vector<int> its(n_mol, 0), its_mixed(n_mol, 0), done(n_mol, 0), conv(n_mol, 0);
    std::vector<double> e_mixed(n_mol, 0.0);
    double t_jk = 0.0;
    int64_t n_builds = 0, n_quart_last = 0;

    // One stage over all molecules; a molecule leaves it when converged or, with switch_dE > 0,
    // once its |dE| < switch_dE.  keep_diis continues the previous stage's DIIS histories.
    std::vector<Diis> diis;
    auto stage = [&](int precision, bool first_full, double switch_dE, bool keep_diis) -> int {
        if (!keep_diis) diis.assign(n_mol, Diis(o.diis_size));
        std::vector<int> active(n_mol, 1);
        for (int m = 0; m < n_mol; ++m) {
            mol[m].precision = precision;
            mol[m].iteration = 0;
        }
        for (int it = 0; it < o.max_iter; ++it) {
            int n_active = 0;
            for (int m = 0; m < n_mol; ++m) n_active += active[m];
            if (!n_active) break;
            std::fill(dDt.begin(), dDt.end(), 0.0);
            std::fill(dDa.begin(), dDa.end(), 0.0);
            std::fill(dDb.begin(), dDb.end(), 0.0);
            std::vector<double> last(n_mol);
#pragma omp parallel for schedule(dynamic, 1)
            for (int m = 0; m < n_mol; ++m) {
                if (!active[m]) continue;
                Scf& s = mol[m];
                last[m] = s.e_elec;
                double* db = any_uhf ? dDb.data() + off[m] : nullptr;
                s.deltas(o.full_builds || (first_full && it == 0), dDt.data() + off[m], dDa.data() + off[m],
                         s.uhf ? db : nullptr);
                // An RHF molecule in a batch with UHF molecules: beta increment = alpha increment.
                if (db && !s.uhf) std::copy(dDa.begin() + off[m], dDa.begin() + off[m + 1], dDb.begin() + off[m]);
            }
            hfx_jk_args a{};
            a.dD_tot = dDt.data();
            a.dD_alpha = dDa.data();
            a.dD_beta = any_uhf ? dDb.data() : nullptr;
            a.Q = nullptr;
            a.schwarz_thresh = o.schwarz_thresh;
            a.density_thresh = o.density_thresh;
            a.precision = precision;
            a.fp32_cut = o.fp32_cut;
            a.flags = o.jk_flags;
            hfx_jk_stats st{};
            const auto tj = clk::now();
            if ((rc = hfx_jk(e, &a, J.data(), Ka.data(), any_uhf ? Kb.data() : nullptr, &st))) return rc;
            t_jk += secs(tj, clk::now());
            n_builds++;
            n_quart_last = st.n_quartets;
#pragma omp parallel for schedule(dynamic, 1)
            for (int m = 0; m < n_mol; ++m) {
                if (!active[m]) continue;
                Scf& s = mol[m];
                s.iteration = it;
                s.apply_jk(J.data() + off[m], Ka.data() + off[m], s.uhf ? Kb.data() + off[m] : nullptr);
                const int slot = its[m] + it;
                if (s.after_build(o.use_diis ? &diis[m] : nullptr, last[m], res ? &res[m] : nullptr, slot)) {
                    active[m] = 0;
                    conv[m] = 1;
                    done[m] = it + 1;
                } else if (switch_dE > 0.0 && it > 0 && s.dE < switch_dE) {
                    active[m] = 0;
                    done[m] = it + 1;
                }
            }
        }
        for (int m = 0; m < n_mol; ++m) {
            if (active[m]) {
                conv[m] = 0;
                done[m] = o.max_iter;
            }
        }
        return HFX_OK;
    };

    const bool final64 = o.precision == HFX_MIXED && o.final_fp64;
    const bool early = final64 && o.switch_tol > 0.0;
    if ((rc = stage(o.precision, false, early ? o.switch_tol : 0.0, false))) return rc;
    for (int m = 0; m < n_mol; ++m) {
        its[m] = done[m];
        its_mixed[m] = done[m];
        e_mixed[m] = mol[m].e_elec + mol[m].e_nuc;
    }
    if (o.precision == HFX_MIXED && o.final_fp64) {
        if ((rc = stage(HFX_FP64, true, 0.0, early))) return rc;
        for (int m = 0; m < n_mol; ++m) its[m] += done[m];
    }
    const double t_total = secs(t0, clk::now());
    for (int m = 0; m < n_mol; ++m) {
        Scf& s = mol[m];
        s.remove_level_shift();
        s.t_jk = t_jk;
        s.n_builds = n_builds;
        s.n_quart_last = n_quart_last;
        s.write_result(&res[m]);
        res[m].iterations = its[m];
        res[m].converged = conv[m];
        res[m].e_total_mixed = e_mixed[m];
        res[m].iterations_mixed = its_mixed[m];
        res[m].time_init_s = secs(t0, t_init);
        res[m].time_total_s = t_total;
    }
    return HFX_OK;
}

int hfx_scf_run_batched(int32_t n_mol, const hfx_basis* bases, int32_t n_atoms, const double* Z,
                        const double* xyz, const int32_t* atom_mol, const int32_t* charge, const int32_t* mult,
                        const hfx_scf_options* opt, int device, hfx_scf_result* res) {
    if (n_mol <= 0 || !bases || !res) return HFX_ERR_INVALID;
    hfx_engine* e = nullptr;
    const auto t0 = clk::now();
    int rc = hfx_create_batched(n_mol, bases, device, &e);
    if (rc) return rc;
    const double t_create = secs(t0, clk::now());
    std::vector<int32_t> n_aos(n_mol);
    for (int m = 0; m < n_mol; ++m) n_aos[m] = bases[m].n_aos;
    rc = hfx_scf_run_batched_engine(e, n_mol, n_aos.data(), n_atoms, Z, xyz, atom_mol, charge, mult, opt, res);
    hfx_destroy(e);
    if (rc == HFX_OK)
        for (int m = 0; m < n_mol; ++m) {
            res[m].time_init_s += t_create;
            res[m].time_total_s += t_create;
        }
    return rc;
}

}  // extern "C"
