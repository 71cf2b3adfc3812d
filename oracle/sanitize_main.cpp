// sanitize_main.cpp - runs the oracle and the package's host SCF driver (hfx_scf.cpp, linalg.hpp)
// under AddressSanitizer and UndefinedBehaviorSanitizer (TEST INFRASTRUCTURE ONLY; built by
// `make sanitize` in this directory, run by tests/test_sanitize.py).
//
// Input file (native-endian): the flattened basis as in include/hfx.h (int32 n_shells, n_prims,
// n_aos, n_norm; int32 l, n_prim, prim_offset, ao_offset, norm_offset [nS]; double centers [3 nS],
// alpha [nP], coeff [nP], norm [n_norm]; int32 cart [3 N]), then int32 n_atoms, double Z [n_atoms],
// double xyz [3 n_atoms], int32 charge, int32 multiplicity.
// Exercises: hfx_create, S/T/V, Schwarz, ERI blocks, a direct J/K build (RHF and UHF densities),
// the direct SCF (DIIS; damping and level shift), the conventional SCF (stored integrals), the
// gradient, and a batched SCF of two copies of the molecule.  Prints the energies; exit 0 on success.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../include/hfx.h"
#include "../include/hfx_scf.h"

template <class T>
static bool rd(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

static int ck(int rc, const char* what) {
    if (rc != HFX_OK) std::fprintf(stderr, "%s: %d %s\n", what, rc, hfx_last_error());
    return rc;
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t h[4];
    if (fread(h, 4, 4, f) != 4) return 2;
    const int nS = h[0], nP = h[1], N = h[2], nNorm = h[3];
    std::vector<int32_t> l, np, po, ao, no, cart, nat, chg;
    std::vector<double> ctr, al, co, nm, Z, xyz;
    bool ok = rd(f, l, nS) && rd(f, np, nS) && rd(f, po, nS) && rd(f, ao, nS) && rd(f, no, nS) &&
              rd(f, ctr, 3 * (size_t)nS) && rd(f, al, nP) && rd(f, co, nP) && rd(f, nm, nNorm) &&
              rd(f, cart, 3 * (size_t)N) && rd(f, nat, 1);
    ok = ok && rd(f, Z, nat[0]) && rd(f, xyz, 3 * (size_t)nat[0]) && rd(f, chg, 2);
    std::fclose(f);
    if (!ok) return 2;
    const int na = nat[0];
    hfx_basis b{nS, nP, N, nNorm, l.data(), np.data(), po.data(), ao.data(), no.data(),
                ctr.data(), al.data(), co.data(), nm.data(), cart.data()};
    hfx_engine* e = nullptr;
    if (ck(hfx_create(&b, 0, &e), "create")) return 1;
    std::vector<double> S((size_t)N * N), T(S), V(S), Q((size_t)nS * nS), J(S), Ka(S), Kb(S), D(S);
    int rc = ck(hfx_overlap(e, S.data()), "S") || ck(hfx_kinetic(e, T.data()), "T") ||
             ck(hfx_nuclear(e, na, Z.data(), xyz.data(), nullptr, V.data()), "V") || ck(hfx_schwarz(e, Q.data()), "Q");
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) D[i + (size_t)j * N] = 0.01 * ((i * 7 + j * 3) % 11) * ((j * 7 + i * 3) % 11);
    std::vector<double> Dt(D);
    for (double& x : Dt) x *= 2.0;
    hfx_jk_args a{Dt.data(), D.data(), nullptr, nullptr, 1e-10, 1e-10, HFX_FP64, 0.0, 0};
    hfx_jk_stats st{};
    if (!rc) rc = ck(hfx_jk(e, &a, J.data(), Ka.data(), nullptr, &st), "jk");
    a.dD_beta = D.data();
    if (!rc) rc = ck(hfx_jk(e, &a, J.data(), Ka.data(), Kb.data(), &st), "jk uhf");
    std::vector<int32_t> q = {nS - 1, 0, nS - 1, 0, 0, 0, 0, 0};
    std::vector<int64_t> off = {0, 0};
    std::vector<double> blk(6 * 6 * 6 * 6 * 2);
    off[1] = 1296;
    if (!rc) rc = ck(hfx_eri_quartets(e, 2, q.data(), off.data(), blk.data()), "eri");
    hfx_scf_options o;
    hfx_scf_default_options(&o);
    o.energy_tol = 1e-8;
    o.max_iter = 60;
    std::vector<double> F((size_t)N * N), Da(F), C(F), eps(N), trace(4 * 100);
    hfx_scf_result r{};
    r.F_alpha = F.data();
    r.D_alpha = Da.data();
    r.C_alpha = C.data();
    r.eps_alpha = eps.data();
    r.trace = trace.data();
    r.trace_cap = 100;
    if (!rc) rc = ck(hfx_scf_run_engine(e, na, Z.data(), xyz.data(), chg[0], chg[1], &o, &r), "scf");
    const double e_direct = r.e_total;
    if (!rc && r.converged) {  // gradient at the converged density
        std::vector<double> P(F), W(F), g(3 * (size_t)na), parts(12 * (size_t)na);
        int ne = -chg[0];
        for (double z : Z) ne += (int)z;
        const int occ = ne / 2;
        for (size_t x = 0; x < P.size(); ++x) P[x] = 2.0 * Da[x];
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                double w = 0.0;
                for (int k = 0; k < occ; ++k) w += 2.0 * eps[k] * C[i + (size_t)k * N] * C[j + (size_t)k * N];
                W[i + (size_t)j * N] = w;
            }
        hfx_grad_args ga{P.data(), W.data(), na, Z.data(), xyz.data(), nullptr, 1e-10, 1e-10, HFX_FP64, 0.0, 0};
        rc = ck(hfx_rhf_gradient(e, &ga, g.data(), parts.data(), nullptr), "gradient");
    }
    o.damp = 30;
    o.max_damp_iter = 5;
    o.level_shift = 0.2;
    o.max_lshift_iter = 5;
    hfx_scf_result r2{};
    if (!rc) rc = ck(hfx_scf_run_engine(e, na, Z.data(), xyz.data(), chg[0], chg[1], &o, &r2), "scf damped");
    hfx_scf_default_options(&o);
    o.energy_tol = 1e-8;
    o.max_iter = 60;
    o.direct = 0;
    hfx_scf_result r3{};
    if (!rc) rc = ck(hfx_scf_run_engine(e, na, Z.data(), xyz.data(), chg[0], chg[1], &o, &r3), "scf conventional");
    o.direct = 1;
    o.unrestricted = 1;
    hfx_scf_result r4{};
    if (!rc) rc = ck(hfx_scf_run_engine(e, na, Z.data(), xyz.data(), chg[0], chg[1], &o, &r4), "scf uhf");
    ck(hfx_destroy(e), "destroy");
    if (!rc) {  // two copies of the molecule through the batched driver
        hfx_basis bb[2] = {b, b};
        std::vector<double> Z2(Z), xyz2(xyz);
        Z2.insert(Z2.end(), Z.begin(), Z.end());
        xyz2.insert(xyz2.end(), xyz.begin(), xyz.end());
        std::vector<int32_t> am(2 * (size_t)na, 0);
        for (int i = na; i < 2 * na; ++i) am[i] = 1;
        std::vector<int32_t> c2 = {chg[0], chg[0]}, m2 = {chg[1], chg[1]};
        hfx_scf_result rb[2] = {};
        rc = ck(hfx_scf_run_batched(2, bb, 2 * na, Z2.data(), xyz2.data(), am.data(), c2.data(), m2.data(), &o, 0, rb),
                "scf batched");
        if (!rc) std::printf("batched %.10f %.10f\n", rb[0].e_total, rb[1].e_total);
    }
    if (rc) return 1;
    std::printf("direct %.10f (%d it) damped %.10f conventional %.10f uhf %.10f quartets %lld\n", e_direct,
                r.iterations, r2.e_total, r3.e_total, r4.e_total, (long long)st.n_quartets);
    return r.converged && r3.converged ? 0 : 1;
}
