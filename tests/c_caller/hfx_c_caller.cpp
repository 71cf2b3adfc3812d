// hfx_c_caller.cpp - a C++ program that drives the engine through include/hfx.h only, the way the
// reference's SCF would after the integration (no Python, no ctypes): it reads a flattened basis
// and a density from a binary file, calls hfx_create / hfx_overlap / hfx_schwarz / hfx_jk (blocking,
// then graph-replayed) / hfx_destroy, and writes S, Q, J and K to another binary file.
//
// Input file (native-endian):  int32 n_shells, n_prims, n_aos, n_norm;
//   int32 l[nS], n_prim[nS], prim_offset[nS], ao_offset[nS], norm_offset[nS];
//   double centers[3 nS], alpha[nP], coeff[nP], norm[n_norm]; int32 cart[3 N];
//   double D[N N] (column-major alpha density; the build uses dD_tot = 2D, dD_alpha = D)
// Output file: double S[N N], Q[nS nS], J[N N], K[N N], J_graph[N N], K_graph[N N]; int64 n_quartets.
// Exit status 0 on success; on failure the hfx_last_error() text goes to stderr.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hfx.h"

namespace {

template <class T>
bool read_n(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

template <class T>
bool write_n(FILE* f, const std::vector<T>& v) {
    return v.empty() || fwrite(v.data(), sizeof(T), v.size(), f) == v.size();
}

int check(int rc, const char* what) {
    if (rc != HFX_OK) std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, hfx_last_error());
    return rc;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    int32_t hdr[4];
    if (fread(hdr, sizeof(int32_t), 4, in) != 4) return 2;
    const int nS = hdr[0], nP = hdr[1], N = hdr[2], nNorm = hdr[3];
    std::vector<int32_t> l, nprim, poff, aoff, noff, cart;
    std::vector<double> ctr, alpha, coeff, norm, D;
    bool ok = read_n(in, l, nS) && read_n(in, nprim, nS) && read_n(in, poff, nS) && read_n(in, aoff, nS) &&
              read_n(in, noff, nS) && read_n(in, ctr, 3 * (size_t)nS) && read_n(in, alpha, nP) &&
              read_n(in, coeff, nP) && read_n(in, norm, nNorm) && read_n(in, cart, 3 * (size_t)N) &&
              read_n(in, D, (size_t)N * N);
    std::fclose(in);
    if (!ok) {
        std::fprintf(stderr, "short input file\n");
        return 2;
    }
    hfx_basis b{nS, nP, N, nNorm, l.data(), nprim.data(), poff.data(), aoff.data(), noff.data(),
                ctr.data(), alpha.data(), coeff.data(), norm.data(), cart.data()};
    hfx_engine* e = nullptr;
    if (check(hfx_create(&b, 0, &e), "hfx_create")) return 1;
    std::vector<double> S((size_t)N * N), Q((size_t)nS * nS), J((size_t)N * N), K((size_t)N * N);
    std::vector<double> Jg((size_t)N * N), Kg((size_t)N * N), Dt((size_t)N * N);
    for (size_t x = 0; x < Dt.size(); ++x) Dt[x] = 2.0 * D[x];
    int rc = check(hfx_overlap(e, S.data()), "hfx_overlap");
    if (!rc) rc = check(hfx_schwarz(e, Q.data()), "hfx_schwarz");
    hfx_jk_args a{Dt.data(), D.data(), nullptr, Q.data(), 1e-10, 1e-10, HFX_FP64, 0.0, 0};
    hfx_jk_stats st{};
    if (!rc) rc = check(hfx_jk(e, &a, J.data(), K.data(), nullptr, &st), "hfx_jk");
    a.flags = HFX_JK_USE_GRAPH;
    for (int rep = 0; rep < 2 && !rc; ++rep) rc = check(hfx_jk(e, &a, Jg.data(), Kg.data(), nullptr, nullptr), "hfx_jk graph");
    if (check(hfx_destroy(e), "hfx_destroy") && !rc) rc = 1;
    if (rc) return 1;
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    const int64_t nq = st.n_quartets;
    ok = write_n(out, S) && write_n(out, Q) && write_n(out, J) && write_n(out, K) && write_n(out, Jg) &&
         write_n(out, Kg) && fwrite(&nq, sizeof(nq), 1, out) == 1;
    std::fclose(out);
    return ok ? 0 : 2;
}
