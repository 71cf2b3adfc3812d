// hfx_oracle.cpp - CPU oracle for the direct-SCF Fock build.  TEST INFRASTRUCTURE ONLY.
//
// A from-scratch C++ restatement of the reference's integral engine and direct J/K loop
// (MalachiF1/Hartree-Fock), exporting the same C ABI as the HIP engine (include/hfx.h) so the
// tests can run both on the same inputs.  Only tests/, the smoke check and the benchmark's
// CPU-baseline leg load this library; the product never links it.
//
// What each part restates (reference file:line):
//   Boys function and its 5e-5 grid ........ src/Boys.cpp:14-105, include/Boys.hpp:5-55
//   primitive-pair data ..................... src/IntegralEngine.cpp:995-1022
//   Hermite E recursion ..................... src/IntegralEngine.cpp:1025-1079 (EBuffer .hpp:68-99)
//   Hermite R recursion ..................... src/IntegralEngine.cpp:1082-1162 (RBuffer .hpp:109-141)
//   per-quartet ERI ......................... src/IntegralEngine.cpp:731-992
//   Schwarz table ........................... src/IntegralEngine.cpp:695-728
//   overlap / kinetic / nuclear ............. src/IntegralEngine.cpp:78-368
//   direct J/K build ........................ src/SCF.cpp:523-744
//
// Pins (tests/): Boys against mpmath; E and R against closed forms; S, T, V and (ss|ss)
// against closed forms; ERIs against an independent Obara-Saika evaluation; J/K against a
// brute-force contraction of the dense ERI tensor; the Szabo-Ostlund H2 values.
// Parity is tolerance parity: the reference's bits depend on Eigen's reduction order and on
// -ffp-contract=fast; this file is built with -ffp-contract=off.

#include "../include/hfx.h"
#include "../hartree-fock_amd/csrc/scf_device.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// ---------------------------------------------------------------------------------------
// Boys function (src/Boys.cpp, include/Boys.hpp).  The reference's grid holds orders 0..33;
// only orders up to kGridOrders-1 are kept here (the same values at the same grid points).
// ---------------------------------------------------------------------------------------
constexpr double kSqrtPiBy2 = 0.8862269254527579;
constexpr double kGridStep = 5.0e-5;
constexpr double kGridMax = 30.0;
constexpr int kGridOrders = 14;
const size_t kGridRows = static_cast<size_t>(kGridMax / kGridStep) + 2;  // rows 0..600001

// T above which the downward recursion from m_max is used, indexed by m_max
// (include/Boys.hpp:19-26; the first entries only, m_max <= 13 here).
constexpr double kTransition[14] = {0.00, 2.35, 2.35, 2.35, 2.35, 2.38, 2.82,
                                    3.05, 3.72, 3.76, 4.01, 4.61, 4.92, 5.15};

double boys_series(unsigned m, double T) {  // Boys::analyticalBoys (Boys.cpp:74-89)
    double sum = 0.0;
    double term = 1.0 / (m + 0.5);
    sum += term;
    for (int i = 1; i < 200; ++i) {
        term *= T / (m + i + 0.5);
        sum += term;
        if (std::abs(term) < 1e-15) break;
    }
    return sum * std::exp(-T) / 2.0;
}

const std::vector<double>& boys_grid() {  // GridManager (Boys.cpp:92-105), built in memory
    static std::vector<double> grid;
    static std::once_flag once;
    std::call_once(once, [] {
        grid.resize(kGridRows * kGridOrders);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)kGridRows; ++i) {
            const double Ti = i * kGridStep;
            for (int j = 0; j < kGridOrders; ++j) grid[i * kGridOrders + j] = boys_series(j, Ti);
        }
    });
    return grid;
}

void boys(unsigned m_max, double T, double* F) {  // Boys::calculateBoys (Boys.cpp:14-71)
    if (T > 30.0) {
        F[0] = kSqrtPiBy2 / std::sqrt(T);
        const double invT = 1.0 / T;
        for (unsigned m = 0; m < m_max; ++m) F[m + 1] = (m + 0.5) * F[m] * invT;
        return;
    }
    const std::vector<double>& g = boys_grid();
    const size_t idx = static_cast<size_t>(T / kGridStep);
    const double Tg = idx * kGridStep;
    const double dT = T - Tg;
    const double half_dT2 = 0.5 * dT * dT;
    if (m_max < 2) {
        const double* row = &g[idx * kGridOrders];
        F[0] = row[0] - dT * row[1] + half_dT2 * row[2];
        if (m_max == 0) return;
        F[1] = row[1] - dT * row[2] + half_dT2 * row[3];
        return;
    }
    const double emT = std::exp(-T);
    if (T <= kTransition[m_max]) {
        const double* row = &g[idx * kGridOrders + m_max];
        F[m_max] = row[0] - dT * row[1] + half_dT2 * row[2];
        const double twoT = 2.0 * T;
        for (unsigned m = m_max; m >= 1; --m) F[m - 1] = (twoT * F[m] + emT) / (2 * (m - 1) + 1);
    } else {
        const double* row = &g[idx * kGridOrders];
        F[0] = row[0] - dT * row[1] + half_dT2 * row[2];
        const double inv2T = 1 / (2 * T);
        for (unsigned m = 0; m < m_max; ++m) F[m + 1] = ((2.0 * m + 1.0) * F[m] - emT) * inv2T;
    }
}

// ---------------------------------------------------------------------------------------
// Primitive-pair data (IntegralEngine.cpp:995-1022).
// ---------------------------------------------------------------------------------------
struct PairData {
    double p, K, P[3], PA[3], PB[3];
};

PairData pair_data(double a, const double* A, double b, const double* B) {
    PairData d;
    d.p = a + b;
    const double oo_p = 1.0 / d.p;
    for (int x = 0; x < 3; ++x) {
        d.P[x] = (a * A[x] + b * B[x]) * oo_p;
        d.PA[x] = d.P[x] - A[x];
        d.PB[x] = d.P[x] - B[x];
    }
    const double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
    const double AmB2 = dx * dx + dy * dy + dz * dz;
    d.K = std::exp(-a * b * oo_p * AmB2);
    return d;
}

// ---------------------------------------------------------------------------------------
// Packed 1-D Hermite coefficients E^{ij}_t (EBuffer, IntegralEngine.hpp:68-99).
// ---------------------------------------------------------------------------------------
struct EBuf {
    unsigned l1 = 0, l2 = 0;
    std::vector<double> v;
    EBuf() = default;
    EBuf(unsigned a, unsigned b) : l1(a), l2(b), v((a + 1) * (b + 1) * (a + b + 2) / 2) {}
    size_t ioff(unsigned i) const { return i * (l2 + 1) * (i + l2 + 1) / 2; }
    size_t joff(unsigned i, unsigned j) const { return j * (i + 1) + j * (j - 1) / 2; }
    double& at(unsigned i, unsigned j, unsigned t) { return v[ioff(i) + joff(i, j) + t]; }
    double at(unsigned i, unsigned j, unsigned t) const { return v[ioff(i) + joff(i, j) + t]; }
};

// IntegralEngine::computeHermiteCoeffs (IntegralEngine.cpp:1025-1079): i-steps with PA at
// j = 0, then j-steps with PB for every i; t is visited downwards inside each step.
void hermite(unsigned l1, unsigned l2, double p, double PA, double PB, EBuf& E) {
    E.at(0, 0, 0) = 1.0;
    const double oo_2p = 0.5 / p;
    for (unsigned i = 1; i <= l1; ++i) {
        const double* prev = &E.v[E.ioff(i - 1)];
        double* cur = &E.v[E.ioff(i)];
        for (int t = (int)i; t >= 0; --t) {
            double val = 0.0;
            if (t <= (int)i - 1) val = PA * prev[t];
            if (t >= 1) val += oo_2p * prev[t - 1];
            if (t + 1 <= (int)i - 1) val += (t + 1) * prev[t + 1];
            cur[t] = val;
        }
    }
    for (unsigned j = 1; j <= l2; ++j) {
        for (unsigned i = 0; i <= l1; ++i) {
            const double* prev = &E.v[E.ioff(i) + E.joff(i, j - 1)];
            double* cur = &E.v[E.ioff(i) + E.joff(i, j)];
            const int top = (int)(i + j);
            for (int t = top; t >= 0; --t) {
                double val = 0.0;
                if (t <= top - 1) val = PB * prev[t];
                if (t >= 1) val += oo_2p * prev[t - 1];
                if (t + 1 <= top - 1) val += (t + 1) * prev[t + 1];
                cur[t] = val;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// Hermite Coulomb integrals R^n_{tuv} in a dense (L+1)^4 buffer (RBuffer,
// IntegralEngine.hpp:109-141; IntegralEngine::computeAuxiliaryIntegrals :1082-1162).
// ---------------------------------------------------------------------------------------
struct RBuf {
    size_t L = 0, su = 0, sv = 0, sn = 0;
    std::vector<double> v;
    RBuf() = default;
    explicit RBuf(size_t l) : L(l), su(l + 1), sv((l + 1) * (l + 1)), sn((l + 1) * (l + 1) * (l + 1)),
                              v(sn * (l + 1)) {}
    double& at(size_t t, size_t u, size_t w, size_t n) { return v[t + u * su + w * sv + n * sn]; }
};

void hermite_r(double p, const double PC[3], const double* F, RBuf& R) {
    const unsigned L = (unsigned)R.L;
    const double m2p = -2.0 * p;
    double pw = 1.0;
    for (unsigned n = 0; n <= L; ++n) {
        R.at(0, 0, 0, n) = pw * F[n];
        pw *= m2p;
    }
    for (int n = (int)L - 1; n >= 0; --n) {
        const unsigned top = L - n;
        for (unsigned t = 1; t <= top; ++t) {
            const double val = PC[0] * R.at(t - 1, 0, 0, n + 1);
            R.at(t, 0, 0, n) = (t >= 2) ? val + (t - 1) * R.at(t - 2, 0, 0, n + 1) : val;
        }
        for (unsigned u = 1; u <= top; ++u) {
            for (unsigned t = 0; t <= top - u; ++t) {
                double val = PC[1] * R.at(t, u - 1, 0, n + 1);
                if (u >= 2) val += (u - 1) * R.at(t, u - 2, 0, n + 1);
                R.at(t, u, 0, n) = val;
            }
        }
        for (unsigned w = 1; w <= top; ++w) {
            for (unsigned u = 0; u <= top - w; ++u) {
                for (unsigned t = 0; t <= top - u - w; ++t) {
                    double val = PC[2] * R.at(t, u, w - 1, n + 1);
                    if (w >= 2) val += (w - 1) * R.at(t, u, w - 2, n + 1);
                    R.at(t, u, w, n) = val;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// Engine state: the flat basis of IntegralEngine::IntegralEngine (IntegralEngine.cpp:13-75).
// ---------------------------------------------------------------------------------------
struct Mol {
    int nS = 0, nP = 0, N = 0;
    std::vector<int> l, nprim, poff, aoff, noff, nao;
    std::vector<double> ctr, alpha, coeff, norm;
    std::vector<int> cart;  // 3 x N
    std::vector<double> Q;  // nS x nS, filled on demand
    bool haveQ = false;
    // hfx_store_eris: the unique integrals in the reference's packed canonical order
    // (ElectronRepulsionTensor, include/Molecule.hpp:10-80) and the Schwarz survivors counted
    std::vector<double> vee;
    int64_t stored_quartets = 0;
};

}  // namespace

struct hfx_engine {
    std::vector<Mol> mols;
    std::vector<int64_t> mat_off;  // packed N_m x N_m block offsets
    std::vector<int64_t> shell_off;
    int64_t nS_tot = 0, N_tot = 0, elems = 0;
    bool stored = false;
};

namespace {

int load_mol(const hfx_basis* b, Mol& m) {
    if (!b || b->n_shells <= 0 || !b->l || !b->n_prim || !b->prim_offset || !b->ao_offset ||
        !b->norm_offset || !b->centers || !b->alpha || !b->coeff || !b->norm || !b->cart)
        return fail(HFX_ERR_INVALID, "hfx_basis: missing field");
    m.nS = b->n_shells;
    m.nP = b->n_prims;
    m.N = b->n_aos;
    m.l.assign(b->l, b->l + m.nS);
    m.nprim.assign(b->n_prim, b->n_prim + m.nS);
    m.poff.assign(b->prim_offset, b->prim_offset + m.nS);
    m.aoff.assign(b->ao_offset, b->ao_offset + m.nS);
    m.noff.assign(b->norm_offset, b->norm_offset + m.nS);
    m.nao.resize(m.nS);
    for (int s = 0; s < m.nS; ++s) {
        if (m.l[s] < 0 || m.l[s] > 4) return fail(HFX_ERR_UNSUPPORTED, "oracle: l out of range");
        m.nao[s] = (m.l[s] + 1) * (m.l[s] + 2) / 2;
    }
    m.ctr.assign(b->centers, b->centers + 3 * m.nS);
    m.alpha.assign(b->alpha, b->alpha + m.nP);
    m.coeff.assign(b->coeff, b->coeff + m.nP);
    m.norm.assign(b->norm, b->norm + b->n_norm);
    m.cart.assign(b->cart, b->cart + 3 * m.N);
    return HFX_OK;
}

// IntegralEngine::electronRepulsion(a,b,c,d, std::vector<double>&) (IntegralEngine.cpp:731-992).
// G is zeroed by the caller and has naoA*naoB*naoC*naoD entries, row-major (a,b,c,d).
void eri_quartet(const Mol& m, int sa, int sb, int sc, int sd, double* G) {
    const int npA = m.nprim[sa], npB = m.nprim[sb], npC = m.nprim[sc], npD = m.nprim[sd];
    const int nA = m.nao[sa], nB = m.nao[sb], nC = m.nao[sc], nD = m.nao[sd];
    const unsigned lA = m.l[sa], lB = m.l[sb], lC = m.l[sc], lD = m.l[sd];
    const int aoA = m.aoff[sa], aoB = m.aoff[sb], aoC = m.aoff[sc], aoD = m.aoff[sd];
    const double *cA = &m.ctr[3 * sa], *cB = &m.ctr[3 * sb], *cC = &m.ctr[3 * sc], *cD = &m.ctr[3 * sd];
    const bool same_pair = (sa == sc && sb == sd);

    // cd primitive pairs with their Hermite coefficients, cached once (:755-786).
    const int ncd = npC * npD;
    std::vector<PairData> pcd(ncd);
    std::vector<EBuf> Ecx(ncd, EBuf(lC, lD)), Ecy(ncd, EBuf(lC, lD)), Ecz(ncd, EBuf(lC, lD));
    for (int pc = 0; pc < npC; ++pc)
        for (int pd = 0; pd < npD; ++pd) {
            const int k = pc * npD + pd;
            pcd[k] = pair_data(m.alpha[m.poff[sc] + pc], cC, m.alpha[m.poff[sd] + pd], cD);
            hermite(lC, lD, pcd[k].p, pcd[k].PA[0], pcd[k].PB[0], Ecx[k]);
            hermite(lC, lD, pcd[k].p, pcd[k].PA[1], pcd[k].PB[1], Ecy[k]);
            hermite(lC, lD, pcd[k].p, pcd[k].PA[2], pcd[k].PB[2], Ecz[k]);
        }
    EBuf Eax(lA, lB), Eay(lA, lB), Eaz(lA, lB);
    const unsigned L = lA + lB + lC + lD;
    RBuf R(L);
    std::vector<double> F(L + 1);
    std::vector<double> prim((size_t)nA * nB * nC * nD);
    const int nCD = nC * nD, nBCD = nB * nCD;

    for (int pa = 0; pa < npA; ++pa) {
        const double ca = m.coeff[m.poff[sa] + pa];
        const double* na = &m.norm[m.noff[sa] + pa * nA];
        for (int pb = 0; pb < npB; ++pb) {
            const double cb = m.coeff[m.poff[sb] + pb];
            const double* nb = &m.norm[m.noff[sb] + pb * nB];
            const PairData ab = pair_data(m.alpha[m.poff[sa] + pa], cA, m.alpha[m.poff[sb] + pb], cB);
            hermite(lA, lB, ab.p, ab.PA[0], ab.PB[0], Eax);
            hermite(lA, lB, ab.p, ab.PA[1], ab.PB[1], Eay);
            hermite(lA, lB, ab.p, ab.PA[2], ab.PB[2], Eaz);
            for (int k = 0; k < ncd; ++k) {
                const PairData& cd = pcd[k];
                const int pc = k / npD, pd = k % npD;
                const double pref = (2.0 * std::pow(M_PI, 2.5) / (ab.p * cd.p * std::sqrt(ab.p + cd.p))) *
                                    ab.K * cd.K;
                const double delta = (ab.p * cd.p) / (ab.p + cd.p);
                const double PQ[3] = {ab.P[0] - cd.P[0], ab.P[1] - cd.P[1], ab.P[2] - cd.P[2]};
                const double T = delta * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                boys(L, T, F.data());
                hermite_r(delta, PQ, F.data(), R);

                for (int a = 0; a < nA; ++a) {
                    const int i = aoA + a;
                    const int bigIb = i * (i + 1) / 2;
                    const int l1 = m.cart[3 * i], m1 = m.cart[3 * i + 1], n1 = m.cart[3 * i + 2];
                    for (int b = 0; b < nB; ++b) {
                        const int j = aoB + b;
                        if (j > i) continue;
                        const int bigI = bigIb + j;
                        const int l2 = m.cart[3 * j], m2 = m.cart[3 * j + 1], n2 = m.cart[3 * j + 2];
                        for (int c = 0; c < nC; ++c) {
                            const int kk = aoC + c;
                            const int bigKb = kk * (kk + 1) / 2;
                            const int l3 = m.cart[3 * kk], m3 = m.cart[3 * kk + 1], n3 = m.cart[3 * kk + 2];
                            for (int d = 0; d < nD; ++d) {
                                const int ll = aoD + d;
                                if (ll > kk) continue;
                                if (same_pair && bigI < bigKb + ll) continue;
                                const int l4 = m.cart[3 * ll], m4 = m.cart[3 * ll + 1], n4 = m.cart[3 * ll + 2];
                                const int l12 = l1 + l2, l34 = l3 + l4, m12 = m1 + m2, m34 = m3 + m4;
                                const int n12 = n1 + n2, n34 = n3 + n4;
                                double acc = 0.0;
                                for (int v1 = 0; v1 <= n12; ++v1) {
                                    double sv2 = 0.0;
                                    for (int v2 = 0; v2 <= n34; ++v2) {
                                        double su1 = 0.0;
                                        for (int u1 = 0; u1 <= m12; ++u1) {
                                            double su2 = 0.0;
                                            for (int u2 = 0; u2 <= m34; ++u2) {
                                                double st1 = 0.0;
                                                for (int t1 = 0; t1 <= l12; ++t1) {
                                                    double st2 = 0.0;
                                                    for (int t2 = 0; t2 <= l34; ++t2) {
                                                        const double sgn = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
                                                        st2 += sgn * Ecx[k].at(l3, l4, t2) *
                                                               R.at(t1 + t2, u1 + u2, v1 + v2, 0);
                                                    }
                                                    st1 += Eax.at(l1, l2, t1) * st2;
                                                }
                                                su2 += Ecy[k].at(m3, m4, u2) * st1;
                                            }
                                            su1 += Eay.at(m1, m2, u1) * su2;
                                        }
                                        sv2 += Ecz[k].at(n3, n4, v2) * su1;
                                    }
                                    acc += Eaz.at(n1, n2, v1) * sv2;
                                }
                                prim[(size_t)a * nBCD + b * nCD + c * nD + d] = acc;
                            }
                        }
                    }
                }

                const double cc = m.coeff[m.poff[sc] + pc];
                const double cdd = m.coeff[m.poff[sd] + pd];
                const double* nc = &m.norm[m.noff[sc] + pc * nC];
                const double* nd = &m.norm[m.noff[sd] + pd * nD];
                const double cprod = ca * cb * cc * cdd;
                for (int a = 0; a < nA; ++a) {
                    const int i = aoA + a;
                    const int bigIb = i * (i + 1) / 2;
                    for (int b = 0; b < nB; ++b) {
                        const int j = aoB + b;
                        if (j > i) continue;
                        const int bigI = bigIb + j;
                        for (int c = 0; c < nC; ++c) {
                            const int kk = aoC + c;
                            const int bigKb = kk * (kk + 1) / 2;
                            for (int d = 0; d < nD; ++d) {
                                const int ll = aoD + d;
                                if (ll > kk) continue;
                                if (same_pair && bigI < bigKb + ll) continue;
                                const size_t idx = (size_t)a * nBCD + b * nCD + c * nD + d;
                                G[idx] += na[a] * nb[b] * nc[c] * nd[d] * cprod * pref * prim[idx];
                            }
                        }
                    }
                }
            }
        }
    }
}

// IntegralEngine::schwartzScreeningMatrix (IntegralEngine.cpp:695-728).
void schwarz(Mol& m) {
    if (m.haveQ) return;
    const int nS = m.nS;
    m.Q.assign((size_t)nS * nS, 0.0);
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int i = 0; i < nS; ++i)
        for (int j = 0; j < nS; ++j) {
            if (j > i) continue;
            const int ni = m.nao[i], nj = m.nao[j];
            std::vector<double> G((size_t)ni * nj * ni * nj, 0.0);
            eri_quartet(m, i, j, i, j, G.data());
            const size_t aoffs = (size_t)nj * ni * nj + nj, boffs = (size_t)ni * nj + 1;
            double mx = 0;
            for (int a = 0; a < ni; ++a)
                for (int b = 0; b < nj; ++b) mx = std::max(mx, G[a * aoffs + b * boffs]);
            m.Q[j + (size_t)i * nS] = mx;
        }
    for (int i = 0; i < nS; ++i)
        for (int j = 0; j < i; ++j) m.Q[i + (size_t)j * nS] = m.Q[j + (size_t)i * nS];
    for (auto& q : m.Q) q = std::sqrt(q);
    m.haveQ = true;
}

// ---------------------------------------------------------------------------------------
// One-electron integrals (IntegralEngine.cpp:78-368).  Each fills the upper shell-pair
// triangle and mirrors it (selfadjointView<Upper>).
// ---------------------------------------------------------------------------------------
void overlap_pair(const Mol& m, int sa, int sb, double* S) {
    const int nA = m.nao[sa], nB = m.nao[sb], N = m.N;
    const unsigned lA = m.l[sa], lB = m.l[sb];
    EBuf Ex(lA, lB), Ey(lA, lB), Ez(lA, lB);
    for (int pa = 0; pa < m.nprim[sa]; ++pa)
        for (int pb = 0; pb < m.nprim[sb]; ++pb) {
            const double a = m.alpha[m.poff[sa] + pa], b = m.alpha[m.poff[sb] + pb];
            const PairData d = pair_data(a, &m.ctr[3 * sa], b, &m.ctr[3 * sb]);
            const double pref = std::pow(M_PI / d.p, 1.5) * d.K;
            hermite(lA, lB, d.p, d.PA[0], d.PB[0], Ex);
            hermite(lA, lB, d.p, d.PA[1], d.PB[1], Ey);
            hermite(lA, lB, d.p, d.PA[2], d.PB[2], Ez);
            const double cp = m.coeff[m.poff[sa] + pa] * m.coeff[m.poff[sb] + pb];
            for (int x = 0; x < nA; ++x) {
                const double na = m.norm[m.noff[sa] + pa * nA + x];
                const int i = m.aoff[sa] + x;
                const int* c1 = &m.cart[3 * i];
                for (int y = 0; y < nB; ++y) {
                    const double nb = m.norm[m.noff[sb] + pb * nB + y];
                    const int j = m.aoff[sb] + y;
                    const int* c2 = &m.cart[3 * j];
                    S[i + (size_t)j * N] += na * nb * cp * pref * Ex.at(c1[0], c2[0], 0) * Ey.at(c1[1], c2[1], 0) *
                                            Ez.at(c1[2], c2[2], 0);
                }
            }
        }
}

void kinetic_pair(const Mol& m, int sa, int sb, double* T) {
    const int nA = m.nao[sa], nB = m.nao[sb], N = m.N;
    const unsigned lA = m.l[sa], lB = m.l[sb];
    EBuf Ex(lA, lB + 2), Ey(lA, lB + 2), Ez(lA, lB + 2);
    for (int pa = 0; pa < m.nprim[sa]; ++pa)
        for (int pb = 0; pb < m.nprim[sb]; ++pb) {
            const double a = m.alpha[m.poff[sa] + pa], b = m.alpha[m.poff[sb] + pb];
            const PairData d = pair_data(a, &m.ctr[3 * sa], b, &m.ctr[3 * sb]);
            const double pref = std::pow(M_PI / d.p, 1.5) * d.K;
            hermite(lA, lB + 2, d.p, d.PA[0], d.PB[0], Ex);
            hermite(lA, lB + 2, d.p, d.PA[1], d.PB[1], Ey);
            hermite(lA, lB + 2, d.p, d.PA[2], d.PB[2], Ez);
            const double cp = m.coeff[m.poff[sa] + pa] * m.coeff[m.poff[sb] + pb];
            for (int x = 0; x < nA; ++x) {
                const int i = m.aoff[sa] + x;
                const int l1 = m.cart[3 * i], m1 = m.cart[3 * i + 1], n1 = m.cart[3 * i + 2];
                const double na = m.norm[m.noff[sa] + pa * nA + x];
                for (int y = 0; y < nB; ++y) {
                    const int j = m.aoff[sb] + y;
                    const int l2 = m.cart[3 * j], m2 = m.cart[3 * j + 1], n2 = m.cart[3 * j + 2];
                    const double nb = m.norm[m.noff[sb] + pb * nB + y];
                    const double sx = Ex.at(l1, l2, 0), sy = Ey.at(m1, m2, 0), sz = Ez.at(n1, n2, 0);
                    const double t1 = b * (2 * (l2 + m2 + n2) + 3) * sx * sy * sz;
                    const double t2 = -2.0 * b * b *
                                      (Ex.at(l1, l2 + 2, 0) * sy * sz + sx * Ey.at(m1, m2 + 2, 0) * sz +
                                       sx * sy * Ez.at(n1, n2 + 2, 0));
                    double t3 = 0.0;
                    if (l2 >= 2) t3 -= l2 * (l2 - 1) * Ex.at(l1, l2 - 2, 0) * sy * sz;
                    if (m2 >= 2) t3 -= m2 * (m2 - 1) * sx * Ey.at(m1, m2 - 2, 0) * sz;
                    if (n2 >= 2) t3 -= n2 * (n2 - 1) * sx * sy * Ez.at(n1, n2 - 2, 0);
#ifdef ORACLE_PLANT_KINETIC_FACTOR
                    // planted mistake for the mutation test: l(l-1) instead of l(l-1)/2
#else
                    t3 *= 0.5;
#endif
                    T[i + (size_t)j * N] += na * nb * cp * pref * (t1 + t2 + t3);
                }
            }
        }
}

void nuclear_pair(const Mol& m, int nat, const double* Z, const double* xyz, int sa, int sb, double* V) {
    const int nA = m.nao[sa], nB = m.nao[sb], N = m.N;
    const unsigned lA = m.l[sa], lB = m.l[sb];
    const unsigned L = lA + lB;
    RBuf R(L);
    std::vector<double> F(L + 1);
    EBuf Ex(lA, lB), Ey(lA, lB), Ez(lA, lB);
    for (int pa = 0; pa < m.nprim[sa]; ++pa)
        for (int pb = 0; pb < m.nprim[sb]; ++pb) {
            const double a = m.alpha[m.poff[sa] + pa], b = m.alpha[m.poff[sb] + pb];
            const PairData d = pair_data(a, &m.ctr[3 * sa], b, &m.ctr[3 * sb]);
            const double pref = (2.0 * M_PI / d.p) * d.K;
            hermite(lA, lB, d.p, d.PA[0], d.PB[0], Ex);
            hermite(lA, lB, d.p, d.PA[1], d.PB[1], Ey);
            hermite(lA, lB, d.p, d.PA[2], d.PB[2], Ez);
            const double cp = m.coeff[m.poff[sa] + pa] * m.coeff[m.poff[sb] + pb];
            for (int at = 0; at < nat; ++at) {
                const double PC[3] = {d.P[0] - xyz[3 * at], d.P[1] - xyz[3 * at + 1], d.P[2] - xyz[3 * at + 2]};
                const double T = d.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                boys(L, T, F.data());
                hermite_r(d.p, PC, F.data(), R);
                for (int x = 0; x < nA; ++x) {
                    const int i = m.aoff[sa] + x;
                    const int l1 = m.cart[3 * i], m1 = m.cart[3 * i + 1], n1 = m.cart[3 * i + 2];
                    const double na = m.norm[m.noff[sa] + pa * nA + x];
                    for (int y = 0; y < nB; ++y) {
                        const int j = m.aoff[sb] + y;
                        const int l2 = m.cart[3 * j], m2 = m.cart[3 * j + 1], n2 = m.cart[3 * j + 2];
                        const double nb = m.norm[m.noff[sb] + pb * nB + y];
                        double sum = 0.0;
                        for (int v = 0; v <= n1 + n2; ++v) {
                            double su = 0.0;
                            for (int u = 0; u <= m1 + m2; ++u) {
                                double st = 0.0;
#ifdef ORACLE_PLANT_NUCLEAR_TRANSPOSE
                                // planted mistake for the mutation test: u and v of R transposed
                                for (int t = 0; t <= l1 + l2; ++t) st += Ex.at(l1, l2, t) * R.at(t, v, u, 0);
#else
                                for (int t = 0; t <= l1 + l2; ++t) st += Ex.at(l1, l2, t) * R.at(t, u, v, 0);
#endif
                                su += Ey.at(m1, m2, u) * st;
                            }
                            sum += Ez.at(n1, n2, v) * su;
                        }
                        V[i + (size_t)j * N] -= na * nb * cp * pref * Z[at] * sum;
                    }
                }
            }
        }
}

void mirror_upper(double* M, int N) {
    for (int c = 0; c < N; ++c)
        for (int r = c + 1; r < N; ++r) M[r + (size_t)c * N] = M[c + (size_t)r * N];
}

// ---------------------------------------------------------------------------------------
// Class indexing and work-list hash shared with the engine's statistics.
// ---------------------------------------------------------------------------------------
int pair_class(int l1, int l2) {
    const int hi = std::max(l1, l2), lo = std::min(l1, l2);
    return hi * (hi + 1) / 2 + lo;
}
int quartet_class(int la, int lb, int lc, int ld) {
    int p = pair_class(la, lb), q = pair_class(lc, ld);
    if (p < q) std::swap(p, q);
    return p * (p + 1) / 2 + q;
}
uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
int ncart(int l) { return (l + 1) * (l + 2) / 2; }
int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
double hsum(int l1, int l2) {  // sum over Cartesian pairs of (x1+x2+1)(y1+y2+1)(z1+z2+1)
    double s = 0;
    for (int i1 = l1; i1 >= 0; --i1)
        for (int j1 = l1 - i1; j1 >= 0; --j1)
            for (int i2 = l2; i2 >= 0; --i2)
                for (int j2 = l2 - i2; j2 >= 0; --j2) {
                    const int k1 = l1 - i1 - j1, k2 = l2 - i2 - j2;
                    s += (i1 + i2 + 1) * (j1 + j2 + 1) * (k1 + k2 + 1);
                }
    return s;
}
// Model FLOPs per primitive quartet, bra = the larger pair class.
double model_flop(int la, int lb, int lc, int ld) {
    if (pair_class(la, lb) < pair_class(lc, ld)) {
        std::swap(la, lc);
        std::swap(lb, ld);
    }
    const int L = la + lb + lc + ld;
    const double c4 = (double)(L + 4) * (L + 3) * (L + 2) * (L + 1) / 24.0;
    return (20 + 3 * (L + 1)) + 3 * c4 + 2.0 * nherm(la + lb) * hsum(lc, ld) +
           2.0 * ncart(lc) * ncart(ld) * hsum(la, lb) + 2.0 * ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
}

bool all_finite(const double* x, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(x[i])) return false;
    return true;
}

// SCF::buildFockMatrixDirect (SCF.cpp:523-744) on one molecule, restricted to bra shells
// a in [a_lo, a_hi) (the whole loop when a_lo = 0, a_hi = nS).  Outputs are the symmetric
// increments J, Ka[, Kb]; F = F_prev + J - K is left to the caller.
void direct_jk(Mol& m, const double* dDt, const double* dDa, const double* dDb, const double* Qext,
               double tauS, double tauD, int a_lo, int a_hi, double* J, double* Ka, double* Kb,
               hfx_jk_stats* st) {
    const int nS = m.nS, N = m.N;
    const bool uhf = dDb != nullptr;
    const double* Q = Qext;
    if (!Q) {
        schwarz(m);
        Q = m.Q.data();
    }
    // Shell-pair density maxima (SCF.cpp:543-575).
    std::vector<double> mt((size_t)nS * nS, 0.0), ma((size_t)nS * nS, 0.0), mb;
    if (uhf) mb.assign((size_t)nS * nS, 0.0);
#pragma omp parallel for collapse(2)
    for (int s1 = 0; s1 < nS; ++s1)
        for (int s2 = 0; s2 < nS; ++s2) {
            if (s2 > s1) continue;
            double vt = 0, va = 0, vb = 0;
            for (int a = m.aoff[s1]; a < m.aoff[s1] + m.nao[s1]; ++a)
                for (int b = m.aoff[s2]; b < m.aoff[s2] + m.nao[s2]; ++b) {
                    vt = std::max(vt, std::abs(dDt[a + (size_t)b * N]));
                    va = std::max(va, std::abs(dDa[a + (size_t)b * N]));
                    if (uhf) vb = std::max(vb, std::abs(dDb[a + (size_t)b * N]));
                }
            mt[s2 + (size_t)s1 * nS] = mt[s1 + (size_t)s2 * nS] = vt;
            ma[s2 + (size_t)s1 * nS] = ma[s1 + (size_t)s2 * nS] = va;
            if (uhf) mb[s2 + (size_t)s1 * nS] = mb[s1 + (size_t)s2 * nS] = vb;
        }

    std::vector<double> Jt((size_t)N * N, 0.0), Kat((size_t)N * N, 0.0), Kbt;
    if (uhf) Kbt.assign((size_t)N * N, 0.0);
    int64_t cls[HFX_N_CLASSES] = {0};
    double clsfl[HFX_N_CLASSES] = {0};
    uint64_t hsum_ = 0, hxor = 0;
    int64_t nsurv = 0;
    double mflop = 0;
    auto QQ = [&](int r, int c) { return Q[r + (size_t)c * nS]; };
    auto MT = [&](int r, int c) { return mt[r + (size_t)c * nS]; };
    auto MA = [&](int r, int c) { return ma[r + (size_t)c * nS]; };
    auto MB = [&](int r, int c) { return mb[r + (size_t)c * nS]; };

#pragma omp parallel
    {
        std::vector<double> Jp((size_t)N * N, 0.0), Kap((size_t)N * N, 0.0), Kbp;
        if (uhf) Kbp.assign((size_t)N * N, 0.0);
        int64_t lcls[HFX_N_CLASSES] = {0};
        double lclsfl[HFX_N_CLASSES] = {0};
        uint64_t lhs = 0, lhx = 0;
        int64_t lns = 0;
        double lfl = 0;
        std::vector<double> G;
#pragma omp for collapse(4) schedule(dynamic, 32) nowait
        for (int a = a_lo; a < a_hi; ++a)
            for (int b = 0; b < nS; ++b)
                for (int c = 0; c < nS; ++c)
                    for (int d = 0; d < nS; ++d) {
                        if (b > a || d > c || (a * (a + 1) / 2 + b) < (c * (c + 1) / 2 + d)) continue;
                        if (QQ(a, b) * QQ(c, d) < tauS) continue;
                        double mx = std::max({2 * MT(a, b), 2 * MT(c, d), MA(a, c), MA(a, d), MA(b, c), MA(b, d)});
                        if (uhf) mx = std::max({mx, MB(a, c), MB(a, d), MB(b, c), MB(b, d)});
                        if (mx < tauD) continue;
                        const int nA = m.nao[a], nB = m.nao[b], nC = m.nao[c], nD = m.nao[d];
                        G.assign((size_t)nA * nB * nC * nD, 0.0);
                        eri_quartet(m, a, b, c, d, G.data());
                        lns++;
                        const int qcls = quartet_class(m.l[a], m.l[b], m.l[c], m.l[d]);
                        lcls[qcls]++;
                        const uint64_t h = splitmix64((((uint64_t)a * nS + b) * nS + c) * nS + d);
                        lhs += h;
                        lhx ^= h;
                        const double qfl = (double)m.nprim[a] * m.nprim[b] * m.nprim[c] * m.nprim[d] *
                                               model_flop(m.l[a], m.l[b], m.l[c], m.l[d]) +
                                           12.0 * nA * nB * nC * nD;
                        lfl += qfl;
                        lclsfl[qcls] += qfl;
                        const size_t nCD = (size_t)nC * nD, nBCD = nB * nCD;
                        // Digestion (SCF.cpp:633-711).
                        for (int A = 0; A < nA; ++A) {
                            const size_t i = A + m.aoff[a];
                            for (int B = 0; B < nB; ++B) {
                                const size_t j = B + m.aoff[b];
                                if (j > i) continue;
                                for (int C = 0; C < nC; ++C) {
                                    const size_t k = C + m.aoff[c];
                                    const size_t lmax = (i == k) ? j : k;
                                    for (int D = 0; D < nD; ++D) {
                                        const size_t l = D + m.aoff[d];
                                        if (l > lmax) continue;
                                        const double eri = G[A * nBCD + B * nCD + C * nD + D];
                                        if (std::abs(eri) < tauS) continue;
                                        const size_t ji = j + i * N, lk = l + k * N, li = l + i * N;
                                        const size_t ki = k + i * N, kj = k + j * N, lj = l + j * N;
                                        const double s_kl = (k == l) ? 1 : 2;
                                        const double s_ij = (i == j) ? 1 : 2;
                                        Jp[ji] += s_kl * dDt[lk] * eri;
                                        Jp[lk] += ((l != j || k != i) ? 1.0 : 0.0) * s_ij * dDt[ji] * eri;
                                        const double s_jk = (j == k) ? 2 : 1;
                                        const double s_jl = (j == l && i != k) ? 2 : 1;
                                        const double s_ik = (i == k && j != l) ? 2 : 1;
                                        const size_t i1 = i < k ? i + k * N : ki;
                                        const size_t i2 = j < l ? j + l * N : lj;
                                        const size_t i3 = j < k ? j + k * N : kj;
                                        const double m1 = (i1 != li) ? 1.0 : 0.0;
                                        const double m2 = (i2 != li && i2 != i1) ? 1.0 : 0.0;
                                        const double m3 = (i3 != li && i3 != i1 && i3 != i2) ? 1.0 : 0.0;
                                        Kap[li] += dDa[kj] * eri;
                                        Kap[i1] += m1 * s_ik * dDa[lj] * eri;
                                        Kap[i2] += m2 * s_jl * dDa[ki] * eri;
                                        Kap[i3] += m3 * s_jk * dDa[li] * eri;
                                        if (uhf) {
                                            Kbp[li] += dDb[kj] * eri;
                                            Kbp[i1] += m1 * s_ik * dDb[lj] * eri;
                                            Kbp[i2] += m2 * s_jl * dDb[ki] * eri;
                                            Kbp[i3] += m3 * s_jk * dDb[li] * eri;
                                        }
                                    }
                                }
                            }
                        }
                    }
#pragma omp critical(oracle_jk)
        {
            for (size_t x = 0; x < Jt.size(); ++x) {
                Jt[x] += Jp[x];
                Kat[x] += Kap[x];
                if (uhf) Kbt[x] += Kbp[x];
            }
            for (int q = 0; q < HFX_N_CLASSES; ++q) {
                cls[q] += lcls[q];
                clsfl[q] += lclsfl[q];
            }
            hsum_ += lhs;
            hxor ^= lhx;
            nsurv += lns;
            mflop += lfl;
        }
    }
    // selfadjointView<Upper> (SCF.cpp:727-728).
    auto sym = [&](std::vector<double>& M, double* out) {
        for (int c = 0; c < N; ++c)
            for (int r = 0; r < N; ++r) out[r + (size_t)c * N] = (r <= c) ? M[r + (size_t)c * N] : M[c + (size_t)r * N];
    };
    sym(Jt, J);
    sym(Kat, Ka);
    if (uhf && Kb) sym(Kbt, Kb);
    if (st) {
        st->n_quartets += nsurv;
        for (int q = 0; q < HFX_N_CLASSES; ++q) {
            st->class_count[q] += cls[q];
            st->class_gflop64[q] += clsfl[q] * 1e-9;
        }
        st->hash_sum += hsum_;
        st->hash_xor ^= hxor;
        st->model_gflop += mflop * 1e-9;
    }
}


// ---------------------------------------------------------------------------------------
// RHF nuclear gradient.  No reference counterpart (the reference computes energies only):
// the analytic derivative of the energy that the functions above define,
//   dE/dX = sum P dh/dX - sum W dS/dX + 1/2 sum Gamma d(mn|ls)/dX + dVnn/dX.
// A Cartesian Gaussian's centre derivative is d/dA_x g_i = 2a g_{i+1} - i g_{i-1}; applied
// to the 1-D Hermite coefficients it gives 2a E^{i+1,j}_t - i E^{i-1,j}_t (and likewise on
// the second centre with b and j).  The operator derivative of the nuclear attraction is
// dR_tuv/dC_x = -R_{t+1,u,v}.  Every centre is differentiated explicitly (no translational
// invariance is used), so sum over atoms of the gradient = 0 is a check, not a construction.
// Pinned by central finite differences of the SCF energy (tests/).
// ---------------------------------------------------------------------------------------

// 1-D coefficient of the base (w = 0), first-centre (w = 1) or second-centre (w = 2)
// derivative; E holds extents (l1+1, l2+1) so every shifted index exists.
double e_der(const EBuf& E, int i, int j, int t, int w, double a, double b) {
    auto at = [&](int ii, int jj) { return (ii < 0 || jj < 0 || t > ii + jj) ? 0.0 : E.at(ii, jj, t); };
    if (w == 0) return at(i, j);
    if (w == 1) return 2.0 * a * at(i + 1, j) - (double)i * at(i - 1, j);
    return 2.0 * b * at(i, j + 1) - (double)j * at(i, j - 1);
}

// (ab|cd) from per-dimension Hermite vectors: e[d][t] for the bra, f[d][t] for the ket.
double hermite_contract(const double* const e[3], const int ne[3], const double* const f[3], const int nf[3],
                        RBuf& R) {
    double acc = 0.0;
    for (int t = 0; t < ne[0]; ++t)
        for (int u = 0; u < ne[1]; ++u)
            for (int v = 0; v < ne[2]; ++v) {
                const double ebra = e[0][t] * e[1][u] * e[2][v];
                if (ebra == 0.0) continue;
                double s = 0.0;
                for (int tt = 0; tt < nf[0]; ++tt)
                    for (int uu = 0; uu < nf[1]; ++uu)
                        for (int vv = 0; vv < nf[2]; ++vv) {
                            const double sg = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
                            s += sg * f[0][tt] * f[1][uu] * f[2][vv] * R.at(t + tt, u + uu, v + vv, 0);
                        }
                acc += ebra * s;
            }
    return acc;
}

// g_out[4][3]: derivative of sum_{block} Gam * (ab|cd) w.r.t. the centres of a, b, c, d.
void eri_quartet_grad(const Mol& m, int sa, int sb, int sc, int sd, const double* Gam, double g_out[4][3]) {
    const int s[4] = {sa, sb, sc, sd};
    int nao[4], l[4], ao[4];
    for (int k = 0; k < 4; ++k) {
        nao[k] = m.nao[s[k]];
        l[k] = m.l[s[k]];
        ao[k] = m.aoff[s[k]];
    }
    const unsigned L = l[0] + l[1] + l[2] + l[3] + 1;
    RBuf R(L);
    std::vector<double> F(L + 1);
    EBuf Eb[3] = {EBuf(l[0] + 1, l[1] + 1), EBuf(l[0] + 1, l[1] + 1), EBuf(l[0] + 1, l[1] + 1)};
    EBuf Ek[3] = {EBuf(l[2] + 1, l[3] + 1), EBuf(l[2] + 1, l[3] + 1), EBuf(l[2] + 1, l[3] + 1)};
    const double* ctr[4];
    for (int k = 0; k < 4; ++k) ctr[k] = &m.ctr[3 * s[k]];
    for (int k = 0; k < 4; ++k)
        for (int x = 0; x < 3; ++x) g_out[k][x] = 0.0;
    const int nB = nao[1], nC = nao[2], nD = nao[3];
    for (int pa = 0; pa < m.nprim[sa]; ++pa)
        for (int pb = 0; pb < m.nprim[sb]; ++pb) {
            const double a = m.alpha[m.poff[sa] + pa], b = m.alpha[m.poff[sb] + pb];
            const PairData ab = pair_data(a, ctr[0], b, ctr[1]);
            for (int x = 0; x < 3; ++x) hermite(l[0] + 1, l[1] + 1, ab.p, ab.PA[x], ab.PB[x], Eb[x]);
            for (int pc = 0; pc < m.nprim[sc]; ++pc)
                for (int pd = 0; pd < m.nprim[sd]; ++pd) {
                    const double c = m.alpha[m.poff[sc] + pc], d = m.alpha[m.poff[sd] + pd];
                    const PairData cd = pair_data(c, ctr[2], d, ctr[3]);
                    for (int x = 0; x < 3; ++x) hermite(l[2] + 1, l[3] + 1, cd.p, cd.PA[x], cd.PB[x], Ek[x]);
                    const double pref = (2.0 * std::pow(M_PI, 2.5) / (ab.p * cd.p * std::sqrt(ab.p + cd.p))) *
                                        ab.K * cd.K;
                    const double delta = (ab.p * cd.p) / (ab.p + cd.p);
                    const double PQ[3] = {ab.P[0] - cd.P[0], ab.P[1] - cd.P[1], ab.P[2] - cd.P[2]};
                    boys(L, delta * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]), F.data());
                    hermite_r(delta, PQ, F.data(), R);
                    const double cprod = m.coeff[m.poff[sa] + pa] * m.coeff[m.poff[sb] + pb] *
                                         m.coeff[m.poff[sc] + pc] * m.coeff[m.poff[sd] + pd] * pref;
                    const double* na = &m.norm[m.noff[sa] + pa * nao[0]];
                    const double* nb = &m.norm[m.noff[sb] + pb * nao[1]];
                    const double* nc = &m.norm[m.noff[sc] + pc * nao[2]];
                    const double* nd = &m.norm[m.noff[sd] + pd * nao[3]];
                    // vectors: [w][dim][t], w = 0 base, 1 first centre, 2 second centre
                    double eb[3][3][8], fk[3][3][8];
                    for (int A = 0; A < nao[0]; ++A)
                        for (int B = 0; B < nB; ++B) {
                            const int* c1 = &m.cart[3 * (ao[0] + A)];
                            const int* c2 = &m.cart[3 * (ao[1] + B)];
                            int nbra[3];
                            for (int x = 0; x < 3; ++x) {
                                nbra[x] = c1[x] + c2[x] + 2;
                                for (int w = 0; w < 3; ++w)
                                    for (int t = 0; t < nbra[x]; ++t) eb[w][x][t] = e_der(Eb[x], c1[x], c2[x], t, w, a, b);
                            }
                            for (int C = 0; C < nC; ++C)
                                for (int D = 0; D < nD; ++D) {
                                    const double gm =
                                        Gam[((size_t)A * nB + B) * nC * nD + (size_t)C * nD + D] * na[A] * nb[B] *
                                        nc[C] * nd[D] * cprod;
                                    if (gm == 0.0) continue;
                                    const int* c3 = &m.cart[3 * (ao[2] + C)];
                                    const int* c4 = &m.cart[3 * (ao[3] + D)];
                                    int nket[3];
                                    for (int x = 0; x < 3; ++x) {
                                        nket[x] = c3[x] + c4[x] + 2;
                                        for (int w = 0; w < 3; ++w)
                                            for (int t = 0; t < nket[x]; ++t)
                                                fk[w][x][t] = e_der(Ek[x], c3[x], c4[x], t, w, c, d);
                                    }
                                    for (int x = 0; x < 3; ++x) {
                                        // bra centres: dimension x takes the derivative vector
                                        for (int w = 1; w <= 2; ++w) {
                                            const double* e3[3] = {eb[0][0], eb[0][1], eb[0][2]};
                                            e3[x] = eb[w][x];
                                            const double* f3[3] = {fk[0][0], fk[0][1], fk[0][2]};
                                            g_out[w - 1][x] += gm * hermite_contract(e3, nbra, f3, nket, R);
                                        }
                                        for (int w = 1; w <= 2; ++w) {
                                            const double* e3[3] = {eb[0][0], eb[0][1], eb[0][2]};
                                            const double* f3[3] = {fk[0][0], fk[0][1], fk[0][2]};
                                            f3[x] = fk[w][x];
                                            g_out[1 + w][x] += gm * hermite_contract(e3, nbra, f3, nket, R);
                                        }
                                    }
                                }
                        }
                }
        }
}

// Two-electron gradient: unique shell quartets (a>=b, c>=d, ab>=cd) with the J/K build's
// Schwarz and density tests on (P, P/2); every quartet counts deg = 2^(#distinct swaps)
// times, and the energy carries 1/2 sum Gamma (mn|ls).
void grad_eri(Mol& m, const double* P, const int* shell_atom, double tauS, double tauD, double* g,
              int64_t* n_quartets) {
    const int nS = m.nS, N = m.N;
    schwarz(m);
    const double* Q = m.Q.data();
    std::vector<double> mt((size_t)nS * nS, 0.0);
    for (int s1 = 0; s1 < nS; ++s1)
        for (int s2 = 0; s2 < nS; ++s2) {
            double v = 0;
            for (int a = m.aoff[s1]; a < m.aoff[s1] + m.nao[s1]; ++a)
                for (int b = m.aoff[s2]; b < m.aoff[s2] + m.nao[s2]; ++b) v = std::max(v, std::abs(P[a + (size_t)b * N]));
            mt[s1 + (size_t)s2 * nS] = v;
        }
    auto MT = [&](int r, int c) { return mt[r + (size_t)c * nS]; };
    auto QQ = [&](int r, int c) { return Q[r + (size_t)c * nS]; };
    const int nat = 1 + *std::max_element(shell_atom, shell_atom + nS);
    std::vector<double> gt(3 * (size_t)nat, 0.0);
    int64_t nq = 0;
#pragma omp parallel
    {
        std::vector<double> gp(3 * (size_t)nat, 0.0), Gam;
        int64_t lq = 0;
#pragma omp for collapse(4) schedule(dynamic, 16) nowait
        for (int a = 0; a < nS; ++a)
            for (int b = 0; b < nS; ++b)
                for (int c = 0; c < nS; ++c)
                    for (int d = 0; d < nS; ++d) {
                        if (b > a || d > c || (a * (a + 1) / 2 + b) < (c * (c + 1) / 2 + d)) continue;
                        if (QQ(a, b) * QQ(c, d) < tauS) continue;
                        // density test of the J/K build with dD_tot = P, dD_alpha = P/2
                        const double mx = std::max({2 * MT(a, b), 2 * MT(c, d), 0.5 * MT(a, c), 0.5 * MT(a, d),
                                                    0.5 * MT(b, c), 0.5 * MT(b, d)});
                        if (mx < tauD) continue;
                        lq++;
                        const int nA = m.nao[a], nB = m.nao[b], nC = m.nao[c], nD = m.nao[d];
                        Gam.assign((size_t)nA * nB * nC * nD, 0.0);
                        for (int A = 0; A < nA; ++A)
                            for (int B = 0; B < nB; ++B)
                                for (int C = 0; C < nC; ++C)
                                    for (int D = 0; D < nD; ++D) {
                                        const size_t i = m.aoff[a] + A, j = m.aoff[b] + B, k = m.aoff[c] + C,
                                                     l = m.aoff[d] + D;
                                        auto p = [&](size_t r, size_t cc) { return P[r + cc * N]; };
                                        Gam[((size_t)A * nB + B) * nC * nD + (size_t)C * nD + D] =
                                            p(i, j) * p(k, l) - 0.25 * (p(i, k) * p(j, l) + p(i, l) * p(j, k));
                                    }
                        double gq[4][3];
                        eri_quartet_grad(m, a, b, c, d, Gam.data(), gq);
                        const double deg = (a != b ? 2.0 : 1.0) * (c != d ? 2.0 : 1.0) * ((a != c || b != d) ? 2.0 : 1.0);
                        const int s4[4] = {a, b, c, d};
                        for (int k = 0; k < 4; ++k)
                            for (int x = 0; x < 3; ++x) gp[3 * shell_atom[s4[k]] + x] += 0.5 * deg * gq[k][x];
                    }
#pragma omp critical(oracle_grad)
        {
            for (size_t x = 0; x < gt.size(); ++x) gt[x] += gp[x];
            nq += lq;
        }
    }
    for (size_t x = 0; x < gt.size(); ++x) g[x] += gt[x];
    if (n_quartets) *n_quartets = nq;
}

// One-electron gradient terms: -W dS, P dT, P dV (basis centres and the Hellmann-Feynman
// operator term).  Every ordered shell pair (a, b) is visited, so each matrix element
// counts once.
void grad_one(const Mol& m, int nat, const double* Z, const double* xyz, const int* shell_atom, const double* P,
              const double* W, double* gS, double* gH) {
    const int N = m.N, nS = m.nS;
#pragma omp parallel
    {
        std::vector<double> lS(3 * (size_t)nat, 0.0), lH(3 * (size_t)nat, 0.0);
#pragma omp for collapse(2) schedule(dynamic)
        for (int sa = 0; sa < nS; ++sa)
            for (int sb = 0; sb < nS; ++sb) {
                const int nA = m.nao[sa], nB = m.nao[sb];
                const int lA = m.l[sa], lB = m.l[sb];
                const int atA = shell_atom[sa], atB = shell_atom[sb];
                EBuf E[3] = {EBuf(lA + 1, lB + 3), EBuf(lA + 1, lB + 3), EBuf(lA + 1, lB + 3)};
                const unsigned L = lA + lB + 1;
                RBuf R(L);
                std::vector<double> F(L + 1);
                for (int pa = 0; pa < m.nprim[sa]; ++pa)
                    for (int pb = 0; pb < m.nprim[sb]; ++pb) {
                        const double a = m.alpha[m.poff[sa] + pa], b = m.alpha[m.poff[sb] + pb];
                        const PairData d = pair_data(a, &m.ctr[3 * sa], b, &m.ctr[3 * sb]);
                        for (int x = 0; x < 3; ++x) hermite(lA + 1, lB + 3, d.p, d.PA[x], d.PB[x], E[x]);
                        const double cp = m.coeff[m.poff[sa] + pa] * m.coeff[m.poff[sb] + pb];
                        const double sq = std::pow(M_PI / d.p, 1.5) * d.K * cp;
                        const double vpref = (2.0 * M_PI / d.p) * d.K * cp;
                        for (int X = 0; X < nA; ++X) {
                            const int i = m.aoff[sa] + X;
                            const int* c1 = &m.cart[3 * i];
                            const double na = m.norm[m.noff[sa] + pa * nA + X];
                            for (int Y = 0; Y < nB; ++Y) {
                                const int j = m.aoff[sb] + Y;
                                const int* c2 = &m.cart[3 * j];
                                const double nb = m.norm[m.noff[sb] + pb * nB + Y];
                                const double w = W[i + (size_t)j * N], pd = P[i + (size_t)j * N];
                                const double f = na * nb;
                                // 1-D overlap factor of dimension x with second index jj; der = 0 base,
                                // 1 d/dA, 2 d/dB.
                                auto s1 = [&](int x, int jj, int der) {
                                    return e_der(E[x], c1[x], jj, 0, der, a, b);
                                };
                                // kinetic energy of the product with one dimension's factors
                                // differentiated (dx < 0: none).
                                auto kin = [&](int dx, int der) {
                                    double sx[3], tx[3];
                                    for (int x = 0; x < 3; ++x) {
                                        const int dd = (x == dx) ? der : 0;
                                        const int l2 = c2[x];
                                        sx[x] = s1(x, l2, dd);
                                        // -1/2 d^2/dx^2 acting on the second function (kinetic_pair)
                                        double t = b * (2 * l2 + 1) * sx[x] - 2.0 * b * b * s1(x, l2 + 2, dd);
                                        if (l2 >= 2) t -= 0.5 * l2 * (l2 - 1) * s1(x, l2 - 2, dd);
                                        tx[x] = t;
                                    }
                                    return tx[0] * sx[1] * sx[2] + sx[0] * tx[1] * sx[2] + sx[0] * sx[1] * tx[2];
                                };
                                for (int x = 0; x < 3; ++x) {
                                    for (int der = 1; der <= 2; ++der) {
                                        double ds = 1.0;
                                        for (int y = 0; y < 3; ++y) ds *= s1(y, c2[y], y == x ? der : 0);
                                        const int at = der == 1 ? atA : atB;
                                        lS[3 * at + x] -= w * f * sq * ds;
                                        lH[3 * at + x] += pd * f * sq * kin(x, der);
                                    }
                                }
                                // nuclear attraction
                                for (int C = 0; C < nat; ++C) {
                                    const double PC[3] = {d.P[0] - xyz[3 * C], d.P[1] - xyz[3 * C + 1],
                                                          d.P[2] - xyz[3 * C + 2]};
                                    boys(L, d.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), F.data());
                                    hermite_r(d.p, PC, F.data(), R);
                                    const double zf = -Z[C] * vpref * f * pd;
                                    double ev[3][3][8];
                                    int ne[3];
                                    for (int x = 0; x < 3; ++x) {
                                        ne[x] = c1[x] + c2[x] + 2;
                                        for (int der = 0; der < 3; ++der)
                                            for (int t = 0; t < ne[x]; ++t) ev[der][x][t] = e_der(E[x], c1[x], c2[x], t, der, a, b);
                                    }
                                    for (int x = 0; x < 3; ++x) {
                                        for (int der = 1; der <= 2; ++der) {
                                            double sum = 0.0;
                                            for (int t = 0; t < ne[0]; ++t)
                                                for (int u = 0; u < ne[1]; ++u)
                                                    for (int v = 0; v < ne[2]; ++v) {
                                                        const double e0 = (x == 0 ? ev[der][0][t] : ev[0][0][t]);
                                                        const double e1 = (x == 1 ? ev[der][1][u] : ev[0][1][u]);
                                                        const double e2 = (x == 2 ? ev[der][2][v] : ev[0][2][v]);
                                                        sum += e0 * e1 * e2 * R.at(t, u, v, 0);
                                                    }
                                            lH[3 * (der == 1 ? atA : atB) + x] += zf * sum;
                                        }
                                        // operator derivative: d/dC_x R_tuv(P - C) = -R_{t+1,u,v}
                                        double sum = 0.0;
                                        for (int t = 0; t < ne[0] - 1; ++t)
                                            for (int u = 0; u < ne[1] - 1; ++u)
                                                for (int v = 0; v < ne[2] - 1; ++v) {
                                                    const int tt = t + (x == 0), uu = u + (x == 1), vv = v + (x == 2);
                                                    sum -= ev[0][0][t] * ev[0][1][u] * ev[0][2][v] * R.at(tt, uu, vv, 0);
                                                }
                                        lH[3 * C + x] += zf * sum;
                                    }
                                }
                            }
                        }
                    }
            }
#pragma omp critical(oracle_grad1)
        for (size_t x = 0; x < lS.size(); ++x) {
            gS[x] += lS[x];
            gH[x] += lH[x];
        }
    }
}

void grad_nuc(int nat, const double* Z, const double* xyz, double* g) {
    for (int A = 0; A < nat; ++A)
        for (int B = 0; B < nat; ++B) {
            if (A == B) continue;
            const double d[3] = {xyz[3 * A] - xyz[3 * B], xyz[3 * A + 1] - xyz[3 * B + 1], xyz[3 * A + 2] - xyz[3 * B + 2]};
            const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            for (int x = 0; x < 3; ++x) g[3 * A + x] -= Z[A] * Z[B] * d[x] / (r * r * r);
        }
}

}  // namespace

namespace {

// Index of the unique element (ij|kl) in the packed canonical tensor (include/Molecule.hpp:42-56).
inline size_t vee_index(size_t i, size_t j, size_t k, size_t l) {
    if (j > i) std::swap(i, j);
    if (l > k) std::swap(k, l);
    size_t I = i * (i + 1) / 2 + j, K = k * (k + 1) / 2 + l;
    if (K > I) std::swap(I, K);
    return I * (I + 1) / 2 + K;
}

// IntegralEngine::electronRepulsionTensor (IntegralEngine.cpp:635-692): the diagonal quartets
// (i j|i j) of every shell pair, then every other canonical quartet with Q(i,j)Q(k,l) >= tau.
// Each block comes from eri_quartet with the reference's zero rules, whose surviving elements
// are exactly the unique AO quartets the shell quartet owns (IntegralEngine.cpp:600-626).
int store_tensor(Mol& m, const double* Qext, double tau) {
    const int nS = m.nS, N = m.N;
    const size_t T = (size_t)N * (N + 1) / 2;
    if (T * (T + 1) / 2 > ((size_t)1 << 31)) return fail(HFX_ERR_NOMEM, "oracle hfx_store_eris: tensor too large");
    const double* Q = Qext;
    if (!Q) {
        schwarz(m);
        Q = m.Q.data();
    }
    m.vee.assign(T * (T + 1) / 2, 0.0);
    std::vector<std::array<int, 4>> qs;
    int64_t counted = 0;
    for (int i = 0; i < nS; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k <= i; ++k)
                for (int l = 0; l <= k; ++l) {
                    if (i * (i + 1) / 2 + j < k * (k + 1) / 2 + l) continue;
                    const bool pass = Q[i + (size_t)j * nS] * Q[k + (size_t)l * nS] >= tau;
                    counted += pass;
                    if (pass || (i == k && j == l)) qs.push_back({i, j, k, l});
                }
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t x = 0; x < qs.size(); ++x) {
        const int a = qs[x][0], b = qs[x][1], c = qs[x][2], d = qs[x][3];
        const int nA = m.nao[a], nB = m.nao[b], nC = m.nao[c], nD = m.nao[d];
        std::vector<double> G((size_t)nA * nB * nC * nD, 0.0);
        eri_quartet(m, a, b, c, d, G.data());
        for (int p = 0; p < nA; ++p)
            for (int q = 0; q < nB; ++q)
                for (int r = 0; r < nC; ++r)
                    for (int t = 0; t < nD; ++t) {
                        const int i = m.aoff[a] + p, j = m.aoff[b] + q, k = m.aoff[c] + r, l = m.aoff[d] + t;
                        if (j > i || l > k) continue;
                        if (a == c && b == d && i * (i + 1) / 2 + j < k * (k + 1) / 2 + l) continue;
                        // every unique element belongs to one shell quartet: no two threads write it
                        m.vee[vee_index(i, j, k, l)] = G[(((size_t)p * nB + q) * nC + r) * nD + t];
                    }
    }
    m.stored_quartets = counted;
    return HFX_OK;
}

// SCF::buildFockMatrix (SCF.cpp:233-521) from the stored tensor: every unique element
// (ij|kl), i >= j, k >= l, ij >= kl, with |eri| >= tau, digested into all of its distinct index
// permutations: J(p,q) += (pq|rs) D_tot(r,s), K(p,s) += (pq|rs) D_x(q,r), each of the 8
// permutations weighted by 1/(number of coincident ones).  Full symmetric outputs.
void stored_jk(const Mol& m, const double* Dt, const double* Da, const double* Db, double tau, double* J,
               double* Ka, double* Kb) {
    const int N = m.N;
    std::fill(J, J + (size_t)N * N, 0.0);
    std::fill(Ka, Ka + (size_t)N * N, 0.0);
    if (Db) std::fill(Kb, Kb + (size_t)N * N, 0.0);
    auto at = [N](int r, int c) { return r + (size_t)c * N; };
    for (int i = 0; i < N; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k <= i; ++k)
                for (int l = 0; l <= k; ++l) {
                    if (i * (i + 1) / 2 + j < k * (k + 1) / 2 + l) continue;
                    const double v0 = m.vee[vee_index(i, j, k, l)];
                    if (std::abs(v0) < tau) continue;
                    double w = 1.0;
                    if (i == j) w *= 0.5;
                    if (k == l) w *= 0.5;
                    if (i == k && j == l) w *= 0.5;
                    const double v = v0 * w;
                    const int P[8][4] = {{i, j, k, l}, {j, i, k, l}, {i, j, l, k}, {j, i, l, k},
                                         {k, l, i, j}, {l, k, i, j}, {k, l, j, i}, {l, k, j, i}};
                    for (const auto& p : P) {
                        J[at(p[0], p[1])] += v * Dt[at(p[2], p[3])];
                        Ka[at(p[0], p[3])] += v * Da[at(p[1], p[2])];
                        if (Db) Kb[at(p[0], p[3])] += v * Db[at(p[1], p[2])];
                    }
                }
}

}  // namespace

// The host SCF driver linked into this library reports through this library's error text; the
// SCF iteration with device linear algebra needs the HIP library and is refused here.
namespace hfx_internal {

int set_error(int code, const std::string& msg) { return fail(code, msg); }

int device_cycles(hfx_engine*, const hfx_scf_options&, const ScfSetup&, hfx_scf_result*) {
    return fail(HFX_ERR_UNSUPPORTED, "device_linalg needs the HIP library");
}

}  // namespace hfx_internal

extern "C" {

const char* hfx_last_error(void) { return g_err.c_str(); }

int hfx_create_batched(int32_t n_mol, const hfx_basis* bases, int /*device*/, hfx_engine** out) {
    if (!out || !bases || n_mol <= 0) return fail(HFX_ERR_INVALID, "hfx_create_batched: bad arguments");
    hfx_engine* e = new hfx_engine();
    e->mols.resize(n_mol);
    for (int i = 0; i < n_mol; ++i) {
        const int rc = load_mol(&bases[i], e->mols[i]);
        if (rc) {
            delete e;
            return rc;
        }
        e->mat_off.push_back(e->elems);
        e->shell_off.push_back(e->nS_tot);
        e->elems += (int64_t)e->mols[i].N * e->mols[i].N;
        e->nS_tot += e->mols[i].nS;
        e->N_tot += e->mols[i].N;
    }
    *out = e;
    return HFX_OK;
}

int hfx_create(const hfx_basis* basis, int device, hfx_engine** out) {
    return hfx_create_batched(1, basis, device, out);
}

int hfx_destroy(hfx_engine* e) {
    delete e;
    return HFX_OK;
}

int hfx_sizes(const hfx_engine* e, int64_t* n_aos, int64_t* n_shells, int64_t* n_mol, int64_t* elems) {
    if (!e) return fail(HFX_ERR_INVALID, "null engine");
    if (n_aos) *n_aos = e->N_tot;
    if (n_shells) *n_shells = e->nS_tot;
    if (n_mol) *n_mol = (int64_t)e->mols.size();
    if (elems) *elems = e->elems;
    return HFX_OK;
}

static int one_electron(hfx_engine* e, int kind, int32_t nat, const double* Z, const double* xyz,
                        const int32_t* atom_mol, double* M) {
    if (!e || !M) return fail(HFX_ERR_INVALID, "null argument");
    for (size_t im = 0; im < e->mols.size(); ++im) {
        const Mol& m = e->mols[im];
        double* out = M + e->mat_off[im];
        std::fill(out, out + (size_t)m.N * m.N, 0.0);
        std::vector<double> z, x;
        for (int a = 0; a < nat; ++a)
            if (!atom_mol || atom_mol[a] == (int)im) {
                z.push_back(Z[a]);
                x.insert(x.end(), xyz + 3 * a, xyz + 3 * a + 3);
            }
#pragma omp parallel for collapse(2) schedule(dynamic)
        for (int i = 0; i < m.nS; ++i)
            for (int j = 0; j < m.nS; ++j) {
                if (j < i) continue;
                if (kind == 0) overlap_pair(m, i, j, out);
                else if (kind == 1) kinetic_pair(m, i, j, out);
                else nuclear_pair(m, (int)z.size(), z.data(), x.data(), i, j, out);
            }
        mirror_upper(out, m.N);
    }
    return HFX_OK;
}

int hfx_overlap(hfx_engine* e, double* S) { return one_electron(e, 0, 0, nullptr, nullptr, nullptr, S); }
int hfx_kinetic(hfx_engine* e, double* T) { return one_electron(e, 1, 0, nullptr, nullptr, nullptr, T); }
int hfx_nuclear(hfx_engine* e, int32_t nat, const double* Z, const double* xyz, const int32_t* atom_mol,
                double* V) {
    if (nat < 0 || (nat > 0 && (!Z || !xyz))) return fail(HFX_ERR_INVALID, "hfx_nuclear: bad atoms");
    return one_electron(e, 2, nat, Z, xyz, atom_mol, V);
}

// On the CPU there is no device memory to avoid: the same integrals as the allocating calls.
int hfx_one_electron_noalloc(hfx_engine* e, int32_t kind, int32_t n_atoms, const double* Z, const double* xyz,
                             const int32_t* atom_mol, double* M) {
    if (kind == 0) return hfx_overlap(e, M);
    if (kind == 1) return hfx_kinetic(e, M);
    if (kind == 2) return hfx_nuclear(e, n_atoms, Z, xyz, atom_mol, M);
    return fail(HFX_ERR_INVALID, "hfx_one_electron_noalloc: kind must be 0, 1 or 2");
}

int hfx_schwarz(hfx_engine* e, double* Q) {
    if (!e || !Q) return fail(HFX_ERR_INVALID, "null argument");
    int64_t off = 0;
    for (auto& m : e->mols) {
        schwarz(m);
        std::copy(m.Q.begin(), m.Q.end(), Q + off);
        off += (int64_t)m.nS * m.nS;
    }
    return HFX_OK;
}

int hfx_eri_quartets(hfx_engine* e, int64_t n, const int32_t* abcd, const int64_t* offsets, double* out) {
    if (!e || e->mols.size() != 1 || (n > 0 && (!abcd || !offsets || !out)))
        return fail(HFX_ERR_INVALID, "hfx_eri_quartets: bad arguments");
    const Mol& m = e->mols[0];
    for (int64_t q = 0; q < n; ++q)
        for (int k = 0; k < 4; ++k)
            if (abcd[4 * q + k] < 0 || abcd[4 * q + k] >= m.nS) return fail(HFX_ERR_INVALID, "shell index");
#pragma omp parallel for schedule(dynamic)
    for (int64_t q = 0; q < n; ++q) {
        const int* s = &abcd[4 * q];
        const size_t sz = (size_t)m.nao[s[0]] * m.nao[s[1]] * m.nao[s[2]] * m.nao[s[3]];
        std::fill(out + offsets[q], out + offsets[q] + sz, 0.0);
        eri_quartet(m, s[0], s[1], s[2], s[3], out + offsets[q]);
    }
    return HFX_OK;
}

// Schwarz tables of a batched engine are concatenated nS_m x nS_m blocks.
static int jk_impl(hfx_engine* e, const hfx_jk_args* args, double* J, double* Ka, double* Kb,
                   hfx_jk_stats* st, int a_lo, int a_hi) {
    if (!e || !args || !args->dD_tot || !args->dD_alpha || !J || !Ka)
        return fail(HFX_ERR_INVALID, "hfx_jk: null argument");
    if (args->dD_beta && !Kb) return fail(HFX_ERR_INVALID, "hfx_jk: dD_beta given without Kb");
    if (!all_finite(args->dD_tot, e->elems) || !all_finite(args->dD_alpha, e->elems) ||
        (args->dD_beta && !all_finite(args->dD_beta, e->elems)))
        return fail(HFX_ERR_NONFINITE, "hfx_jk: non-finite density");
    if (st) std::memset(st, 0, sizeof(*st));
    if (args->flags & HFX_JK_STORED) {
        if (!e->stored) return fail(HFX_ERR_INVALID, "hfx_jk: HFX_JK_STORED without stored integrals");
        if (args->precision != HFX_FP64) return fail(HFX_ERR_UNSUPPORTED, "hfx_jk: HFX_JK_STORED needs HFX_FP64");
        for (size_t im = 0; im < e->mols.size(); ++im) {
            const Mol& m = e->mols[im];
            const int64_t o = e->mat_off[im];
            stored_jk(m, args->dD_tot + o, args->dD_alpha + o, args->dD_beta ? args->dD_beta + o : nullptr,
                      args->schwarz_thresh, J + o, Ka + o, args->dD_beta ? Kb + o : nullptr);
            if (st) st->n_quartets += m.stored_quartets;
        }
        return HFX_OK;
    }
    int64_t qoff = 0;
    for (size_t im = 0; im < e->mols.size(); ++im) {
        Mol& m = e->mols[im];
        const int64_t o = e->mat_off[im];
        const double* Qm = args->Q ? args->Q + qoff : nullptr;
        const int lo = std::max(0, a_lo), hi = std::min(m.nS, a_hi);
        direct_jk(m, args->dD_tot + o, args->dD_alpha + o, args->dD_beta ? args->dD_beta + o : nullptr, Qm,
                  args->schwarz_thresh, args->density_thresh, lo, std::max(lo, hi), J + o, Ka + o,
                  args->dD_beta ? Kb + o : nullptr, st);
        qoff += (int64_t)m.nS * m.nS;
    }
    return HFX_OK;
}

int hfx_jk(hfx_engine* e, const hfx_jk_args* args, double* J, double* Ka, double* Kb, hfx_jk_stats* st) {
    return jk_impl(e, args, J, Ka, Kb, st, 0, 1 << 30);
}

// Launch parameters of device kernels: nothing to set on the CPU (checked like the engine does).
int hfx_set_launch(hfx_engine* e, int32_t blocks_per_cu, int32_t chunk_len, int32_t streams) {
    if (!e || blocks_per_cu < 0 || blocks_per_cu > 16 || (chunk_len != 0 && (chunk_len < 256 || chunk_len > 65536)) ||
        streams < 0 || streams > 4)
        return fail(HFX_ERR_INVALID, "hfx_set_launch: bad arguments");
    return HFX_OK;
}

// The CPU oracle has no device and no stream: the "device" buffers of the async variant are host
// buffers and the build completes before the call returns (HFX_JK_DETERMINISTIC is accepted and
// changes nothing here).
int hfx_jk_async(hfx_engine* e, const hfx_jk_args* args, double* J, double* Ka, double* Kb, void* stream) {
    (void)stream;
    if (!args) return fail(HFX_ERR_INVALID, "hfx_jk_async: null argument");
    if (!(args->flags & HFX_JK_DEVICE_PTRS) || (args->flags & HFX_JK_CLASS_TIMING))
        return fail(HFX_ERR_INVALID, "hfx_jk_async: needs HFX_JK_DEVICE_PTRS and no HFX_JK_CLASS_TIMING");
    hfx_jk_args a = *args;
    a.flags &= ~HFX_JK_DEVICE_PTRS;
    return jk_impl(e, &a, J, Ka, Kb, nullptr, 0, 1 << 30);
}

int hfx_store_eris(hfx_engine* e, const double* Q, double schwarz_thresh, int64_t* n_quartets, int64_t* n_bytes) {
    if (!e) return fail(HFX_ERR_INVALID, "hfx_store_eris: null engine");
    if (!(schwarz_thresh >= 0.0) || !std::isfinite(schwarz_thresh))
        return fail(HFX_ERR_INVALID, "hfx_store_eris: bad schwarz_thresh");
    e->stored = false;
    int64_t n = 0, bytes = 0, qoff = 0;
    for (Mol& m : e->mols) {
        const int rc = store_tensor(m, Q ? Q + qoff : nullptr, schwarz_thresh);
        if (rc) {
            for (Mol& x : e->mols) std::vector<double>().swap(x.vee);
            return rc;
        }
        n += m.stored_quartets;
        bytes += (int64_t)(m.vee.size() * sizeof(double));
        qoff += (int64_t)m.nS * m.nS;
    }
    e->stored = true;
    if (n_quartets) *n_quartets = n;
    if (n_bytes) *n_bytes = bytes;
    return HFX_OK;
}

int hfx_release_eris(hfx_engine* e) {
    if (!e) return fail(HFX_ERR_INVALID, "hfx_release_eris: null engine");
    for (Mol& m : e->mols) std::vector<double>().swap(m.vee);
    e->stored = false;
    return HFX_OK;
}

int hfx_rhf_gradient(hfx_engine* e, const hfx_grad_args* a, double* grad, double* parts, hfx_grad_stats* st) {
    if (!e || !a || !grad || !a->P || !a->W || a->n_atoms <= 0 || !a->Z || !a->xyz)
        return fail(HFX_ERR_INVALID, "hfx_rhf_gradient: null argument");
    if (e->mols.size() != 1) return fail(HFX_ERR_INVALID, "hfx_rhf_gradient: single-molecule engines only");
    if (a->flags & HFX_JK_DEVICE_PTRS) return fail(HFX_ERR_INVALID, "oracle: host pointers only");
    Mol& m = e->mols[0];
    const int nat = a->n_atoms;
    if (!all_finite(a->P, e->elems) || !all_finite(a->W, e->elems))
        return fail(HFX_ERR_NONFINITE, "hfx_rhf_gradient: non-finite density");
    std::vector<int> sat(m.nS);
    for (int s = 0; s < m.nS; ++s) {
        if (a->shell_atom) {
            sat[s] = a->shell_atom[s];
        } else {
            sat[s] = -1;
            for (int A = 0; A < nat && sat[s] < 0; ++A)
                if (m.ctr[3 * s] == a->xyz[3 * A] && m.ctr[3 * s + 1] == a->xyz[3 * A + 1] &&
                    m.ctr[3 * s + 2] == a->xyz[3 * A + 2])
                    sat[s] = A;
        }
        if (sat[s] < 0 || sat[s] >= nat) return fail(HFX_ERR_INVALID, "hfx_rhf_gradient: shell not on an atom");
    }
    const size_t n3 = 3 * (size_t)nat;
    std::vector<double> g[4];
    for (auto& v : g) v.assign(n3, 0.0);
    grad_nuc(nat, a->Z, a->xyz, g[0].data());
    grad_one(m, nat, a->Z, a->xyz, sat.data(), a->P, a->W, g[1].data(), g[2].data());
    int64_t nq = 0;
    grad_eri(m, a->P, sat.data(), a->schwarz_thresh, a->density_thresh, g[3].data(), &nq);
    for (size_t x = 0; x < n3; ++x) grad[x] = g[0][x] + g[1][x] + g[2][x] + g[3][x];
    if (parts)
        for (int k = 0; k < 4; ++k) std::copy(g[k].begin(), g[k].end(), parts + k * n3);
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->n_quartets = nq;
    }
    return HFX_OK;
}

// --- oracle-only entry points (tests) -----------------------------------------------------

// The J/K loop restricted to bra shells a in [a_lo, a_hi): a bounded sample of a build.
int oracle_jk_range(hfx_engine* e, const hfx_jk_args* args, int32_t a_lo, int32_t a_hi, double* J, double* Ka,
                    double* Kb, hfx_jk_stats* st) {
    return jk_impl(e, args, J, Ka, Kb, st, a_lo, a_hi);
}

void oracle_boys(int32_t m_max, double T, double* F) { boys((unsigned)m_max, T, F); }
double oracle_boys_series(int32_t m, double T) { return boys_series((unsigned)m, T); }

// Packed E^{ij}_t for one direction; out has (l1+1)(l2+1)(l1+l2+2)/2 entries.
void oracle_hermite(int32_t l1, int32_t l2, double p, double PA, double PB, double* out) {
    EBuf E(l1, l2);
    hermite(l1, l2, p, PA, PB, E);
    std::copy(E.v.begin(), E.v.end(), out);
}
// Index of E^{ij}_t in the packed buffer above.
int32_t oracle_hermite_index(int32_t l1, int32_t l2, int32_t i, int32_t j, int32_t t) {
    EBuf E;
    E.l1 = l1;
    E.l2 = l2;
    return (int32_t)(E.ioff(i) + E.joff(i, j) + t);
}
// R^0_{tuv} in a dense (L+1)^3 array, index t + u(L+1) + v(L+1)^2.
void oracle_hermite_r(int32_t L, double p, const double* PC, double* out) {
    std::vector<double> F(L + 1);
    const double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
    boys(L, T, F.data());
    RBuf R(L);
    hermite_r(p, PC, F.data(), R);
    std::copy(R.v.begin(), R.v.begin() + R.sn, out);
}
double oracle_model_flop(int32_t la, int32_t lb, int32_t lc, int32_t ld) { return model_flop(la, lb, lc, ld); }

}  // extern "C"
