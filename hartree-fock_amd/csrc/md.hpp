// md.hpp - McMurchie-Davidson building blocks for CDNA4, fully unrolled at compile time.
//
// Everything here is templated on the angular momenta so that every array is indexed by a
// constant expression and lives in VGPRs:
//   hermite_e   1-D Hermite expansion coefficients E^{ij}_t of a Gaussian product
//               (the recursion of the reference's computeHermiteCoeffs,
//               src/IntegralEngine.cpp:1025-1079, restated per direction)
//   hermite_r   Hermite Coulomb integrals R^0_{tuv}, built in place level by level from
//               R^n_{000} = (-2 alpha)^n F_n(T) (src/IntegralEngine.cpp:1082-1162)
//   boys        F_m(T), m = 0..L: asymptotic form above T = 30 as the reference
//               (src/Boys.cpp:16-23); below, an NT-term Taylor expansion around the nearest
//               point of a 0.1-spaced table (kept in LDS), then downward recursion.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace md {

template <int I>
using ic = std::integral_constant<int, I>;

template <class F, int... Is>
__device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(ic<Is>{}), ...);
}
// Calls f(ic<0>), ..., f(ic<N-1>) in order.
template <int N, class F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (N > 0) sfor_impl(f, std::make_integer_sequence<int, N>{});
}

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int tri(int m) { return m * (m + 1) / 2; }
__host__ __device__ constexpr int tet(int s) { return s * (s + 1) * (s + 2) / 6; }
__host__ __device__ constexpr int nherm(int L) { return tet(L + 1); }
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }

// Cartesian exponents of component c of a shell of angular momentum l, in the reference's
// AO order (src/Basis.cpp:47-59): i from l down to 0, j from l-i down to 0, k = l-i-j.
__host__ __device__ constexpr int cart_i(int l, int c) {
    int idx = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (idx == c) return i;
            ++idx;
        }
    return 0;
}
__host__ __device__ constexpr int cart_j(int l, int c) {
    int idx = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (idx == c) return j;
            ++idx;
        }
    return 0;
}
__host__ __device__ constexpr int cart_k(int l, int c) { return l - cart_i(l, c) - cart_j(l, c); }

// Packed Hermite index of (t,u,v): all triples of total s = t+u+v follow those of total s-1.
__host__ __device__ constexpr int hidx(int t, int u, int v) { return tet(t + u + v) + tri(u + v) + v; }
// Inverse within one level s: entry k has u+v = m with tri(m) <= k < tri(m+1).
__host__ __device__ constexpr int lev_m(int k) {
    int m = 0;
    while (tri(m + 1) <= k) ++m;
    return m;
}
// Full inverse of hidx.
__host__ __device__ constexpr int h_level(int h) {
    int s = 0;
    while (tet(s + 1) <= h) ++s;
    return s;
}
__host__ __device__ constexpr int h_v(int h) { return (h - tet(h_level(h))) - tri(lev_m(h - tet(h_level(h)))); }
__host__ __device__ constexpr int h_u(int h) { return lev_m(h - tet(h_level(h))) - h_v(h); }
__host__ __device__ constexpr int h_t(int h) { return h_level(h) - lev_m(h - tet(h_level(h))); }

// sum_{k<N} term(ic<k>), starting from the first term (no added zero).
template <int N, class F>
__device__ __forceinline__ auto sum_of(F&& term) {
    auto s = term(ic<0>{});
    sfor<N - 1>([&](auto K_) { s += term(ic<decltype(K_)::value + 1>{}); });
    return s;
}

// E[i][j][t] for i <= LA, j <= LB, t <= i+j, one Cartesian direction.
template <int LA, int LB, class R>
struct Ecoef {
    R e[LA + 1][LB + 1][LA + LB + 1];
};

template <int LA, int LB, class R>
__device__ __forceinline__ void hermite_e(R oo2p, R PA, R PB, Ecoef<LA, LB, R>& E) {
    E.e[0][0][0] = R(1);
    sfor<LA>([&](auto I_) {
        constexpr int i = decltype(I_)::value + 1;
        sfor<i + 1>([&](auto T_) {
            constexpr int t = decltype(T_)::value;
            R v;
            if constexpr (t <= i - 1) {
                v = PA * E.e[i - 1][0][t];
                if constexpr (t >= 1) v += oo2p * E.e[i - 1][0][t - 1];
            } else {
                v = oo2p * E.e[i - 1][0][t - 1];
            }
            if constexpr (t + 1 <= i - 1) v += R(t + 1) * E.e[i - 1][0][t + 1];
            E.e[i][0][t] = v;
        });
    });
    sfor<LB>([&](auto J_) {
        constexpr int j = decltype(J_)::value + 1;
        sfor<LA + 1>([&](auto I_) {
            constexpr int i = decltype(I_)::value;
            constexpr int top = i + j;
            sfor<top + 1>([&](auto T_) {
                constexpr int t = decltype(T_)::value;
                R v;
                if constexpr (t <= top - 1) {
                    v = PB * E.e[i][j - 1][t];
                    if constexpr (t >= 1) v += oo2p * E.e[i][j - 1][t - 1];
                } else {
                    v = oo2p * E.e[i][j - 1][t - 1];
                }
                if constexpr (t + 1 <= top - 1) v += R(t + 1) * E.e[i][j - 1][t + 1];
                E.e[i][j][t] = v;
            });
        });
    });
}

// R[hidx(t,u,v)] = R^0_{tuv} for t+u+v <= L, given base[n] = scale * (-2 alpha)^n F_n(T).
template <int L, class R>
__device__ __forceinline__ void hermite_r(const R (&base)[L + 1], R X, R Y, R Z, R (&Rt)[nherm(L)]) {
    Rt[0] = base[L];
    sfor<L>([&](auto N_) {
        constexpr int n = L - 1 - decltype(N_)::value;
        sfor<L - n>([&](auto S_) {
            constexpr int s = L - n - decltype(S_)::value;  // level s = L-n, ..., 1
            sfor<tri(s + 1)>([&](auto K_) {
                constexpr int k = decltype(K_)::value;
                constexpr int m = lev_m(k);
                constexpr int v = k - tri(m);
                constexpr int u = m - v;
                constexpr int t = s - m;
                R val;
                if constexpr (t > 0) {
                    val = X * Rt[hidx(t - 1, u, v)];
                    if constexpr (t > 1) val += R(t - 1) * Rt[hidx(t - 2, u, v)];
                } else if constexpr (u > 0) {
                    val = Y * Rt[hidx(t, u - 1, v)];
                    if constexpr (u > 1) val += R(u - 1) * Rt[hidx(t, u - 2, v)];
                } else {
                    val = Z * Rt[hidx(t, u, v - 1)];
                    if constexpr (v > 1) val += R(v - 1) * Rt[hidx(t, u, v - 2)];
                }
                Rt[hidx(t, u, v)] = val;
            });
        });
        Rt[0] = base[n];
    });
}

// ---------------------------------------------------------------------------------------
// Boys function.  The tables hold F_m(i h), m = 0..16, and e^{-i h} at the grid points of
// BoysGrid (h = 0.1 up to T = 30 in FP64, h = 0.05 in FP32).  A kernel whose highest order is
// L copies the Taylor coefficients of the orders it evaluates into LDS (BoysSlice).
// ---------------------------------------------------------------------------------------
constexpr int BOYS_POINTS = 301;  // FP64 grid
constexpr int BOYS_ORDERS = 18;   // columns per grid point: F_0 .. F_16, then e^{-T_i}
constexpr int BOYS_EXP_COL = 17;
constexpr bool BOYS_SELECT_FP32 = true;
constexpr bool BOYS_SELECT_FP64 = false;

template <class R>
struct BoysTerms;
template <>
struct BoysTerms<double> {
    static constexpr int nt = 8;  // |dT| <= 0.05: remainder < 1e-15 relative
};
template <>
struct BoysTerms<float> {
    static constexpr int nt = 4;  // |dT| <= 0.025: remainder < 2e-8 relative
};

template <class R>
struct BoysGrid;
template <>
struct BoysGrid<double> {
    static constexpr int points = BOYS_POINTS;
    static constexpr double step = 0.1, inv_step = 10.0;
};
template <>
struct BoysGrid<float> {
    static constexpr int points = 601;
    static constexpr float step = 0.05f, inv_step = 20.0f;
};
__host__ __device__ constexpr int boys_points(bool fp64) { return fp64 ? 301 : 601; }

// Layout of the LDS slice of a kernel whose highest order is L.  An FP64 slice used with the
// downward recursion also carries e^{-T_i} (its last column).  Taylor coefficients are stored
// divided by k! (one FMA per term), one run of nt per order evaluated from the table; the FP64
// slice of L = 1 keeps the orders shared (F_0..F_8 once), which halves its LDS.
__host__ __device__ constexpr bool boys_slice_texp(int L, bool fp64) { return L >= 2 && fp64; }
__host__ __device__ constexpr bool boys_slice_pre(int L, bool fp64) { return !(L == 1 && fp64); }
__host__ __device__ constexpr int boys_slice_nord(int L, bool fp64) {
    const int nt = fp64 ? 8 : 4;
    const int ntab = L <= 1 ? L + 1 : 1;  // orders evaluated from the table
    return (boys_slice_pre(L, fp64) ? ntab * nt : (L <= 1 ? L + nt : nt)) + (boys_slice_texp(L, fp64) ? 1 : 0);
}

template <int L, class R>
struct BoysSlice {
    static constexpr int nt = BoysTerms<R>::nt;
    static constexpr int first = (L <= 1) ? 0 : L;
    static constexpr bool texp = boys_slice_texp(L, sizeof(R) == 8);
    static constexpr bool pre = boys_slice_pre(L, sizeof(R) == 8);
    static constexpr int nord = boys_slice_nord(L, sizeof(R) == 8);
    static_assert(nt == (sizeof(R) == 8 ? 8 : 4), "boys_slice_nord assumes these term counts");
    static_assert(BoysGrid<R>::points == boys_points(sizeof(R) == 8), "grid sizes");
};


// 1/sqrt(x): the hardware estimate, refined by two Newton steps in FP64 (relative error of
// a few ulp); the single-precision estimate is used as is in FP32.
template <class R>
__device__ __forceinline__ R dev_rsqrt(R x);
template <>
__device__ __forceinline__ double dev_rsqrt<double>(double x) {
    double r = __builtin_amdgcn_rsq(x);
    double e = fma(-x * r, r, 1.0);
    r = fma(0.5 * r, e, r);
    e = fma(-x * r, r, 1.0);
    return fma(0.5 * r, e, r);
}
template <>
__device__ __forceinline__ float dev_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <class R>
__device__ __forceinline__ R dev_exp(R x) { return exp(x); }
template <>
__device__ __forceinline__ float dev_exp<float>(float x) { return __expf(x); }

// Copy the slice a kernel of total angular momentum L needs into LDS.
template <int L, class R>
__device__ __forceinline__ void load_boys(const R* __restrict__ g, R* lds) {
    using S = BoysSlice<L, R>;
    constexpr R inv_fact[10] = {R(1.0), R(1.0), R(1.0 / 2), R(1.0 / 6), R(1.0 / 24), R(1.0 / 120), R(1.0 / 720),
                                R(1.0 / 5040), R(1.0 / 40320), R(1.0 / 362880)};
    for (int x = threadIdx.x; x < BoysGrid<R>::points * S::nord; x += blockDim.x) {
        const int i = x / S::nord, m = x - i * S::nord;
        R v;
        if (S::texp && m == S::nord - 1) {
            v = g[i * BOYS_ORDERS + BOYS_EXP_COL];
        } else if (S::pre) {
            const int o = m / S::nt, k = m - o * S::nt;  // order first + o, Taylor term k
            v = g[i * BOYS_ORDERS + S::first + o + k] * inv_fact[k];
        } else {
            v = g[i * BOYS_ORDERS + S::first + m];
        }
        lds[x] = v;
    }
}

// Sum of row[k] x^k, k < NT (coefficients already divided by k!).
template <int NT, class R>
__device__ __forceinline__ R horner(const R* row, R x) {
    R t = row[NT - 1];
    sfor<NT - 1>([&](auto K_) {
        constexpr int k = NT - 2 - decltype(K_)::value;
        t = row[k] + x * t;
    });
    return t;
}

// Taylor sum of row[0..nt) at offset x = T0 - T: sum_k row[k] x^k / k!.
template <int NT, class R>
__device__ __forceinline__ R taylor(const R* row, R x) {
    R t = row[NT - 1];
    sfor<NT - 1>([&](auto K_) {
        constexpr int k = NT - 2 - decltype(K_)::value;
        t = row[k] + (x * R(1.0 / (k + 1))) * t;
    });
    return t;
}

// tab points at order S::first of grid point 0; consecutive grid points are STRIDE apart
// (STRIDE = S::nord for an LDS slice, BOYS_ORDERS for the full table in global memory).
// Asymptotic form above T = 30, as the reference (src/Boys.cpp:16-23): F_0 = sqrt(pi/4T) and
// the upward recursion without the e^{-T} term.
template <int L, class R>
__device__ __forceinline__ void boys_asymptotic(R T, R (&F)[L + 1]) {
    const R rt = dev_rsqrt(T);
    F[0] = R(0.8862269254527579) * rt;
    const R invT = rt * rt;
    sfor<L>([&](auto M_) {
        constexpr int m = decltype(M_)::value;
        F[m + 1] = (R(m) + R(0.5)) * F[m] * invT;
    });
}

// Taylor expansion around the nearest grid point, for 0 <= T <= 30 (larger T is clamped to
// the last point, which keeps the result finite).
template <int L, class R, int STRIDE>
__device__ __forceinline__ void boys_table(R T, const R* tab, R (&F)[L + 1]) {
    using S = BoysSlice<L, R>;
    using G = BoysGrid<R>;
    int i = (int)(T * G::inv_step + R(0.5));
    i = i < 0 ? 0 : (i > G::points - 1 ? G::points - 1 : i);  // NaN input stays in bounds
    const R x = R(i) * G::step - T;
    const R* row = tab + i * STRIDE;
    constexpr bool slice = STRIDE == S::nord;  // LDS slice (else the full table in global memory)
    if constexpr (L <= 1) {
        sfor<L + 1>([&](auto M_) {
            constexpr int m = decltype(M_)::value;
            if constexpr (slice && S::pre)
                F[m] = horner<S::nt>(row + m * S::nt, x);
            else
                F[m] = taylor<S::nt>(row + m, x);
        });
    } else {
        if constexpr (slice && S::pre)
            F[L] = horner<S::nt>(row, x);
        else
            F[L] = taylor<S::nt>(row, x);
        R e;
        if constexpr (S::texp && slice) {
            // e^{-T} = e^{-T_i} e^{x} with |x| <= 0.05: nine Taylor terms of e^x (remainder
            // below 6e-18 relative)
            constexpr R c[9] = {R(1.0), R(1.0), R(1.0 / 2), R(1.0 / 6), R(1.0 / 24), R(1.0 / 120), R(1.0 / 720),
                                R(1.0 / 5040), R(1.0 / 40320)};
            R ex = horner<9>(c, x);
            e = row[S::nord - 1] * ex;
        } else {
            e = dev_exp(-T);
        }
        const R twoT = T + T;
        sfor<L>([&](auto M_) {
            constexpr int m = L - decltype(M_)::value;  // m = L .. 1
            F[m - 1] = (twoT * F[m] + e) * R(1.0 / (2 * m - 1));
        });
    }
}

// Whether boys() evaluates both forms and selects per lane (no divergent branch) or branches.
template <class R>
struct BoysSelect {
    static constexpr bool value = BOYS_SELECT_FP32;
};
template <>
struct BoysSelect<double> {
    static constexpr bool value = BOYS_SELECT_FP64;
};

template <int L, class R, int STRIDE = BoysSlice<L, R>::nord>
__device__ __forceinline__ void boys(R T, const R* tab, R (&F)[L + 1]) {
    if constexpr (BoysSelect<R>::value) {
        R Fa[L + 1];
        boys_asymptotic<L>(T, Fa);
        boys_table<L, R, STRIDE>(T < R(30) ? T : R(30), tab, F);
        const bool far = T > R(30);
        sfor<L + 1>([&](auto M_) {
            constexpr int m = decltype(M_)::value;
            F[m] = far ? Fa[m] : F[m];
        });
    } else {
        if (T > R(30)) {
            boys_asymptotic<L>(T, F);
            return;
        }
        boys_table<L, R, STRIDE>(T, tab, F);
    }
}

}  // namespace md

namespace md {
template <class F, int... Is>
inline void sfor_host_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(ic<Is>{}), ...);
}
// Host-side counterpart of sfor (launch loops over compile-time chunk indices).
template <int N, class F>
inline void sfor_host(F&& f) {
    if constexpr (N > 0) sfor_host_impl(f, std::make_integer_sequence<int, N>{});
}
}  // namespace md
