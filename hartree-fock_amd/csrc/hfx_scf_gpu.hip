// hfx_scf_gpu.hip - the SCF iteration on the device (hfx_scf_options.device_linalg = 1).
//
// The cycle of the host driver (hfx_scf.cpp; the reference's SCF::run, src/SCF.cpp:23-136,
// SCF::diagonalizeAndUpdate, :747-835, DIIS, src/DIIS.cpp:11-81, inverseSqrtMatrix,
// src/Utils.cpp:117-138) with every matrix resident in HBM:
//   Fock build ............ hfx_jk_async on the driver's stream, device pointers
//   F = F_prev + J - K .... elementwise kernel
//   DIIS .................. e = F D S - S D F from one DGEMM pair and a transpose-difference
//                           kernel (F, D, S symmetric: S D F = (F D S)^T); a ring of error
//                           vectors and Fock matrices; the Gram matrix of the ring in one DGEMM;
//                           the bordered least-squares system solved by Householder QR with
//                           column pivoting in a one-thread kernel, as the host driver solves
//                           it; the extrapolation kernel
//   diagonalization ....... X^T F X (+ level shift) by DGEMM, rocSOLVER DSYEVD, C = X V
//   density, energy ....... C_occ C_occ^T by DGEMM, damping, and one fixed-order reduction of
//                           tr(D (h + F)), |D_tot - D_tot,prev|_F and |e|_F
// The host keeps the control flow: the damping and level-shift decisions and the convergence
// test, from one read-back of a few scalars per iteration.  rocBLAS and rocSOLVER are loaded
// with dlopen by the first such run in a process (time_linalg_load_s), so that programs that
// never ask for this mode do not pay their load time.
#include <hip/hip_runtime.h>
#include <dlfcn.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "scf_device.hpp"

namespace {

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

using hfx_internal::set_error;

// ---------------------------------------------------------------------------------------
// rocBLAS / rocSOLVER, loaded on first use; one handle per device, kept for the process.
// ---------------------------------------------------------------------------------------
struct RocLib {
    bool tried = false, ok = false;
    std::string err;
    decltype(&rocblas_create_handle) create_handle = nullptr;
    decltype(&rocblas_set_stream) set_stream = nullptr;
    decltype(&rocblas_dgemm) dgemm = nullptr;
    decltype(&rocsolver_dsyevd) dsyevd = nullptr;
    std::map<int, rocblas_handle> handles;
    std::map<int, std::mutex> busy;  // one run at a time per device (a handle is not re-entrant)
};
RocLib g_roc;
std::mutex g_roc_mu;

template <class F>
bool sym(void* lib, const char* name, F& f) {
    f = reinterpret_cast<F>(dlsym(lib, name));
    return f != nullptr;
}

// The libraries and this device's handle; the seconds spent here go to *load_s.
int roc_for_device(int device, rocblas_handle* h, std::mutex** busy, double* load_s) {
    std::lock_guard<std::mutex> g(g_roc_mu);
    const auto t0 = clk::now();
    if (!g_roc.tried) {
        g_roc.tried = true;
        void* hb = dlopen("librocblas.so.5", RTLD_NOW | RTLD_GLOBAL);
        if (!hb) hb = dlopen("librocblas.so", RTLD_NOW | RTLD_GLOBAL);
        void* hs = hb ? dlopen("librocsolver.so.0", RTLD_NOW | RTLD_GLOBAL) : nullptr;
        if (hb && !hs) hs = dlopen("librocsolver.so", RTLD_NOW | RTLD_GLOBAL);
        if (!hb || !hs) {
            const char* d = dlerror();
            g_roc.err = std::string("cannot load rocBLAS/rocSOLVER: ") + (d ? d : "not found");
        } else if (!sym(hb, "rocblas_create_handle", g_roc.create_handle) ||
                   !sym(hb, "rocblas_set_stream", g_roc.set_stream) || !sym(hb, "rocblas_dgemm", g_roc.dgemm) ||
                   !sym(hs, "rocsolver_dsyevd", g_roc.dsyevd)) {
            g_roc.err = "rocBLAS/rocSOLVER lack an expected entry point";
        } else {
            g_roc.ok = true;
        }
    }
    if (!g_roc.ok) return set_error(HFX_ERR_UNSUPPORTED, "device_linalg: " + g_roc.err);
    auto it = g_roc.handles.find(device);
    if (it == g_roc.handles.end()) {
        rocblas_handle nh = nullptr;
        if (g_roc.create_handle(&nh) != rocblas_status_success)
            return set_error(HFX_ERR_DEVICE, "device_linalg: rocblas_create_handle failed");
        it = g_roc.handles.emplace(device, nh).first;
    }
    *h = it->second;
    *busy = &g_roc.busy[device];
    *load_s = secs(t0, clk::now());
    return HFX_OK;
}

#define GCHK(x)                                                                                      \
    do {                                                                                             \
        hipError_t err_ = (x);                                                                       \
        if (err_ != hipSuccess)                                                                      \
            return set_error(HFX_ERR_DEVICE, std::string("device_linalg: ") + #x + ": " + hipGetErrorString(err_)); \
    } while (0)
#define RCHK(x)                                                                                      \
    do {                                                                                             \
        const rocblas_status st_ = (x);                                                              \
        if (st_ != rocblas_status_success)                                                           \
            return set_error(HFX_ERR_DEVICE, std::string("device_linalg: ") + #x + ": status " + std::to_string((int)st_)); \
    } while (0)

// ---------------------------------------------------------------------------------------
// Kernels.  Matrices are n x n column-major; nn = n * n.
// ---------------------------------------------------------------------------------------
constexpr int BLK = 256;
constexpr int SUM_BLOCKS = 128;  // partial sums of the per-iteration reduction
constexpr int DIIS_MAX = 15;     // ring slots the one-thread solve handles (system of 16)

struct Grid {
    dim3 g, b;
};
Grid grid_for(long long n) { return {dim3((unsigned)std::max<long long>(1, std::min<long long>(1024, (n + BLK - 1) / BLK))), dim3(BLK)}; }

#define GRID_STRIDE(i, n) \
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

// out = a + sb * b + sc * c (b, c may be null); out may alias a
__global__ void k_axpby(double* out, const double* a, const double* b, const double* c, double sb, double sc,
                        long long n) {
    GRID_STRIDE(i, n) {
        double v = a[i];
        if (b) v += sb * b[i];
        if (c) v += sc * c[i];
        out[i] = v;
    }
}

// out = ax * x + ay * y
__global__ void k_lin2(double* out, const double* x, const double* y, double ax, double ay, long long n) {
    GRID_STRIDE(i, n) out[i] = ax * x[i] + ay * y[i];
}

// e = A - A^T (+ add); e must not alias A
__global__ void k_antisym(double* e, const double* A, const double* add, int n) {
    const long long nn = (long long)n * n;
    GRID_STRIDE(i, nn) {
        const long long r = i % n, c = i / n;
        double v = A[i] - A[c + r * n];
        if (add) v += add[i];
        e[i] = v;
    }
}

// columns c < c0 set to zero, the others scaled by f (the level shift on the virtuals)
__global__ void k_shift_cols(double* A, int n, int c0, double f) {
    const long long nn = (long long)n * n;
    GRID_STRIDE(i, nn) A[i] = (i / n) < c0 ? 0.0 : A[i] * f;
}

// V(:, c) /= sqrt(w_c), zero for w_c <= 0 (inverseSqrtMatrix, src/Utils.cpp:117-138)
__global__ void k_rsqrt_cols(double* V, const double* w, int n) {
    const long long nn = (long long)n * n;
    GRID_STRIDE(i, nn) {
        const double wc = w[i / n];
        V[i] = wc > 0.0 ? V[i] / sqrt(wc) : 0.0;
    }
}

// HOMO/LUMO mixing of the UHF guess (SCF.cpp:165-231): columns h and h+1 of C
__global__ void k_mix_homo_lumo(double* C, int n, int h, double k, double f) {
    GRID_STRIDE(r, (long long)n) {
        const double homo = C[r + (long long)h * n], lumo = C[r + (long long)(h + 1) * n];
        C[r + (long long)h * n] = f * (homo + k * lumo);
        C[r + (long long)(h + 1) * n] = f * (lumo - k * homo);
    }
}

// FIFO position -> ring slot of the DIIS history
struct Slots {
    int s[DIIS_MAX];
};

// F = sum_k c_k F_ring[slot_k] (DIIS.cpp:44-48)
__global__ void k_extrapolate(double* F, const double* ring, long long nn, const double* c, int m, Slots sl) {
    GRID_STRIDE(i, nn) {
        double v = 0.0;
        for (int k = 0; k < m; ++k) v += c[k] * ring[sl.s[k] * nn + i];
        F[i] = v;
    }
}

// The bordered DIIS system (DIIS.cpp:52-72) over the m newest error vectors, built from their
// Gram matrix G (ring slots x ring slots), solved in the least-squares sense by Householder QR
// with column pivoting (Golub & Van Loan Alg. 5.4.1; the host driver's lstsq_qrcp), columns
// beyond the numerical rank set to zero.  One thread: the system is at most 16 x 16.
__global__ void k_diis_solve(const double* G, int nring, int m, Slots sl, double* c_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    constexpr int MX = DIIS_MAX + 1;
    const int M = m + 1;
    double B[MX * MX], b[MX], cn[MX], v[MX], y[MX], x[MX];
    int perm[MX];
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) B[i + j * M] = G[sl.s[i] + sl.s[j] * nring];
    for (int i = 0; i < M; ++i) {
        B[(M - 1) + i * M] = -1.0;
        B[i + (M - 1) * M] = -1.0;
        b[i] = 0.0;
    }
    B[(M - 1) + (M - 1) * M] = 0.0;
    b[M - 1] = -1.0;
    for (int j = 0; j < M; ++j) {
        perm[j] = j;
        double s = 0.0;
        for (int i = 0; i < M; ++i) s += B[i + j * M] * B[i + j * M];
        cn[j] = s;
    }
    double maxpiv = 0.0;
    int rank = M;
    const double thr = 2.220446049250313e-16 * M;  // machine epsilon x order, as the host solve
    for (int k = 0; k < M; ++k) {
        int piv = k;
        for (int j = k + 1; j < M; ++j)
            if (cn[j] > cn[piv]) piv = j;
        if (piv != k) {
            for (int i = 0; i < M; ++i) {
                const double t = B[i + k * M];
                B[i + k * M] = B[i + piv * M];
                B[i + piv * M] = t;
            }
            const double t = cn[k];
            cn[k] = cn[piv];
            cn[piv] = t;
            const int p = perm[k];
            perm[k] = perm[piv];
            perm[piv] = p;
        }
        double s = 0.0;
        for (int i = k; i < M; ++i) s += B[i + k * M] * B[i + k * M];
        const double nrm = sqrt(s);
        if (k == 0) maxpiv = nrm;
        if (nrm <= thr * maxpiv || nrm == 0.0) {
            rank = k;
            break;
        }
        const double x0 = B[k + k * M];
        const double alpha = (x0 >= 0.0) ? -nrm : nrm;
        v[0] = x0 - alpha;
        for (int i = k + 1; i < M; ++i) v[i - k] = B[i + k * M];
        double vtv = 0.0;
        for (int i = 0; i < M - k; ++i) vtv += v[i] * v[i];
        const double tau = (vtv > 0.0) ? 2.0 / vtv : 0.0;
        for (int j = k; j < M; ++j) {
            double dot = 0.0;
            for (int i = k; i < M; ++i) dot += v[i - k] * B[i + j * M];
            dot *= tau;
            for (int i = k; i < M; ++i) B[i + j * M] -= dot * v[i - k];
        }
        double dot = 0.0;
        for (int i = k; i < M; ++i) dot += v[i - k] * b[i];
        dot *= tau;
        for (int i = k; i < M; ++i) b[i] -= dot * v[i - k];
        for (int j = k + 1; j < M; ++j) {
            const double r = B[k + j * M];
            cn[j] = fmax(0.0, cn[j] - r * r);
        }
    }
    for (int k = 0; k < M; ++k) y[k] = 0.0;
    for (int k = rank - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < rank; ++j) s -= B[k + j * M] * y[j];
        y[k] = s / B[k + k * M];
    }
    for (int k = 0; k < M; ++k) x[perm[k]] = y[k];
    for (int k = 0; k < m; ++k) c_out[k] = x[k];
}

// Per-block partial sums of tr(Da (h + Fa)), tr(Db (h + Fb)), |Dt - Dt_prev|^2 and |e|^2 (the
// matrices are symmetric, so tr(D M) = sum D .* M); k_finish adds the partials in block order,
// so the scalars do not depend on scheduling.
__global__ __launch_bounds__(BLK) void k_sums(const double* Da, const double* Db, const double* h, const double* Fa,
                                              const double* Fb, const double* Dt, const double* Dtp,
                                              const double* err, long long nn, double* part) {
    __shared__ double sm[4][BLK];
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    GRID_STRIDE(i, nn) {
        s0 += Da[i] * (h[i] + Fa[i]);
        s1 += Db[i] * (h[i] + Fb[i]);
        const double d = Dt[i] - Dtp[i];
        s2 += d * d;
        if (err) s3 += err[i] * err[i];
    }
    sm[0][threadIdx.x] = s0;
    sm[1][threadIdx.x] = s1;
    sm[2][threadIdx.x] = s2;
    sm[3][threadIdx.x] = s3;
    __syncthreads();
    for (int o = BLK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 4; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[threadIdx.x * gridDim.x + blockIdx.x] = sm[threadIdx.x][0];
}

// out[0..3] the four sums; out[4..7] = eps_a[na-1], eps_a[na], eps_b[nb-1], eps_b[nb] (0 where
// the index is outside 0..n-1); out[8] = the eigensolvers' info codes (0: converged)
__global__ void k_finish(const double* part, int nparts, const double* wa, const double* wb, int n, int na, int nb,
                         const int* info, int n_info, double* out) {
    const int k = threadIdx.x;
    if (k < 4) {
        double s = 0.0;
        for (int b = 0; b < nparts; ++b) s += part[k * nparts + b];
        out[k] = s;
    } else if (k < 8) {
        const double* w = k < 6 ? wa : wb;
        const int idx = (k < 6 ? na : nb) - 1 + ((k - 4) & 1);
        out[k] = (idx >= 0 && idx < n) ? w[idx] : 0.0;
    } else if (k == 8) {
        int bad = 0;
        for (int i = 0; i < n_info; ++i) bad |= info[i] != 0;
        out[8] = bad;
    }
}

struct DevMem {
    std::vector<void*> ptrs;
    template <class T>
    T* get(size_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~DevMem() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

// ---------------------------------------------------------------------------------------
// The driver: the host driver's Scf, with matrices in HBM and one read-back per iteration.
// ---------------------------------------------------------------------------------------
struct DevScf {
    hfx_engine* e = nullptr;
    hfx_scf_options o{};
    int n = 0, na = 0, nb = 0;
    long long nn = 0;
    bool uhf = false;
    int precision = HFX_FP64;
    const double* Qh = nullptr;
    double e_nuc = 0, e_elec = 0, dE = 0, dD = 0, diis_err = 0, damp_c = 0;
    bool ls_a = false, ls_b = false;
    int iteration = 0;
    double gap_a = std::numeric_limits<double>::infinity(), gap_b = std::numeric_limits<double>::infinity();
    double t_jk = 0, t_diag = 0;
    int64_t n_builds = 0;

    rocblas_handle blas = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_c = nullptr;
    double* host_scal = nullptr;  // pinned, 16 doubles
    DevMem mem;
    double *S = nullptr, *X = nullptr, *h = nullptr;
    double *Fa = nullptr, *Fb = nullptr, *Fap = nullptr, *Fbp = nullptr, *Ca = nullptr, *Cb = nullptr;
    double *Da = nullptr, *Db = nullptr, *Dap = nullptr, *Dbp = nullptr, *Dt = nullptr, *Dtp = nullptr;
    double *dDt = nullptr, *dDa = nullptr, *dDb = nullptr, *J = nullptr, *Ka = nullptr, *Kb = nullptr;
    double *T1 = nullptr, *T2 = nullptr, *T3 = nullptr, *T4 = nullptr, *ea = nullptr;
    double *wa = nullptr, *wb = nullptr, *ework = nullptr, *part = nullptr, *scal = nullptr;
    int* info = nullptr;
    // DIIS history (DIIS.cpp): a ring of cap slots, count entries from slot head on
    double *ring_e = nullptr, *ring_fa = nullptr, *ring_fb = nullptr, *G = nullptr, *coef = nullptr;
    const double* err_new = nullptr;  // the newest error vector, for its norm
    int cap = 0, count = 0, head = 0;

    ~DevScf() {
        for (hipEvent_t ev : {ev_a, ev_b, ev_c})
            if (ev) (void)hipEventDestroy(ev);
        if (host_scal) (void)hipHostFree(host_scal);
        if (blas) (void)g_roc.set_stream(blas, nullptr);  // the handle outlives this stream
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }

    template <class K, class... A>
    void launch(K kern, long long work, A... args) {
        const Grid gr = grid_for(work);
        hipLaunchKernelGGL(kern, gr.g, gr.b, 0, s, args...);
    }

    int alloc(int diis_cap) {
        cap = diis_cap;
        const size_t m = (size_t)nn;
        double** mats[] = {&S, &X, &h, &Fa, &Fb, &Fap, &Fbp, &Ca, &Cb, &Da, &Db, &Dap, &Dbp, &Dt, &Dtp,
                           &dDt, &dDa, &dDb, &J, &Ka, &Kb, &T1, &T2, &T3, &T4, &ea};
        for (double** p : mats)
            if (!(*p = mem.get<double>(m))) return set_error(HFX_ERR_NOMEM, "device_linalg: hipMalloc failed");
        wa = mem.get<double>(n);
        wb = mem.get<double>(n);
        ework = mem.get<double>(n);
        part = mem.get<double>(4 * SUM_BLOCKS);
        scal = mem.get<double>(16);
        info = mem.get<int>(2);
        const int slots = std::max(1, cap);
        ring_e = mem.get<double>(m * slots);
        ring_fa = mem.get<double>(m * slots);
        ring_fb = mem.get<double>(m * slots);
        G = mem.get<double>((size_t)slots * slots);
        coef = mem.get<double>(DIIS_MAX + 1);
        if (!wa || !wb || !ework || !part || !scal || !info || !ring_e || !ring_fa || !ring_fb || !G || !coef)
            return set_error(HFX_ERR_NOMEM, "device_linalg: hipMalloc failed");
        GCHK(hipMemsetAsync(ring_e, 0, m * slots * sizeof(double), s));  // unused slots enter G as zeros
        GCHK(hipMemsetAsync(info, 0, 2 * sizeof(int), s));
        GCHK(hipHostMalloc((void**)&host_scal, 16 * sizeof(double), hipHostMallocDefault));
        return HFX_OK;
    }

    int copy(double* dst, const double* src) {
        GCHK(hipMemcpyAsync(dst, src, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
        return HFX_OK;
    }
    int zero(double* dst) {
        GCHK(hipMemsetAsync(dst, 0, nn * sizeof(double), s));
        return HFX_OK;
    }
    int gemm(bool ta, bool tb, int m, int nc, int k, double alpha, const double* A, int lda, const double* B, int ldb,
             double beta, double* C, int ldc) {
        RCHK(g_roc.dgemm(blas, ta ? rocblas_operation_transpose : rocblas_operation_none,
                         tb ? rocblas_operation_transpose : rocblas_operation_none, m, nc, k, &alpha, A, lda, B, ldb,
                         &beta, C, ldc));
        return HFX_OK;
    }
    int mm(const double* A, const double* B, double* C, bool ta = false, bool tb = false) {
        return gemm(ta, tb, n, n, n, 1.0, A, n, B, n, 0.0, C, n);
    }
    int syevd(double* A, double* w, int* inf) {  // eigenvalues ascending, eigenvectors over A
        RCHK(g_roc.dsyevd(blas, rocblas_evect_original, rocblas_fill_lower, n, A, n, w, ework, inf));
        return HFX_OK;
    }
    int density(const double* C, int nocc, double* D) {  // C_occ C_occ^T
        if (nocc <= 0) return zero(D);
        return gemm(false, true, n, n, nocc, 1.0, C, n, C, n, 0.0, D, n);
    }

    // X = S^-1/2 (inverseSqrtMatrix, src/Utils.cpp:117-138)
    int inv_sqrt() {
        int rc;
        if ((rc = copy(T1, S)) || (rc = syevd(T1, wa, info)) || (rc = copy(T2, T1))) return rc;  // T1 = V
        launch(k_rsqrt_cols, nn, T2, (const double*)wa, n);                                      // T2 = V w^-1/2
        return mm(T2, T1, X, false, true);
    }

    // X^T F X (+ level shift on the virtuals of Cold) -> eigenpairs (w, C = X V) (SCF.cpp:775-790).
    // C may be Cold: the shifted matrix is formed before C is written.
    int diag(const double* F, bool shift, const double* Cold, int nocc, double* C, double* w, int* inf) {
        int rc;
        if ((rc = mm(F, X, T1)) || (rc = mm(X, T1, T2, true, false))) return rc;  // T2 = X^T F X
        if (shift) {
            // F' += (X S C) diag(shift on virtuals) (X S C)^T  (SCF.cpp:781-786)
            if ((rc = mm(S, Cold, T1)) || (rc = mm(X, T1, T3)) || (rc = copy(T4, T3))) return rc;
            launch(k_shift_cols, nn, T4, n, nocc, o.level_shift);
            if ((rc = gemm(false, true, n, n, n, 1.0, T4, n, T3, n, 1.0, T2, n))) return rc;
        }
        if ((rc = syevd(T2, w, inf))) return rc;
        return mm(X, T2, C);
    }

    int apply_jk() {  // F = F_prev + J - K; F_prev = F (SCF.cpp:731-743)
        int rc;
        launch(k_axpby, nn, Fa, (const double*)Fap, (const double*)J, (const double*)Ka, 1.0, -1.0, nn);
        if ((rc = copy(Fap, Fa))) return rc;
        if (uhf) {
            launch(k_axpby, nn, Fb, (const double*)Fbp, (const double*)J, (const double*)Kb, 1.0, -1.0, nn);
            return copy(Fbp, Fb);
        }
        if ((rc = copy(Fb, Fa))) return rc;
        return copy(Fbp, Fap);
    }

    // The density increments (SCF.cpp:536-541; with full, from zero: F_prev = h, D_prev = 0),
    // the build on this stream, and the Fock update.
    int build_fock(bool full) {
        int rc;
        if (full) {
            if ((rc = zero(Dtp)) || (rc = zero(Dap)) || (rc = zero(Dbp)) || (rc = copy(Fap, h)) || (rc = copy(Fbp, h)))
                return rc;
        }
        launch(k_axpby, nn, dDt, (const double*)Dt, (const double*)Dtp, (const double*)nullptr, -1.0, 0.0, nn);
        launch(k_axpby, nn, dDa, (const double*)Da, (const double*)Dap, (const double*)nullptr, -1.0, 0.0, nn);
        if (uhf) launch(k_axpby, nn, dDb, (const double*)Db, (const double*)Dbp, (const double*)nullptr, -1.0, 0.0, nn);
        hfx_jk_args a{};
        a.dD_tot = dDt;
        a.dD_alpha = dDa;
        a.dD_beta = uhf ? dDb : nullptr;
        a.Q = Qh;
        a.schwarz_thresh = o.schwarz_thresh;
        a.density_thresh = o.density_thresh;
        a.precision = precision;
        a.fp32_cut = o.fp32_cut;
        a.flags = (o.jk_flags & ~(HFX_JK_CLASS_TIMING | HFX_JK_NO_STATS)) | HFX_JK_DEVICE_PTRS;
        GCHK(hipEventRecord(ev_a, s));
        if ((rc = hfx_jk_async(e, &a, J, Ka, uhf ? Kb : nullptr, (void*)s))) return rc;
        GCHK(hipEventRecord(ev_b, s));
        n_builds++;
        return apply_jk();
    }

    // DIIS::update and DIIS::extrapolate (DIIS.cpp:11-49) on e = (Fa Da S - S Da Fa) + (Fb Db S - S Db Fb)
    int diis_step() {
        int rc;
        const int slot = count < cap ? (head + count) % cap : head;
        double* es = ring_e + (size_t)slot * nn;
        if ((rc = mm(Fa, Da, T1)) || (rc = mm(T1, S, T2))) return rc;
        launch(k_antisym, nn, ea, (const double*)T2, (const double*)nullptr, n);
        if ((rc = mm(Fb, Db, T1)) || (rc = mm(T1, S, T2))) return rc;
        launch(k_antisym, nn, es, (const double*)T2, (const double*)ea, n);
        if ((rc = copy(ring_fa + (size_t)slot * nn, Fa)) || (rc = copy(ring_fb + (size_t)slot * nn, Fb))) return rc;
        if (count < cap)
            ++count;
        else
            head = (head + 1) % cap;
        err_new = es;
        // the Gram matrix of every ring slot (unused slots hold zeros and are not read)
        if ((rc = gemm(true, false, cap, cap, (int)nn, 1.0, ring_e, (int)nn, ring_e, (int)nn, 0.0, G, cap))) return rc;
        Slots sl{};
        for (int k = 0; k < count; ++k) sl.s[k] = (head + k) % cap;
        hipLaunchKernelGGL(k_diis_solve, dim3(1), dim3(64), 0, s, (const double*)G, cap, count, sl, coef);
        launch(k_extrapolate, nn, Fa, (const double*)ring_fa, nn, (const double*)coef, count, sl);
        launch(k_extrapolate, nn, Fb, (const double*)ring_fb, nn, (const double*)coef, count, sl);
        return HFX_OK;
    }
    void diis_reset() { count = head = 0; }

    // The reduction and its read-back: e_elec, |dD|, the DIIS error, the HOMO-LUMO gaps.
    int readback(bool with_err) {
        hipLaunchKernelGGL(k_sums, dim3(SUM_BLOCKS), dim3(BLK), 0, s, (const double*)Da, (const double*)Db,
                           (const double*)h, (const double*)Fa, (const double*)Fb, (const double*)Dt,
                           (const double*)Dtp, with_err ? err_new : (const double*)nullptr, nn, part);
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, s, (const double*)part, SUM_BLOCKS, (const double*)wa,
                           (const double*)wb, n, na, nb, (const int*)info, 2, scal);
        GCHK(hipMemcpyAsync(host_scal, scal, 9 * sizeof(double), hipMemcpyDeviceToHost, s));
        GCHK(hipEventRecord(ev_c, s));
        GCHK(hipEventSynchronize(ev_c));
        const double* r = host_scal;
        if (r[8] != 0.0) return set_error(HFX_ERR_DEVICE, "device_linalg: rocsolver_dsyevd did not converge");
        e_elec = 0.5 * r[0] + 0.5 * r[1];  // SCF.cpp:833-834
        dD = std::sqrt(r[2]);
        diis_err = with_err ? std::sqrt(r[3]) : 0.0;
        const double inf = std::numeric_limits<double>::infinity();
        gap_a = (na < n && na > 0) ? r[5] - r[4] : inf;
        gap_b = (nb < n && nb > 0) ? r[7] - r[6] : inf;
        return HFX_OK;
    }

    int initial_guess() {  // SCF.cpp:165-231
        int rc;
        if ((rc = copy(Fa, h)) || (rc = copy(Fap, h)) || (rc = diag(Fa, false, nullptr, na, Ca, wa, info)) ||
            (rc = zero(Dap)) || (rc = density(Ca, na, Da)) || (rc = zero(Dbp)))
            return rc;
        if (uhf) {
            if (o.guess_mix > 0 && na >= 1 && na < n) {
                const double k = o.guess_mix / 10.0;
                launch(k_mix_homo_lumo, n, Ca, n, na - 1, k, 1.0 / std::sqrt(1 + k * k));
                if ((rc = density(Ca, na, Da))) return rc;
            }
            if ((rc = copy(Fb, h)) || (rc = copy(Fbp, h)) || (rc = diag(Fb, false, nullptr, nb, Cb, wb, info + 1)) ||
                (rc = density(Cb, nb, Db)))
                return rc;
        } else {
            if ((rc = copy(Fb, Fa)) || (rc = copy(Fbp, Fap)) || (rc = copy(Cb, Ca)) || (rc = copy(Db, Da))) return rc;
            GCHK(hipMemcpyAsync(wb, wa, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        if ((rc = zero(Dtp))) return rc;
        launch(k_axpby, nn, Dt, (const double*)Da, (const double*)Db, (const double*)nullptr, 1.0, 0.0, nn);
        return readback(false);
    }

    int diagonalize_and_update() {  // SCF.cpp:747-835
        if (o.damp > 0 && o.max_damp_iter != 0 && o.max_damp_iter > iteration && (dE > o.stop_damp_thresh || iteration == 0))
            damp_c = o.damp / 100.0;
        else
            damp_c = 0.0;
        if (o.level_shift > 0 && iteration < o.max_lshift_iter && (dE > o.stop_lshift_thresh || iteration == 0)) {
            ls_a = gap_a < o.lshift_gap_tol;
            ls_b = gap_b < o.lshift_gap_tol;
        } else {
            ls_a = ls_b = false;
        }
        int rc;
        // D = (1 - c) C_occ C_occ^T + c D_old
        if ((rc = diag(Fa, ls_a, Ca, na, Ca, wa, info)) || (rc = copy(Dap, Da)) || (rc = density(Ca, na, T1))) return rc;
        launch(k_lin2, nn, Da, (const double*)T1, (const double*)Dap, 1.0 - damp_c, damp_c, nn);
        if ((rc = copy(Dbp, Db))) return rc;
        if (uhf) {
            if ((rc = diag(Fb, ls_b, Cb, nb, Cb, wb, info + 1)) || (rc = density(Cb, nb, T1))) return rc;
            launch(k_lin2, nn, Db, (const double*)T1, (const double*)Dbp, 1.0 - damp_c, damp_c, nn);
        } else {
            if ((rc = copy(Cb, Ca)) || (rc = copy(Db, Da))) return rc;
            GCHK(hipMemcpyAsync(wb, wa, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        if ((rc = copy(Dtp, Dt))) return rc;
        launch(k_axpby, nn, Dt, (const double*)Da, (const double*)Db, (const double*)nullptr, 1.0, 0.0, nn);
        return HFX_OK;
    }

    // One SCF cycle after the previous one's read-back (SCF.cpp:51-91): returns converged.
    int cycle(bool full, bool use_diis, hfx_scf_result* res, int slot, bool* conv) {
        const double last = e_elec;
        int rc;
        if ((rc = build_fock(full))) return rc;
        if (use_diis && (rc = diis_step())) return rc;
        if ((rc = diagonalize_and_update()) || (rc = readback(use_diis))) return rc;
        float ms_jk = 0.f, ms_all = 0.f;
        GCHK(hipEventElapsedTime(&ms_jk, ev_a, ev_b));
        GCHK(hipEventElapsedTime(&ms_all, ev_a, ev_c));
        t_jk += 1e-3 * ms_jk;
        t_diag += 1e-3 * (ms_all - ms_jk);
        dE = std::abs(e_elec - last);
        if (res && res->trace && slot < res->trace_cap) {
            res->trace[4 * slot + 0] = e_elec;
            res->trace[4 * slot + 1] = dE;
            res->trace[4 * slot + 2] = dD;
            res->trace[4 * slot + 3] = diis_err;
        }
        *conv = dE < o.energy_tol && dD < o.density_tol && (!use_diis || diis_err < o.diis_error_tol);
        return HFX_OK;
    }

    // One stage of cycles (the host driver's Scf::cycles): with switch_dE > 0 the stage also ends,
    // not converged, once |dE| < switch_dE after its first iteration.
    int cycles(bool use_diis, bool first_full, int& it_done, bool& conv, hfx_scf_result* res, int trace_base,
               double switch_dE) {
        conv = false;
        for (iteration = 0; iteration < o.max_iter; ++iteration) {
            bool c = false;
            const int rc = cycle(o.full_builds || (first_full && iteration == 0), use_diis, res, trace_base + iteration, &c);
            if (rc) return rc;
            if (c) {
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

    int write_result(hfx_scf_result* res) {  // the host driver's write_result, from HBM
        res->e_electronic = e_elec;
        res->e_nuclear = e_nuc;
        res->e_total = e_elec + e_nuc;
        res->last_dE = dE;
        res->last_dD = dD;
        res->last_diis_err = diis_err;
        res->time_jk_s = t_jk;
        res->time_diag_s = t_diag;
        res->n_jk_builds = n_builds;
        res->n_quartets_last = 0;  // the stream-ordered builds return no statistics
        const std::pair<double*, const double*> mats[] = {{res->F_alpha, Fa}, {res->D_alpha, Da}, {res->C_alpha, Ca},
                                                          {res->F_beta, Fb},  {res->D_beta, Db},  {res->C_beta, Cb},
                                                          {res->h_core, h},   {res->S, S}};
        for (const auto& m : mats)
            if (m.first) GCHK(hipMemcpyAsync(m.first, m.second, nn * sizeof(double), hipMemcpyDeviceToHost, s));
        std::vector<double> ea_h(n), eb_h(n);
        GCHK(hipMemcpyAsync(ea_h.data(), wa, n * sizeof(double), hipMemcpyDeviceToHost, s));
        GCHK(hipMemcpyAsync(eb_h.data(), wb, n * sizeof(double), hipMemcpyDeviceToHost, s));
        GCHK(hipStreamSynchronize(s));
        // remove the level shift from the virtual orbital energies (SCF.cpp:93-107)
        if (ls_a)
            for (int i = na; i < n; ++i) ea_h[i] -= o.level_shift;
        if (ls_b)
            for (int i = nb; i < n; ++i) eb_h[i] -= o.level_shift;
        if (res->eps_alpha) std::memcpy(res->eps_alpha, ea_h.data(), n * sizeof(double));
        if (res->eps_beta) std::memcpy(res->eps_beta, eb_h.data(), n * sizeof(double));
        return HFX_OK;
    }
};

}  // namespace

namespace hfx_internal {

int device_cycles(hfx_engine* e, const hfx_scf_options& o, const ScfSetup& su, hfx_scf_result* res) {
    const auto t0 = clk::now();
    if (o.use_diis && (o.diis_size < 1 || o.diis_size > DIIS_MAX))
        return set_error(HFX_ERR_INVALID, "device_linalg: diis_size must be 1.." + std::to_string(DIIS_MAX));
    int dev = 0;
    {
        const int rc = hfx_internal::engine_device(e, &dev);
        if (rc) return rc;
    }
    GCHK(hipSetDevice(dev));
    rocblas_handle blas = nullptr;
    std::mutex* busy = nullptr;
    double load_s = 0.0;
    {
        const int rc = roc_for_device(dev, &blas, &busy, &load_s);
        if (rc) return rc;
    }
    std::lock_guard<std::mutex> run_lock(*busy);
    DevScf d;
    d.e = e;
    d.o = o;
    d.n = su.n;
    d.nn = (long long)su.n * su.n;
    d.na = su.na;
    d.nb = su.nb;
    d.uhf = su.uhf;
    d.precision = su.precision;
    d.e_nuc = su.e_nuc;
    d.Qh = su.Q;
    d.blas = blas;
    GCHK(hipStreamCreateWithFlags(&d.s, hipStreamNonBlocking));
    GCHK(hipEventCreate(&d.ev_a));
    GCHK(hipEventCreate(&d.ev_b));
    GCHK(hipEventCreate(&d.ev_c));
    RCHK(g_roc.set_stream(blas, d.s));
    int rc = d.alloc(o.use_diis ? o.diis_size : 0);
    if (rc) return rc;
    GCHK(hipMemcpyAsync(d.S, su.S, d.nn * sizeof(double), hipMemcpyHostToDevice, d.s));
    GCHK(hipMemcpyAsync(d.h, su.h, d.nn * sizeof(double), hipMemcpyHostToDevice, d.s));
    if ((rc = d.inv_sqrt()) || (rc = d.initial_guess())) return rc;
    d.dD = 0.0;
    d.diis_err = 0.0;
    const auto t_init = clk::now();

    int its = 0;
    bool conv = false;
    const bool final64 = d.precision == HFX_MIXED && o.final_fp64;
    const bool early = final64 && o.switch_tol > 0.0;
    if ((rc = d.cycles(o.use_diis != 0, false, its, conv, res, 0, early ? o.switch_tol : 0.0))) return rc;
    int total = its;
    res->e_total_mixed = d.e_elec + d.e_nuc;
    res->iterations_mixed = its;
    if (final64) {
        // a full FP64 build of the current density, then FP64 to convergence; after an early
        // switch the DIIS history of the mixed builds is kept, otherwise DIIS starts afresh
        d.precision = HFX_FP64;
        if (!early) d.diis_reset();
        int its2 = 0;
        if ((rc = d.cycles(o.use_diis != 0, true, its2, conv, res, its, 0.0))) return rc;
        total += its2;
    }
    if ((rc = d.write_result(res))) return rc;
    res->iterations = total;
    res->converged = conv ? 1 : 0;
    res->time_init_s = secs(t0, t_init);
    res->time_linalg_load_s = load_s;
    res->n_mo = su.n;
    return HFX_OK;
}

}  // namespace hfx_internal
