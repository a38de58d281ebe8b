// Internal declarations of libmultiscat_amd: complex arithmetic, device buffers, the
// problem layout shared by every kernel, and the kernel launchers.
//
// Layout in HBM (all instances of a batch concatenated, "compact"):
//   state vectors   x[off_b*nz + j*nz + k]   condition b, channel j (solve order), node k
//   channel tables  chx/chy/e/c/fac          per channel, concatenated in the same order
//   potential table V[g*nz + k]              g = ix*nky + iy (storage order), node k
//   eigensystem     Q[z*nz + m], QT[m*nz + z], lam[m]  (H0 = Q diag(lam) Q^T, ascending)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace msx {

template <typename T>
struct cx {
    T re, im;
};

template <typename T>
__host__ __device__ inline cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T>
__host__ __device__ inline cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__host__ __device__ inline cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__host__ __device__ inline cx<T> operator*(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ inline cx<T> operator*(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__host__ __device__ inline cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
template <typename T>
__host__ __device__ inline T norm2(cx<T> a) { return a.re * a.re + a.im * a.im; }
template <typename T>
__host__ __device__ inline cx<T> cdiv(cx<T> a, cx<T> b) {
    // Smith's algorithm keeps the quotient finite for tiny/huge |b|.
    if (fabs(b.re) >= fabs(b.im)) {
        T r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    T r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
// acc += a * b
template <typename T>
__host__ __device__ inline void cfma(cx<T> &acc, cx<T> a, cx<T> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}
// acc += conj(a) * b
template <typename T>
__host__ __device__ inline void cfma_conj(cx<T> &acc, cx<T> a, cx<T> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(-a.im, b.re, acc.im);
}

// ---- error handling -------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define MSX_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            int _code = (_e == hipErrorOutOfMemory) ? 7 : 6;                               \
            throw ::msx::Error(_code, std::string(#expr) + ": " + hipGetErrorString(_e) +  \
                                          " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
        }                                                                                  \
    } while (0)

#define MSX_LAUNCH_CHECK() MSX_HIP(hipGetLastError())

// Raise a kernel's dynamic-LDS limit to `bytes` (needed above 64 KiB). The limit covers the
// dynamic part only; static __shared__ arrays of the kernel count on top of it.
inline void set_dynamic_lds(const void *kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // do not leave a sticky error for the next launch check
        throw Error(6, "kernel needs " + std::to_string(bytes) +
                           " bytes of LDS, more than the device grants: " + hipGetErrorString(e));
    }
}

// ---- device buffers -------------------------------------------------------------------
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error(e == hipErrorOutOfMemory ? 7 : 6,
                        "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                            hipGetErrorString(e));
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t s = 0) {
        MSX_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *h, size_t count, hipStream_t s = 0) const {
        MSX_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// ---- problem descriptor shared by kernels ----------------------------------------------
// Per-instance metadata lives in device arrays; kernels index them by blockIdx.
struct Layout {
    int nkx, nky, nz;
    int n_inst;          // conditions in the batch
    int max_nc;          // largest channel count
    const int *nc;       // [n_inst]
    const long long *off;  // [n_inst] channel offset of each instance
    const int *chx, *chy;  // [total channels]
    long long total = 0;   // total channels of the batch (0: not given, at most n_inst * max_nc)
    // Order of the state vectors inside each instance's block [off*nz, (off+nc)*nz):
    // 0 channel-major (element (j, z) at (off + j)*nz + z, the batch layout above), 1
    // node-major (at off*nz + z*nc + j; the FP32 inner vectors of the mixed solver, so that
    // the coupling reads and writes one node's channels as whole cache lines).
    int nm = 0;
};

// Position of element (channel j, node z) of the instance at channel offset off with nc
// channels, in either vector order.
__host__ __device__ inline long long vec_index(const Layout &L, long long off, int nc, int j, int z) {
    return L.nm ? off * L.nz + (long long)z * nc + j : (off + j) * L.nz + z;
}

// ---- kernel launchers (msx_kernels.hip) ------------------------------------------------
// Coupling modes of the per-z channel sum y_j = sum_i V[G_j - G_i] x_i.
enum CouplingMode { COUPLE_UPPER = 0, COUPLE_FULL = 1, COUPLE_LOWER = 2 };

template <typename T>
void launch_coupling(const Layout &L, const cx<T> *V, const cx<T> *x, cx<T> *y, int mode,
                     const int *active, hipStream_t s);

// x <- L^-1 x in place for every active instance (sequential over channels, one workgroup per
// instance). If `out` is non-null, also writes out = base + sign * (L^-1 x).
void launch_lower_solve(const Layout &L, const cx<double> *V, const double *Q,
                        const double *QT, const double *lam, const double *e,
                        const cx<double> *c, const double *fac, cx<double> *x,
                        const cx<double> *base, cx<double> *out, double sign,
                        const int *active, hipStream_t s);

// y = D^-1 x for all channels of active instances (eigenbasis GEMMs + Sherman-Morrison). If
// `base` is non-null: y = base + sign * D^-1 x.
// lam and e stay FP64 in every precision so that e_j + lam_m never cancels in FP32.
template <typename T>
void launch_dinv(const Layout &L, const T *Q, const T *QT, const double *lam, const double *e,
                 const cx<T> *c, const T *fac, const cx<T> *x, const cx<T> *base, cx<T> *y,
                 T sign, const int *active, hipStream_t s);

// Eigensystem zero-padded to nzp (a multiple of 16) for the matrix-core kernels.
template <typename T>
struct EigTables {
    int nzp;
    const T *Qp, *QTp;   // [nzp][nzp]
    const double *lam;   // [nz]
};

// Split-FP16 planes of the eigenvectors for the FP32 inner operator's D^-1 (msx_diag_h3.hip):
// Q^T and Q (columns reordered for the second product), each as FP16 hi and lo, nq x (nq + 8)
// halves apiece, and Q[nz-1][m] in FP32 (the Sherman-Morrison row).
struct H3Planes {
    const _Float16 *planes = nullptr;  // [4][nq][nq + 8]: P1 hi, P1 lo, P2 hi, P2 lo
    int nq = 0;                        // 32 ceil(nz / 32), 0 when nz > 224
    const float *qlast = nullptr;      // [nz]
    int p2_lds_rows = 0;  // planes-in-L2 kernel: leading rows of P2 hi held in LDS (its launcher)
};
int h3_padded_size(int nz);
size_t h3_plane_halves(int nq);
void launch_h3_planes(int nz, int nq, const double *Q, _Float16 *planes, hipStream_t s);
// Whether the split-FP16 kernel takes this layout (node-major, nz <= 160, LDS large enough).
bool diag_h3_fits(const Layout &L, const H3Planes &H);
// y = D^-1 x, or y = base + sign D^-1 x, node-major FP32 vectors and factors; cpw: channels
// per wavefront, 16 (or 0) or 8.
void launch_diag_h3(const Layout &L, const H3Planes &H, const double *lam, const double *e,
                    const cx<float> *c, const float *fac, const cx<float> *x,
                    const cx<float> *base, cx<float> *y, float sign, const int *active,
                    hipStream_t s, int cpw = 0);

// MFMA form of the diagonal blocks (inverse = 1: D^-1, inverse = 0: D), nz <= 320.
// channels_per_wave (FP64, nzp <= 160): 8 or 16 channels per wavefront, 0 = the built-in 16.
bool diag_mfma_supported(int nz);
template <typename T>
void launch_diag_blocks(const Layout &L, const EigTables<T> &E, const double *e, const cx<T> *c,
                        const T *fac, const cx<T> *x, const cx<T> *base, cx<T> *y, T sign,
                        int inverse, const int *active, hipStream_t s, int channels_per_wave = 0);
void launch_pad_square(int n, int np, const double *src, double *dst, hipStream_t s);
void launch_pad_square(int n, int np, const double *src, float *dst, hipStream_t s);

// y = D x (the diagonal blocks themselves: (H0 + e_j) x_j - c_j x_j(N) e_N), or
// y = base + D x when base is given (vector ALU, channel-major vectors).
void launch_apply_diag(const Layout &L, const double *Q, const double *QT, const double *lam,
                       const double *e, const cx<double> *c, const cx<double> *x,
                       cx<double> *y, const int *active, hipStream_t s,
                       const cx<double> *base = nullptr);

// Eigensystem of a symmetric n x n matrix (one workgroup, Householder + implicit QL).
// ticks (nullable, 8 entries): wall-clock stamps at the start, after the tridiagonalisation,
// after the accumulation of Q_H, after the QL iteration and at the end; then the wall-clock
// ticks lane 0 spent chasing bulges, the number of QL sweeps and of rotations.
void launch_eigensystem(int n, double *A, double *Vec, double *lam_sorted, double *Q_sorted,
                        double *QT_sorted, int *info, long long *ticks, hipStream_t s);

// The same eigensystem by Householder tridiagonalisation (in LDS once the active block fits)
// and divide and conquer (2 <= n <= 1024); A is overwritten, work holds
// eig_dc_workspace(n) doubles. ticks: [0] start, [1] = [2] end of the Householder stage,
// [3] start and [4] end of the back-transformation ([5..7] zero).
bool eig_dc_supported(int n);
size_t eig_dc_workspace(int n);
void launch_eigensystem_dc(int n, double *A, double *work, double *lam_sorted,
                           double *Q_sorted, double *QT_sorted, int *info, long long *ticks,
                           hipStream_t s);

// The same factors on the matrix cores (FP64) when nzp <= 320; false if the size does not fit.
bool launch_factors_mfma(long long total_channels, int nz, const EigTables<double> &E,
                         const double *e, double *fac, hipStream_t s);

// fac[j][z] = sum_m Q[z][m] Q[N-1][m] / (e_j + lam_m) for all channels.
void launch_preconditioner_factors_qt(long long total_channels, int nz, const double *Q,
                                      const double *QT, const double *lam, const double *e,
                                      double *fac, hipStream_t s);

// A = T + diag(Re V_00) (row-major n x n) from the device table and kinetic matrix.
void launch_build_h0(int nz, const double *T, const cx<double> *V, double *A, hipStream_t s);

// Vt[g] = V[-g] (the table of the transposed coupling).
void launch_flip_table(int nkx, int nky, int nz, const cx<double> *V, cx<double> *Vt,
                       hipStream_t s);

// Source vector: zero everywhere except s[spec_b][N-1] = bsrc_b.
void launch_set_source(const Layout &L, const int *spec, const cx<double> *bsrc, cx<double> *s,
                       hipStream_t st);
// d = D^-1 s for that source without touching the other channels' blocks (Sherman-Morrison
// with the channel-major preconditioner factors; d is zero outside the specular channels), in
// the vector order of L; d32 (optional) receives the same vector in FP32.
void launch_specular_dinv(const Layout &L, const int *spec, const cx<double> *bsrc,
                          const cx<double> *c, const double *fac, cx<double> *d, hipStream_t st,
                          cx<float> *d32 = nullptr);
// FP64 state from channel-major to node-major order (to_nm) or back, src != dst.
void launch_state_order(const Layout &L, const cx<double> *src, cx<double> *dst, bool to_nm,
                        hipStream_t s);

// Full coupling of a state nonzero only on the specular channels (y_j = V[G_j - G_s] x_s).
template <typename T>
void launch_specular_coupling(const Layout &L, const cx<T> *V, const int *spec, const cx<T> *x,
                              cx<T> *y, const int *active, hipStream_t s);

// Adjoint source g_j(N-1) = 4i dLdP_j b_j conj(S_j), S_j = 2i b_j psi_j(N-1) + a0 delta_spec.
void launch_adjoint_source(const Layout &L, const int *spec, const cx<double> *psi,
                           const cx<double> *b, const cx<double> *a0, const double *dLdP,
                           cx<double> *g, hipStream_t st);

// Accumulate a (FP32) correction into an FP64 vector: x += d.
void launch_add_float(long long n, const cx<float> *d, cx<double> *x, hipStream_t s);
// dst[i] += src[i] for i < n.
void launch_add_ints(int n, const int *src, int *dst, hipStream_t s);

// Precision conversions of channel-major state vectors.
void launch_to_float(long long n, const cx<double> *src, cx<float> *dst, hipStream_t s);
// rho = s - D x - V x (FP64 sums, FP32 result in the vector order of Lo): the residual of the
// scattering system before the block-Jacobi preconditioner, with D_j x_j from H0p (the
// symmetrised H0 zero-padded to nzp, FP64 matrix cores) and V x given (msx_diag.hip); s the
// scattering source (src null) or the vector src in the order of L.
bool split_residual_supported(int nzp);
void launch_split_residual(const Layout &L, const Layout &Lo, const double *H0p, int nzp,
                           const double *e, const cx<double> *c, const int *spec,
                           const cx<double> *bsrc, const cx<double> *x, const cx<double> *vx,
                           cx<float> *rho, const int *active, hipStream_t s,
                           const cx<double> *src = nullptr);
// FP64 channel-major -> FP32 in the vector order of L (a tiled transpose when L.nm); n is the
// element count of the batch.
void launch_to_float_vec(const Layout &L, long long n, const cx<double> *src, cx<float> *dst,
                         hipStream_t s);
// FP64 channel-major real table (one value per channel and node) -> FP32 node-major.
void launch_to_float_real_nm(const Layout &L, const double *src, float *dst, hipStream_t s);
void launch_to_float_real(long long n, const double *src, float *dst, hipStream_t s);

// P_j = |2i b_j psi_j(N) + a0 delta_spec|^2 (psi in the vector order of L)
void launch_intensities(const Layout &L, const cx<double> *psi, const cx<double> *b,
                        const cx<double> *a0, const int *spec, double *P, hipStream_t s);

// Gradient of the coupling: grad[g][z] += conj(-sum_{j,i: G_j-G_i = g, i != j} lam_j psi_i)
// and grad[0][z] += Re(-sum_j lam_j psi_j) (real part only). chmap[b][g]: channel of
// instance b at grid point g, -1 if none.
void launch_potential_gradient_map(const Layout &L, const int *chmap, const cx<double> *lam_adj,
                                   const cx<double> *psi, cx<double> *grad, hipStream_t s);

}  // namespace msx
