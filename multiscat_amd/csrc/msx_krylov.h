// Batched GMRES state and the launchers of msx_krylov.hip.
#pragma once

#include "msx_internal.h"

namespace msx {

constexpr int KR_MAX_BASIS = 2048;  // longest Krylov cycle the kernels accept

// Where each instance's vector lives inside a concatenated state vector, and how the
// vector kernels split it: grid = (nchunk, n_inst), chunk elements per block.
struct KLayout {
    int n_inst;
    int kmax;               // restart length (Hessenberg is (kmax+1) x (kmax+1) per instance)
    int nchunk;             // blocks per instance
    long long chunk;        // elements per block
    const long long *start; // [n_inst] first element of each instance
    const long long *len;   // [n_inst] element count of each instance
    int *active;            // [n_inst] 1 while the instance iterates in this cycle
};

// Per-instance GMRES recurrence state (device pointers).
struct KState {
    cx<double> *H;   // [n_inst][(kmax+1)^2] column-major
    cx<double> *g;   // [n_inst][kmax+1]
    cx<double> *cs;  // [n_inst][kmax+1]
    cx<double> *sn;  // [n_inst][kmax+1]
    cx<double> *y;   // [n_inst][kmax+1]
    double *beta, *beta0, *resid, *hnorm;
    int *kconv, *iters, *cycle_iters, *converged;
    double eps;
    int window;
    const double *eps_inst = nullptr;  // [n_inst] per-instance tolerance, overrides eps if set
    // Arnoldi on F with A = hshift I + hscale F (the Krylov spaces of A and F are the same):
    // each Hessenberg column of F enters as hshift e_k + hscale h (1 and 0: F is A itself)
    double hscale = 1.0, hshift = 0.0;
};

// dst[b] = src[b * stride].re for b < n (dst may be pinned host memory mapped for the device)
void k_gather_re(int n, int stride, const cx<double> *src, double *dst, hipStream_t s);

// Chunk partials of V^H w (and |w|^2 with with_norm) in partial, reduced into h; with h null
// the partials are left for a consumer that sums them itself (pin of k_axpy, k_axpy_dots).
template <typename T>
void k_dots(const KLayout &K, const cx<T> *const *vptr, int kcount, const cx<T> *w, int with_norm,
            double *partial, cx<double> *h, int accumulate, hipStream_t s);
// with_norm: chunk partials of |w|^2 (after the update) in slot kcount, reduced into
// hnorm_dst[kcount] if hnorm_dst is given. pin (kcount <= 16): the coefficients are the sums
// of these chunk partials (slots 0 .. kcount-1), formed by every workgroup, and also stored to
// hout; h is then unused. partial may be pin's buffer (the norm goes to slot kcount).
template <typename T>
void k_axpy(const KLayout &K, const cx<T> *const *vptr, int kcount, const cx<double> *h, cx<T> *w,
            int with_norm, double *partial, cx<double> *hnorm_dst, hipStream_t s,
            const double *pin = nullptr, cx<double> *hout = nullptr);
// w -= sum_{j<kcount} v_j h_j, then h2 = V^H w, in one pass over the basis (kcount <= 16;
// false, with nothing launched, above that). pin / hout as for k_axpy (pin must not be
// partial: this pass writes its own partials there); h2 null leaves the new partials unreduced.
// with_norm: |w|^2 partials of the updated w in slot kcount too.
template <typename T>
bool k_axpy_dots(const KLayout &K, const cx<T> *const *vptr, int kcount, const cx<double> *h,
                 cx<T> *w, double *partial, cx<double> *h2, hipStream_t s,
                 const double *pin = nullptr, cx<double> *hout = nullptr, int with_norm = 0);
// From k_axpy_dots' partials (with_norm): mask[b] = 1 where an iterating instance's second
// projection exceeds tau |w| (it takes the second update), else 0 with h2 stored as zero.
void k_reorth_select(const KLayout &K, int kcount, const double *partial, int *mask,
                     cx<double> *h2, double tau, hipStream_t s);
template <typename T, typename TX>
void k_update(const KLayout &K, const cx<T> *const *vptr, const int *kk, const cx<double> *y,
              cx<TX> *x, const int *mask, hipStream_t s);
// k_update for a basis in node-major order (L.nm) and x in the channel-major batch order;
// kmax: the Krylov dimension that strides y.
template <typename T>
void k_update_nm(const Layout &L, int kmax, const cx<T> *const *vptr, const int *kk,
                 const cx<double> *y, cx<double> *x, hipStream_t s);
template <typename T>
void k_scale(const KLayout &K, const cx<T> *src, cx<T> *dst, const double *sv, const int *mask,
             hipStream_t s);
template <typename T>
void k_axpby(const KLayout &K, const cx<T> *a, const cx<T> *c, T alpha, cx<T> *dst,
             const int *mask, hipStream_t s);
// norm_partial (optional): the chunk partials of |w|^2 in slot k + 1 of a partial buffer
// (k_axpy with_norm and no hnorm_dst), summed here instead of read from hcol[k + 1].
void k_givens(const KLayout &K, int k, const cx<double> *hcol, const cx<double> *hcol2,
              const KState &S, int *count, hipStream_t s,
              int *host_flags = nullptr, const double *norm_partial = nullptr);
void k_backsolve(const KLayout &K, const KState &S, const int *mask, hipStream_t s);
// Starts a cycle from the residual norms nrm2; on the first cycle the convergence reference
// beta0 is the residual norm, or sqrt(ref_nrm2) when given (an initial guess: then an
// instance whose residual is already below eps * beta0 stops at once).
void k_cycle_start(const KLayout &K, const KState &S, const cx<double> *nrm2, int first,
                   int *count, hipStream_t s, const cx<double> *ref_nrm2 = nullptr,
                   int *host_flags = nullptr);
void k_reset(const KLayout &K, const KState &S, int *enabled, hipStream_t s);
void k_reactivate(const KLayout &K, const KState &S, const int *enabled, hipStream_t s);

}  // namespace msx
