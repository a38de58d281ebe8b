// Symmetric eigensystem H0 = Q diag(lam) Q^T (ascending) on the GPU in FP64, one workgroup.
//
// The reference diagonalises H0 = T + diag(Re V_00) once per solve with an EISPACK-style
// Householder + QL solver (lib/multiscat_fortran/fortran/diagsub.f90:5-284, called from
// multiscat_gmres.f90:251-253). Any exact symmetric eigensolver gives the same eigenpairs
// (up to eigenvector signs); this one is the textbook pair
//   (1) Householder tridiagonalisation A = Q_H T Q_H^T (Golub & Van Loan, Matrix
//       Computations, sec. 8.3.1), every reflector applied as a parallel rank-2 update;
//   (2) implicit QL iteration with shifts on the tridiagonal T (Bowdler, Martin, Reinsch and
//       Wilkinson, Numer. Math. 11 (1968) 293), where one lane runs the scalar bulge chase of a
//       sweep and records its rotations, and then every lane applies the whole rotation
//       sequence to its own column of Z (stored transposed, so the updates are coalesced).
// The eigenvalues are then ranked ascending and the eigenvectors permuted accordingly.
#include <cfloat>

#include "msx_internal.h"

namespace msx {

constexpr int EIG_THREADS = 1024;
constexpr int EIG_GROUPS = 4;  // row groups of the matrix-vector products

__device__ double eig_block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    return t;
}

// A (n x n row-major, symmetric) is destroyed; W is an n x n workspace.
__global__ __launch_bounds__(EIG_THREADS) void eig_tridiag_ql_kernel(int n, double *A, double *W,
                                                                     double *lam_sorted,
                                                                     double *Q_sorted,
                                                                     double *QT_sorted,
                                                                     int *info) {
    extern __shared__ double es[];
    double *v = es;                      // [n]
    double *w = v + n;                   // [n]
    double *pp = w + n;                  // [EIG_GROUPS][n]
    double *d = pp + EIG_GROUPS * n;     // [n]
    double *e = d + n;                   // [n]
    double *cs = e + n;                  // [n]
    double *sn = cs + n;                 // [n]
    double *bet = sn + n;                // [n]
    __shared__ double red[EIG_THREADS / 64];
    __shared__ int s_hi, s_stop, s_done, s_fail;
    const int tid = threadIdx.x, NT = blockDim.x;
    const int grp = tid / (NT / EIG_GROUPS), gl = tid % (NT / EIG_GROUPS), GS = NT / EIG_GROUPS;

    // ---- (1) Householder tridiagonalisation ------------------------------------------
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1, r0 = k + 1;
        double *rowk = A + (long long)k * n + r0;  // x = A[k][k+1 ..]
        double loc = 0;
        for (int i = tid; i < m; i += NT) loc += rowk[i] * rowk[i];
        const double sigma = eig_block_sum(loc, red);
        const double x0 = rowk[0];
        if (!(sigma - x0 * x0 > 0.0)) {  // already tridiagonal in this column
            if (tid == 0) {
                e[k] = x0;
                bet[k] = 0.0;
            }
            __syncthreads();
            continue;
        }
        const double alpha = x0 > 0 ? -sqrt(sigma) : sqrt(sigma);
        const double beta = 1.0 / (sigma - x0 * alpha);
        for (int i = tid; i < m; i += NT) {
            const double vi = i == 0 ? x0 - alpha : rowk[i];
            v[i] = vi;
        }
        __syncthreads();
        // p = beta A22 v (A22 symmetric: p_j = sum_i A[r0+i][r0+j] v_i, coalesced in j)
        {
            const int ib = (m * grp) / EIG_GROUPS, ie = (m * (grp + 1)) / EIG_GROUPS;
            for (int j = gl; j < m; j += GS) {
                double acc = 0;
                for (int i = ib; i < ie; ++i) acc = fma(A[(long long)(r0 + i) * n + r0 + j], v[i], acc);
                pp[grp * n + j] = acc;
            }
        }
        __syncthreads();
        double kl = 0;
        for (int j = tid; j < m; j += NT) {
            double pj = 0;
            for (int g = 0; g < EIG_GROUPS; ++g) pj += pp[g * n + j];
            pj *= beta;
            w[j] = pj;
            kl += v[j] * pj;
        }
        const double K = 0.5 * beta * eig_block_sum(kl, red);
        for (int j = tid; j < m; j += NT) w[j] -= K * v[j];
        __syncthreads();
        // A22 -= v w^T + w v^T
        for (long long idx = tid; idx < (long long)m * m; idx += NT) {
            const int i = (int)(idx / m), j = (int)(idx - (long long)i * m);
            A[(long long)(r0 + i) * n + r0 + j] -= v[i] * w[j] + w[i] * v[j];
        }
        // keep v in row k (no longer read) for the accumulation of Q_H
        for (int i = tid; i < m; i += NT) rowk[i] = v[i];
        if (tid == 0) {
            e[k] = alpha;
            bet[k] = beta;
        }
        __syncthreads();
    }
    for (int k = tid; k < n; k += NT) d[k] = A[(long long)k * n + k];
    if (tid == 0) {
        if (n >= 2) {
            e[n - 2] = A[(long long)(n - 2) * n + n - 1];
            bet[n - 2] = 0.0;
        }
        e[n - 1] = 0.0;
        bet[n - 1] = 0.0;
        s_fail = 0;
    }
    __syncthreads();

    // ---- Q_H = H_0 H_1 ... H_{n-3} (row-major in W), accumulated backwards -----------------
    for (long long idx = tid; idx < (long long)n * n; idx += NT) {
        const int i = (int)(idx / n), j = (int)(idx - (long long)i * n);
        W[idx] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int k = n - 3; k >= 0; --k) {
        const double beta = bet[k];
        if (beta == 0.0) continue;
        const int m = n - k - 1, r0 = k + 1;
        for (int i = tid; i < m; i += NT) v[i] = A[(long long)k * n + r0 + i];
        __syncthreads();
        {
            const int ib = (m * grp) / EIG_GROUPS, ie = (m * (grp + 1)) / EIG_GROUPS;
            for (int j = gl; j < m; j += GS) {
                double acc = 0;
                for (int i = ib; i < ie; ++i) acc = fma(v[i], W[(long long)(r0 + i) * n + r0 + j], acc);
                pp[grp * n + j] = acc;
            }
        }
        __syncthreads();
        for (int j = tid; j < m; j += NT) {
            double u = 0;
            for (int g = 0; g < EIG_GROUPS; ++g) u += pp[g * n + j];
            w[j] = beta * u;
        }
        __syncthreads();
        for (long long idx = tid; idx < (long long)m * m; idx += NT) {
            const int i = (int)(idx / m), j = (int)(idx - (long long)i * m);
            W[(long long)(r0 + i) * n + r0 + j] -= v[i] * w[j];
        }
        __syncthreads();
    }
    // Z^T into A (A's contents are no longer needed)
    double *ZT = A;
    for (long long idx = tid; idx < (long long)n * n; idx += NT) {
        const int i = (int)(idx / n), j = (int)(idx - (long long)i * n);
        ZT[(long long)j * n + i] = W[idx];
    }
    __syncthreads();

    // ---- (2) implicit QL with shifts on (d, e), rotations applied to the rows of Z^T -----
    for (int l = 0; l < n; ++l) {
        for (int iter = 0;; ++iter) {
            if (tid == 0) {
                int m;
                for (m = l; m < n - 1; ++m) {
                    const double dd = fabs(d[m]) + fabs(d[m + 1]);
                    if (fabs(e[m]) <= DBL_EPSILON * dd) break;
                }
                if (m == l) {
                    s_done = 1;
                } else if (iter >= 60) {
                    s_done = 1;
                    s_fail = 1;
                } else {
                    s_done = 0;
                    double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                    double r = hypot(g, 1.0);
                    g = d[m] - d[l] + e[l] / (g + copysign(r, g));
                    double s = 1.0, c = 1.0, p = 0.0;
                    int i;
                    bool underflow = false;
                    for (i = m - 1; i >= l; --i) {
                        const double f = s * e[i], b = c * e[i];
                        r = hypot(f, g);
                        e[i + 1] = r;
                        if (r == 0.0) {
                            d[i + 1] -= p;
                            e[m] = 0.0;
                            underflow = true;
                            break;
                        }
                        s = f / r;
                        c = g / r;
                        g = d[i + 1] - p;
                        r = (d[i] - g) * s + 2.0 * c * b;
                        p = s * r;
                        d[i + 1] = g + p;
                        g = c * r - b;
                        cs[i] = c;
                        sn[i] = s;
                    }
                    s_hi = m - 1;
                    s_stop = underflow ? i + 1 : l;
                    if (!underflow) {
                        d[l] -= p;
                        e[l] = g;
                        e[m] = 0.0;
                    }
                }
            }
            __syncthreads();
            if (s_done) break;
            const int hi = s_hi, lo = s_stop;
            for (int k = tid; k < n; k += NT) {
                double zi1 = ZT[(long long)(hi + 1) * n + k];
                for (int i = hi; i >= lo; --i) {
                    const double c = cs[i], s = sn[i];
                    const double zi = ZT[(long long)i * n + k];
                    ZT[(long long)(i + 1) * n + k] = s * zi + c * zi1;
                    zi1 = c * zi - s * zi1;
                }
                ZT[(long long)lo * n + k] = zi1;
            }
            __syncthreads();
        }
    }

    // ---- ascending order ---------------------------------------------------------------------
    for (int i = tid; i < n; i += NT) {
        const double li = d[i];
        int rank = 0;
        for (int j = 0; j < n; ++j) rank += (d[j] < li) || (d[j] == li && j < i);
        lam_sorted[rank] = li;
        for (int z = 0; z < n; ++z) {
            const double val = ZT[(long long)i * n + z];
            Q_sorted[(long long)z * n + rank] = val;
            QT_sorted[(long long)rank * n + z] = val;
        }
    }
    if (tid == 0) *info = s_fail;
}

void launch_eigensystem(int n, double *A, double *Vec, double *lam_sorted, double *Q_sorted,
                        double *QT_sorted, int *info, hipStream_t s) {
    if (n < 1 || n > 2048) throw Error(3, "eigensystem: 1 <= n <= 2048 required");
    const size_t shm = sizeof(double) * (size_t)n * (7 + EIG_GROUPS);
    set_dynamic_lds((const void *)eig_tridiag_ql_kernel, shm);
    hipLaunchKernelGGL(eig_tridiag_ql_kernel, dim3(1), dim3(EIG_THREADS), shm, s, n, A, Vec,
                       lam_sorted, Q_sorted, QT_sorted, info);
    MSX_LAUNCH_CHECK();
}

}  // namespace msx
