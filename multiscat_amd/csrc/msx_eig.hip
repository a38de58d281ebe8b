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
//       sweep and records its rotations while the other wavefronts apply the previous sweep's
//       rotation sequence, each lane to its own column of Z (stored transposed, so the
//       updates are coalesced).
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
                                                                     int *info,
                                                                     long long *ticks) {
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
    __shared__ int s_fail;
    const int tid = threadIdx.x, NT = blockDim.x;
    const int grp = tid / (NT / EIG_GROUPS), gl = tid % (NT / EIG_GROUPS), GS = NT / EIG_GROUPS;

    if (ticks && tid == 0) ticks[0] = wall_clock64();
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

    if (ticks && tid == 0) ticks[1] = wall_clock64();
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

    if (ticks && tid == 0) ticks[2] = wall_clock64();
    // ---- (2) implicit QL with shifts on (d, e), rotations applied to the rows of Z^T -----
    // Lane 0 runs the scalar bulge chases; a sweep only reads and writes (d, e), so while lane 0
    // chases sweep s+1 into one rotation buffer, wavefronts 1.. apply sweep s from the other
    // buffer to the columns of Z^T. pp is free after the tridiagonalisation and holds the
    // second buffer.
    double *cbuf[2] = {cs, pp}, *sbuf[2] = {sn, pp + n};
    int ql_l = 0, ql_iter = 0;  // lane 0's position in the QL iteration
    // One sweep of the iteration into (c_out, s_out); false when every eigenvalue has converged.
    // Run by all 64 lanes of wavefront 0 (identical ql_l, ql_iter in every lane): the search
    // for the first negligible e[m], m >= l, tests 64 indices per LDS round trip and takes the
    // lowest hit of a ballot; lane 0 alone chases the bulge and writes hi, lo.
    const int lane = tid & 63;
    auto ql_sweep = [&](double *c_out, double *s_out, int &hi, int &lo) -> bool {
        while (ql_l < n) {
            const int l = ql_l;
            int m = n - 1;
            for (int base = l; base < n - 1; base += 64) {
                const int mm = base + lane;
                bool small = false;
                if (mm < n - 1) small = fabs(e[mm]) <= DBL_EPSILON * (fabs(d[mm]) + fabs(d[mm + 1]));
                const unsigned long long hits = __ballot(small);
                if (hits) {
                    m = base + __ffsll(hits) - 1;
                    break;
                }
            }
            if (m == l) {
                ++ql_l;
                ql_iter = 0;
                continue;
            }
            if (ql_iter >= 60) {
                s_fail = 1;
                ++ql_l;
                ql_iter = 0;
                continue;
            }
            ++ql_iter;
            if (lane != 0) return true;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            double dn = d[m];  // d[i + 1] of the current step (read, not yet rewritten)
            double ei = e[m - 1], di = d[m - 1];
            int i;
            bool underflow = false;
            for (i = m - 1; i >= l; --i) {
                const int inext = i > l ? i - 1 : l;  // in range; unused at i == l
                const double e_next = e[inext], d_next = d[inext];
                const double f = s * ei, b = c * ei;
                r = sqrt(f * f + g * g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] = dn - p;
                    e[m] = 0.0;
                    underflow = true;
                    break;
                }
                const double rinv = 1.0 / r;
                s = f * rinv;
                c = g * rinv;
                g = dn - p;
                r = (di - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                c_out[i] = c;
                s_out[i] = s;
                dn = di;
                ei = e_next;
                di = d_next;
            }
            hi = m - 1;
            lo = underflow ? i + 1 : l;
            if (!underflow) {
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
            return true;
        }
        return false;
    };
    __shared__ int s_more[2], s_hi2[2], s_lo2[2];
    long long chase_cycles = 0, n_sweeps = 0, n_steps = 0;  // wavefront 0's diagnostics
    auto timed_sweep = [&](int slot) -> int {
        const long long t0 = ticks ? wall_clock64() : 0;
        const int more = ql_sweep(cbuf[slot], sbuf[slot], s_hi2[slot], s_lo2[slot]);
        if (ticks && tid == 0) {
            chase_cycles += wall_clock64() - t0;
            if (more) {
                ++n_sweeps;
                n_steps += s_hi2[slot] - s_lo2[slot] + 1;
            }
        }
        return more;
    };
    if (tid < 64) {
        const int more = timed_sweep(0);
        if (tid == 0) s_more[0] = more;
    }
    __syncthreads();
    constexpr int QC = 16;  // rows of Z^T loaded ahead per column
    for (int cur = 0;; cur ^= 1) {
        if (!s_more[cur]) break;
        const int hi = s_hi2[cur], lo = s_lo2[cur];
        if (tid < 64) {
            const int more = timed_sweep(cur ^ 1);
            if (tid == 0) s_more[cur ^ 1] = more;
        } else {
            const double *cb = cbuf[cur], *sb = sbuf[cur];
            for (int k = tid - 64; k < n; k += NT - 64) {
                double zi1 = ZT[(long long)(hi + 1) * n + k];
                for (int i0 = hi; i0 >= lo; i0 -= QC) {
                    const int cnt = min(QC, i0 - lo + 1);
                    double z[QC];
#pragma unroll
                    for (int t = 0; t < QC; ++t)
                        if (t < cnt) z[t] = ZT[(long long)(i0 - t) * n + k];
#pragma unroll
                    for (int t = 0; t < QC; ++t) {
                        if (t < cnt) {
                            const int i = i0 - t;
                            const double c = cb[i], s = sb[i];
                            ZT[(long long)(i + 1) * n + k] = s * z[t] + c * zi1;
                            zi1 = c * z[t] - s * zi1;
                        }
                    }
                }
                ZT[(long long)lo * n + k] = zi1;
            }
        }
        __syncthreads();
    }

    if (ticks && tid == 0) ticks[3] = wall_clock64();
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
    if (tid == 0) {
        *info = s_fail;
        if (ticks) {
            ticks[4] = wall_clock64();
            ticks[5] = chase_cycles;
            ticks[6] = n_sweeps;
            ticks[7] = n_steps;
        }
    }
}

void launch_eigensystem(int n, double *A, double *Vec, double *lam_sorted, double *Q_sorted,
                        double *QT_sorted, int *info, long long *ticks, hipStream_t s) {
    if (n < 1 || n > 2048) throw Error(3, "eigensystem: 1 <= n <= 2048 required");
    const size_t shm = sizeof(double) * (size_t)n * (7 + EIG_GROUPS);
    set_dynamic_lds((const void *)eig_tridiag_ql_kernel, shm);
    hipLaunchKernelGGL(eig_tridiag_ql_kernel, dim3(1), dim3(EIG_THREADS), shm, s, n, A, Vec,
                       lam_sorted, Q_sorted, QT_sorted, info, ticks);
    MSX_LAUNCH_CHECK();
}

}  // namespace msx
