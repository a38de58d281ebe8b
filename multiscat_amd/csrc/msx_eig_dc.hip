// Eigensystem of H0 = Q diag(lam) Q^T by Householder tridiagonalisation and divide and conquer.
//
// The reference diagonalises H0 with an EISPACK-style Householder + QL pair
// (lib/multiscat_fortran/fortran/diagsub.f90:5-284, called from multiscat_gmres.f90:251-253);
// the eigenpairs are the same up to eigenvector signs. The QL stage is a serial bulge chase; here
// the tridiagonal (d, e) from the Householder stage is solved by Cuppen's divide and conquer
// instead, which is parallel at every step:
//   * Householder reduction A = Q_H T Q_H^T, Q_H = H_0 ... H_{n-3} (reflectors kept, Q_H never
//     formed);
//   * tearing: T = diag(T1, T2) + |beta| u u^T at the boundary p of every node of a balanced
//     binary tree whose leaves are single rows (d_{p-1} and d_p lose |e_{p-1}|), so the leaves
//     are the 1 x 1 problems (d_i', e_i) (J. J. M. Cuppen, Numer. Math. 36 (1981) 177);
//   * merge of two solved children: D + rho z z^T with z the last row of Q1 and the first row of
//     Q2 (times sign(beta)), scaled so that ||z|| = 1 and rho = 2 |beta|; deflation of small
//     z_i and of close pole pairs by a Givens rotation (tolerance 8 eps max(|D|, rho));
//   * secular roots 1 + rho sum z_i^2 / (d_i - lam) = 0, one per thread, each kept as
//     lam = d_o + tau from its nearer pole d_o so that d_i - lam = (d_i - d_o) - tau carries no
//     cancellation; iteration on the two-pole rational model of the equation (the
//     "middle way" of R.-C. Li, LAPACK Working Note 89, 1994) inside a bisection bracket;
//   * eigenvectors from the weights recomputed from the roots (M. Gu and S. C. Eisenstat, SIAM
//     J. Matrix Anal. Appl. 16 (1995) 172), z_i^2 = prod_k (lam_k - d_i) / (rho prod_{j!=i}
//     (d_j - d_i)), which makes them orthogonal to working precision; the product
//     Q-hat U of the children's eigenvectors with the secular eigenvectors runs on the FP64
//     matrix cores.
// Every tree level is one launch with one workgroup per node. The Householder stage holds its
// trailing block in LDS once it fits, and the back-transformation applies its reflectors to
// the tridiagonal eigenvectors one column per lane group.
#include <algorithm>
#include <cfloat>

#include "msx_internal.h"
#include "msx_mfma.h"

namespace msx {

constexpr int DC_MAX_N = 1024;

bool eig_dc_supported(int n) { return n >= 2 && n <= DC_MAX_N; }

// first row of node k at depth t of the tree over n rows
__host__ __device__ inline int dc_row(int n, int t, int k) { return (int)(((long long)k * n) >> t); }

// Leaves: every row is its own 1 x 1 problem after the tearing at all n - 1 boundaries.
__global__ void dc_leaves_kernel(int n, const double *__restrict__ d, const double *__restrict__ e,
                                 double *__restrict__ lam, double *__restrict__ S) {
    const long long nn = (long long)n * n;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < nn;
         idx += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(idx / n), j = (int)(idx - (long long)i * n);
        S[idx] = i == j ? 1.0 : 0.0;
        if (i == j) {
            double v = d[i];
            if (i > 0) v -= fabs(e[i - 1]);
            if (i + 1 < n) v -= fabs(e[i]);
            lam[i] = v;
        }
    }
}

// 1 / x from the hardware reciprocal and two Newton steps (about 1 ulp, a third of the
// instructions of a correctly rounded division); used in the O(K^2) loops of the merge.
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

// Lanes per secular root (and per weight and norm): each takes every DC_G-th term and the
// partial results meet by butterfly shuffles, which give every lane of the group the same
// bits (each step adds the same two operands), so the group steers one iteration.
constexpr int DC_G = 4;
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int o = 1; o < DC_G; o <<= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ double group_prod(double v) {
#pragma unroll
    for (int o = 1; o < DC_G; o <<= 1) v *= __shfl_xor(v, o);
    return v;
}

// Root k of g(lam) = 1 + rho sum_i z_i^2 / (dl_i - lam) (K poles dl ascending, rho > 0): the
// origin pole o and tau with lam = dl[o] + tau. Called by the DC_G lanes of a group (gl their
// index in it).
__device__ void dc_secular_root(int k, int K, const double *dl, const double *z, double rho,
                                int gl, int &o_out, double &tau_out) {
    double lo, hi;
    int o;
    if (k < K - 1) {
        // g increases from -inf to +inf on (dl_k, dl_k+1): its sign at the midpoint tells which
        // pole is nearer
        const double mid = 0.5 * (dl[k + 1] - dl[k]);
        double f = 0.0;
        for (int i = gl; i < K; i += DC_G) f += rho * z[i] * z[i] * rcp_nr((dl[i] - dl[k]) - mid);
        f = 1.0 + group_sum(f);
        if (f >= 0.0) {
            o = k;
            lo = 0.0;
            hi = mid;
        } else {
            o = k + 1;
            lo = -mid;
            hi = 0.0;
        }
    } else {
        // the last root lies in (dl_K, dl_K + rho ||z||^2]
        double zz = 0.0;
        for (int i = gl; i < K; i += DC_G) zz += z[i] * z[i];
        o = k;
        lo = 0.0;
        hi = rho * group_sum(zz);
    }
    double tau = 0.5 * (lo + hi);
    const double dlo = dl[o];
    const double Dk = dl[k] - dlo, Dk1 = k < K - 1 ? dl[k + 1] - dlo : 0.0;
    for (int it = 0; it < 80; ++it) {
        // psi: poles left of the root (terms <= 0), phi: poles right of it (terms >= 0)
        double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
        for (int i = gl; i < K; i += DC_G) {
            const double r = rcp_nr((dl[i] - dlo) - tau);
            const double term = rho * z[i] * (z[i] * r);
            if (i <= k) {
                psi += term;
                dpsi += term * r;
            } else {
                phi += term;
                dphi += term * r;
            }
        }
        psi = group_sum(psi);
        dpsi = group_sum(dpsi);
        phi = group_sum(phi);
        dphi = group_sum(dphi);
        const double g = 1.0 + psi + phi;
        const double err = 8.0 * (phi - psi + 1.0) + K * (phi - psi);
        if (fabs(g) <= DBL_EPSILON * err) break;
        if (g < 0.0) lo = tau;
        else hi = tau;
        if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi))) break;
        // root of the model c0 + b1 / (d1 - eta) + b2 / (d2 - eta) that matches g and g' at tau
        const double d1 = Dk - tau;
        double eta = 0.0;
        bool ok = false;
        if (k < K - 1) {
            const double d2 = Dk1 - tau;
            const double b1 = dpsi * d1 * d1, b2 = dphi * d2 * d2;
            const double c0 = g - b1 / d1 - b2 / d2;
            const double A = c0, B = c0 * (d1 + d2) + b1 + b2, C = d1 * d2 * g;
            const double q = B + copysign(sqrt(fmax(B * B - 4.0 * A * C, 0.0)), B);
            if (q != 0.0) {
                eta = 2.0 * C / q;
                ok = d1 < eta && eta < d2 && lo < tau + eta && tau + eta < hi;
            }
            if (!ok && A != 0.0) {
                eta = q / (2.0 * A);
                ok = d1 < eta && eta < d2 && lo < tau + eta && tau + eta < hi;
            }
        } else {
            const double b1 = dpsi * d1 * d1;
            const double c0 = g - b1 / d1;
            if (c0 > 0.0) {
                eta = d1 + b1 / c0;
                ok = lo < tau + eta && tau + eta < hi;
            }
        }
        tau = ok ? tau + eta : 0.5 * (lo + hi);
    }
    o_out = o;
    tau_out = tau;
}

constexpr int DC_LDS_DOUBLES = 10, DC_LDS_INTS = 7;  // arrays of length m in the merge kernel

// Merge of the two children of node blockIdx.x at depth t. S (n x n, row-major) holds in its
// diagonal block [a, b) x [a, b) the eigenvectors of the node (columns, ascending eigenvalues
// in lam[a .. b)); Wq and Wu are n x n workspaces used in the same block.
// S, Wq and Wu are disjoint (restrict), so the copies below keep their loads in flight.
__global__ __launch_bounds__(1024) void dc_merge_kernel(int n, int t, const double *__restrict__ e,
                                                        double *lam, double *__restrict__ S,
                                                        double *__restrict__ Wq,
                                                        double *__restrict__ Wu) {
    const int node = blockIdx.x;
    const int a = dc_row(n, t, node), b = dc_row(n, t, node + 1);
    const int m = b - a;
    if (m < 2) return;  // a single row: its child's eigenpair is already in place
    const int p = dc_row(n, t + 1, 2 * node + 1);
    const int m1 = p - a;  // 1 <= m1 < m for every node of two or more rows
    const int tid = threadIdx.x, NT = blockDim.x;
    extern __shared__ double dcs[];
    double *D = dcs, *z = D + m, *dl = z + m, *zz = dl + m, *tau = zz + m, *zh = tau + m;
    double *inrm = zh + m, *val = inrm + m, *ru = val + m, *rw = ru + m;
    int *perm = reinterpret_cast<int *>(rw + m), *nd = perm + m, *dfl = nd + m, *org = dfl + m;
    int *rank = org + m, *rp = rank + m, *rj = rp + m;
    __shared__ int s_K, s_R;
    const double beta = e[p - 1];
    const double rho = 2.0 * fabs(beta), sg = beta >= 0.0 ? 1.0 : -1.0;
    const double r2 = 0.70710678118654752440;
    const int wave = tid >> 6, lane = tid & 63, NW = NT >> 6;
    // element (r, c) of the node's block of an n x n matrix (n <= 1024: 32-bit offsets)
    auto at = [&](double *M, int r, int c) -> double & { return M[(a + r) * n + a + c]; };

    // (1) merged order of the children's eigenvalues (ties: first child first) and z
    for (int c = tid; c < m; c += NT) {
        const double v = lam[a + c];
        int pos;
        if (c < m1) {
            int lo = 0, hi = m - m1;
            while (lo < hi) {
                const int md = (lo + hi) >> 1;
                if (lam[p + md] < v) lo = md + 1;
                else hi = md;
            }
            pos = c + lo;
        } else {
            int lo = 0, hi = m1;
            while (lo < hi) {
                const int md = (lo + hi) >> 1;
                if (lam[a + md] <= v) lo = md + 1;
                else hi = md;
            }
            pos = (c - m1) + lo;
        }
        perm[pos] = c;
        D[pos] = v;
        z[pos] = c < m1 ? r2 * at(S, m1 - 1, c) : sg * r2 * at(S, m1, c);
    }
    __syncthreads();
    // (2) Q-hat = diag(Q1, Q2) with its columns in merged order, a row's loads (up to 16
    // columns per lane) issued before its stores, by the wavefronts other than the first
    // (whose first thread runs the deflation scan (3) meanwhile) when there are several
    const int cw0 = NW > 1 ? 1 : 0;
    for (int r = wave - cw0; wave >= cw0 && r < m; r += NW - cw0) {
        constexpr int QC = 16;  // m <= 1024
        double v[QC];
#pragma unroll
        for (int k = 0; k < QC; ++k) {
            const int q = lane + 64 * k;
            const int c = q < m ? perm[q] : 0;
            v[k] = (q < m && (r < m1) == (c < m1)) ? at(S, r, c) : 0.0;
        }
#pragma unroll
        for (int k = 0; k < QC; ++k) {
            const int q = lane + 64 * k;
            if (q < m) at(Wq, r, q) = v[k];
        }
    }
    // (3) deflation (one thread: the pair test chains along the non-deflated entries). The
    // close-pole test |u w (d_j - d_prev)| <= tol, u = z_prev / t, w = z_j / t, t^2 = z_prev^2 +
    // z_j^2, is evaluated as |z_prev z_j (d_j - d_prev)| <= tol t^2 (no square root or division
    // unless the pair rotates), the current pair's values stay in registers and the next
    // entry is loaded ahead: the loop is this kernel's serial part
    if (tid == 0) {
        const double tol = 8.0 * DBL_EPSILON * fmax(fmax(fabs(D[0]), fabs(D[m - 1])), rho);
        int K = 0, nq = 0, R = 0, prev = -1;
        double zp = 0.0, Dp = 0.0;  // z and D of prev, as updated
        double zn = z[0], Dn = D[0];
        for (int j = 0; j < m; ++j) {
            const double zj = zn, Dj = Dn;
            if (j + 1 < m) {
                zn = z[j + 1];
                Dn = D[j + 1];
            }
            if (rho * fabs(zj) <= tol) {
                dfl[nq++] = j;
                continue;
            }
            if (prev < 0) {
                prev = j;
                zp = zj;
                Dp = Dj;
                continue;
            }
            if (fabs(zp * zj * (Dj - Dp)) <= tol * (zp * zp + zj * zj)) {
                // rotate columns (prev, j) so that z_prev = 0: prev deflates
                const double tt = hypot(zp, zj);
                const double u = zp / tt, w = zj / tt;
                rp[R] = prev;
                rj[R] = j;
                ru[R] = u;
                rw[R] = w;
                ++R;
                D[prev] = w * w * Dp + u * u * Dj;
                const double dj_new = u * u * Dp + w * w * Dj;
                D[j] = dj_new;
                z[j] = tt;
                z[prev] = 0.0;
                dfl[nq++] = prev;
                zp = tt;
                Dp = dj_new;
            } else {
                nd[K++] = prev;
                zp = zj;
                Dp = Dj;
            }
            prev = j;
        }
        if (prev >= 0) nd[K++] = prev;
        s_K = K;
        s_R = R;
    }
    __syncthreads();
    const int K = s_K, R = s_R;
    // (4) the deflating rotations on the rows of Q-hat, in order
    for (int r = tid; r < m; r += NT) {
        double *row = &at(Wq, r, 0);
        for (int q = 0; q < R; ++q) {
            const double x = row[rp[q]], y = row[rj[q]];
            row[rp[q]] = rw[q] * x - ru[q] * y;
            row[rj[q]] = ru[q] * x + rw[q] * y;
        }
    }
    for (int i = tid; i < K; i += NT) {
        dl[i] = D[nd[i]];
        zz[i] = z[nd[i]];
    }
    __syncthreads();
    // (5) secular roots, DC_G lanes per root; eigenvalues of the node, non-deflated first,
    // then deflated
    const int gq = tid / DC_G, gl = tid % DC_G, NG = NT / DC_G;
    for (int q = gq; q < K; q += NG) {
        int o;
        double tq;
        dc_secular_root(q, K, dl, zz, rho, gl, o, tq);
        if (gl == 0) {
            org[q] = o;
            tau[q] = tq;
            val[q] = dl[o] + tq;
        }
    }
    for (int q = tid; q < m - K; q += NT) val[K + q] = D[dfl[q]];
    __syncthreads();
    // (6) recomputed weights (DC_G lanes each), ranks of the eigenvalues
    for (int i = gq; i < K; i += NG) {
        const double di = dl[i];
        double prod = 1.0;
        for (int q = gl; q < K - 1; q += DC_G) {
            // (lam_q - d_i) / (d_q - d_i) for q < i, / (d_q+1 - d_i) for q >= i: both in (0, 1)
            const double num = -((di - dl[org[q]]) - tau[q]);
            const double den = (q < i ? dl[q] : dl[q + 1]) - di;
            prod *= num * rcp_nr(den);
        }
        prod = group_prod(prod) * (-((di - dl[org[K - 1]]) - tau[K - 1]) / rho);
        if (gl == 0) zh[i] = copysign(sqrt(fmax(prod, 0.0)), zz[i]);
    }
    for (int s = tid; s < m; s += NT) {
        const double v = val[s];
        int r = 0;
        for (int q = 0; q < m; ++q) r += (val[q] < v) || (val[q] == v && q < s);
        rank[s] = r;
    }
    __syncthreads();
    // (7) ascending eigenvalues; norms of the secular eigenvectors u_q(i) = zh_i / (d_i - lam_q)
    for (int s = tid; s < m; s += NT) lam[a + rank[s]] = val[s];
    for (int q = gq; q < K; q += NG) {
        const double base = dl[org[q]], tq = tau[q];
        double ss = 0.0;
        for (int i = gl; i < K; i += DC_G) {
            const double u = zh[i] * rcp_nr((dl[i] - base) - tq);
            ss += u * u;
        }
        ss = group_sum(ss);
        if (gl == 0) inrm[q] = 1.0 / sqrt(ss);
    }
    __syncthreads();
    // (8) U (K x K) into Wu; deflated eigenvectors are columns of Q-hat
    for (int i = wave; i < K; i += NW) {
        const double zi = zh[i], di = dl[i];
        for (int q = lane; q < K; q += 64) at(Wu, i, q) = zi * inrm[q] * rcp_nr((di - dl[org[q]]) - tau[q]);
    }
    const int nd_cnt = m - K;
    for (int r = wave; r < m; r += NW)
        for (int q = lane; q < nd_cnt; q += 64) at(S, r, rank[K + q]) = at(Wq, r, dfl[q]);
    __syncthreads();
    // (9) S[:, rank_q] = Q-hat[:, nd] U[:, q], 16 x 16 tiles per wavefront on the matrix cores
    using M = Mfma<double>;
    const int l16 = lane & 15, g = lane >> 4;
    const int tr = (m + 15) >> 4, tc = (K + 15) >> 4;
    for (int tile = wave; tile < tr * tc; tile += NW) {
        const int r0 = (tile / tc) * 16, q0 = (tile % tc) * 16;
        const int ra = r0 + l16, qb = q0 + l16;
        const double *arow = &at(Wq, ra < m ? ra : m - 1, 0);
        M::acc acc = {0.0, 0.0, 0.0, 0.0};
        // eight k-steps at a time, their loads unconditional at clamped indices (the masks are
        // applied after the loads) and all issued before the products: loads under
        // per-element conditions became branches that waited for each load in turn
        const bool rok = ra < m, qok = qb < K;
        const int qbc = qok ? qb : 0;
        for (int i0 = 0; i0 < K; i0 += 32) {
            double av[8], bv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = i0 + 4 * u + g, ic = i < K ? i : 0;
                const double a_raw = arow[nd[ic]], b_raw = at(Wu, ic, qbc);
                av[u] = (rok && i < K) ? a_raw : 0.0;
                bv[u] = (qok && i < K) ? b_raw : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = M::op(av[u], bv[u], acc);
        }
        if (qb < K) {
            const int col = rank[qb];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = r0 + M::row(g, v);
                if (r < m) at(S, r, col) = acc[v];
            }
        }
    }
}

// ---- Householder stage -------------------------------------------------------------------
// Forward order (Golub & Van Loan sec. 8.3.1, the order of LAPACK's dsytrd with the lower
// triangle and of the reference's routine): row k = 0 .. n-3 is reduced by the reflector
// H_k = I - u u^T / h acting on coordinates k+1 .. n-1, so the active block is the trailing one.
// H0 is graded at both ends (the kinetic matrix of the Lobatto basis is largest at the end
// nodes); reducing from the top keeps the low-lying eigenvectors an order of magnitude closer
// to LAPACK's than the bottom-up order does. The block is kept in full (both triangles) so
// that the matrix-vector product and the rank-2 update are plain row sweeps: in place in the
// input matrix while it is large, then in LDS (odd row stride) once rows k .. n-1 fit. u is
// kept in row k (columns k+1 .. n-1), which the later steps no longer touch.
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

constexpr int TRED_WAVES = 16;

// One reduction step: row k of the matrix M, whose element (r, c) (r, c >= b0) sits at
// M[(r - b0) * ld + (c - b0)] (LDS or global). pp: LDS partial sums [TRED_WAVES][64]; q: LDS
// vector of length >= m. Three workgroup barriers; the two reductions are computed
// redundantly by the wavefronts that need them (same order, same value), and u_0 is
// substituted in registers and written into row k after the last barrier.
__device__ __forceinline__ void tred_step(int n, int k, double *M, int b0, int ld, double *pp,
                                          double *q, double *h_out, double *e_out) {
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m = n - k - 1;  // trailing block size, coordinates k+1 .. n-1 as t = 0 .. m-1
    double *x = M + (long long)(k - b0) * ld + (k + 1 - b0);  // x[t] = A[k][k+1+t]
    double *B = M + (long long)(k + 1 - b0) * ld + (k + 1 - b0);  // B[r*ld + c] = A22[r][c]
    const double f = x[0];
    double loc = 0.0;
    for (int t = 1 + lane; t < m; t += 64) loc += x[t] * x[t];
    const double sigma = wave_sum(loc);
    if (!(sigma > 0.0)) {  // row k is already tridiagonal
        if (tid == 0) {
            e_out[k] = f;
            h_out[k] = 0.0;
        }
        __syncthreads();
        return;
    }
    const double g = f > 0.0 ? -sqrt(sigma + f * f) : sqrt(sigma + f * f);
    const double h = sigma + f * f - f * g;  // ||u||^2 / 2 with u = x - g e_0
    const double u0 = f - g, hinv = 1.0 / h;
    auto u = [&](int t) { return t == 0 ? u0 : x[t]; };
    // p = A22 u = A22^T u (the block is exactly symmetric): wavefront w takes a slice of 64
    // columns (one per lane, so a row segment is one coalesced read) and a group of rows
    const int CS = (m + 63) >> 6, RGP = TRED_WAVES / CS;  // column slices, row groups
    {
        const int cs = wave % CS, rgp = wave / CS;
        if (rgp < RGP) {
            const int c = 64 * cs + lane;
            const int rb = (m * rgp) / RGP, re = (m * (rgp + 1)) / RGP;
            double acc0 = 0.0, acc1 = 0.0;
            if (c < m) {
                const double *col = B + c;
                int r = rb;
                for (; r + 1 < re; r += 2) {
                    acc0 = fma(col[(long long)r * ld], u(r), acc0);
                    acc1 = fma(col[(long long)(r + 1) * ld], u(r + 1), acc1);
                }
                if (r < re) acc0 = fma(col[(long long)r * ld], u(r), acc0);
            }
            pp[wave * 64 + lane] = acc0 + acc1;
        }
    }
    __syncthreads();
    // wavefront 0: p from the partial sums, K = u^T p / (2 h^2), q = p / h - K u
    if (wave == 0) {
        double kl = 0.0;
        for (int c = lane; c < m; c += 64) {
            const int cs = c >> 6;
            double pc = 0.0;
            for (int g2 = 0; g2 < RGP; ++g2) pc += pp[(g2 * CS + cs) * 64 + (c & 63)];
            q[c] = pc;
            kl += u(c) * pc;
        }
        const double K = 0.5 * wave_sum(kl) * hinv * hinv;
        for (int c = lane; c < m; c += 64) q[c] = q[c] * hinv - K * u(c);
    }
    __syncthreads();
    // A22 -= u q^T + q u^T, one row per wavefront, lanes along the row; both triangles are
    // stored, and the update keeps them bitwise equal (products rounded separately, then one
    // commutative add)
    {
        constexpr int KT = 4;  // columns per lane and pass (256 columns a pass)
        for (int c0 = 0; c0 < m; c0 += 64 * KT) {
            double uk[KT], qk[KT];
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                const int c = c0 + lane + 64 * t;
                uk[t] = c < m ? u(c) : 0.0;
                qk[t] = c < m ? q[c] : 0.0;
            }
            for (int r = wave; r < m; r += TRED_WAVES) {
                const double ur = u(r), qr = q[r];
                double *row = B + (long long)r * ld;
#pragma unroll
                for (int t = 0; t < KT; ++t) {
                    const int c = c0 + lane + 64 * t;
                    if (c < m) row[c] = __dsub_rn(row[c], __dadd_rn(__dmul_rn(ur, qk[t]), __dmul_rn(qr, uk[t])));
                }
            }
        }
    }
    if (tid == 0) {
        e_out[k] = g;
        h_out[k] = h;
    }
    __syncthreads();
    if (tid == 0) x[0] = u0;  // the reflector is stored in row k, which no later step reads
}

// Steps 0 .. k_end - 1 of the reduction of A (n x n row-major, symmetric, both triangles
// stored; overwritten) in place in global memory, for n > 173, where the trailing block of
// the first steps does not fit LDS: reflector u_k into row k of A (columns k+1 .. n-1), h[k]
// (0: no reflector) and e[k]; rows and columns k_end .. n-1 are left to
// eig_tred_packed_kernel.
__global__ __launch_bounds__(1024) void eig_tred_kernel(int n, double *A, double *h_out,
                                                        double *e_out, long long *ticks,
                                                        int k_end) {
    extern __shared__ double tl[];
    double *pp = tl, *q = pp + TRED_WAVES * 64;
    if (ticks && threadIdx.x == 0) ticks[0] = wall_clock64();
    for (int k = 0; k + 2 < n && k < k_end; ++k) tred_step(n, k, A, 0, n, pp, q, h_out, e_out);
}

// Packed form: the lower triangle of A packed by rows in LDS for the whole reduction when
// n <= 173 (element (r, c), c <= r, at r (r + 1) / 2 + c: 90 KB at n = 150, where the full
// matrix would not fit), and for its last n - k0 rows (n - k0 <= 189) above that. Every wavefront forms the
// reflector itself from column k (u in registers, lane j holding coordinates j, j + 64,
// j + 128, and in a private LDS copy that the product reads as broadcasts), and after the
// partial sums of p = A22 u every wavefront completes p, K and q for all coordinates the same
// way, so a step has two workgroup barriers. q[r] for a wave-uniform row r comes from the
// owning lane by readlane. Same formulas as tred_step.
constexpr int TP_CH = 3;  // coordinate chunks of 64 per lane (n <= 192)

__device__ __forceinline__ double readlane_d(double x, int l) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Sum over the wavefront by DPP within rows of 16 lanes (quad_perm [1,0,3,2] and [2,3,0,1],
// half-row mirror, row mirror: every lane of a row then holds the row's sum, each step adding
// the same two operands in both lanes) and the four row sums read out in a fixed order: the
// same bits in every lane and every wavefront. A few cycles per step where the shuffle form
// waits on an LDS permute twice a step.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
// Sum over a DPP row of 16 lanes, the same bits in all of them.
__device__ __forceinline__ double row16_sum(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return v;
}
// v[r] for a wave-uniform coordinate r of a lane-distributed vector: all three chunks are read
// (scalar results) and one is selected, since a select among the registers before the
// readlane becomes an indexed array in scratch memory
__device__ __forceinline__ double coord(const double (&v)[TP_CH], int r) {
    const int s = r >> 6, l = r & 63;
    const double a = readlane_d(v[0], l), b = readlane_d(v[1], l), c = readlane_d(v[2], l);
    return s == 0 ? a : (s == 1 ? b : c);
}

//
// k0 > 0: steps 0 .. k0 - 1 ran in global memory (eig_tred_kernel with k_end = k0), and the
// triangle holds rows and columns k0 .. n-1 only (element (r, c) at P(r, c) =
// (r - k0)(r - k0 + 1) / 2 + c - k0). SHU: one copy of u shared by the wavefronts instead of one
// a wavefront (every wavefront writes the same values before it reads them, and the step's
// barriers keep a step's reads from the next step's writes), which leaves LDS for a larger
// triangle.
template <bool SHU>
__global__ __launch_bounds__(1024) void eig_tred_packed_kernel(int n, double *A, double *h_out,
                                                               double *d_out, double *e_out,
                                                               long long *ticks, int k0) {
    extern __shared__ double tl[];
    constexpr int UCOPIES = SHU ? 1 : TRED_WAVES;
    // partial sums [TRED_WAVES][64], copies of u [UCOPIES][64 TP_CH], triangle
    double *pp = tl, *Lp = pp + TRED_WAVES * 64 + UCOPIES * 64 * TP_CH;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    double *uw = pp + TRED_WAVES * 64 + (SHU ? 0 : wave * 64 * TP_CH);
    if (ticks && tid == 0 && k0 == 0) ticks[0] = wall_clock64();
    auto P = [k0](int r, int c) {
        const int rr = r - k0;
        return ((rr * (rr + 1)) >> 1) + c - k0;
    };
    for (int r = k0 + wave; r < n; r += TRED_WAVES)
        for (int c = k0 + lane; c <= r; c += 64) Lp[P(r, c)] = A[(long long)r * n + c];
    __syncthreads();
    for (int k = k0; k + 2 < n; ++k) {
        const int m = n - k - 1, b = k + 1;  // block coordinates t = 0 .. m-1: rows b + t
        // x[t] = A[b + t][k] (column k below the diagonal), u = x - g e_0
        double u[TP_CH];
        double loc = 0.0;
#pragma unroll
        for (int s = 0; s < TP_CH; ++s) {
            const int t = lane + 64 * s;
            u[s] = t < m ? Lp[P(b + t, k)] : 0.0;
            if (t >= 1) loc = fma(u[s], u[s], loc);
        }
        const double sigma = wave_sum_dpp(loc);
        const double f = readlane_d(u[0], 0);
        if (!(sigma > 0.0)) {  // column k is already reduced: no reflector, no update
            if (tid == 0) {
                e_out[k] = f;
                h_out[k] = 0.0;
            }
            continue;
        }
        const double g = f > 0.0 ? -sqrt(sigma + f * f) : sqrt(sigma + f * f);
        const double h = sigma + f * f - f * g;  // ||u||^2 / 2
        const double hinv = 1.0 / h;
        if (lane == 0) u[0] = f - g;
#pragma unroll
        for (int s = 0; s < TP_CH; ++s) uw[lane + 64 * s] = u[s];  // read by this wavefront only
        if (wave == 0) {  // the reflector into row k of A for the back-transformation
#pragma unroll
            for (int s = 0; s < TP_CH; ++s) {
                const int t = lane + 64 * s;
                if (t < m) A[(long long)k * n + b + t] = u[s];
            }
        }
        if (tid == 0) {
            e_out[k] = g;
            h_out[k] = h;
        }
        // p = A22 u: wavefront (cs, rgp) sums the rows of group rgp for the 64 columns of
        // chunk cs, eight independent row loads at a time. A22[r][c] is L[c][r] for rows above
        // the chunk (down the packed row of c), L[r][c] below it (along the packed row of r),
        // and either within it
        const int CS = (m + 63) >> 6, RGP = TRED_WAVES / CS;
        const int cs = wave % CS, rgp = wave / CS;
        if (rgp < RGP) {
            const int c = 64 * cs + lane, cc = c < m ? c : m - 1;  // past m: clamped, unused
            const double *colp = Lp + P(b + cc, b);  // colp[r] = A22[r][cc] for r <= cc
            const int rb = (m * rgp) / RGP, re = (m * (rgp + 1)) / RGP;
            const int c0 = 64 * cs, c1 = c0 + 64;  // rows [c0, c1): mixed
            double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            auto run = [&](int r0, int r1, auto elem) {
                int r = r0;
                for (; r + 8 <= r1; r += 8) {
                    double v[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[j] = elem(r + j);
#pragma unroll
                    for (int j = 0; j < 8; ++j) acc[j] = fma(v[j], uw[r + j], acc[j]);
                }
                for (; r < r1; ++r) acc[0] = fma(elem(r), uw[r], acc[0]);
            };
            run(rb, min(re, c0), [&](int r) { return colp[r]; });
            run(max(rb, c0), min(re, c1),
                [&](int r) { return r >= cc ? Lp[P(b + r, b + cc)] : colp[r]; });
            run(max(rb, c1), re, [&](int r) { return Lp[P(b + r, b + cc)]; });
            pp[wave * 64 + lane] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) +
                                   ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        }
        __syncthreads();
        // every wavefront: p, K = u^T p / (2 h^2) and q = p / h - K u for its lanes
        double q[TP_CH];
        double kl = 0.0;
#pragma unroll
        for (int s = 0; s < TP_CH; ++s) {
            double pc = 0.0;
            if (s < CS)
                for (int g2 = 0; g2 < RGP; ++g2) pc += pp[(g2 * CS + s) * 64 + lane];
            q[s] = lane + 64 * s < m ? pc : 0.0;
            kl = fma(u[s], q[s], kl);
        }
        const double K = 0.5 * wave_sum_dpp(kl) * hinv * hinv;
#pragma unroll
        for (int s = 0; s < TP_CH; ++s) q[s] = q[s] * hinv - K * u[s];
        // A22 -= u q^T + q u^T on the packed rows (columns 0 .. r of block row r; one stored
        // copy, so fused), two rows per wavefront at a time, both loaded before either is
        // stored (four at a time measured slower)
        for (int r = wave; r < m; r += 2 * TRED_WAVES) {
            const int r2 = r + TRED_WAVES;
            double *row = Lp + P(b + r, b);
            double *row2 = Lp + P(b + (r2 < m ? r2 : r), b);
            double a1[TP_CH], a2[TP_CH];
#pragma unroll
            for (int s = 0; s < TP_CH; ++s) {
                const int c = lane + 64 * s;
                a1[s] = c <= r ? row[c] : 0.0;
                a2[s] = (r2 < m && c <= r2) ? row2[c] : 0.0;
            }
            const double ur = uw[r], qr = coord(q, r);
            const double ur2 = r2 < m ? uw[r2] : 0.0, qr2 = r2 < m ? coord(q, r2) : 0.0;
#pragma unroll
            for (int s = 0; s < TP_CH; ++s) {
                const int c = lane + 64 * s;
                if (c <= r) row[c] = fma(-ur, q[s], fma(-qr, u[s], a1[s]));
                if (r2 < m && c <= r2) row2[c] = fma(-ur2, q[s], fma(-qr2, u[s], a2[s]));
            }
        }
        __syncthreads();
    }
    for (int r = tid; r < n; r += blockDim.x) d_out[r] = r < k0 ? A[(long long)r * n + r] : Lp[P(r, r)];
    if (tid == 0) {
        if (n >= 2) {
            e_out[n - 2] = Lp[P(n - 1, n - 2)];
            h_out[n - 2] = 0.0;
        }
        e_out[n - 1] = 0.0;
        h_out[n - 1] = 0.0;
        if (ticks) ticks[1] = ticks[2] = wall_clock64();
    }
}

// Q = H_0 H_1 ... H_{n-3} S: G lanes own one column of S (R rows each, in registers) and apply
// the reflectors from the last to the first, with no workgroup barrier, each reflector's
// loads and reciprocal issued while the previous one is applied. Also copies the
// eigenvalues, sets info = 1 if any is not finite, and Q^T.
template <int G, int R>
__global__ __launch_bounds__(256) void eig_backtransform_kernel(
    int n, const double *__restrict__ U, const double *__restrict__ h,
    const double *__restrict__ S, const double *__restrict__ lam, double *__restrict__ lam_sorted,
    double *__restrict__ Q, double *__restrict__ QT, int *info, long long *ticks) {
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && tid == 0 && ticks) ticks[3] = wall_clock64();
    const int col = (blockIdx.x * 256 + tid) / G, gl = tid % G;
    if (col < n) {
        double y[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const int r = gl + G * t;
            y[t] = r < n ? S[(long long)r * n + col] : 0.0;
        }
        // reflector k + 1 is loaded while reflector k is applied (u[r] for r = k+1 .. n-1 in
        // row k of U; h = 0: no reflector, w = 0)
        double un[R], hin = 0.0;
        auto load_refl = [&](int k) {
            const double *u = U + (long long)k * n;
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const int r = gl + G * t;
                un[t] = (r > k && r < n) ? u[r] : 0.0;
            }
            const double hk = h[k];
            hin = hk != 0.0 ? 1.0 / hk : 0.0;
        };
        if (n >= 3) load_refl(n - 3);
        for (int k = n - 3; k >= 0; --k) {
            double ur[R];
#pragma unroll
            for (int t = 0; t < R; ++t) ur[t] = un[t];
            const double hinv = hin;
            if (k > 0) load_refl(k - 1);
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < R; ++t) s = fma(ur[t], y[t], s);
            if constexpr (G == 16) {
                s = row16_sum(s);
            } else {
#pragma unroll
                for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
            }
            const double w = s * hinv;
#pragma unroll
            for (int t = 0; t < R; ++t) y[t] = fma(-ur[t], w, y[t]);
        }
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const int r = gl + G * t;
            if (r < n) {
                Q[(long long)r * n + col] = y[t];
                QT[(long long)col * n + r] = y[t];
            }
        }
    }
    if (blockIdx.x == 0) {
        __shared__ int bad;
        if (tid == 0) bad = 0;
        __syncthreads();
        for (int i = tid; i < n; i += blockDim.x) {
            const double v = lam[i];
            lam_sorted[i] = v;
            if (!isfinite(v)) bad = 1;
        }
        __syncthreads();
        if (tid == 0) {
            *info = bad;
            if (ticks) ticks[4] = wall_clock64();
        }
    }
}

template <int G, int R>
static void launch_backtransform(int n, const double *U, const double *h, const double *S,
                                 const double *lam, double *lam_sorted, double *Q, double *QT,
                                 int *info, long long *ticks, hipStream_t s) {
    const unsigned blocks = (unsigned)(((long long)n * G + 255) / 256);
    hipLaunchKernelGGL((eig_backtransform_kernel<G, R>), dim3(blocks), dim3(256), 0, s, n, U, h,
                       S, lam, lam_sorted, Q, QT, info, ticks);
    MSX_LAUNCH_CHECK();
}

size_t eig_dc_workspace(int n) { return 3 * (size_t)n * n + 4 * (size_t)n; }

void launch_eigensystem_dc(int n, double *A, double *work, double *lam_sorted,
                           double *Q_sorted, double *QT_sorted, int *info, long long *ticks,
                           hipStream_t s) {
    if (!eig_dc_supported(n)) throw Error(3, "divide and conquer: 2 <= n <= 1024 required");
    const size_t n2 = (size_t)n * n;
    double *S = work, *Wq = S + n2, *Wu = Wq + n2, *d = Wu + n2, *e = d + n, *lam = e + n;
    double *h = lam + n;
    if (ticks) MSX_HIP(hipMemsetAsync(ticks, 0, 8 * sizeof(long long), s));
    // Householder stage: the packed lower triangle in LDS throughout when it fits 150 KiB
    // (n <= 173); up to n = 1024, the first k0 steps work in place in A and the packed kernel
    // takes the last n - k0 rows (at most 189: the triangle with one shared copy of u fits
    // 150 KiB, and a lane's three chunks cover the 188 block coordinates)
    constexpr size_t lds_cap = 150 * 1024;
    auto packed_bytes = [](int rows, int ucopies) {
        return sizeof(double) * ((size_t)TRED_WAVES * 64 + (size_t)ucopies * 64 * TP_CH +
                                 (size_t)rows * (rows + 1) / 2);
    };
    if (n <= 64 * TP_CH && packed_bytes(n, TRED_WAVES) <= lds_cap) {
        const size_t shm = packed_bytes(n, TRED_WAVES);
        set_dynamic_lds((const void *)eig_tred_packed_kernel<false>, shm);
        hipLaunchKernelGGL(eig_tred_packed_kernel<false>, dim3(1), dim3(64 * TRED_WAVES), shm, s,
                           n, A, h, d, e, ticks, 0);
        MSX_LAUNCH_CHECK();
    } else {
        int k0 = std::max(0, n - 1 - 64 * TP_CH);
        while (packed_bytes(n - k0, 1) > lds_cap) ++k0;
        if (k0 > 0) {
            const size_t fixed = sizeof(double) * ((size_t)TRED_WAVES * 64 + n);
            set_dynamic_lds((const void *)eig_tred_kernel, fixed);
            hipLaunchKernelGGL(eig_tred_kernel, dim3(1), dim3(64 * TRED_WAVES), fixed, s, n, A,
                               h, e, ticks, k0);
            MSX_LAUNCH_CHECK();
        }
        const size_t shm = packed_bytes(n - k0, 1);
        set_dynamic_lds((const void *)eig_tred_packed_kernel<true>, shm);
        hipLaunchKernelGGL(eig_tred_packed_kernel<true>, dim3(1), dim3(64 * TRED_WAVES), shm, s,
                           n, A, h, d, e, ticks, k0);
        MSX_LAUNCH_CHECK();
    }
    const unsigned leaf_blocks = (unsigned)std::min<size_t>((n2 + 255) / 256, 4096);
    hipLaunchKernelGGL(dc_leaves_kernel, dim3(leaf_blocks), dim3(256), 0, s, n, d, e, lam, S);
    MSX_LAUNCH_CHECK();
    int L = 0;
    while ((1 << L) < n) ++L;
    const size_t per = DC_LDS_DOUBLES * sizeof(double) + DC_LDS_INTS * sizeof(int);
    for (int t = L - 1; t >= 0; --t) {
        const int mmax = (n + (1 << t) - 1) >> t;
        const size_t shm = per * (size_t)mmax;
        const int threads = mmax <= 32 ? 64 : (mmax <= 128 ? 256 : 1024);
        set_dynamic_lds((const void *)dc_merge_kernel, shm);
        hipLaunchKernelGGL(dc_merge_kernel, dim3(1u << t), dim3(threads), shm, s, n, t, e, lam, S,
                           Wq, Wu);
        MSX_LAUNCH_CHECK();
    }
    // back-transformation with the reflectors left in A: G lanes per column, R = ceil(n / G)
    if (n <= 64) launch_backtransform<16, 4>(n, A, h, S, lam, lam_sorted, Q_sorted, QT_sorted, info, ticks, s);
    else if (n <= 128) launch_backtransform<16, 8>(n, A, h, S, lam, lam_sorted, Q_sorted, QT_sorted, info, ticks, s);
    else if (n <= 192) launch_backtransform<16, 12>(n, A, h, S, lam, lam_sorted, Q_sorted, QT_sorted, info, ticks, s);
    else if (n <= 256) launch_backtransform<16, 16>(n, A, h, S, lam, lam_sorted, Q_sorted, QT_sorted, info, ticks, s);
    else if (n <= 512) launch_backtransform<64, 8>(n, A, h, S, lam, lam_sorted, Q_sorted, QT_sorted, info, ticks, s);
    else launch_backtransform<64, 16>(n, A, h, S, lam, lam_sorted, Q_sorted, QT_sorted, info, ticks, s);
}

}  // namespace msx
