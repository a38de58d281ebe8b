/*
 * multiscat_oracle.c -- CPU restatement (FP64) of the reference close-coupling
 * solve. TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's
 * cpu_baseline leg load it as the checker. The product never links it.
 *
 * What it restates (reference files relative to the reference repo root):
 *   - upper-block coupling apply      lib/multiscat_fortran/fortran/multiscat_gmres.f90:620-639
 *   - block forward substitution L^-1 multiscat_gmres.f90:641-687
 *   - preconditioner factors f_j      multiscat_gmres.f90:255-271
 *   - scattering operator (flag)      multiscat_gmres.f90:452-475
 *   - unrestarted MGS-GMRES, window   multiscat_gmres.f90:477-618
 *   - complex Givens rotation         multiscat_gmres.f90:689-716
 * with one extension the reference's SciPy backend defines: an explicit channel
 * list (any subset of the table grid, in any order; multiscat/multiscat/_scipy.py:414-443).
 * The eigendecomposition of H0 = T + diag(Re V00) is an input here (the
 * reference's own rs/tred2/tql2 is a textbook symmetric eigensolver,
 * diagsub.f90:5-284; any exact eigensolver is a valid stand-in).
 *
 * Layouts (all row-major, "C order"):
 *   V[g*nz + k]        g = ix*nky + iy, table of V_G(z_k) in natural units
 *   psi[j*nz + k]      channel j of the solve order, node k
 *   Q[z*nz + m]        eigenvector m of H0 at node z;  lam[m] its eigenvalue
 *   fac[j*nz + k]      f_j = (H0 + e_j)^-1 e_N  (preconditioner factors)
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

typedef double _Complex cplx;

typedef struct {
    int nkx, nky, nz, nc;
    const cplx *V;
    const int *chx, *chy; /* storage indices of each channel */
    const double *e;      /* channel energies e_j             */
    const cplx *c;        /* outgoing log-derivative terms c_j */
    const double *Q, *lam, *fac;
} op_t;

static inline int wrap(int d, int n) { int r = d % n; return r < 0 ? r + n : r; }

static inline const cplx *vrow(const op_t *op, int j, int i) {
    int dx = wrap(op->chx[j] - op->chx[i], op->nkx);
    int dy = wrap(op->chy[j] - op->chy[i], op->nky);
    return op->V + ((size_t)dx * op->nky + dy) * op->nz;
}

/* y_j = sum_{i>j} V[G_j-G_i] x_i  (multiscat_gmres.f90:620-639). x and y may alias
 * because row j reads only rows i > j, which are still untouched. */
static void apply_upper(const op_t *op, const cplx *x, cplx *y) {
    const int nz = op->nz, nc = op->nc;
    cplx *acc = malloc(sizeof(cplx) * nz);
    for (int j = 0; j < nc; ++j) {
        for (int k = 0; k < nz; ++k) acc[k] = 0;
        for (int i = j + 1; i < nc; ++i) {
            const cplx *v = vrow(op, j, i);
            const cplx *xi = x + (size_t)i * nz;
            for (int k = 0; k < nz; ++k) acc[k] += v[k] * xi[k];
        }
        memcpy(y + (size_t)j * nz, acc, sizeof(cplx) * nz);
    }
    free(acc);
}

/* In place: x <- L^-1 x, forward substitution over channels with the
 * eigenbasis solve and Sherman-Morrison boundary fix (multiscat_gmres.f90:641-687). */
static void solve_lower(const op_t *op, cplx *x) {
    const int nz = op->nz, nc = op->nc;
    cplx *t = malloc(sizeof(cplx) * nz);
    for (int j = 0; j < nc; ++j) {
        cplx *xj = x + (size_t)j * nz;
        for (int i = 0; i < j; ++i) {
            const cplx *v = vrow(op, j, i);
            const cplx *xi = x + (size_t)i * nz;
            for (int k = 0; k < nz; ++k) xj[k] -= v[k] * xi[k];
        }
        const double ej = op->e[j];
        for (int m = 0; m < nz; ++m) {
            cplx s = 0;
            for (int z = 0; z < nz; ++z) s += xj[z] * op->Q[(size_t)z * nz + m];
            t[m] = s / (ej + op->lam[m]);
        }
        for (int z = 0; z < nz; ++z) {
            cplx s = 0;
            const double *qz = op->Q + (size_t)z * nz;
            for (int m = 0; m < nz; ++m) s += qz[m] * t[m];
            xj[z] = s;
        }
        const double *fj = op->fac + (size_t)j * nz;
        const cplx cj = op->c[j];
        const cplx f = xj[nz - 1] * cj / (1.0 - fj[nz - 1] * cj);
        for (int z = 0; z < nz; ++z) xj[z] += f * fj[z];
    }
    free(t);
}

/* f_j = Q (Q[N,:] / (e_j + lam))  (multiscat_gmres.f90:255-271). */
void oracle_preconditioner_factors(int nz, int nc, const double *Q, const double *lam,
                                   const double *e, double *fac) {
    double *g = malloc(sizeof(double) * nz);
    for (int j = 0; j < nc; ++j) {
        for (int m = 0; m < nz; ++m) g[m] = Q[(size_t)(nz - 1) * nz + m] / (e[j] + lam[m]);
        for (int z = 0; z < nz; ++z) {
            double s = 0;
            for (int m = 0; m < nz; ++m) s += Q[(size_t)z * nz + m] * g[m];
            fac[(size_t)j * nz + z] = s;
        }
    }
    free(g);
}

/* x <- x + L^-1 U x; with flag 1 additionally x <- x - L^-1 U x
 * (multiscat_gmres.f90:452-475). work holds n elements. */
static void apply_operator(const op_t *op, int flag, cplx *x, cplx *work) {
    const size_t n = (size_t)op->nc * op->nz;
    memcpy(work, x, sizeof(cplx) * n);
    apply_upper(op, x, x);
    solve_lower(op, x);
    for (size_t i = 0; i < n; ++i) x[i] = work[i] + x[i];
    if (flag == 1) {
        memcpy(work, x, sizeof(cplx) * n);
        apply_upper(op, x, x);
        solve_lower(op, x);
        for (size_t i = 0; i < n; ++i) x[i] = work[i] - x[i];
    }
}

/* LINPACK-style complex Givens rotation (multiscat_gmres.f90:689-716):
 * returns c, s with [c, conj(s); -s, conj(c)] (a, b)^T = (r, 0)^T, r real. */
static void givens(cplx a, cplx b, cplx *c, cplx *s, double *r_out) {
    const double aa = cabs(a), ab = cabs(b);
    const cplx rho = aa > ab ? a : b;
    const double scale = aa + ab;
    if (scale == 0.0) { *c = 1.0; *s = 0.0; *r_out = 0.0; return; }
    const cplx as = a / scale, bs = b / scale;
    double r = scale * sqrt(creal(as * conj(as) + bs * conj(bs)));
    if (creal(rho) < 0.0) r = -r;
    *c = conj(a / r);
    *s = b / r;
    *r_out = r;
}

static void apply_upper_only(const op_t *op, cplx *x) { apply_upper(op, x, x); }

/*
 * Unrestarted GMRES on (I + L^-1 U) psi = L^-1 b (left-preconditioned by
 * I - L^-1 U when flag == 1), x0 = 0, b = b_spec at node N of channel `spec`
 * (multiscat_gmres.f90:477-618). Stops when |g_{k+1}| < eps |g_1| holds for
 * `window` consecutive iterations or on breakdown. Returns 0 on convergence,
 * 1 if maxiter was reached (ierr=2 at the boundary), -1 on allocation failure.
 */
int oracle_gmres(int nkx, int nky, int nz, const cplx *V, int nc, const int *chx,
                 const int *chy, const double *e, const cplx *c, const double *Q,
                 const double *lam, const double *fac, int spec, double b_re, double b_im, int flag,
                 double eps, int maxiter, int window, cplx *psi, int *iterations,
                 double *residual_history) {
    op_t op = {nkx, nky, nz, nc, V, chx, chy, e, c, Q, lam, fac};
    const cplx b_spec = b_re + I * b_im;
    const size_t n = (size_t)nc * nz;
    int cap = 64 < maxiter + 1 ? 64 : maxiter + 1;
    cplx *basis = malloc(sizeof(cplx) * n * (size_t)cap);
    cplx *x = malloc(sizeof(cplx) * n), *w = malloc(sizeof(cplx) * n);
    const size_t L = (size_t)maxiter + 1;
    cplx *h = calloc(L * L, sizeof(cplx));
    cplx *g = calloc(L, sizeof(cplx)), *zc = calloc(L, sizeof(cplx));
    cplx *co = calloc(L, sizeof(cplx)), *si = calloc(L, sizeof(cplx));
    if (!basis || !x || !w || !h || !g || !zc || !co || !si) return -1;
#define H(i, j) h[(size_t)(j) * L + (i)]

    /* r0 = M L^-1 b */
    memset(x, 0, sizeof(cplx) * n);
    x[(size_t)spec * nz + nz - 1] = b_spec;
    solve_lower(&op, x);
    if (flag == 1) {
        memcpy(w, x, sizeof(cplx) * n);
        apply_upper_only(&op, x);
        solve_lower(&op, x);
        for (size_t i = 0; i < n; ++i) x[i] = w[i] - x[i];
    }
    double xnorm = 0;
    for (size_t i = 0; i < n; ++i) xnorm += creal(conj(x[i]) * x[i]);
    xnorm = sqrt(xnorm);
    g[0] = xnorm;
    const double r0 = xnorm > DBL_MIN ? xnorm : DBL_MIN;

    int kconv = 0, kk = 0;
    for (int k = 0; k < maxiter; ++k) {
        if (k >= cap) {
            int nc2 = cap * 2 < maxiter + 1 ? cap * 2 : maxiter + 1;
            cplx *nb = realloc(basis, sizeof(cplx) * n * (size_t)nc2);
            if (!nb) return -1;
            basis = nb;
            cap = nc2;
        }
        for (size_t i = 0; i < n; ++i) x[i] /= xnorm;
        cplx *vk = basis + (size_t)k * n;
        memcpy(vk, x, sizeof(cplx) * n);
        apply_operator(&op, flag, x, w);
        for (int j = 0; j <= k; ++j) { /* modified Gram-Schmidt */
            const cplx *vj = basis + (size_t)j * n;
            cplx s = 0;
            for (size_t i = 0; i < n; ++i) s += conj(vj[i]) * x[i];
            H(j, k) = s;
            for (size_t i = 0; i < n; ++i) x[i] -= vj[i] * s;
        }
        xnorm = 0;
        for (size_t i = 0; i < n; ++i) xnorm += creal(conj(x[i]) * x[i]);
        xnorm = sqrt(xnorm);
        H(k + 1, k) = xnorm;
        for (int j = 0; j < k; ++j) {
            cplx t = co[j] * H(j, k) + conj(si[j]) * H(j + 1, k);
            H(j + 1, k) = conj(co[j]) * H(j + 1, k) - si[j] * H(j, k);
            H(j, k) = t;
        }
        double r;
        givens(H(k, k), H(k + 1, k), &co[k], &si[k], &r);
        H(k, k) = r;
        H(k + 1, k) = 0;
        g[k + 1] = -si[k] * g[k];
        g[k] = co[k] * g[k];
        const double res = cabs(g[k + 1]);
        if (residual_history) residual_history[k] = res / r0;
        kconv = res < eps * r0 ? kconv + 1 : 0;
        kk = k + 1;
        if (kconv == window || xnorm == 0.0) break;
    }
    /* back substitution R z = g and x = sum_j v_j z_j */
    for (int j = 0; j < kk; ++j) zc[j] = g[j];
    for (int j = kk - 1; j >= 0; --j) {
        zc[j] /= H(j, j);
        for (int i = 0; i < j; ++i) zc[i] -= H(i, j) * zc[j];
    }
    memset(psi, 0, sizeof(cplx) * n);
    for (int j = 0; j < kk; ++j) {
        const cplx *vj = basis + (size_t)j * n;
        for (size_t i = 0; i < n; ++i) psi[i] += vj[i] * zc[j];
    }
    *iterations = kk;
    const int info = (kconv < window && xnorm > 0.0) ? 1 : 0;
#undef H
    free(basis); free(x); free(w); free(h); free(g); free(zc); free(co); free(si);
    return info;
}

/* Debug faces of the two operator blocks (same conventions as above). */
void oracle_apply_upper(int nkx, int nky, int nz, const cplx *V, int nc, const int *chx,
                        const int *chy, const cplx *x, cplx *y) {
    op_t op = {nkx, nky, nz, nc, V, chx, chy, 0, 0, 0, 0, 0};
    apply_upper(&op, x, y);
}

void oracle_solve_lower(int nkx, int nky, int nz, const cplx *V, int nc, const int *chx,
                        const int *chy, const double *e, const cplx *c, const double *Q,
                        const double *lam, const double *fac, cplx *x) {
    op_t op = {nkx, nky, nz, nc, V, chx, chy, e, c, Q, lam, fac};
    solve_lower(&op, x);
}

/* Full off-diagonal coupling y_j = sum_{i != j} V[G_j - G_i] x_i (the unsplit form,
 * multiscat/multiscat/_scipy_von_neumann.py:28-37). */
void oracle_apply_coupling(int nkx, int nky, int nz, const cplx *V, int nc, const int *chx,
                           const int *chy, const cplx *x, cplx *y) {
    op_t op = {nkx, nky, nz, nc, V, chx, chy, 0, 0, 0, 0, 0};
    for (int j = 0; j < nc; ++j) {
        cplx *yj = y + (size_t)j * nz;
        for (int k = 0; k < nz; ++k) yj[k] = 0;
        for (int i = 0; i < nc; ++i) {
            if (i == j) continue;
            const cplx *v = vrow(&op, j, i);
            const cplx *xi = x + (size_t)i * nz;
            for (int k = 0; k < nz; ++k) yj[k] += v[k] * xi[k];
        }
    }
}
