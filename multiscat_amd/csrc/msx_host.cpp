// Host-side input builders of the drop-in group (the reference's scatsub.f90 helpers):
// Gauss-Lobatto rule, Lobatto kinetic matrix, channel energies e_G and boundary waves.
// Closed forms, restated:
//   GLL nodes: +-1 and the roots of P'_N (N = n-1); weights 2 / (N (N+1) P_N(x)^2)
//              (lib/multiscat_fortran/fortran/scatsub.f90:229-305)
//   kinetic:   T_ij = sum_k W_k u_i'(x_k) u_j'(x_k), u_i = l_i / sqrt(W_i), of an (nz+1)-node
//              rule with its first node dropped (scatsub.f90:134-226)
//   e_G:       |K + G|^2 - |k|^2, FFT-ordered G (scatsub.f90:54-86)
//   a, b, c:   exp(-2i dk zmax), sqrt(dk) exp(-i dk zmax) / sqrt(W_N), i dk / W_N for open
//              channels; 0, 0, -dk / W_N for closed ones (scatsub.f90:102-132, 307-335)
#include <cmath>
#include <vector>

#include "../../include/multiscat_amd.h"

namespace {

const double kPi = 3.14159265358979323846;

// P_N(x) and P_{N-1}(x) by Bonnet's recursion.
void legendre(int N, double x, double &p, double &pm) {
    double a = 1.0, b = x;
    if (N == 0) {
        p = 1.0;
        pm = 0.0;
        return;
    }
    for (int k = 2; k <= N; ++k) {
        const double c = ((2.0 * k - 1.0) * x * b - (k - 1.0) * a) / k;
        a = b;
        b = c;
    }
    p = b;
    pm = a;
}

// Ascending GLL nodes t and weights w on [-1, 1]; pN = P_N(t).
void gll(int n, std::vector<double> &t, std::vector<double> &w, std::vector<double> &pN) {
    const int N = n - 1;
    t.assign(n, 0.0);
    w.assign(n, 0.0);
    pN.assign(n, 0.0);
    t[0] = -1.0;
    t[N] = 1.0;
    for (int i = 1; i < N; ++i) {
        double x = -std::cos(kPi * i / N);
        for (int it = 0; it < 100; ++it) {
            double p, pm;
            legendre(N, x, p, pm);
            const double om = 1.0 - x * x;
            const double dp = N * (pm - x * p) / om;
            const double ddp = (2.0 * x * dp - N * (N + 1.0) * p) / om;
            const double step = dp / ddp;
            x -= step;
            if (std::fabs(step) < 1e-16) break;
        }
        t[i] = x;
    }
    for (int i = 1; i < N; ++i) {  // exact mirror symmetry
        const double s = 0.5 * (t[i] - t[N - i]);
        t[i] = s;
        t[N - i] = -s;
    }
    if (n % 2 == 1) t[N / 2] = 0.0;
    for (int i = 0; i < n; ++i) {
        double p, pm;
        legendre(N, t[i], p, pm);
        pN[i] = p;
        w[i] = 2.0 / (N * (N + 1.0) * p * p);
    }
}

void lobatto(int n, double zmin, double zmax, std::vector<double> &x, std::vector<double> &W,
             std::vector<double> &t, std::vector<double> &pN) {
    std::vector<double> w;
    gll(n, t, w, pN);
    const double half = 0.5 * (zmax - zmin);
    x.resize(n);
    W.resize(n);
    for (int i = 0; i < n; ++i) {
        x[i] = zmin + half * (t[i] + 1.0);
        W[i] = w[i] * half;
    }
}

}  // namespace

extern "C" {

int msx_get_lobatto_weights(double zmin, double zmax, int node_count, double *w, double *x) {
    if (node_count <= 0) return MSX_ERR_DIMS;
    if (!w || !x) return MSX_ERR_INVALID;
    if (node_count == 1) {
        x[0] = zmin;
        w[0] = std::sqrt(zmax - zmin);
        return MSX_OK;
    }
    std::vector<double> xs, W, t, pN;
    lobatto(node_count, zmin, zmax, xs, W, t, pN);
    for (int i = 0; i < node_count; ++i) {
        x[i] = xs[i];
        w[i] = std::sqrt(W[i]);
    }
    return MSX_OK;
}

int msx_get_parallel_kinetic_energy(double zmin, double zmax, int nz, double *out) {
    if (nz <= 0) return MSX_ERR_DIMS;
    if (!out) return MSX_ERR_INVALID;
    const int n = nz + 1, N = n - 1;
    std::vector<double> x, W, t, pN;
    lobatto(n, zmin, zmax, x, W, t, pN);
    // D[k][i] = l_i'(x_k) (spectral-element form at GLL nodes), scaled to [zmin, zmax]
    const double scale = 2.0 / (zmax - zmin);
    std::vector<double> U((size_t)n * n);  // U[k][i] = u_i'(x_k)
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < n; ++i) {
            double d;
            if (k == i) d = (k == 0) ? -N * (N + 1.0) / 4.0 : (k == N ? N * (N + 1.0) / 4.0 : 0.0);
            else d = pN[k] / (pN[i] * (t[k] - t[i]));
            U[(size_t)k * n + i] = d * scale / std::sqrt(W[i]);
        }
    for (int i = 1; i < n; ++i)
        for (int j = 1; j <= i; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += W[k] * U[(size_t)k * n + i] * U[(size_t)k * n + j];
            out[(size_t)(i - 1) + (size_t)nz * (j - 1)] = s;  // column-major, symmetric
            out[(size_t)(j - 1) + (size_t)nz * (i - 1)] = s;
        }
    return MSX_OK;
}

int msx_get_perpendicular_kinetic_difference(double kx, double ky, double kz, int nx, int ny,
                                             double ax, double ay, double bx, double by,
                                             double *out) {
    if (nx <= 0 || ny <= 0) return MSX_ERR_DIMS;
    if (!out) return MSX_ERR_INVALID;
    const double area = std::fabs(ax * by - ay * bx);
    if (!(area > 0)) return MSX_ERR_INVALID;
    const double f = 2.0 * kPi / area;
    const double gax = by * f, gay = -bx * f, gbx = -ay * f, gby = ax * f;
    const double k2 = kx * kx + ky * ky + kz * kz;
    for (int i = 0; i < nx; ++i) {
        const int mx = i > (nx - 1) / 2 ? i - nx : i;
        for (int j = 0; j < ny; ++j) {
            const int my = j > (ny - 1) / 2 ? j - ny : j;
            const double gx = gax * mx + gbx * my, gy = gay * mx + gby * my;
            out[(size_t)i + (size_t)nx * j] = (kx + gx) * (kx + gx) + (ky + gy) * (ky + gy) - k2;
        }
    }
    return MSX_OK;
}

int msx_get_abc_arrays(double zmin, double zmax, int nx, int ny, const double *e, int nz,
                       double *a, double *b, double *c) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return MSX_ERR_DIMS;
    if (!e || !a || !b || !c) return MSX_ERR_INVALID;
    std::vector<double> x, W, t, pN;
    lobatto(nz + 1, zmin, zmax, x, W, t, pN);
    const double wl = std::sqrt(W[nz]);  // sqrt weight of the last node
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const size_t idx = (size_t)i * ny + j;
            const double en = e[(size_t)i + (size_t)nx * j];
            const double dk = std::sqrt(std::fabs(en));
            if (en < 0.0) {
                const double th = dk * zmax;
                a[2 * idx] = std::cos(2.0 * th);
                a[2 * idx + 1] = -std::sin(2.0 * th);
                b[2 * idx] = std::sqrt(dk) * std::cos(th) / wl;
                b[2 * idx + 1] = -std::sqrt(dk) * std::sin(th) / wl;
                c[2 * idx] = 0.0;
                c[2 * idx + 1] = dk / (wl * wl);
            } else {
                a[2 * idx] = a[2 * idx + 1] = 0.0;
                b[2 * idx] = b[2 * idx + 1] = 0.0;
                c[2 * idx] = -dk / (wl * wl);
                c[2 * idx + 1] = 0.0;
            }
        }
    return MSX_OK;
}

}  // extern "C"
