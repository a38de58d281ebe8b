// All-channel diagonal blocks on the matrix cores.
//
//   inverse: y_j = D_j^-1 x_j = Q diag(1/(e_j + lam)) Q^T x_j + f_j g_j(N) c_j / (1 - f_j(N) c_j)
//   forward: y_j = D_j x_j    = Q diag(e_j + lam) Q^T x_j - c_j x_j(N) e_N
// (the reference applies D_j^-1 channel by channel inside its forward substitution,
// lib/multiscat_fortran/fortran/multiscat_gmres.f90:666-685; its SciPy backend applies D^-1
// to all channels at once, multiscat/multiscat/_scipy.py:249-259,336-354).
//
// A block owns TJ channels of one instance: their real and imaginary parts form 2 TJ real
// rows in LDS, and the two products with the real matrices Q and Q^T are 16x16x4 MFMA GEMMs
// (Q, Q^T zero-padded to nzp, a multiple of 16, in HBM; B fragments are 16 consecutive
// elements of a row). The 1/(e_j + lam_m) scaling is applied to the first product in
// registers, with e and lam kept in FP64 so the denominator never cancels in FP32.
#include "msx_mfma.h"
#include "msx_internal.h"

namespace msx {

constexpr int DIAG_MAXNT = 5;  // N tiles per wavefront: nzp <= 4 * 5 * 16 = 320

template <typename T, int TJ>
__global__ __launch_bounds__(256) void diag_mfma_kernel(Layout L, EigTables<T> E,
                                                        const double *__restrict__ e,
                                                        const cx<T> *__restrict__ c,
                                                        const T *__restrict__ fac,
                                                        const cx<T> *x, const cx<T> *base,
                                                        cx<T> *y, T sign, int inverse,
                                                        const int *__restrict__ active) {
    constexpr int MR = 2 * TJ, MT = MR / 16;
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int nc = L.nc[b];
    const int j0 = blockIdx.x * TJ;
    if (j0 >= nc) return;
    const int nz = L.nz, nzp = E.nzp, S = lds_stride_a(nzp);
    const long long off = L.off[b];
    const int rows = min(TJ, nc - j0);
    extern __shared__ double diag_smem[];
    T *buf = reinterpret_cast<T *>(diag_smem);  // [MR][S]: re rows, then im rows
    __shared__ double s_e[TJ];
    __shared__ cx<T> s_c[TJ], s_x0[TJ];
    for (int idx = threadIdx.x; idx < TJ * nzp; idx += blockDim.x) {
        const int r = idx / nzp, z = idx - r * nzp;
        cx<T> v = mk<T>(0, 0);
        if (r < rows && z < nz) v = x[(off + j0 + r) * nz + z];
        buf[r * S + z] = v.re;
        buf[(TJ + r) * S + z] = v.im;
    }
    if (threadIdx.x < TJ) {
        const int r = threadIdx.x;
        const bool ok = r < rows;
        s_e[r] = ok ? e[off + j0 + r] : 0.0;
        s_c[r] = ok ? c[off + j0 + r] : mk<T>(0, 0);
        s_x0[r] = ok ? x[(off + j0 + r) * nz + nz - 1] : mk<T>(0, 0);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int l16 = lane & 15, k4 = lane >> 4;
    const int NTL = nzp / 16;
    typename Mfma<T>::acc acc[DIAG_MAXNT][MT];
    for (int pass = 0; pass < 2; ++pass) {
        const T *Bm = pass == 0 ? E.Qp : E.QTp;
#pragma unroll
        for (int t = 0; t < DIAG_MAXNT; ++t)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[t][mt] = typename Mfma<T>::acc{0, 0, 0, 0};
        for (int kk = 0; kk < nzp / 4; ++kk) {
            const int k = kk * 4 + k4;
            T a[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) a[mt] = buf[(mt * 16 + l16) * S + k];
#pragma unroll
            for (int t = 0; t < DIAG_MAXNT; ++t) {
                const int nt = wave + 4 * t;
                if (nt < NTL) {
                    const T bv = Bm[(long long)k * nzp + nt * 16 + l16];
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt) acc[t][mt] = Mfma<T>::op(a[mt], bv, acc[t][mt]);
                }
            }
        }
        __syncthreads();  // every wavefront has finished reading buf
#pragma unroll
        for (int t = 0; t < DIAG_MAXNT; ++t) {
            const int nt = wave + 4 * t;
            if (nt >= NTL) continue;
            const int col = nt * 16 + l16;
            const double lm = col < nz ? E.lam[col] : 0.0;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = mt * 16 + Mfma<T>::row(k4, v);
                    T val = acc[t][mt][v];
                    if (pass == 0) {
                        const int r = row < TJ ? row : row - TJ;
                        const double d = s_e[r] + lm;
                        val = col < nz ? (T)(inverse ? (double)val / d : (double)val * d) : T(0);
                    }
                    buf[row * S + col] = val;
                }
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < rows * nz; idx += blockDim.x) {
        const int r = idx / nz, z = idx - r * nz;
        const long long g = (off + j0 + r) * nz + z;
        cx<T> v = mk(buf[r * S + z], buf[(TJ + r) * S + z]);
        if (inverse) {
            const T *fj = fac + (off + j0 + r) * nz;
            const cx<T> cj = s_c[r];
            const cx<T> den = mk<T>(1, 0) - cj * fj[nz - 1];
            const cx<T> yN = mk(buf[r * S + nz - 1], buf[(TJ + r) * S + nz - 1]);
            v = v + cdiv(yN * cj, den) * fj[z];
        } else if (z == nz - 1) {
            v = v - s_c[r] * s_x0[r];
        }
        y[g] = base ? base[g] + v * sign : v;
    }
}

template <typename T, int TJ>
static void launch_diag_mfma(const Layout &L, const EigTables<T> &E, const double *e,
                             const cx<T> *c, const T *fac, const cx<T> *x, const cx<T> *base,
                             cx<T> *y, T sign, int inverse, const int *active, hipStream_t s) {
    const size_t shm = sizeof(T) * 2 * TJ * (size_t)lds_stride_a(E.nzp);
    set_dynamic_lds((const void *)diag_mfma_kernel<T, TJ>, shm);
    dim3 grid((L.max_nc + TJ - 1) / TJ, L.n_inst);
    hipLaunchKernelGGL((diag_mfma_kernel<T, TJ>), grid, dim3(256), shm, s, L, E, e, c, fac, x, base,
                       y, sign, inverse, active);
    MSX_LAUNCH_CHECK();
}

bool diag_mfma_supported(int nz) { return nz <= 4 * DIAG_MAXNT * 16; }

template <typename T>
void launch_diag_blocks(const Layout &L, const EigTables<T> &E, const double *e, const cx<T> *c,
                        const T *fac, const cx<T> *x, const cx<T> *base, cx<T> *y, T sign,
                        int inverse, const int *active, hipStream_t s) {
    if (sizeof(T) == 4) launch_diag_mfma<T, 32>(L, E, e, c, fac, x, base, y, sign, inverse, active, s);
    else launch_diag_mfma<T, 16>(L, E, e, c, fac, x, base, y, sign, inverse, active, s);
}
template void launch_diag_blocks<float>(const Layout &, const EigTables<float> &, const double *,
                                        const cx<float> *, const float *, const cx<float> *,
                                        const cx<float> *, cx<float> *, float, int, const int *,
                                        hipStream_t);
template void launch_diag_blocks<double>(const Layout &, const EigTables<double> &,
                                         const double *, const cx<double> *, const double *,
                                         const cx<double> *, const cx<double> *, cx<double> *,
                                         double, int, const int *, hipStream_t);

// Zero-padded copies: dst[i][j] = src[i][j] for i, j < n, 0 up to np.
template <typename T>
__global__ void pad_square_kernel(int n, int np, const double *__restrict__ src, T *__restrict__ dst) {
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < np * np; idx += gridDim.x * blockDim.x) {
        const int i = idx / np, j = idx - i * np;
        dst[idx] = (i < n && j < n) ? (T)src[(long long)i * n + j] : T(0);
    }
}

void launch_pad_square(int n, int np, const double *src, double *dst, hipStream_t s) {
    hipLaunchKernelGGL(pad_square_kernel<double>, dim3(256), dim3(256), 0, s, n, np, src, dst);
    MSX_LAUNCH_CHECK();
}
void launch_pad_square(int n, int np, const double *src, float *dst, hipStream_t s) {
    hipLaunchKernelGGL(pad_square_kernel<float>, dim3(256), dim3(256), 0, s, n, np, src, dst);
    MSX_LAUNCH_CHECK();
}

}  // namespace msx
