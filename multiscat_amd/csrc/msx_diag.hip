// All-channel diagonal blocks on the matrix cores.
//
//   inverse: y_j = D_j^-1 x_j = Q diag(1/(e_j + lam)) Q^T x_j + f_j g_j(N) c_j / (1 - f_j(N) c_j)
//   forward: y_j = D_j x_j    = Q diag(e_j + lam) Q^T x_j - c_j x_j(N) e_N
// (the reference applies D_j^-1 channel by channel inside its forward substitution,
// lib/multiscat_fortran/fortran/multiscat_gmres.f90:666-685; its SciPy backend applies D^-1
// to all channels at once, multiscat/multiscat/_scipy.py:249-259,336-354).
//
// Every wavefront owns a column tile of 16 real rows: the real and imaginary parts of 16
// channels (FP32, two column tiles) or of 8 channels (FP64, one tile). It computes
//   W^T = Q^T X^T        (M = m, K = z, N = the tile's rows; 16x16x4 MFMA)
// scales W^T by 1/(e_j + lam_m) in registers (e and lam in FP64, so the denominator never
// cancels in FP32), and then
//   Y^T = Q W^T          (M = z, K = m)
// with the accumulators of the first product as the B operand of the second: register v of
// lane group g of m tile mt holds W^T[16 mt + 4 g + v] (the A rows of the first product are
// permuted, Mfma::perm, so that this holds for both accumulator layouts), and the second
// product takes the K index of lane group g in k-step (mt, v) to be m = 16 mt + 4 g + v,
// reading the matching four consecutive columns of Q as its A operand. No LDS, no barrier:
// the A fragments of Q and Q^T (zero-padded to nzp, a multiple of 16) are 16-byte loads that
// stay in L2, shared by all wavefronts.
#include "msx_mfma.h"
#include "msx_internal.h"

namespace msx {

constexpr int DG_WAVES = 4;  // wavefronts per workgroup (independent)

// CC consecutive elements p[0..CC) of one 16-byte load (CC = 4 floats or 2 doubles).
template <int CC>
__device__ __forceinline__ void load16(const float *p, float (&a)[CC]) {
    static_assert(CC == 4, "");
    const float4 v = *reinterpret_cast<const float4 *>(p);
    a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w;
}
template <int CC>
__device__ __forceinline__ void load16(const double *p, double (&a)[CC]) {
    static_assert(CC == 2, "");
    const double2 v = *reinterpret_cast<const double2 *>(p);
    a[0] = v.x; a[1] = v.y;
}

// NR = 2: column j of the tile pair is channel j (tile 0 real parts, tile 1 imaginary parts).
// NR = 1: column j is channel j / 2, part j % 2. NT: tiles per dimension, a compile-time bound
// >= nzp / 16 so that the inner loops unroll with static register indices. Tiles past
// nzp / 16 read clamped (valid) addresses of Q and Q^T; the scaling is 0 for m >= nz, so they
// add nothing, and their outputs are never stored.
//
// Both products run as software pipelines over stages of CC k-steps (one 16-byte load per A
// fragment): the loads of stage s + 1 are issued into the other register buffer before the
// MFMAs of stage s, and scheduling barriers keep the compiler from sinking them.
template <typename T, int NR, int NT>
__global__ __launch_bounds__(64 * DG_WAVES, NT <= 10 ? 2 : 1) void diag_gemm_kernel(
    Layout L, EigTables<T> E, const double *__restrict__ e, const cx<T> *__restrict__ c,
    const T *__restrict__ fac, const cx<T> *x, const cx<T> *base, cx<T> *y, T sign, int inverse,
    const int *__restrict__ active) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int CH = NR == 2 ? 16 : 8;    // channels per wavefront
    // z tiles per group of the second product: a divisor of NT where the registers allow
    constexpr int ZG = NT % 4 == 0 ? 4 : ((NR == 2 || sizeof(T) == 8) ? 2 : (NT % 5 == 0 ? 5 : 4));
    constexpr int CC = 16 / sizeof(T);      // k-steps per stage
    constexpr int NH = 4 / CC;              // stages per 16-wide K block
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int nc = L.nc[b];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j0 = (blockIdx.x * DG_WAVES + wave) * CH;
    if (j0 >= nc) return;
    const int nz = L.nz, nzp = E.nzp, NTL = nzp >> 4;
    const int l16 = lane & 15, g = lane >> 4;
    const int ch = NR == 2 ? l16 : (l16 >> 1);
    const int part = NR == 2 ? 0 : (l16 & 1);
    const int jc = j0 + ch;
    const bool live = jc < nc;
    const long long cidx = L.off[b] + (live ? jc : 0);
    const long long row0 = cidx * nz;
    const int pa = M::perm(l16);
    auto tile = [&](int t) { return t < NTL ? t : NTL - 1; };

    // ---- W^T = Q^T X^T ------------------------------------------------------------------
    acc_t acc1[NT][NR];
#pragma unroll
    for (int mt = 0; mt < NT; ++mt)
#pragma unroll
        for (int r = 0; r < NR; ++r) acc1[mt][r] = acc_t{0, 0, 0, 0};
    T a1[2][NT][CC], x1[2][NR][CC];
    // stage st: K block st / NH, k-steps (st % NH) * CC .. + CC; blocks past nzp / 16 load
    // zeros for X
    auto load1 = [&](int st, T (&a)[NT][CC], T (&xb)[NR][CC]) {
        const int s = st / NH, h = st - s * NH;
        const int zb = 16 * s + 4 * g + h * CC;
        const int zq = 16 * tile(s) + 4 * g + h * CC;
#pragma unroll
        for (int cc = 0; cc < CC; ++cc) {
            const int z = zb + cc;
            const bool ok = live && z < nz;
            const cx<T> v = x[row0 + (ok ? z : 0)];
            if constexpr (NR == 2) {
                xb[0][cc] = ok ? v.re : T(0);
                xb[1][cc] = ok ? v.im : T(0);
            } else {
                xb[0][cc] = ok ? (part ? v.im : v.re) : T(0);
            }
        }
#pragma unroll
        for (int mt = 0; mt < NT; ++mt)
            load16<CC>(E.QTp + (long long)(16 * tile(mt) + pa) * nzp + zq, a[mt]);
    };
    load1(0, a1[0], x1[0]);
#pragma unroll
    for (int st = 0; st < NT * NH; ++st) {
        const int cur = st & 1;
        if (st + 1 < NT * NH) load1(st + 1, a1[cur ^ 1], x1[cur ^ 1]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int cc = 0; cc < CC; ++cc)
#pragma unroll
            for (int mt = 0; mt < NT; ++mt)
#pragma unroll
                for (int r = 0; r < NR; ++r)
                    acc1[mt][r] = M::op(a1[cur][mt][cc], x1[cur][r][cc], acc1[mt][r]);
        __builtin_amdgcn_sched_barrier(0);
    }

    // ---- scale by 1/(e_j + lam_m) (or e_j + lam_m); y0(N) = Q[N-1][:] W for the correction
    const double ej = e[cidx];
    T yn[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) yn[r] = T(0);
#pragma unroll
    for (int mt = 0; mt < NT; ++mt) {
        const int m0 = 16 * mt + 4 * g;
        T qh[NH][CC];
#pragma unroll
        for (int h = 0; h < NH; ++h)
            load16<CC>(E.Qp + (long long)(nz - 1) * nzp + 16 * tile(mt) + 4 * g + h * CC, qh[h]);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const T qv = qh[v / CC][v % CC];
            const int m = m0 + v;
            const double d = ej + E.lam[m < nz ? m : nz - 1];
            const T sc = m < nz ? (inverse ? T(1) / T(d) : T(d)) : T(0);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                acc1[mt][r][v] *= sc;
                yn[r] += qv * acc1[mt][r][v];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        yn[r] += __shfl_xor(yn[r], 16);
        yn[r] += __shfl_xor(yn[r], 32);
    }
    // Sherman-Morrison coefficient (inverse) or the boundary term c_j x_j(N) (forward)
    cx<T> corr = mk<T>(0, 0);
    {
        T other = T(0);
        if constexpr (NR == 1) other = __shfl_xor(yn[0], 1);
        if (live) {
            const cx<T> cj = c[cidx];
            if (inverse) {
                cx<T> y0n;
                if constexpr (NR == 2) y0n = mk(yn[0], yn[1]);
                else y0n = part ? mk(other, yn[0]) : mk(yn[0], other);
                const cx<T> den = mk<T>(1, 0) - cj * fac[row0 + nz - 1];
                corr = cdiv(y0n * cj, den);
            } else {
                corr = cj * x[row0 + nz - 1];
            }
        }
    }

    // ---- Y^T = Q W^T, ZG z tiles at a time, then the correction and the store -------------
    for (int zt0 = 0; zt0 < NTL; zt0 += ZG) {
        acc_t acc2[ZG][NR];
#pragma unroll
        for (int zg = 0; zg < ZG; ++zg)
#pragma unroll
            for (int r = 0; r < NR; ++r) acc2[zg][r] = acc_t{0, 0, 0, 0};
        T a2[2][ZG][CC];
        // stage st: m tile st / NH, registers v = (st % NH) * CC .. + CC of acc1
        auto load2 = [&](int st, T (&a)[ZG][CC]) {
            const int mt = st / NH, h = st - mt * NH;
#pragma unroll
            for (int zg = 0; zg < ZG; ++zg)
                load16<CC>(E.Qp + (long long)(16 * tile(zt0 + zg) + pa) * nzp + 16 * tile(mt) +
                               4 * g + h * CC,
                           a[zg]);
        };
        load2(0, a2[0]);
#pragma unroll
        for (int st = 0; st < NT * NH; ++st) {
            const int cur = st & 1, mt = st / NH, h = st % NH;
            if (st + 1 < NT * NH) load2(st + 1, a2[cur ^ 1]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int cc = 0; cc < CC; ++cc)
#pragma unroll
                for (int zg = 0; zg < ZG; ++zg)
#pragma unroll
                    for (int r = 0; r < NR; ++r)
                        acc2[zg][r] = M::op(a2[cur][zg][cc], acc1[mt][r][h * CC + cc], acc2[zg][r]);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int zg = 0; zg < ZG; ++zg) {
            if (!live || zt0 + zg >= NTL) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int z = 16 * (zt0 + zg) + 4 * g + v;
                if (z >= nz) continue;
                const long long idx = row0 + z;
                if constexpr (NR == 2) {
                    cx<T> val = mk(acc2[zg][0][v], acc2[zg][1][v]);
                    if (inverse) val = val + corr * fac[idx];
                    else if (z == nz - 1) val = val - corr;
                    y[idx] = base ? base[idx] + val * sign : val;
                } else {
                    T val = acc2[zg][0][v];
                    const T cp = part ? corr.im : corr.re;
                    if (inverse) val += cp * fac[idx];
                    else if (z == nz - 1) val -= cp;
                    T *yp = reinterpret_cast<T *>(y + idx) + part;
                    *yp = base ? reinterpret_cast<const T *>(base + idx)[part] + val * sign : val;
                }
            }
        }
    }
}

template <typename T, int NR, int NT>
static void launch_diag_gemm(const Layout &L, const EigTables<T> &E, const double *e,
                             const cx<T> *c, const T *fac, const cx<T> *x, const cx<T> *base,
                             cx<T> *y, T sign, int inverse, const int *active, hipStream_t s) {
    constexpr int CH = NR == 2 ? 16 : 8;
    if (E.nzp % 16 != 0 || E.nzp > 16 * NT || E.nzp < L.nz)
        throw Error(3, "diagonal blocks: padded size does not fit the kernel");
    dim3 grid((L.max_nc + CH * DG_WAVES - 1) / (CH * DG_WAVES), L.n_inst);
    hipLaunchKernelGGL((diag_gemm_kernel<T, NR, NT>), grid, dim3(64 * DG_WAVES), 0, s, L, E, e,
                       c, fac, x, base, y, sign, inverse, active);
    MSX_LAUNCH_CHECK();
}

bool diag_mfma_supported(int nz) { return nz <= 320; }

template <typename T>
void launch_diag_blocks(const Layout &L, const EigTables<T> &E, const double *e, const cx<T> *c,
                        const T *fac, const cx<T> *x, const cx<T> *base, cx<T> *y, T sign,
                        int inverse, const int *active, hipStream_t s) {
    if (!diag_mfma_supported(L.nz)) throw Error(3, "diagonal blocks on the matrix cores: nz <= 320");
    // the smallest instantiated tile count that covers nzp / 16
    const int nt = E.nzp / 16;
#define MSX_DIAG(NR_, NT_)                                                                  \
    if (nt <= NT_) {                                                                        \
        launch_diag_gemm<T, NR_, NT_>(L, E, e, c, fac, x, base, y, sign, inverse, active, s); \
        return;                                                                             \
    }
    if constexpr (sizeof(T) == 4) {
        MSX_DIAG(2, 4) MSX_DIAG(2, 7) MSX_DIAG(2, 8) MSX_DIAG(2, 10)
    } else {
        MSX_DIAG(1, 4) MSX_DIAG(1, 7) MSX_DIAG(1, 8) MSX_DIAG(1, 10)
    }
    MSX_DIAG(1, 13) MSX_DIAG(1, 16) MSX_DIAG(1, 19) MSX_DIAG(1, 20)
#undef MSX_DIAG
    throw Error(3, "diagonal blocks on the matrix cores: nz <= 320");
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
