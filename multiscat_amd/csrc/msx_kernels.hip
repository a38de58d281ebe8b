// Operator kernels of the close-coupling solve on gfx950 (64-wide wavefronts).
//
//   coupling        y_j(z) = sum_i V[G_j - G_i](z) x_i(z) over i>j / i!=j / i<j
//                   (lib/multiscat_fortran/fortran/multiscat_gmres.f90:620-639, 654-660)
//   lower solve     x <- L^-1 x: forward substitution over channels with the eigenbasis
//                   solve and the Sherman-Morrison boundary fix (multiscat_gmres.f90:641-687)
//   dinv / diag     D^-1 x and D x for all channels at once (multiscat/multiscat/_scipy.py:249-354)
//   eigensystem     H0 = Q diag(lam) Q^T, ascending (multiscat_gmres.f90:247-253; diagsub.f90)
//   factors         f_j = (H0 + e_j)^-1 e_N (multiscat_gmres.f90:255-271)
//   intensities     P_j = |2i b_j psi_j(N) + a_0 delta_spec|^2 (multiscat/multiscat/_multiscat.py:40-60)
//   gradient        dL/dV_G(z) from the forward and adjoint states
//
// Work decomposition used throughout: the 64 lanes of a wavefront run over the node index z
// (contiguous in memory, so every load is coalesced) and everything that is uniform across
// the wavefront -- channel indices, G differences, table offsets -- is scalar work.
#include <algorithm>

#include "msx_internal.h"

namespace msx {

__device__ __forceinline__ int wrap_mod(int d, int n) { return d < 0 ? d + n : d; }

// ---------------------------------------------------------------------------------------
// Coupling sum. Block = 256 threads = 4 waves; each wave owns RJ consecutive channels j and
// 64 nodes z; loops over the source channels i. V[(dx*nky+dy)*nz + z] is a coalesced row.
// ---------------------------------------------------------------------------------------
template <typename T, int RJ>
__global__ __launch_bounds__(256) void coupling_kernel(Layout L, const cx<T> *__restrict__ V,
                                                       const cx<T> *__restrict__ x,
                                                       cx<T> *__restrict__ y, int mode,
                                                       const int *__restrict__ active) {
    const int b = blockIdx.z;
    if (active && !active[b]) return;
    const int nc = L.nc[b];
    const long long off = L.off[b];
    const int nz = L.nz, nkx = L.nkx, nky = L.nky;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j0 = (blockIdx.y * 4 + wave) * RJ;
    if (j0 >= nc) return;
    const int z = blockIdx.x * 64 + lane;
    const int zl = z < nz ? z : nz - 1;  // clamp loads; stores are guarded
    const int *chx = L.chx + off, *chy = L.chy + off;
    const cx<T> *xb = x + off * nz;

    int gxj[RJ], gyj[RJ];
    cx<T> acc[RJ];
#pragma unroll
    for (int r = 0; r < RJ; ++r) {
        const int j = j0 + r < nc ? j0 + r : nc - 1;
        gxj[r] = chx[j];
        gyj[r] = chy[j];
        acc[r] = mk<T>(0, 0);
    }
    const int jlast = (j0 + RJ < nc ? j0 + RJ : nc) - 1;
    int ibeg = 0, iend = nc;
    if (mode == COUPLE_UPPER) ibeg = j0 + 1;
    if (mode == COUPLE_LOWER) iend = jlast;  // i < j <= jlast
    for (int i = ibeg; i < iend; ++i) {
        const cx<T> xi = xb[(long long)i * nz + zl];
        const int gxi = chx[i], gyi = chy[i];
#pragma unroll
        for (int r = 0; r < RJ; ++r) {
            const int j = j0 + r;
            bool use = j < nc;
            if (mode == COUPLE_UPPER) use = use && (i > j);
            else if (mode == COUPLE_LOWER) use = use && (i < j);
            else use = use && (i != j);
            if (use) {
                const int dx = wrap_mod(gxj[r] - gxi, nkx);
                const int dy = wrap_mod(gyj[r] - gyi, nky);
                const cx<T> v = V[((long long)dx * nky + dy) * nz + zl];
                cfma(acc[r], v, xi);
            }
        }
    }
    if (z < nz) {
#pragma unroll
        for (int r = 0; r < RJ; ++r) {
            const int j = j0 + r;
            if (j < nc) y[(off + j) * nz + z] = acc[r];
        }
    }
}

template <typename T>
void launch_coupling(const Layout &L, const cx<T> *V, const cx<T> *x, cx<T> *y, int mode,
                     const int *active, hipStream_t s) {
    constexpr int RJ = 4;
    if (L.nm) throw Error(3, "direct coupling: channel-major vectors only");
    dim3 grid((L.nz + 63) / 64, (L.max_nc + 4 * RJ - 1) / (4 * RJ), L.n_inst);
    hipLaunchKernelGGL((coupling_kernel<T, RJ>), grid, dim3(256), 0, s, L, V, x, y, mode, active);
    MSX_LAUNCH_CHECK();
}
template void launch_coupling<double>(const Layout &, const cx<double> *, const cx<double> *,
                                      cx<double> *, int, const int *, hipStream_t);
template void launch_coupling<float>(const Layout &, const cx<float> *, const cx<float> *,
                                     cx<float> *, int, const int *, hipStream_t);

// Full coupling of a state that is nonzero only on each instance's specular channel s:
// y_j(z) = V[G_j - G_s](z) x_s(z) for j != s, y_s = 0. A wavefront per channel, lanes over z
// (channel-major vectors), or lanes over channels (node-major).
template <typename T>
__global__ __launch_bounds__(256) void specular_coupling_kernel(Layout L, const cx<T> *__restrict__ V,
                                                                const int *__restrict__ spec,
                                                                const cx<T> *__restrict__ x,
                                                                cx<T> *__restrict__ y,
                                                                const int *__restrict__ active) {
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int nc = L.nc[b], nz = L.nz, s = spec[b];
    const long long off = L.off[b];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gxs = L.chx[off + s], gys = L.chy[off + s];
    if (L.nm) {
        // node-major vectors: lanes over channels, workgroups over nodes
        for (int j = threadIdx.x; j < nc; j += blockDim.x) {
            const int dx = wrap_mod(L.chx[off + j] - gxs, L.nkx);
            const int dy = wrap_mod(L.chy[off + j] - gys, L.nky);
            const cx<T> *v = V + ((long long)dx * L.nky + dy) * nz;
            for (int z = blockIdx.x; z < nz; z += gridDim.x) {
                const long long row = off * nz + (long long)z * nc;
                y[row + j] = j == s ? mk<T>(0, 0) : v[z] * x[row + s];
            }
        }
        return;
    }
    const cx<T> *xs = x + (off + s) * nz;
    for (int j = blockIdx.x * 4 + wave; j < nc; j += gridDim.x * 4) {
        const int dx = wrap_mod(L.chx[off + j] - gxs, L.nkx);
        const int dy = wrap_mod(L.chy[off + j] - gys, L.nky);
        const cx<T> *v = V + ((long long)dx * L.nky + dy) * nz;
        cx<T> *yj = y + (off + j) * nz;
        for (int z = lane; z < nz; z += 64) yj[z] = j == s ? mk<T>(0, 0) : v[z] * xs[z];
    }
}

template <typename T>
void launch_specular_coupling(const Layout &L, const cx<T> *V, const int *spec, const cx<T> *x,
                              cx<T> *y, const int *active, hipStream_t s) {
    const unsigned bx = (unsigned)std::max(1, std::min((L.max_nc + 3) / 4, 64));
    hipLaunchKernelGGL((specular_coupling_kernel<T>), dim3(bx, L.n_inst), dim3(256), 0, s, L, V,
                       spec, x, y, active);
    MSX_LAUNCH_CHECK();
}
template void launch_specular_coupling<float>(const Layout &, const cx<float> *, const int *,
                                              const cx<float> *, cx<float> *, const int *,
                                              hipStream_t);

// ---------------------------------------------------------------------------------------
// Block forward substitution x <- L^-1 x in channel blocks [j0, j1) of LS_BLOCK channels:
// lower_solve_kernel solves one block, one workgroup (8 waves) per instance, sequential over
// its channels, with the coupling to the earlier channels of the same block; before it,
// lower_update_kernel has subtracted the coupling to every earlier block from the block's
// right-hand side (all later channels at once, many workgroups). The chain of dependent
// per-channel steps stays the reference's, but the sum over earlier channels, most of each
// step's work, leaves the sequential kernel. Per channel j of the block:
//   (1) r(z) = x_j(z) - sum_{j0<=i<j} V[G_j-G_i](z) psi_i(z)   waves split i, lanes run over z
//   (2) t_m = sum_z r(z) Q[z][m] / (e_j + lam_m)           waves split z, lanes run over m
//   (3) g(z) = sum_m QT[m][z] t_m                          waves split m, lanes run over z
//   (4) psi_j = g + f_j g(N) c_j / (1 - f_j(N) c_j)
// Partial sums of the waves meet in LDS. psi_i (i < j) are re-read from global memory: they
// were written by this workgroup, and __syncthreads orders those writes before the reads.
// ---------------------------------------------------------------------------------------
constexpr int LS_WAVES = 8;
constexpr int LS_MAXMZ = 9;  // nz <= 576 (LDS: 17 vectors of nz complex values)
constexpr int LS_BLOCK = 32;

// MZ = ceil(nz / 64) values per lane; ZB rows of Q / Q^T loaded together. Two reduction
// buffers alternate so that a step reads the previous step's partial sums while writing its
// own, and each wavefront sums the partials of its own rows into a shared vector (vs) that
// it alone reads next: four workgroup barriers per channel.
template <int MZ>
__global__ __launch_bounds__(512) void lower_solve_kernel(
    Layout L, const cx<double> *__restrict__ V, const double *__restrict__ Q,
    const double *__restrict__ QT, const double *__restrict__ lam, const double *__restrict__ e,
    const cx<double> *__restrict__ c, const double *__restrict__ fac, cx<double> *x,
    const cx<double> *base, cx<double> *out, double sign, const int *__restrict__ active,
    int j0, int j1) {
    constexpr int ZB = MZ <= 4 ? 8 : 4;
    const int b = blockIdx.x;
    if (active && !active[b]) return;
    extern __shared__ double smem[];
    const int nz = L.nz, nkx = L.nkx, nky = L.nky;
    cx<double> *redA = reinterpret_cast<cx<double> *>(smem);  // [LS_WAVES][nz]
    cx<double> *redB = redA + LS_WAVES * nz;                  // [LS_WAVES][nz]
    cx<double> *vs = redB + LS_WAVES * nz;                    // [nz]

    const int nc = L.nc[b];
    const long long off = L.off[b];
    const int *chx = L.chx + off, *chy = L.chy + off;
    cx<double> *xb = x + off * nz;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int zchunk = (nz + LS_WAVES - 1) / LS_WAVES;
    const int rb = wave * zchunk, re = min(nz, rb + zchunk);  // this wavefront's rows
    const int jend = min(j1, nc);
    for (int j = j0; j < jend; ++j) {
        // (1) coupling to the already solved channels of this block -> redA
        cx<double> acc[MZ];
#pragma unroll
        for (int m = 0; m < MZ; ++m) acc[m] = mk(0.0, 0.0);
        const int gxj = chx[j], gyj = chy[j];
        for (int i = j0 + wave; i < j; i += LS_WAVES) {
            const int dx = wrap_mod(gxj - chx[i], nkx);
            const int dy = wrap_mod(gyj - chy[i], nky);
            const cx<double> *vrow = V + ((long long)dx * nky + dy) * nz;
            const cx<double> *xi = xb + (long long)i * nz;
#pragma unroll
            for (int m = 0; m < MZ; ++m) {
                const int z = lane + 64 * m;
                if (z < nz) cfma(acc[m], vrow[z], xi[z]);
            }
        }
#pragma unroll
        for (int m = 0; m < MZ; ++m) {
            const int z = lane + 64 * m;
            if (z < nz) redA[wave * nz + z] = acc[m];
        }
        __syncthreads();
        // (2) r = x_j - sum_w redA over this wavefront's rows -> vs; t = Q^T r (rows z of this
        // wavefront, lanes over m) -> redB
        const cx<double> *xj = xb + (long long)j * nz;
        for (int z = rb + lane; z < re; z += 64) {
            cx<double> r = xj[z];
#pragma unroll
            for (int w = 0; w < LS_WAVES; ++w) r = r - redA[w * nz + z];
            vs[z] = r;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this wavefront's vs stores landed
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int m = 0; m < MZ; ++m) acc[m] = mk(0.0, 0.0);
        for (int z0 = rb; z0 < re; z0 += ZB) {
            double q[ZB][MZ];
            cx<double> r[ZB];
#pragma unroll
            for (int u = 0; u < ZB; ++u) {
                const int z = z0 + u;
                const bool ok = z < re;
                const int zc = ok ? z : rb;
                r[u] = ok ? vs[zc] : mk(0.0, 0.0);
                const double *qrow = Q + (long long)zc * nz;
#pragma unroll
                for (int m = 0; m < MZ; ++m) {
                    const int mm = lane + 64 * m;
                    q[u][m] = mm < nz ? qrow[mm] : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < ZB; ++u)
#pragma unroll
                for (int m = 0; m < MZ; ++m) {
                    acc[m].re = fma(r[u].re, q[u][m], acc[m].re);
                    acc[m].im = fma(r[u].im, q[u][m], acc[m].im);
                }
        }
#pragma unroll
        for (int m = 0; m < MZ; ++m) {
            const int mm = lane + 64 * m;
            if (mm < nz) redB[wave * nz + mm] = acc[m];
        }
        __syncthreads();
        // (3) t_m = sum_w redB / (e_j + lam_m) over this wavefront's eigenvectors -> vs;
        // g = Q t (eigenvectors of this wavefront, lanes over z) -> redA
        const double ej = e[off + j];
        for (int mm = rb + lane; mm < re; mm += 64) {
            cx<double> t = mk(0.0, 0.0);
#pragma unroll
            for (int w = 0; w < LS_WAVES; ++w) t = t + redB[w * nz + mm];
            vs[mm] = t * (1.0 / (ej + lam[mm]));
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int m = 0; m < MZ; ++m) acc[m] = mk(0.0, 0.0);
        for (int m0 = rb; m0 < re; m0 += ZB) {
            double q[ZB][MZ];
            cx<double> t[ZB];
#pragma unroll
            for (int u = 0; u < ZB; ++u) {
                const int mm = m0 + u;
                const bool ok = mm < re;
                const int mc = ok ? mm : rb;
                t[u] = ok ? vs[mc] : mk(0.0, 0.0);
                const double *qtrow = QT + (long long)mc * nz;
#pragma unroll
                for (int m = 0; m < MZ; ++m) {
                    const int z = lane + 64 * m;
                    q[u][m] = z < nz ? qtrow[z] : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < ZB; ++u)
#pragma unroll
                for (int m = 0; m < MZ; ++m) {
                    acc[m].re = fma(t[u].re, q[u][m], acc[m].re);
                    acc[m].im = fma(t[u].im, q[u][m], acc[m].im);
                }
        }
#pragma unroll
        for (int m = 0; m < MZ; ++m) {
            const int z = lane + 64 * m;
            if (z < nz) redA[wave * nz + z] = acc[m];
        }
        __syncthreads();
        // (4) g = sum_w redA; the Sherman-Morrison boundary correction (every thread forms the
        // factor from g(N) itself: eight broadcast reads, no extra barrier)
        const double *fj = fac + (off + j) * nz;
        cx<double> gN = mk(0.0, 0.0);
#pragma unroll
        for (int w = 0; w < LS_WAVES; ++w) gN = gN + redA[w * nz + nz - 1];
        const cx<double> cj = c[off + j];
        const cx<double> f = cdiv(gN * cj, mk(1.0, 0.0) - cj * fj[nz - 1]);
        for (int z = threadIdx.x; z < nz; z += blockDim.x) {
            cx<double> g = mk(0.0, 0.0);
#pragma unroll
            for (int w = 0; w < LS_WAVES; ++w) g = g + redA[w * nz + z];
            const cx<double> psi = g + f * fj[z];
            const long long idx = (off + j) * nz + z;
            xb[(long long)j * nz + z] = psi;
            if (out) out[idx] = base[idx] + psi * sign;
        }
        __syncthreads();
    }
}

// x_k(z) -= sum_{j0<=i<j1} V[G_k - G_i](z) psi_i(z) for every channel k >= j1 of each active
// instance (psi_i = x_i, already solved): one wavefront per target channel, lanes over z.
constexpr int LU_WAVES = 4;
__global__ __launch_bounds__(64 * LU_WAVES) void lower_update_kernel(
    Layout L, const cx<double> *__restrict__ V, cx<double> *x, int j0, int j1,
    const int *__restrict__ active) {
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int nc = L.nc[b];
    const int k = j1 + blockIdx.x * LU_WAVES + (threadIdx.x >> 6);
    if (k >= nc) return;
    const int nz = L.nz, nkx = L.nkx, nky = L.nky, lane = threadIdx.x & 63;
    const long long off = L.off[b];
    const int *chx = L.chx + off, *chy = L.chy + off;
    cx<double> *xb = x + off * nz;
    const int gxk = chx[k], gyk = chy[k];
    const int iend = min(j1, nc);
    for (int z0 = 0; z0 < nz; z0 += 64 * 4) {
        cx<double> acc[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[m] = mk(0.0, 0.0);
        for (int i = j0; i < iend; ++i) {
            const int dx = wrap_mod(gxk - chx[i], nkx);
            const int dy = wrap_mod(gyk - chy[i], nky);
            const cx<double> *vrow = V + ((long long)dx * nky + dy) * nz;
            const cx<double> *xi = xb + (long long)i * nz;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int z = z0 + lane + 64 * m;
                if (z < nz) cfma(acc[m], vrow[z], xi[z]);
            }
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int z = z0 + lane + 64 * m;
            if (z < nz) xb[(long long)k * nz + z] = xb[(long long)k * nz + z] - acc[m];
        }
    }
}

template <int MZ>
static void launch_lower_block(const Layout &L, const cx<double> *V, const double *Q,
                               const double *QT, const double *lam, const double *e,
                               const cx<double> *c, const double *fac, cx<double> *x,
                               const cx<double> *base, cx<double> *out, double sign,
                               const int *active, hipStream_t s, int j0, int j1) {
    const size_t shm = sizeof(cx<double>) * (size_t)L.nz * (2 * LS_WAVES + 1);
    set_dynamic_lds((const void *)lower_solve_kernel<MZ>, shm);
    hipLaunchKernelGGL(lower_solve_kernel<MZ>, dim3(L.n_inst), dim3(64 * LS_WAVES), shm, s, L, V,
                       Q, QT, lam, e, c, fac, x, base, out, sign, active, j0, j1);
}

void launch_lower_solve(const Layout &L, const cx<double> *V, const double *Q, const double *QT,
                        const double *lam, const double *e, const cx<double> *c,
                        const double *fac, cx<double> *x, const cx<double> *base,
                        cx<double> *out, double sign, const int *active, hipStream_t s) {
    if (L.nz > 576) throw Error(3, "lower solve supports nz <= 576");
    const int mz = (L.nz + 63) / 64;
    for (int j0 = 0; j0 < L.max_nc; j0 += LS_BLOCK) {
        const int j1 = std::min(L.max_nc, j0 + LS_BLOCK);
        switch (mz) {
#define MSX_LS(M)                                                                             \
    case M:                                                                                   \
        launch_lower_block<M>(L, V, Q, QT, lam, e, c, fac, x, base, out, sign, active, s, j0, j1); \
        break;
            MSX_LS(1) MSX_LS(2) MSX_LS(3) MSX_LS(4) MSX_LS(5)
            MSX_LS(6) MSX_LS(7) MSX_LS(8) MSX_LS(9)
#undef MSX_LS
            default: throw Error(3, "lower solve supports nz <= 576");
        }
        MSX_LAUNCH_CHECK();
        if (j1 < L.max_nc) {
            const unsigned gx = (unsigned)((L.max_nc - j1 + LU_WAVES - 1) / LU_WAVES);
            hipLaunchKernelGGL(lower_update_kernel, dim3(gx, L.n_inst), dim3(64 * LU_WAVES), 0, s,
                               L, V, x, j0, j1, active);
            MSX_LAUNCH_CHECK();
        }
    }
}

// ---------------------------------------------------------------------------------------
// All-channel diagonal blocks. inverse = 1: y = D^-1 x with
//   D_j^-1 x = Q diag(1/(e_j+lam)) Q^T x + f_j (g(N) c_j / (1 - f_j(N) c_j)),
// inverse = 0: y = D x = Q diag(e_j+lam) Q^T x - c_j x(N) e_N.
// One block = TJ channels of one instance; the rows live in LDS, lanes run over the output
// index of each of the two GEMMs with Q.
// ---------------------------------------------------------------------------------------
template <typename T, int TJ>
__global__ __launch_bounds__(256) void diag_kernel(Layout L, const T *__restrict__ Q,
                                                   const T *__restrict__ QT,
                                                   const double *__restrict__ lam,
                                                   const double *__restrict__ e,
                                                   const cx<T> *__restrict__ c,
                                                   const T *__restrict__ fac, const cx<T> *x,
                                                   const cx<T> *base, cx<T> *y,
                                                   T sign, int inverse,
                                                   const int *__restrict__ active) {
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int nc = L.nc[b];
    const int j0 = blockIdx.x * TJ;
    if (j0 >= nc) return;
    const int nz = L.nz;
    const long long off = L.off[b];
    extern __shared__ double smem_raw[];
    cx<T> *Xs = reinterpret_cast<cx<T> *>(smem_raw);  // [TJ][nz]
    cx<T> *Ts = Xs + TJ * nz;                           // [TJ][nz]
    __shared__ double s_e[TJ];
    __shared__ cx<T> s_c[TJ], s_x0[TJ];
    const int rows = min(TJ, nc - j0);
    for (int idx = threadIdx.x; idx < TJ * nz; idx += blockDim.x) {
        const int r = idx / nz, z = idx - r * nz;
        Xs[idx] = r < rows ? x[(off + j0 + r) * nz + z] : mk<T>(0, 0);
    }
    if (threadIdx.x < TJ) {
        const int r = threadIdx.x;
        s_e[r] = r < rows ? e[off + j0 + r] : 0.0;
        s_c[r] = r < rows ? c[off + j0 + r] : mk<T>(0, 0);
    }
    __syncthreads();
    if (threadIdx.x < TJ) s_x0[threadIdx.x] = Xs[threadIdx.x * nz + nz - 1];
    for (int m = threadIdx.x; m < nz; m += blockDim.x) {
        cx<T> acc[TJ];
#pragma unroll
        for (int r = 0; r < TJ; ++r) acc[r] = mk<T>(0, 0);
        for (int z = 0; z < nz; ++z) {
            const T q = Q[(long long)z * nz + m];
#pragma unroll
            for (int r = 0; r < TJ; ++r) {
                const cx<T> xv = Xs[r * nz + z];
                acc[r].re = fma(xv.re, q, acc[r].re);
                acc[r].im = fma(xv.im, q, acc[r].im);
            }
        }
        const double lm = lam[m];
#pragma unroll
        for (int r = 0; r < TJ; ++r) {
            const double d = s_e[r] + lm;
            Ts[r * nz + m] = acc[r] * (T)(inverse ? 1.0 / d : d);
        }
    }
    __syncthreads();
    for (int z = threadIdx.x; z < nz; z += blockDim.x) {
        cx<T> acc[TJ];
#pragma unroll
        for (int r = 0; r < TJ; ++r) acc[r] = mk<T>(0, 0);
        for (int m = 0; m < nz; ++m) {
            const T q = QT[(long long)m * nz + z];
#pragma unroll
            for (int r = 0; r < TJ; ++r) {
                const cx<T> tv = Ts[r * nz + m];
                acc[r].re = fma(tv.re, q, acc[r].re);
                acc[r].im = fma(tv.im, q, acc[r].im);
            }
        }
#pragma unroll
        for (int r = 0; r < TJ; ++r) Xs[r * nz + z] = acc[r];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < rows * nz; idx += blockDim.x) {
        const int r = idx / nz, z = idx - r * nz;
        const long long g = (off + j0 + r) * nz + z;
        cx<T> v = Xs[idx];
        if (inverse) {
            const T *fj = fac + (off + j0 + r) * nz;
            const cx<T> cj = s_c[r];
            const cx<T> den = mk<T>(1, 0) - cj * fj[nz - 1];
            const cx<T> f = cdiv(Xs[r * nz + nz - 1] * cj, den);
            v = v + f * fj[z];
        } else if (z == nz - 1) {
            v = v - s_c[r] * s_x0[r];
        }
        y[g] = base ? base[g] + v * sign : v;
    }
}

template <typename T>
void launch_dinv_impl(const Layout &L, const T *Q, const T *QT, const double *lam,
                      const double *e,
                      const cx<T> *c, const T *fac, const cx<T> *x, const cx<T> *base, cx<T> *y,
                      T sign, int inverse, const int *active, hipStream_t s) {
    const size_t row_bytes = sizeof(cx<T>) * (size_t)L.nz * 2;
    if (row_bytes * 4 > 150 * 1024) throw Error(3, "diag kernel: nz too large");
    if (row_bytes * 8 <= 64 * 1024) {
        dim3 grid((L.max_nc + 7) / 8, L.n_inst);
        hipLaunchKernelGGL((diag_kernel<T, 8>), grid, dim3(256), row_bytes * 8, s, L, Q, QT, lam,
                           e, c, fac, x, base, y, sign, inverse, active);
    } else {
        set_dynamic_lds((const void *)diag_kernel<T, 4>, row_bytes * 4);
        dim3 grid((L.max_nc + 3) / 4, L.n_inst);
        hipLaunchKernelGGL((diag_kernel<T, 4>), grid, dim3(256), row_bytes * 4, s, L, Q, QT, lam,
                           e, c, fac, x, base, y, sign, inverse, active);
    }
    MSX_LAUNCH_CHECK();
}

template <typename T>
void launch_dinv(const Layout &L, const T *Q, const T *QT, const double *lam, const double *e,
                 const cx<T> *c, const T *fac, const cx<T> *x, const cx<T> *base, cx<T> *y,
                 T sign, const int *active, hipStream_t s) {
    if (L.nm) throw Error(3, "vector D^-1: channel-major vectors only");
    launch_dinv_impl<T>(L, Q, QT, lam, e, c, fac, x, base, y, sign, 1, active, s);
}
template void launch_dinv<double>(const Layout &, const double *, const double *,
                                  const double *, const double *, const cx<double> *,
                                  const double *, const cx<double> *, const cx<double> *,
                                  cx<double> *, double, const int *, hipStream_t);
template void launch_dinv<float>(const Layout &, const float *, const float *, const double *,
                                 const double *, const cx<float> *, const float *,
                                 const cx<float> *, const cx<float> *, cx<float> *, float,
                                 const int *, hipStream_t);

void launch_apply_diag(const Layout &L, const double *Q, const double *QT, const double *lam,
                       const double *e, const cx<double> *c, const cx<double> *x,
                       cx<double> *y, const int *active, hipStream_t s, const cx<double> *base) {
    if (L.nm) throw Error(3, "vector D: channel-major vectors only");
    launch_dinv_impl<double>(L, Q, QT, lam, e, c, nullptr, x, base, y, 1.0, 0, active, s);
}

// ---------------------------------------------------------------------------------------
// Preconditioner factors: one block per channel; g_m = Q[N-1][m] / (e_j + lam_m) in LDS,
// then fac_j(z) = sum_m QT[m][z] g_m with lanes over z.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void factors_kernel(long long total, int nz,
                                                      const double *__restrict__ Q,
                                                      const double *__restrict__ QT,
                                                      const double *__restrict__ lam,
                                                      const double *__restrict__ e,
                                                      double *__restrict__ fac) {
    extern __shared__ double g[];
    const long long j = blockIdx.x;
    if (j >= total) return;
    const double ej = e[j];
    for (int m = threadIdx.x; m < nz; m += blockDim.x)
        g[m] = Q[(long long)(nz - 1) * nz + m] / (ej + lam[m]);
    __syncthreads();
    for (int z = threadIdx.x; z < nz; z += blockDim.x) {
        double acc = 0;
        for (int m = 0; m < nz; ++m) acc = fma(QT[(long long)m * nz + z], g[m], acc);
        fac[j * nz + z] = acc;
    }
}

void launch_preconditioner_factors_qt(long long total_channels, int nz, const double *Q,
                                      const double *QT, const double *lam, const double *e,
                                      double *fac, hipStream_t s) {
    if (total_channels <= 0) return;
    hipLaunchKernelGGL(factors_kernel, dim3((unsigned)total_channels), dim3(256),
                       sizeof(double) * nz, s, total_channels, nz, Q, QT, lam, e, fac);
    MSX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Element-wise helpers.
// ---------------------------------------------------------------------------------------
__global__ void to_float_kernel(long long n, const cx<double> *__restrict__ src,
                                cx<float> *__restrict__ dst) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = mk((float)src[i].re, (float)src[i].im);
}
__global__ void to_float_real_kernel(long long n, const double *__restrict__ src,
                                     float *__restrict__ dst) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}
static unsigned ew_blocks(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}
void launch_to_float(long long n, const cx<double> *src, cx<float> *dst, hipStream_t s) {
    hipLaunchKernelGGL(to_float_kernel, dim3(ew_blocks(n)), dim3(256), 0, s, n, src, dst);
    MSX_LAUNCH_CHECK();
}
void launch_to_float_real(long long n, const double *src, float *dst, hipStream_t s) {
    hipLaunchKernelGGL(to_float_real_kernel, dim3(ew_blocks(n)), dim3(256), 0, s, n, src, dst);
    MSX_LAUNCH_CHECK();
}

// FP64 channel-major -> FP32 node-major, one 32-channel x 32-node tile per workgroup
// (blockIdx.x: tile, blockIdx.y: instance): reads run along z, writes along j, through LDS.
// Complex vectors and the real factor table alike.
constexpr int TT = 32;
__device__ __forceinline__ cx<float> narrow(const cx<double> &v) { return mk((float)v.re, (float)v.im); }
__device__ __forceinline__ float narrow(double v) { return (float)v; }
template <typename S, typename D>
__global__ __launch_bounds__(256) void to_float_nm_kernel(Layout L, const S *__restrict__ src,
                                                          D *__restrict__ dst) {
    __shared__ D tile[TT][TT + 1];
    const int b = blockIdx.y;
    const int nc = L.nc[b], nz = L.nz;
    const int tz = (nz + TT - 1) / TT;
    const int j0 = (blockIdx.x / tz) * TT, z0 = (blockIdx.x % tz) * TT;
    if (j0 >= nc) return;
    const long long off = L.off[b];
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;
    for (int jj = ty; jj < TT; jj += 256 / TT) {
        const int j = j0 + jj, z = z0 + tx;
        if (j < nc && z < nz) tile[jj][tx] = narrow(src[(off + j) * nz + z]);
    }
    __syncthreads();
    for (int zz = ty; zz < TT; zz += 256 / TT) {
        const int z = z0 + zz, j = j0 + tx;
        if (j < nc && z < nz) dst[off * nz + (long long)z * nc + j] = tile[tx][zz];
    }
}

template <typename S, typename D>
static void launch_nm(const Layout &L, const S *src, D *dst, hipStream_t s) {
    const unsigned tiles = (unsigned)(((L.max_nc + TT - 1) / TT) * ((L.nz + TT - 1) / TT));
    hipLaunchKernelGGL((to_float_nm_kernel<S, D>), dim3(std::max(tiles, 1u), L.n_inst), dim3(256), 0,
                       s, L, src, dst);
    MSX_LAUNCH_CHECK();
}

void launch_to_float_vec(const Layout &L, long long n, const cx<double> *src, cx<float> *dst,
                         hipStream_t s) {
    if (!L.nm) return launch_to_float(n, src, dst, s);
    launch_nm(L, src, dst, s);
}

void launch_to_float_real_nm(const Layout &L, const double *src, float *dst, hipStream_t s) {
    launch_nm(L, src, dst, s);
}

// FP64 state between the two vector orders of each instance block, one 32-channel x 32-node
// tile per workgroup through LDS: reads and writes both run along the contiguous index.
template <bool TO_NM>
__global__ __launch_bounds__(256) void state_order_kernel(Layout L, const cx<double> *__restrict__ src,
                                                          cx<double> *__restrict__ dst) {
    __shared__ cx<double> tile[TT][TT + 1];  // [j - j0][z - z0]
    const int b = blockIdx.y;
    const int nc = L.nc[b], nz = L.nz;
    const int tz = (nz + TT - 1) / TT;
    const int j0 = (blockIdx.x / tz) * TT, z0 = (blockIdx.x % tz) * TT;
    if (j0 >= nc) return;
    const long long cm = L.off[b] * nz, nmb = L.off[b] * nz;  // block start in both orders
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;
    if constexpr (TO_NM) {
        for (int jj = ty; jj < TT; jj += 256 / TT) {
            const int j = j0 + jj, z = z0 + tx;
            if (j < nc && z < nz) tile[jj][tx] = src[cm + (long long)j * nz + z];
        }
        __syncthreads();
        for (int zz = ty; zz < TT; zz += 256 / TT) {
            const int z = z0 + zz, j = j0 + tx;
            if (j < nc && z < nz) dst[nmb + (long long)z * nc + j] = tile[tx][zz];
        }
    } else {
        for (int zz = ty; zz < TT; zz += 256 / TT) {
            const int z = z0 + zz, j = j0 + tx;
            if (j < nc && z < nz) tile[tx][zz] = src[nmb + (long long)z * nc + j];
        }
        __syncthreads();
        for (int jj = ty; jj < TT; jj += 256 / TT) {
            const int j = j0 + jj, z = z0 + tx;
            if (j < nc && z < nz) dst[cm + (long long)j * nz + z] = tile[jj][tx];
        }
    }
}

void launch_state_order(const Layout &L, const cx<double> *src, cx<double> *dst, bool to_nm,
                        hipStream_t s) {
    if (src == dst) throw Error(3, "state order change: source and destination must differ");
    const unsigned tiles = (unsigned)(((L.max_nc + TT - 1) / TT) * ((L.nz + TT - 1) / TT));
    if (to_nm)
        hipLaunchKernelGGL(state_order_kernel<true>, dim3(std::max(tiles, 1u), L.n_inst), dim3(256),
                           0, s, L, src, dst);
    else
        hipLaunchKernelGGL(state_order_kernel<false>, dim3(std::max(tiles, 1u), L.n_inst), dim3(256),
                           0, s, L, src, dst);
    MSX_LAUNCH_CHECK();
}

__global__ void intensities_kernel(Layout L, const cx<double> *__restrict__ psi,
                                   const cx<double> *__restrict__ b,
                                   const cx<double> *__restrict__ a0,
                                   const int *__restrict__ spec, double *__restrict__ P) {
    const int inst = blockIdx.y;
    const int nc = L.nc[inst];
    const long long off = L.off[inst];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nc; j += gridDim.x * blockDim.x) {
        const long long g = off + j;
        const cx<double> two_i = mk(0.0, 2.0);
        cx<double> s = two_i * b[g] * psi[vec_index(L, off, nc, j, L.nz - 1)];
        if (j == spec[inst]) s = s + a0[inst];
        P[g] = norm2(s);
    }
}

void launch_intensities(const Layout &L, const cx<double> *psi, const cx<double> *b,
                        const cx<double> *a0, const int *spec, double *P, hipStream_t s) {
    dim3 grid((L.max_nc + 255) / 256, L.n_inst);
    hipLaunchKernelGGL(intensities_kernel, grid, dim3(256), 0, s, L, psi, b, a0, spec, P);
    MSX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Potential gradient. For table entry g = (gx, gy) != 0 and node z:
//   grad[g][z] = conj( -sum_b sum_j lam_j(z) psi_{i(j)}(z) ),  G_i = G_j - g (mod grid)
// and for g = 0 the real part of -sum_b sum_j lam_j(z) psi_j(z). chmap[b][g] gives the
// channel of instance b sitting at grid point g (-1 if none). One block per g, lanes over z.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gradient_kernel(Layout L, const int *__restrict__ chmap,
                                                       const cx<double> *__restrict__ lam_adj,
                                                       const cx<double> *__restrict__ psi,
                                                       cx<double> *__restrict__ grad) {
    const int g = blockIdx.y;
    const int nkx = L.nkx, nky = L.nky, nz = L.nz;
    const int gx = g / nky, gy = g - gx * nky;
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int zl = z < nz ? z : nz - 1;
    const int nG = nkx * nky;
    cx<double> acc = mk(0.0, 0.0);
    for (int b = 0; b < L.n_inst; ++b) {
        const int nc = L.nc[b];
        const long long off = L.off[b];
        const int *map = chmap + (long long)b * nG;
        for (int j = 0; j < nc; ++j) {
            const int sx = wrap_mod(L.chx[off + j] - gx, nkx);
            const int sy = wrap_mod(L.chy[off + j] - gy, nky);
            const int i = map[sx * nky + sy];
            if (i < 0) continue;
            cfma(acc, lam_adj[(off + j) * nz + zl], psi[(off + i) * nz + zl]);
        }
    }
    if (z < nz) {
        cx<double> r;
        if (g == 0) r = mk(-acc.re, 0.0);
        else r = mk(-acc.re, acc.im);  // conj(-acc)
        grad[(long long)g * nz + z] = grad[(long long)g * nz + z] + r;
    }
}

void launch_potential_gradient_map(const Layout &L, const int *chmap, const cx<double> *lam_adj,
                                   const cx<double> *psi, cx<double> *grad, hipStream_t s) {
    dim3 grid((L.nz + 255) / 256, L.nkx * L.nky);
    hipLaunchKernelGGL(gradient_kernel, grid, dim3(256), 0, s, L, chmap, lam_adj, psi, grad);
    MSX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Small set-up kernels.
// ---------------------------------------------------------------------------------------
__global__ void build_h0_kernel(int n, const double *__restrict__ T,
                                const cx<double> *__restrict__ V, double *__restrict__ A) {
    const long long nn = (long long)n * n;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < nn;
         idx += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(idx / n), j = (int)(idx - (long long)i * n);
        // symmetrised T (the eigensolver assumes exact symmetry)
        double a = 0.5 * (T[(long long)i * n + j] + T[(long long)j * n + i]);
        if (i == j) a += V[i].re;  // Re V_00(z_i): g = 0 row of the table
        A[idx] = a;
    }
}

void launch_build_h0(int nz, const double *T, const cx<double> *V, double *A, hipStream_t s) {
    hipLaunchKernelGGL(build_h0_kernel, dim3(ew_blocks((long long)nz * nz)), dim3(256), 0, s, nz,
                       T, V, A);
    MSX_LAUNCH_CHECK();
}

__global__ void flip_table_kernel(int nkx, int nky, int nz, const cx<double> *__restrict__ V,
                                  cx<double> *__restrict__ Vt) {
    const long long total = (long long)nkx * nky * nz;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long g = idx / nz;
        const int z = (int)(idx - g * nz);
        const int gx = (int)(g / nky), gy = (int)(g - (long long)gx * nky);
        const int fx = (nkx - gx) % nkx, fy = (nky - gy) % nky;
        Vt[idx] = V[((long long)fx * nky + fy) * nz + z];
    }
}

void launch_flip_table(int nkx, int nky, int nz, const cx<double> *V, cx<double> *Vt,
                       hipStream_t s) {
    hipLaunchKernelGGL(flip_table_kernel, dim3(ew_blocks((long long)nkx * nky * nz)), dim3(256),
                       0, s, nkx, nky, nz, V, Vt);
    MSX_LAUNCH_CHECK();
}

// One block per instance: zero the instance's rows, then place the source.
__global__ void set_source_kernel(Layout L, const int *__restrict__ spec,
                                  const cx<double> *__restrict__ bsrc, cx<double> *__restrict__ s) {
    const int b = blockIdx.x;
    const long long beg = L.off[b] * L.nz, len = (long long)L.nc[b] * L.nz;
    for (long long i = threadIdx.x; i < len; i += blockDim.x) s[beg + i] = mk(0.0, 0.0);
    __syncthreads();
    if (threadIdx.x == 0) s[(L.off[b] + spec[b]) * L.nz + L.nz - 1] = bsrc[b];
}

void launch_set_source(const Layout &L, const int *spec, const cx<double> *bsrc, cx<double> *s,
                       hipStream_t st) {
    hipLaunchKernelGGL(set_source_kernel, dim3(L.n_inst), dim3(256), 0, st, L, spec, bsrc, s);
    MSX_LAUNCH_CHECK();
}

// d = D^-1 s for the scattering source s (b at node N of each instance's specular channel, zero
// elsewhere). D_j = H0 + e_j - c_j e_N e_N^T, so by Sherman-Morrison D_j^-1 e_N =
// f_j / (1 - c_j f_j(N)) with the preconditioner factor f_j = (H0 + e_j)^-1 e_N: only the
// specular channel of d is nonzero.
__global__ void specular_dinv_kernel(Layout L, const int *__restrict__ spec,
                                     const cx<double> *__restrict__ bsrc,
                                     const cx<double> *__restrict__ c,
                                     const double *__restrict__ fac, cx<double> *__restrict__ d,
                                     cx<float> *__restrict__ d32) {
    const int b = blockIdx.y;
    const int nc = L.nc[b];
    const long long nz = L.nz, beg = L.off[b] * nz, len = (long long)nc * nz;
    const int jsp = spec[b];
    const long long js = L.off[b] + jsp, sbeg = js * nz;  // the factors are channel-major
    const cx<double> coef = cdiv(bsrc[b], mk(1.0, 0.0) - c[js] * fac[sbeg + nz - 1]);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len;
         i += (long long)gridDim.x * blockDim.x) {
        // (channel, node) of position i of the instance block in the vector order of L
        long long j, z;
        if (L.nm) {
            z = i / nc;
            j = i - z * nc;
        } else {
            j = i / nz;
            z = i - j * nz;
        }
        const cx<double> v = j == jsp ? coef * fac[sbeg + z] : mk(0.0, 0.0);
        d[beg + i] = v;
        if (d32) d32[beg + i] = mk((float)v.re, (float)v.im);
    }
}

void launch_specular_dinv(const Layout &L, const int *spec, const cx<double> *bsrc,
                          const cx<double> *c, const double *fac, cx<double> *d, hipStream_t st,
                          cx<float> *d32) {
    const long long per = (long long)L.max_nc * L.nz;
    const unsigned bx = (unsigned)std::min<long long>((per + 255) / 256, 64);
    hipLaunchKernelGGL(specular_dinv_kernel, dim3(std::max(bx, 1u), L.n_inst), dim3(256), 0, st, L,
                       spec, bsrc, c, fac, d, d32);
    MSX_LAUNCH_CHECK();
}

__global__ void adjoint_source_kernel(Layout L, const int *__restrict__ spec,
                                      const cx<double> *__restrict__ psi,
                                      const cx<double> *__restrict__ b,
                                      const cx<double> *__restrict__ a0,
                                      const double *__restrict__ dLdP,
                                      cx<double> *__restrict__ g) {
    const int inst = blockIdx.x;
    const int nz = L.nz, nc = L.nc[inst];
    const long long off = L.off[inst];
    for (long long i = threadIdx.x; i < (long long)nc * nz; i += blockDim.x)
        g[off * nz + i] = mk(0.0, 0.0);
    __syncthreads();
    for (int j = threadIdx.x; j < nc; j += blockDim.x) {
        const long long ch = off + j;
        const cx<double> two_i = mk(0.0, 2.0);
        cx<double> S = two_i * b[ch] * psi[ch * nz + nz - 1];
        if (j == spec[inst]) S = S + a0[inst];
        const cx<double> four_i = mk(0.0, 4.0 * dLdP[ch]);
        g[ch * nz + nz - 1] = four_i * b[ch] * conj(S);
    }
}

void launch_adjoint_source(const Layout &L, const int *spec, const cx<double> *psi,
                           const cx<double> *b, const cx<double> *a0, const double *dLdP,
                           cx<double> *g, hipStream_t st) {
    hipLaunchKernelGGL(adjoint_source_kernel, dim3(L.n_inst), dim3(256), 0, st, L, spec, psi, b,
                       a0, dLdP, g);
    MSX_LAUNCH_CHECK();
}

__global__ void add_float_kernel(long long n, const cx<float> *__restrict__ d,
                                 cx<double> *__restrict__ x) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = mk(x[i].re + (double)d[i].re, x[i].im + (double)d[i].im);
}

void launch_add_float(long long n, const cx<float> *d, cx<double> *x, hipStream_t s) {
    hipLaunchKernelGGL(add_float_kernel, dim3(ew_blocks(n)), dim3(256), 0, s, n, d, x);
    MSX_LAUNCH_CHECK();
}

__global__ void add_ints_kernel(int n, const int *__restrict__ src, int *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] += src[i];
}

void launch_add_ints(int n, const int *src, int *dst, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(add_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, src, dst);
    MSX_LAUNCH_CHECK();
}

}  // namespace msx
