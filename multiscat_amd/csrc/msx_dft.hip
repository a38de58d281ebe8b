// Full channel coupling as a circular 2-D convolution on the G grid, on the matrix cores.
//
// The unsplit coupling y_j(z) = sum_{i != j} V[(G_j - G_i) mod grid](z) x_i(z)
// (lib/multiscat_fortran/fortran/multiscat_gmres.f90:632-635 wrap the G difference modulo the
// table; multiscat/multiscat/_scipy_von_neumann.py:28-37 applies it unsplit) is, for every
// node z, the circular convolution of the table V(z) (with V_00 set to zero, which removes
// i = j) with the state scattered onto the nkx x nky grid. With the DFT matrices
// F_x[k][g] = exp(-2 pi i k g / nkx), F_y likewise:
//     Y = conj(F_x) ( Vhat(z) .* (F_x X F_y) ) conj(F_y),   Vhat = F_x V'(z) F_y / (nkx nky).
// Each of the four products is a small complex GEMM (NP x NP, NP = 16 or 32 >= nkx, nky,
// zero padded) computed with 16x16x4 MFMA instructions: four real MFMAs per complex step.
//
// One workgroup = one instance and ZG nodes. The state is scattered from the compact
// channel layout into LDS slabs, each wavefront runs the four GEMMs of its slabs out of LDS
// (F_x, F_y stay resident in LDS for the whole workgroup), and the result is gathered back.
// LDS row strides are chosen so that the A-fragment reads (16 rows x 4 columns) and the
// B-fragment reads (4 rows x 16 columns) of a wavefront hit 64 distinct banks.
#include "msx_dft.h"
#include "msx_mfma.h"

namespace msx {

template <int NP>
struct DftGeom {
    static constexpr int NT = NP / 16;             // 16 x 16 output tiles per side
    static constexpr int SA = NP == 16 ? 20 : 36;  // row stride of A-read buffers
    static constexpr int SB = NP == 16 ? 16 : 48;  // row stride of B-read buffers
    static constexpr int PLANE = NP * (SA > SB ? SA : SB);
};

// C (NT x NT tiles, re/im) = A B.  A: rows x k from (ar, ai) with stride sa; B: k x cols from
// (br, bi) with stride sb. conj_a / conj_b negate the imaginary part of the operand.
// 16x16x4 operand layout: A[i][k] at lane i + 16 k, B[k][j] at lane j + 16 k; the result
// element (i, j) sits in lane j + 16 (i / 4), register i % 4.
template <typename T, int NP>
__device__ __forceinline__ void cgemm(const T *ar, const T *ai, int sa, const T *br, const T *bi,
                                      int sb, bool conj_a, bool conj_b,
                                      typename Mfma<T>::acc (&cr)[DftGeom<NP>::NT][DftGeom<NP>::NT],
                                      typename Mfma<T>::acc (&ci)[DftGeom<NP>::NT][DftGeom<NP>::NT]) {
    constexpr int NT = DftGeom<NP>::NT;
    const int lane = threadIdx.x & 63;
    const int l16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj) {
            cr[ti][tj] = typename Mfma<T>::acc{0, 0, 0, 0};
            ci[ti][tj] = typename Mfma<T>::acc{0, 0, 0, 0};
        }
#pragma unroll
    for (int kk = 0; kk < NP / 4; ++kk) {
        const int k = kk * 4 + k4;
        T xr[NT], xi[NT], yr[NT], yi[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            xr[t] = ar[(t * 16 + l16) * sa + k];
            xi[t] = ai[(t * 16 + l16) * sa + k];
            if (conj_a) xi[t] = -xi[t];
            yr[t] = br[k * sb + t * 16 + l16];
            yi[t] = bi[k * sb + t * 16 + l16];
            if (conj_b) yi[t] = -yi[t];
        }
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj) {
                cr[ti][tj] = Mfma<T>::op(xr[ti], yr[tj], cr[ti][tj]);
                cr[ti][tj] = Mfma<T>::op(-xi[ti], yi[tj], cr[ti][tj]);
                ci[ti][tj] = Mfma<T>::op(xr[ti], yi[tj], ci[ti][tj]);
                ci[ti][tj] = Mfma<T>::op(xi[ti], yr[tj], ci[ti][tj]);
            }
    }
}

template <typename T, int NP>
__device__ __forceinline__ void store_tiles(T *dr, T *di, int stride,
                                            typename Mfma<T>::acc (&cr)[DftGeom<NP>::NT][DftGeom<NP>::NT],
                                            typename Mfma<T>::acc (&ci)[DftGeom<NP>::NT][DftGeom<NP>::NT]) {
    constexpr int NT = DftGeom<NP>::NT;
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, g = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = ti * 16 + Mfma<T>::row(g, v), c = tj * 16 + col;
                dr[r * stride + c] = cr[ti][tj][v];
                di[r * stride + c] = ci[ti][tj][v];
            }
}

template <typename T, int NP, int ZG>
__global__ __launch_bounds__(256) void dft_coupling_kernel(Layout L, const cx<T> *__restrict__ Fx,
                                                           const cx<T> *__restrict__ Fy,
                                                           const cx<T> *__restrict__ Vhat,
                                                           const cx<T> *__restrict__ x,
                                                           cx<T> *__restrict__ y,
                                                           const int *__restrict__ active) {
    using G = DftGeom<NP>;
    constexpr int NT = G::NT, SA = G::SA, SB = G::SB;
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int nz = L.nz, nky = L.nky;
    const int z0 = blockIdx.x * ZG;
    extern __shared__ double dft_smem[];
    T *fxr = reinterpret_cast<T *>(dft_smem);
    T *fxi = fxr + NP * SA;
    T *fyr = fxi + NP * SA;
    T *fyi = fyr + NP * SB;
    T *slabs = fyi + NP * SB;  // ZG x {re, im} x PLANE

    for (int idx = threadIdx.x; idx < NP * NP; idx += blockDim.x) {
        const int r = idx / NP, c = idx - r * NP;
        const cx<T> fx = Fx[idx], fy = Fy[idx];
        fxr[r * SA + c] = fx.re;
        fxi[r * SA + c] = fx.im;
        fyr[r * SB + c] = fy.re;
        fyi[r * SB + c] = fy.im;
    }
    for (int idx = threadIdx.x; idx < ZG * 2 * G::PLANE; idx += blockDim.x) slabs[idx] = T(0);
    __syncthreads();

    const int nc = L.nc[b];
    const long long off = L.off[b];
    const int *chx = L.chx + off, *chy = L.chy + off;
    // scatter: slab (zz) holds X[gx][gy] in the B layout
    for (int idx = threadIdx.x; idx < nc * ZG; idx += blockDim.x) {
        const int j = idx / ZG, zz = idx - j * ZG;
        const int z = z0 + zz;
        if (z >= nz) continue;
        const cx<T> v = x[(off + j) * nz + z];
        T *sr = slabs + (size_t)zz * 2 * G::PLANE;
        const int p = chx[j] * SB + chy[j];
        sr[p] = v.re;
        sr[G::PLANE + p] = v.im;
    }
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    typename Mfma<T>::acc cr[NT][NT], ci[NT][NT];
    for (int zz = wave; zz < ZG; zz += 4) {
        const int z = z0 + zz;
        if (z >= nz) break;
        T *sr = slabs + (size_t)zz * 2 * G::PLANE;
        T *si = sr + G::PLANE;
        // (1) Y = F_x X
        cgemm<T, NP>(fxr, fxi, SA, sr, si, SB, false, false, cr, ci);
        __builtin_amdgcn_wave_barrier();
        store_tiles<T, NP>(sr, si, SA, cr, ci);
        __builtin_amdgcn_wave_barrier();
        // (2) Z = Y F_y, then Z .* Vhat(z)
        cgemm<T, NP>(sr, si, SA, fyr, fyi, SB, false, false, cr, ci);
        {
            const cx<T> *vh = Vhat + (size_t)z * NP * NP;
            const int col = lane & 15, g = lane >> 4;
#pragma unroll
            for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const cx<T> h = vh[(ti * 16 + Mfma<T>::row(g, v)) * NP + tj * 16 + col];
                        const T a = cr[ti][tj][v], c = ci[ti][tj][v];
                        cr[ti][tj][v] = a * h.re - c * h.im;
                        ci[ti][tj][v] = a * h.im + c * h.re;
                    }
        }
        __builtin_amdgcn_wave_barrier();
        store_tiles<T, NP>(sr, si, SB, cr, ci);
        __builtin_amdgcn_wave_barrier();
        // (3) W = conj(F_x) Zhat
        cgemm<T, NP>(fxr, fxi, SA, sr, si, SB, true, false, cr, ci);
        __builtin_amdgcn_wave_barrier();
        store_tiles<T, NP>(sr, si, SA, cr, ci);
        __builtin_amdgcn_wave_barrier();
        // (4) out = W conj(F_y)
        cgemm<T, NP>(sr, si, SA, fyr, fyi, SB, false, true, cr, ci);
        __builtin_amdgcn_wave_barrier();
        store_tiles<T, NP>(sr, si, SB, cr, ci);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < nc * ZG; idx += blockDim.x) {
        const int j = idx / ZG, zz = idx - j * ZG;
        const int z = z0 + zz;
        if (z >= nz) continue;
        const T *sr = slabs + (size_t)zz * 2 * G::PLANE;
        const int p = chx[j] * SB + chy[j];
        y[(off + j) * nz + z] = mk(sr[p], sr[G::PLANE + p]);
    }
    (void)nky;
}

template <typename T, int NP, int ZG>
static void launch_dft_impl(const Layout &L, const DftTables<T> &D, const cx<T> *x, cx<T> *y,
                            const int *active, hipStream_t s) {
    using G = DftGeom<NP>;
    const size_t shm = sizeof(T) * ((size_t)NP * G::SA * 2 + (size_t)NP * G::SB * 2 +
                                    (size_t)ZG * 2 * G::PLANE);
    set_dynamic_lds((const void *)dft_coupling_kernel<T, NP, ZG>, shm);
    dim3 grid((L.nz + ZG - 1) / ZG, L.n_inst);
    hipLaunchKernelGGL((dft_coupling_kernel<T, NP, ZG>), grid, dim3(256), shm, s, L, D.Fx, D.Fy,
                       D.Vhat, x, y, active);
    MSX_LAUNCH_CHECK();
}

template <typename T>
void launch_dft_coupling(const Layout &L, const DftTables<T> &D, const cx<T> *x, cx<T> *y,
                         const int *active, hipStream_t s) {
    constexpr int ZG = sizeof(T) == 4 ? 8 : 4;
    if (D.np == 16) launch_dft_impl<T, 16, ZG>(L, D, x, y, active, s);
    else if (D.np == 32) launch_dft_impl<T, 32, ZG>(L, D, x, y, active, s);
    else throw Error(3, "DFT coupling supports tables up to 32 x 32");
}
template void launch_dft_coupling<float>(const Layout &, const DftTables<float> &,
                                         const cx<float> *, cx<float> *, const int *, hipStream_t);
template void launch_dft_coupling<double>(const Layout &, const DftTables<double> &,
                                          const cx<double> *, cx<double> *, const int *,
                                          hipStream_t);

// ---------------------------------------------------------------------------------------
// Set-up: DFT matrices (FP64 twiddles, k g reduced mod n) and Vhat(z) = F_x V'(z) F_y / nG.
// ---------------------------------------------------------------------------------------
__global__ void dft_matrix_kernel(int n, int np, cx<double> *F) {
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < np * np; idx += gridDim.x * blockDim.x) {
        const int k = idx / np, g = idx - k * np;
        cx<double> v = mk(0.0, 0.0);
        if (k < n && g < n) {
            const long long kg = ((long long)k * g) % n;
            double sn, cs;
            sincospi(-2.0 * (double)kg / (double)n, &sn, &cs);
            v = mk(cs, sn);
        }
        F[idx] = v;
    }
}

// tmp[z][kx][gy] = sum_gx Fx[kx][gx] V'[gx][gy][z]   (V' = V with V_00 = 0)
__global__ void vhat_stage1_kernel(int nkx, int nky, int nz, int np, const cx<double> *__restrict__ V,
                                   const cx<double> *__restrict__ Fx, cx<double> *__restrict__ tmp) {
    const long long total = (long long)nz * np * nky;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int gy = (int)(idx % nky);
        const int kx = (int)((idx / nky) % np);
        const int z = (int)(idx / ((long long)nky * np));
        cx<double> acc = mk(0.0, 0.0);
        if (kx < nkx)
            for (int gx = 0; gx < nkx; ++gx) {
                if (gx == 0 && gy == 0) continue;
                cfma(acc, Fx[kx * np + gx], V[((long long)gx * nky + gy) * nz + z]);
            }
        tmp[((long long)z * np + kx) * nky + gy] = acc;
    }
}

// Vhat[z][kx][ky] = sum_gy tmp[z][kx][gy] Fy[ky][gy] / (nkx nky)
__global__ void vhat_stage2_kernel(int nkx, int nky, int nz, int np, const cx<double> *__restrict__ tmp,
                                   const cx<double> *__restrict__ Fy, cx<double> *__restrict__ Vhat) {
    const long long total = (long long)nz * np * np;
    const double scale = 1.0 / ((double)nkx * nky);
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int ky = (int)(idx % np);
        const int kx = (int)((idx / np) % np);
        const int z = (int)(idx / ((long long)np * np));
        cx<double> acc = mk(0.0, 0.0);
        if (kx < nkx && ky < nky)
            for (int gy = 0; gy < nky; ++gy)
                cfma(acc, tmp[((long long)z * np + kx) * nky + gy], Fy[ky * np + gy]);
        Vhat[idx] = acc * scale;
    }
}

__global__ void cast_cx_kernel(long long n, const cx<double> *__restrict__ a, cx<float> *__restrict__ b) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        b[i] = mk((float)a[i].re, (float)a[i].im);
}

int dft_padded_size(int nkx, int nky) {
    const int m = nkx > nky ? nkx : nky;
    if (m <= 16) return 16;
    if (m <= 32) return 32;
    return 0;
}

void build_dft_tables(int nkx, int nky, int nz, const cx<double> *V, DftStore &st, hipStream_t s) {
    const int np = dft_padded_size(nkx, nky);
    st.np = np;
    if (np == 0) return;
    const size_t np2 = (size_t)np * np;
    if (st.Fx.n != np2) {
        st.Fx.alloc(np2);
        st.Fy.alloc(np2);
        st.Fxf.alloc(np2);
        st.Fyf.alloc(np2);
    }
    if (st.Vhat.n != np2 * nz) {
        st.Vhat.alloc(np2 * nz);
        st.Vhatf.alloc(np2 * nz);
        st.tmp.alloc((size_t)nz * np * nky);
    }
    hipLaunchKernelGGL(dft_matrix_kernel, dim3(8), dim3(256), 0, s, nkx, np, st.Fx.p);
    hipLaunchKernelGGL(dft_matrix_kernel, dim3(8), dim3(256), 0, s, nky, np, st.Fy.p);
    MSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(vhat_stage1_kernel, dim3(1024), dim3(256), 0, s, nkx, nky, nz, np, V, st.Fx.p,
                       st.tmp.p);
    hipLaunchKernelGGL(vhat_stage2_kernel, dim3(1024), dim3(256), 0, s, nkx, nky, nz, np, st.tmp.p,
                       st.Fy.p, st.Vhat.p);
    MSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(cast_cx_kernel, dim3(64), dim3(256), 0, s, (long long)np2, st.Fx.p, st.Fxf.p);
    hipLaunchKernelGGL(cast_cx_kernel, dim3(64), dim3(256), 0, s, (long long)np2, st.Fy.p, st.Fyf.p);
    hipLaunchKernelGGL(cast_cx_kernel, dim3(1024), dim3(256), 0, s, (long long)(np2 * nz), st.Vhat.p,
                       st.Vhatf.p);
    MSX_LAUNCH_CHECK();
}

}  // namespace msx
