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
// Each wavefront transforms one node at a time in its own LDS slab with F_x and F_y held in
// registers, so every MFMA reads one operand from LDS; the grid runs over instances fastest so
// that workgroups resident together share the Vhat slabs of the same nodes in L2. For a square
// table F_y = F_x (the DFT matrix is symmetric), so one copy serves both, and the kernel then
// forms each complex product from three real ones (T1 = A_r B_r, T2 = A_i B_i,
// T3 = (A_r + A_i)(B_r + B_i); re = T1 - T2, im = T3 - T1 - T2), 384 MFMAs per node instead of
// 512 (relative difference to the four-product form: 5e-7 in FP32, 9e-16 in FP64). With
// node-major vectors (Layout::nm) the channels of one node are contiguous, so the loads of
// X(z) and the stores of Y(z) are whole cache lines. The coupling kernel runs all four
// transforms as right products with transposed stores (see the slab products below): every
// LDS access of a transform is a 16-byte run (8 reads and 8 stores per lane and transform
// instead of 32 and 32 single words; 2.5% faster). The gradient kernels keep the left/right
// products with single-word accesses: LDS row stride 36 (= 36 mod 64 words) at NP = 32, so the
// A-fragment reads (16 rows x 4 columns) and the accumulator stores (rows 4g + v, 16 columns)
// hit 64 distinct banks and the B-fragment reads (4 rows x 16 columns) two.
#include "msx_dft.h"
#include "msx_mfma.h"

#include <algorithm>
#include <type_traits>

namespace msx {

// Workgroups go to the device's 8 XCDs (each with its own L2) round-robin by linear index.
// xcd_order maps a 1-D launch of 8 ceil(total / 8) workgroups onto work items 0 .. total-1 so
// that consecutive items land on one XCD in dispatch order (false: an item past the end).
__device__ __forceinline__ bool xcd_order(int total, int &item) {
    const int L = blockIdx.x, S = (total + 7) >> 3;
    item = (L & 7) * S + (L >> 3);
    return item < total;
}

template <int NP>
struct DftGeom {
    static constexpr int NT = NP / 16;             // 16 x 16 output tiles per side
    static constexpr int SA = NP == 16 ? 20 : 36;  // row stride of A-read buffers
    static constexpr int SB = NP == 16 ? 16 : 36;  // row stride of B-read buffers
    static constexpr int PLANE = NP * (SA > SB ? SA : SB);
};

// C (NT x NT tiles, re/im) = F B with the A operand F held in registers: fr/fi[ti][kk] is
// F[ti*16 + l16][kk*4 + k4] (lane l16 + 16 k4); B (k x cols) is read from LDS (br, bi) with
// row stride sb. conj_f negates the imaginary part of F.
// 16x16x4 operand layout: A[i][k] at lane i + 16 k, B[k][j] at lane j + 16 k; the result
// element (i, j) sits in lane j + 16 g, register v with i = Mfma::row(g, v).
template <typename T, int NP>
__device__ __forceinline__ void cgemm_reg_a(const T (&fr)[DftGeom<NP>::NT][NP / 4],
                                            const T (&fi)[DftGeom<NP>::NT][NP / 4], bool conj_f,
                                            const T *br, const T *bi, int sb,
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
        T yr[NT], yi[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            yr[t] = br[k * sb + t * 16 + l16];
            yi[t] = bi[k * sb + t * 16 + l16];
        }
        // negations fall on the transient LDS operand, never on the register-resident F:
        // F y:       cr += fr yr + fi (-yi),  ci += fr yi + fi yr
        // conj(F) y: cr += fr yr + fi yi,     ci += fr yi + fi (-yr)
        T yr2[NT], yi2[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            yi2[t] = conj_f ? yi[t] : -yi[t];
            yr2[t] = conj_f ? -yr[t] : yr[t];
        }
#pragma unroll
        for (int ti = 0; ti < NT; ++ti) {
            const T xr = fr[ti][kk], xi = fi[ti][kk];
#pragma unroll
            for (int tj = 0; tj < NT; ++tj) {
                cr[ti][tj] = Mfma<T>::op(xr, yr[tj], cr[ti][tj]);
                cr[ti][tj] = Mfma<T>::op(xi, yi2[tj], cr[ti][tj]);
                ci[ti][tj] = Mfma<T>::op(xr, yi[tj], ci[ti][tj]);
                ci[ti][tj] = Mfma<T>::op(xi, yr2[tj], ci[ti][tj]);
            }
        }
    }
}

// C = A F with A (rows x k) read from LDS (ar, ai; row stride sa) and the B operand F in
// registers: fr/fi[tj][kk] is F[kk*4 + k4][tj*16 + l16]. conj_f negates Im F.
template <typename T, int NP>
__device__ __forceinline__ void cgemm_reg_b(const T *ar, const T *ai, int sa,
                                            const T (&fr)[DftGeom<NP>::NT][NP / 4],
                                            const T (&fi)[DftGeom<NP>::NT][NP / 4], bool conj_f,
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
        T xr[NT], xi[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            xr[t] = ar[(t * 16 + l16) * sa + k];
            xi[t] = ai[(t * 16 + l16) * sa + k];
        }
        // x F:       cr += xr fr + (-xi) fi,  ci += xr fi + xi fr
        // x conj(F): cr += xr fr + xi fi,     ci += (-xr) fi + xi fr
        T xn[NT], xm[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            xn[t] = conj_f ? xi[t] : -xi[t];
            xm[t] = conj_f ? -xr[t] : xr[t];
        }
#pragma unroll
        for (int tj = 0; tj < NT; ++tj) {
            const T yr = fr[tj][kk], yi = fi[tj][kk];
#pragma unroll
            for (int ti = 0; ti < NT; ++ti) {
                cr[ti][tj] = Mfma<T>::op(xr[ti], yr, cr[ti][tj]);
                cr[ti][tj] = Mfma<T>::op(xn[ti], yi, cr[ti][tj]);
                ci[ti][tj] = Mfma<T>::op(xm[ti], yi, ci[ti][tj]);
                ci[ti][tj] = Mfma<T>::op(xi[ti], yr, ci[ti][tj]);
            }
        }
    }
}

// The three real products of C = F Y (F in registers as the A operand, Y from LDS) or of
// C = conj(F) Y (CONJ: the sum operand is Y_r - Y_i); combine3 forms C from them.
template <typename T, int NP, bool CONJ>
__device__ __forceinline__ void cgemm3_reg_a(const T (&fr)[DftGeom<NP>::NT][NP / 4],
                                             const T (&fi)[DftGeom<NP>::NT][NP / 4],
                                             const T *br, const T *bi, int sb,
                                             typename Mfma<T>::acc (&c1)[DftGeom<NP>::NT][DftGeom<NP>::NT],
                                             typename Mfma<T>::acc (&c2)[DftGeom<NP>::NT][DftGeom<NP>::NT],
                                             typename Mfma<T>::acc (&c3)[DftGeom<NP>::NT][DftGeom<NP>::NT]) {
    constexpr int NT = DftGeom<NP>::NT;
    const int lane = threadIdx.x & 63;
    const int l16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj)
            c1[ti][tj] = c2[ti][tj] = c3[ti][tj] = typename Mfma<T>::acc{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < NP / 4; ++kk) {
        const int k = kk * 4 + k4;
        T yr[NT], yi[NT], ys[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            yr[t] = br[k * sb + t * 16 + l16];
            yi[t] = bi[k * sb + t * 16 + l16];
            ys[t] = CONJ ? yr[t] - yi[t] : yr[t] + yi[t];
        }
#pragma unroll
        for (int ti = 0; ti < NT; ++ti) {
            const T xr = fr[ti][kk], xi = fi[ti][kk], xs = xr + xi;
#pragma unroll
            for (int tj = 0; tj < NT; ++tj) {
                c1[ti][tj] = Mfma<T>::op(xr, yr[tj], c1[ti][tj]);
                c2[ti][tj] = Mfma<T>::op(xi, yi[tj], c2[ti][tj]);
                c3[ti][tj] = Mfma<T>::op(xs, ys[tj], c3[ti][tj]);
            }
        }
    }
}

// The three real products of C = X F (X from LDS as the A operand, F in registers) or of
// C = X conj(F) (CONJ: the sum operand is X_r - X_i).
template <typename T, int NP, bool CONJ>
__device__ __forceinline__ void cgemm3_reg_b(const T *ar, const T *ai, int sa,
                                             const T (&fr)[DftGeom<NP>::NT][NP / 4],
                                             const T (&fi)[DftGeom<NP>::NT][NP / 4],
                                             typename Mfma<T>::acc (&c1)[DftGeom<NP>::NT][DftGeom<NP>::NT],
                                             typename Mfma<T>::acc (&c2)[DftGeom<NP>::NT][DftGeom<NP>::NT],
                                             typename Mfma<T>::acc (&c3)[DftGeom<NP>::NT][DftGeom<NP>::NT]) {
    constexpr int NT = DftGeom<NP>::NT;
    const int lane = threadIdx.x & 63;
    const int l16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj)
            c1[ti][tj] = c2[ti][tj] = c3[ti][tj] = typename Mfma<T>::acc{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < NP / 4; ++kk) {
        const int k = kk * 4 + k4;
        T xr[NT], xi[NT], xs[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            xr[t] = ar[(t * 16 + l16) * sa + k];
            xi[t] = ai[(t * 16 + l16) * sa + k];
            xs[t] = CONJ ? xr[t] - xi[t] : xr[t] + xi[t];
        }
#pragma unroll
        for (int tj = 0; tj < NT; ++tj) {
            const T yr = fr[tj][kk], yi = fi[tj][kk], ysum = yr + yi;
#pragma unroll
            for (int ti = 0; ti < NT; ++ti) {
                c1[ti][tj] = Mfma<T>::op(xr[ti], yr, c1[ti][tj]);
                c2[ti][tj] = Mfma<T>::op(xi[ti], yi, c2[ti][tj]);
                c3[ti][tj] = Mfma<T>::op(xs[ti], ysum, c3[ti][tj]);
            }
        }
    }
}

// (re, im) of the product into (c1, c2): plain re = T1 - T2, im = T3 - T1 - T2;
// conjugated re = T1 + T2, im = T1 - T2 - T3.
template <typename T, int NP, bool CONJ>
__device__ __forceinline__ void combine3(typename Mfma<T>::acc (&c1)[DftGeom<NP>::NT][DftGeom<NP>::NT],
                                         typename Mfma<T>::acc (&c2)[DftGeom<NP>::NT][DftGeom<NP>::NT],
                                         const typename Mfma<T>::acc (&c3)[DftGeom<NP>::NT][DftGeom<NP>::NT]) {
    constexpr int NT = DftGeom<NP>::NT;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const T a = c1[ti][tj][v], b = c2[ti][tj][v], c = c3[ti][tj][v];
                const T sum = a + b, dif = a - b;
                c1[ti][tj][v] = CONJ ? sum : dif;
                c2[ti][tj][v] = CONJ ? dif - c : c - sum;
            }
}

// Compile-time choice between two register arrays of the same shape.
template <bool FIRST, class A>
__device__ __forceinline__ A &pick(A &a, A &b) {
    if constexpr (FIRST) return a;
    else return b;
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

// ---- the coupling kernel's slab products ---------------------------------------------------
// Every transform is a right product C = A F~ (F~ = F or conj(F), in registers as the B
// operand) with A read from the wavefront's LDS slab by rows, and its result is stored
// transposed, slab[j][i] = C[i][j]: with F symmetric, F X = (X^T F)^T, so alternating the
// roles of rows and columns turns the four transforms of Y = F_x X F_y into four such
// products (the scatter writes X^T, the gather reads Y^T). The k index of MFMA step kk in lane
// group k4 is k4 KK + kk (the products sum over k in any order), so a lane reads the KK
// entries of its A row as 16-byte runs; the accumulator of lane (j, g) holds rows 4 g .. 4 g + 3
// of column j (FP64 through the row permutation of the A operand), one 16-byte run (two in
// FP64) of the transposed store. Row stride S: 4 (mod 64) words times an odd number for FP32,
// so the 16 lanes of a group touch distinct 16-byte bank quads in reads and stores alike.
template <typename T, int NP>
struct DftSlab {
    static constexpr int NT = NP / 16;  // 16 x 16 tiles per side
    static constexpr int KK = NP / 4;   // k-steps per product
    static constexpr int S = sizeof(T) == 4 ? (NP == 16 ? 20 : 36) : (NP == 16 ? 18 : 34);
    static constexpr int PLANE = NP * S;
};

// x[0 .. KK) = p[0 .. KK), p 16-byte aligned
template <int KK>
__device__ __forceinline__ void slab_row(const float *p, float (&x)[KK]) {
#pragma unroll
    for (int q = 0; q < KK / 4; ++q) {
        const float4 v = reinterpret_cast<const float4 *>(p)[q];
        x[4 * q] = v.x; x[4 * q + 1] = v.y; x[4 * q + 2] = v.z; x[4 * q + 3] = v.w;
    }
}
template <int KK>
__device__ __forceinline__ void slab_row(const double *p, double (&x)[KK]) {
#pragma unroll
    for (int q = 0; q < KK / 2; ++q) {
        const double2 v = reinterpret_cast<const double2 *>(p)[q];
        x[2 * q] = v.x; x[2 * q + 1] = v.y;
    }
}

// C = A F (CONJ: A conj(F)), A from the slab planes (ar, ai), F in registers:
// fr/fi[tj][kk] = F[k4 KK + kk][16 tj + l16]. M3: three real products (T1 = Ar Fr, T2 = Ai Fi,
// T3 = (Ar +- Ai)(Fr + Fi)) combined by combine3; else four.
template <typename T, int NP, bool CONJ, bool M3>
__device__ __forceinline__ void slab_product(const T *ar, const T *ai,
                                             const T (&fr)[DftSlab<T, NP>::NT][NP / 4],
                                             const T (&fi)[DftSlab<T, NP>::NT][NP / 4],
                                             typename Mfma<T>::acc (&cr)[DftSlab<T, NP>::NT][DftSlab<T, NP>::NT],
                                             typename Mfma<T>::acc (&ci)[DftSlab<T, NP>::NT][DftSlab<T, NP>::NT],
                                             typename Mfma<T>::acc (&c3)[DftSlab<T, NP>::NT][DftSlab<T, NP>::NT]) {
    using D = DftSlab<T, NP>;
    constexpr int NT = D::NT, KK = D::KK, S = D::S;
    using acc_t = typename Mfma<T>::acc;
    const int lane = threadIdx.x & 63;
    const int l16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj) cr[ti][tj] = ci[ti][tj] = c3[ti][tj] = acc_t{0, 0, 0, 0};
    T xr[NT][KK], xi[NT][KK];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int row = 16 * t + Mfma<T>::perm(l16);
        slab_row<KK>(ar + row * S + k4 * KK, xr[t]);
        slab_row<KK>(ai + row * S + k4 * KK, xi[t]);
    }
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
        if constexpr (M3) {
            T xs[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) xs[t] = CONJ ? xr[t][kk] - xi[t][kk] : xr[t][kk] + xi[t][kk];
#pragma unroll
            for (int tj = 0; tj < NT; ++tj) {
                const T yr = fr[tj][kk], yi = fi[tj][kk], ys = yr + yi;
#pragma unroll
                for (int ti = 0; ti < NT; ++ti) {
                    cr[ti][tj] = Mfma<T>::op(xr[ti][kk], yr, cr[ti][tj]);
                    ci[ti][tj] = Mfma<T>::op(xi[ti][kk], yi, ci[ti][tj]);
                    c3[ti][tj] = Mfma<T>::op(xs[ti], ys, c3[ti][tj]);
                }
            }
        } else {
            // x F:       cr += xr fr - xi fi,  ci += xr fi + xi fr
            // x conj(F): cr += xr fr + xi fi,  ci += xi fr - xr fi
            T xn[NT], xm[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                xn[t] = CONJ ? xi[t][kk] : -xi[t][kk];
                xm[t] = CONJ ? -xr[t][kk] : xr[t][kk];
            }
#pragma unroll
            for (int tj = 0; tj < NT; ++tj) {
                const T yr = fr[tj][kk], yi = fi[tj][kk];
#pragma unroll
                for (int ti = 0; ti < NT; ++ti) {
                    cr[ti][tj] = Mfma<T>::op(xr[ti][kk], yr, cr[ti][tj]);
                    cr[ti][tj] = Mfma<T>::op(xn[ti], yi, cr[ti][tj]);
                    ci[ti][tj] = Mfma<T>::op(xm[ti], yi, ci[ti][tj]);
                    ci[ti][tj] = Mfma<T>::op(xi[ti][kk], yr, ci[ti][tj]);
                }
            }
        }
    }
    if constexpr (M3) combine3<T, NP, CONJ>(cr, ci, c3);
}

// slab[j][i] = C[i][j]: lane (j, g) of tile (ti, tj) holds rows 16 ti + 4 g .. + 3 of column
// 16 tj + j, contiguous in the transposed slab
template <typename T, int NP>
__device__ __forceinline__ void slab_store_t(T *dr, T *di,
                                             const typename Mfma<T>::acc (&cr)[DftSlab<T, NP>::NT][DftSlab<T, NP>::NT],
                                             const typename Mfma<T>::acc (&ci)[DftSlab<T, NP>::NT][DftSlab<T, NP>::NT]) {
    using D = DftSlab<T, NP>;
    constexpr int NT = D::NT, S = D::S;
    using acc_t = typename Mfma<T>::acc;
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, g = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj) {
            const int a = (16 * tj + col) * S + 16 * ti + 4 * g;
            if constexpr (sizeof(T) == 4) {
                *reinterpret_cast<acc_t *>(dr + a) = cr[ti][tj];
                *reinterpret_cast<acc_t *>(di + a) = ci[ti][tj];
            } else {  // 16-byte halves (a row run starts on a 16-byte boundary only)
                reinterpret_cast<double2 *>(dr + a)[0] = make_double2(cr[ti][tj][0], cr[ti][tj][1]);
                reinterpret_cast<double2 *>(dr + a)[1] = make_double2(cr[ti][tj][2], cr[ti][tj][3]);
                reinterpret_cast<double2 *>(di + a)[0] = make_double2(ci[ti][tj][0], ci[ti][tj][1]);
                reinterpret_cast<double2 *>(di + a)[1] = make_double2(ci[ti][tj][2], ci[ti][tj][3]);
            }
        }
}

// One workgroup = one instance and zpb nodes (node-major: blockIdx.x the instance, so that
// consecutive workgroups share the Vhat slabs of the same nodes in L2, blockIdx.y the node
// group; channel-major: the XCD-aware order below). Each wavefront owns an LDS slab and
// processes its nodes independently: scatter X(z)^T onto the grid, the four transforms as
// slab products with F_x / F_y held in registers for the whole kernel, the pointwise product
// with Vhat(z) in registers after the second, gather Y(z)^T back to the compact layout. No
// workgroup barrier. The channel values of a wavefront's next node are loaded into registers
// while the current node's transforms run; the channels' grid positions are computed once.
// SQ: square table (F_y = F_x); NM: node-major vectors.
template <typename T, int NP, int WPB, bool SQ, bool NM>
__global__ __launch_bounds__(64 * WPB, sizeof(T) == 4 ? 2 : 1) void dft_coupling_kernel(
    Layout L, const cx<T> *__restrict__ Fx, const cx<T> *__restrict__ Fy,
    const cx<T> *__restrict__ Vhat, const cx<T> *__restrict__ x, cx<T> *__restrict__ y,
    const int *__restrict__ active, int zpb, const double *__restrict__ x_norm,
    cx<T> *__restrict__ x_copy) {
    using G = DftSlab<T, NP>;
    using acc_t = typename Mfma<T>::acc;
    constexpr int NT = G::NT, S = G::S, KK = G::KK;
    constexpr int KC = NP * NP / 64;  // channels per lane (nc <= NP * NP)
    // three real products per complex one when one F in registers serves both transforms
    // (register budget)
    constexpr bool M3 = SQ;
    // NM: instances fastest (workgroups resident together share the Vhat slabs of their
    // nodes); channel-major: the node groups of one instance consecutive on one XCD (128-byte
    // lines of the state hold several nodes of a channel, so neighbouring groups share lines)
    int b, zg;
    if constexpr (NM) {
        b = blockIdx.x;
        zg = blockIdx.y;
    } else {
        const int nzb = (L.nz + zpb - 1) / zpb;
        int item;
        if (!xcd_order(L.n_inst * nzb, item)) return;
        b = item / nzb;
        zg = item - b * nzb;
    }
    if (active && !active[b]) return;
    const int nz = L.nz;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int l16 = lane & 15, k4 = lane >> 4;
    extern __shared__ double dft_smem[];
    T *sr = reinterpret_cast<T *>(dft_smem) + (size_t)wave * 2 * G::PLANE;
    T *si = sr + G::PLANE;

    // F_x and F_y as B operands: f[t][kk] = F[k4 KK + kk][16 t + l16]
    T fxr[NT][KK], fxi[NT][KK], fyr_[NT][KK], fyi_[NT][KK];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int kk = 0; kk < KK; ++kk) {
            const int e = (k4 * KK + kk) * NP + t * 16 + l16;
            const cx<T> f = Fx[e];
            fxr[t][kk] = f.re;
            fxi[t][kk] = f.im;
            if constexpr (!SQ) {
                const cx<T> g = Fy[e];
                fyr_[t][kk] = g.re;
                fyi_[t][kk] = g.im;
            }
        }
    auto &fyr = pick<SQ>(fxr, fyr_);
    auto &fyi = pick<SQ>(fxi, fyi_);

    const int nc = L.nc[b];
    const long long off = L.off[b];
    __shared__ int s_pos[NP * NP];  // slab position of each channel of the instance (X^T)
    for (int j = threadIdx.x; j < nc; j += blockDim.x) s_pos[j] = L.chy[off + j] * S + L.chx[off + j];
    __syncthreads();
    const int col = l16, g = k4;
    cx<T> xv[KC], vh[NT][NT][4];
    // element offsets fit in 32 bits (the launcher checks), so the loads and stores use a
    // uniform base and one 32-bit offset register; lanes past nc read the instance's first
    // element (a valid address) and store nothing
    const int row_l = (int)(off + lane) * nz, row_step = 64 * nz, row_0 = (int)off * nz;
    auto vidx = [&](int k, int z) {
        const int j = lane + 64 * k;
        if constexpr (NM) return row_0 + z * nc + (j < nc ? j : 0);
        else return (j < nc ? row_l + k * row_step : row_0) + z;
    };
    // with x_copy: the input is x / x_norm[b] (a Krylov vector before its normalisation),
    // scaled where it is scattered and stored to x_copy there (each element is scattered
    // exactly once; doing it at the load would make the prefetch wait for its data)
    const T sc = x_copy ? (T)(1.0 / x_norm[b]) : T(1);
    auto load_x = [&](int z) {
#pragma unroll
        for (int k = 0; k < KC; ++k) xv[k] = x[vidx(k, z)];
    };
    // Vhat(z)[kx][ky] in the accumulator layout of the second product (rows kx = 16 ti + 4 g + v)
    auto load_v = [&](int z) {
        const cx<T> *vz = Vhat + (size_t)z * NP * NP;
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    vh[ti][tj][v] = vz[(ti * 16 + 4 * g + v) * NP + tj * 16 + col];
    };
    acc_t cr[NT][NT], ci[NT][NT], c3[NT][NT];
    using plain_t = std::integral_constant<bool, false>;
    using conj_t = std::integral_constant<bool, true>;
    const int z_first = zg * zpb + wave;
    const int z_end = min(nz, (zg + 1) * zpb);
    if (z_first < z_end) load_x(z_first);
    // the first node's data waited for here: with loads pending on entry, the loop head
    // waited for every outstanding operation, the previous node's stores included
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    for (int z = z_first; z < z_end; z += WPB) {
        {
            float4 *p4 = reinterpret_cast<float4 *>(sr);  // both planes, 16-byte stores
            constexpr int N4 = 2 * G::PLANE * (int)sizeof(T) / 16;
            for (int idx = lane; idx < N4; idx += 64) p4[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int j = lane + 64 * k;
            if (j < nc) {
                const int p = s_pos[j];
                cx<T> v = xv[k];
                if (x_copy) {
                    v = mk(v.re * sc, v.im * sc);
                    x_copy[vidx(k, z)] = v;
                }
                sr[p] = v.re;
                si[p] = v.im;
            }
        }
        const bool more = z + WPB < z_end;
        if (more) load_x(z + WPB);  // consumed by the scatter above
        __builtin_amdgcn_wave_barrier();
        // (1) X^T F_x = (F_x X)^T, stored transposed: the slab holds Y1 = F_x X
        slab_product<T, NP, false, M3>(sr, si, fxr, fxi, cr, ci, c3);
        __builtin_amdgcn_wave_barrier();
        slab_store_t<T, NP>(sr, si, cr, ci);
        __builtin_amdgcn_wave_barrier();
        // (2) Y1 F_y, then .* Vhat(z) (the slab is loaded just before this product; the
        // barrier keeps the scheduler from moving the loads down to their use after it); the
        // transposed store leaves Z^T
        load_v(z);
        __builtin_amdgcn_sched_barrier(0);
        slab_product<T, NP, false, M3>(sr, si, fyr, fyi, cr, ci, c3);
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const cx<T> h = vh[ti][tj][v];
                    const T a = cr[ti][tj][v], c = ci[ti][tj][v];
                    cr[ti][tj][v] = a * h.re - c * h.im;
                    ci[ti][tj][v] = a * h.im + c * h.re;
                }
        __builtin_amdgcn_wave_barrier();
        slab_store_t<T, NP>(sr, si, cr, ci);
        __builtin_amdgcn_wave_barrier();
        // (3) Z^T conj(F_x) = (conj(F_x) Z)^T, stored transposed: the slab holds W
        slab_product<T, NP, true, M3>(sr, si, fxr, fxi, cr, ci, c3);
        __builtin_amdgcn_wave_barrier();
        slab_store_t<T, NP>(sr, si, cr, ci);
        __builtin_amdgcn_wave_barrier();
        // (4) W conj(F_y) = out, stored transposed: the slab holds out^T
        slab_product<T, NP, true, M3>(sr, si, fyr, fyi, cr, ci, c3);
        __builtin_amdgcn_wave_barrier();
        slab_store_t<T, NP>(sr, si, cr, ci);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int j = lane + 64 * k;
            if (j < nc) {
                const int p = s_pos[j];
                y[vidx(k, z)] = mk(sr[p], si[p]);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------------------
// Tables beyond 32 x 32 (padded to NP = 48, 64 or 80): one workgroup of NT = NP / 16
// wavefronts per (instance, node); wavefront w owns rows [16 w, 16 w + 16) of the NP x NP grid
// in one LDS slab. NP^2 DFT coefficients are too many registers to hold and, at NP = 80, too
// much LDS beside the slab, so the matrices are read from L2 as the MFMA operands. A left
// product F Y needs every row of Y: each wavefront forms its row strip in registers, the
// workgroup waits, then the strips are stored; a right product Y F reads only the wavefront's
// own strip. Square tables use the three-product complex form with one matrix (F_y = F_x).
// Channel-major vectors; the node of a workgroup runs fastest in an XCD's dispatch order, so
// the nodes of one channel that share a 128-byte line meet in that XCD's L2.
// ---------------------------------------------------------------------------------------
template <int NT>
struct BigGeom {
    static constexpr int NP = 16 * NT;
    static constexpr int S = NP + 4;  // row stride of the slab
    static constexpr int PLANE = NP * S;
};

// (cr, ci) = F Y (CONJ: conj(F) Y) for this wavefront's row strip ti: F[ti*16 + i][k] from
// global memory (row-major NP x NP), Y[k][j] from the slab.
template <typename T, int NT, bool CONJ, bool M3>
__device__ __forceinline__ void big_left(const cx<T> *__restrict__ F, const T *br, const T *bi,
                                         typename Mfma<T>::acc (&cr)[NT],
                                         typename Mfma<T>::acc (&ci)[NT]) {
    using G = BigGeom<NT>;
    using acc_t = typename Mfma<T>::acc;
    const int lane = threadIdx.x & 63, l16 = lane & 15, k4 = lane >> 4, ti = threadIdx.x >> 6;
    acc_t c3[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) cr[t] = ci[t] = c3[t] = acc_t{0, 0, 0, 0};
    const cx<T> *frow = F + (size_t)(ti * 16 + l16) * G::NP;
#pragma unroll 4
    for (int kk = 0; kk < G::NP / 4; ++kk) {
        const int k = kk * 4 + k4;
        const cx<T> f = frow[k];
        T yr[NT], yi[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            yr[t] = br[k * G::S + t * 16 + l16];
            yi[t] = bi[k * G::S + t * 16 + l16];
        }
        if constexpr (M3) {
            const T fs = f.re + f.im;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                cr[t] = Mfma<T>::op(f.re, yr[t], cr[t]);
                ci[t] = Mfma<T>::op(f.im, yi[t], ci[t]);
                c3[t] = Mfma<T>::op(fs, CONJ ? yr[t] - yi[t] : yr[t] + yi[t], c3[t]);
            }
        } else {
            // F y: re += fr yr - fi yi, im += fr yi + fi yr; conj(F) y: re += fr yr + fi yi,
            // im += fr yi - fi yr
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                cr[t] = Mfma<T>::op(f.re, yr[t], cr[t]);
                cr[t] = Mfma<T>::op(f.im, CONJ ? yi[t] : -yi[t], cr[t]);
                ci[t] = Mfma<T>::op(f.re, yi[t], ci[t]);
                ci[t] = Mfma<T>::op(f.im, CONJ ? -yr[t] : yr[t], ci[t]);
            }
        }
    }
    if constexpr (M3) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const T a = cr[t][v], b = ci[t][v], c = c3[t][v];
                cr[t][v] = CONJ ? a + b : a - b;
                ci[t][v] = CONJ ? a - b - c : c - a - b;
            }
    }
}

// (cr, ci) = Y F (CONJ: Y conj(F)) for this wavefront's row strip: Y[ti*16 + i][k] from the
// slab, F[k][j] from global memory.
template <typename T, int NT, bool CONJ, bool M3>
__device__ __forceinline__ void big_right(const T *ar, const T *ai, const cx<T> *__restrict__ F,
                                          typename Mfma<T>::acc (&cr)[NT],
                                          typename Mfma<T>::acc (&ci)[NT]) {
    using G = BigGeom<NT>;
    using acc_t = typename Mfma<T>::acc;
    const int lane = threadIdx.x & 63, l16 = lane & 15, k4 = lane >> 4, ti = threadIdx.x >> 6;
    acc_t c3[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) cr[t] = ci[t] = c3[t] = acc_t{0, 0, 0, 0};
    const int arow = (ti * 16 + l16) * G::S;
#pragma unroll 4
    for (int kk = 0; kk < G::NP / 4; ++kk) {
        const int k = kk * 4 + k4;
        const T xr = ar[arow + k], xi = ai[arow + k];
        cx<T> f[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) f[t] = F[(size_t)k * G::NP + t * 16 + l16];
        if constexpr (M3) {
            const T xs = CONJ ? xr - xi : xr + xi;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                cr[t] = Mfma<T>::op(xr, f[t].re, cr[t]);
                ci[t] = Mfma<T>::op(xi, f[t].im, ci[t]);
                c3[t] = Mfma<T>::op(xs, f[t].re + f[t].im, c3[t]);
            }
        } else {
            // x F: re += xr fr - xi fi, im += xr fi + xi fr; x conj(F): re += xr fr + xi fi,
            // im += xi fr - xr fi
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                cr[t] = Mfma<T>::op(xr, f[t].re, cr[t]);
                cr[t] = Mfma<T>::op(CONJ ? xi : -xi, f[t].im, cr[t]);
                ci[t] = Mfma<T>::op(CONJ ? -xr : xr, f[t].im, ci[t]);
                ci[t] = Mfma<T>::op(xi, f[t].re, ci[t]);
            }
        }
    }
    if constexpr (M3) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const T a = cr[t][v], b = ci[t][v], c = c3[t][v];
                cr[t][v] = CONJ ? a + b : a - b;
                ci[t][v] = CONJ ? a - b - c : c - a - b;
            }
    }
}

template <typename T, int NT>
__device__ __forceinline__ void big_store(T *sr, T *si, const typename Mfma<T>::acc (&cr)[NT],
                                          const typename Mfma<T>::acc (&ci)[NT]) {
    using G = BigGeom<NT>;
    const int lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4, ti = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = ti * 16 + Mfma<T>::row(g, v), c = t * 16 + col;
            sr[r * G::S + c] = cr[t][v];
            si[r * G::S + c] = ci[t][v];
        }
}

template <typename T, int NT, bool SQ>
__global__ __launch_bounds__(64 * NT) void dft_coupling_big_kernel(
    Layout L, const cx<T> *__restrict__ Fx, const cx<T> *__restrict__ Fy,
    const cx<T> *__restrict__ Vhat, const cx<T> *__restrict__ x, cx<T> *__restrict__ y,
    const int *__restrict__ active, const double *__restrict__ x_norm, cx<T> *__restrict__ x_copy) {
    using G = BigGeom<NT>;
    using acc_t = typename Mfma<T>::acc;
    constexpr bool M3 = SQ;
    const int nz = L.nz;
    int item;
    if (!xcd_order(L.n_inst * nz, item)) return;
    const int b = item / nz, z = item - b * nz;
    if (active && !active[b]) return;
    extern __shared__ double dft_smem[];
    T *sr = reinterpret_cast<T *>(dft_smem);
    T *si = sr + G::PLANE;
    const int nc = L.nc[b];
    const long long off = L.off[b];
    const T sc = x_copy ? (T)(1.0 / x_norm[b]) : T(1);
    const int lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4, ti = threadIdx.x >> 6;
    {
        acc_t *p4 = reinterpret_cast<acc_t *>(sr);  // both planes
        constexpr int N4 = 2 * G::PLANE * (int)sizeof(T) / (int)sizeof(acc_t);
        for (int idx = threadIdx.x; idx < N4; idx += blockDim.x) p4[idx] = acc_t{0, 0, 0, 0};
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nc; j += blockDim.x) {
        const int p = L.chx[off + j] * G::S + L.chy[off + j];
        const long long idx = (off + j) * nz + z;
        cx<T> v = x[idx];
        if (x_copy) {
            v = mk(v.re * sc, v.im * sc);
            x_copy[idx] = v;
        }
        sr[p] = v.re;
        si[p] = v.im;
    }
    __syncthreads();
    acc_t cr[NT], ci[NT];
    const cx<T> *FyR = SQ ? Fx : Fy;
    // (1) Y = F_x X
    big_left<T, NT, false, M3>(Fx, sr, si, cr, ci);
    __syncthreads();
    big_store<T, NT>(sr, si, cr, ci);
    __builtin_amdgcn_wave_barrier();
    // (2) Z = Y F_y, then Z .* Vhat(z), back into this wavefront's strip
    big_right<T, NT, false, M3>(sr, si, FyR, cr, ci);
    {
        const cx<T> *vz = Vhat + (size_t)z * G::NP * G::NP;
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const cx<T> h = vz[(ti * 16 + Mfma<T>::row(g, v)) * G::NP + t * 16 + col];
                const T a = cr[t][v], c = ci[t][v];
                cr[t][v] = a * h.re - c * h.im;
                ci[t][v] = a * h.im + c * h.re;
            }
    }
    __builtin_amdgcn_wave_barrier();
    big_store<T, NT>(sr, si, cr, ci);
    __syncthreads();
    // (3) W = conj(F_x) Zhat
    big_left<T, NT, true, M3>(Fx, sr, si, cr, ci);
    __syncthreads();
    big_store<T, NT>(sr, si, cr, ci);
    __builtin_amdgcn_wave_barrier();
    // (4) out = W conj(F_y)
    big_right<T, NT, true, M3>(sr, si, FyR, cr, ci);
    __builtin_amdgcn_wave_barrier();
    big_store<T, NT>(sr, si, cr, ci);
    __syncthreads();
    for (int j = threadIdx.x; j < nc; j += blockDim.x) {
        const int p = L.chx[off + j] * G::S + L.chy[off + j];
        y[(off + j) * nz + z] = mk(sr[p], si[p]);
    }
}

template <typename T, int NT>
static void launch_dft_big(const Layout &L, const DftTables<T> &D, const cx<T> *x, cx<T> *y,
                           const int *active, hipStream_t s, const double *x_norm, cx<T> *x_copy) {
    using G = BigGeom<NT>;
    if (L.nm) throw Error(3, "DFT coupling beyond 32 x 32: channel-major vectors only");
    const size_t shm = sizeof(T) * 2 * (size_t)G::PLANE;
    const long long items = (long long)L.n_inst * L.nz;
    if (items * 8 >= (1LL << 31)) throw Error(3, "DFT coupling: too many (instance, node) pairs");
    const bool sq = L.nkx == L.nky;
    const dim3 grid((unsigned)(8 * ((items + 7) / 8)));
    if (sq) {
        set_dynamic_lds((const void *)dft_coupling_big_kernel<T, NT, true>, shm);
        hipLaunchKernelGGL((dft_coupling_big_kernel<T, NT, true>), grid, dim3(64 * NT), shm, s, L,
                           D.Fx, D.Fy, D.Vhat, x, y, active, x_norm, x_copy);
    } else {
        set_dynamic_lds((const void *)dft_coupling_big_kernel<T, NT, false>, shm);
        hipLaunchKernelGGL((dft_coupling_big_kernel<T, NT, false>), grid, dim3(64 * NT), shm, s, L,
                           D.Fx, D.Fy, D.Vhat, x, y, active, x_norm, x_copy);
    }
    MSX_LAUNCH_CHECK();
}

int dft_block_nodes(const Layout &L, bool fp64, int active_hint, int nodes_per_wave) {
    const int WPB = fp64 ? 2 : 4;
    int zpb = (nodes_per_wave > 0 ? nodes_per_wave : 4) * WPB;
    if (active_hint >= 0) {
        const long long waves = (long long)active_hint * ((L.nz + zpb - 1) / zpb) * WPB;
        if (waves < 8192) zpb = WPB;
    }
    return zpb;
}

template <typename T, int NP, int WPB, bool SQ, bool NM>
static void launch_dft_impl(const Layout &L, const DftTables<T> &D, const cx<T> *x, cx<T> *y,
                            const int *active, hipStream_t s, int active_hint,
                            const double *x_norm, cx<T> *x_copy, int nodes_per_wave) {
    using G = DftSlab<T, NP>;
    const size_t shm = sizeof(T) * (size_t)WPB * 2 * G::PLANE;
    set_dynamic_lds((const void *)dft_coupling_kernel<T, NP, WPB, SQ, NM>, shm);
    const long long tot = L.total > 0 ? L.total : (long long)L.max_nc * L.n_inst;
    if ((long long)L.nz * (tot + 64) >= (1LL << 31))
        throw Error(3, "DFT coupling: batch state beyond 2^31 elements");
    // four nodes per wavefront (the next node's data is prefetched during the current one);
    // one node per wavefront when so few instances iterate that four-node groups would leave
    // the device short of wavefronts (fewer than about eight per SIMD)
    const int zpb = dft_block_nodes(L, sizeof(T) == 8, active_hint, nodes_per_wave);
    const int nzb = (L.nz + zpb - 1) / zpb;
    const dim3 grid = NM ? dim3(L.n_inst, nzb) : dim3(8 * ((L.n_inst * nzb + 7) / 8));
    hipLaunchKernelGGL((dft_coupling_kernel<T, NP, WPB, SQ, NM>), grid, dim3(64 * WPB), shm, s, L,
                       D.Fx, D.Fy, D.Vhat, x, y, active, zpb, x_norm, x_copy);
    MSX_LAUNCH_CHECK();
}

template <typename T, int NP, int WPB>
static void launch_dft_np(const Layout &L, const DftTables<T> &D, const cx<T> *x, cx<T> *y,
                          const int *active, hipStream_t s, int active_hint, const double *x_norm,
                          cx<T> *x_copy, int npw) {
    const bool sq = L.nkx == L.nky;
    if (L.nm) {
        if (sq) launch_dft_impl<T, NP, WPB, true, true>(L, D, x, y, active, s, active_hint, x_norm, x_copy, npw);
        else launch_dft_impl<T, NP, WPB, false, true>(L, D, x, y, active, s, active_hint, x_norm, x_copy, npw);
        return;
    }
    if (sq) launch_dft_impl<T, NP, WPB, true, false>(L, D, x, y, active, s, active_hint, x_norm, x_copy, npw);
    else launch_dft_impl<T, NP, WPB, false, false>(L, D, x, y, active, s, active_hint, x_norm, x_copy, npw);
}

template <typename T>
void launch_dft_coupling(const Layout &L, const DftTables<T> &D, const cx<T> *x, cx<T> *y,
                         const int *active, hipStream_t s, int active_hint, const double *x_norm,
                         cx<T> *x_copy, int nodes_per_wave) {
    constexpr int WPB = sizeof(T) == 4 ? 4 : 2;  // LDS per wavefront: 9 KB FP32, 18 KB FP64
    if ((x_norm == nullptr) != (x_copy == nullptr))
        throw Error(3, "DFT coupling: x_norm and x_copy go together");
    if (nodes_per_wave < 0 || nodes_per_wave > 16) throw Error(3, "DFT coupling: 0..16 nodes per wave");
    if (D.np == 16) launch_dft_np<T, 16, WPB>(L, D, x, y, active, s, active_hint, x_norm, x_copy, nodes_per_wave);
    else if (D.np == 32) launch_dft_np<T, 32, WPB>(L, D, x, y, active, s, active_hint, x_norm, x_copy, nodes_per_wave);
    else if (D.np == 48) launch_dft_big<T, 3>(L, D, x, y, active, s, x_norm, x_copy);
    else if (D.np == 64) launch_dft_big<T, 4>(L, D, x, y, active, s, x_norm, x_copy);
    else if (D.np == 80) launch_dft_big<T, 5>(L, D, x, y, active, s, x_norm, x_copy);
    else throw Error(3, "DFT coupling supports tables up to 80 x 80");
}
template void launch_dft_coupling<float>(const Layout &, const DftTables<float> &,
                                         const cx<float> *, cx<float> *, const int *, hipStream_t,
                                         int, const double *, cx<float> *, int);
template void launch_dft_coupling<double>(const Layout &, const DftTables<double> &,
                                          const cx<double> *, cx<double> *, const int *,
                                          hipStream_t, int, const double *, cx<double> *, int);

// ---------------------------------------------------------------------------------------
// Gradient of the coupling with respect to the table: the adjoint of the convolution above,
// a circular cross-correlation on the G grid,
//     C[g](z) = sum_b sum_j lam_j(z) psi_i(z),   G_i = G_j - g (mod grid),
// which per condition and node is F^-1( F(Lam) .* F(Psi') ): Lam the adjoint state on the
// grid at G_j, Psi' the forward state at the reflected points -G_i. The products are summed
// over conditions in the transformed domain and inverted once per node, so a (condition,
// node) pair costs two forward transforms (four small complex GEMMs on the matrix cores)
// instead of nc^2 gathered products. FP64 (the adjoint's precision), channel-major states.
// dft_gradient_kernel: one wavefront per node, conditions [y * per_chunk, ...) of chunk y,
// the summed product written to part[y][z][kx][ky] (the accumulator tiles' element order);
// a chunk's node blocks run consecutively on one XCD.
// dft_gradient_finish_kernel: one wavefront per node sums the chunks, applies
// conj(F_x) . conj(F_y) / (nkx nky) and adds conj(-C) (Re(-C) at g = 0, where only Re V_00
// enters the problem) to grad[g][z].
// ---------------------------------------------------------------------------------------
template <int NP, bool SQ>
struct DftRegs {
    static constexpr int NT = DftGeom<NP>::NT, KK = NP / 4;
    double fxr[NT][KK], fxi[NT][KK], fyr[NT][KK], fyi[NT][KK];
    __device__ __forceinline__ void load(const cx<double> *Fx, const cx<double> *Fy) {
        const int lane = threadIdx.x & 63, l16 = lane & 15, k4 = lane >> 4;
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int kk = 0; kk < KK; ++kk) {
                const cx<double> f = Fx[(t * 16 + l16) * NP + kk * 4 + k4];
                fxr[t][kk] = f.re;
                fxi[t][kk] = f.im;
                if constexpr (!SQ) {
                    const cx<double> g = Fy[(kk * 4 + k4) * NP + t * 16 + l16];
                    fyr[t][kk] = g.re;
                    fyi[t][kk] = g.im;
                }
            }
    }
};

// The 2-D transform of the grid in (sr, si) (row stride SB): (cr, ci) = F_x X F_y, or with
// CONJ = conj(F_x) X conj(F_y); the slab is overwritten by the intermediate.
template <int NP, bool SQ, bool CONJ>
__device__ __forceinline__ void dft2d(DftRegs<NP, SQ> &F, double *sr, double *si,
                                      typename Mfma<double>::acc (&cr)[DftGeom<NP>::NT][DftGeom<NP>::NT],
                                      typename Mfma<double>::acc (&ci)[DftGeom<NP>::NT][DftGeom<NP>::NT]) {
    using G = DftGeom<NP>;
    if constexpr (SQ) {
        typename Mfma<double>::acc c3[G::NT][G::NT];
        cgemm3_reg_a<double, NP, CONJ>(F.fxr, F.fxi, sr, si, G::SB, cr, ci, c3);
        combine3<double, NP, CONJ>(cr, ci, c3);
        __builtin_amdgcn_wave_barrier();
        store_tiles<double, NP>(sr, si, G::SA, cr, ci);
        __builtin_amdgcn_wave_barrier();
        cgemm3_reg_b<double, NP, CONJ>(sr, si, G::SA, F.fxr, F.fxi, cr, ci, c3);
        combine3<double, NP, CONJ>(cr, ci, c3);
    } else {
        cgemm_reg_a<double, NP>(F.fxr, F.fxi, CONJ, sr, si, G::SB, cr, ci);
        __builtin_amdgcn_wave_barrier();
        store_tiles<double, NP>(sr, si, G::SA, cr, ci);
        __builtin_amdgcn_wave_barrier();
        cgemm_reg_b<double, NP>(sr, si, G::SA, F.fyr, F.fyi, CONJ, cr, ci);
    }
}

template <int NP>
__device__ __forceinline__ void zero_slab(double *sr) {
    using acc_t = typename Mfma<double>::acc;
    acc_t *p4 = reinterpret_cast<acc_t *>(sr);  // both planes (contiguous)
    constexpr int N4 = 2 * DftGeom<NP>::PLANE * (int)sizeof(double) / (int)sizeof(acc_t);
    for (int idx = threadIdx.x & 63; idx < N4; idx += 64) p4[idx] = acc_t{0, 0, 0, 0};
}

template <int NP, int WPB, bool SQ>
__global__ __launch_bounds__(64 * WPB, 1) void dft_gradient_kernel(
    Layout L, const cx<double> *__restrict__ Fx, const cx<double> *__restrict__ Fy,
    const cx<double> *__restrict__ lam, const cx<double> *__restrict__ psi,
    cx<double> *__restrict__ part, int per_chunk, int n_chunks) {
    using G = DftGeom<NP>;
    using acc_t = typename Mfma<double>::acc;
    constexpr int NT = G::NT, SA = G::SA, SB = G::SB;
    const int nz = L.nz, nkx = L.nkx, nky = L.nky;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, g = lane >> 4;
    // node blocks of a chunk consecutive on one XCD: neighbouring nodes of a channel share the
    // 128-byte lines of the channel-major states
    const int zblocks = (nz + WPB - 1) / WPB;
    int item;
    if (!xcd_order(zblocks * n_chunks, item)) return;
    const int chunk = item / zblocks;
    const int z = (item - chunk * zblocks) * WPB + wave;
    if (z >= nz) return;  // no workgroup barrier below: a wavefront may leave on its own
    extern __shared__ double dft_smem[];
    double *sr = dft_smem + (size_t)wave * 4 * G::PLANE;
    double *si = sr + G::PLANE;
    double *hr = si + G::PLANE;  // F(Lam) of the current condition, accumulator element order
    double *hi = hr + G::PLANE;
    DftRegs<NP, SQ> F;
    F.load(Fx, Fy);
    acc_t ar[NT][NT], ai[NT][NT], cr[NT][NT], ci[NT][NT];
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj) ar[ti][tj] = ai[ti][tj] = acc_t{0, 0, 0, 0};
    const int b0 = chunk * per_chunk;
    const int b1 = min(L.n_inst, b0 + per_chunk);
    using plain_t = std::integral_constant<bool, false>;
    for (int b = b0; b < b1; ++b) {
        const int nc = L.nc[b];
        const long long off = L.off[b];
        // F(Lam): the adjoint state at the channels' own grid points
        zero_slab<NP>(sr);
        __builtin_amdgcn_wave_barrier();
        for (int j = lane; j < nc; j += 64) {
            const int p = L.chx[off + j] * SB + L.chy[off + j];
            const cx<double> v = lam[(off + j) * nz + z];
            sr[p] = v.re;
            si[p] = v.im;
        }
        __builtin_amdgcn_wave_barrier();
        dft2d<NP, SQ, plain_t::value>(F, sr, si, cr, ci);
        __builtin_amdgcn_wave_barrier();
        store_tiles<double, NP>(hr, hi, SA, cr, ci);
        // F(Psi'): the forward state at the reflected grid points -G_i
        zero_slab<NP>(sr);
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < nc; i += 64) {
            const int rx = L.chx[off + i], ry = L.chy[off + i];
            const int p = (rx == 0 ? 0 : nkx - rx) * SB + (ry == 0 ? 0 : nky - ry);
            const cx<double> v = psi[(off + i) * nz + z];
            sr[p] = v.re;
            si[p] = v.im;
        }
        __builtin_amdgcn_wave_barrier();
        dft2d<NP, SQ, plain_t::value>(F, sr, si, cr, ci);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = ti * 16 + Mfma<double>::row(g, v), c = tj * 16 + col;
                    const double lr = hr[r * SA + c], li = hi[r * SA + c];
                    const double pr = cr[ti][tj][v], pi = ci[ti][tj][v];
                    ar[ti][tj][v] = fma(lr, pr, fma(-li, pi, ar[ti][tj][v]));
                    ai[ti][tj][v] = fma(lr, pi, fma(li, pr, ai[ti][tj][v]));
                }
        __builtin_amdgcn_wave_barrier();
    }
    cx<double> *out = part + ((size_t)chunk * nz + z) * NP * NP;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = ti * 16 + Mfma<double>::row(g, v), c = tj * 16 + col;
                out[r * NP + c] = mk(ar[ti][tj][v], ai[ti][tj][v]);
            }
}

template <int NP, int WPB, bool SQ>
__global__ __launch_bounds__(64 * WPB, 1) void dft_gradient_finish_kernel(
    Layout L, const cx<double> *__restrict__ Fx, const cx<double> *__restrict__ Fy,
    const cx<double> *__restrict__ part, int n_chunks, cx<double> *__restrict__ grad) {
    using G = DftGeom<NP>;
    using acc_t = typename Mfma<double>::acc;
    constexpr int NT = G::NT, SB = G::SB;
    const int nz = L.nz, nky = L.nky, nG = L.nkx * L.nky;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int z = blockIdx.x * WPB + wave;
    if (z >= nz) return;
    extern __shared__ double dft_smem[];
    double *sr = dft_smem + (size_t)wave * 2 * G::PLANE;
    double *si = sr + G::PLANE;
    DftRegs<NP, SQ> F;
    F.load(Fx, Fy);
    for (int idx = lane; idx < NP * NP; idx += 64) {
        cx<double> acc = mk(0.0, 0.0);
        for (int c = 0; c < n_chunks; ++c) acc = acc + part[((size_t)c * nz + z) * NP * NP + idx];
        const int r = idx / NP, cc = idx - r * NP;
        sr[r * SB + cc] = acc.re;
        si[r * SB + cc] = acc.im;
    }
    __builtin_amdgcn_wave_barrier();
    acc_t cr[NT][NT], ci[NT][NT];
    using conj_t = std::integral_constant<bool, true>;
    dft2d<NP, SQ, conj_t::value>(F, sr, si, cr, ci);
    __builtin_amdgcn_wave_barrier();
    store_tiles<double, NP>(sr, si, SB, cr, ci);
    __builtin_amdgcn_wave_barrier();
    const double scale = 1.0 / (double)nG;
    for (int gidx = lane; gidx < nG; gidx += 64) {
        const int gx = gidx / nky, gy = gidx - gx * nky;
        const double re = sr[gx * SB + gy] * scale, im = si[gx * SB + gy] * scale;
        const cx<double> add = gidx == 0 ? mk(-re, 0.0) : mk(-re, im);  // conj(-C)
        cx<double> &dst = grad[(size_t)gidx * nz + z];
        dst = dst + add;
    }
}

template <int NP, bool SQ>
static void launch_dft_gradient_np(const Layout &L, const DftTables<double> &D,
                                   const cx<double> *lam, const cx<double> *psi, cx<double> *grad,
                                   DBuf<cx<double>> &part, hipStream_t s) {
    using G = DftGeom<NP>;
    constexpr int WPB = 2;
    const int zblocks = (L.nz + WPB - 1) / WPB;
    // enough workgroups to fill the device (about 1,024 of two wavefronts), at most one
    // chunk per condition
    const int want = std::max(1, (1024 + zblocks - 1) / zblocks);
    const int n_chunks = std::min(L.n_inst, want);
    const int per_chunk = (L.n_inst + n_chunks - 1) / n_chunks;
    const int used = (L.n_inst + per_chunk - 1) / per_chunk;
    const size_t need = (size_t)used * L.nz * NP * NP;
    if (part.n < need) part.alloc(need);
    const size_t shm = sizeof(double) * (size_t)WPB * 4 * G::PLANE;
    set_dynamic_lds((const void *)dft_gradient_kernel<NP, WPB, SQ>, shm);
    hipLaunchKernelGGL((dft_gradient_kernel<NP, WPB, SQ>), dim3(8 * ((zblocks * used + 7) / 8)),
                       dim3(64 * WPB), shm, s, L, D.Fx, D.Fy, lam, psi, part.p, per_chunk, used);
    MSX_LAUNCH_CHECK();
    const size_t shm2 = sizeof(double) * (size_t)WPB * 2 * G::PLANE;
    set_dynamic_lds((const void *)dft_gradient_finish_kernel<NP, WPB, SQ>, shm2);
    hipLaunchKernelGGL((dft_gradient_finish_kernel<NP, WPB, SQ>), dim3(zblocks), dim3(64 * WPB),
                       shm2, s, L, D.Fx, D.Fy, part.p, used, grad);
    MSX_LAUNCH_CHECK();
}

void launch_dft_gradient(const Layout &L, const DftTables<double> &D, const cx<double> *lam,
                         const cx<double> *psi, cx<double> *grad, DBuf<cx<double>> &part,
                         hipStream_t s) {
    if (L.nm) throw Error(3, "DFT gradient: channel-major states only");
    const bool sq = L.nkx == L.nky;
    if (D.np == 16) {
        if (sq) launch_dft_gradient_np<16, true>(L, D, lam, psi, grad, part, s);
        else launch_dft_gradient_np<16, false>(L, D, lam, psi, grad, part, s);
    } else if (D.np == 32) {
        if (sq) launch_dft_gradient_np<32, true>(L, D, lam, psi, grad, part, s);
        else launch_dft_gradient_np<32, false>(L, D, lam, psi, grad, part, s);
    } else {
        throw Error(3, "DFT gradient supports tables up to 32 x 32");
    }
}

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
    if (m <= 48) return 48;
    if (m <= 64) return 64;
    if (m <= 80) return 80;
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
