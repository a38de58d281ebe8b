// All-channel inverse diagonal blocks of the mixed solver's FP32 inner operator with split-FP16
// operands on the matrix cores:
//
//   y_j = D_j^-1 x_j = Q diag(1/(e_j + lam)) Q^T x_j + f_j g_j(N) c_j / (1 - f_j(N) c_j)
//
// (the reference applies D_j^-1 channel by channel inside its forward substitution,
// lib/multiscat_fortran/fortran/multiscat_gmres.f90:666-685; its SciPy backend applies D^-1
// to all channels at once, multiscat/multiscat/_scipy.py:249-259,336-354).
//
// Every FP32 operand a of the two products is carried as a pair of FP16 values scaled by a
// power of two s per matrix column: s a = hi + lo, hi = fp16(s a), lo = fp16(s a - hi), so the
// pair holds s a to 2^-22 relative (FP32: 2^-24). A product then takes three 16x16x32 FP16
// MFMAs with FP32 accumulation, A_hi B_hi + A_hi B_lo + A_lo B_hi (the dropped A_lo B_lo is
// below 2^-22 of the product), and the column scale is removed exactly from the FP32
// accumulators; the MFMA rate is 16 times FP32's. On the 31 x 31, nz = 150 hexagonal table
// (H0's spectrum spans 2e6, so eigenvector entries span many decades) D^-1 of random vectors
// differs from an FP64 solve by 4e-7 to 1e-5 normwise, as the FP32-operand kernel does (the
// FP32 reciprocals and the Sherman-Morrison quotient dominate), and the mixed solver takes
// the same sweeps and iterations with either kernel.
//
// A wavefront owns 16 channels: their real and imaginary parts are two 16-column tiles.
//   W^T = Q^T X^T   A = Q^T (m x z, from LDS), B = X^T (z x column), K = z
//   W'  = W^T scaled by 1/(e_j + lam_m) in FP32 (e and lam summed in FP64), 0 for m >= nz,
//         plus the Sherman-Morrison term corr_j Q^T f_j (Q^T f_j = Q[N-1][m] / (e_j + lam_m),
//         corr_j from g_j(N) = Q[N-1][:] W' and the table's f_j(N)), so that
//   Y^T = Q W'      A = Q (z x m), B = W' (m x column), K = m, is D^-1 x itself: the
//         factor vectors f_j are not read (the 8-channel form below adds corr_j f_j after
//         the second product instead)
// K runs in blocks of 32; lane group g (lane / 16) supplies the 8 K values of its operands.
// Product 1 takes z = 32 s + 8 g + u in block s. The B operand of product 2 is formed in place
// from product 1's accumulators: lane group g of m tile mt holds rows 16 mt + 4 g + v, so in
// block p it takes m = 32 p + 16 h + 4 g + v (h = 0, 1 for tiles 2p, 2p + 1; v = 0..3) as its
// 8 K values, and the Q plane of product 2 is stored with its columns in that order (plane
// column 32 p + 8 g + 4 h + v holds Q[z][32 p + 16 h + 4 g + v]), one 16-byte read per lane.
//
// Planes (built with the table from the FP64 eigenvectors): P1 = Q^T rows m, columns z; P2 =
// Q rows z, columns m in the order above; each of 2^15 Q as hi and lo (|Q| <= 1: the scale
// keeps hi below FP16's 65504 and moves lo out of FP16's subnormal range for entries down to
// 2^-18, below which the pair still holds Q to 2^-39 absolute; unscaled, lo of a typical
// entry ~0.08 was already subnormal), NQ = 32 ceil(nz / 32) columns of
// stride NQ + 8 halves (16-byte reads of 8 lanes fall in distinct banks). P1 hi, P2 hi and
// (when it fits) P1 lo sit in LDS, rows below nz only: rows past nz of P1 give W rows that the
// scaling zeroes, rows past nz of P2 give outputs that are never stored, so reads there are
// clamped to row nz - 1. P2 lo (and P1 lo when LDS is short) is read from L2, a z group ahead.
#include "msx_internal.h"
#include "msx_mfma.h"
#include "msx_split.h"

#include <type_traits>

namespace msx {


constexpr int H3_WAVES = 8;
constexpr int H3_ZG = 2;  // z tiles per group of the second product
constexpr float H3_QSCALE = 32768.0f;  // the planes hold 2^15 Q

__device__ __forceinline__ v4f mfma_h(h8 a, h8 b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Power of two that brings a column maximum m >= 0 into [2^14, 2^15) (FP16 holds 65504); a
// zero or tiny maximum gets at most 2^115, which keeps the scaled values finite.
__device__ __forceinline__ float h3_scale(float m) {
    int e = 0;
    (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
    return ldexpf(1.0f, 15 - max(e, -100));
}
// The same scale as an exponent k (scale 2^k), whose removal together with the planes' 2^15
// is the exact 2^(-k-15)
__device__ __forceinline__ int h3_scale_exp(float m) {
    int e = 0;
    (void)frexpf(m, &e);
    return 15 - max(e, -100);
}

// a / b for the Sherman-Morrison quotient: Smith's form (the larger of |b.re|, |b.im| divides)
// with hardware reciprocals and selects instead of branches and IEEE divisions
__device__ __forceinline__ cx<float> cdiv_rcp(cx<float> a, cx<float> b) {
    const bool sw = fabsf(b.re) < fabsf(b.im);
    const float p = sw ? b.im : b.re, q = sw ? b.re : b.im;
    const float u = sw ? a.im : a.re, w = sw ? a.re : a.im;
    const float r = q * __builtin_amdgcn_rcpf(p);
    const float id = __builtin_amdgcn_rcpf(fmaf(q, r, p));
    const float re = fmaf(w, r, u) * id, im = fmaf(-u, r, w) * id;
    return mk(re, sw ? -im : im);
}

__device__ __forceinline__ float wave_max4(float v) {
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}
__device__ __forceinline__ float wave_sum4(float v) {
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}

// First tile of wavefront `wave` of a persistent launch with `waves` wavefronts per workgroup,
// which takes gridDim.x * waves consecutive tiles per round. With a tile for every wavefront,
// a workgroup takes consecutive tiles and consecutive workgroups of one XCD (workgroups are
// dealt to the 8 XCDs in turn) take consecutive runs of them: the rows of neighbouring tiles
// share cache lines (a tile's row of one node is 64 or 128 bytes at any alignment), which are
// then fetched into one L2 rather than two. With fewer tiles, tile t goes to wavefront
// t / gridDim.x of workgroup t % gridDim.x, which gives each tile its own SIMD.
__device__ __forceinline__ int first_tile(int n_tiles, int waves, int wave) {
    const int G = gridDim.x;
    if (n_tiles < G * waves) return wave * G + blockIdx.x;
    const int b = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
    return b * waves + wave;
}

// P1LO: P1 lo in LDS (else read from L2 like P2 lo). BASE: y = base + sign D^-1 x.
// Node-major vectors and factors only (element (j, z) of an instance at off nz + z nc + j).
template <int KT, bool P1LO, bool BASE>
__global__ __launch_bounds__(64 * H3_WAVES) void diag_h3_kernel(
    Layout L, H3Planes H, const double *__restrict__ lam_g, const double *__restrict__ e,
    const cx<float> *__restrict__ c, const float *__restrict__ fac, const cx<float> *x,
    const cx<float> *base, cx<float> *y, float sign, const int *__restrict__ active) {
    constexpr int NQ = 32 * KT, MT = 2 * KT, S = NQ + 8;
    extern __shared__ double h3_smem[];
    __shared__ int wave_sum[H3_WAVES];
    const int nz = L.nz, n_inst = L.n_inst;
    double *lam = h3_smem;                                    // [NQ], 0 past nz
    float *qn = reinterpret_cast<float *>(lam + NQ);           // [NQ]: Q[nz - 1][m]
    _Float16 *p1h = reinterpret_cast<_Float16 *>(qn + NQ);    // [nz][S]
    _Float16 *p2h = p1h + nz * S;                              // [nz][S]
    _Float16 *p1l = p2h + nz * S;                              // [nz][S] (P1LO)
    int *tile_start = reinterpret_cast<int *>(p1l + (P1LO ? nz * S : 0));  // [n_inst + 1]
    int *inst_nc = tile_start + n_inst + 1;                    // [n_inst]
    int *inst_off = inst_nc + n_inst;                          // [n_inst]
    const _Float16 *g1l = H.planes + (size_t)NQ * S;           // P1 lo in global memory
    const _Float16 *g2l = H.planes + (size_t)3 * NQ * S;       // P2 lo in global memory
    {
        // exclusive scan of the active instances' 16-channel tiles: per-thread serial sums
        // over a contiguous chunk, then a scan across the threads
        const int per = (n_inst + blockDim.x - 1) / blockDim.x;
        const int i0 = threadIdx.x * per, i1 = min(n_inst, i0 + per);
        int local = 0;
        for (int i = i0; i < i1; ++i) local += (active && !active[i]) ? 0 : (L.nc[i] + 15) / 16;
        int incl = local;
        const int ln = threadIdx.x & 63;
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o);
            if (ln >= o) incl += t;
        }
        if (ln == 63) wave_sum[threadIdx.x >> 6] = incl;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) before += wave_sum[w];
        int run = before + incl - local;
        for (int i = i0; i < i1; ++i) {
            const int nci = L.nc[i];
            tile_start[i] = run;
            inst_nc[i] = nci;
            inst_off[i] = (int)L.off[i];
            run += (active && !active[i]) ? 0 : (nci + 15) / 16;
        }
        if (threadIdx.x == blockDim.x - 1) tile_start[n_inst] = before + incl;
    }
    for (int m = threadIdx.x; m < NQ; m += blockDim.x) {
        lam[m] = m < nz ? lam_g[m] : 0.0;
        qn[m] = m < nz ? H.qlast[m] : 0.0f;
    }
    {
        // rows 0..nz-1 of the LDS planes, 16 bytes at a time (rows are contiguous at stride S
        // in both copies)
        const int n16 = nz * S / 8;
        const uint4 *s1h = reinterpret_cast<const uint4 *>(H.planes);
        const uint4 *s1l = reinterpret_cast<const uint4 *>(g1l);
        const uint4 *s2h = reinterpret_cast<const uint4 *>(H.planes + (size_t)2 * NQ * S);
        for (int i = threadIdx.x; i < n16; i += blockDim.x) {
            reinterpret_cast<uint4 *>(p1h)[i] = s1h[i];
            reinterpret_cast<uint4 *>(p2h)[i] = s2h[i];
            if (P1LO) reinterpret_cast<uint4 *>(p1l)[i] = s1l[i];
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int l16 = lane & 15, g = lane >> 4;
    const int n_tiles = tile_start[n_inst];
    const int t_step = gridDim.x * H3_WAVES;

    // One tile: its instance's channel count and this lane's channel; element indices in 32
    // bits (the launcher checks that the batch's vectors fit), nodes past nz clamped to the
    // channel's last node (a clamp of the index, which depends on the tile, rather than of z,
    // which the compiler would hoist)
    struct Tile {
        int nc, cidx, vb, vlast;
        bool live;
        double ej;
        cx<float> cj;
        float facN;
    };
    float xr[KT][8], xi[KT][8];
    // Sets up tile t and issues its loads: the lane's channel scalars and X at z = 32 s + 8 g + u
    // (past nz: the last node, which P1's zero columns cancel). The loads of the next tile are
    // issued from inside the current one, after its last products and before its last stores,
    // so that their wait includes neither those stores nor the setup.
    auto begin_tile = [&](int t, Tile &T) {
        int lo = 0, hi = n_inst - 1;  // last instance with tile_start <= t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tile_start[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const int j0 = (t - tile_start[lo]) * 16;
        T.nc = inst_nc[lo];
        const int jc = j0 + l16;
        T.live = jc < T.nc;
        T.cidx = inst_off[lo] + (T.live ? jc : 0);
        T.vb = inst_off[lo] * nz + (T.live ? jc : 0);
        T.vlast = T.vb + (nz - 1) * T.nc;
        T.ej = e[T.cidx];
        T.cj = c[T.cidx];
        T.facN = fac[(unsigned)T.vlast];
#pragma unroll
        for (int s = 0; s < KT; ++s)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int z = 32 * s + 8 * g + u;
                const cx<float> v = x[(unsigned)min(T.vb + z * T.nc, T.vlast)];
                xr[s][u] = v.re;
                xi[s][u] = v.im;
            }
    };

    int t = first_tile(n_tiles, H3_WAVES, wave);
    Tile cur, nxt;
    if (t < n_tiles) begin_tile(t, cur);
#pragma unroll 1
    while (t < n_tiles) {
        const int t_next = t + t_step;
        auto vi = [&](int z) -> unsigned { return (unsigned)min(cur.vb + z * cur.nc, cur.vlast); };
        // (nodes known to lie below nz: no clamp)
        auto vu = [&](int z) -> unsigned { return (unsigned)(cur.vb + z * cur.nc); };

        // ---- column maxima and split of the tile's X ----
        float mr = 0.f, mi = 0.f;
#pragma unroll
        for (int s = 0; s < KT; ++s)
#pragma unroll
            for (int u = 0; u < 8; u += 2) {
                mr = max3abs(xr[s][u], xr[s][u + 1], mr);
                mi = max3abs(xi[s][u], xi[s][u + 1], mi);
            }
        const int exr = h3_scale_exp(wave_max4(mr)), exi = h3_scale_exp(wave_max4(mi));
        const float sxr = ldexpf(1.0f, exr), sxi = ldexpf(1.0f, exi);
        h8 xhr[KT], xlr[KT], xhi[KT], xli[KT];
#pragma unroll
        for (int s = 0; s < KT; ++s) {
            split8s(xr[s], sxr, xhr[s], xlr[s]);
            split8s(xi[s], sxi, xhi[s], xli[s]);
        }

        // ---- W^T = Q^T X^T ----
        v4f acc1[MT][2];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc1[mt][0] = acc1[mt][1] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < KT; ++s) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int row = min(16 * mt + l16, nz - 1);
                const h8 ah = *reinterpret_cast<const h8 *>(p1h + row * S + 32 * s + 8 * g);
                const h8 al = P1LO ? *reinterpret_cast<const h8 *>(p1l + row * S + 32 * s + 8 * g)
                                   : *reinterpret_cast<const h8 *>(g1l + row * S + 32 * s + 8 * g);
                acc1[mt][0] = mfma_h(ah, xhr[s], acc1[mt][0]);
                acc1[mt][1] = mfma_h(ah, xhi[s], acc1[mt][1]);
                acc1[mt][0] = mfma_h(ah, xlr[s], acc1[mt][0]);
                acc1[mt][1] = mfma_h(ah, xli[s], acc1[mt][1]);
                acc1[mt][0] = mfma_h(al, xhr[s], acc1[mt][0]);
                acc1[mt][1] = mfma_h(al, xhi[s], acc1[mt][1]);
            }
        }

        // ---- W' = W / (e_j + lam_m) with the column scales removed; y0(N) = Q[N-1][:] W';
        // 1 / (e_j + lam_m) from the FP64 sum by the hardware reciprocal, 0 past nz ----
        // powers of two (exact): the column scales and the planes' 2^15
        const float ixr = ldexpf(1.0f, -exr - 15), ixi = ldexpf(1.0f, -exi - 15);
        float yn0 = 0.f, yn1 = 0.f, wr = 0.f, wi = 0.f;
        float qs[MT][4];  // Q[N-1][m] / (e_j + lam_m), the Q^T of f_j = D0^-1 e_N
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = 16 * mt + 4 * g + v;
                // unconditional loads (lam is 0 past nz), then a select: a condition around
                // the load made every element a branch
                const float r = __builtin_amdgcn_rcpf((float)(cur.ej + lam[m]));
                const float sc = m < nz ? r : 0.f;
                const float q = qn[m];
                qs[mt][v] = q * sc;
                acc1[mt][0][v] *= sc * ixr;
                acc1[mt][1][v] *= sc * ixi;
                yn0 += q * acc1[mt][0][v];
                yn1 += q * acc1[mt][1][v];
            }
        }
        yn0 = wave_sum4(yn0);
        yn1 = wave_sum4(yn1);
        // D^-1 x = D0^-1 x + corr f_j, corr = c y0(N) / (1 - c f_j(N)), f_j = Q (Q^T f_j):
        // the correction joins W' before the second product, which then gives D^-1 x itself
        cx<float> corr = mk(0.f, 0.f);
        if (cur.live) corr = cdiv_rcp(mk(yn0, yn1) * cur.cj, mk(1.0f, 0.0f) - cur.cj * cur.facN);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                acc1[mt][0][v] = fmaf(corr.re, qs[mt][v], acc1[mt][0][v]);
                acc1[mt][1][v] = fmaf(corr.im, qs[mt][v], acc1[mt][1][v]);
            }
#pragma unroll
            for (int v = 0; v < 4; v += 2) {
                wr = max3abs(acc1[mt][0][v], acc1[mt][0][v + 1], wr);
                wi = max3abs(acc1[mt][1][v], acc1[mt][1][v + 1], wi);
            }
        }
        // P2 lo fragments of the first z group, read during the split of W'
        h8 a2l[KT][H3_ZG];
#pragma unroll
        for (int zg = 0; zg < H3_ZG; ++zg) {
            const int zr = min(16 * zg + l16, nz - 1);
#pragma unroll
            for (int p = 0; p < KT; ++p)
                a2l[p][zg] = *reinterpret_cast<const h8 *>(g2l + zr * S + 32 * p + 8 * g);
        }
        const int ewr = h3_scale_exp(wave_max4(wr)), ewi = h3_scale_exp(wave_max4(wi));
        const float swr = ldexpf(1.0f, ewr), swi = ldexpf(1.0f, ewi);
        h8 whr[KT], wlr[KT], whi[KT], wli[KT];
#pragma unroll
        for (int p = 0; p < KT; ++p) {
            float ar[8], ai[8];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                ar[v] = acc1[2 * p][0][v];
                ar[4 + v] = acc1[2 * p + 1][0][v];
                ai[v] = acc1[2 * p][1][v];
                ai[4 + v] = acc1[2 * p + 1][1][v];
            }
            split8s(ar, swr, whr[p], wlr[p]);
            split8s(ai, swi, whi[p], wli[p]);
        }
        const float iwr = ldexpf(1.0f, -ewr - 15), iwi = ldexpf(1.0f, -ewi - 15);

        // ---- Y^T = Q W' by groups of H3_ZG z tiles; base values one group ahead at clamped
        // addresses (unused values are never stored) ----
        cx<float> bz[H3_ZG][4];
        auto load_group = [&](int zt) {
            if constexpr (BASE) {
#pragma unroll
                for (int zg = 0; zg < H3_ZG; ++zg)
#pragma unroll
                    for (int v = 0; v < 4; ++v) bz[zg][v] = base[vi(16 * (zt + zg) + 4 * g + v)];
            }
        };
        // one group: its products, then (the last group) the next tile's loads, then its
        // results and stores, then (the others) the next group's factor and base loads
        auto group = [&](int zt0, auto last_tag) {
            constexpr bool LAST = decltype(last_tag)::value;
            int zrow[H3_ZG];
#pragma unroll
            for (int zg = 0; zg < H3_ZG; ++zg) zrow[zg] = min(16 * (zt0 + zg) + l16, nz - 1);
            v4f acc2[H3_ZG][2];
#pragma unroll
            for (int zg = 0; zg < H3_ZG; ++zg) acc2[zg][0] = acc2[zg][1] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int p = 0; p < KT; ++p)
#pragma unroll
                for (int zg = 0; zg < H3_ZG; ++zg) {
                    const h8 ah = *reinterpret_cast<const h8 *>(p2h + zrow[zg] * S + 32 * p + 8 * g);
                    acc2[zg][0] = mfma_h(ah, whr[p], acc2[zg][0]);
                    acc2[zg][1] = mfma_h(ah, whi[p], acc2[zg][1]);
                    acc2[zg][0] = mfma_h(ah, wlr[p], acc2[zg][0]);
                    acc2[zg][1] = mfma_h(ah, wli[p], acc2[zg][1]);
                }
#pragma unroll
            for (int p = 0; p < KT; ++p)
#pragma unroll
                for (int zg = 0; zg < H3_ZG; ++zg) {
                    acc2[zg][0] = mfma_h(a2l[p][zg], whr[p], acc2[zg][0]);
                    acc2[zg][1] = mfma_h(a2l[p][zg], whi[p], acc2[zg][1]);
                }
            if constexpr (LAST) {
                // after every product of the tile (the scheduler would otherwise move the
                // loads above the products, where W and the new X would both need registers)
                __builtin_amdgcn_sched_barrier(0);
                // unconditional (past the last tile: the last tile again, unused), so that the
                // current X's registers are free for the whole tile: under a condition the old
                // values would have to survive until it
                begin_tile(min(t_next, n_tiles - 1), nxt);
            } else {
                // the next group's P2 lo fragments
#pragma unroll
                for (int zg = 0; zg < H3_ZG; ++zg) {
                    const int zr = min(16 * (zt0 + H3_ZG + zg) + l16, nz - 1);
#pragma unroll
                    for (int p = 0; p < KT; ++p)
                        a2l[p][zg] = *reinterpret_cast<const h8 *>(g2l + zr * S + 32 * p + 8 * g);
                }
            }
            cx<float> out[H3_ZG][4];
#pragma unroll
            for (int zg = 0; zg < H3_ZG; ++zg)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const cx<float> val = mk(acc2[zg][0][v] * iwr, acc2[zg][1][v] * iwi);
                    out[zg][v] = BASE ? bz[zg][v] + val * sign : val;
                }
            if (16 * (zt0 + H3_ZG) <= nz) {
                if (cur.live) {
#pragma unroll
                    for (int zg = 0; zg < H3_ZG; ++zg)
#pragma unroll
                        for (int v = 0; v < 4; ++v) y[vu(16 * (zt0 + zg) + 4 * g + v)] = out[zg][v];
                }
            } else {
#pragma unroll
                for (int zg = 0; zg < H3_ZG; ++zg)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int z = 16 * (zt0 + zg) + 4 * g + v;
                        if (cur.live && zt0 + zg < MT && z < nz) y[vi(z)] = out[zg][v];
                    }
            }
            if constexpr (!LAST) load_group(zt0 + H3_ZG);
            // keeps the loads here: without a possible write after them, the compiler merges
            // them with the loads before the loop into one load at the point of use
            asm volatile("" ::: "memory");
        };
        load_group(0);
#pragma unroll 1
        for (int zt0 = 0; zt0 < MT - H3_ZG; zt0 += H3_ZG) group(zt0, std::false_type{});
        group(MT - H3_ZG, std::true_type{});
        cur = nxt;
        t = t_next;
    }
}

// The same D^-1 with 8 channels per wavefront (H8_WAVES = 12 wavefronts a workgroup, three per
// SIMD): one 16-column tile whose column l16 is the real (l16 even) or imaginary (odd) part of
// channel l16 / 2, so the accumulators, operands and W take half the registers of the
// 16-channel form (the planes in LDS feed half as many products per read). Loads and stores
// are 4-byte parts of the complex elements; the two parts of a channel meet by a lane swap for
// the Sherman-Morrison correction.
constexpr int H8_WAVES = 12;
constexpr int H8_ZG = 1;
// nq 192 and 224 (nz up to 224): the four planes do not fit LDS, so all of them are read from
// L2 (each A fragment once per tile), and the workgroup holds 8 wavefronts of up to 256 VGPRs
// (189 and 213 used; nq 256 would spill)
constexpr int H8L_WAVES = 8;

// L2P: every plane read from L2 (P1LO false; LDS holds lam, the Q row and the tile scan only),
// or with P1HL all but P1 hi, which LDS holds (nz rows: 93 KB at nz = 200), and the first
// H.p2_lds_rows rows of P2 hi (a multiple of 16: whole row tiles of the second product), so a
// tile reads less than three planes from L2 instead of four.
template <int KT, bool P1LO, bool BASE, bool L2P = false, int WAVES = H8_WAVES, bool P1HL = false>
__global__ __launch_bounds__(64 * WAVES) void diag_h8_kernel(
    Layout L, H3Planes H, const double *__restrict__ lam_g, const double *__restrict__ e,
    const cx<float> *__restrict__ c, const float *__restrict__ fac, const cx<float> *x,
    const cx<float> *base, cx<float> *y, float sign, const int *__restrict__ active) {
    static_assert(!(L2P && P1LO), "with the planes in L2, P1 lo is there too");
    static_assert(L2P || !P1HL, "P1HL is a form of the planes-in-L2 kernel");
    constexpr int LPLANES = L2P ? (P1HL ? 1 : 0) : 2;  // hi planes held in LDS
    constexpr int NQ = 32 * KT, MT = 2 * KT, S = NQ + 8;
    extern __shared__ double h3_smem[];
    __shared__ int wave_sum[WAVES];
    const int nz = L.nz, n_inst = L.n_inst;
    double *lam = h3_smem;                                    // [NQ], 0 past nz
    float *qn = reinterpret_cast<float *>(lam + NQ);           // [NQ]: Q[nz - 1][m]
    _Float16 *lp = reinterpret_cast<_Float16 *>(qn + NQ);     // LDS planes
    const int p2r = P1HL ? H.p2_lds_rows : 0;
    _Float16 *p2part = lp + LPLANES * nz * S;                  // [p2r][S]: P2 hi rows (P1HL)
    _Float16 *p1l = p2part + p2r * S;                          // [nz][S] (P1LO)
    const _Float16 *p1h = LPLANES >= 1 ? lp : H.planes;                                  // [nz][S]
    const _Float16 *p2h = LPLANES == 2 ? lp + nz * S : H.planes + (size_t)2 * NQ * S;  // [nz][S]
    int *tile_start = reinterpret_cast<int *>(p1l + (P1LO ? nz * S : 0));  // [n_inst + 1]
    int *inst_nc = tile_start + n_inst + 1;                    // [n_inst]
    int *inst_off = inst_nc + n_inst;                          // [n_inst]
    const _Float16 *g1l = H.planes + (size_t)NQ * S;           // P1 lo in global memory
    const _Float16 *g2l = H.planes + (size_t)3 * NQ * S;       // P2 lo in global memory
    {
        // exclusive scan of the active instances' 8-channel tiles (as the 16-channel kernel)
        const int per = (n_inst + blockDim.x - 1) / blockDim.x;
        const int i0 = threadIdx.x * per, i1 = min(n_inst, i0 + per);
        int local = 0;
        for (int i = i0; i < i1; ++i) local += (active && !active[i]) ? 0 : (L.nc[i] + 7) / 8;
        int incl = local;
        const int ln = threadIdx.x & 63;
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o);
            if (ln >= o) incl += t;
        }
        if (ln == 63) wave_sum[threadIdx.x >> 6] = incl;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) before += wave_sum[w];
        int run = before + incl - local;
        for (int i = i0; i < i1; ++i) {
            const int nci = L.nc[i];
            tile_start[i] = run;
            inst_nc[i] = nci;
            inst_off[i] = (int)L.off[i];
            run += (active && !active[i]) ? 0 : (nci + 7) / 8;
        }
        if (threadIdx.x == blockDim.x - 1) tile_start[n_inst] = before + incl;
    }
    for (int m = threadIdx.x; m < NQ; m += blockDim.x) {
        lam[m] = m < nz ? lam_g[m] : 0.0;
        qn[m] = m < nz ? H.qlast[m] : 0.0f;
    }
    if constexpr (LPLANES > 0) {
        const int n16 = nz * S / 8;
        const uint4 *s1h = reinterpret_cast<const uint4 *>(H.planes);
        const uint4 *s1l = reinterpret_cast<const uint4 *>(g1l);
        const uint4 *s2h = reinterpret_cast<const uint4 *>(H.planes + (size_t)2 * NQ * S);
        for (int i = threadIdx.x; i < n16; i += blockDim.x) {
            reinterpret_cast<uint4 *>(lp)[i] = s1h[i];
            if (LPLANES == 2) reinterpret_cast<uint4 *>(lp + nz * S)[i] = s2h[i];
            if (P1LO) reinterpret_cast<uint4 *>(p1l)[i] = s1l[i];
        }
        if (P1HL)
            for (int i = threadIdx.x; i < p2r * S / 8; i += blockDim.x)
                reinterpret_cast<uint4 *>(p2part)[i] = s2h[i];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int l16 = lane & 15, g = lane >> 4;
    const int jl = l16 >> 1, part = l16 & 1;  // column l16: channel jl's real or imaginary part
    const int n_tiles = tile_start[n_inst];
    const float *xf = reinterpret_cast<const float *>(x);
    const float *bf = reinterpret_cast<const float *>(base);
    float *yf = reinterpret_cast<float *>(y);
#pragma unroll 1
    for (int t = first_tile(n_tiles, WAVES, wave); t < n_tiles; t += gridDim.x * WAVES) {
        int lo = 0, hi = n_inst - 1;  // last instance with tile_start <= t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tile_start[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const int nc = inst_nc[lo];
        const int jc = (t - tile_start[lo]) * 8 + jl;
        const bool live = jc < nc;
        const int cidx = inst_off[lo] + (live ? jc : 0);
        const int vb = inst_off[lo] * nz + (live ? jc : 0);
        const int vlast = vb + (nz - 1) * nc;
        // float index of this lane's part of element (channel, z); nodes past nz: the last node
        auto vf = [&](int z) -> int { return 2 * min(vb + z * nc, vlast) + part; };
        const double ej = e[cidx];
        const cx<float> cj = c[cidx];
        const float facN = fac[vlast];

        // ---- the tile's X (this lane's part), z = 32 s + 8 g + u; column maximum, split ----
        float xv[KT][8];
#pragma unroll
        for (int s = 0; s < KT; ++s)
#pragma unroll
            for (int u = 0; u < 8; ++u) xv[s][u] = xf[vf(32 * s + 8 * g + u)];
        float mx = 0.f;
#pragma unroll
        for (int s = 0; s < KT; ++s)
#pragma unroll
            for (int u = 0; u < 8; u += 2) mx = max3abs(xv[s][u], xv[s][u + 1], mx);
        const float sx = h3_scale(wave_max4(mx));
        h8 xh[KT], xl[KT];
#pragma unroll
        for (int s = 0; s < KT; ++s) split8s(xv[s], sx, xh[s], xl[s]);

        // ---- W^T = Q^T X^T ----
        v4f acc1[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc1[mt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < KT; ++s)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int row = min(16 * mt + l16, nz - 1);
                const h8 ah = *reinterpret_cast<const h8 *>(p1h + row * S + 32 * s + 8 * g);
                const h8 al = P1LO ? *reinterpret_cast<const h8 *>(p1l + row * S + 32 * s + 8 * g)
                                   : *reinterpret_cast<const h8 *>(g1l + row * S + 32 * s + 8 * g);
                acc1[mt] = mfma_h(ah, xh[s], acc1[mt]);
                acc1[mt] = mfma_h(ah, xl[s], acc1[mt]);
                acc1[mt] = mfma_h(al, xh[s], acc1[mt]);
            }

        // P2 lo fragments of the first z group, read during the scaling and the split of W
        h8 a2l[KT][H8_ZG];
#pragma unroll
        for (int zg = 0; zg < H8_ZG; ++zg) {
            const int zr = min(16 * zg + l16, nz - 1);
#pragma unroll
            for (int p = 0; p < KT; ++p)
                a2l[p][zg] = *reinterpret_cast<const h8 *>(g2l + zr * S + 32 * p + 8 * g);
        }

        // ---- W' = W / (e_j + lam_m) with the column scale removed; y0(N) = Q[N-1][:] W' ----
        const float ix = 1.0f / (sx * H3_QSCALE);
        float yn = 0.f, wm = 0.f;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = 16 * mt + 4 * g + v;
                const float r = __builtin_amdgcn_rcpf((float)(ej + lam[m]));
                const float sc = m < nz ? r : 0.f;
                acc1[mt][v] *= sc * ix;
                yn += qn[m] * acc1[mt][v];
            }
#pragma unroll
            for (int v = 0; v < 4; v += 2) wm = max3abs(acc1[mt][v], acc1[mt][v + 1], wm);
        }
        yn = wave_sum4(yn);
        // the channel's other part sits in the neighbouring lane (l16 ^ 1)
        const float yo = __int_as_float(
            __builtin_amdgcn_mov_dpp(__float_as_int(yn), 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
        const cx<float> y0 = part ? mk(yo, yn) : mk(yn, yo);
        // D^-1 x = D0^-1 x + corr f_j, corr = c y0(N) / (1 - c f_j(N)); this lane's part of it
        cx<float> corr = mk(0.f, 0.f);
        if (live) corr = cdiv(y0 * cj, mk(1.0f, 0.0f) - cj * facN);
        const float cp = part ? corr.im : corr.re;
        const float sw = h3_scale(wave_max4(wm));
        h8 wh[KT], wl[KT];
#pragma unroll
        for (int p = 0; p < KT; ++p) {
            float a[8];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                a[v] = acc1[2 * p][v];
                a[4 + v] = acc1[2 * p + 1][v];
            }
            split8s(a, sw, wh[p], wl[p]);
        }
        const float iw = 1.0f / (sw * H3_QSCALE);

        // ---- Y^T = Q W' by groups of H8_ZG z tiles; factors and base parts one group ahead
        // at clamped addresses (unused values are never stored) ----
        float fz[H8_ZG][4], bz[H8_ZG][4];
        auto load_group = [&](int zt) {
#pragma unroll
            for (int zg = 0; zg < H8_ZG; ++zg)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int z = 16 * (zt + zg) + 4 * g + v;
                    fz[zg][v] = fac[min(vb + z * nc, vlast)];
                    if constexpr (BASE) bz[zg][v] = bf[vf(z)];
                }
        };
        load_group(0);
#pragma unroll 1
        for (int zt0 = 0; zt0 < MT; zt0 += H8_ZG) {
            int zrow[H8_ZG];
#pragma unroll
            for (int zg = 0; zg < H8_ZG; ++zg) zrow[zg] = min(16 * (zt0 + zg) + l16, nz - 1);
            v4f acc2[H8_ZG];
#pragma unroll
            for (int zg = 0; zg < H8_ZG; ++zg) acc2[zg] = v4f{0.f, 0.f, 0.f, 0.f};
            auto product2 = [&](const _Float16 *src) {
#pragma unroll
                for (int p = 0; p < KT; ++p)
#pragma unroll
                    for (int zg = 0; zg < H8_ZG; ++zg) {
                        const h8 ah = *reinterpret_cast<const h8 *>(src + zrow[zg] * S + 32 * p + 8 * g);
                        acc2[zg] = mfma_h(ah, wh[p], acc2[zg]);
                        acc2[zg] = mfma_h(ah, wl[p], acc2[zg]);
                        acc2[zg] = mfma_h(a2l[p][zg], wh[p], acc2[zg]);
                    }
            };
            // P2 hi rows of this group from LDS when they are all there (the group index is
            // the same in every lane)
            if (P1HL && 16 * (zt0 + H8_ZG) <= p2r) product2(p2part);
            else product2(p2h);
            // the next group's P2 lo fragments (past the last group: clamped rows, unused)
#pragma unroll
            for (int zg = 0; zg < H8_ZG; ++zg) {
                const int zr = min(16 * (zt0 + H8_ZG + zg) + l16, nz - 1);
#pragma unroll
                for (int p = 0; p < KT; ++p)
                    a2l[p][zg] = *reinterpret_cast<const h8 *>(g2l + zr * S + 32 * p + 8 * g);
            }
            float out[H8_ZG][4];
#pragma unroll
            for (int zg = 0; zg < H8_ZG; ++zg)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float val = fmaf(cp, fz[zg][v], acc2[zg][v] * iw);
                    out[zg][v] = BASE ? fmaf(sign, val, bz[zg][v]) : val;
                }
            if (16 * (zt0 + H8_ZG) <= nz) {
                if (live) {
#pragma unroll
                    for (int zg = 0; zg < H8_ZG; ++zg)
#pragma unroll
                        for (int v = 0; v < 4; ++v) yf[vf(16 * (zt0 + zg) + 4 * g + v)] = out[zg][v];
                }
            } else {
#pragma unroll
                for (int zg = 0; zg < H8_ZG; ++zg)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int z = 16 * (zt0 + zg) + 4 * g + v;
                        if (live && zt0 + zg < MT && z < nz) yf[vf(z)] = out[zg][v];
                    }
            }
            load_group(zt0 + H8_ZG);  // past the last group: clamped, unused
            // keeps the loads here (without a possible write after them, the compiler merges
            // them with the loads before the loop into one load at the point of use)
            asm volatile("" ::: "memory");
        }
    }
}

// Planes from the FP64 eigenvectors Q[z][m] (nz x nz row-major): 2^15 Q as P1 hi, P1 lo,
// P2 hi, P2 lo, each nq rows of stride nq + 8 halves, zero outside the nz x nz block.
__global__ void h3_planes_kernel(int nz, int nq, const double *__restrict__ Q, _Float16 *planes) {
    const int S = nq + 8;
    const size_t plane = (size_t)nq * S;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nq * S; idx += gridDim.x * blockDim.x) {
        const int r = idx / S, col = idx - r * S;
        // P1[m = r][z = col] = Q[z][m]
        const double v1 = (r < nz && col < nz) ? (double)H3_QSCALE * Q[(size_t)col * nz + r] : 0.0;
        // P2[z = r][col] = Q[z][m(col)], m = 32 p + 16 h + 4 g + v for col = 32 p + 8 g + 4 h + v
        const int p = col >> 5, rr = col & 31, gg = rr >> 3, h = (rr >> 2) & 1, v = rr & 3;
        const int m = 32 * p + 16 * h + 4 * gg + v;
        const double v2 = (r < nz && col < nq && m < nz) ? (double)H3_QSCALE * Q[(size_t)r * nz + m] : 0.0;
        const _Float16 h1 = (_Float16)v1, h2 = (_Float16)v2;
        planes[idx] = h1;
        planes[plane + idx] = (_Float16)(v1 - (double)h1);
        planes[2 * plane + idx] = h2;
        planes[3 * plane + idx] = (_Float16)(v2 - (double)h2);
    }
}

int h3_padded_size(int nz) { return nz <= 224 ? (nz + 31) / 32 * 32 : 0; }

size_t h3_plane_halves(int nq) { return 4 * (size_t)nq * (nq + 8); }

void launch_h3_planes(int nz, int nq, const double *Q, _Float16 *planes, hipStream_t s) {
    hipLaunchKernelGGL(h3_planes_kernel, dim3(64), dim3(256), 0, s, nz, nq, Q, planes);
    MSX_LAUNCH_CHECK();
}

namespace {
size_t h3_lds_bytes(int nq, int nz, int n_inst, bool p1lo) {
    const size_t S = (size_t)nq + 8;
    return 12 * (size_t)nq + (p1lo ? 3 : 2) * (size_t)nz * S * 2 +
           (3 * (size_t)n_inst + 1) * sizeof(int);
}
// LDS of the planes-in-L2 form (nq > 160): lam, the Q row and the tile scan, and with p1h
// P1 hi (nz rows) and p2_rows rows of P2 hi
size_t h8l_lds_bytes(int nq, int n_inst, int nz = 0, bool p1h = false, int p2_rows = 0) {
    return 12 * (size_t)nq + (p1h ? (size_t)(nz + p2_rows) * (nq + 8) * 2 : 0) +
           (3 * (size_t)n_inst + 1) * sizeof(int);
}
size_t lds_limit() {
    static size_t lim = 0;
    if (lim == 0) {
        int dev = 0, v = 0;
        MSX_HIP(hipGetDevice(&dev));
        MSX_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        lim = (size_t)v;
    }
    return lim;
}
int cu_count() {
    static int n = 0;
    if (n == 0) {
        int dev = 0;
        MSX_HIP(hipGetDevice(&dev));
        MSX_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    }
    return n;
}

template <int KT, bool P1LO, bool BASE>
void launch_h3(const Layout &L, const H3Planes &H, const double *lam, const double *e,
               const cx<float> *c, const float *fac, const cx<float> *x, const cx<float> *base,
               cx<float> *y, float sign, const int *active, size_t shm, hipStream_t s, int cpw) {
    if (cpw == 8) {
        set_dynamic_lds((const void *)diag_h8_kernel<KT, P1LO, BASE>, shm);
        const long long tiles_max = (long long)L.n_inst * ((L.max_nc + 7) / 8);
        const int grid = (int)std::min<long long>(cu_count(), (tiles_max + H8_WAVES - 1) / H8_WAVES);
        hipLaunchKernelGGL((diag_h8_kernel<KT, P1LO, BASE>), dim3(std::max(grid, 1)),
                           dim3(64 * H8_WAVES), shm, s, L, H, lam, e, c, fac, x, base, y, sign,
                           active);
        MSX_LAUNCH_CHECK();
        return;
    }
    set_dynamic_lds((const void *)diag_h3_kernel<KT, P1LO, BASE>, shm);
    const long long tiles_max = (long long)L.n_inst * ((L.max_nc + 15) / 16);
    const int grid = (int)std::min<long long>(cu_count(), (tiles_max + H3_WAVES - 1) / H3_WAVES);
    hipLaunchKernelGGL((diag_h3_kernel<KT, P1LO, BASE>), dim3(std::max(grid, 1)),
                       dim3(64 * H3_WAVES), shm, s, L, H, lam, e, c, fac, x, base, y, sign, active);
    MSX_LAUNCH_CHECK();
}

template <int KT, bool BASE, bool P1HL>
void launch_h8l(const Layout &L, const H3Planes &H, const double *lam, const double *e,
                const cx<float> *c, const float *fac, const cx<float> *x, const cx<float> *base,
                cx<float> *y, float sign, const int *active, size_t shm, hipStream_t s) {
    const void *kern = (const void *)diag_h8_kernel<KT, false, BASE, true, H8L_WAVES, P1HL>;
    set_dynamic_lds(kern, shm);
    const long long tiles_max = (long long)L.n_inst * ((L.max_nc + 7) / 8);
    const int grid = (int)std::min<long long>(cu_count(), (tiles_max + H8L_WAVES - 1) / H8L_WAVES);
    hipLaunchKernelGGL((diag_h8_kernel<KT, false, BASE, true, H8L_WAVES, P1HL>),
                       dim3(std::max(grid, 1)), dim3(64 * H8L_WAVES), shm, s, L, H, lam, e, c, fac,
                       x, base, y, sign, active);
    MSX_LAUNCH_CHECK();
}
}  // namespace

bool diag_h3_fits(const Layout &L, const H3Planes &H) {
    if (!H.planes || H.nq == 0 || !L.nm || L.nz > H.nq || H.nq > 224) return false;
    // 32-bit element indices, including vb + z nc for the padded nodes z < nq before the clamp
    const long long total = L.total > 0 ? L.total : (long long)L.n_inst * L.max_nc;
    if (total * L.nz + (long long)H.nq * L.max_nc >= (1LL << 31)) return false;
    if (H.nq > 160) return h8l_lds_bytes(H.nq, L.n_inst) <= lds_limit();
    return h3_lds_bytes(H.nq, L.nz, L.n_inst, false) <= lds_limit();
}

void launch_diag_h3(const Layout &L, const H3Planes &H, const double *lam, const double *e,
                    const cx<float> *c, const float *fac, const cx<float> *x,
                    const cx<float> *base, cx<float> *y, float sign, const int *active,
                    hipStream_t s, int cpw) {
    if (!diag_h3_fits(L, H))
        throw Error(3, "split-FP16 diagonal blocks: shape or LDS size outside the kernel");
    if (cpw != 0 && cpw != 8 && cpw != 16)
        throw Error(3, "split-FP16 diagonal blocks: 8 or 16 channels per wavefront");
    if (H.nq > 160) {
        // planes in L2 (P1 hi in LDS where it fits), 8 channels per wavefront whatever cpw asks
        const bool p1h = h8l_lds_bytes(H.nq, L.n_inst, L.nz, true) <= lds_limit();
        // and as many 16-row tiles of P2 hi as the rest of LDS holds
        int p2_rows = 0;
        while (p1h && p2_rows + 16 <= L.nz &&
               h8l_lds_bytes(H.nq, L.n_inst, L.nz, true, p2_rows + 16) + 256 <= lds_limit())
            p2_rows += 16;
        const size_t shm = h8l_lds_bytes(H.nq, L.n_inst, L.nz, p1h, p2_rows);
        H3Planes Hp = H;
        Hp.p2_lds_rows = p2_rows;
#define MSX_H8L(KT_)                                                                           \
    if (H.nq == 32 * KT_) {                                                                    \
        if (p1h) {                                                                             \
            if (base) return launch_h8l<KT_, true, true>(L, Hp, lam, e, c, fac, x, base, y, sign, active, shm, s); \
            return launch_h8l<KT_, false, true>(L, Hp, lam, e, c, fac, x, base, y, sign, active, shm, s);        \
        }                                                                                      \
        if (base) return launch_h8l<KT_, true, false>(L, Hp, lam, e, c, fac, x, base, y, sign, active, shm, s); \
        return launch_h8l<KT_, false, false>(L, Hp, lam, e, c, fac, x, base, y, sign, active, shm, s);        \
    }
        MSX_H8L(6) MSX_H8L(7)
#undef MSX_H8L
        throw Error(3, "split-FP16 diagonal blocks: padded size not instantiated");
    }
    const bool p1lo = h3_lds_bytes(H.nq, L.nz, L.n_inst, true) <= lds_limit();
    const size_t shm = h3_lds_bytes(H.nq, L.nz, L.n_inst, p1lo);
#define MSX_H3(KT_)                                                                            \
    if (H.nq == 32 * KT_) {                                                                    \
        if (p1lo) {                                                                            \
            if (base) return launch_h3<KT_, true, true>(L, H, lam, e, c, fac, x, base, y, sign, active, shm, s, cpw); \
            return launch_h3<KT_, true, false>(L, H, lam, e, c, fac, x, base, y, sign, active, shm, s, cpw);         \
        }                                                                                      \
        if (base) return launch_h3<KT_, false, true>(L, H, lam, e, c, fac, x, base, y, sign, active, shm, s, cpw);    \
        return launch_h3<KT_, false, false>(L, H, lam, e, c, fac, x, base, y, sign, active, shm, s, cpw);            \
    }
    MSX_H3(1) MSX_H3(2) MSX_H3(3) MSX_H3(4) MSX_H3(5)
#undef MSX_H3
    throw Error(3, "split-FP16 diagonal blocks: padded size not instantiated");
}

}  // namespace msx
