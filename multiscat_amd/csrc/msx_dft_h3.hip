// Full channel coupling of the mixed solver's FP32 inner operator as the DFT convolution
//     Y = conj(F) ( Vhat(z) .* (F X F) ) conj(F),   Vhat = F V'(z) F / (n n), V' = V, V_00 = 0
// (square tables; the unsplit coupling y_j(z) = sum_{i != j} V[(G_j - G_i) mod grid](z) x_i(z),
// lib/multiscat_fortran/fortran/multiscat_gmres.f90:632-635 wrap the G difference modulo the
// table; multiscat/multiscat/_scipy_von_neumann.py:28-37 applies it unsplit) with the four
// 32 x 32 complex products on split-FP16 operands.
//
// Structure as in the FP32 kernel (msx_dft.hip): one workgroup per (instance, group of nodes)
// of node-major vectors; each wavefront owns an LDS slab (two FP32 planes, 32 rows of stride
// 36 floats) and runs its nodes one at a time: scatter X(z)^T onto the grid, four right
// products C = A F~ (A the slab by rows, F~ = F or conj(F) in registers, the result stored
// transposed, slab[j][i] = C[i][j]: with F symmetric, F X = (X^T F)^T), the pointwise product
// with Vhat(z) after the second, gather Y(z)^T.
//
// Each product is 2 x 2 output tiles of one 16x16x32 FP16 MFMA step (K = 32 = the padded
// grid): a complex product takes four real products (Cr = Ar Fr -+ Ai Fi, Ci = Ar (+-Fi) +
// Ai Fr; negated copies of Fi's planes avoid sign flips), each of three FP16 MFMAs on split
// operands, a = hi + lo: Ahi Fhi + Ahi Flo + Alo Fhi. F is held as 2^15 F (|F| = 1, so every
// lo is a normal FP16 number down to entries of 2^-18); A is scaled by one power of two per
// node and product so that its largest entry lies in [2^13, 2^14) (the sum planes of a
// complex entry stay below FP16's 65504). The powers of two are accumulated as an exponent
// and removed at the gather. Per node the error is normwise, like the FP32 kernel's, whose
// transforms also mix every channel of the node: entries below 2^-17 of the node's largest
// keep absolute precision 2^-38 of it, every other entry 22 significant bits.
// 192 FP16 MFMAs per node against the FP32 kernel's 384 FP32 ones (16 times as many
// flops per cycle per MFMA).
#include "msx_dft.h"
#include "msx_mfma.h"
#include "msx_split.h"

namespace msx {

namespace {

constexpr int HD_NP = 32;        // padded grid (tables up to 32 x 32)
constexpr int HD_S = 36;         // slab row stride in floats (16-byte rows, conflict-free)
constexpr int HD_PLANE = HD_NP * HD_S;
constexpr int HD_WPB = 4;        // wavefronts per workgroup
constexpr int HD_FEXP = 15;      // F is held as 2^15 F

__device__ __forceinline__ v4f mfma16(h8 a, h8 b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Maximum of a non-negative value over the wavefront (uniform result): non-negative floats
// order as their bit patterns, so the 16-lane rows reduce by integer maxima through DPP
// (quad swaps, half-row and row mirrors; no LDS round trips), then the four rows by readlane.
__device__ __forceinline__ float wave_max64(float v) {
    int x = __builtin_bit_cast(int, v);
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false));  // row_half_mirror
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, false));  // row_mirror
    const int m = max(max(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16)),
                      max(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48)));
    return __builtin_bit_cast(float, m);
}

// The split planes of one operand row tile: lane (i, kg) holds A[16 ti + i][8 kg .. 8 kg + 8)
struct ATile {
    h8 rh, rl, ih, il;
};

// F~ = F (CONJ false) or conj(F): C += A F~ for the 2 x 2 output tiles.
//   plain: Cr = Ar Fr - Ai Fi, Ci = Ar Fi + Ai Fr
//   conj:  Cr = Ar Fr + Ai Fi, Ci = Ai Fr - Ar Fi
template <bool CONJ>
__device__ __forceinline__ void tile_products(const ATile (&a)[2], const h8 (&frh)[2],
                                              const h8 (&frl)[2], const h8 (&fih)[2],
                                              const h8 (&fil)[2], const h8 (&nfih)[2],
                                              const h8 (&nfil)[2], v4f (&cr)[2][2],
                                              v4f (&ci)[2][2]) {
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
            v4f r = {0.f, 0.f, 0.f, 0.f}, m = {0.f, 0.f, 0.f, 0.f};
            // Ar Fr
            r = mfma16(a[ti].rh, frh[tj], r);
            r = mfma16(a[ti].rh, frl[tj], r);
            r = mfma16(a[ti].rl, frh[tj], r);
            if constexpr (!CONJ) {
                // - Ai Fi, Ar Fi, Ai Fr
                r = mfma16(a[ti].ih, nfih[tj], r);
                r = mfma16(a[ti].ih, nfil[tj], r);
                r = mfma16(a[ti].il, nfih[tj], r);
                m = mfma16(a[ti].rh, fih[tj], m);
                m = mfma16(a[ti].rh, fil[tj], m);
                m = mfma16(a[ti].rl, fih[tj], m);
            } else {
                // + Ai Fi, - Ar Fi, Ai Fr
                r = mfma16(a[ti].ih, fih[tj], r);
                r = mfma16(a[ti].ih, fil[tj], r);
                r = mfma16(a[ti].il, fih[tj], r);
                m = mfma16(a[ti].rh, nfih[tj], m);
                m = mfma16(a[ti].rh, nfil[tj], m);
                m = mfma16(a[ti].rl, nfih[tj], m);
            }
            m = mfma16(a[ti].ih, frh[tj], m);
            m = mfma16(a[ti].ih, frl[tj], m);
            m = mfma16(a[ti].il, frh[tj], m);
            cr[ti][tj] = r;
            ci[ti][tj] = m;
        }
}

// Reads the slab's A operand (rows 16 ti + l16, columns 8 kg .. + 8 of both planes), scales it
// by 2^e (e chosen from the largest |entry| of the whole 32 x 32 grid) and splits it; returns e.
__device__ __forceinline__ int load_split(const float *sr, const float *si, ATile (&a)[2]) {
    const int lane = threadIdx.x & 63, l16 = lane & 15, kg = lane >> 4;
    float xr[2][8], xi[2][8];
    float mx = 0.f;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
        const int o = (16 * ti + l16) * HD_S + 8 * kg;
        const float4 r0 = *reinterpret_cast<const float4 *>(sr + o);
        const float4 r1 = *reinterpret_cast<const float4 *>(sr + o + 4);
        const float4 i0 = *reinterpret_cast<const float4 *>(si + o);
        const float4 i1 = *reinterpret_cast<const float4 *>(si + o + 4);
        xr[ti][0] = r0.x; xr[ti][1] = r0.y; xr[ti][2] = r0.z; xr[ti][3] = r0.w;
        xr[ti][4] = r1.x; xr[ti][5] = r1.y; xr[ti][6] = r1.z; xr[ti][7] = r1.w;
        xi[ti][0] = i0.x; xi[ti][1] = i0.y; xi[ti][2] = i0.z; xi[ti][3] = i0.w;
        xi[ti][4] = i1.x; xi[ti][5] = i1.y; xi[ti][6] = i1.z; xi[ti][7] = i1.w;
#pragma unroll
        for (int u = 0; u < 8; ++u) mx = max3abs(xr[ti][u], xi[ti][u], mx);
    }
    mx = wave_max64(mx);
    int e2 = 0;
    (void)frexpf(mx, &e2);  // mx = f 2^e2, f in [0.5, 1); the scaled maximum lies in [2^13, 2^14)
    const int e = 14 - max(e2, -100);
    const float sc = ldexpf(1.0f, e);
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
        split8s(xr[ti], sc, a[ti].rh, a[ti].rl);
        split8s(xi[ti], sc, a[ti].ih, a[ti].il);
    }
    return e;
}

// slab[j][i] = C[i][j]: lane (j, g) of tile (ti, tj) holds rows 16 ti + 4 g .. + 3 of
// column 16 tj + j
__device__ __forceinline__ void store_t(float *dr, float *di, const v4f (&cr)[2][2],
                                        const v4f (&ci)[2][2]) {
    const int lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
            const int a = (16 * tj + col) * HD_S + 16 * ti + 4 * g;
            *reinterpret_cast<v4f *>(dr + a) = cr[ti][tj];
            *reinterpret_cast<v4f *>(di + a) = ci[ti][tj];
        }
}

__global__ __launch_bounds__(64 * HD_WPB, 2) void dft_h3_kernel(
    Layout L, const cx<float> *__restrict__ Fx, const cx<float> *__restrict__ Vhat,
    const cx<float> *__restrict__ x, cx<float> *__restrict__ y, const int *__restrict__ active,
    int zpb, const double *__restrict__ x_norm, cx<float> *__restrict__ x_copy) {
    constexpr int KC = HD_NP * HD_NP / 64;  // channels per lane (nc <= 1024)
    const int b = blockIdx.x, zg = blockIdx.y;
    if (active && !active[b]) return;
    const int nz = L.nz;
    // the wavefront index is uniform (readfirstlane makes that visible to the compiler, so node
    // indices and the pointers built from them stay in scalar registers)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int l16 = lane & 15, kg = lane >> 4;
    extern __shared__ double hd_smem[];
    float *sr = reinterpret_cast<float *>(hd_smem) + (size_t)wave * 2 * HD_PLANE;
    float *si = sr + HD_PLANE;

    // B operands: lane (j, kg) of column tile tj holds 2^15 F[8 kg + t][16 tj + j], t = 0..7,
    // as hi and lo, and the negated imaginary planes
    h8 frh[2], frl[2], fih[2], fil[2], nfih[2], nfil[2];
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
        float vr[8], vi[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const cx<float> f = Fx[(8 * kg + t) * HD_NP + 16 * tj + l16];
            vr[t] = f.re;
            vi[t] = f.im;
        }
        const float fs = (float)(1 << HD_FEXP);
        split8s(vr, fs, frh[tj], frl[tj]);
        split8s(vi, fs, fih[tj], fil[tj]);
        split8s(vi, -fs, nfih[tj], nfil[tj]);
    }

    const int nc = L.nc[b];
    const long long off = L.off[b];
    __shared__ int s_pos[HD_NP * HD_NP];  // slab position of each channel (X^T: row chy)
    for (int j = threadIdx.x; j < nc; j += blockDim.x) s_pos[j] = L.chy[off + j] * HD_S + L.chx[off + j];
    __syncthreads();
    const int g = kg, col = l16;
    // Addressing: a uniform pointer to the node's channels (node-major: contiguous) and the
    // lane's byte offset 8 lane + 512 (k mod 8) (32 bits; the pointer advances by 512 channels
    // every eight blocks, so the rest fits the instructions' offset field) for channel
    // j = lane + 64 k; blocks of 64
    // channels inside nc take no lane mask, blocks past nc are skipped (uniform branches)
    const size_t base0 = (size_t)off * nz;
    const uint32_t lb = (uint32_t)lane * (uint32_t)sizeof(cx<float>);
    auto full = [&](int k) { return 64 * k + 64 <= nc; };
    auto part = [&](int k) { return 64 * k < nc && lane + 64 * k < nc; };
    const float xs = x_copy ? (float)(1.0 / x_norm[b]) : 1.0f;
    cx<float> xv[KC];
    auto load_x = [&](int z) {
        const cx<float> *xz = x + base0 + (size_t)z * nc;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            if (full(k)) xv[k] = at(xz + 512 * (k / 8), lb + 512u * (k % 8));
            else if (part(k)) xv[k] = at(xz + 512 * (k / 8), lb + 512u * (k % 8));
        }
    };
    const int z_first = zg * zpb + wave;
    const int z_end = min(nz, (zg + 1) * zpb);
    if (z_first < z_end) load_x(z_first);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the first node's data before the loop
    for (int z = z_first; z < z_end; z += HD_WPB) {
        {
            float4 *p4 = reinterpret_cast<float4 *>(sr);  // both planes, 16-byte stores
            constexpr int N4 = 2 * HD_PLANE / 4;
            for (int idx = lane; idx < N4; idx += 64) p4[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __builtin_amdgcn_wave_barrier();
        cx<float> *cz = x_copy ? x_copy + base0 + (size_t)z * nc : nullptr;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            if (full(k) || part(k)) {
                const int p = s_pos[lane + 64 * k];
                cx<float> v = xv[k];
                if (x_copy) {
                    v = mk(v.re * xs, v.im * xs);
                    at(cz + 512 * (k / 8), lb + 512u * (k % 8)) = v;
                }
                sr[p] = v.re;
                si[p] = v.im;
            }
        }
        const bool more = z + HD_WPB < z_end;
        if (more) load_x(z + HD_WPB);  // consumed by the scatter above
        __builtin_amdgcn_wave_barrier();
        ATile a[2];
        v4f cr[2][2], ci[2][2];
        int ex = 4 * HD_FEXP;  // exponent of the scale the slab carries after the four products
        // (1) X^T F = (F X)^T, stored transposed: the slab holds F X
        ex += load_split(sr, si, a);
        tile_products<false>(a, frh, frl, fih, fil, nfih, nfil, cr, ci);
        __builtin_amdgcn_wave_barrier();
        store_t(sr, si, cr, ci);
        __builtin_amdgcn_wave_barrier();
        // (2) (F X) F, then .* Vhat(z) in the accumulator layout (rows kx = 16 ti + 4 g + v);
        // the transposed store leaves Z^T
        cx<float> vh[2][2][4];
        {
            const cx<float> *vz = Vhat + (size_t)z * HD_NP * HD_NP;
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        vh[ti][tj][v] = vz[(ti * 16 + 4 * g + v) * HD_NP + tj * 16 + col];
        }
        ex += load_split(sr, si, a);
        tile_products<false>(a, frh, frl, fih, fil, nfih, nfil, cr, ci);
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const cx<float> h = vh[ti][tj][v];
                    const float re = cr[ti][tj][v], im = ci[ti][tj][v];
                    cr[ti][tj][v] = re * h.re - im * h.im;
                    ci[ti][tj][v] = re * h.im + im * h.re;
                }
        __builtin_amdgcn_wave_barrier();
        store_t(sr, si, cr, ci);
        __builtin_amdgcn_wave_barrier();
        // (3) Z^T conj(F) = (conj(F) Z)^T, stored transposed
        ex += load_split(sr, si, a);
        tile_products<true>(a, frh, frl, fih, fil, nfih, nfil, cr, ci);
        __builtin_amdgcn_wave_barrier();
        store_t(sr, si, cr, ci);
        __builtin_amdgcn_wave_barrier();
        // (4) W conj(F) = out, stored transposed: the slab holds 2^ex out^T
        ex += load_split(sr, si, a);
        tile_products<true>(a, frh, frl, fih, fil, nfih, nfil, cr, ci);
        __builtin_amdgcn_wave_barrier();
        store_t(sr, si, cr, ci);
        __builtin_amdgcn_wave_barrier();
        cx<float> *yz = y + base0 + (size_t)z * nc;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            if (full(k) || part(k)) {
                const int p = s_pos[lane + 64 * k];
                at(yz + 512 * (k / 8), lb + 512u * (k % 8)) = mk(ldexpf(sr[p], -ex), ldexpf(si[p], -ex));
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace

bool dft_h3_fits(const Layout &L, const DftTables<float> &D) {
    if (D.np != HD_NP || !L.nm || L.nkx != L.nky || L.max_nc > HD_NP * HD_NP) return false;
    // byte offsets of the batch's FP32 vectors in 32 bits
    const long long tot = L.total > 0 ? L.total : (long long)L.max_nc * L.n_inst;
    return (long long)L.nz * (tot + 64) * (long long)sizeof(cx<float>) < (1LL << 32);
}

void launch_dft_coupling_h3(const Layout &L, const DftTables<float> &D, const cx<float> *x,
                            cx<float> *y, const int *active, hipStream_t s, int active_hint,
                            const double *x_norm, cx<float> *x_copy, int nodes_per_wave) {
    if (!dft_h3_fits(L, D))
        throw Error(3, "split-FP16 DFT coupling: node-major vectors and a square table of at "
                       "most 32 x 32");
    if ((x_norm == nullptr) != (x_copy == nullptr))
        throw Error(3, "DFT coupling: x_norm and x_copy go together");
    if (nodes_per_wave < 0 || nodes_per_wave > 16) throw Error(3, "DFT coupling: 0..16 nodes per wave");
    const size_t shm = sizeof(float) * (size_t)HD_WPB * 2 * HD_PLANE;
    // nodes per workgroup as the FP32 kernel takes them (four wavefronts per workgroup)
    const int zpb = dft_block_nodes(L, false, active_hint, nodes_per_wave);
    const int nzb = (L.nz + zpb - 1) / zpb;
    hipLaunchKernelGGL(dft_h3_kernel, dim3(L.n_inst, nzb), dim3(64 * HD_WPB), shm, s, L, D.Fx,
                       D.Vhat, x, y, active, zpb, x_norm, x_copy);
    MSX_LAUNCH_CHECK();
}

}  // namespace msx
