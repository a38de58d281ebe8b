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
__global__ __launch_bounds__(64 * DG_WAVES, (NT <= 10 && !(NR == 2 && sizeof(T) == 8)) ? 2 : 1) void diag_gemm_kernel(
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
    if (L.nm) throw Error(3, "diagonal blocks (global-memory Q): channel-major vectors only");
    dim3 grid((L.max_nc + CH * DG_WAVES - 1) / (CH * DG_WAVES), L.n_inst);
    hipLaunchKernelGGL((diag_gemm_kernel<T, NR, NT>), grid, dim3(64 * DG_WAVES), 0, s, L, E, e,
                       c, fac, x, base, y, sign, inverse, active);
    MSX_LAUNCH_CHECK();
}

// FP32 form with Q^T resident in LDS (nzp <= 160: 105 KB). One persistent workgroup per CU
// stages Q^T once; each of its DL_WAVES wavefronts then loops over (instance, 16-channel)
// tiles with the same two products as diag_gemm_kernel, reading the A fragments of both
// products from LDS (the second product reads Q = (Q^T)^T by columns) so that only the
// channel data, the factors and the result move through the memory hierarchy; the channel
// data are prefetched two stages ahead. Row stride 164 (= 36 mod 64 words) keeps the 16-row
// b128 reads of the first product and the 4-row x 16-column b32 reads of the second
// conflict-free. Work balance: every workgroup scans the active instances' tile counts into
// LDS, and the wavefronts stride over the compact list of active tiles (a binary search maps
// a tile to its instance), so each gets the same number of tiles to within one.
// The Sherman-Morrison column f_j = D0^-1 e_N is read from its table (summed in FP64 when
// the table was built: forming it here from Q and 1 / (e_j + lam_m) in FP32 loses up to
// 1e-3 of f_j(N) to cancellation when e_j + lam_m nearly vanishes, which cost the mixed
// solver an extra refinement sweep).
// NM: node-major vectors and factors (Layout::nm), where the 16 channels of a tile at one
// node are one 128-byte line (64 bytes of factors). INV (D^-1 or D) and BASE (y = base + sign * D^-1 x or y = D^-1 x) are
// template parameters so that every load and store is straight-line code: under a runtime
// condition the loads of the epilogue became branches, and the compiler then waited for all
// outstanding memory operations (the previous stores included) before each store.
constexpr int DL_WAVES = 8;
constexpr int DL_STRIDE = 164;
constexpr int DL_MAX_INST = 4096;  // instances the in-LDS tile scan covers

template <int NT, bool NM, bool INV, bool BASE>
__global__ __launch_bounds__(64 * DL_WAVES) void diag_lds_kernel(
    Layout L, EigTables<float> E, const double *__restrict__ e, const cx<float> *__restrict__ c,
    const float *__restrict__ fac, const cx<float> *x, const cx<float> *base, cx<float> *y,
    float sign, const int *__restrict__ active) {
    using M = Mfma<float>;
    using acc_t = M::acc;
    constexpr int ZG = NT % 4 == 0 ? 4 : 2;
    extern __shared__ double dl_smem[];
    double *lam = dl_smem;                                    // [nzp], 0 past nz
    float *qt = reinterpret_cast<float *>(dl_smem + 16 * NT);  // [nzp][DL_STRIDE]
    __shared__ int tile_start[DL_MAX_INST + 1];               // exclusive scan of active tiles
    __shared__ int inst_nc[DL_MAX_INST], inst_off[DL_MAX_INST];  // channel counts and offsets
    __shared__ int wave_sum[DL_WAVES];
    const int nzp = E.nzp, nz = L.nz, NTL = nzp >> 4;
    const int n_inst = L.n_inst;
    {
        // per-thread serial sums over a contiguous chunk, then a scan across the threads
        const int per = (n_inst + blockDim.x - 1) / blockDim.x;
        const int i0 = threadIdx.x * per, i1 = min(n_inst, i0 + per);
        int local = 0;
        for (int i = i0; i < i1; ++i)
            local += (active && !active[i]) ? 0 : (L.nc[i] + 15) / 16;
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
    for (int m = threadIdx.x; m < 16 * NT; m += blockDim.x) lam[m] = m < nz ? E.lam[m] : 0.0;
    // Q^T rows with the columns past nzp zeroed up to an even number of 16-column blocks
    // (the first product runs its stages in pairs)
    const int ncol = 16 * ((NTL + 1) & ~1);
    for (int idx = threadIdx.x; idx < nzp * ncol / 4; idx += blockDim.x) {
        const int r = idx / (ncol / 4), c4 = idx - r * (ncol / 4);
        const float4 v = 4 * c4 < nzp ? reinterpret_cast<const float4 *>(E.QTp + (long long)r * nzp)[c4]
                                      : make_float4(0.f, 0.f, 0.f, 0.f);
        *reinterpret_cast<float4 *>(qt + r * DL_STRIDE + 4 * c4) = v;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int l16 = lane & 15, g = lane >> 4;
    auto tile = [&](int t) { return t < NTL ? t : NTL - 1; };
    const int n_tiles = tile_start[n_inst];
#pragma unroll 1
    // tile t to wavefront slot t: the first gridDim.x tiles to wavefront 0 of every workgroup,
    // the next to wavefront 1, ...: a launch with few active tiles gives each its own SIMD
    // (36 instead of 60 us for one tile per wavefront, where two wavefronts of one workgroup
    // shared a SIMD's matrix core)
    for (int t = wave * gridDim.x + blockIdx.x; t < n_tiles; t += gridDim.x * DL_WAVES) {
        int lo = 0, hi = n_inst - 1;  // last instance with tile_start <= t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tile_start[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const int b = lo;
        const int j0 = (t - tile_start[b]) * 16;
        const int nc = inst_nc[b];
        if (j0 >= nc) continue;  // never taken; also keeps the register allocation spill-free
        const int jc = j0 + l16;
        const bool live = jc < nc;
        const long long cidx = inst_off[b] + (live ? jc : 0);
        const long long row0 = cidx * nz;  // channel-major row (vectors and factors unless NM)
        const long long vb = (long long)inst_off[b] * nz + (live ? jc : 0);
        // vector element of this lane's channel at node z
        auto vi = [&](int z) -> long long {
            if constexpr (NM) return vb + (long long)z * nc;
            else return row0 + z;
        };
        // per-channel scalars of the later phases, loaded now so that the first product
        // covers their latency
        const double ej = e[cidx];
        const cx<float> cj = c[cidx];
        const float facN = INV ? fac[vi(nz - 1)] : 0.0f;
        const cx<float> xN = INV ? mk(0.0f, 0.0f) : x[vi(nz - 1)];

        // ---- W^T = Q^T X^T, stages in pairs with the X data of each pair's second stage in
        // flight during the first: the two register buffers keep fixed names, so each stage
        // waits only for its own loads ----
        acc_t acc1[NT][2];
#pragma unroll
        for (int mt = 0; mt < NT; ++mt) acc1[mt][0] = acc1[mt][1] = acc_t{0, 0, 0, 0};
        // every lane loads from a clamped, valid address and nothing is masked: Q^T is zero
        // past column nz, so the values loaded for z >= nz add nothing, and a lane past the
        // instance's last channel only feeds its own output column, which is never stored
        auto load_x = [&](int s, cx<float> (&q)[4]) {
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                const int z = 16 * s + 4 * g + cc;
                q[cc] = x[vi(z < nz ? z : 0)];
            }
        };
        auto stage = [&](int s, const cx<float> (&q)[4]) {
            const int zq = 16 * s + 4 * g;
            float4 a[NT];
#pragma unroll
            for (int mt = 0; mt < NT; ++mt)
                a[mt] = *reinterpret_cast<const float4 *>(qt + (16 * tile(mt) + l16) * DL_STRIDE + zq);
#pragma unroll
            for (int cc = 0; cc < 4; ++cc)
#pragma unroll
                for (int mt = 0; mt < NT; ++mt) {
                    const float av = cc == 0 ? a[mt].x : cc == 1 ? a[mt].y : cc == 2 ? a[mt].z : a[mt].w;
                    acc1[mt][0] = M::op(av, q[cc].re, acc1[mt][0]);
                    acc1[mt][1] = M::op(av, q[cc].im, acc1[mt][1]);
                }
        };
        // an unconditional pair body over an even stage count (Q^T is zero in the extra
        // block): each stage then waits for the four loads of its own buffer only (peeling
        // the last pair, whose loads are unused, measured no faster and spills)
        cx<float> xa[4], xb[4];
        load_x(0, xa);
        load_x(1, xb);
#pragma unroll 1
        for (int s = 0; s < NTL; s += 2) {
            stage(s, xa);
            __builtin_amdgcn_sched_barrier(0);
            load_x(s + 2, xa);  // past the last block: clamped, unused
            stage(s + 1, xb);
            __builtin_amdgcn_sched_barrier(0);
            load_x(s + 3, xb);
        }

        // ---- scaling and y0(N) = Q[N-1][:] W (Q[N-1][m] = Q^T[m][N-1]); 1 / (e_j + lam_m)
        // from the FP64 sum by the hardware reciprocal (1 ulp), 0 past nz ----
        float yn0 = 0.f, yn1 = 0.f;
#pragma unroll
        for (int mt = 0; mt < NT; ++mt) {
            double lv[4];
            float qv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                lv[v] = lam[16 * mt + 4 * g + v];
                qv[v] = qt[(16 * tile(mt) + 4 * g + v) * DL_STRIDE + nz - 1];
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = 16 * mt + 4 * g + v;
                const float df = (float)(ej + lv[v]);
                const float sc = m < nz ? (INV ? __builtin_amdgcn_rcpf(df) : df) : 0.f;
                acc1[mt][0][v] *= sc;
                acc1[mt][1][v] *= sc;
                yn0 += qv[v] * acc1[mt][0][v];
                yn1 += qv[v] * acc1[mt][1][v];
            }
        }
        yn0 += __shfl_xor(yn0, 16);
        yn0 += __shfl_xor(yn0, 32);
        yn1 += __shfl_xor(yn1, 16);
        yn1 += __shfl_xor(yn1, 32);
        // D^-1 x = D0^-1 x + corr f_j, corr = c y0(N) / (1 - c f_j(N)); D x = D0 x - c x(N) e_N
        cx<float> corr = mk(0.f, 0.f);
        if (live) corr = INV ? cdiv(mk(yn0, yn1) * cj, mk(1.0f, 0.0f) - cj * facN) : cj * xN;

        // ---- Y^T = Q W^T: A[i][k] = Q[16 zt + i][m] = Q^T[m][16 zt + i] ----
        // factors and base values one z-group ahead, at clamped addresses (unused values are
        // never stored), issued after the previous group's stores: a whole group of products
        // then covers both (issued before the stores, they could not reuse the registers of
        // the values being stored, and the copy at the loop's end waited for them at once)
        float fz[ZG][4];
        cx<float> bz[ZG][4];
        auto load_group = [&](int zt) {
#pragma unroll
            for (int zg = 0; zg < ZG; ++zg)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int z = 16 * (zt + zg) + 4 * g + v;
                    const int zc = zt + zg < NTL && z < nz ? z : 0;
                    if constexpr (INV) fz[zg][v] = fac[vi(zc)];
                    if constexpr (BASE) bz[zg][v] = base[vi(zc)];
                }
        };
        load_group(0);
#pragma unroll 1
        for (int zt0 = 0; zt0 < NTL; zt0 += ZG) {
            acc_t acc2[ZG][2];
#pragma unroll
            for (int zg = 0; zg < ZG; ++zg) acc2[zg][0] = acc2[zg][1] = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int mt = 0; mt < NT; ++mt) {
                float a[ZG][4];
#pragma unroll
                for (int zg = 0; zg < ZG; ++zg)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        a[zg][v] = qt[(16 * tile(mt) + 4 * g + v) * DL_STRIDE +
                                      16 * tile(zt0 + zg) + l16];
#pragma unroll
                for (int v = 0; v < 4; ++v)
#pragma unroll
                    for (int zg = 0; zg < ZG; ++zg) {
                        acc2[zg][0] = M::op(a[zg][v], acc1[mt][0][v], acc2[zg][0]);
                        acc2[zg][1] = M::op(a[zg][v], acc1[mt][1][v], acc2[zg][1]);
                    }
            }
            // all of the group's results first, then its stores back to back
            cx<float> out[ZG][4];
#pragma unroll
            for (int zg = 0; zg < ZG; ++zg)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int z = 16 * (zt0 + zg) + 4 * g + v;
                    cx<float> val = mk(acc2[zg][0][v], acc2[zg][1][v]);
                    if (INV) val = val + corr * fz[zg][v];
                    else if (z == nz - 1) val = val - corr;
                    out[zg][v] = BASE ? bz[zg][v] + val * sign : val;
                }
            // a group inside [0, nz) stores under the lane mask alone (one block of stores:
            // conditions per element split them into blocks, and the waits the compiler
            // places at the joins held later stores back for earlier ones)
            if (16 * (zt0 + ZG) <= nz) {
                if (live) {
#pragma unroll
                    for (int zg = 0; zg < ZG; ++zg)
#pragma unroll
                        for (int v = 0; v < 4; ++v) y[vi(16 * (zt0 + zg) + 4 * g + v)] = out[zg][v];
                }
            } else {
#pragma unroll
                for (int zg = 0; zg < ZG; ++zg)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int z = 16 * (zt0 + zg) + 4 * g + v;
                        if (live && zt0 + zg < NTL && z < nz) y[vi(z)] = out[zg][v];
                    }
            }
            load_group(zt0 + ZG);  // past the last group: clamped, unused
            // keeps the loads here: without a possible write after them, the compiler merges
            // them with the loads before the loop into one load at the point of use
            asm volatile("" ::: "memory");
        }
    }
}

template <int NT, bool NM, bool INV, bool BASE>
static void launch_diag_lds(const Layout &L, const EigTables<float> &E, const double *e,
                            const cx<float> *c, const float *fac, const cx<float> *x,
                            const cx<float> *base, cx<float> *y, float sign, const int *active,
                            hipStream_t s) {
    if (E.nzp % 16 != 0 || E.nzp > 16 * NT || E.nzp < L.nz)
        throw Error(3, "diagonal blocks: padded size does not fit the kernel");
    const size_t shm = sizeof(double) * 16 * NT + sizeof(float) * (size_t)E.nzp * DL_STRIDE;
    set_dynamic_lds((const void *)diag_lds_kernel<NT, NM, INV, BASE>, shm);
    static int n_cu = 0;
    if (n_cu == 0) {
        int dev = 0;
        MSX_HIP(hipGetDevice(&dev));
        MSX_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    if (L.n_inst > DL_MAX_INST) throw Error(3, "diagonal blocks: too many instances for one launch");
    if ((L.total > 0 ? L.total : (long long)L.n_inst * L.max_nc) >= (1LL << 31))
        throw Error(3, "diagonal blocks: channel offsets past 32 bits");
    const long long tiles_max = (long long)L.n_inst * ((L.max_nc + 15) / 16);
    const int grid = (int)std::min<long long>(n_cu, (tiles_max + DL_WAVES - 1) / DL_WAVES);
    hipLaunchKernelGGL((diag_lds_kernel<NT, NM, INV, BASE>), dim3(std::max(grid, 1)),
                       dim3(64 * DL_WAVES), shm, s, L, E, e, c, fac, x, base, y, sign, active);
    MSX_LAUNCH_CHECK();
}

// The runtime choices of the LDS form as template parameters: node-major vectors come from
// the mixed solver's inner iteration, which applies D^-1 only.
template <int NT>
static void launch_diag_lds_any(const Layout &L, const EigTables<float> &E, const double *e,
                                const cx<float> *c, const float *fac, const cx<float> *x,
                                const cx<float> *base, cx<float> *y, float sign, int inverse,
                                const int *active, hipStream_t s) {
#define MSX_LDS(NM_, INV_, BASE_) \
    return launch_diag_lds<NT, NM_, INV_, BASE_>(L, E, e, c, fac, x, base, y, sign, active, s)
    if (L.nm) {
        if (!inverse) throw Error(3, "diagonal blocks: node-major vectors with D^-1 only");
        if (base) MSX_LDS(true, true, true);
        MSX_LDS(true, true, false);
    }
    if (inverse) {
        if (base) MSX_LDS(false, true, true);
        MSX_LDS(false, true, false);
    }
    if (base) MSX_LDS(false, false, true);
    MSX_LDS(false, false, false);
#undef MSX_LDS
}

// Unpreconditioned residual of the scattering system for the mixed solver's refinement:
//   rho_j = s_j - D_j x_j - (V x)_j,   D_j x_j = (H0 + e_j) x_j - c_j x_j(N) e_N,
// s the scattering source (b at node N of each instance's specular channel; src null) or a
// given source vector src (the adjoint's), V x given (the FP64 coupling, of the flipped table
// for the transposed system: D is complex symmetric), x, V x and src in the vector order of L,
// written in FP32 in the vector order of
// Lo (the inner solver's). H0 x_j is one FP64 product on the matrix cores (D^-1 takes two):
// 16 channels (real and imaginary parts as two column tiles) share RT row tiles of Y^T = H0 X^T
// per wavefront, so NT / RT wavefronts cover a channel tile (for NT = 10, two of five tiles:
// 80 accumulator registers instead of 160, and two or three wavefronts per SIMD hide the
// latency of the H0 fragments, read from L2, and of the epilogue). The product runs as the
// first product of diag_gemm_kernel does with Q^T, without the row permutation, so register v
// of lane group g of tile mt holds row 16 mt + 4 v + g. Every term is FP64 until the sum.
template <int NT, int RT>
__global__ __launch_bounds__(64 * DG_WAVES, 2) void split_residual_kernel(
    Layout L, Layout Lo, const double *__restrict__ H0p, int nzp, const double *__restrict__ e,
    const cx<double> *__restrict__ c, const int *__restrict__ spec,
    const cx<double> *__restrict__ bsrc, const cx<double> *__restrict__ x,
    const cx<double> *__restrict__ vx, const cx<double> *__restrict__ src,
    cx<float> *__restrict__ rho, const int *__restrict__ active) {
    static_assert(NT % RT == 0, "row tiles split evenly over the wavefronts of a channel tile");
    constexpr int NHALF = NT / RT;  // wavefronts per channel tile
    using M = Mfma<double>;
    using acc_t = typename M::acc;
    constexpr int CC = 2, NH = 2;  // k-steps per 16-byte stage, stages per 16-wide K block
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int nc = L.nc[b];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wid = blockIdx.x * DG_WAVES + wave;
    const int j0 = (wid / NHALF) * 16;
    const int mt0 = (wid % NHALF) * RT;  // first row tile of this wavefront
    if (j0 >= nc) return;
    const int nz = L.nz, NTL = nzp >> 4;
    if (mt0 >= NTL) return;  // every row tile of this wavefront lies past nzp
    const int l16 = lane & 15, g = lane >> 4;
    const int jc = j0 + l16;
    const bool live = jc < nc;
    const long long off = L.off[b];
    const long long cidx = off + (live ? jc : 0);
    // element (this lane's channel, z) of x and V x in the vector order of L: node-major puts
    // the 16 channels of a wavefront at one node in one 256-byte run
    const long long vb = L.nm ? off * nz + (live ? jc : 0) : cidx * nz;
    const long long vstep = L.nm ? nc : 1;
    auto vi = [&](int z) { return vb + z * vstep; };
    auto tile = [&](int t) { return t < NTL ? t : NTL - 1; };

    acc_t acc[RT][2];
#pragma unroll
    for (int mt = 0; mt < RT; ++mt) acc[mt][0] = acc[mt][1] = acc_t{0, 0, 0, 0};
    double a1[2][RT][CC], x1[2][2][CC];
    // stage st: K block st / NH, k-steps (st % NH) * CC .. + CC. H0p is zero past column nz,
    // so X values loaded (at clamped addresses) for z >= nz add nothing; blocks past nzp / 16
    // repeat the last block's H0 columns, and their X is zeroed.
    auto load1 = [&](int st, double (&a)[RT][CC], double (&xb)[2][CC]) {
        const int s = st / NH, h = st - s * NH;
        const int zb = 16 * s + 4 * g + h * CC;
        const int zq = 16 * tile(s) + 4 * g + h * CC;
#pragma unroll
        for (int cc = 0; cc < CC; ++cc) {
            const int z = zb + cc;
            const bool ok = s < NTL;
            const cx<double> v = x[vi(z < nz ? z : 0)];
            xb[0][cc] = ok ? v.re : 0.0;
            xb[1][cc] = ok ? v.im : 0.0;
        }
#pragma unroll
        for (int mt = 0; mt < RT; ++mt)
            load16<CC>(H0p + (long long)(16 * tile(mt0 + mt) + l16) * nzp + zq, a[mt]);
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
            for (int mt = 0; mt < RT; ++mt)
#pragma unroll
                for (int r = 0; r < 2; ++r) acc[mt][r] = M::op(a1[cur][mt][cc], x1[cur][r][cc], acc[mt][r]);
        __builtin_amdgcn_sched_barrier(0);
    }

    if (!live) return;
    const double ej = e[cidx];
    const cx<double> cj = c[cidx], xN = x[vi(nz - 1)];
    const bool is_spec = jc == spec[b];
    const cx<double> bs = bsrc[b];
    // per row tile: its x and (V x) values loaded together at clamped addresses (the next
    // tile's before this tile's stores), and a tile inside [0, nz) stored with no condition.
    // With loads and stores under per-element conditions the compiler waited for every
    // outstanding memory operation before each store: 40 round trips per tile.
    cx<double> xz[4], vz[4], sz[4];
    auto load_t = [&](int mt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int z = 16 * mt + M::row(g, v);
            const long long i = vi(mt < NTL && z < nz ? z : 0);
            xz[v] = x[i];
            vz[v] = vx[i];
            sz[v] = src ? src[i] : mk(0.0, 0.0);
        }
    };
    load_t(mt0);
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int mt = mt0 + t;
        cx<float> out[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int z = 16 * mt + M::row(g, v);
            cx<double> dx = mk(acc[t][0][v] + ej * xz[v].re, acc[t][1][v] + ej * xz[v].im);
            if (z == nz - 1) dx = dx - cj * xN;
            cx<double> r = sz[v] - dx - vz[v];
            if (!src && is_spec && z == nz - 1) r = r + bs;
            out[v] = mk((float)r.re, (float)r.im);
        }
        if (t + 1 < RT) load_t(mt + 1);
        if (mt < NTL && 16 * mt + 16 <= nz) {
#pragma unroll
            for (int v = 0; v < 4; ++v) rho[vec_index(Lo, off, nc, jc, 16 * mt + M::row(g, v))] = out[v];
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int z = 16 * mt + M::row(g, v);
                if (mt < NTL && z < nz) rho[vec_index(Lo, off, nc, jc, z)] = out[v];
            }
        }
    }
}

bool split_residual_supported(int nzp) { return nzp % 16 == 0 && nzp <= 224; }

void launch_split_residual(const Layout &L, const Layout &Lo, const double *H0p, int nzp,
                           const double *e, const cx<double> *c, const int *spec,
                           const cx<double> *bsrc, const cx<double> *x, const cx<double> *vx,
                           cx<float> *rho, const int *active, hipStream_t s,
                           const cx<double> *src) {
    if (!split_residual_supported(nzp) || nzp < L.nz)
        throw Error(3, "split residual: nz <= 224 required");
    // wavefronts: (channel tiles) x (NT / RT row groups), DG_WAVES per workgroup
    auto grid = [&](int halves) {
        const long long waves = (long long)((L.max_nc + 15) / 16) * halves;
        return dim3((unsigned)((waves + DG_WAVES - 1) / DG_WAVES), L.n_inst);
    };
    if (nzp <= 64)
        hipLaunchKernelGGL((split_residual_kernel<4, 4>), grid(1), dim3(64 * DG_WAVES), 0, s, L,
                           Lo, H0p, nzp, e, c, spec, bsrc, x, vx, src, rho, active);
    else if (nzp <= 160)
        hipLaunchKernelGGL((split_residual_kernel<10, 5>), grid(2), dim3(64 * DG_WAVES), 0, s, L,
                           Lo, H0p, nzp, e, c, spec, bsrc, x, vx, src, rho, active);
    else
        hipLaunchKernelGGL((split_residual_kernel<14, 7>), grid(2), dim3(64 * DG_WAVES), 0, s, L,
                           Lo, H0p, nzp, e, c, spec, bsrc, x, vx, src, rho, active);
    MSX_LAUNCH_CHECK();
}

// Preconditioner factors fac[j][z] = ((H0 + e_j)^-1 e_N)(z) = sum_m Q[z][m] Q[N-1][m] / (e_j + lam_m)
// for every channel j (multiscat_gmres.f90:255-271), FP64 on the matrix cores: the first
// product of the diagonal blocks is trivial here (Q^T e_N is the last row of Q), so each
// wavefront builds W[m][j] = Q[N-1][m] / (e_j + lam_m) for 16 channels directly as the B
// operand of Y^T = Q W (same K order and A-row permutation as diag_gemm_kernel).
template <int NT>
__global__ __launch_bounds__(256) void factors_mfma_kernel(long long total, int nz, EigTables<double> E,
                                                           const double *__restrict__ e,
                                                           double *__restrict__ fac) {
    using M = Mfma<double>;
    using acc_t = M::acc;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int l16 = lane & 15, g = lane >> 4;
    const long long j0 = ((long long)blockIdx.x * 4 + wave) * 16;
    if (j0 >= total) return;
    const long long j = j0 + l16;
    const bool live = j < total;
    const double ej = e[live ? j : j0];
    const int nzp = E.nzp, NTL = nzp >> 4;
    const int pa = M::perm(l16);
    auto tile = [&](int t) { return t < NTL ? t : NTL - 1; };
    // W with every load unconditional at clamped indices: Q is zero past column nz (and the
    // extra tiles repeat the last one), so the quotient needs no select. Under a condition the
    // 40 loads and divisions became branches waited for one at a time.
    double w[NT][4];
#pragma unroll
    for (int mt = 0; mt < NT; ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = 16 * tile(mt) + 4 * g + v;
            const bool ok = 16 * mt + 4 * g + v < nz;
            const double q = ok ? E.Qp[(long long)(nz - 1) * nzp + m] : 0.0;
            const double lv = E.lam[m < nz ? m : nz - 1];
            w[mt][v] = q / (ok ? ej + lv : 1.0);
        }
    for (int zt = 0; zt < NTL; ++zt) {
        acc_t acc = acc_t{0, 0, 0, 0};
        // all of the row tile's A fragments loaded before its products
        double2 a01[NT], a23[NT];
#pragma unroll
        for (int mt = 0; mt < NT; ++mt) {
            const double *qa = E.Qp + (long long)(16 * zt + pa) * nzp + 16 * tile(mt) + 4 * g;
            a01[mt] = reinterpret_cast<const double2 *>(qa)[0];
            a23[mt] = reinterpret_cast<const double2 *>(qa)[1];
        }
#pragma unroll
        for (int mt = 0; mt < NT; ++mt) {
            acc = M::op(a01[mt].x, w[mt][0], acc);
            acc = M::op(a01[mt].y, w[mt][1], acc);
            acc = M::op(a23[mt].x, w[mt][2], acc);
            acc = M::op(a23[mt].y, w[mt][3], acc);
        }
        if (!live) continue;
        if (16 * zt + 16 <= nz) {
#pragma unroll
            for (int v = 0; v < 4; ++v) fac[j * nz + 16 * zt + 4 * g + v] = acc[v];
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int z = 16 * zt + 4 * g + v;
                if (z < nz) fac[j * nz + z] = acc[v];
            }
        }
    }
}

bool launch_factors_mfma(long long total_channels, int nz, const EigTables<double> &E,
                         const double *e, double *fac, hipStream_t s) {
    if (total_channels <= 0) return true;
    if (E.nzp % 16 != 0 || E.nzp < nz || E.nzp > 320) return false;
    const unsigned blocks = (unsigned)((total_channels + 63) / 64);
    if (E.nzp <= 160)
        hipLaunchKernelGGL(factors_mfma_kernel<10>, dim3(blocks), dim3(256), 0, s, total_channels,
                           nz, E, e, fac);
    else
        hipLaunchKernelGGL(factors_mfma_kernel<20>, dim3(blocks), dim3(256), 0, s, total_channels,
                           nz, E, e, fac);
    MSX_LAUNCH_CHECK();
    return true;
}

bool diag_mfma_supported(int nz) { return nz <= 320; }

template <typename T>
void launch_diag_blocks(const Layout &L, const EigTables<T> &E, const double *e, const cx<T> *c,
                        const T *fac, const cx<T> *x, const cx<T> *base, cx<T> *y, T sign,
                        int inverse, const int *active, hipStream_t s, int channels_per_wave) {
    if (!diag_mfma_supported(L.nz)) throw Error(3, "diagonal blocks on the matrix cores: nz <= 320");
    if (channels_per_wave != 0 && channels_per_wave != 8 && channels_per_wave != 16)
        throw Error(3, "diagonal blocks: 0, 8 or 16 channels per wavefront");
    // the smallest instantiated tile count that covers nzp / 16
    const int nt = E.nzp / 16;
#define MSX_DIAG(NR_, NT_)                                                                  \
    if (nt <= NT_) {                                                                        \
        launch_diag_gemm<T, NR_, NT_>(L, E, e, c, fac, x, base, y, sign, inverse, active, s); \
        return;                                                                             \
    }
    if constexpr (sizeof(T) == 4) {
        // Q^T in LDS up to nzp = 160
        if (nt <= 7) return launch_diag_lds_any<7>(L, E, e, c, fac, x, base, y, sign, inverse, active, s);
        if (nt <= 10) return launch_diag_lds_any<10>(L, E, e, c, fac, x, base, y, sign, inverse, active, s);
    } else {
        // 16 channels per wavefront up to nzp = 160 unless 8 are asked for: each Q fragment
        // then feeds two products
        if (channels_per_wave != 8) {
            MSX_DIAG(2, 4) MSX_DIAG(2, 7) MSX_DIAG(2, 8) MSX_DIAG(2, 10)
        } else {
            MSX_DIAG(1, 4) MSX_DIAG(1, 7) MSX_DIAG(1, 8) MSX_DIAG(1, 10)
        }
    }
    MSX_DIAG(1, 13) MSX_DIAG(1, 16) MSX_DIAG(1, 19) MSX_DIAG(1, 20)
#undef MSX_DIAG
    throw Error(3, "diagonal blocks on the matrix cores: nz <= 320");
}
template void launch_diag_blocks<float>(const Layout &, const EigTables<float> &, const double *,
                                        const cx<float> *, const float *, const cx<float> *,
                                        const cx<float> *, cx<float> *, float, int, const int *,
                                        hipStream_t, int);
template void launch_diag_blocks<double>(const Layout &, const EigTables<double> &,
                                         const double *, const cx<double> *, const double *,
                                         const cx<double> *, const cx<double> *, cx<double> *,
                                         double, int, const int *, hipStream_t, int);

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
