// Batched GMRES building blocks (Arnoldi with classical Gram-Schmidt run twice, complex
// Givens rotations, per-instance convergence masks), for FP64 and FP32 Krylov vectors.
// Reductions always accumulate in FP64; the Hessenberg matrix, rotations and residual
// recurrence are FP64 whatever the vector precision.
//
// The reference's Arnoldi is modified Gram-Schmidt with a full back substitution each step
// and a 3-iteration convergence window (lib/multiscat_fortran/fortran/multiscat_gmres.f90:549-606);
// CGS2 produces the same Hessenberg matrix up to rounding and needs two passes over the
// basis instead of k dependent ones.
#include "msx_krylov.h"

#include <algorithm>

namespace msx {

constexpr int KR_THREADS = 256;

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block-wide sum of (re, im) pairs; every thread gets the result.
__device__ inline void block_sum2(double &a, double &b, double *red) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) { red[2 * wave] = a; red[2 * wave + 1] = b; }
    __syncthreads();
    a = 0; b = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { a += red[2 * w]; b += red[2 * w + 1]; }
}

// partial[(b*(kmax+1) + j)*nchunk + chunk] = sum over the chunk of conj(v_j) w   (j < kcount)
// partial[(b*(kmax+1) + kcount)*nchunk + chunk] = sum |w|^2           (if with_norm)
template <typename T>
__global__ __launch_bounds__(KR_THREADS) void dots_kernel(KLayout K, const cx<T> *const *vptr,
                                                          int kcount, const cx<T> *__restrict__ w,
                                                          int with_norm,
                                                          double *__restrict__ partial) {
    __shared__ double red[2 * (KR_THREADS / 64)];
    const int b = blockIdx.y;
    if (K.active && !K.active[b]) return;
    const long long n = K.len[b], base = K.start[b];
    const long long cbeg = (long long)blockIdx.x * K.chunk;
    const long long cend = cbeg + K.chunk < n ? cbeg + K.chunk : n;
    const long long slot = (long long)b * (K.kmax + 1);
    for (int j = 0; j < kcount + (with_norm ? 1 : 0); ++j) {
        double sr = 0, si = 0;
        if (j < kcount) {
            const cx<T> *v = vptr[j] + base;
            for (long long i = cbeg + threadIdx.x; i < cend; i += KR_THREADS) {
                const cx<T> a = v[i], c = w[base + i];
                sr += (double)a.re * c.re + (double)a.im * c.im;
                si += (double)a.re * c.im - (double)a.im * c.re;
            }
        } else {
            for (long long i = cbeg + threadIdx.x; i < cend; i += KR_THREADS) {
                const cx<T> c = w[base + i];
                sr += (double)c.re * c.re + (double)c.im * c.im;
            }
        }
        block_sum2(sr, si, red);
        if (threadIdx.x == 0) {
            double *p = partial + ((slot + j) * K.nchunk + blockIdx.x) * 2;
            p[0] = cbeg < n ? sr : 0.0;
            p[1] = cbeg < n ? si : 0.0;
        }
    }
}

// Block sums of the pairs j < np of KM register pairs (the pair at index npos, if >= 0, is
// (nr, 0) instead), written as dots_kernel writes them: the same wave sums and the same in-order
// sum over the wavefronts as block_sum2, with one barrier for all pairs.
template <int KM>
__device__ inline void write_pair_sums(const double (&sr)[KM], const double (&si)[KM], int kcount,
                                       bool with_norm, double nr, double *red, double *partial,
                                       long long slot, int nchunk, bool live) {
    constexpr int NW = KR_THREADS / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < KM; ++j) {
        if (j >= kcount) break;
        const double a = wave_sum(sr[j]), c = wave_sum(si[j]);
        if (lane == 0) {
            red[(wave * (KM + 1) + j) * 2] = a;
            red[(wave * (KM + 1) + j) * 2 + 1] = c;
        }
    }
    if (with_norm) {
        const double a = wave_sum(nr), c = wave_sum(0.0);
        if (lane == 0) {
            red[(wave * (KM + 1) + kcount) * 2] = a;
            red[(wave * (KM + 1) + kcount) * 2 + 1] = c;
        }
    }
    __syncthreads();
    const int j = threadIdx.x;
    if (j < kcount + (with_norm ? 1 : 0)) {
        double ro = 0, io = 0;
        for (int w = 0; w < NW; ++w) {
            ro += red[(w * (KM + 1) + j) * 2];
            io += red[(w * (KM + 1) + j) * 2 + 1];
        }
        double *p = partial + ((slot + j) * nchunk + blockIdx.x) * 2;
        p[0] = live ? ro : 0.0;
        p[1] = live ? io : 0.0;
    }
}

// dots_kernel for kcount (+ the norm) <= KM with the sums in registers: w is read once per
// element, the kcount basis values of an element are loaded together, and one barrier serves
// all the block sums (dots_kernel walks the basis with a block reduction, two barriers, per
// vector). Same per-thread element order and the same reductions as dots_kernel.
template <typename T, int KM>
__global__ __launch_bounds__(KR_THREADS) void dots_reg_kernel(KLayout K, const cx<T> *const *vptr,
                                                              int kcount,
                                                              const cx<T> *__restrict__ w,
                                                              int with_norm,
                                                              double *__restrict__ partial) {
    __shared__ double red[(KR_THREADS / 64) * (KM + 1) * 2];
    const int b = blockIdx.y;
    if (K.active && !K.active[b]) return;
    const long long n = K.len[b], base = K.start[b];
    const long long cbeg = (long long)blockIdx.x * K.chunk;
    const long long cend = cbeg + K.chunk < n ? cbeg + K.chunk : n;
    const long long slot = (long long)b * (K.kmax + 1);
    const cx<T> *vb[KM];
#pragma unroll
    for (int j = 0; j < KM; ++j) vb[j] = j < kcount ? vptr[j] + base : nullptr;
    double sr[KM], si[KM], nr = 0.0;
#pragma unroll
    for (int j = 0; j < KM; ++j) sr[j] = si[j] = 0.0;
    for (long long i = cbeg + threadIdx.x; i < cend; i += KR_THREADS) {
        const cx<T> c = w[base + i];
        cx<T> v[KM];
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (j < kcount) v[j] = vb[j][i];
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (j < kcount) {
                sr[j] += (double)v[j].re * c.re + (double)v[j].im * c.im;
                si[j] += (double)v[j].re * c.im - (double)v[j].im * c.re;
            }
        if (with_norm) nr += (double)c.re * c.re + (double)c.im * c.im;
    }
    write_pair_sums<KM>(sr, si, kcount, with_norm != 0, nr, red, partial, slot, K.nchunk, cbeg < n);
}

// h[b*(kmax+1) + j] (+)= sum_chunk partial   for first <= j < count
// One wavefront per sum (wavefront w of the block takes j = first + w, first + w + 4, ...):
// its lanes load the chunk partials side by side and add them in a fixed butterfly, so the
// loads of a sum are in flight together (one lane summing a chunk list in turn waited on them
// one after another: 8 to 16 us a launch).
constexpr int RP_WAVES = 4;
__global__ __launch_bounds__(64 * RP_WAVES) void reduce_partials_kernel(
    KLayout K, int first, int count, const double *__restrict__ partial,
    cx<double> *__restrict__ h, int accumulate) {
    const int b = blockIdx.x;
    if (K.active && !K.active[b]) return;
    const long long slot = (long long)b * (K.kmax + 1);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = first + wave; j < count; j += RP_WAVES) {
        const double2 *p = reinterpret_cast<const double2 *>(partial + (slot + j) * K.nchunk * 2);
        double sr = 0, si = 0;
        for (int c = lane; c < K.nchunk; c += 64) {
            const double2 v = p[c];
            sr += v.x;
            si += v.y;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            sr += __shfl_xor(sr, o);
            si += __shfl_xor(si, o);
        }
        if (lane == 0) {
            cx<double> &dst = h[slot + j];
            if (accumulate) dst = mk(dst.re + sr, dst.im + si);
            else dst = mk(sr, si);
        }
    }
}

// The coefficients of an update for instance b, in LDS as T: h[slot + j], or (pin given) the
// sums of the chunk partials pin[(slot + j) nchunk + c] formed here, in the order of
// reduce_partials_kernel (lane-strided, then the same butterfly), so that every workgroup of the
// instance holds the same values and they equal the reduction kernel's bit for bit; workgroup 0
// also stores the sums to hout[slot + j] for the Givens kernel. This spares a reduction launch
// (and the gap after it) between a projection and the update that consumes it.
template <typename T>
__device__ __forceinline__ void load_coefficients(const KLayout &K, long long slot, int kcount,
                                                  const cx<double> *__restrict__ h,
                                                  const double *__restrict__ pin,
                                                  cx<double> *__restrict__ hout, cx<T> *hs) {
    if (pin) {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int j = wave; j < kcount; j += KR_THREADS / 64) {
            const double2 *p = reinterpret_cast<const double2 *>(pin + (slot + j) * K.nchunk * 2);
            double sr = 0, si = 0;
            for (int c = lane; c < K.nchunk; c += 64) {
                const double2 v = p[c];
                sr += v.x;
                si += v.y;
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                sr += __shfl_xor(sr, o);
                si += __shfl_xor(si, o);
            }
            if (lane == 0) {
                hs[j] = mk((T)sr, (T)si);
                if (hout && blockIdx.x == 0) hout[slot + j] = mk(sr, si);
            }
        }
    } else {
        for (int j = threadIdx.x; j < kcount; j += KR_THREADS) {
            const cx<double> hv = h[slot + j];
            hs[j] = mk((T)hv.re, (T)hv.im);
        }
    }
    __syncthreads();
}

// w -= sum_{j<kcount} v_j h_j ; optionally norm partials of the updated w in slot kcount
template <typename T>
__global__ __launch_bounds__(KR_THREADS) void axpy_kernel(KLayout K, const cx<T> *const *vptr,
                                                          int kcount,
                                                          const cx<double> *__restrict__ h,
                                                          cx<T> *__restrict__ w, int with_norm,
                                                          double *__restrict__ partial) {
    __shared__ cx<T> hs[KR_MAX_BASIS];
    __shared__ double red[2 * (KR_THREADS / 64)];
    const int b = blockIdx.y;
    if (K.active && !K.active[b]) return;
    const long long n = K.len[b], base = K.start[b];
    const long long slot = (long long)b * (K.kmax + 1);
    for (int j = threadIdx.x; j < kcount; j += KR_THREADS) {
        const cx<double> hv = h[slot + j];
        hs[j] = mk((T)hv.re, (T)hv.im);
    }
    __syncthreads();
    const long long cbeg = (long long)blockIdx.x * K.chunk;
    const long long cend = cbeg + K.chunk < n ? cbeg + K.chunk : n;
    double nrm = 0;
    for (long long i = cbeg + threadIdx.x; i < cend; i += KR_THREADS) {
        cx<T> acc = w[base + i];
        for (int j = 0; j < kcount; ++j) {
            const cx<T> v = vptr[j][base + i];
            const cx<T> hj = hs[j];
            acc.re -= v.re * hj.re - v.im * hj.im;
            acc.im -= v.re * hj.im + v.im * hj.re;
        }
        w[base + i] = acc;
        nrm += (double)acc.re * acc.re + (double)acc.im * acc.im;
    }
    if (with_norm) {
        double dummy = 0;
        block_sum2(nrm, dummy, red);
        if (threadIdx.x == 0) {
            double *p = partial + ((slot + kcount) * K.nchunk + blockIdx.x) * 2;
            p[0] = cbeg < n ? nrm : 0.0;
            p[1] = 0.0;
        }
    }
}

// axpy_kernel for kcount <= KM: the basis pointers held in registers and the kcount basis values
// of an element loaded together before its updates (the runtime-length loop over the basis
// issued them one at a time). Same per-thread order and reduction as axpy_kernel.
template <typename T, int KM>
__global__ __launch_bounds__(KR_THREADS) void axpy_reg_kernel(KLayout K, const cx<T> *const *vptr,
                                                              int kcount,
                                                              const cx<double> *__restrict__ h,
                                                              cx<T> *__restrict__ w, int with_norm,
                                                              double *__restrict__ partial,
                                                              const double *__restrict__ pin,
                                                              cx<double> *__restrict__ hout) {
    __shared__ cx<T> hs[KM];
    __shared__ double red[2 * (KR_THREADS / 64)];
    const int b = blockIdx.y;
    if (K.active && !K.active[b]) return;
    const long long n = K.len[b], base = K.start[b];
    const long long slot = (long long)b * (K.kmax + 1);
    load_coefficients<T>(K, slot, kcount, h, pin, hout, hs);
    const cx<T> *vb[KM];
#pragma unroll
    for (int j = 0; j < KM; ++j) vb[j] = j < kcount ? vptr[j] + base : nullptr;
    const long long cbeg = (long long)blockIdx.x * K.chunk;
    const long long cend = cbeg + K.chunk < n ? cbeg + K.chunk : n;
    double nrm = 0;
    for (long long i = cbeg + threadIdx.x; i < cend; i += KR_THREADS) {
        cx<T> v[KM];
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (j < kcount) v[j] = vb[j][i];
        cx<T> acc = w[base + i];
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (j < kcount) {
                const cx<T> hj = hs[j];
                acc.re -= v[j].re * hj.re - v[j].im * hj.im;
                acc.im -= v[j].re * hj.im + v[j].im * hj.re;
            }
        w[base + i] = acc;
        nrm += (double)acc.re * acc.re + (double)acc.im * acc.im;
    }
    if (with_norm) {
        double dummy = 0;
        block_sum2(nrm, dummy, red);
        if (threadIdx.x == 0) {
            double *p = partial + ((slot + kcount) * K.nchunk + blockIdx.x) * 2;
            p[0] = cbeg < n ? nrm : 0.0;
            p[1] = 0.0;
        }
    }
}

// The first update of CGS2 and the second projection in one pass over the basis:
//   w <- w - sum_{j<kcount} v_j h_j,  then partial sums of conj(v_j) w (as dots_kernel writes
// them), with the kcount <= KM basis values of each element held in registers between the two.
// Same per-thread element order and block reductions as axpy_kernel then dots_kernel, so the
// results are bitwise those of the two separate passes. with_norm: |w|^2 partials of the
// updated w in slot kcount as well.
template <typename T, int KM>
__global__ __launch_bounds__(KR_THREADS) void axpy_dots_kernel(KLayout K,
                                                               const cx<T> *const *vptr,
                                                               int kcount,
                                                               const cx<double> *__restrict__ h,
                                                               cx<T> *__restrict__ w,
                                                               double *__restrict__ partial,
                                                               const double *__restrict__ pin,
                                                               cx<double> *__restrict__ hout,
                                                               int with_norm) {
    __shared__ cx<T> hs[KM];
    __shared__ double red[(KR_THREADS / 64) * (KM + 1) * 2];
    const int b = blockIdx.y;
    if (K.active && !K.active[b]) return;
    const long long n = K.len[b], base = K.start[b];
    const long long slot = (long long)b * (K.kmax + 1);
    load_coefficients<T>(K, slot, kcount, h, pin, hout, hs);
    const cx<T> *vb[KM];
#pragma unroll
    for (int j = 0; j < KM; ++j) vb[j] = j < kcount ? vptr[j] + base : nullptr;
    const long long cbeg = (long long)blockIdx.x * K.chunk;
    const long long cend = cbeg + K.chunk < n ? cbeg + K.chunk : n;
    double sr[KM], si[KM], nr = 0.0;
#pragma unroll
    for (int j = 0; j < KM; ++j) sr[j] = si[j] = 0.0;
    for (long long i = cbeg + threadIdx.x; i < cend; i += KR_THREADS) {
        cx<T> v[KM];
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (j < kcount) v[j] = vb[j][i];
        cx<T> acc = w[base + i];
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (j < kcount) {
                const cx<T> hj = hs[j];
                acc.re -= v[j].re * hj.re - v[j].im * hj.im;
                acc.im -= v[j].re * hj.im + v[j].im * hj.re;
            }
        w[base + i] = acc;
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (j < kcount) {
                sr[j] += (double)v[j].re * acc.re + (double)v[j].im * acc.im;
                si[j] += (double)v[j].re * acc.im - (double)v[j].im * acc.re;
            }
        if (with_norm) nr += (double)acc.re * acc.re + (double)acc.im * acc.im;
    }
    write_pair_sums<KM>(sr, si, kcount, with_norm != 0, nr, red, partial, slot, K.nchunk, cbeg < n);
}

// Selective second Gram-Schmidt pass: after the first pass, partial holds the chunk partials
// of the second projection h2 = V^H w (slots 0 .. kcount-1) and of |w|^2 (slot kcount). An
// iterating instance takes the second update (mask[b] = 1) only if |h2| > tau |w|, i.e. the
// first pass left the new vector's projection on the basis above tau; otherwise mask[b] = 0
// and its h2 is stored as zero (the update and the Hessenberg column then use h1 alone).
__global__ __launch_bounds__(64 * RP_WAVES) void reorth_select_kernel(
    KLayout K, int kcount, const double *__restrict__ partial, int *__restrict__ mask,
    cx<double> *__restrict__ h2, double tau2) {
    __shared__ double acc[RP_WAVES];
    __shared__ double nrm;
    const int b = blockIdx.x;
    if (!K.active[b]) {
        if (threadIdx.x == 0) mask[b] = 0;
        return;
    }
    const long long slot = (long long)b * (K.kmax + 1);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double s2 = 0.0;
    for (int j = wave; j <= kcount; j += RP_WAVES) {
        const double2 *p = reinterpret_cast<const double2 *>(partial + (slot + j) * K.nchunk * 2);
        double sr = 0, si = 0;
        for (int c = lane; c < K.nchunk; c += 64) {
            const double2 v = p[c];
            sr += v.x;
            si += v.y;
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        if (j < kcount) s2 += sr * sr + si * si;
        else if (lane == 0) nrm = sr;
    }
    if (lane == 0) acc[wave] = s2;
    __syncthreads();
    const double tot = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    const int take = tot > tau2 * nrm ? 1 : 0;
    if (threadIdx.x == 0) mask[b] = take;
    if (!take)
        for (int j = threadIdx.x; j < kcount; j += blockDim.x) h2[slot + j] = mk(0.0, 0.0);
}

// x += sum_{j<kk_b} v_j y_j  (kk_b per instance)
template <typename T, typename TX>
__global__ __launch_bounds__(KR_THREADS) void update_kernel(KLayout K, const cx<T> *const *vptr,
                                                            const int *__restrict__ kk,
                                                            const cx<double> *__restrict__ y,
                                                            cx<TX> *__restrict__ x,
                                                            const int *__restrict__ mask) {
    __shared__ cx<double> ys[KR_MAX_BASIS];
    const int b = blockIdx.y;
    if (mask && !mask[b]) return;
    const int kb = kk[b];
    if (kb <= 0) return;
    const long long n = K.len[b], base = K.start[b];
    const long long slot = (long long)b * (K.kmax + 1);
    for (int j = threadIdx.x; j < kb; j += KR_THREADS) ys[j] = y[slot + j];
    __syncthreads();
    const long long cbeg = (long long)blockIdx.x * K.chunk;
    const long long cend = cbeg + K.chunk < n ? cbeg + K.chunk : n;
    for (long long i = cbeg + threadIdx.x; i < cend; i += KR_THREADS) {
        cx<double> acc = mk((double)x[base + i].re, (double)x[base + i].im);
        for (int j = 0; j < kb; ++j) {
            const cx<T> v = vptr[j][base + i];
            const cx<double> yj = ys[j];
            acc.re += (double)v.re * yj.re - (double)v.im * yj.im;
            acc.im += (double)v.re * yj.im + (double)v.im * yj.re;
        }
        x[base + i] = mk((TX)acc.re, (TX)acc.im);
    }
}

// x += sum_{j<kk_b} v_j y_j with the basis in node-major order (Layout::nm) and x in the
// channel-major batch order: one 32-channel x 32-node tile per workgroup (blockIdx.x: tile,
// blockIdx.y: instance). The sums run along channels (the basis vectors' contiguous index),
// cross over in LDS, and update x along nodes (its contiguous index).
template <typename T>
__global__ __launch_bounds__(KR_THREADS) void update_nm_kernel(Layout L, int kmax,
                                                               const cx<T> *const *vptr,
                                                               const int *__restrict__ kk,
                                                               const cx<double> *__restrict__ y,
                                                               cx<double> *__restrict__ x) {
    constexpr int TT = 32;
    __shared__ cx<double> tile[TT][TT + 1];
    const int b = blockIdx.y;
    const int kb = kk[b];
    if (kb <= 0) return;
    const int nc = L.nc[b], nz = L.nz;
    const int tz = (nz + TT - 1) / TT;
    const int j0 = (blockIdx.x / tz) * TT, z0 = (blockIdx.x % tz) * TT;
    if (j0 >= nc) return;
    const long long off = L.off[b];
    const cx<double> *yb = y + (long long)b * (kmax + 1);
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;
    for (int zz = ty; zz < TT; zz += KR_THREADS / TT) {
        const int z = z0 + zz, j = j0 + tx;
        cx<double> acc = mk(0.0, 0.0);
        if (z < nz && j < nc) {
            const long long i = off * nz + (long long)z * nc + j;
            for (int k = 0; k < kb; ++k) {
                const cx<T> v = vptr[k][i];
                const cx<double> yk = yb[k];
                acc.re += (double)v.re * yk.re - (double)v.im * yk.im;
                acc.im += (double)v.re * yk.im + (double)v.im * yk.re;
            }
        }
        tile[zz][tx] = acc;
    }
    __syncthreads();
    for (int jj = ty; jj < TT; jj += KR_THREADS / TT) {
        const int j = j0 + jj, z = z0 + tx;
        if (j < nc && z < nz) {
            const long long i = (off + j) * nz + z;
            x[i] = x[i] + tile[tx][jj];
        }
    }
}

// dst = src * (1 / s_b)  for instances with mask set and s_b > 0
template <typename T>
__global__ __launch_bounds__(KR_THREADS) void scale_kernel(KLayout K, const cx<T> *src, cx<T> *dst,
                                                           const double *__restrict__ s,
                                                           const int *__restrict__ mask) {
    const int b = blockIdx.y;
    if (mask && !mask[b]) return;
    const double sb = s[b];
    if (!(sb > 0)) return;
    const T inv = (T)(1.0 / sb);
    const long long n = K.len[b], base = K.start[b];
    const long long cbeg = (long long)blockIdx.x * K.chunk;
    const long long cend = cbeg + K.chunk < n ? cbeg + K.chunk : n;
    for (long long i = cbeg + threadIdx.x; i < cend; i += KR_THREADS) {
        const cx<T> v = src[base + i];
        dst[base + i] = mk(v.re * inv, v.im * inv);
    }
}

// dst = a + alpha * c (alpha real), all instances with mask
template <typename T>
__global__ __launch_bounds__(KR_THREADS) void axpby_kernel(KLayout K, const cx<T> *__restrict__ a,
                                                           const cx<T> *__restrict__ c, T alpha,
                                                           cx<T> *__restrict__ dst,
                                                           const int *__restrict__ mask) {
    const int b = blockIdx.y;
    if (mask && !mask[b]) return;
    const long long n = K.len[b], base = K.start[b];
    const long long cbeg = (long long)blockIdx.x * K.chunk;
    const long long cend = cbeg + K.chunk < n ? cbeg + K.chunk : n;
    for (long long i = cbeg + threadIdx.x; i < cend; i += KR_THREADS) {
        const cx<T> u = a[base + i], v = c[base + i];
        dst[base + i] = mk(u.re + alpha * v.re, u.im + alpha * v.im);
    }
}

// LINPACK-style complex Givens rotation (multiscat_gmres.f90:689-716).
__device__ inline void givens(cx<double> a, cx<double> b, cx<double> &c, cx<double> &s,
                              double &r) {
    const double aa = sqrt(norm2(a)), ab = sqrt(norm2(b));
    const cx<double> rho = aa > ab ? a : b;
    const double scale = aa + ab;
    if (scale == 0.0) {
        c = mk(1.0, 0.0);
        s = mk(0.0, 0.0);
        r = 0.0;
        return;
    }
    const cx<double> as = a * (1.0 / scale), bs = b * (1.0 / scale);
    r = scale * sqrt(norm2(as) + norm2(bs));
    if (rho.re < 0.0) r = -r;
    c = conj(a * (1.0 / r));
    s = b * (1.0 / r);
}

// Block total of `alive` into *count; with host_flags (pinned host memory, n_inst + 1 ints)
// also the active flags and the count, written straight to the host (no copy launch).
__device__ void finish_count(const KLayout &K, int alive, int *red, int *count, int *host_flags) {
    for (int o = 32; o > 0; o >>= 1) alive += __shfl_xor(alive, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = alive;
    __syncthreads();
    int t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    if (threadIdx.x == 0) *count = t;
    if (host_flags) {
        for (int b = threadIdx.x; b < K.n_inst; b += blockDim.x) host_flags[b] = K.active[b];
        if (threadIdx.x == 0) host_flags[K.n_inst] = t;
        __threadfence_system();
    }
}

// One thread per instance: finish column k of the Hessenberg matrix (hcol + hcol2, or hcol alone
// when hcol2 is null), rotate, update the
// residual recurrence and the convergence window. Writes the number of instances still
// active to *count.
__global__ void givens_kernel(KLayout K, int k, const cx<double> *__restrict__ hcol,
                              const cx<double> *__restrict__ hcol2, KState S,
                              int *__restrict__ count, int *host_flags,
                              const double *__restrict__ norm_partial) {
    __shared__ int red[KR_THREADS / 64];
    int alive = 0;
    for (int b = threadIdx.x; b < K.n_inst; b += blockDim.x) {
        if (!K.active[b]) continue;
        const long long slot = (long long)b * (K.kmax + 1);
        cx<double> *H = S.H + (long long)b * (K.kmax + 1) * (K.kmax + 1);  // column-major
#define HH(i, j) H[(long long)(j) * (K.kmax + 1) + (i)]
        for (int i = 0; i <= k; ++i) {
            const cx<double> h = hcol2 ? hcol[slot + i] + hcol2[slot + i] : hcol[slot + i];
            HH(i, k) = mk(S.hscale * h.re, S.hscale * h.im);
        }
        HH(k, k).re += S.hshift;
        // |w|^2: from the chunk partials of slot k + 1 (summed here, in chunk order, as the
        // reduction kernel would), or already reduced into hcol
        double n2;
        if (norm_partial) {
            // sixteen loads in flight at a time, summed in chunk order
            const double *p = norm_partial + (slot + k + 1) * K.nchunk * 2;
            n2 = 0.0;
            int c = 0;
            for (; c + 16 <= K.nchunk; c += 16) {
                double v[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) v[i] = p[2 * (c + i)];
#pragma unroll
                for (int i = 0; i < 16; ++i) n2 += v[i];
            }
            for (; c < K.nchunk; ++c) n2 += p[2 * c];
        } else {
            n2 = hcol[slot + k + 1].re;
        }
        const double hn = sqrt(n2);
        HH(k + 1, k) = mk(S.hscale * hn, 0.0);
        cx<double> *cs = S.cs + slot, *sn = S.sn + slot, *g = S.g + slot;
        for (int i = 0; i < k; ++i) {
            const cx<double> t = cs[i] * HH(i, k) + conj(sn[i]) * HH(i + 1, k);
            HH(i + 1, k) = conj(cs[i]) * HH(i + 1, k) - sn[i] * HH(i, k);
            HH(i, k) = t;
        }
        double r;
        givens(HH(k, k), HH(k + 1, k), cs[k], sn[k], r);
        HH(k, k) = mk(r, 0.0);
        HH(k + 1, k) = mk(0.0, 0.0);
        g[k + 1] = mk(0.0, 0.0) - sn[k] * g[k];
        g[k] = cs[k] * g[k];
#undef HH
        const double res = sqrt(norm2(g[k + 1]));
        const double ref = S.beta0[b] > 2.2250738585072014e-308 ? S.beta0[b] : 2.2250738585072014e-308;
        S.resid[b] = res / ref;
        const double eps = S.eps_inst ? S.eps_inst[b] : S.eps;
        S.kconv[b] = res < eps * ref ? S.kconv[b] + 1 : 0;
        S.cycle_iters[b] = k + 1;
        S.iters[b] += 1;
        S.hnorm[b] = hn;
        if (S.kconv[b] >= S.window || hn == 0.0) {
            K.active[b] = 0;
            S.converged[b] = 1;
        } else {
            alive += 1;
        }
    }
    finish_count(K, alive, red, count, host_flags);
}

// One thread per instance: back substitution R y = g over the cycle's kk columns.
__global__ void backsolve_kernel(KLayout K, KState S, const int *__restrict__ mask) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= K.n_inst || (mask && !mask[b])) return;
    const int kk = S.cycle_iters[b];
    const long long slot = (long long)b * (K.kmax + 1);
    const cx<double> *H = S.H + (long long)b * (K.kmax + 1) * (K.kmax + 1);
    cx<double> *y = S.y + slot;
    for (int j = 0; j < kk; ++j) y[j] = S.g[slot + j];
    for (int j = kk - 1; j >= 0; --j) {
        y[j] = cdiv(y[j], H[(long long)j * (K.kmax + 1) + j]);
        for (int i = 0; i < j; ++i) y[i] = y[i] - H[(long long)j * (K.kmax + 1) + i] * y[j];
    }
}

// Start of a cycle: beta = ||r0||, g = (beta, 0, ...), kconv kept, cycle_iters = 0.
// Instances with beta == 0 are already solved.
__global__ void cycle_start_kernel(KLayout K, KState S, const cx<double> *__restrict__ nrm2,
                                   int first_cycle, int *__restrict__ count,
                                   const cx<double> *__restrict__ ref_nrm2, int *host_flags) {
    __shared__ int red[KR_THREADS / 64];
    int alive = 0;
    for (int b = threadIdx.x; b < K.n_inst; b += blockDim.x) {
        S.cycle_iters[b] = 0;  // inactive instances must not replay an older cycle's update
        if (!K.active[b]) continue;
        const long long slot = (long long)b * (K.kmax + 1);
        const double beta = sqrt(nrm2[slot].re);
        S.beta[b] = beta;
        if (first_cycle) S.beta0[b] = ref_nrm2 ? sqrt(ref_nrm2[slot].re) : beta;
        S.g[slot] = mk(beta, 0.0);
        S.cycle_iters[b] = 0;
        const double eps = S.eps_inst ? S.eps_inst[b] : S.eps;
        if (!(beta > 0.0) || (ref_nrm2 && first_cycle && beta < eps * S.beta0[b])) {
            K.active[b] = 0;
            S.converged[b] = 1;
            S.resid[b] = S.beta0[b] > 0.0 ? beta / S.beta0[b] : 0.0;
        } else {
            alive += 1;
        }
    }
    finish_count(K, alive, red, count, host_flags);
}

// Start of a solve: remember which instances take part, clear their counters.
__global__ void gmres_reset_kernel(KLayout K, KState S, int *__restrict__ enabled) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < K.n_inst; b += gridDim.x * blockDim.x) {
        enabled[b] = K.active[b];
        S.converged[b] = 0;
        S.kconv[b] = 0;
        S.iters[b] = 0;
        S.cycle_iters[b] = 0;
        S.resid[b] = 0.0;
    }
}

// Restart: instances that take part and have not converged iterate again.
__global__ void reactivate_kernel(KLayout K, KState S, const int *__restrict__ enabled) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < K.n_inst; b += gridDim.x * blockDim.x)
        K.active[b] = enabled[b] && !S.converged[b];
}

void k_reset(const KLayout &K, const KState &S, int *enabled, hipStream_t s) {
    hipLaunchKernelGGL(gmres_reset_kernel, dim3((K.n_inst + 255) / 256), dim3(256), 0, s, K, S,
                       enabled);
    MSX_LAUNCH_CHECK();
}

void k_reactivate(const KLayout &K, const KState &S, const int *enabled, hipStream_t s) {
    hipLaunchKernelGGL(reactivate_kernel, dim3((K.n_inst + 255) / 256), dim3(256), 0, s, K, S,
                       enabled);
    MSX_LAUNCH_CHECK();
}

// dst[b] = src[b * stride].re (dst may be pinned host memory mapped for the device)
__global__ void gather_re_kernel(int n, int stride, const cx<double> *__restrict__ src,
                                 double *__restrict__ dst) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) dst[b] = src[(long long)b * stride].re;
}

// ---- launchers ----------------------------------------------------------------------------
static dim3 kgrid(const KLayout &K) { return dim3(K.nchunk, K.n_inst); }

void k_gather_re(int n, int stride, const cx<double> *src, double *dst, hipStream_t s) {
    hipLaunchKernelGGL(gather_re_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, stride, src, dst);
    MSX_LAUNCH_CHECK();
}

template <typename T>
void k_dots(const KLayout &K, const cx<T> *const *vptr, int kcount, const cx<T> *w,
            int with_norm, double *partial, cx<double> *h, int accumulate, hipStream_t s) {
    const int np = kcount + (with_norm ? 1 : 0);
    if (np <= 8)
        hipLaunchKernelGGL((dots_reg_kernel<T, 8>), kgrid(K), dim3(KR_THREADS), 0, s, K, vptr,
                           kcount, w, with_norm, partial);
    else if (np <= 16)
        hipLaunchKernelGGL((dots_reg_kernel<T, 16>), kgrid(K), dim3(KR_THREADS), 0, s, K, vptr,
                           kcount, w, with_norm, partial);
    else
        hipLaunchKernelGGL((dots_kernel<T>), kgrid(K), dim3(KR_THREADS), 0, s, K, vptr, kcount, w,
                           with_norm, partial);
    MSX_LAUNCH_CHECK();
    if (!h) return;  // the consumer sums the partials itself
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(K.n_inst), dim3(64 * RP_WAVES), 0, s, K, 0,
                       kcount + (with_norm ? 1 : 0), partial, h, accumulate);
    MSX_LAUNCH_CHECK();
}

template <typename T>
void k_axpy(const KLayout &K, const cx<T> *const *vptr, int kcount, const cx<double> *h,
            cx<T> *w, int with_norm, double *partial, cx<double> *hnorm_dst, hipStream_t s,
            const double *pin, cx<double> *hout) {
    if (kcount > KR_MAX_BASIS) throw Error(3, "Krylov dimension exceeds the kernel limit");
    if (pin && kcount > 16) throw Error(3, "in-kernel coefficient sums cover 16 basis vectors");
    if (kcount <= 8)
        hipLaunchKernelGGL((axpy_reg_kernel<T, 8>), kgrid(K), dim3(KR_THREADS), 0, s, K, vptr,
                           kcount, h, w, with_norm, partial, pin, hout);
    else if (kcount <= 16)
        hipLaunchKernelGGL((axpy_reg_kernel<T, 16>), kgrid(K), dim3(KR_THREADS), 0, s, K, vptr,
                           kcount, h, w, with_norm, partial, pin, hout);
    else
        hipLaunchKernelGGL((axpy_kernel<T>), kgrid(K), dim3(KR_THREADS), 0, s, K, vptr, kcount, h,
                           w, with_norm, partial);
    MSX_LAUNCH_CHECK();
    if (with_norm && hnorm_dst) {
        // the norm partials sit in slot kcount; reduce only that slot into hnorm_dst (without
        // hnorm_dst the consumer sums them itself)
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(K.n_inst), dim3(64 * RP_WAVES), 0, s, K,
                           kcount, kcount + 1, partial, hnorm_dst, 0);
        MSX_LAUNCH_CHECK();
    }
}

void k_reorth_select(const KLayout &K, int kcount, const double *partial, int *mask,
                     cx<double> *h2, double tau, hipStream_t s) {
    hipLaunchKernelGGL(reorth_select_kernel, dim3(K.n_inst), dim3(64 * RP_WAVES), 0, s, K, kcount,
                       partial, mask, h2, tau * tau);
    MSX_LAUNCH_CHECK();
}

template <typename T>
bool k_axpy_dots(const KLayout &K, const cx<T> *const *vptr, int kcount, const cx<double> *h,
                 cx<T> *w, double *partial, cx<double> *h2, hipStream_t s, const double *pin,
                 cx<double> *hout, int with_norm) {
    if (kcount <= 8)
        hipLaunchKernelGGL((axpy_dots_kernel<T, 8>), kgrid(K), dim3(KR_THREADS), 0, s, K, vptr,
                           kcount, h, w, partial, pin, hout, with_norm);
    else if (kcount <= 16)
        hipLaunchKernelGGL((axpy_dots_kernel<T, 16>), kgrid(K), dim3(KR_THREADS), 0, s, K, vptr,
                           kcount, h, w, partial, pin, hout, with_norm);
    else
        return false;
    MSX_LAUNCH_CHECK();
    if (!h2) return true;  // the consumer sums the partials itself
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(K.n_inst), dim3(64 * RP_WAVES), 0, s, K, 0, kcount,
                       partial, h2, 0);
    MSX_LAUNCH_CHECK();
    return true;
}

template <typename T>
void k_update_nm(const Layout &L, int kmax, const cx<T> *const *vptr, const int *kk,
                 const cx<double> *y, cx<double> *x, hipStream_t s) {
    const unsigned tiles = (unsigned)(((L.max_nc + 31) / 32) * ((L.nz + 31) / 32));
    hipLaunchKernelGGL((update_nm_kernel<T>), dim3(std::max(tiles, 1u), L.n_inst), dim3(KR_THREADS),
                       0, s, L, kmax, vptr, kk, y, x);
    MSX_LAUNCH_CHECK();
}
template void k_update_nm<float>(const Layout &, int, const cx<float> *const *, const int *,
                                 const cx<double> *, cx<double> *, hipStream_t);
template void k_update_nm<double>(const Layout &, int, const cx<double> *const *, const int *,
                                  const cx<double> *, cx<double> *, hipStream_t);

template <typename T, typename TX>
void k_update(const KLayout &K, const cx<T> *const *vptr, const int *kk, const cx<double> *y,
              cx<TX> *x, const int *mask, hipStream_t s) {
    hipLaunchKernelGGL((update_kernel<T, TX>), kgrid(K), dim3(KR_THREADS), 0, s, K, vptr, kk, y, x,
                       mask);
    MSX_LAUNCH_CHECK();
}

template <typename T>
void k_scale(const KLayout &K, const cx<T> *src, cx<T> *dst, const double *sv, const int *mask,
             hipStream_t s) {
    hipLaunchKernelGGL((scale_kernel<T>), kgrid(K), dim3(KR_THREADS), 0, s, K, src, dst, sv, mask);
    MSX_LAUNCH_CHECK();
}

template <typename T>
void k_axpby(const KLayout &K, const cx<T> *a, const cx<T> *c, T alpha, cx<T> *dst,
             const int *mask, hipStream_t s) {
    hipLaunchKernelGGL((axpby_kernel<T>), kgrid(K), dim3(KR_THREADS), 0, s, K, a, c, alpha, dst,
                       mask);
    MSX_LAUNCH_CHECK();
}

void k_givens(const KLayout &K, int k, const cx<double> *hcol, const cx<double> *hcol2,
              const KState &S, int *count, hipStream_t s, int *host_flags,
              const double *norm_partial) {
    hipLaunchKernelGGL(givens_kernel, dim3(1), dim3(KR_THREADS), 0, s, K, k, hcol, hcol2, S,
                       count, host_flags, norm_partial);
    MSX_LAUNCH_CHECK();
}

void k_backsolve(const KLayout &K, const KState &S, const int *mask, hipStream_t s) {
    hipLaunchKernelGGL(backsolve_kernel, dim3((K.n_inst + 63) / 64), dim3(64), 0, s, K, S, mask);
    MSX_LAUNCH_CHECK();
}

void k_cycle_start(const KLayout &K, const KState &S, const cx<double> *nrm2, int first,
                   int *count, hipStream_t s, const cx<double> *ref_nrm2, int *host_flags) {
    hipLaunchKernelGGL(cycle_start_kernel, dim3(1), dim3(KR_THREADS), 0, s, K, S, nrm2, first,
                       count, ref_nrm2, host_flags);
    MSX_LAUNCH_CHECK();
}

#define MSX_INST(T)                                                                           \
    template void k_dots<T>(const KLayout &, const cx<T> *const *, int, const cx<T> *, int,      \
                            double *, cx<double> *, int, hipStream_t);                          \
    template void k_axpy<T>(const KLayout &, const cx<T> *const *, int, const cx<double> *,     \
                            cx<T> *, int, double *, cx<double> *, hipStream_t, const double *,  \
                            cx<double> *);                                                      \
    template bool k_axpy_dots<T>(const KLayout &, const cx<T> *const *, int,                    \
                                 const cx<double> *, cx<T> *, double *, cx<double> *,           \
                                 hipStream_t, const double *, cx<double> *, int);               \
    template void k_scale<T>(const KLayout &, const cx<T> *, cx<T> *, const double *,           \
                             const int *, hipStream_t);                                         \
    template void k_axpby<T>(const KLayout &, const cx<T> *, const cx<T> *, T, cx<T> *,          \
                             const int *, hipStream_t);
MSX_INST(double)
MSX_INST(float)
template void k_update<double, double>(const KLayout &, const cx<double> *const *, const int *,
                                       const cx<double> *, cx<double> *, const int *,
                                       hipStream_t);
template void k_update<float, float>(const KLayout &, const cx<float> *const *, const int *,
                                     const cx<double> *, cx<float> *, const int *, hipStream_t);
template void k_update<float, double>(const KLayout &, const cx<float> *const *, const int *,
                                      const cx<double> *, cx<double> *, const int *,
                                      hipStream_t);

}  // namespace msx
