// DFT (convolution) form of the full channel coupling: tables and launchers (msx_dft.hip).
#pragma once

#include "msx_internal.h"

namespace msx {

// Device tables of one precision: F_x, F_y (np x np, zero padded) and
// Vhat[z][kx][ky] = (F_x V'(z) F_y)[kx][ky] / (nkx nky), V' = V with V_00 = 0.
template <typename T>
struct DftTables {
    int np;
    const cx<T> *Fx, *Fy, *Vhat;
};

struct DftStore {
    int np = 0;  // 0: table too large for the DFT kernel
    DBuf<cx<double>> Fx, Fy, Vhat, tmp;
    DBuf<cx<float>> Fxf, Fyf, Vhatf;
    DftTables<double> d64() const { return {np, Fx.p, Fy.p, Vhat.p}; }
    DftTables<float> d32() const { return {np, Fxf.p, Fyf.p, Vhatf.p}; }
};

// Padded transform size (16, 32, 48, 64 or 80) for an nkx x nky table, 0 if larger than 80.
int dft_padded_size(int nkx, int nky);

// Builds the DFT matrices and Vhat from the FP64 table V[g][z] (both precisions).
void build_dft_tables(int nkx, int nky, int nz, const cx<double> *V, DftStore &st, hipStream_t s);

// y = full off-diagonal coupling of x (compact channel layout), every active instance.
// active_hint: an upper bound of the number of active instances (-1: unknown), used to
// spread few instances over more workgroups. With x_norm and x_copy: the coupling of
// x / x_norm[b] (per instance), whose elements are also stored to x_copy as they are scaled
// (a Krylov vector normalised on its way into the operator). nodes_per_wave: nodes each
// wavefront of a workgroup handles while every instance iterates (0: the built-in 4).
template <typename T>
void launch_dft_coupling(const Layout &L, const DftTables<T> &D, const cx<T> *x, cx<T> *y,
                         const int *active, hipStream_t s, int active_hint = -1,
                         const double *x_norm = nullptr, cx<T> *x_copy = nullptr,
                         int nodes_per_wave = 0);

// The same coupling (FP32 vectors) with the four products on split-FP16 operands
// (msx_dft_h3.hip): node-major vectors and square tables of at most 32 x 32 only.
bool dft_h3_fits(const Layout &L, const DftTables<float> &D);
void launch_dft_coupling_h3(const Layout &L, const DftTables<float> &D, const cx<float> *x,
                            cx<float> *y, const int *active, hipStream_t s, int active_hint = -1,
                            const double *x_norm = nullptr, cx<float> *x_copy = nullptr,
                            int nodes_per_wave = 0);

// Nodes per workgroup of the DFT coupling's launch for a given active-instance hint and node
// grouping (the only launch parameter that depends on host-side state).
int dft_block_nodes(const Layout &L, bool fp64, int active_hint, int nodes_per_wave);

// Adds dL/dV from the coupling to grad[g][z] (complex FP64, storage order): conj(-C[g](z)),
// Re(-C[0](z)) at g = 0, with C[g](z) = sum_b sum_j lam_j(z) psi_i(z) over the channel pairs
// with G_i = G_j - g (mod grid), as a cross-correlation through the DFT tables (np 16 or 32).
// part: workspace of per-chunk partial sums, grown as needed. Channel-major states.
void launch_dft_gradient(const Layout &L, const DftTables<double> &D, const cx<double> *lam,
                         const cx<double> *psi, cx<double> *grad, DBuf<cx<double>> &part,
                         hipStream_t s);

}  // namespace msx
