// krylov.h -- right-preconditioned BiCGStab on the cell stencil.
//
// Restates the iteration of pbicgstab (source/solvers/pbicgstab.f90:21-96):
//   p = r + beta*(p - omega*v);  v = A M^-1 p;  alpha = rho/(r0.v)
//   x += alpha M^-1 p;  s = r - alpha v;  t = A M^-1 s;  omega = (t.s)/(t.t)
//   x += omega M^-1 s;  r = s - omega t
// with M = diag(ap) (Jacobi, applied inline) or an explicit preconditioner
// whose z = M^-1 p and M^-1 s are computed between the kernels (PREC = 1,
// one system), and the standard beta = (rho/rho_old)(alpha/omega).
// The reference's beta denominator carries +1e-15 (pbicgstab.f90:43), which
// stalls the recurrence once rho < ~1e-13, so the tight solve instead restarts
// the shadow residual when rho breaks down.  Up to NS systems that share the
// off-diagonal coefficients (u and v momentum) advance together so that every
// coefficient is read once per sweep.
//
// Reductions without atomics: every Krylov kernel runs on the same fixed grid
// of NB blocks (each block walks a fixed set of cells, grid_walk below).  A kernel that produces dot
// products writes one partial per block; the kernel that consumes them sums
// all NB partials in block order in every block (identical result everywhere),
// and block 0 publishes the scalars for later kernels into slots indexed by
// the iteration parity, so no block reads a slot that the same launch writes.
#pragma once
#include "common.h"

namespace cms {

struct KryScal {
    double rho[2];     // r0.r at the start of iteration it (slot it&1)
    double r0n[2];     // ||r0|| of the current shadow residual
    double alpha[2];   // alpha of iteration it
    double omega[2];   // omega of iteration it
    double thr;        // stopping threshold on ||r||2
    double rmom;       // normalized residual RMS of the solve prologue
    double rnorm;      // ||r|| when the solve stopped
    int iters[2];      // iterations completed at the start of iteration it
    int done;          // solve finished
    int final_iters;
};

// Per-block partial sums, slot-major: pr[2*NB] (r.r | r0.r, or r.r | res.res
// for the prologue), pv[NB] (r0.v), pt[2*NB] (t.s | t.t).
template <int NS>
struct KryRed {
    double *pr[NS], *pv[NS], *pt[NS];
};

struct Idx {
    int c, i, j;
};

// Fixed-grid mapping of a cell list (EllOp): block b (XCD-remapped) owns a
// contiguous chunk of the list; its threads stride through the chunk.
struct Chunk {
    int begin, end;
};
__device__ __forceinline__ Chunk block_chunk(int n) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int per = ((n + (int)gridDim.x - 1) / (int)gridDim.x + KBLOCK - 1) / KBLOCK * KBLOCK;
    Chunk ch;
    ch.begin = min(n, b * per);
    ch.end = min(n, ch.begin + per);
    return ch;
}

// Cells of this block on the fixed Krylov grid of a structured grid: strips
// KW_STRIP columns wide; the rows of a strip go in groups of RP (one pass of
// the block) to its blocks in turn, so in every pass the blocks of a strip
// cover one run of consecutive rows.  Consecutive blocks run on one XCD
// (xcd_block), so the rows above and below a block's rows are the ones its
// neighbours load at the same time and meet in that XCD's L2; a thread walks
// one column.  Every active cell is visited once, whatever the grid size.
constexpr int KW_STRIP = 256;
template <class F>
__device__ __forceinline__ void grid_walk(const Geo &g, F &&f) {
    constexpr int RP = KBLOCK / KW_STRIP;                 // rows per pass
    const int G = (int)gridDim.x;
    const int sx = (g.nx + KW_STRIP - 1) / KW_STRIP;
    const int bands = max(1, G / sx);                     // blocks per strip
    const int nbx = G / bands;                            // strips in parallel (>= 1)
    const int b = xcd_block(blockIdx.x, G);
    const int bx = b / bands, by = b % bands;
    if (bx >= nbx) return;
    const int ngroups = (g.ny + RP - 1) / RP;
    const int tx = (int)threadIdx.x % KW_STRIP, ty = (int)threadIdx.x / KW_STRIP;
    for (int s = bx; s < sx; s += nbx) {
        const int col = s * KW_STRIP + tx;
        if (col >= g.nx) continue;
        for (int grp = by; grp < ngroups; grp += bands) {
            const int row = grp * RP + ty;
            if (row >= g.ny) continue;
            Idx x;
            x.i = col;
            x.j = row;
            x.c = cidx(g, col, row);
            f(x);
        }
    }
}

// Matrix operators: y_i = ap_i z_i - sum_k a_k(i) z_nb(k).  apply() is the
// Jacobi-preconditioned form, z = M^-1 x = x/ap with 1/ap stored (dinv), so
// y_i = x_i - sum_k a_k(i) x_nb dinv_nb; z = 0 on ghost cells.  C is the type
// of the coefficients and of the result.
template <class C>
struct StencilOpT {
    Geo g;
    const C *aN, *aE, *aS, *aW;
    __device__ __forceinline__ int n() const { return g.nact; }
    template <class F>
    __device__ __forceinline__ void each(F &&f) const { grid_walk(g, f); }
    __device__ __forceinline__ Idx at(int t) const {
        Idx x;
        x.j = t / g.nx;
        x.i = t - x.j * g.nx;
        x.c = cidx(g, x.i, x.j);
        return x;
    }
    // V: the vectors' type (double; float in the inner solves of the FP32 form)
    template <class V>
    __device__ __forceinline__ V apply(const V *xv, const V *dinv, const Idx &x) const {
        const int c = x.c;
        const int cN = c + g.nxp, cE = c + 1, cS = c - g.nxp, cW = c - 1;
        const V zN = x.j == g.ny - 1 ? V(0) : xv[cN] * dinv[cN];
        const V zE = x.i == g.nx - 1 ? V(0) : xv[cE] * dinv[cE];
        const V zS = x.j == 0 ? V(0) : xv[cS] * dinv[cS];
        const V zW = x.i == 0 ? V(0) : xv[cW] * dinv[cW];
        const V sum = (((V)aN[c] * zN + (V)aE[c] * zE) + (V)aS[c] * zS) + (V)aW[c] * zW;
        return xv[c] - sum;
    }
    // y_i = ap_i z_i - sum_k a_k(i) z_nb(k) for an explicit preconditioned
    // vector z (FP32, zero on the ghost ring), evaluated in C.
    __device__ __forceinline__ C apply_z(const float *z, const C *ap, const Idx &x) const {
        const int c = x.c;
        const C zN = x.j == g.ny - 1 ? C(0) : (C)z[c + g.nxp];
        const C zE = x.i == g.nx - 1 ? C(0) : (C)z[c + 1];
        const C zS = x.j == 0 ? C(0) : (C)z[c - g.nxp];
        const C zW = x.i == 0 ? C(0) : (C)z[c - 1];
        const C sum = ((aN[c] * zN + aE[c] * zE) + aS[c] * zS) + aW[c] * zW;
        return ap[c] * (C)z[c] - sum;
    }
};
using StencilOp = StencilOpT<double>;
using StencilOp32 = StencilOpT<float>;

// The reference's general cell list: acoef(4,ncellsD) and cell2cell(4,ncellsD)
// (0-based, ghosts numbered from ncells, -1 for no face).
constexpr int ELL_MAX_FACES = 8;   // nmaxfaces: 4 Cartesian, 6 telescoping (read_tel, geospatial.F90:433)
struct EllOp {
    int ncells, nf;      // cells, faces per cell (row stride of c2c and a)
    const int *c2c;
    const double *a;
    __device__ __forceinline__ int n() const { return ncells; }
    template <class F>
    __device__ __forceinline__ void each(F &&f) const {
        const Chunk ch = block_chunk(ncells);
        for (int t = ch.begin + threadIdx.x; t < ch.end; t += KBLOCK) f(at(t));
    }
    __device__ __forceinline__ Idx at(int t) const {
        Idx x;
        x.c = t;
        x.i = x.j = 0;
        return x;
    }
    __device__ __forceinline__ double apply(const double *xv, const double *dinv, const Idx &x) const {
        const int c = x.c;
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < ELL_MAX_FACES; k++) {
            if (k >= nf) break;
            const int nb = c2c[nf * c + k];
            if (nb >= 0 && nb < ncells) sum += a[nf * c + k] * (xv[nb] * dinv[nb]);
        }
        return xv[c] - sum;
    }
};

// Vectors of NS systems in precision T (double, or float for the inner solve
// of the FP32 form); scalars and reductions are FP64 in both.
template <int NS, class T = double>
struct KrySys {
    T *x[NS], *r[NS], *r0[NS], *p[NS], *v[NS], *t[NS];
    T *ap_w[NS];       // diagonal ap of each system (written by the prologue)
    T *dinv_w[NS];     // 1/ap (Jacobi form)
    float *zp, *zs;    // M^-1 p and M^-1 s of system 0 (explicit preconditioner only)
    KryScal *sc;       // [NS]
    KryRed<NS> red;
};

// Restarted GMRES(m) (pgmres, source/solvers/pgmres.f90:2-146: GMRES(3) with
// restarts) for one system with an explicit right preconditioner z = M^-1 v.
// Arnoldi by classical Gram-Schmidt with the dot products fused into the
// SpMV; Givens rotations and the back substitution in one workgroup; each
// cycle ends with x += M^-1 V y and the true residual r -= A M^-1 V y, whose
// norm decides the restart.
constexpr int GM_M = 20;
struct GmScal {
    KryScal ks;                     // ks.done stops the Arnoldi launches; thr, rmom, rnorm
    KryScal ks_final;               // never done: the cycle's closing V-cycle
    double H[(GM_M + 1) * GM_M];    // Hessenberg, column j at H[(GM_M + 1) * j], rotated in place
    double cs[GM_M], sn[GM_M], g[GM_M + 1], y[GM_M];
    double beta, hinv;
    int j;                          // Arnoldi steps of this cycle
    int iters;                      // steps over all cycles
    int cycle_end;
};
struct GmSys {
    double *V[GM_M + 1];
    double *x, *r;
    const double *ap;
    float *z;
    double *part;                   // (GM_M + 2) slots of nb partial sums
    GmScal *sc;
};

__device__ __forceinline__ bool lead() { return blockIdx.x == 0 && threadIdx.x == 0; }

// Called by the kernel that starts a solve (its block 0 only writes the slot
// no other block of that launch reads).
__device__ __forceinline__ void kry_reset(KryScal &sc) {
    if (lead()) { sc.done = 0; sc.final_iters = 0; }
}

// Totals of NV partial arrays of nb entries, summed in block order; every
// thread of every block obtains the same values.
template <int NV>
__device__ __forceinline__ void sum_partials(const double *const *src, int nb, double (&out)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; k++) out[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += KBLOCK)
#pragma unroll
        for (int k = 0; k < NV; k++) out[k] += src[k][b];
    block_sum<NV, KBLOCK>(out);
}

// Block totals of v written as this block's partials.
template <int NV>
__device__ __forceinline__ void put_partials(double (&v)[NV], double *const *dst) {
    block_sum<NV, KBLOCK>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) dst[k][blockIdx.x] = v[k];
}

}  // namespace cms
