// krylov.h -- right-preconditioned BiCGStab on the 5-point cell stencil.
//
// Restates the iteration of pbicgstab (source/solvers/pbicgstab.f90:21-96):
//   p = r + beta*(p - omega*v);  v = A M^-1 p;  alpha = rho/(r0.v)
//   x += alpha M^-1 p;  s = r - alpha v;  t = A M^-1 s;  omega = (t.s)/(t.t)
//   x += omega M^-1 s;  r = s - omega t
// with M = diag(ap) (Jacobi) and the standard beta = (rho/rho_old)(alpha/omega);
// the reference's beta denominator carries +1e-15 (pbicgstab.f90:43), which
// stalls the recurrence once rho < ~1e-13, so the tight solve restarts the
// shadow residual on rho breakdown instead.  Up to NS systems that share the
// off-diagonal coefficients (u and v momentum) advance together so each
// coefficient is read once per sweep.
#pragma once
#include "common.h"

namespace cms {

struct KryScal {
    double rho;        // r0.r of the current residual
    double alpha, omega, beta;
    double r0norm;     // ||r|| at the start (or at the last restart)
    double thr;        // stopping threshold on ||r||2
    double rnorm;
    double rmom;       // normalized residual RMS of the solve prologue
    double red[4];     // reduction results
    int done, iters, restart, first;
};

// Matrix operators: y_i = ap_i z_i - sum_k a_k(i) z_nb(k), z = M^-1 x = x/ap,
// z = 0 on ghost cells.  cell() maps a thread to its row.
struct Idx {
    int c, i, j;
};

// Structured 5-point stencil on the padded grid (directed coefficients).
struct StencilOp {
    Geo g;
    const double *aN, *aE, *aS, *aW;
    __device__ __forceinline__ bool cell(Idx &x) const { return act_thread(g, x.i, x.j, x.c); }
    __device__ __forceinline__ int n() const { return g.nact; }
    __device__ __forceinline__ double apply(const double *xv, const double *ap, const Idx &x) const {
        const int c = x.c;
        const double zc = xv[c] / ap[c];
        const int cN = c + g.nxp, cE = c + 1, cS = c - g.nxp, cW = c - 1;
        const double zN = x.j == g.ny - 1 ? 0.0 : xv[cN] / ap[cN];
        const double zE = x.i == g.nx - 1 ? 0.0 : xv[cE] / ap[cE];
        const double zS = x.j == 0 ? 0.0 : xv[cS] / ap[cS];
        const double zW = x.i == 0 ? 0.0 : xv[cW] / ap[cW];
        const double sum = ((aN[c] * zN + aE[c] * zE) + aS[c] * zS) + aW[c] * zW;
        return ap[c] * zc - sum;
    }
};

// The reference's general cell list: acoef(4,ncellsD) and cell2cell(4,ncellsD)
// (0-based, ghosts numbered from ncells, -1 for no face).
struct EllOp {
    int ncells;
    const int *c2c;
    const double *a;
    __device__ __forceinline__ bool cell(Idx &x) const {
        x.c = blockIdx.x * blockDim.x + threadIdx.x;
        x.i = x.j = 0;
        return x.c < ncells;
    }
    __device__ __forceinline__ int n() const { return ncells; }
    __device__ __forceinline__ double apply(const double *xv, const double *ap, const Idx &x) const {
        const int c = x.c;
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int nb = c2c[4 * c + k];
            if (nb >= 0 && nb < ncells) sum += a[4 * c + k] * (xv[nb] / ap[nb]);
        }
        return ap[c] * (xv[c] / ap[c]) - sum;
    }
};

template <int NS>
struct KrySys {
    double *x[NS], *r[NS], *r0[NS], *p[NS], *v[NS], *t[NS];
    double *ap_w[NS];  // diagonal ap of each system (written by the prologue)
    KryScal *sc;       // [NS]
};

}  // namespace cms
