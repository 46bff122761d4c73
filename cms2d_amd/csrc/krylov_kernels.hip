// krylov_kernels.hip -- solve() prologue and Jacobi-preconditioned BiCGStab
// sweeps on the structured 5-point operator (restating solvers.f90:205-304
// and pbicgstab.f90:21-96; see krylov.h for the recurrence).
#include "flow_state.h"

namespace cms {

#ifndef SMALL
#define SMALL F4(1.0e-15)
#endif

template <int NS>
__device__ __forceinline__ bool kdone(const KrySys<NS> &K) {
#pragma unroll
    for (int q = 0; q < NS; q++)
        if (!K.sc[q].done) return false;
    return true;
}

// solve() prologue for u (n=2) and v (n=3), relax(2)=relax(3)=0.8
// (solvers.f90:232-240), then the BiCGStab start (pbicgstab.f90:21-31).
__global__ void __launch_bounds__(BLOCK) k_uv_prologue(Geo g, St s, Par P, KrySys<2> K) {
    int i = 0, j = 0, c = 0;
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    if (act_thread(g, i, j, c)) {
        const double a[4] = {s.aN[c], s.aE[c], s.aS[c], s.aW[c]};
        const int nb[4] = {c + g.nxp, c + 1, c - g.nxp, c - 1};
        const double asum = ((a[0] + a[1]) + a[2]) + a[3];
        const double rinv = 1.0 / F4(0.8), relaxm = 1.0 - F4(0.8);
        const int w = s.iwet[c];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            double *phi = q ? s.v : s.u;
            double *ss = q ? s.sv : s.su;
            const double *sp = q ? s.spv : s.spu;
            double ap = (asum - sp[c]) * rinv;
            double ph = phi[c];
            double ssv = ss[c] + relaxm * ph * ap;
            ss[c] = ssv;
            double anb = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) anb += a[k] * phi[nb[k]];
            double res = w * (anb + ssv - ap * ph) / ap;
            double r = anb + ssv - ap * ph + SMALL;
            (q ? s.resv : s.resu)[c] = res;
            K.ap_w[q][c] = ap;
            K.r[q][c] = r; K.r0[q][c] = r; K.p[q][c] = 0.0; K.v[q][c] = 0.0;
            red[2 * q] = res * res;
            red[2 * q + 1] = r * r;
        }
    }
    if (grid_sum<4>(red, s.partials, s.counter) && threadIdx.x == 0) {
        kry_start(K.sc[0], red[0], red[1], g.nact, P);
        kry_start(K.sc[1], red[2], red[3], g.nact, P);
    }
}

// p = r + beta (p - omega v); on a restart the shadow residual becomes r.
template <class Op, int NS>
__global__ void __launch_bounds__(BLOCK) k_bicg_p(Op A, KrySys<NS> K) {
    if (kdone(K)) return;
    Idx x;
    if (!A.cell(x)) return;
    const int c = x.c;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        if (K.sc[q].done) continue;
        const double beta = K.sc[q].beta, om = K.sc[q].omega;
        const double r = K.r[q][c];
        if (K.sc[q].restart) K.r0[q][c] = r;
        K.p[q][c] = r + beta * (K.p[q][c] - om * K.v[q][c]);
    }
}

// y = A M^-1 x with M = diag(ap), z = 0 on ghost cells.
// MODE 1: x = p, y -> v, reduce r0.v, then alpha = rho / (r0.v).
// MODE 2: x = s (held in r), y -> t, reduce t.s and t.t, then omega = (t.s)/(t.t).
template <class Op, int NS, int MODE>
__global__ void __launch_bounds__(BLOCK) k_bicg_spmv(Op A, KrySys<NS> K, double *partials, unsigned *counter) {
    if (kdone(K)) return;
    constexpr int NV = MODE == 1 ? NS : 2 * NS;
    double red[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) red[q] = 0.0;
    Idx x;
    if (A.cell(x)) {
        const int c = x.c;
#pragma unroll
        for (int q = 0; q < NS; q++) {
            if (K.sc[q].done) continue;
            const double *xv = MODE == 1 ? K.p[q] : K.r[q];
            const double y = A.apply(xv, K.ap_w[q], x);
            if (MODE == 1) {
                K.v[q][c] = y;
                red[q] = K.r0[q][c] * y;
            } else {
                K.t[q][c] = y;
                red[2 * q] = y * xv[c];
                red[2 * q + 1] = y * y;
            }
        }
    }
    if (grid_sum<NV>(red, partials, counter) && threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NS; q++) {
            KryScal &sc = K.sc[q];
            if (sc.done) continue;
            if (MODE == 1) {
                if (red[q] == 0.0) sc.done = 1;
                else sc.alpha = sc.rho / red[q];
            } else {
                double ts = red[2 * q], tt = red[2 * q + 1];
                sc.omega = tt > 0.0 ? ts / tt : 0.0;
            }
        }
    }
}

// x += alpha M^-1 p ; s = r - alpha v (s overwrites r)
template <class Op, int NS>
__global__ void __launch_bounds__(BLOCK) k_bicg_s(Op A, KrySys<NS> K) {
    if (kdone(K)) return;
    Idx x;
    if (!A.cell(x)) return;
    const int c = x.c;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        if (K.sc[q].done) continue;
        const double al = K.sc[q].alpha;
        K.x[q][c] = K.x[q][c] + al * (K.p[q][c] / K.ap_w[q][c]);
        K.r[q][c] = K.r[q][c] - al * K.v[q][c];
    }
}

// x += omega M^-1 s ; r = s - omega t ; convergence, restart and next beta
template <class Op, int NS>
__global__ void __launch_bounds__(BLOCK) k_bicg_r(Op A, KrySys<NS> K, double *partials, unsigned *counter,
                                                  int kmax) {
    if (kdone(K)) return;
    double red[2 * NS];
#pragma unroll
    for (int q = 0; q < 2 * NS; q++) red[q] = 0.0;
    Idx x;
    if (A.cell(x)) {
        const int c = x.c;
#pragma unroll
        for (int q = 0; q < NS; q++) {
            if (K.sc[q].done) continue;
            const double om = K.sc[q].omega;
            const double sv = K.r[q][c];
            K.x[q][c] = K.x[q][c] + om * (sv / K.ap_w[q][c]);
            const double r = sv - om * K.t[q][c];
            K.r[q][c] = r;
            red[2 * q] = r * r;
            red[2 * q + 1] = K.r0[q][c] * r;
        }
    }
    if (grid_sum<2 * NS>(red, partials, counter) && threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NS; q++) {
            KryScal &sc = K.sc[q];
            if (sc.done) continue;
            sc.iters += 1;
            sc.first = 0;
            const double nr = sqrt(red[2 * q]);
            sc.rnorm = nr;
            const double r0r = red[2 * q + 1];
            if (sc.omega == 0.0 || nr <= sc.thr || sc.iters >= kmax) {
                sc.done = 1;
            } else if (fabs(r0r) <= 1e-12 * nr * sc.r0norm) {
                sc.restart = 1;
                sc.rho = nr * nr;
                sc.r0norm = nr;
                sc.beta = 0.0;
            } else {
                sc.restart = 0;
                sc.beta = (r0r / sc.rho) * (sc.alpha / sc.omega);
                sc.rho = r0r;
            }
        }
    }
}

// solve() prologue on the reference's general cell list (solvers.f90:232-240)
// and the BiCGStab start, for the drop-in cms2d_solve.
__global__ void __launch_bounds__(BLOCK) k_ell_prologue(EllOp A, const int *iwet, double *ss, const double *sp,
                                                        double *res, const double *phi, double relax, Par P,
                                                        KrySys<1> K, double *partials, unsigned *counter) {
    double red[2] = {0.0, 0.0};
    Idx x;
    if (A.cell(x)) {
        const int c = x.c;
        const double rinv = 1.0 / relax, relaxm = 1.0 - relax;
        double asum = 0.0, anb = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int nb = A.c2c[4 * c + k];
            if (nb < 0) continue;
            asum += A.a[4 * c + k];
            anb += A.a[4 * c + k] * phi[nb];
        }
        double ap = (asum - sp[c]) * rinv;
        double ph = phi[c];
        double ssv = ss[c] + relaxm * ph * ap;
        ss[c] = ssv;
        double rs = iwet[c] * (anb + ssv - ap * ph) / ap;
        res[c] = rs;
        double r = anb + ssv - ap * ph + SMALL;
        K.ap_w[0][c] = ap;
        K.r[0][c] = r; K.r0[0][c] = r; K.p[0][c] = 0.0; K.v[0][c] = 0.0;
        red[0] = rs * rs;
        red[1] = r * r;
    }
    if (grid_sum<2>(red, partials, counter) && threadIdx.x == 0) kry_start(K.sc[0], red[0], red[1], A.ncells, P);
}

}  // namespace cms
