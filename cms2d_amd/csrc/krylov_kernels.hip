// krylov_kernels.hip -- solve() prologue and Jacobi-preconditioned BiCGStab
// sweeps (restating solvers.f90:205-304 and pbicgstab.f90:21-96; krylov.h
// describes the recurrence and the atomic-free reductions).
//
// Iteration `it` is five launches on the fixed grid:
//   k_bicg_p(it)     finish iteration it-1 from its r.r and r0.r partials
//                    (stop test, restart, beta); p = r + beta (p - omega v)
//   k_bicg_spmv1(it) v = A M^-1 p, partials r0.v
//   k_bicg_s(it)     alpha = rho / r0.v;  x += alpha M^-1 p;  s = r - alpha v
//   k_bicg_spmv2(it) t = A M^-1 s, partials t.s and t.t
//   k_bicg_r(it)     omega = t.s / t.t;  x += omega M^-1 s;  r = s - omega t,
//                    partials r.r and r0.r
#include "flow_state.h"

namespace cms {

#ifndef SMALL
#define SMALL F4(1.0e-15)
#endif

template <int NS, class T>
__device__ __forceinline__ bool kdone(const KrySys<NS, T> &K) {
#pragma unroll
    for (int q = 0; q < NS; q++)
        if (!K.sc[q].done) return false;
    return true;
}

// solve() prologue for u (n=2) and v (n=3), relax(2)=relax(3)=0.8
// (solvers.f90:232-240), then the BiCGStab start (pbicgstab.f90:21-31).
__global__ void __launch_bounds__(KBLOCK) k_uv_prologue(Geo g, St s, KrySys<2> K) {
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    grid_walk(g, [&](const Idx &x_) {
        const int i = x_.i, j = x_.j, c = x_.c;
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
            K.dinv_w[q][c] = 1.0 / ap;
            K.r[q][c] = r; K.r0[q][c] = r;
            red[2 * q] += r * r;
            red[2 * q + 1] += res * res;
        }
    });
    double *const dst[4] = {K.red.pr[0], K.red.pr[0] + gridDim.x, K.red.pr[1], K.red.pr[1] + gridDim.x};
    put_partials<4>(red, dst);
    kry_reset(K.sc[0]);
    kry_reset(K.sc[1]);
}

// Finishes iteration it-1 (or starts the solve when it == 0) and, unless
// fin_only, updates p.  Launched with fin_only on one block at the end of a
// batch so the host can read the state; the next batch repeats the same
// finish (same inputs, same result) before its first p update.
template <class Op, int NS, class T = double>
__global__ void __launch_bounds__(KBLOCK) k_bicg_p(Op A, KrySys<NS, T> K, int it, int nb, int kmax, double rtol,
                                                  double atol, int fin_only) {
    double tot[2 * NS];
    const double *src[2 * NS];
#pragma unroll
    for (int q = 0; q < NS; q++) { src[2 * q] = K.red.pr[q]; src[2 * q + 1] = K.red.pr[q] + nb; }
    sum_partials<2 * NS>(src, nb, tot);
    bool act[NS];
    int rst[NS];
    double beta[NS], om[NS];
    bool any = false;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        KryScal &sc = K.sc[q];
        act[q] = false; rst[q] = 0; beta[q] = 0.0; om[q] = 1.0;
        if (sc.done) continue;
        const double rr = tot[2 * q], nr = sqrt(rr);
        if (it == 0) {
            const double thr = fmax(rtol * nr, atol * sqrt((double)A.n()));
            act[q] = !(nr <= thr);
            if (lead()) {
                sc.rmom = sqrt(tot[2 * q + 1] / (double)(float)A.n());   // sqrt(adak/real(ncells))
                sc.thr = thr;
                sc.rho[0] = rr;
                sc.r0n[0] = nr;
                sc.iters[0] = 0;
                sc.rnorm = nr;
                if (!act[q]) { sc.done = 1; sc.final_iters = 0; }
            }
        } else {
            const int pq = (it - 1) & 1, cq = it & 1;
            const double r0r = tot[2 * q + 1];
            const int iters = sc.iters[pq] + 1;
            const double rho_p = sc.rho[pq], r0n_p = sc.r0n[pq], al = sc.alpha[pq], omp = sc.omega[pq];
            double rho = rho_p, r0n = r0n_p;
            if (omp == 0.0 || nr <= sc.thr || iters >= kmax) {
                act[q] = false;
            } else if (fabs(r0r) <= 1e-12 * nr * r0n_p) {
                rst[q] = 1; rho = rr; r0n = nr; beta[q] = 0.0; act[q] = true;
            } else {
                beta[q] = (r0r / rho_p) * (al / omp); rho = r0r; act[q] = true;
            }
            om[q] = omp;
            if (lead()) {
                sc.rnorm = nr;
                sc.iters[cq] = iters;
                if (act[q]) { sc.rho[cq] = rho; sc.r0n[cq] = r0n; }
                else { sc.done = 1; sc.final_iters = iters; }
            }
        }
        any = any || act[q];
    }
    if (fin_only || !any) return;
    A.each([&](const Idx &x_) {
        const int c = x_.c;
#pragma unroll
        for (int q = 0; q < NS; q++) {
            if (!act[q]) continue;
            const T r = K.r[q][c];
            if (rst[q]) K.r0[q][c] = r;
            // the first iteration starts from p = r (p and v of an earlier solve are not read)
            K.p[q][c] = it == 0 ? r : r + (T)beta[q] * (K.p[q][c] - (T)om[q] * K.v[q][c]);
        }
    });
}

// MODE 1: v = A M^-1 p, partials r0.v.  MODE 2: t = A M^-1 s (s held in r),
// partials t.s and t.t.  PREC 1: M^-1 p and M^-1 s are the explicit zp, zs.
template <class Op, int NS, int MODE, int PREC = 0, class T = double>
__global__ void __launch_bounds__(KBLOCK) k_bicg_spmv(Op A, KrySys<NS, T> K) {
    static_assert(PREC == 0 || NS == 1, "explicit preconditioner: one system");
    if (kdone(K)) return;
    constexpr int NV = MODE == 1 ? NS : 2 * NS;
    double red[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) red[q] = 0.0;
    bool act[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) act[q] = !K.sc[q].done;
    A.each([&](const Idx &x) {
        const int c = x.c;
#pragma unroll
        for (int q = 0; q < NS; q++) {
            if (!act[q]) continue;
            const T *xv = MODE == 1 ? K.p[q] : K.r[q];
            T y;
            if constexpr (PREC == 1) y = A.apply_z(MODE == 1 ? K.zp : K.zs, K.ap_w[q], x);
            else y = A.apply(xv, K.dinv_w[q], x);
            const double yd = y;
            if (MODE == 1) {
                K.v[q][c] = y;
                red[q] += (double)K.r0[q][c] * yd;
            } else {
                K.t[q][c] = y;
                red[2 * q] += yd * (double)xv[c];
                red[2 * q + 1] += yd * yd;
            }
        }
    });
    double *dst[NV];
#pragma unroll
    for (int q = 0; q < NS; q++) {
        if (MODE == 1) dst[q] = K.red.pv[q];
        else { dst[2 * q] = K.red.pt[q]; dst[2 * q + 1] = K.red.pt[q] + gridDim.x; }
    }
    put_partials<NV>(red, dst);
}

// alpha = rho / r0.v ; x += alpha M^-1 p ; s = r - alpha v (s overwrites r).
// PREC 1 defers the x update to k_bicg_r.
template <class Op, int NS, int PREC = 0, class T = double>
__global__ void __launch_bounds__(KBLOCK) k_bicg_s(Op A, KrySys<NS, T> K, int it, int nb) {
    if (kdone(K)) return;
    double tot[NS];
    const double *src[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) src[q] = K.red.pv[q];
    sum_partials<NS>(src, nb, tot);
    bool act[NS];
    double al[NS];
    bool any = false;
    const int cq = it & 1;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        KryScal &sc = K.sc[q];
        act[q] = !sc.done;
        al[q] = 0.0;
        if (!act[q]) continue;
        if (tot[q] == 0.0) {                       // r0.v = 0: stop before the update
            act[q] = false;
            if (lead()) { sc.done = 1; sc.final_iters = sc.iters[cq]; }
            continue;
        }
        al[q] = sc.rho[cq] / tot[q];
        if (lead()) sc.alpha[cq] = al[q];
        any = true;
    }
    if (!any) return;
    A.each([&](const Idx &x_) {
        const int c = x_.c;
#pragma unroll
        for (int q = 0; q < NS; q++) {
            if (!act[q]) continue;
            if constexpr (PREC == 0) K.x[q][c] = K.x[q][c] + al[q] * (K.p[q][c] * K.dinv_w[q][c]);
            K.r[q][c] = K.r[q][c] - (T)al[q] * K.v[q][c];
        }
    });
}

// omega = t.s / t.t ; x += omega M^-1 s ; r = s - omega t ; partials r.r, r0.r
// PREC 1: x += alpha zp + omega zs.
template <class Op, int NS, int PREC = 0, class T = double>
__global__ void __launch_bounds__(KBLOCK) k_bicg_r(Op A, KrySys<NS, T> K, int it, int nb) {
    if (kdone(K)) return;
    double tot[2 * NS];
    const double *src[2 * NS];
#pragma unroll
    for (int q = 0; q < NS; q++) { src[2 * q] = K.red.pt[q]; src[2 * q + 1] = K.red.pt[q] + nb; }
    sum_partials<2 * NS>(src, nb, tot);
    bool act[NS];
    double om[NS], al[NS];
    const int cq = it & 1;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        act[q] = !K.sc[q].done;
        const double ts = tot[2 * q], tt = tot[2 * q + 1];
        om[q] = tt > 0.0 ? ts / tt : 0.0;
        al[q] = K.sc[q].alpha[cq];
        if (act[q] && lead()) K.sc[q].omega[cq] = om[q];
    }
    double red[2 * NS];
#pragma unroll
    for (int q = 0; q < 2 * NS; q++) red[q] = 0.0;
    A.each([&](const Idx &x_) {
        const int c = x_.c;
#pragma unroll
        for (int q = 0; q < NS; q++) {
            if (!act[q]) continue;
            const T sv = K.r[q][c];
            if constexpr (PREC == 1)
                K.x[q][c] = K.x[q][c] + ((T)al[q] * (T)K.zp[c] + (T)om[q] * (T)K.zs[c]);
            else
                K.x[q][c] = K.x[q][c] + om[q] * (sv * K.dinv_w[q][c]);
            const T r = sv - (T)om[q] * K.t[q][c];
            K.r[q][c] = r;
            red[2 * q] += (double)r * (double)r;
            red[2 * q + 1] += (double)K.r0[q][c] * (double)r;
        }
    });
    double *dst[2 * NS];
#pragma unroll
    for (int q = 0; q < NS; q++) { dst[2 * q] = K.red.pr[q]; dst[2 * q + 1] = K.red.pr[q] + gridDim.x; }
    put_partials<2 * NS>(red, dst);
}

// solve() prologue on the reference's general cell list (solvers.f90:232-240)
// and the BiCGStab start, for the drop-in cms2d_solve.
// Clears n 32-bit words with vector stores (used instead of a memset node).
__global__ void k_zero_words(unsigned int *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// bias: the reference's small number added to the starting residual
// (pbicgstab.f90), 0 for the transposed solve of the adjoint, which must stay
// linear in its right-hand side.
__global__ void __launch_bounds__(KBLOCK) k_ell_prologue(EllOp A, const int *iwet, double *ss, const double *sp,
                                                        double *res, const double *phi, double relax,
                                                        KrySys<1> K, double bias) {
    double red[2] = {0.0, 0.0};
    const Chunk ch = block_chunk(A.ncells);
    const double rinv = 1.0 / relax, relaxm = 1.0 - relax;
    for (int c = ch.begin + threadIdx.x; c < ch.end; c += KBLOCK) {
        double asum = 0.0, anb = 0.0;
#pragma unroll
        for (int k = 0; k < ELL_MAX_FACES; k++) {
            if (k >= A.nf) break;
            const int nbc = A.c2c[A.nf * c + k];
            if (nbc < 0) continue;
            asum += A.a[A.nf * c + k];
            anb += A.a[A.nf * c + k] * phi[nbc];
        }
        double ap = (asum - sp[c]) * rinv;
        double ph = phi[c];
        double ssv = ss[c] + relaxm * ph * ap;
        ss[c] = ssv;
        double rs = iwet[c] * (anb + ssv - ap * ph) / ap;
        res[c] = rs;
        double r = anb + ssv - ap * ph + bias;
        K.ap_w[0][c] = ap;
        K.dinv_w[0][c] = 1.0 / ap;
        K.r[0][c] = r; K.r0[0][c] = r;
        red[0] += r * r;
        red[1] += rs * rs;
    }
    double *const dst[2] = {K.red.pr[0], K.red.pr[0] + gridDim.x};
    put_partials<2>(red, dst);
    kry_reset(K.sc[0]);
}

}  // namespace cms

namespace cms {

// --- FP32 Krylov inside FP64 iterative refinement ----------------------------
// Outer sweep: d = inner FP32 BiCGStab solve of A d = r (FP32 matrix and
// vectors, multigrid preconditioner); x += d and r -= A d in FP64; stop when
// the FP64 ||r|| meets the FP64 solve's own threshold.

// Inner start from the FP64 residual: r = r0 = (float) r64, x = 0 (p starts from r).
template <class Op>
__global__ void __launch_bounds__(KBLOCK) k_ref_start(Op A, KrySys<1, float> Ki, const double *r64) {
    double red[2] = {0.0, 0.0};
    A.each([&](const Idx &x_) {
        const int c = x_.c;
        const float r = (float)r64[c];
        Ki.r[0][c] = r; Ki.r0[0][c] = r; Ki.x[0][c] = 0.0f;
        red[0] += (double)r * (double)r;
    });
    red[1] = red[0];
    double *const dst[2] = {Ki.red.pr[0], Ki.red.pr[0] + gridDim.x};
    put_partials<2>(red, dst);
    kry_reset(Ki.sc[0]);
}

// x += d, r -= A d (FP64 operator on the FP32 correction), partials ||r||^2.
template <class Op>
__global__ void __launch_bounds__(KBLOCK) k_ref_update(Op A, KrySys<1, double> K, const float *d) {
    double red[2] = {0.0, 0.0};
    A.each([&](const Idx &x) {
        const int c = x.c;
        K.x[0][c] = K.x[0][c] + (double)d[c];
        const double r = K.r[0][c] - A.apply_z(d, K.ap_w[0], x);
        K.r[0][c] = r;
        red[0] += r * r;
    });
    double *const dst[2] = {K.red.pr[0], K.red.pr[0] + gridDim.x};
    put_partials<2>(red, dst);
}

// The u and v solves of the FP32 form (two systems, Jacobi): FP32 copies of
// the shared coefficients and of each system's diagonal.
__global__ void __launch_bounds__(KBLOCK) k_uv_to32(StencilOp A, KrySys<2, double> K, float *aN, float *aE, float *aS,
                                                   float *aW, KrySys<2, float> Ki) {
    A.each([&](const Idx &x) {
        const int c = x.c;
        aN[c] = (float)A.aN[c]; aE[c] = (float)A.aE[c]; aS[c] = (float)A.aS[c]; aW[c] = (float)A.aW[c];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            Ki.ap_w[q][c] = (float)K.ap_w[q][c];
            Ki.dinv_w[q][c] = (float)K.dinv_w[q][c];
        }
    });
}
// Start of an inner FP32 sweep of the two systems: r = FP64 residual (0 for
// a system whose FP64 solve has stopped), x = 0.
template <class Op>
__global__ void __launch_bounds__(KBLOCK) k_ref_start2(Op A, KrySys<2, float> Ki, KrySys<2, double> K) {
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    bool live[2];
#pragma unroll
    for (int q = 0; q < 2; q++) live[q] = !K.sc[q].done;
    A.each([&](const Idx &x_) {
        const int c = x_.c;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const float r = live[q] ? (float)K.r[q][c] : 0.0f;
            Ki.r[q][c] = r; Ki.r0[q][c] = r; Ki.x[q][c] = 0.0f;
            red[2 * q] += (double)r * (double)r;
        }
    });
    red[1] = red[0];
    red[3] = red[2];
    double *const dst[4] = {Ki.red.pr[0], Ki.red.pr[0] + gridDim.x, Ki.red.pr[1], Ki.red.pr[1] + gridDim.x};
    put_partials<4>(red, dst);
    kry_reset(Ki.sc[0]);
    kry_reset(Ki.sc[1]);
}
// x += d, r -= A d for each running system (FP64 operator on the FP32
// correction), partials ||r||^2.
template <class Op>
__global__ void __launch_bounds__(KBLOCK) k_ref_update2(Op A, KrySys<2, double> K, KrySys<2, float> Ki) {
    double red[2] = {0.0, 0.0};
    bool live[2];
#pragma unroll
    for (int q = 0; q < 2; q++) live[q] = !K.sc[q].done;
    A.each([&](const Idx &x) {
        const int c = x.c;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (!live[q]) continue;
            const float *d = Ki.x[q];
            K.x[q][c] = K.x[q][c] + (double)d[c];
            const double r = K.r[q][c] - A.apply_z(d, K.ap_w[q], x);
            K.r[q][c] = r;
            red[q] += r * r;
        }
    });
    double *const dst[2] = {K.red.pr[0], K.red.pr[1]};
    put_partials<2>(red, dst);
}
// One block: the FP64 stop tests of the two systems after a sweep.
__global__ void __launch_bounds__(KBLOCK) k_ref_check2(KrySys<2, double> K, int nb) {
    double tot[2];
    const double *src[2] = {K.red.pr[0], K.red.pr[1]};
    sum_partials<2>(src, nb, tot);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (K.sc[q].done) continue;
            const double nr = sqrt(tot[q]);
            K.sc[q].rnorm = nr;
            if (nr <= K.sc[q].thr) K.sc[q].done = 1;
        }
}

// One block: the FP64 stop test of the outer sweep.
__global__ void __launch_bounds__(KBLOCK) k_ref_check(KrySys<1, double> K, int nb) {
    double tot[1];
    const double *src[1] = {K.red.pr[0]};
    sum_partials<1>(src, nb, tot);
    if (threadIdx.x == 0) {
        const double nr = sqrt(tot[0]);
        K.sc[0].rnorm = nr;
        if (nr <= K.sc[0].thr) K.sc[0].done = 1;
    }
}

}  // namespace cms

namespace cms {

// --- GMRES(m) -----------------------------------------------------------------

// Start of a cycle (one block): beta = ||r|| from partials pr0 (pr1: res.res
// of the solve prologue on the first cycle, for rmom).
__global__ void __launch_bounds__(KBLOCK) k_gm_begin(GmSys G, const double *pr0, const double *pr1, int nb,
                                                     int first, int n, double rtol, double atol) {
    double tot[2];
    const double *src[2] = {pr0, pr1 ? pr1 : pr0};
    sum_partials<2>(src, nb, tot);
    if (threadIdx.x != 0) return;
    GmScal &s = *G.sc;
    const double beta = sqrt(tot[0]);
    if (first) {
        s.ks.thr = fmax(rtol * beta, atol * sqrt((double)n));
        s.ks.rmom = sqrt(tot[1] / (double)(float)n);
        s.iters = 0;
    }
    s.beta = beta;
    s.g[0] = beta;
    s.j = 0;
    s.cycle_end = 0;
    s.ks.rnorm = beta;
    s.ks.done = beta <= s.ks.thr ? 1 : 0;
    s.ks.final_iters = s.iters;
    s.ks_final.done = 0;
}

template <class Op>
__global__ void __launch_bounds__(KBLOCK) k_gm_v0(Op A, GmSys G) {
    if (G.sc->ks.done) return;
    const double bi = 1.0 / G.sc->beta;
    A.each([&](const Idx &x_) {
        const int c = x_.c;
        G.V[0][c] = G.r[c] * bi;
    });
}

// w = A z into V[j+1]; partial sums of (w, V[i]), i <= j, in slot i.
template <class Op>
__global__ void __launch_bounds__(KBLOCK) k_gm_spmv_dot(Op A, GmSys G, int j) {
    if (G.sc->ks.done) return;
    double acc[GM_M + 1];
#pragma unroll
    for (int i = 0; i <= GM_M; i++) acc[i] = 0.0;
    A.each([&](const Idx &x) {
        const double w = A.apply_z(G.z, G.ap, x);
        G.V[j + 1][x.c] = w;
#pragma unroll
        for (int i = 0; i <= GM_M; i++)
            if (i <= j) acc[i] += w * G.V[i][x.c];
    });
#pragma unroll
    for (int i = 0; i <= GM_M; i++) {
        if (i <= j) {                   // j is uniform: every thread takes the same branches
            double v[1] = {acc[i]};
            double *const dst[1] = {G.part + (size_t)i * gridDim.x};
            put_partials<1>(v, dst);
        }
    }
}

// One block: column j of H from the partials.
__global__ void __launch_bounds__(KBLOCK) k_gm_hred(GmSys G, int j, int nb) {
    if (G.sc->ks.done) return;
    for (int i = 0; i <= j; i++) {
        double tot[1];
        const double *src[1] = {G.part + (size_t)i * nb};
        sum_partials<1>(src, nb, tot);
        if (threadIdx.x == 0) G.sc->H[(GM_M + 1) * j + i] = tot[0];
    }
}

// w -= sum_i H[i,j] V[i]; partial ||w||^2 in slot GM_M + 1.
template <class Op>
__global__ void __launch_bounds__(KBLOCK) k_gm_orth(Op A, GmSys G, int j) {
    if (G.sc->ks.done) return;
    double h[GM_M + 1];
#pragma unroll
    for (int i = 0; i <= GM_M; i++) h[i] = i <= j ? G.sc->H[(GM_M + 1) * j + i] : 0.0;
    double acc[1] = {0.0};
    A.each([&](const Idx &x_) {
        const int c = x_.c;
        double w = G.V[j + 1][c];
#pragma unroll
        for (int i = 0; i <= GM_M; i++)
            if (i <= j) w -= h[i] * G.V[i][c];
        G.V[j + 1][c] = w;
        acc[0] += w * w;
    });
    double *const dst[1] = {G.part + (size_t)(GM_M + 1) * gridDim.x};
    put_partials<1>(acc, dst);
}

// One block: H[j+1,j] = ||w||, Givens rotations, residual estimate, stop test.
__global__ void __launch_bounds__(KBLOCK) k_gm_finish(GmSys G, int j, int nb) {
    if (G.sc->ks.done) return;
    double tot[1];
    const double *src[1] = {G.part + (size_t)(GM_M + 1) * nb};
    sum_partials<1>(src, nb, tot);
    if (threadIdx.x != 0) return;
    GmScal &s = *G.sc;
    double *Hc = s.H + (GM_M + 1) * j;
    const double nw = sqrt(tot[0]);
    Hc[j + 1] = nw;
    for (int i = 0; i < j; i++) {
        const double a = Hc[i], b = Hc[i + 1];
        Hc[i] = s.cs[i] * a + s.sn[i] * b;
        Hc[i + 1] = -s.sn[i] * a + s.cs[i] * b;
    }
    const double a = Hc[j], b = Hc[j + 1], rho = hypot(a, b);
    const double c = rho > 0.0 ? a / rho : 1.0, sn = rho > 0.0 ? b / rho : 0.0;
    Hc[j] = rho;
    Hc[j + 1] = 0.0;
    s.cs[j] = c;
    s.sn[j] = sn;
    s.g[j + 1] = -sn * s.g[j];
    s.g[j] = c * s.g[j];
    s.iters += 1;
    s.j = j + 1;
    s.hinv = nw > 0.0 ? 1.0 / nw : 0.0;
    const double res = fabs(s.g[j + 1]);
    s.ks.rnorm = res;
    if (res <= s.ks.thr || j + 1 == GM_M || nw == 0.0) {
        s.ks.done = 1;
        s.cycle_end = 1;
        s.ks.final_iters = s.iters;
    }
}

template <class Op>
__global__ void __launch_bounds__(KBLOCK) k_gm_scale(Op A, GmSys G, int j) {
    if (G.sc->ks.done) return;
    const double hi = G.sc->hinv;
    A.each([&](const Idx &x_) {
        const int c = x_.c;
        G.V[j + 1][c] *= hi;
    });
}

// One block: y = R^-1 g over the cycle's k steps.
__global__ void k_gm_solve(GmSys G) {
    if (threadIdx.x != 0) return;
    GmScal &s = *G.sc;
    const int k = s.j;
    for (int i = k - 1; i >= 0; i--) {
        double v = s.g[i];
        for (int l = i + 1; l < k; l++) v -= s.H[(GM_M + 1) * l + i] * s.y[l];
        const double d = s.H[(GM_M + 1) * i + i];
        s.y[i] = d != 0.0 ? v / d : 0.0;
    }
}

// u = sum_{i<k} y_i V[i] into V[k].
template <class Op>
__global__ void __launch_bounds__(KBLOCK) k_gm_combine(Op A, GmSys G, int k) {
    double y[GM_M];
#pragma unroll
    for (int i = 0; i < GM_M; i++) y[i] = i < k ? G.sc->y[i] : 0.0;
    A.each([&](const Idx &x_) {
        const int c = x_.c;
        double u = 0.0;
#pragma unroll
        for (int i = 0; i < GM_M; i++)
            if (i < k) u += y[i] * G.V[i][c];
        G.V[k][c] = u;
    });
}

// x += z, r -= A z (z = M^-1 V y); partials r.r in slot 0.
template <class Op>
__global__ void __launch_bounds__(KBLOCK) k_gm_update(Op A, GmSys G) {
    double acc[1] = {0.0};
    A.each([&](const Idx &x) {
        const int c = x.c;
        G.x[c] = G.x[c] + (double)G.z[c];
        const double r = G.r[c] - A.apply_z(G.z, G.ap, x);
        G.r[c] = r;
        acc[0] += r * r;
    });
    double *const dst[1] = {G.part};
    put_partials<1>(acc, dst);
}

}  // namespace cms
