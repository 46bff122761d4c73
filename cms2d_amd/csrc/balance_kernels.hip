// balance_kernels.hip -- the domain volume balance of a time step
// (flow_balance, source/flow/flow_imp.F90:2260-2372, two-level time scheme,
// no rain or evaporation) and the step's extreme water level, velocities and
// Courant number over the wet cells (flow_step_stat,
// source/flow/flow.F90:1912-2008), evaluated on the device at the end of
// every step.  Each block of a fixed grid writes its partial sums and maxima;
// one block combines them in a fixed order and updates the running totals,
// so the results do not depend on scheduling.  Compiled after
// flow_kernels.hip and tel_kernels.hip (fdir, TelG, TelS, m_ncf).
#include "flow_state.h"

namespace cms {

// Running totals and the last step's values (device memory, part of the
// checkpointed state).
struct BalAcc {
    double stg, bnd, flux, bal, net, err;                // this step [m^3]; err [%]
    double cstg, cbnd, cflux, cbal, cnet, cerr, cnorm;   // cumulative
    double xv[4];     // extremes: p (largest |p|), u, v (largest |.|), Courant number
    double xc[4];     // their cells (0-based, exact in a double)
    double steps;
};
// per block: 2 sums, then (key, value, cell) of the 4 extremes
constexpr int BAL_PART = 2 + 12;

// flow_step_stat keeps the first cell whose key is strictly larger than the
// best so far, starting from key 0 at cell 0: the largest key wins and ties go
// to the smaller cell, so a combination in any order gives the serial result.
struct BalMax {
    double k, v, c;
};
__device__ __forceinline__ void bal_take(BalMax &a, double k, double v, double c) {
    if (k > a.k || (k == a.k && c < a.c)) { a.k = k; a.v = v; a.c = c; }
}
__device__ __forceinline__ void bal_init(double (&sum)[2], BalMax (&m)[4]) {
    sum[0] = sum[1] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) m[q] = BalMax{0.0, 0.0, 0.0};
}
// flow_step_stat's tests for one wet cell: |p|, |u|, |v| and the Courant
// number dt max((|u| + c)/dx, (|v| + c)/dy) with c = sqrt(g h).
__device__ __forceinline__ void bal_cell(BalMax (&m)[4], double p, double u, double v, double h, double dx,
                                         double dy, double dt, double cell) {
    bal_take(m[0], fabs(p), p, cell);
    bal_take(m[1], fabs(u), u, cell);
    bal_take(m[2], fabs(v), v, cell);
    const double cel = sqrt(GRAV * h);
    const double cr = dt * fmax((fabs(u) + cel) / dx, (fabs(v) + cel) / dy);
    bal_take(m[3], cr, cr, cell);
}
// Block reduction: every thread ends with the block's sums and maxima.
template <int NV>
__device__ __forceinline__ void bal_reduce(double (&sum)[NV], BalMax (&m)[4]) {
    block_sum<NV, KBLOCK>(sum);
    __shared__ BalMax sh[4][KBLOCK / 64];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double k = __shfl_xor(m[q].k, off, 64), v = __shfl_xor(m[q].v, off, 64),
                         c = __shfl_xor(m[q].c, off, 64);
            bal_take(m[q], k, v, c);
        }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) sh[q][w] = m[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
        m[q] = sh[q][0];
        for (int x = 1; x < KBLOCK / 64; x++) bal_take(m[q], sh[q][x].k, sh[q][x].v, sh[q][x].c);
    }
}
__device__ __forceinline__ void bal_put(const double (&sum)[2], const BalMax (&m)[4], double *part) {
    if (threadIdx.x != 0) return;
    double *p = part + (size_t)blockIdx.x * BAL_PART;
    p[0] = sum[0];
    p[1] = sum[1];
#pragma unroll
    for (int q = 0; q < 4; q++) { p[2 + 3 * q] = m[q].k; p[3 + 3 * q] = m[q].v; p[4 + 3 * q] = m[q].c; }
}

// Cartesian grid: storage area (h - h1) and dt times the sum of each cell's
// directed face fluxes (N, E, S, W) over the active cells, extremes over the
// wet ones; cells are numbered row-major from the south-west.
__global__ void __launch_bounds__(KBLOCK) k_bal_part(Geo g, St s, Par P, double *part) {
    double sum[2];
    BalMax m[4];
    bal_init(sum, m);
    grid_walk(g, [&](const Idx &x) {
        const int c = x.c;
        sum[0] += g.area * (s.h[c] - s.h1[c]);
        sum[1] += P.dtime * (((s.fluxN[c] + s.fluxE[c]) + (-s.fluxN[c - g.nxp])) + (-s.fluxE[c - 1]));
        if (s.iwet[c] == 1)
            bal_cell(m, s.p[c], s.u[c], s.v[c], s.h[c], g.dx, g.dy, P.dtime, (double)(x.j * g.nx + x.i));
    });
    bal_reduce(sum, m);
    bal_put(sum, m, part);
}

// Telescoping grid: the same over the cell list (faces in slot order).
__global__ void __launch_bounds__(KBLOCK) k_tl_bal_part(TelG T, TelS s, Par P, double *part) {
    double sum[2];
    BalMax m[4];
    bal_init(sum, m);
    const Chunk ch = block_chunk(T.n);
    for (int i = ch.begin + threadIdx.x; i < ch.end; i += KBLOCK) {
        const double a = T.dx[i] * T.dy[i];
        sum[0] += a * (s.h[i] - s.h1[i]);
        const int nf = m_ncf(T.meta[i]);
        double f = 0.0;
        for (int k = 0; k < nf; k++) f += s.flux[(size_t)k * T.nd + i];
        sum[1] += P.dtime * f;
        if (s.iwet[i] == 1) bal_cell(m, s.p[i], s.u[i], s.v[i], s.h[i], T.dx[i], T.dy[i], P.dtime, (double)i);
    }
    bal_reduce(sum, m);
    bal_put(sum, m, part);
}

// One block: the totals of the nb partials, the forcing string's efflux
// `bnd` (dt times the outward flux of its faces, summed by the caller's
// threads), the running sums and the percentage errors of flow_balance.
__device__ __forceinline__ void bal_finish(const double *part, int nb, double bnd, BalAcc *acc) {
    double sum[3] = {0.0, 0.0, bnd};
    BalMax m[4];
#pragma unroll
    for (int q = 0; q < 4; q++) m[q] = BalMax{0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += KBLOCK) {
        const double *p = part + (size_t)b * BAL_PART;
        sum[0] += p[0];
        sum[1] += p[1];
#pragma unroll
        for (int q = 0; q < 4; q++) bal_take(m[q], p[2 + 3 * q], p[3 + 3 * q], p[4 + 3 * q]);
    }
    bal_reduce(sum, m);
    if (threadIdx.x != 0) return;
    BalAcc &A = *acc;
    A.stg = sum[0];
    A.flux = sum[1];
    A.bnd = sum[2];
    A.cstg += A.stg;
    A.cflux += A.flux;
    A.cbnd += A.bnd;
    A.bal = A.stg + A.bnd;
    A.net = A.stg + A.flux;
    A.cbal = A.cstg + A.cbnd;
    A.cnet = A.cstg + A.cflux;
    A.err = 100.0 * A.bal / fmax(fabs(A.stg) + fabs(A.bnd), 1.0e-10);
    A.cnorm = fmax(A.cnorm, fabs(A.cstg));
    A.cerr = 100.0 * A.cbal / fmax(A.cnorm, 1.0e-10);
#pragma unroll
    for (int q = 0; q < 4; q++) { A.xv[q] = m[q].v; A.xc[q] = m[q].c; }
    A.steps += 1.0;
}

// Cartesian: the strings are the active cells along the water-level and flux
// sides, their faces the sides' ghost faces.
__global__ void __launch_bounds__(KBLOCK) k_bal_fin(Geo g, St s, Par P, const double *part, int nb, BalAcc *acc) {
    double bnd = 0.0;
    for (int k = 0; k < 4; k++) {
        if (g.side[k] == 0) continue;      // water-level and flux strings
        const int n = (k == KN || k == KS) ? g.nx : g.ny;
        for (int t = threadIdx.x; t < n; t += KBLOCK) {
            const int i = k == KE ? g.nx - 1 : k == KW ? 0 : t;
            const int j = k == KN ? g.ny - 1 : k == KS ? 0 : t;
            bnd += P.dtime * fdir(s.fluxE, s.fluxN, g, cidx(g, i, j), k);
        }
    }
    bal_finish(part, nb, bnd, acc);
}

__global__ void __launch_bounds__(KBLOCK) k_tl_bal_fin(TelG T, TelS s, Par P, const double *part, int nb,
                                                       BalAcc *acc) {
    double bnd = 0.0;
    for (int j = threadIdx.x; j < T.nstr; j += KBLOCK)
        bnd += P.dtime * s.flux[(size_t)T.str_k[j] * T.nd + T.str_c[j]];
    for (int j = threadIdx.x; j < T.nq; j += KBLOCK)
        bnd += P.dtime * s.flux[(size_t)T.q_k[j] * T.nd + T.q_c[j]];
    bal_finish(part, nb, bnd, acc);
}

}  // namespace cms

namespace cms {

// Flow statistics (stat_update_flow, source/statistics/statistics.F90:476-720):
// per-cell accumulators over the steps of a window, in the order of stat_def.
enum { S_ETAMAX, S_TETAMAX, S_ETAMIN, S_TETAMIN, S_UVMAX, S_UMAX, S_VMAX, S_TUVMAX, S_ETAMEAN, S_UR, S_VR,
       S_UVCURV, S_HYDPER, S_RSUAVG, S_RSVAVG, S_RSPAVG, S_RSUMAX, S_RSVMAX, S_RSPMAX, S_ETAMAXGRAD, S_UVMAXGRAD,
       S_FRXMAX, S_FRYMAX, S_UMAXCURVX, S_UMAXCURVY, S_VMAXCURVX, S_VMAXCURVY, S_CFLXMAX, S_CFLYMAX,
       S_QCFLXMAX, S_QCFLYMAX, S_QFRXMAX, S_QFRYMAX, NSTAT };

// One step of every active cell: the wet-cell extremes, sums and gradient
// terms (the second derivatives with the wet-cell gradient operator, d2xy2d),
// then the Courant and quality terms of every cell.  A new minimum anywhere
// sets *flag: the reference assigns the time to the whole tetamin array.
__global__ void k_stat(Geo g, St s, Par P, double *acc, size_t npad, double timehrs, int *flag) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    auto S = [&](int q) -> double & { return acc[(size_t)q * npad + c]; };
    const double u = s.u[c], v = s.v[c], h = s.h[c];
    if (s.iwet[c] != 0) {
        const double eta = s.eta[c], uv = s.uv[c];
        if (eta > S(S_ETAMAX)) { S(S_ETAMAX) = eta; S(S_TETAMAX) = timehrs; }
        if (eta < S(S_ETAMIN)) { S(S_ETAMIN) = eta; *flag = 1; }
        if (uv > S(S_UVMAX)) { S(S_UVMAX) = uv; S(S_UMAX) = u; S(S_VMAX) = v; S(S_TUVMAX) = timehrs; }
        S(S_ETAMEAN) += eta;
        S(S_UR) += u;
        S(S_VR) += v;
        // streamwise_curvature (flow_lib.F90:256-276)
        S(S_UVCURV) += (u * u * s.dvx[c] - v * v * s.duy[c] + u * v * (s.dvy[c] - s.dux[c])) /
                       pow(fmax(uv, F4(1.e-6)), 1.5);
        S(S_HYDPER) += (double)s.iwet[c];
        S(S_RSUAVG) += s.resu[c];
        S(S_RSVAVG) += s.resv[c];
        S(S_RSPAVG) += s.resp[c];
        S(S_RSPMAX) = fmax(S(S_RSPMAX), s.resp[c]);
        S(S_RSUMAX) = fmax(S(S_RSUMAX), s.resu[c]);
        S(S_RSVMAX) = fmax(S(S_RSVMAX), s.resv[c]);
        double val = sqrt(s.dpx[c] * s.dpx[c] + s.dpy[c] * s.dpy[c]) * GRAVINV;
        S(S_ETAMAXGRAD) = fmax(S(S_ETAMAXGRAD), val);
        val = sqrt(s.dux[c] * s.dux[c] + s.duy[c] * s.duy[c]);
        S(S_UVMAXGRAD) = fmax(S(S_UVMAXGRAD), val);
        val = sqrt(s.dvx[c] * s.dvx[c] + s.dvy[c] * s.dvy[c]);
        S(S_UVMAXGRAD) = fmax(S(S_UVMAXGRAD), val);
        const double cel = sqrt(GRAV * h);
        S(S_FRXMAX) = fmax(S(S_FRXMAX), u / cel);
        S(S_FRYMAX) = fmax(S(S_FRYMAX), v / cel);
        const double *vars[4] = {s.dux, s.duy, s.dvx, s.dvy};
        double gx[4], gy[4];
        cbgg_grad(g, s.iwet, c, 4, vars, gx, gy);
        S(S_UMAXCURVX) = fmax(S(S_UMAXCURVX), gx[0]);
        S(S_UMAXCURVY) = fmax(S(S_UMAXCURVY), gy[1]);
        S(S_VMAXCURVX) = fmax(S(S_VMAXCURVX), gx[2]);
        S(S_VMAXCURVY) = fmax(S(S_VMAXCURVY), gy[3]);
    }
    const double cel = sqrt(GRAV * h);
    S(S_CFLXMAX) = fmax(S(S_CFLXMAX), (fabs(u) + cel) * P.dtime / g.dx);
    S(S_CFLYMAX) = fmax(S(S_CFLYMAX), (fabs(v) + cel) * P.dtime / g.dy);
    S(S_QCFLXMAX) = fmax(S(S_QCFLXMAX), g.dx / P.dtime / (fabs(u) + sqrt(GRAV * h)));
    S(S_QCFLYMAX) = fmax(S(S_QCFLYMAX), g.dy / P.dtime / (fabs(v) + sqrt(GRAV * h)));
    S(S_QFRXMAX) = fmax(S(S_QFRXMAX), g.dx / P.dtime * fabs(u) / sqrt(GRAV * h));
    S(S_QFRYMAX) = fmax(S(S_QFRYMAX), g.dy / P.dtime * fabs(v) / sqrt(GRAV * h));
}

// tetamin = timehrs on every cell after a step that set a new minimum; the
// window's reset (all accumulators 0, etamin 1e20, the times kept) when close.
__global__ void k_stat_fix(Geo g, double *acc, size_t npad, double timehrs, const int *flag, int close) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npad) return;
    if (*flag) acc[(size_t)S_TETAMIN * npad + t] = timehrs;
    if (!close) return;
    for (int q = 0; q < NSTAT; q++) {
        if (q == S_TETAMAX || q == S_TETAMIN || q == S_TUVMAX) continue;
        acc[(size_t)q * npad + t] = q == S_ETAMIN ? F4(1.0e20) : 0.0;
    }
}

}  // namespace cms
