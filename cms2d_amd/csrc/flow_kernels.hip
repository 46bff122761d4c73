// flow_kernels.hip -- gfx950 kernels of CMS-Flow's implicit hydrodynamic step
// (erdc/cms2d 5.4, source/flow/flow_imp.F90:29-312) on a uniform Cartesian grid.
//
// Layout: every cell field is a padded row-major array over the active cells
// plus a one-cell ghost ring (common.h).  The reference's ghost cells (one per
// missing neighbour, geospatial.F90:2877-2930) are exactly the ring cells.
// Quantities that the reference stores twice per face with the same value
// (hk, visk, pk, zbk) or opposite values (flux) are stored once per face:
// xE[c] is the face between c and c+1, xN[c] the face between c and c+nxp.
// Matrix coefficients are directed (aN,aE,aS,aW per cell) as in acoef(k,i).
//
// Arithmetic follows the reference expression by expression (build with
// -ffp-contract=off) so that the non-iterative kernels reproduce the FP64
// oracle bit for bit; the Krylov solves agree to their stopping tolerance.
#include "flow_state.h"

namespace cms {

#define SMALL F4(1.0e-15)
#define RELAX_UV F4(0.8)
#define FACPP F4(0.2)
#define CVISCON F4(1.0e-6)
#define CVISBOT F4(0.0667)
#define CVISMAX 20.0
#define CVISWAV 0.5
#define CVISWAVBRK F4(0.1)
#define CFRICWAV F4(0.65)
#define RHOW 1025.0

__device__ __constant__ int c_kopp[4] = {KS, KW, KN, KE};

// ---------------------------------------------------------------- helpers
// comp_lib.f90:40-109 (ndsch 0 zero, 1 upwind, 2 hybrid, 3 power law,
// 4 exponential; every other scheme takes upwind coefficients,
// flow_imp.F90:120-126)
__device__ __forceinline__ double schmcoef(int ndsch, double dk, double fk) {
    double absfk = fabs(fk), peck;
    switch (ndsch) {
    case 0:
        return 0.0;
    case 2:
        peck = absfk / fmax(dk, F4(1.0e-6));
        return dk * fmax(0.0, 1.0 - 0.5 * peck) + fmax(0.0, -fk);
    case 3:
        peck = absfk / fmax(fmax(dk, absfk), F4(1.0e-6));
        return dk * (1.0 - pow(peck, 5.0)) + fmax(0.0, -fk);
    case 4:
        peck = absfk / fmax(fmax(dk, F4(1.e-6)), absfk);
        return dk * peck / fmax(exp(peck) - 1.0, F4(1.0e-6)) + fmax(0.0, -fk);
    default:
        return dk + fmax(0.0, -fk);
    }
}
// fric_lib.f90:162-194
__device__ __forceinline__ double wall_coef(double yp, double y0) {
    double y0lim = fmin(fmax(y0, F4(1.0e-6)), yp * F4(0.9999));
    double t = F4(0.4) / log(yp / y0lim);
    return t * t;
}
// fric_lib.f90:944-996
__device__ __forceinline__ double man2drag(double cman, double h) {
    double hh = fmax(fmin(10.0, h), F4(0.1) * cman);
    return GRAV * cman * cman / pow(hh, F4(0.3333));
}
__device__ __forceinline__ double drag2length(double cd, double h) {
    double z0 = h / exp(F4(0.4) / sqrt(cd) + 1.0);
    return fmax(z0, F4(1.0e-6));
}
// fric_lib.f90:916-941, 999-1023
__device__ __forceinline__ double drag2man(double cd, double h, double hmax) {
    return sqrt(cd * pow(fmin(hmax, h), F4(0.3333)) / GRAV);
}
__device__ __forceinline__ double length2drag(double z0, double h) {
    double z0lim = fmax(fmin(z0, 0.5 * h), F4(1.0e-5));
    double t = F4(0.4) / (log(h / z0lim) - 1.0);
    return t * t;
}
// The per-law part of fric_eval (friction.F90:638-679): the law's input
// (cfrict, z0, cbf or coefman) stays, the other coefficients follow from it
// and the depth; returns the drag coefficient.
__device__ __forceinline__ double fric_law(const Par &P, const St &s, int c, double h) {
    double cd;
    switch (P.mbedfric) {
    case 1:
        cd = s.cfrict[c];
        s.coefman[c] = drag2man(cd, h, 10.0);
        s.z0[c] = drag2length(cd, h);
        break;
    case 3:
        cd = length2drag(s.z0[c], h);
        s.cfrict[c] = cd;
        s.coefman[c] = drag2man(cd, h, 10.0);
        break;
    case 4:   // uelwc is the previous evaluation's
        cd = s.cbf[c] * h / fmax(s.uelwc[c], F4(1.0e-15));
        s.cfrict[c] = cd;
        s.coefman[c] = drag2man(cd, h, h);
        s.z0[c] = drag2length(cd, h);
        break;
    default:
        cd = man2drag(s.coefman[c], h);
        s.cfrict[c] = cd;
        if (cd != 0.0) s.z0[c] = drag2length(cd, h);
    }
    return cd;
}

// directed face values of cell c
__device__ __forceinline__ double fdir(const double *E, const double *N, const Geo &g, int c, int k) {
    return k == KN ? N[c] : k == KE ? E[c] : k == KS ? -N[c - g.nxp] : -E[c - 1];
}
__device__ __forceinline__ double fsym(const double *E, const double *N, const Geo &g, int c, int k) {
    return k == KN ? N[c] : k == KE ? E[c] : k == KS ? N[c - g.nxp] : E[c - 1];
}
// is the neighbour of active cell (i,j) in direction k a ghost (ring) cell?
__device__ __forceinline__ bool nb_ghost(const Geo &g, int i, int j, int k) {
    return k == KN ? j == g.ny - 1 : k == KE ? i == g.nx - 1 : k == KS ? j == 0 : i == 0;
}
// Ring cell -> side it lies on (-1 for corners and active cells).
__device__ __forceinline__ int ring_side(const Geo &g, int i, int j) {
    bool ci = (i == -1 || i == g.nx), cj = (j == -1 || j == g.ny);
    if (ci && cj) return -1;
    if (j == g.ny) return KN;
    if (i == g.nx) return KE;
    if (j == -1) return KS;
    if (i == -1) return KW;
    return -1;
}
__device__ __forceinline__ void pad_ij(const Geo &g, int t, int &i, int &j) {
    j = t / g.nxp - 1;
    i = t - (j + 1) * g.nxp - 1;
}

// Cell-based Green-Gauss gradient with wet-only stencils (der_grad_cbgg with
// gow%ibc=1, der_lib.f90:272-378, screened and re-centred as :360-376), then
// der_grad_eval (:726-839).  Weights are rebuilt from iwet on every call,
// which is what der_update keeps them equal to (derivatives.f90:204-296).
__device__ __forceinline__ void cbgg_grad(const Geo &g, const uint8_t *iwet, int c, int nv,
                                          const double *const *var, double *gx, double *gy) {
    if (!iwet[c]) {
        for (int q = 0; q < nv; q++) { gx[q] = 0.0; gy[q] = 0.0; }
        return;
    }
    double wx[5] = {0, 0, 0, 0, 0}, wy[5] = {0, 0, 0, 0, 0};
    int ix[5];
    ix[0] = c;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int nb = nbr(g, c, k);
        ix[k + 1] = nb;
        if (iwet[nb] == 1) {
            wx[0] += g.dsx[k] * (1.0 - g.fintp[k]);
            wx[k + 1] = g.dsx[k] * g.fintp[k];
            wy[0] += g.dsy[k] * (1.0 - g.fintp[k]);
            wy[k + 1] = g.dsy[k] * g.fintp[k];
        } else {
            wx[0] += g.dsx[k];
            wy[0] += g.dsy[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 5; k++) { wx[k] /= g.area; wy[k] /= g.area; }
    const double tolw = F4(1.0e-8);
    int nx_ = 0, ny_ = 0;
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        if (fabs(wx[k]) > tolw) { nx_++; sx += wx[k]; }
        if (fabs(wy[k]) > tolw) { ny_++; sy += wy[k]; }
    }
    double mx = nx_ ? sx / (double)nx_ : 0.0, my = ny_ ? sy / (double)ny_ : 0.0;
    for (int q = 0; q < nv; q++) {
        double ax = 0.0, ay = 0.0;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            if (fabs(wx[k]) > tolw) ax += (wx[k] - mx) * var[q][ix[k]];
            if (fabs(wy[k]) > tolw) ay += (wy[k] - my) * var[q][ix[k]];
        }
        gx[q] = ax;
        gy[q] = ay;
    }
}

// ------------------------------------------------------- wetting and drying
// flow_wetdry(0), flow.F90:1257-1292: depth and wet flag on every cell.
__global__ void k_wetdry_h(Geo g, St s, Par P) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    int i, j;
    pad_ij(g, t, i, j);
    double h = s.p[t] * GRAVINV - s.zb[t];
    uint8_t w = 1;
    if (h < P.hdry) { h = fmax(h, P.hmin); w = 0; }
    if (!active_ij(g, i, j)) {
        int sd = ring_side(g, i, j);
        if (sd < 0 || g.side[sd] == 0) {
            w = 0;                                  // corner or wall ghost
        } else {
            // forcing string: the ghost takes its cell's flag (flow.F90:1606-1613)
            const int c = sd == KN ? t - g.nxp : sd == KE ? t - 1 : sd == KS ? t + g.nxp : t + 1;
            w = s.p[c] * GRAVINV - s.zb[c] < P.hdry ? 0 : 1;
        }
    }
    s.h[t] = h;
    s.iwet[t] = w;
}

// Narrow-channel removal, flow.F90:1342-1455.  Every rule dries a cell when
// fewer of its neighbours are wet, and cells only ever dry, so any order of
// in-place updates repeated until a pass changes nothing reaches the same
// fixed point as the reference's sequential sweep.
__global__ void k_narrow(Geo g, St s, Par P) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (!s.iwet[c]) return;
    const uint8_t *w = s.iwet;
    int n = w[c + g.nxp], e = w[c + 1], so = w[c - g.nxp], we = w[c - 1];
    double h = s.h[c];
    bool dry = false;
    if (n + e + so + we <= 1) dry = true;
    if (((n + so) == 0 || (e + we) == 0) && h < P.hdry1) dry = true;
    int ss2 = j > 0 ? w[c - 2 * g.nxp] : 0;
    int ww2 = i > 0 ? w[c - 2] : 0;
    int ee2 = i < g.nx - 1 ? w[c + 2] : 0;
    if (n + ss2 == 0 && h < P.hdry2) dry = true;     // knss (tested twice in the reference)
    if (e + ww2 == 0 && h < P.hdry2) dry = true;     // keww
    if (we + ee2 == 0 && h < P.hdry2) dry = true;    // keew
    if (dry) {
        s.iwet[c] = 0;
        s.flags[0] = 1;
    }
}

// flow_pond (flow.F90:1807-1876): wet cells not reached from a forcing
// boundary are dried.  The reference seeds the boundary-string cells and their
// ghosts and spreads from reached ACTIVE cells to wet neighbours until nothing
// changes, so the reached set is every wet cell in a connected component that
// holds a seed, where two neighbours are connected when both are wet and at
// least one is active (ghosts receive but do not pass on).  The components
// are labelled in one pass instead of by repeated sweeps:
//   k_pond_local    union-find inside 32x32 tiles in LDS (root = smallest
//                   padded index of the tile-local component)
//   k_pond_merge    union across tile edges on the global label array
//   k_pond_compress label = root (smallest padded index of the component)
//   k_pond_mark     reached[root] = 1 for every wet seed
//   k_pond_final    wet cells whose root was not reached are dried
// Labels only ever decrease towards the component minimum, so the result does
// not depend on the order in which the atomic unions land.
constexpr int PT = 32;

__device__ __forceinline__ bool pond_link(const Geo &g, const uint8_t *iwet, int c, int d) {
    if (!iwet[c] || !iwet[d]) return false;
    int i, j, k, l;
    pad_ij(g, c, i, j);
    pad_ij(g, d, k, l);
    return active_ij(g, i, j) || active_ij(g, k, l);
}

__device__ __forceinline__ int lds_find(int *lab, int x) {
    for (;;) {
        int y = __hip_atomic_load(&lab[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (y == x) return x;
        x = y;
    }
}
__device__ __forceinline__ void lds_union(int *lab, int a, int b) {
    for (;;) {
        a = lds_find(lab, a);
        b = lds_find(lab, b);
        if (a == b) return;
        if (a < b) { int t = a; a = b; b = t; }             // link the larger root a under b
        int old = atomicMin(&lab[a], b);
        if (old == a) return;
        a = old;
    }
}
__device__ __forceinline__ int glb_find(int *P, int x) {
    for (;;) {
        int y = __hip_atomic_load(&P[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (y == x) return x;
        x = y;
    }
}
__device__ __forceinline__ void glb_union(int *P, int a, int b) {
    for (;;) {
        a = glb_find(P, a);
        b = glb_find(P, b);
        if (a == b) return;
        if (a < b) { int t = a; a = b; b = t; }
        int old = atomicMin(&P[a], b);
        if (old == a) return;
        a = old;
    }
}

__global__ void __launch_bounds__(256) k_pond_local(Geo g, St s) {
    __shared__ int lab[PT * PT];
    const int X0 = blockIdx.x * PT, Y0 = blockIdx.y * PT;
    for (int q = threadIdx.x; q < PT * PT; q += blockDim.x) {
        const int X = X0 + q % PT, Y = Y0 + q / PT;
        lab[q] = (X < g.nxp && Y < g.nyp && s.iwet[Y * g.nxp + X]) ? q : -1;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < PT * PT; q += blockDim.x) {
        const int lx = q % PT, ly = q / PT, X = X0 + lx, Y = Y0 + ly;
        if (X >= g.nxp || Y >= g.nyp) continue;
        const int c = Y * g.nxp + X;
        if (lx > 0 && pond_link(g, s.iwet, c, c - 1)) lds_union(lab, q, q - 1);
        if (ly > 0 && pond_link(g, s.iwet, c, c - g.nxp)) lds_union(lab, q, q - PT);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < PT * PT; q += blockDim.x) {
        const int X = X0 + q % PT, Y = Y0 + q / PT;
        if (X >= g.nxp || Y >= g.nyp) continue;
        const int c = Y * g.nxp + X;
        int r = -1;
        if (lab[q] >= 0) {
            const int rq = lds_find(lab, q);
            r = (Y0 + rq / PT) * g.nxp + (X0 + rq % PT);
        }
        s.plab[c] = r;
        s.ipond[c] = 0;
    }
}

// Edges that cross a tile boundary: threads [0, nv) take the west edges of
// tile columns 1.., threads [nv, nv+nh) the south edges of tile rows 1..
__global__ void k_pond_merge(Geo g, St s) {
    const int ntx = (g.nxp + PT - 1) / PT, nty = (g.nyp + PT - 1) / PT;
    const int nv = (ntx - 1) * g.nyp, nh = (nty - 1) * g.nxp;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int c, d;
    if (t < nv) {
        const int Y = t / (ntx - 1), X = (t % (ntx - 1) + 1) * PT;
        c = Y * g.nxp + X;
        d = c - 1;
    } else if (t < nv + nh) {
        const int u = t - nv, X = u % g.nxp, Y = (u / g.nxp + 1) * PT;
        c = Y * g.nxp + X;
        d = c - g.nxp;
    } else {
        return;
    }
    if (pond_link(g, s.iwet, c, d)) glb_union(s.plab, c, d);
}

__global__ void k_pond_compress(Geo g, St s) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    if (s.plab[t] >= 0) s.plab[t] = glb_find(s.plab, t);
}

// Seeds: boundary-string cells and the ghosts across their boundary faces.
__global__ void k_pond_mark(Geo g, St s) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    int i, j;
    pad_ij(g, t, i, j);
    bool seed;
    if (active_ij(g, i, j)) {
        seed = (j == g.ny - 1 && g.side[KN]) || (i == g.nx - 1 && g.side[KE]) || (j == 0 && g.side[KS]) ||
               (i == 0 && g.side[KW]);
    } else {
        const int sd = ring_side(g, i, j);
        seed = sd >= 0 && g.side[sd];
    }
    if (seed && s.iwet[t]) s.ipond[s.plab[t]] = 1;
}

__global__ void k_pond_final(Geo g, St s) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    if (s.iwet[t] == 1 && !s.ipond[s.plab[t]]) s.iwet[t] = 0;
}

// Corner flags, newly wet count, and zero flux/visk on faces not wet on both
// sides (flow.F90:1459-1525).
__global__ void k_wetdry_post(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const uint8_t *w = s.iwet;
    uint8_t ic = 0;
    if (w[c]) {
        int nd = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) nd += w[nbr(g, c, k)] == 0;
        if (nd > 1) ic = 1;
    }
    s.icorner[c] = ic;
    if (w[c] == 1 && s.iwet1[c] == 0) atomicAdd(&s.flags[2], 1);
    if (w[c] * w[c + 1] == 0) { s.fluxE[c] = 0.0; s.viskE[c] = 0.0; }
    if (w[c] * w[c + g.nxp] == 0) { s.fluxN[c] = 0.0; s.viskN[c] = 0.0; }
    if (i == 0 && w[c] * w[c - 1] == 0) { s.fluxE[c - 1] = 0.0; s.viskE[c - 1] = 0.0; }
    if (j == 0 && w[c] * w[c - g.nxp] == 0) { s.fluxN[c - g.nxp] = 0.0; s.viskN[c - g.nxp] = 0.0; }
}

// Dry-to-wet initialisation (flow.F90:1477-1491), one sweep, every read from
// the values before the sweep.
__global__ void k_d2w_sweep(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (!(s.iwet[c] == 1 && s.iwet1[c] == 0)) return;
    double us = 0.0, vs = 0.0, viss = 0.0;
    int ns = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int nb = nbr(g, c, k);
        us += s.u[nb]; vs += s.v[nb]; viss += s.vis[nb]; ns += s.iwet[nb];
    }
    double den = (double)(float)ns + SMALL;
    s.tu[c] = us / den / F4(1.25);
    s.tv[c] = vs / den / F4(1.25);
    s.tvis[c] = viss / den;
}
__global__ void k_d2w_commit(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (!(s.iwet[c] == 1 && s.iwet1[c] == 0)) return;
    s.u[c] = s.tu[c]; s.v[c] = s.tv[c]; s.vis[c] = s.tvis[c];
}

// ------------------------------------------------------------ momentum
// flow_srcsnk_static, flow_imp.F90:336-403 (ntsch=1, wave forcing when noptset>=3)
__global__ void k_srcsnk(Geo g, St s, Par P) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    double gdi = 1.0 / (GRAV * P.dtime), dti = 1.0 / P.dtime;
    double fac1 = g.area * s.h1[c] * dti;
    s.spuv0[c] = -fac1;
    double su0 = fac1 * s.u1[c], sv0 = fac1 * s.v1[c];
    s.sppp0[c] = -g.area * gdi;
    s.sspp0[c] = -s.p1[c] * s.sppp0[c];
    if (P.noptset >= 3) {
        su0 = su0 + s.wavestrx[c] * g.area;
        sv0 = sv0 + s.wavestry[c] * g.area;
    }
    if (P.tauwx != 0.0 || P.tauwy != 0.0) {   // constant wind, flow_imp.F90:386-394
        su0 = su0 + P.tauwx * g.area;
        sv0 = sv0 + P.tauwy * g.area;
    }
    s.ssu0[c] = su0;
    s.ssv0[c] = sv0;
}

// coeffsourcesink_uv (flow_imp.F90:444-566) + bound_uv (boundaries.F90:3310-3490)
// with bnd_vel_openwse (bnd_vel.f90:35-95) for harmonic water-level sides.
// NVD deferred corrections on a Cartesian grid (flow_imp.F90:140-147):
// defcorhlpavec (ndsch 5, computational.F90:337-414) and defcorgammavec
// (ndsch 6-9, :495-581) with the limiters of comp_lib.f90:112-208.  The
// deferred flux of a face is evaluated from its owner (the S/W cell, which
// lists it in kxyface) with the owner's outward flux: C the upwind cell, D
// the downwind one, U the cell upwind of C (a ghost included for the gamma
// family; HLPA skips the face when U is a ghost or dry).
__device__ __forceinline__ double nvd_gamma(int ndsch, double x) {
    switch (ndsch) {
    case 6: return (x > 0.0 && x <= 0.5) ? 2.0 * x : (x > 0.5 && x <= 1.0) ? 1.0 : 0.0;
    case 7:
        return (x > 0.0 && x <= 0.375)  ? 1.5 * x / (1.0 - x)
               : (x > 0.375 && x <= 0.75) ? (3.0 - 2.0 * x) / (4.0 - 4.0 * x)
               : (x > 0.75 && x <= 1.0)   ? 1.5
                                          : 0.0;
    case 8:
        return (x > 0.0 && x <= 0.25)  ? 2.5 * x / (1.0 - x)
               : (x > 0.25 && x <= 0.75) ? (3.0 - 2.0 * x) / (4.0 - 4.0 * x)
               : (x > 0.75 && x <= 1.0)  ? 1.5
                                         : 0.0;
    case 9:
        return (x > 0.0 && x <= F4(0.1667))  ? 5.0 * x / (1.0 - x)
               : (x > F4(0.1667) && x <= 0.5) ? 1.0
               : (x > 0.5 && x <= 0.75)       ? 0.5 / (1.0 - x)
               : (x > 0.75 && x <= 1.0)       ? 2.0
                                              : 0.0;
    default: return 0.0;
    }
}
// Deferred flux (u, v) of face k (KN or KE) owned by cell (i, j).
__device__ __forceinline__ void nvd_face(const Geo &g, const St &s, int ndsch, int i, int j, int k, double &fu,
                                         double &fv) {
    fu = fv = 0.0;
    const int c = cidx(g, i, j);
    if (!s.iwet[c] || nb_ghost(g, i, j, k)) return;
    const int nck = nbr(g, c, k);
    if (!s.iwet[nck]) return;
    const double F = k == KN ? s.fluxN[c] : s.fluxE[c];
    int cc, cd, cu;
    bool u_ghost;
    if (F >= 0.0) {
        cc = c; cd = nck; cu = nbr(g, c, c_kopp[k]);
        u_ghost = nb_ghost(g, i, j, c_kopp[k]);
    } else {
        const int ni = k == KE ? i + 1 : i, nj = k == KN ? j + 1 : j;
        cc = nck; cd = c; cu = nbr(g, nck, k);
        u_ghost = nb_ghost(g, ni, nj, k);
    }
    if (ndsch == 5 && (u_ghost || !s.iwet[cu])) return;
    const double tiny = F4(1.0e-20);
    const double uc = s.u[cc], ud = s.u[cd], uu = s.u[cu];
    const double vc = s.v[cc], vd = s.v[cd], vu = s.v[cu];
    if (fabs(ud - uu) > tiny) {
        const double un = (uc - uu) / (ud - uu);
        if (un > 0.0 && un < 1.0) {
            if (ndsch == 5) fu = F * (ud - uc) * un;
            else fu = F * nvd_gamma(ndsch, un) * ((g.fintp[k] * s.u[nck] + (1.0 - g.fintp[k]) * s.u[c]) - uc);
        }
    }
    if (fabs(vd - vu) > tiny) {
        const double vn = (vc - vu) / (vd - vu);
        if (vn > 0.0 && vn < 1.0) {
            if (ndsch == 5) fv = F * (vd - vc) * vn;
            else fv = F * nvd_gamma(ndsch, vn) * ((g.fintp[k] * s.v[nck] + (1.0 - g.fintp[k]) * s.v[c]) - vc);
        }
    }
}
// Sums over each cell's faces (N, E, S, W, the reference's face order) into
// tu, tv; k_mom subtracts them from su, sv.
__global__ void k_defcor(Geo g, St s, Par P) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    double nu, nv, eu, ev, su_ = 0.0, sv_ = 0.0, wu = 0.0, wv = 0.0;
    nvd_face(g, s, P.ndsch, i, j, KN, nu, nv);
    nvd_face(g, s, P.ndsch, i, j, KE, eu, ev);
    if (j > 0) { nvd_face(g, s, P.ndsch, i, j - 1, KN, su_, sv_); su_ = -su_; sv_ = -sv_; }
    if (i > 0) { nvd_face(g, s, P.ndsch, i - 1, j, KE, wu, wv); wu = -wu; wv = -wv; }
    s.tu[c] = ((nu + eu) + su_) + wu;
    s.tv[c] = ((nv + ev) + sv_) + wv;
}

__global__ void k_mom(Geo g, St s, Par P) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const double A = g.area, fc = s.fc[c];
    double u = s.u[c], v = s.v[c], h = s.h[c];
    double su = s.ssu0[c] + (fc * v - s.dpx[c]) * h * A;
    double sv = s.ssv0[c] + (-fc * u - s.dpy[c]) * h * A;
    if (P.ndsch >= 5 && P.ndsch <= 9) {   // deferred corrections of k_defcor
        su = su - s.tu[c];
        sv = sv - s.tv[c];
    }
    double cbc = 1.0 * s.cfrict[c] * s.uelwc[c] * A;          // cmb=1 (no bed-slope factor)
    double sp = s.spuv0[c] - cbc;
    double a[4], hk[4], fl[4], vk[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        hk[k] = fsym(s.hkE, s.hkN, g, c, k);
        vk[k] = fsym(s.viskE, s.viskN, g, c, k);
        fl[k] = fdir(s.fluxE, s.fluxN, g, c, k);
        double ddk = vk[k] * hk[k] * g.dsxy[k];
        a[k] = schmcoef(P.ndsch, ddk, fl[k]);
    }
    const int wet = s.iwet[c];
    // flux sides first (bound_uv's order): bnd_vel_flux (bnd_vel.f90:2-32), the
    // upwind coefficient times the ghost velocity moves into the sources
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!(nb_ghost(g, i, j, k) && g.side[k] == 2)) continue;
        const int gc = nbr(g, c, k);
        const double ab = schmcoef(1, vk[k] * hk[k] * g.dsxy[k], fl[k]);
        su += ab * s.u[gc];
        sv += ab * s.v[gc];
        sp -= ab;
        a[k] = 0.0;
    }
    // harmonic water-level sides: fold the boundary face into su/sv/sp
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!(nb_ghost(g, i, j, k) && g.side[k] == 1)) continue;
        int gc = nbr(g, c, k);
        double ddk = vk[k] * hk[k] * g.dsxy[k];
        double ab = schmcoef(1, ddk, fl[k]);
        if (fabs(g.fny[k]) > F4(0.9999)) {
            s.u[gc] = wet * u * h / s.h[gc];
            double vg = g.fny[k] * fl[k] / (g.ds[k] * hk[k]);
            s.v[gc] = vg;
            sv += ab * vg;
        } else {
            double ug = g.fnx[k] * fl[k] / (g.ds[k] * hk[k]);
            s.u[gc] = ug;
            s.v[gc] = wet * v * h / s.h[gc];
            su += ab * ug;
        }
        sp -= ab;
        a[k] = 0.0;
    }
    double spu, spv;
    if (P.wallfric) {
        if (wet == 1) {
            spu = sp;
            spv = sp;
            double z0w = s.z0[c] * 1.0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (s.iwet[nbr(g, c, k)] != 0) continue;
                if (fabs(g.fny[k]) > F4(0.999))
                    spu = spu - g.ds[k] * hk[k] * wall_coef(fabs(g.ry[k]), z0w) * fabs(u);
                else
                    spv = spv - g.ds[k] * hk[k] * wall_coef(fabs(g.rx[k]), z0w) * fabs(v);
            }
        } else {
            s.u[c] = 0.0; s.v[c] = 0.0;
            spu = -1.0; spv = -1.0; sp = -1.0;
            su = 0.0; sv = 0.0;
        }
    } else {
        s.u[c] = wet * u;
        s.v[c] = wet * v;
        sp = wet * sp + (double)(wet - 1);
        spu = sp; spv = sp;
        su = wet * su; sv = wet * sv;
    }
    s.su[c] = su; s.sv[c] = sv; s.sp[c] = sp; s.spu[c] = spu; s.spv[c] = spv;
    s.aN[c] = a[KN]; s.aE[c] = a[KE]; s.aS[c] = a[KS]; s.aW[c] = a[KW];
}

// ------------------------------------------------------- pressure correction
// coefsourcesink_pp, flow_imp.F90:608-621
__global__ void k_pp_prep(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    double asum = ((s.aN[c] + s.aE[c]) + s.aS[c]) + s.aW[c];
    int w = s.iwet[c];
    double apu = w * RELAX_UV / fmax(asum - s.sp[c], F4(1.0e-6));
    double apa = apu * g.area;
    s.apuareap[c] = apa;
    s.sumu[c] = asum * apu;
    s.Hu[c] = w * (s.u[c] + apa * s.dpx[c] * s.h[c]);
    s.Hv[c] = w * (s.v[c] + apa * s.dpy[c] * s.h[c]);
}

// pp coefficient of the face of `own` in direction k (flow_imp.F90:628-647)
__device__ __forceinline__ double pp_coef(const Geo &g, const St &s, int own, int oi, int oj, int k) {
    int nck = nbr(g, own, k);
    double hk = fsym(s.hkE, s.hkN, g, own, k);
    if (s.iwet[own] * s.iwet[nck] == 0) return 0.0;
    if (nb_ghost(g, oi, oj, k))
        return hk * hk * g.ds[k] * s.apuareap[own] / (1.0 - s.sumu[own]) / g.dn[k];
    double fk = g.fintp[k], fp = 1.0 - fk;
    double apf = fk * s.apuareap[nck] + fp * s.apuareap[own];
    double suf = fk * s.sumu[nck] + fp * s.sumu[own];
    return hk * hk * g.ds[k] * apf / (1.0 - suf) / g.dn[k];
}
// fluxbnd (boundaries.F90:4740-4764): the outward flux of a flux-side face
// from its ghost's velocity
__device__ __forceinline__ double bnd_flux(const Geo &g, const St &s, int c, int k) {
    const int gc = nbr(g, c, k);
    const double hk = fsym(s.hkE, s.hkN, g, c, k);
    return (double)s.iwet[c] * g.ds[k] * hk * (g.fnx[k] * s.u[gc] + g.fny[k] * s.v[gc]);
}
// Rhie-Chow face flux of `own` in direction k (flow_imp_flux.f90:382-406);
// on flux sides the bndflux that follows (flow_imp.F90:697) overrides it
__device__ __forceinline__ double rc_flux(const Geo &g, const St &s, int own, int oi, int oj, int k) {
    int nck = nbr(g, own, k);
    if (nb_ghost(g, oi, oj, k) && g.side[k] == 2) return bnd_flux(g, s, own, k);
    if (s.iwet[own] * s.iwet[nck] == 0) return 0.0;
    double hk = fsym(s.hkE, s.hkN, g, own, k);
    if (nb_ghost(g, oi, oj, k))
        return hk * g.ds[k] * ((g.fnx[k] * s.Hu[own] + g.fny[k] * s.Hv[own])
                               - s.apuareap[own] * hk * (s.p[nck] - s.p[own]) / g.dn[k]);
    double fk = g.fintp[k], fp = 1.0 - fk;
    double apf = fk * s.apuareap[nck] + fp * s.apuareap[own];
    double Huf = fk * s.Hu[nck] + fp * s.Hu[own];
    double Hvf = fk * s.Hv[nck] + fp * s.Hv[own];
    return hk * g.ds[k] * ((g.fnx[k] * Huf + g.fny[k] * Hvf) - apf * hk * (s.p[nck] - s.p[own]) / g.dn[k]);
}

__global__ void k_pp_coef_flux(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    // coefficients: each face evaluated from its owner (the cell that lists it
    // first in kxyface: the S/W cell of an interior face, the active cell of a
    // boundary face)
    double aN = pp_coef(g, s, c, i, j, KN);
    double aE = pp_coef(g, s, c, i, j, KE);
    double aS = j > 0 ? pp_coef(g, s, c - g.nxp, i, j - 1, KN) : pp_coef(g, s, c, i, j, KS);
    double aW = i > 0 ? pp_coef(g, s, c - 1, i - 1, j, KE) : pp_coef(g, s, c, i, j, KW);
    s.aN[c] = aN; s.aE[c] = aE; s.aS[c] = aS; s.aW[c] = aW;
    s.fluxN[c] = rc_flux(g, s, c, i, j, KN);
    s.fluxE[c] = rc_flux(g, s, c, i, j, KE);
    if (j == 0) s.fluxN[c - g.nxp] = -rc_flux(g, s, c, i, j, KS);
    if (i == 0) s.fluxE[c - 1] = -rc_flux(g, s, c, i, j, KW);
}

// bnd_wse_adjust (boundaries.F90:4620-4686) for one harmonic side: setup from
// the radiation-stress forcing integrated outward from the centre cell.
// The string's level before set-up: (1 - ramp) 0 + ramp (w + wsevar) with the
// temporally constant gradient term wsevar = dist . (gx, gy) from the string's
// middle cell (bnd_eval :4121-4122,4148; bnd_init :1892-1901).
__global__ void __launch_bounds__(1024) k_wse_adjust(Geo g, St s, Par P, int side, double w, double ramp, double gx,
                                                     double gy, int noadj) {
    extern __shared__ double term[];
    const int n = (side == KE || side == KW) ? g.ny : g.nx;
    double *wse = s.wseadj + side * P.lmax;
    const int im0 = n / 2 - 1;
    const bool adjust = n >= 3 && !noadj;
    const int iref = n / 2 - 1 >= 0 ? n / 2 - 1 : 0;
    auto base = [&](int im) {
        const double dist = (double)(im - iref) * ((side == KE || side == KW) ? g.dy : g.dx);
        const double wv = (side == KE || side == KW) ? 0.0 * gx + dist * gy : dist * gx + 0.0 * gy;
        return (1.0 - ramp) * 0.0 + ramp * (w + wv);
    };
    for (int im = threadIdx.x; im < n; im += blockDim.x) {
        double tv = 0.0;
        if (adjust && im != im0) {
            int im1 = im > im0 ? im - 1 : im + 1;
            int ci, cj, ci1, cj1;
            if (side == KE || side == KW) { ci = ci1 = (side == KE ? g.nx - 1 : 0); cj = im; cj1 = im1; }
            else { cj = cj1 = (side == KN ? g.ny - 1 : 0); ci = im; ci1 = im1; }
            int c = cidx(g, ci, cj), c1 = cidx(g, ci1, cj1);
            double fx = 0.0, fx1 = 0.0, fy = 0.0, fy1 = 0.0;
            if (P.noptset >= 3) { fx = s.wavestrx[c]; fx1 = s.wavestrx[c1]; fy = s.wavestry[c]; fy1 = s.wavestry[c1]; }
            if (P.tauwx != 0.0 || P.tauwy != 0.0) {   // forcex/forcey add the wind stress, boundaries.F90:3978-3982
                fx = fx + P.tauwx; fx1 = fx1 + P.tauwx; fy = fy + P.tauwy; fy1 = fy1 + P.tauwy;
            }
            double xd = (ci - ci1) * g.dx, yd = (cj - cj1) * g.dy;
            double sx = xd * (fx + fx1) / (s.h[c] + s.h[c1]) * GRAVINV;
            double sy = yd * (fy + fy1) / (s.h[c] + s.h[c1]) * GRAVINV;
            if (fabs(g.ry[side]) >= F4(0.999)) tv = sx;
            else if (fabs(g.rx[side]) >= F4(0.999)) tv = sy;
            else tv = sx + sy;
        }
        term[im] = tv;
    }
    __syncthreads();
    if (threadIdx.x == 0 && adjust) {
        // sequential prefix in the reference's order, carried in a register;
        // terms are read 16 at a time so the LDS loads overlap
        constexpr int U = 16;
        term[im0] = 0.0;
        double acc = 0.0;
        int im = im0 + 1;
        for (; im + U <= n; im += U) {
            double v[U];
#pragma unroll
            for (int k = 0; k < U; k++) v[k] = term[im + k];
#pragma unroll
            for (int k = 0; k < U; k++) { acc = acc + v[k]; term[im + k] = acc; }
        }
        for (; im < n; im++) { acc = acc + term[im]; term[im] = acc; }
        acc = 0.0;
        im = im0 - 1;
        for (; im - U + 1 >= 0; im -= U) {
            double v[U];
#pragma unroll
            for (int k = 0; k < U; k++) v[k] = term[im - k];
#pragma unroll
            for (int k = 0; k < U; k++) { acc = acc + v[k]; term[im - k] = acc; }
        }
        for (; im >= 0; im--) { acc = acc + term[im]; term[im] = acc; }
    }
    __syncthreads();
    for (int im = threadIdx.x; im < n; im += blockDim.x) wse[im] = adjust ? base(im) + term[im] : base(im);
}

// pp sources (flow_imp.F90:700-712) + bound_pp (boundaries.F90:3493-3646) +
// solve prologue for n=1 (solvers.f90:232-240, relax(1)=1) + BiCGStab start.
// Runs on the fixed Krylov grid and leaves r.r / res.res partials.
__global__ void __launch_bounds__(KBLOCK) k_pp_src(Geo g, St s, Par P, KrySys<1> K) {
    double red[2] = {0.0, 0.0};
    grid_walk(g, [&](const Idx &x_) {
        const int i = x_.i, j = x_.j, c = x_.c;
        double sp = s.sppp0[c];
        double fs = ((s.fluxN[c] + s.fluxE[c]) + (-s.fluxN[c - g.nxp])) + (-s.fluxE[c - 1]);
        double su = s.sspp0[c] - fs + s.p[c] * sp;
        double a[4] = {s.aN[c], s.aE[c], s.aS[c], s.aW[c]};
        const int wet = s.iwet[c];
        if (wet == 0) {
            int sumwet = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) if (!nb_ghost(g, i, j, k)) sumwet += s.iwet[nbr(g, c, k)];
            double aploc = 0.0;
            su = 0.0;
            if (sumwet >= 1) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    int nb = nbr(g, c, k);
                    a[k] = s.iwet[nb];
                    aploc += a[k];
                    su += a[k] * s.p[nb];
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    int nb = nbr(g, c, k);
                    if (nb_ghost(g, i, j, k)) a[k] = 0.0;
                    else { a[k] = g.dsxy[k]; aploc += a[k]; su += a[k] * s.p[nb]; }
                }
            }
            su -= aploc * s.p[c];
            sp = 0.0;
        }
#pragma unroll
        for (int k = 0; k < 4; k++)      // flux sides: closed faces (boundaries.F90:3544-3551)
            if (nb_ghost(g, i, j, k) && g.side[k] == 2) a[k] = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!(nb_ghost(g, i, j, k) && g.side[k] == 1)) continue;
            int gc = nbr(g, c, k);
            int im = (k == KE || k == KW) ? j : i;
            double wse = s.wseadj[k * P.lmax + im];
            double ppg = wse * GRAV - s.p[gc];
            s.pp[gc] = ppg;
            su += a[k] * ppg;
            sp -= a[k];
            a[k] = 0.0;
        }
        s.aN[c] = a[KN]; s.aE[c] = a[KE]; s.aS[c] = a[KS]; s.aW[c] = a[KW];
        s.sp[c] = sp;
        // solve prologue, relax(1) = 1
        double asum = ((a[0] + a[1]) + a[2]) + a[3];
        double ap = (asum - sp) * 1.0;
        double phi = s.pp[c];
        double ss = su + 0.0 * phi * ap;
        s.su[c] = ss;
        double anb = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) anb += a[k] * s.pp[nbr(g, c, k)];
        double res = wet * (anb + ss - ap * phi) / ap;
        s.resp[c] = res;
        double r = anb + ss - ap * phi + SMALL;
        K.ap_w[0][c] = ap;
        K.dinv_w[0][c] = 1.0 / ap;
        K.r[0][c] = r; K.r0[0][c] = r;
        red[0] += r * r;
        red[1] += res * res;
    });
    double *const dst[2] = {K.red.pr[0], K.red.pr[0] + gridDim.x};
    put_partials<2>(red, dst);
    kry_reset(K.sc[0]);
}

// pvfc (flow_imp.F90:856-908): wall ghosts copy pp, gradient of pp, velocity
// and face-flux corrections.
__global__ void k_pvfc1(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (nb_ghost(g, i, j, k) && g.side[k] != 1) s.pp[nbr(g, c, k)] = s.pp[c];   // walls, flux sides
    const double *vars[1] = {s.pp};
    double gx, gy;
    cbgg_grad(g, s.iwet, c, 1, vars, &gx, &gy);
    s.dppx[c] = gx;
    s.dppy[c] = gy;
    if (s.iwet[c] == 0) return;
    double val = s.h[c] * s.apuareap[c] / (1.0 - s.sumu[c] + SMALL);
    s.u[c] = s.u[c] - val * gx;
    s.v[c] = s.v[c] - val * gy;
    const double ppc = s.pp[c];
    int nb = c + g.nxp;
    if (s.iwet[nb]) s.fluxN[c] = s.fluxN[c] - s.aN[c] * (s.pp[nb] - ppc);
    nb = c + 1;
    if (s.iwet[nb]) s.fluxE[c] = s.fluxE[c] - s.aE[c] * (s.pp[nb] - ppc);
    if (j == 0) {
        nb = c - g.nxp;
        if (s.iwet[nb]) { double f = -s.fluxN[nb]; f = f - s.aS[c] * (s.pp[nb] - ppc); s.fluxN[nb] = -f; }
    }
    if (i == 0) {
        nb = c - 1;
        if (s.iwet[nb]) { double f = -s.fluxE[nb]; f = f - s.aW[c] * (s.pp[nb] - ppc); s.fluxE[nb] = -f; }
    }
}

// pvfc (flow_imp.F90:965-969) on every cell incl. ghosts
__global__ void k_pvfc2(Geo g, St s, Par P) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    double pp = s.pp[t];
    double p = s.p[t] + 1.0 * pp;
    s.p[t] = p;
    s.h[t] = fmax(P.hmin, p * GRAVINV - s.zb[t]);
    s.pp[t] = FACPP * pp;
}

// pvfc (flow_imp.F90:972-995): wall ghosts take the pressure of their cell,
// flux-side ghosts also the depth that follows from it
__global__ void k_ghost_p(Geo g, St s, Par P) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int nr = 2 * g.nx + 2 * g.ny;
    if (t >= nr) return;
    int i, j, k;
    if (t < g.nx) { i = t; j = g.ny - 1; k = KN; }
    else if (t < 2 * g.nx) { i = t - g.nx; j = 0; k = KS; }
    else if (t < 2 * g.nx + g.ny) { i = g.nx - 1; j = t - 2 * g.nx; k = KE; }
    else { i = 0; j = t - 2 * g.nx - g.ny; k = KW; }
    if (g.side[k] == 1) return;
    int c = cidx(g, i, j), gc = nbr(g, c, k);
    s.p[gc] = s.p[c];
    if (g.side[k] == 2) s.h[gc] = fmax(P.hmin, s.p[gc] * GRAVINV - s.zb[gc]);
}

// flow_grad_interp (flow_imp.F90:1106-1127) + fric_eval (friction.F90:615-797,
// Manning) + flow_eddyvis cell values (flow_eddyvis.f90:128-160, mturbul=5).
__global__ void k_grad_fric_vis(Geo g, St s, Par P) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const double *vars[2] = {s.u, s.v};
    double gx[2], gy[2];
    cbgg_grad(g, s.iwet, c, 2, vars, gx, gy);
    s.dux[c] = gx[0]; s.duy[c] = gy[0]; s.dvx[c] = gx[1]; s.dvy[c] = gy[1];
    // interp_scal_cell2face(p, ibc=1)
    const double pc = s.p[c];
    const int wc = s.iwet[c];
    double sx = 0.0, sy = 0.0, hkv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int nb = nbr(g, c, k);
        int wn = s.iwet[nb];
        double pk;
        if (wc * wn == 1) pk = g.fintp[k] * s.p[nb] + (1.0 - g.fintp[k]) * pc;
        else if (wc == 1) pk = pc;
        else if (wn == 1) pk = s.p[nb];
        else pk = 0.5 * (pc + s.p[nb]);
        sx += g.dsx[k] * pk;
        sy += g.dsy[k] * pk;
        hkv[k] = fmax(P.hmin, pk * GRAVINV - fsym(s.zbkE, s.zbkN, g, c, k));
    }
    s.dpx[c] = sx / g.area;
    s.dpy[c] = sy / g.area;
    s.hkN[c] = hkv[KN];
    s.hkE[c] = hkv[KE];
    if (j == 0) s.hkN[c - g.nxp] = hkv[KS];
    if (i == 0) s.hkE[c - 1] = hkv[KW];
    // fric_eval
    double h = s.h[c], u = s.u[c], v = s.v[c];
    double cd = fric_law(P, s, c, h);
    double Um, Uwc, taum;
    if (P.noptset >= 3) {
        double Um2 = u * u + v * v;
        Um = sqrt(Um2);
        Uwc = sqrt(Um2 + CFRICWAV * 0.0 * 0.0);
        taum = RHOW * cd * Um * Uwc;
    } else {
        Um = sqrt(u * u + v * v);
        Uwc = Um;
        taum = RHOW * cd * Um * Um;
    }
    double bsvel = sqrt(taum / RHOW + F4(1.0e-15));
    s.uv[c] = Um; s.uelwc[c] = Uwc; s.bsvel[c] = bsvel;
    // eddy viscosity, cell value (flow_eddyvis.f90:23-160): TURBULENCE_MODEL
    // CONSTANT (0), FALCONER (1), PARABOLIC (2), SUBGRID-WU (3), SUBGRID (5);
    // the coefficients keep their mturbul-5 defaults (flow.F90:81-92)
    auto waves = [&](double vv) {
        if (P.noptset >= 3) {
            vv = vv + CVISWAV * 0.0 * 0.0;
            vv = vv + CVISWAVBRK * h * pow(fmax(0.0, F4(1.0e-15)) / RHOW, F4(0.33333));
        }
        return vv;
    };
    double vis;
    if (P.mturbul == 0) {
        vis = CVISCON;
    } else if (P.mturbul == 1) {
        vis = fmin(waves(CVISCON + CVISBOT * cd * Um * h), CVISMAX);
    } else if (P.mturbul == 2) {
        vis = fmin(waves(CVISCON + CVISBOT * bsvel * h), CVISMAX);
    } else if (P.mturbul == 3) {
        if (wc == 1) {
            const double ss12 = 0.5 * (gy[0] + gx[1]);
            const double stens = 2.0 * (gx[0] * gx[0] + 2.0 * ss12 * ss12 + gy[1] * gy[1]);
            const double b = CVISBOT * bsvel * h;
            vis = fmin(waves(CVISCON + sqrt(P.c2a * P.c2a * stens + b * b)), CVISMAX);
        } else {
            vis = F4(1.0e-6);
        }
    } else if (wc == 1) {
        double stens = sqrt(2.0 * (gx[0] * gx[0] + gy[1] * gy[1]) + (gy[0] + gx[1]) * (gy[0] + gx[1]));
        vis = fmin(waves(CVISCON + CVISBOT * bsvel * h + P.c2a * stens), CVISMAX);
    } else {
        vis = F4(1.0e-6);
    }
    s.visb[c] = vis;
    // ghosts: forcing strings copy the cell value (bndcopy2ghost); walls keep theirs
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!nb_ghost(g, i, j, k)) continue;
        int gc = nbr(g, c, k);
        s.visb[gc] = g.side[k] != 0 ? vis : s.vis[gc];
    }
}

// wave_rad_stress_fl (waves/wave_rad_stress_fl.f90:2-59), part 1: the
// radiation stresses (divided by the density) of each active cell from the
// wave height H, length L and unit direction (ux, uy) with the current depth:
//   kh = min(2 pi h / L, 10), n = 1/2 + kh / sinh(2 kh), E = g H^2 / 16,
//   Sxx = E (n (1 + ux^2) - 1/2), Syy = E (n (1 + uy^2) - 1/2), Sxy = E n uy uy
// (the reference's uy*uy, :31).  Ghost cells on the grid edge take their
// cell's values (the reference reads its uninitialised stack array there).
__global__ void k_wave_stress(Geo g, St s, const double *H, const double *L, const double *ux, const double *uy,
                              double *sxx, double *sxy, double *syy) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const double twopi = 2.0 * 3.141592653589793;   // const_def twopi = 2 acos(-1), definitions.F90:86-87
    double kh = twopi / L[c] * s.h[c];
    kh = fmin(kh, 10.0);
    const double cn = 0.5 + kh / sinh(2.0 * kh);
    const double e = 0.0625 * GRAV * H[c] * H[c];
    const double vxx = e * (cn * (1.0 + ux[c] * ux[c]) - 0.5);
    const double vyy = e * (cn * (1.0 + uy[c] * uy[c]) - 0.5);
    const double vxy = e * cn * uy[c] * uy[c];
    sxx[c] = vxx; sxy[c] = vxy; syy[c] = vyy;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!nb_ghost(g, i, j, k)) continue;
        const int gc = nbr(g, c, k);
        sxx[gc] = vxx; sxy[gc] = vxy; syy[gc] = vyy;
    }
}

// Part 2: wavestrx = -dSxx/dx - dSxy/dy, wavestry = -dSxy/dx - dSyy/dy with the
// wet-only gradient; counts components that are NaN or above 10 in bad[0] (the
// reference's error stop, :42-56).
__global__ void k_wave_force(Geo g, St s, const double *sxx, const double *sxy, const double *syy, double *fx,
                             double *fy, int *bad) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const double *vars[3] = {sxx, sxy, syy};
    double gx[3], gy[3];
    cbgg_grad(g, s.iwet, c, 3, vars, gx, gy);
    const double wx = -gx[0] - gy[1], wy = -gx[1] - gy[2];
    fx[c] = wx;
    fy[c] = wy;
    if (isnan(wx) || isnan(wy) || fabs(wx) > 10.0 || fabs(wy) > 10.0) atomicAdd(bad, 1);
}

// flow_eddyvis boundary fix-ups (flow_eddyvis.f90:168-191, Cartesian mesh),
// race-free: every read is of the values before the loop.  A wet cell next to
// a dry cell or a ghost takes the value of its opposite neighbour; that value
// is also written to the dry/ghost neighbour, the last writer in cell order
// (the highest-numbered wet neighbour: N, then E, then W, then S) winning.
__device__ __forceinline__ double vis_from(const Geo &g, const double *vb, int w, int k) {
    // wet cell w, boundary face k -> value from the cell opposite face k
    double vissum = vb[nbr(g, w, c_kopp[k])] - CVISCON;
    return CVISCON + vissum / (1.0 + SMALL);
}
__global__ void k_vis_fix(Geo g, St s) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    int i, j;
    pad_ij(g, t, i, j);
    const uint8_t *w = s.iwet;
    const double *vb = s.visb;
    double val = vb[t];
    if (active_ij(g, i, j)) {
        if (w[t] == 1) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                int nb = nbr(g, t, k);
                if (w[nb] == 0 || nb_ghost(g, i, j, k)) val = vis_from(g, vb, t, k);
            }
        } else {
            // writers in increasing cell order: S, W, E, N -> last wet one wins
            const int order[4] = {KS, KW, KE, KN};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int k = order[q];
                if (nb_ghost(g, i, j, k)) continue;
                int nb = nbr(g, t, k);
                if (w[nb] == 1) val = vis_from(g, vb, nb, c_kopp[k]);
            }
        }
    } else {
        int sd = ring_side(g, i, j);
        if (sd < 0) return;
        int own = nbr(g, t, c_kopp[sd]);
        if (w[own] == 1) val = vis_from(g, vb, own, sd);
    }
    s.vis[t] = val;
}

// corner wet nodes (flow_eddyvis.f90:194-206), reading the fixed-up values
__global__ void k_vis_corner(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (s.icorner[c] != 1) return;
    double cid = 0.0, vs = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int nb = nbr(g, c, k);
        cid += s.iwet[nb];
        vs += (s.vis[nb] - CVISCON) * s.iwet[nb];
    }
    s.visc[c] = CVISCON + vs / (cid + SMALL);
}

// ---------------------------------------------- the reference's cell order
// With serial_order the three loops above follow the reference run on one
// thread: cells in increasing number (the padded index keeps that order for
// active cells; ghost cells are numbered after them), faces N, E, S, W, each
// read seeing every write made before it.  The writes of a loop form a
// recurrence along chains of cells; a pass evaluates every value from the
// previous pass's values (the first pass from the values before the loop),
// and passes repeat until one changes nothing, after which every value
// follows its whole chain.  *flag is set by a pass that changed a value.

__device__ __forceinline__ bool is_act(const Geo &g, int t) {
    int i, j;
    pad_ij(g, t, i, j);
    return active_ij(g, i, j);
}

// Dry-to-wet start values (flow.F90:1477-1491), one pass of one sweep: the
// neighbours that are dry-to-wet cells numbered below c give this sweep's
// values (pu, pv, pvis of the previous pass), the others the values from
// before the sweep.
__global__ void k_d2w_serial(Geo g, St s, const double *pu, const double *pv, const double *pvis, double *nu,
                             double *nv, double *nvis, int first, int *flag) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (!(s.iwet[c] == 1 && s.iwet1[c] == 0)) return;
    double us = 0.0, vs = 0.0, viss = 0.0;
    int ns = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int nb = nbr(g, c, k);
        const bool now = !first && nb < c && is_act(g, nb) && s.iwet[nb] == 1 && s.iwet1[nb] == 0;
        us += now ? pu[nb] : s.u[nb];
        vs += now ? pv[nb] : s.v[nb];
        viss += now ? pvis[nb] : s.vis[nb];
        ns += s.iwet[nb];
    }
    const double den = (double)(float)ns + SMALL;
    const double a = us / den / F4(1.25), b = vs / den / F4(1.25), v = viss / den;
    nu[c] = a;
    nv[c] = b;
    nvis[c] = v;
    if (!first && (a != pu[c] || b != pv[c] || v != pvis[c])) *flag = 1;
}
__global__ void k_d2w_serial_commit(Geo g, St s, const double *nu, const double *nv, const double *nvis) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (!(s.iwet[c] == 1 && s.iwet1[c] == 0)) return;
    s.u[c] = nu[c]; s.v[c] = nv[c]; s.vis[c] = nvis[c];
}

// Eddy-viscosity fix-up next to dry cells (flow_eddyvis.f90:168-191): the
// value written at face k of wet cell c (to c and to the dry or ghost cell
// across it) from the cell o across the opposite face as it stands when the
// loop reaches (c, k).  pf, nf: [4*npad] face-major values of the previous and
// this pass.
__device__ __forceinline__ bool fix_face(const Geo &g, const St &s, int c, int i, int j, int k) {
    return s.iwet[nbr(g, c, k)] == 0 || nb_ghost(g, i, j, k);
}
__device__ __forceinline__ int last_fix_face(const Geo &g, const St &s, int c) {
    int i, j;
    pad_ij(g, c, i, j);
    int kl = -1;
    for (int k = 0; k < 4; k++)
        if (fix_face(g, s, c, i, j, k)) kl = k;
    return kl;
}
// vis of cell o when the loop is at (c, k), from the previous pass's face values
__device__ double vis_serial_read(const Geo &g, const St &s, const double *pf, int o, int c, int k) {
    const double *vb = s.visb;
    const size_t np = (size_t)g.nxp * g.nyp;
    int oi, oj;
    pad_ij(g, o, oi, oj);
    if (active_ij(g, oi, oj) && s.iwet[o] == 1) {
        if (o < c) {
            const int kl = last_fix_face(g, s, o);
            if (kl >= 0) return pf[kl * np + o];
        }
        return vb[o];
    }
    // a dry or ghost cell: the last of its wet neighbours' writes before (c, k)
    const long long now = 4ll * c + k;
    long long best = -1;
    double val = vb[o];
    for (int d = 0; d < 4; d++) {
        const int mi = oi + (d == KE) - (d == KW), mj = oj + (d == KN) - (d == KS);
        if (!active_ij(g, mi, mj)) continue;
        const int m = (mj + 1) * g.nxp + mi + 1;
        if (s.iwet[m] != 1) continue;
        const int km = c_kopp[d];
        const long long key = 4ll * m + km;
        if (key < now && key > best) { best = key; val = pf[km * np + m]; }
    }
    return val;
}
__global__ void k_vis_fix_serial(Geo g, St s, const double *pf, double *nf, int first, int *flag) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (s.iwet[c] != 1) return;
    const size_t np = (size_t)g.nxp * g.nyp;
    for (int k = 0; k < 4; k++) {
        if (!fix_face(g, s, c, i, j, k)) continue;
        const int o = nbr(g, c, c_kopp[k]);
        const double r = first ? s.visb[o] : vis_serial_read(g, s, pf, o, c, k);
        const double v = CVISCON + (r - CVISCON) / (1.0 + SMALL);
        nf[k * np + c] = v;
        if (!first && v != pf[k * np + c]) *flag = 1;
    }
}
// vis after the loop from the settled face values: a wet cell keeps its last
// face's value, a dry or ghost cell its last writer's (the highest-numbered
// wet neighbour), every other cell its value before the loop.
__global__ void k_vis_fix_serial_commit(Geo g, St s, const double *pf) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    const size_t np = (size_t)g.nxp * g.nyp;
    int i, j;
    pad_ij(g, t, i, j);
    double val = s.visb[t];
    if (active_ij(g, i, j) && s.iwet[t] == 1) {
        const int kl = last_fix_face(g, s, t);
        if (kl >= 0) val = pf[kl * np + t];
    } else if (active_ij(g, i, j) || ring_side(g, i, j) >= 0) {
        long long best = -1;
        for (int d = 0; d < 4; d++) {
            const int mi = i + (d == KE) - (d == KW), mj = j + (d == KN) - (d == KS);
            if (!active_ij(g, mi, mj)) continue;
            const int m = (mj + 1) * g.nxp + mi + 1;
            if (s.iwet[m] != 1) continue;
            const int km = c_kopp[d];
            const long long key = 4ll * m + km;
            if (key > best) { best = key; val = pf[km * np + m]; }
        }
    }
    s.vis[t] = val;
}
// Corner wet nodes (flow_eddyvis.f90:194-206): corner cells numbered below c
// give their value of this loop (pc of the previous pass).
__global__ void k_vis_corner_serial(Geo g, St s, const double *pc, double *nc, int first, int *flag) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (s.icorner[c] != 1) return;
    double cid = 0.0, vs = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int nb = nbr(g, c, k);
        const bool now = !first && nb < c && is_act(g, nb) && s.icorner[nb] == 1;
        cid += s.iwet[nb];
        vs += ((now ? pc[nb] : s.vis[nb]) - CVISCON) * s.iwet[nb];
    }
    const double v = CVISCON + vs / (cid + SMALL);
    nc[c] = v;
    if (!first && v != pc[c]) *flag = 1;
}

// interp_scal_cell2face(vis, ibc=3) -> visk on owned faces; commits corners.
__device__ __forceinline__ double vis_final(const Geo &g, const St &s, int x) {
    int j = x / g.nxp - 1, i = x - (j + 1) * g.nxp - 1;
    return (active_ij(g, i, j) && s.icorner[x] == 1) ? s.visc[x] : s.vis[x];
}
__global__ void k_visk(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const uint8_t *w = s.iwet;
    double vc = vis_final(g, s, c);
    int nb = c + g.nxp;
    s.viskN[c] = (w[c] * w[nb] == 1) ? g.fintp[KN] * vis_final(g, s, nb) + (1.0 - g.fintp[KN]) * vc : 0.0;
    nb = c + 1;
    s.viskE[c] = (w[c] * w[nb] == 1) ? g.fintp[KE] * vis_final(g, s, nb) + (1.0 - g.fintp[KE]) * vc : 0.0;
    if (j == 0) {
        nb = c - g.nxp;
        s.viskN[nb] = (w[c] * w[nb] == 1) ? g.fintp[KS] * vis_final(g, s, nb) + (1.0 - g.fintp[KS]) * vc : 0.0;
    }
    if (i == 0) {
        nb = c - 1;
        s.viskE[nb] = (w[c] * w[nb] == 1) ? g.fintp[KW] * vis_final(g, s, nb) + (1.0 - g.fintp[KW]) * vc : 0.0;
    }
    if (s.icorner[c] == 1) s.vis[c] = s.visc[c];
}

// velface_eta (flow_imp.F90:1296-1346, Cartesian) + final fric_eval
__global__ void k_velface(Geo g, St s, Par P) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    double h = s.h[c], u = s.u[c], v = s.v[c];
    if (h > P.hmin && s.iwet[c] == 1) {
        double umin = 1.0e6, umax = -1.0e6, vmin = 1.0e6, vmax = -1.0e6;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double vel = fdir(s.fluxE, s.fluxN, g, c, k) / (fsym(s.hkE, s.hkN, g, c, k) * g.ds[k]);
            double uf = g.fnx[k] * vel, vf = g.fny[k] * vel;
            umin = fmin(umin, uf); umax = fmax(umax, uf);
            vmin = fmin(vmin, vf); vmax = fmax(vmax, vf);
        }
        u = fmin(fmax(u, umin), umax);
        v = fmin(fmax(v, vmin), vmax);
        s.eta[c] = s.p[c] * GRAVINV;
    } else {
        h = fmax(P.hmin, h);
        u = 0.0; v = 0.0;
        s.eta[c] = -999.0;
        s.h[c] = h;
    }
    s.u[c] = u; s.v[c] = v;
    double cd = fric_law(P, s, c, h);
    double Um = sqrt(u * u + v * v), Uwc = Um, taum;
    if (P.noptset >= 3) { Uwc = sqrt(u * u + v * v + CFRICWAV * 0.0 * 0.0); taum = RHOW * cd * Um * Uwc; }
    else taum = RHOW * cd * Um * Um;
    s.uv[c] = Um; s.uelwc[c] = Uwc;
    s.bsvel[c] = sqrt(taum / RHOW + F4(1.0e-15));
}

// flow_update, CMS_Flow.f90:312-350
__global__ void k_update(Geo g, St s) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    s.iwet1[t] = s.iwet[t];
    s.u1[t] = s.u[t]; s.v1[t] = s.v[t]; s.p1[t] = s.p[t]; s.h1[t] = s.h[t];
    s.fluxE1[t] = s.fluxE[t]; s.fluxN1[t] = s.fluxN[t];
    s.pp[t] = 0.0; s.dppx[t] = 0.0; s.dppy[t] = 0.0;
    s.maxeta[t] = fmax(s.eta[t], s.maxeta[t]);   // maximum water level, CMS_Flow.f90:352-355
}

// flow_recalculate (flow_imp.F90:1161-1173): back to the previous time level.
__global__ void k_recalc(Geo g, St s) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    s.u[t] = s.u1[t]; s.v[t] = s.v1[t]; s.p[t] = s.p1[t]; s.h[t] = s.h1[t];
    s.pp[t] = 0.0; s.dppx[t] = 0.0; s.dppy[t] = 0.0;
    s.iwet[t] = s.iwet1[t];
    s.fluxE[t] = s.fluxE1[t]; s.fluxN[t] = s.fluxN1[t];
}

// flow_convflux (flow.F90:1879-1910): face fluxes from cell velocities,
// fintp-weighted normal velocity times hk*ds, zero unless both sides are wet.
__device__ __forceinline__ double conv_face(const Geo &g, const St &s, int c, int nb, int k, double hk) {
    const double uni = g.fnx[k] * s.u[c] + g.fny[k] * s.v[c];
    const double unk = g.fnx[k] * s.u[nb] + g.fny[k] * s.v[nb];
    return (double)(s.iwet[c] * s.iwet[nb]) * g.ds[k] * hk * (g.fintp[k] * unk + (1.0 - g.fintp[k]) * uni);
}
// bndflux (boundaries.F90:4359-4414) on the faces of a flux side
__device__ __forceinline__ void bnd_fluxes(const Geo &g, const St &s, int i, int j, int c) {
    if (j == g.ny - 1 && g.side[KN] == 2) s.fluxN[c] = bnd_flux(g, s, c, KN);
    if (i == g.nx - 1 && g.side[KE] == 2) s.fluxE[c] = bnd_flux(g, s, c, KE);
    if (j == 0 && g.side[KS] == 2) s.fluxN[c - g.nxp] = -bnd_flux(g, s, c, KS);
    if (i == 0 && g.side[KW] == 2) s.fluxE[c - 1] = -bnd_flux(g, s, c, KW);
}
__global__ void k_convflux(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    s.fluxN[c] = conv_face(g, s, c, c + g.nxp, KN, s.hkN[c]);
    s.fluxE[c] = conv_face(g, s, c, c + 1, KE, s.hkE[c]);
    if (j == 0) s.fluxN[c - g.nxp] = -conv_face(g, s, c, c - g.nxp, KS, s.hkN[c - g.nxp]);
    if (i == 0) s.fluxE[c - 1] = -conv_face(g, s, c, c - 1, KW, s.hkE[c - 1]);
    bnd_fluxes(g, s, i, j, c);
}
__global__ void k_bndflux(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    bnd_fluxes(g, s, i, j, c);
}

// bndriver_vel (boundaries.F90:4259-4356) on flux side k: the weights
// iwet(i) iwet(ghost) hk^cmvel / max(n, 1e-5) along (cos, sin) on the ghosts,
// their flux sum sumw = sum ds hk (u fnx + v fny) in string order (one
// thread, as the reference rounds it), then the ghosts' velocities and the
// faces' fluxes scaled to the string's inflow qflux.  One workgroup; the
// string's terms in LDS (n doubles).
__global__ void __launch_bounds__(1024) k_river(Geo g, St s, int k, double qflux, double cmvel, double qcos,
                                                double qsin, const double *qdev) {
    extern __shared__ double term[];
    __shared__ double qnorm;
    const int n = (k == KN || k == KS) ? g.nx : g.ny;
    auto cell = [&](int m) {
        const int i = k == KE ? g.nx - 1 : k == KW ? 0 : m;
        const int j = k == KN ? g.ny - 1 : k == KS ? 0 : m;
        return cidx(g, i, j);
    };
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        const int c = cell(m), gc = nbr(g, c, k);
        const double hk = fsym(s.hkE, s.hkN, g, c, k);
        const double val = (double)(s.iwet[c] * s.iwet[gc]) * pow(hk, cmvel) / fmax(s.coefman[c], F4(1.0e-5));
        const double ug = val * qcos, vg = val * qsin;
        s.u[gc] = ug;
        s.v[gc] = vg;
        term[m] = g.ds[k] * hk * (ug * g.fnx[k] + vg * g.fny[k]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // sequential sum in string order; terms read 16 at a time so the LDS loads overlap
        constexpr int U = 16;
        double sumw = 0.0;
        int m = 0;
        for (; m + U <= n; m += U) {
            double t[U];
#pragma unroll
            for (int q = 0; q < U; q++) t[q] = term[m + q];
#pragma unroll
            for (int q = 0; q < U; q++) sumw = sumw + t[q];
        }
        for (; m < n; m++) sumw = sumw + term[m];
        const double q = qdev ? qdev[0] : qflux;      // a rating curve's inflow of this step
        qnorm = q / fmax(fabs(sumw), F4(1.0e-6));
    }
    __syncthreads();
    const double qn = qnorm;
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        const int c = cell(m), gc = nbr(g, c, k);
        const double hk = fsym(s.hkE, s.hkN, g, c, k);
        const double ug = s.u[gc] * qn, vg = s.v[gc] * qn;
        s.u[gc] = ug;
        s.v[gc] = vg;
        const double f = (double)(s.iwet[c] * s.iwet[gc]) * g.ds[k] * hk * (ug * g.fnx[k] + vg * g.fny[k]);
        if (k == KN) s.fluxN[c] = f;
        else if (k == KE) s.fluxE[c] = f;
        else if (k == KS) s.fluxN[c - g.nxp] = -f;
        else s.fluxE[c - 1] = -f;
    }
}

// interp_scal_cell2face(zb, ibc=0) on owned faces (prestart.f90:320)
__global__ void k_zbk(Geo g, St s) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    s.zbkN[c] = g.fintp[KN] * s.zb[c + g.nxp] + (1.0 - g.fintp[KN]) * s.zb[c];
    s.zbkE[c] = g.fintp[KE] * s.zb[c + 1] + (1.0 - g.fintp[KE]) * s.zb[c];
    if (j == 0) s.zbkN[c - g.nxp] = g.fintp[KS] * s.zb[c - g.nxp] + (1.0 - g.fintp[KS]) * s.zb[c];
    if (i == 0) s.zbkE[c - 1] = g.fintp[KW] * s.zb[c - 1] + (1.0 - g.fintp[KW]) * s.zb[c];
}

// Stage-flow (rating) curve of flux side k (bnd_eval, boundaries.F90:4084-4094):
// the mean depth of the string's cells (summed in string order), then
// plagr_fit of order 1 from a bisection search over the m stages (constant
// beyond the table), times the ramp; q[0] <- the step's inflow, the table in
// q[1..m] (stages) and q[1+m..2m] (flows).
__global__ void k_river_rating(Geo g, St s, int k, int m, double ramp, double *q) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int n = (k == KN || k == KS) ? g.nx : g.ny;
    double stage = 0.0;
    for (int t = 0; t < n; t++) {
        const int i = k == KE ? g.nx - 1 : k == KW ? 0 : t;
        const int j = k == KN ? g.ny - 1 : k == KS ? 0 : t;
        stage = stage + s.h[cidx(g, i, j)];
    }
    stage = stage / (double)n;
    const double *x = q + 1, *v = q + 1 + m;
    double w;
    if (stage < x[0]) w = v[0];
    else if (stage > x[m - 1]) w = v[m - 1];
    else {
        int i = 1, jj = m;
        while (jj > i + 1) {
            const int kk = (i + jj) / 2;
            if (stage < x[kk - 1]) jj = kk;
            else i = kk;
        }
        double lb0 = 1.0, lb1 = 1.0;
        lb0 = lb0 * (stage - x[i]) / (x[i - 1] - x[i]);
        lb1 = lb1 * (stage - x[i - 1]) / (x[i] - x[i - 1]);
        const double sum = (0.0 + lb0) + lb1;
        w = 0.0;
        w += lb0 / sum * v[i - 1];
        w += lb1 / sum * v[i];
    }
    q[0] = ramp * w;
}

__global__ void k_set_int(int *p, int n, int v) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

}  // namespace cms
