// tel_kernels.hip -- the implicit step on a telescoping cell list (read_tel,
// source/geospatial/geospatial.F90:403-795): every cell has up to six faces,
// a coarse cell facing two finer cells on a side.  One thread per cell; face
// arrays are slot-major, a[k*nd + i] for face k of cell i, so the threads of a
// wavefront read slot k of consecutive cells together.  A face shared by two
// active cells is evaluated by its owner (the cell that lists it first in the
// reference's no-repeat list kxyface, geospatial.F90:1528-1552) and written to
// both directed entries.  Face geometry (ds, dn, fintp, r, rp) is recomputed
// from the centres and sizes of the two cells, as geo_init_cart computes it
// (:1469-1611).  Joint cells (a neighbour more than 20 % larger or smaller,
// :1490-1505) take the skewness terms of the reference: reconstructed fluxes
// (flow_imp_flux.f90:2-169), deferred momentum and pp corrections
// (computational.F90:149-264, flow_imp.F90:718-789), pvfc flux fixes
// (:912-962), reconstructed face values with a monotonicity clip
// (interp_lib.f90:2853-2873) and deferred gradient corrections
// (der_lib.f90:381-504,726-839).
#pragma once
#include "flow_state.h"

namespace cms {

constexpr int TF = 6;   // nmaxfaces of a telescoping grid (geospatial.F90:433)

// Per-cell bits: ncface, the face directions (0..3 = N,E,S,W), llec2llec, the
// faces the cell owns, joint, and the multigrid smoothing flags.
__device__ __forceinline__ int m_ncf(uint64_t m) { return (int)(m & 7u); }
__device__ __forceinline__ int m_dir(uint64_t m, int k) { return (int)((m >> (3 + 2 * k)) & 3u); }
__device__ __forceinline__ int m_llec(uint64_t m, int k) { return (int)((m >> (15 + 3 * k)) & 7u); }
__device__ __forceinline__ bool m_own(uint64_t m, int k) { return (m >> (33 + k)) & 1u; }
__device__ __forceinline__ bool m_joint(uint64_t m) { return (m >> 39) & 1u; }
__device__ __forceinline__ bool m_smooth(uint64_t m) { return (m >> 40) & 1u; }
__device__ __forceinline__ bool m_near(uint64_t m) { return (m >> 41) & 1u; }
__device__ __forceinline__ bool m_str(uint64_t m, int k) { return (m >> (42 + k)) & 1u; }
__device__ __forceinline__ bool m_qstr(uint64_t m, int k) { return (m >> (48 + k)) & 1u; }   // flux string

// Topology and geometry (device pointers).
struct TelG {
    int n, nd, ng;               // active cells, cells + ghosts, ghosts
    const int *c2c;              // [TF*nd] slot-major, -1 none
    const uint64_t *meta;        // [n]
    const double *x, *y, *dx, *dy;   // [nd]
    const int *gown, *gk;        // [ng]: owner cell and its face of ghost n+g
    const uint8_t *gwall;        // [ng]: 1 wall ghost, 0 water-level ghost, 2 flux-string ghost
    const int *base;             // [n]: base cell j*nbx+i
    const int *bstart;           // [nbx*nby+1]: cells of base cell b
    int nbx, nby;
    double dxb, dyb, areaavg;
    const int *joint;            // joint cells, ascending
    int njoint;
    int skewcor;                 // skewness corrections (and the gradient corrections) on
    const int *str_c, *str_k;    // water-level string, string order
    int nstr;
    const int *q_c, *q_k;        // flux string, string order
    int nq;
};
__device__ __forceinline__ int tnb(const TelG &T, int i, int k) { return T.c2c[k * T.nd + i]; }

// Geometry of face k of cell i toward nb in direction d (geo_init_cart
// :1569-1611, interp_coef_cell2face interp_lib.f90:2686-2704).
struct FGeo {
    double ds, dn, fnx, fny, rx, ry, rpx, rpy, dsx, dsy, dsxy, fintp;
};
__device__ __forceinline__ FGeo fgeo(const TelG &T, int i, int nb, int d) {
    FGeo f;
    const double xi = T.x[i], yi = T.y[i], dxi = T.dx[i], dyi = T.dy[i];
    const double xn = T.x[nb], yn = T.y[nb], dxn = T.dx[nb], dyn = T.dy[nb];
    if (d == KN || d == KS) f.ds = fmin(dxi, dxn);
    else f.ds = fmin(dyi, dyn);
    switch (d) {
    case KN: f.fnx = 0; f.fny = 1; f.dn = fabs(yn - yi); f.rx = xn - xi; f.ry = dyi / 2.0; break;
    case KE: f.fnx = 1; f.fny = 0; f.dn = fabs(xn - xi); f.rx = dxi / 2.0; f.ry = yn - yi; break;
    case KS: f.fnx = 0; f.fny = -1; f.dn = fabs(yn - yi); f.rx = xn - xi; f.ry = -dyi / 2.0; break;
    default: f.fnx = -1; f.fny = 0; f.dn = fabs(xn - xi); f.rx = -dxi / 2.0; f.ry = yn - yi; break;
    }
    f.rpx = f.rx - f.fnx * fabs(f.rx);
    f.rpy = f.ry - f.fny * fabs(f.ry);
    f.dsx = f.fnx * f.ds;
    f.dsy = f.fny * f.ds;
    f.dsxy = f.ds / f.dn;
    if (d == KN || d == KS) f.fintp = dyi / (dyi + dyn);
    else f.fintp = dxi / (dxi + dxn);
    return f;
}
// The perpendicular offset rp used by the skewness terms of face (i, k):
// rpy on x faces, rpx on y faces.  Ghost faces carry none.
__device__ __forceinline__ double rp_of(const FGeo &f, int d) { return (d == KE || d == KW) ? f.rpy : f.rpx; }

// Device state of a telescoping flow.
struct TelS {
    double *u, *v, *p, *h, *pp, *zb, *vis, *visb, *visc, *u1, *v1, *p1, *h1, *eta, *maxeta;
    double *coefman, *cfrict, *z0, *uv, *uelwc, *bsvel, *cbf, *fc, *wavestrx, *wavestry;
    double *su, *sv, *sp, *spu, *spv, *ssu0, *ssv0, *spuv0, *sppp0, *sspp0;
    double *apuareap, *sumu, *Hu, *Hv, *resu, *resv, *resp, *dpx, *dpy, *tu, *tv, *tvis;
    double *flux, *flux1, *hk, *pk, *visk, *zbk, *am, *apc;   // [TF*nd]
    double *wseadj;              // [nstr]
    uint8_t *iwet, *iwet1, *icorner, *ipond;
    int *plab, *flags;
};

// Cell-gradient pair with the previous evaluation kept for the deferred
// corrections of joint cells (the reference reads the output arrays' old
// values, der_lib.f90:755-770): gx, gy are written, gxm, gym are read.
struct GradIO {
    const double *gxm, *gym;
    double *gx, *gy;
};

__device__ __forceinline__ bool tel_thread(int n, int &i) {
    i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    return i < n;
}

// --------------------------------------------------------------- gradients
// der_grad_cbgg (der_lib.f90:272-378) with screen_weights (:681-715) for cell
// i, wet-only (ibc 1) or all cells (ibc 0); then, on joint cells with the
// corrections on, der_grad_cbggcr (:381-504) applied to the previous
// perpendicular derivative.  NV fields at once.
template <int NV>
__device__ __forceinline__ void tel_grad(const TelG &T, const uint8_t *iwet, int i, int ibc, const double *const *var,
                                         const GradIO *io, double (&gx)[NV], double (&gy)[NV]) {
    const uint64_t m = T.meta[i];
    const int nf = m_ncf(m);
#pragma unroll
    for (int q = 0; q < NV; q++) { gx[q] = 0.0; gy[q] = 0.0; }
    if (iwet[i] == 0 && ibc == 1) return;
    const double area = T.dx[i] * T.dy[i];
    double wx[TF + 1], wy[TF + 1];
    int ix[TF + 1];
    wx[0] = 0.0; wy[0] = 0.0; ix[0] = i;
#pragma unroll
    for (int k = 0; k < TF; k++) {
        wx[k + 1] = 0.0; wy[k + 1] = 0.0; ix[k + 1] = i;
        if (k >= nf) continue;
        const int nb = tnb(T, i, k), d = m_dir(m, k);
        const FGeo f = fgeo(T, i, nb, d);
        ix[k + 1] = nb;
        if (iwet[nb] == 1 || ibc == 0) {
            wx[0] += f.dsx * (1.0 - f.fintp);
            wx[k + 1] = f.dsx * f.fintp;
            wy[0] += f.dsy * (1.0 - f.fintp);
            wy[k + 1] = f.dsy * f.fintp;
        } else {
            wx[0] += f.dsx;
            wy[0] += f.dsy;
        }
    }
    const double tolw = F4(1.0e-8);
    int ncx = 0, ncy = 0;
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int k = 0; k <= TF; k++) {
        if (k > nf) continue;
        wx[k] /= area; wy[k] /= area;
        if (fabs(wx[k]) > tolw) { ncx++; sx += wx[k]; }
        if (fabs(wy[k]) > tolw) { ncy++; sy += wy[k]; }
    }
    // significant weights less their mean sum (zero gradient of a constant)
    const double cx = ncx ? sx / (double)ncx : 0.0, cy = ncy ? sy / (double)ncy : 0.0;
#pragma unroll
    for (int k = 0; k <= TF; k++) {
        if (k > nf) continue;
        if (fabs(wx[k]) > tolw) {
            const double w = wx[k] - cx;
#pragma unroll
            for (int q = 0; q < NV; q++) gx[q] += w * var[q][ix[k]];
        }
        if (fabs(wy[k]) > tolw) {
            const double w = wy[k] - cy;
#pragma unroll
            for (int q = 0; q < NV; q++) gy[q] += w * var[q][ix[k]];
        }
    }
    if (!(T.skewcor && m_joint(m))) return;
    // der_grad_cbggcr: x faces to a neighbour of another height, y faces to
    // one of another width; screened like the Green-Gauss weights
    double wxy[TF + 1], wyx[TF + 1];
    int igx[TF + 1], igy[TF + 1];
#pragma unroll
    for (int k = 0; k <= TF; k++) { wxy[k] = 0.0; wyx[k] = 0.0; igx[k] = i; igy[k] = i; }
#pragma unroll
    for (int k = 0; k < TF; k++) {
        if (k >= nf || !m_own(m, k)) continue;
        const int nb = tnb(T, i, k), d = m_dir(m, k);
        const bool xd = d == KE || d == KW;
        if (!xd) igy[k + 1] = nb;
        if (nb >= T.n) continue;
        if (xd ? fabs(T.dy[i] / T.dy[nb] - 1.0) < F4(0.1) : fabs(T.dx[i] / T.dx[nb] - 1.0) < F4(0.1)) continue;
        if (iwet[nb] == 0 && ibc == 1) continue;
        const FGeo f = fgeo(T, i, nb, d);
        const FGeo b = fgeo(T, nb, i, (d + 2) & 3);
        if (xd) {
            wxy[0] += f.dsx * (1.0 - f.fintp) * f.rpy / area;
            wxy[k + 1] = f.dsx * f.fintp * b.rpy / area;
            igx[k + 1] = nb;
        } else {
            wyx[0] += f.dsy * (1.0 - f.fintp) * f.rpx / area;
            wyx[k + 1] = f.dsy * f.fintp * b.rpx / area;
        }
    }
#pragma unroll
    for (int k = 0; k <= TF; k++) {
        if (k > nf) continue;
        if (fabs(wxy[k]) > tolw)
#pragma unroll
            for (int q = 0; q < NV; q++) gx[q] += wxy[k] * io[q].gym[igx[k]];
        if (fabs(wyx[k]) > tolw)
#pragma unroll
            for (int q = 0; q < NV; q++) gy[q] += wyx[k] * io[q].gxm[igy[k]];
    }
}

// Gradients of NV fields with the wet-only operator, written through io.
template <int NV>
__global__ void k_tl_grad(TelG T, const uint8_t *iwet, int ibc, const double *v0, const double *v1, const double *v2,
                          GradIO io0, GradIO io1, GradIO io2) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const double *var[3] = {v0, v1, v2};
    const GradIO io[3] = {io0, io1, io2};
    double gx[NV], gy[NV];
    tel_grad<NV>(T, iwet, i, ibc, var, io, gx, gy);
#pragma unroll
    for (int q = 0; q < NV; q++) { io[q].gx[i] = gx[q]; io[q].gy[i] = gy[q]; }
}

// ------------------------------------------------------------ wet and dry
// flow_wetdry(0) first part (flow.F90:1250-1275) on every cell; wall ghosts dry,
// the water-level string's ghosts take their cell's flag (:1606-1613)
__global__ void k_tl_wetdry_h(TelG T, TelS s, Par P) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.nd) return;
    double h = s.p[t] * GRAVINV - s.zb[t];
    uint8_t w;
    if (h < P.hdry) { h = fmax(h, P.hmin); w = 0; }
    else w = 1;
    if (t >= T.n) {
        const int c = T.gown[t - T.n];
        w = T.gwall[t - T.n] == 1 ? 0 : (s.p[c] * GRAVINV - s.zb[c] < P.hdry ? 0 : 1);
    }
    s.h[t] = h;
    s.iwet[t] = w;
}

// Narrow-channel removal (flow.F90:1342-1455) in place until no cell changes:
// every rule dries a cell with too few wet neighbours and none re-wets, so the
// repeated parallel passes reach the sequential sweeps' fixed point.
__device__ __forceinline__ int tl_wdir(const TelG &T, const uint8_t *w, int c, int d) {
    if (c >= T.n) return 0;                          // a ghost lists no directions
    const uint64_t m = T.meta[c];
    int s = 0;
    for (int k = 0; k < m_ncf(m); k++)
        if (m_dir(m, k) == d) s += w[tnb(T, c, k)];
    return s;
}
__device__ __forceinline__ int tl_wdir2(const TelG &T, const uint8_t *w, int c, int d) {
    const uint64_t m = T.meta[c];
    int s = 0;
    for (int k = 0; k < m_ncf(m); k++)
        if (m_dir(m, k) == d) s += tl_wdir(T, w, tnb(T, c, k), d);
    return s;
}
__global__ void k_tl_narrow(TelG T, TelS s, Par P) {
    int i;
    if (!tel_thread(T.n, i)) return;
    uint8_t *w = s.iwet;
    if (w[i] == 0) return;
    const uint64_t m = T.meta[i];
    int ksum = 0;
    for (int k = 0; k < m_ncf(m); k++) ksum += w[tnb(T, i, k)];
    const double h = s.h[i];
    bool dry = ksum <= 1;
    const int kns = tl_wdir(T, w, i, KN) + tl_wdir(T, w, i, KS);
    const int kew = tl_wdir(T, w, i, KE) + tl_wdir(T, w, i, KW);
    dry = dry || ((kns == 0 || kew == 0) && h < P.hdry1);
    dry = dry || (tl_wdir(T, w, i, KN) + tl_wdir2(T, w, i, KS) == 0 && h < P.hdry2);
    dry = dry || (tl_wdir(T, w, i, KE) + tl_wdir2(T, w, i, KW) == 0 && h < P.hdry2);
    dry = dry || (tl_wdir(T, w, i, KW) + tl_wdir2(T, w, i, KE) == 0 && h < P.hdry2);
    if (dry) {
        w[i] = 0;
        s.flags[0] = 1;
    }
}

// Ponding (flow_pond, flow.F90:1807-1876): wet cells in a connected component
// without a cell of the forcing string (or its ghost) are dried.  Components
// by union-find over the faces whose two cells are wet.
__device__ __forceinline__ int tl_find(int *P, int x) {
    for (;;) {
        int y = __hip_atomic_load(&P[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (y == x) return x;
        x = y;
    }
}
__device__ __forceinline__ void tl_union(int *P, int a, int b) {
    for (;;) {
        a = tl_find(P, a);
        b = tl_find(P, b);
        if (a == b) return;
        if (a < b) { int t = a; a = b; b = t; }
        int old = atomicMin(&P[a], b);
        if (old == a) return;
        a = old;
    }
}
// Tiles of base cells (each cell, ghosts with their owner, in one tile):
// components inside a tile are labelled in LDS first, so only the faces that
// cross a tile edge take global unions.  Labels decrease towards the smallest
// cell number of the component whatever order the unions land in.
struct TelTiles {
    const int *tstart, *tcell;   // cells of each tile, ascending [ntile+1], [nd]
    const int *tile, *loc;       // tile of each cell and its place in the tile's list [nd]
};
__device__ __forceinline__ bool tl_pond_link(const TelG &T, const uint8_t *w, int i, int nb) {
    return w[i] && w[nb] && (nb >= T.n || nb > i);
}
__device__ __forceinline__ int tl_lds_find(int *lab, int x) {
    for (;;) {
        int y = __hip_atomic_load(&lab[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (y == x) return x;
        x = y;
    }
}
__device__ __forceinline__ void tl_lds_union(int *lab, int a, int b) {
    for (;;) {
        a = tl_lds_find(lab, a);
        b = tl_lds_find(lab, b);
        if (a == b) return;
        if (a < b) { int t = a; a = b; b = t; }
        int old = atomicMin(&lab[a], b);
        if (old == a) return;
        a = old;
    }
}
// One workgroup per tile (dynamic LDS: one int per cell of the largest tile).
__global__ void __launch_bounds__(1024) k_tl_pond_local(TelG T, TelS s, TelTiles Q) {
    extern __shared__ int lab[];
    const int t = blockIdx.x, q0 = Q.tstart[t], m = Q.tstart[t + 1] - q0;
    const uint8_t *w = s.iwet;
    for (int q = threadIdx.x; q < m; q += blockDim.x) lab[q] = w[Q.tcell[q0 + q]] ? q : -1;
    __syncthreads();
    for (int q = threadIdx.x; q < m; q += blockDim.x) {
        const int i = Q.tcell[q0 + q];
        if (i >= T.n || !w[i]) continue;
        const uint64_t mt = T.meta[i];
        for (int k = 0; k < m_ncf(mt); k++) {
            const int nb = tnb(T, i, k);
            if (Q.tile[nb] == t && tl_pond_link(T, w, i, nb)) tl_lds_union(lab, q, Q.loc[nb]);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < m; q += blockDim.x) {
        const int i = Q.tcell[q0 + q];
        s.plab[i] = lab[q] >= 0 ? Q.tcell[q0 + tl_lds_find(lab, q)] : -1;
        s.ipond[i] = 0;
    }
}
// Global unions over the faces that cross a tile edge.
__global__ void k_tl_pond_merge(TelG T, TelS s, TelTiles Q) {
    int i;
    if (!tel_thread(T.n, i)) return;
    if (!s.iwet[i]) return;
    const uint64_t m = T.meta[i];
    for (int k = 0; k < m_ncf(m); k++) {
        const int nb = tnb(T, i, k);
        if (Q.tile[nb] != Q.tile[i] && tl_pond_link(T, s.iwet, i, nb)) tl_union(s.plab, i, nb);
    }
}
__global__ void k_tl_pond_compress(TelG T, TelS s) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.nd) return;
    if (s.plab[t] >= 0) s.plab[t] = tl_find(s.plab, t);
}
__global__ void k_tl_pond_mark(TelG T, TelS s, const int *sc, const int *sk, int ns) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns) return;
    const int i = sc[j], nb = tnb(T, i, sk[j]);
    if (s.iwet[i]) s.ipond[s.plab[i]] = 1;
    if (s.iwet[nb]) s.ipond[s.plab[nb]] = 1;
}
__global__ void k_tl_pond_final(TelG T, TelS s) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.nd) return;
    if (s.iwet[t] == 1 && !s.ipond[s.plab[t]]) s.iwet[t] = 0;
}

// Corner flags, newly wet count, flux and visk zeroed on faces not wet on
// both sides (flow.F90:1459-1525).
__global__ void k_tl_wetdry_post(TelG T, TelS s) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const uint8_t *w = s.iwet;
    const uint64_t m = T.meta[i];
    const int nf = m_ncf(m);
    uint8_t ic = 0;
    if (w[i]) {
        int ndry = 0;
        for (int k = 0; k < nf; k++) ndry += w[tnb(T, i, k)] == 0;
        if (ndry > 1) ic = 1;
    }
    s.icorner[i] = ic;
    if (w[i] == 1 && s.iwet1[i] == 0) atomicAdd(&s.flags[2], 1);
    for (int k = 0; k < nf; k++) {
        const int nb = tnb(T, i, k);
        if (w[i] * w[nb] == 0) {
            s.flux[k * T.nd + i] = 0.0;
            s.visk[k * T.nd + i] = 0.0;
        }
    }
}

// Dry-to-wet initialisation (flow.F90:1477-1491), one sweep reading the
// values from before the sweep.
__global__ void k_tl_d2w_sweep(TelG T, TelS s) {
    int i;
    if (!tel_thread(T.n, i)) return;
    if (!(s.iwet[i] == 1 && s.iwet1[i] == 0)) return;
    const uint64_t m = T.meta[i];
    double us = 0.0, vs = 0.0, viss = 0.0;
    int ns = 0;
    for (int k = 0; k < m_ncf(m); k++) {
        const int nb = tnb(T, i, k);
        us += s.u[nb]; vs += s.v[nb]; viss += s.vis[nb]; ns += s.iwet[nb];
    }
    const double den = (double)(float)ns + SMALL;
    s.tu[i] = us / den / F4(1.25);
    s.tv[i] = vs / den / F4(1.25);
    s.tvis[i] = viss / den;
}
__global__ void k_tl_d2w_commit(TelG T, TelS s) {
    int i;
    if (!tel_thread(T.n, i)) return;
    if (!(s.iwet[i] == 1 && s.iwet1[i] == 0)) return;
    s.u[i] = s.tu[i]; s.v[i] = s.tv[i]; s.vis[i] = s.tvis[i];
}

// ----------------------------------------------------------------- momentum
// flow_srcsnk_static (flow_imp.F90:336-403, ntsch 1)
__global__ void k_tl_srcsnk(TelG T, TelS s, Par P) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const double area = T.dx[i] * T.dy[i];
    const double gdi = 1.0 / (GRAV * P.dtime), dti = 1.0 / P.dtime;
    const double fac1 = area * s.h1[i] * dti;
    double ssu = fac1 * s.u1[i], ssv = fac1 * s.v1[i];
    s.spuv0[i] = -fac1;
    s.sppp0[i] = -area * gdi;
    s.sspp0[i] = -s.p1[i] * s.sppp0[i];
    if (P.noptset >= 3) {
        ssu = ssu + s.wavestrx[i] * area;
        ssv = ssv + s.wavestry[i] * area;
    }
    if (P.tauwx != 0.0 || P.tauwy != 0.0) {
        ssu = ssu + P.tauwx * area;
        ssv = ssv + P.tauwy * area;
    }
    s.ssu0[i] = ssu;
    s.ssv0[i] = ssv;
}

// coeffsourcesink_uv (flow_imp.F90:409-574): sources and every directed
// coefficient of the cell.
__global__ void k_tl_mom(TelG T, TelS s, Par P) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const double area = T.dx[i] * T.dy[i];
    const double h = s.h[i];
    s.su[i] = s.ssu0[i] + (s.fc[i] * s.v[i] - s.dpx[i]) * h * area;
    s.sv[i] = s.ssv0[i] + (-s.fc[i] * s.u[i] - s.dpy[i]) * h * area;
    const double cb = 1.0 * s.cfrict[i] * s.uelwc[i] * area;
    s.sp[i] = s.spuv0[i] - cb;
    const uint64_t m = T.meta[i];
    for (int k = 0; k < m_ncf(m); k++) {
        const int f = k * T.nd + i, nb = tnb(T, i, k);
        const FGeo g = fgeo(T, i, nb, m_dir(m, k));
        const double ddk = s.visk[f] * s.hk[f] * g.dsxy;
        s.am[f] = schmcoef(P.ndsch, ddk, s.flux[f]);
    }
}

// defcorparagradvec (computational.F90:149-264) on joint cells: the deferred
// skewness correction of the momentum sources.  The value of a face is the
// one its owner computes (negated on the other side).
__device__ __forceinline__ bool tl_skew_face(const TelG &T, const TelS &s, const double *dux, const double *duy,
                                             const double *dvx, const double *dvy, int o, int k, double &us,
                                             double &vs) {
    const uint64_t m = T.meta[o];
    if (!m_joint(m) || s.iwet[o] == 0) return false;
    const int nb = tnb(T, o, k), d = m_dir(m, k);
    if (nb >= T.n || s.iwet[nb] == 0) return false;
    const FGeo f = fgeo(T, o, nb, d), b = fgeo(T, nb, o, (d + 2) & 3);
    const bool xd = d == KE || d == KW;
    const double rpi = rp_of(f, d), rpn = rp_of(b, d);
    const int fo = k * T.nd + o;
    const double *du = xd ? duy : dux, *dv = xd ? dvy : dvx;
    if (fabs(rpi) > F4(1.0e-6)) {
        const double val = (s.am[fo] + s.flux[fo]) * rpi;
        us = val * du[o]; vs = val * dv[o];
        return true;
    }
    if (fabs(rpn) > F4(1.0e-6)) {
        const double val = -s.am[fo] * rpn;
        us = val * du[nb]; vs = val * dv[nb];
        return true;
    }
    return false;
}
__global__ void k_tl_defcor_skew(TelG T, TelS s, const double *dux, const double *duy, const double *dvx,
                                 const double *dvy) {
    int jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= T.njoint) return;
    const int i = T.joint[jj];
    const uint64_t m = T.meta[i];
    double su = 0.0, sv = 0.0;
    for (int k = 0; k < m_ncf(m); k++) {
        double us = 0.0, vs = 0.0;
        if (m_own(m, k)) {
            tl_skew_face(T, s, dux, duy, dvx, dvy, i, k, us, vs);
        } else {
            const int nb = tnb(T, i, k);
            if (nb < T.n && tl_skew_face(T, s, dux, duy, dvx, dvy, nb, m_llec(m, k), us, vs)) { us = -us; vs = -vs; }
        }
        su += us;
        sv += vs;
    }
    s.su[i] = s.su[i] - su;
    s.sv[i] = s.sv[i] - sv;
}

// bound_uv (boundaries.F90:3310-3490) with bnd_vel_openwse (bnd_vel.f90:35-95)
// on the water-level string faces, wall friction (fric_lib.f90:162-194), dry cells.
__global__ void k_tl_bound_uv(TelG T, TelS s, Par P) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const uint64_t m = T.meta[i];
    const int nf = m_ncf(m);
    double sp = s.sp[i];
    // flux string first (bound_uv's order): bnd_vel_flux (bnd_vel.f90:2-32)
    for (int k = 0; k < nf; k++) {
        if (!m_qstr(m, k)) continue;
        const int f = k * T.nd + i, nb = tnb(T, i, k);
        const FGeo g = fgeo(T, i, nb, m_dir(m, k));
        const double a = schmcoef(1, s.visk[f] * s.hk[f] * g.dsxy, s.flux[f]);
        s.su[i] = s.su[i] + a * s.u[nb];
        s.sv[i] = s.sv[i] + a * s.v[nb];
        sp = sp - a;
        s.am[f] = 0.0;
    }
    for (int k = 0; k < nf; k++) {
        if (!m_str(m, k)) continue;
        const int f = k * T.nd + i, nb = tnb(T, i, k);
        const FGeo g = fgeo(T, i, nb, m_dir(m, k));
        const double ddk = s.visk[f] * s.hk[f] * g.dsxy;
        const double a = schmcoef(1, ddk, s.flux[f]);
        if (fabs(g.fny) > F4(0.9999)) {
            s.u[nb] = s.iwet[i] * s.u[i] * s.h[i] / s.h[nb];
            const double vg = g.fny * s.flux[f] / (g.ds * s.hk[f]);
            s.v[nb] = vg;
            s.sv[i] = s.sv[i] + a * vg;
        } else {
            const double ug = g.fnx * s.flux[f] / (g.ds * s.hk[f]);
            s.u[nb] = ug;
            s.v[nb] = s.iwet[i] * s.v[i] * s.h[i] / s.h[nb];
            s.su[i] = s.su[i] + a * ug;
        }
        sp = sp - a;
        s.am[f] = 0.0;
    }
    const int w = s.iwet[i];
    if (P.wallfric) {
        if (w == 1) {
            double spu = sp, spv = sp;
            for (int k = 0; k < nf; k++) {
                const int f = k * T.nd + i, nb = tnb(T, i, k);
                if (s.iwet[nb] != 0) continue;
                const FGeo g = fgeo(T, i, nb, m_dir(m, k));
                const double z0w = s.z0[i] * 1.0;
                if (fabs(g.fny) > F4(0.999)) spu -= g.ds * s.hk[f] * wall_coef(fabs(g.ry), z0w) * fabs(s.u[i]);
                else spv -= g.ds * s.hk[f] * wall_coef(fabs(g.rx), z0w) * fabs(s.v[i]);
            }
            s.sp[i] = sp; s.spu[i] = spu; s.spv[i] = spv;
        } else {
            s.u[i] = 0.0; s.v[i] = 0.0;
            s.spu[i] = -1.0; s.spv[i] = -1.0; s.sp[i] = -1.0;
            s.su[i] = 0.0; s.sv[i] = 0.0;
        }
    } else {
        s.u[i] = w * s.u[i];
        s.v[i] = w * s.v[i];
        const double spw = w * sp + (double)(w - 1);
        s.sp[i] = spw; s.spu[i] = spw; s.spv[i] = spw;
        s.su[i] = w * s.su[i];
        s.sv[i] = w * s.sv[i];
    }
}

// Cell-list operator for the Krylov kernels: y_i = ap_i z_i - sum_k a_k(i) z_nb
// over active neighbours (z = 0 on ghost cells, pbicgstab.f90:55,76).
template <class CA>
struct TelOpT {
    int nc, nd;
    const int *c2c;
    const uint64_t *meta;
    const CA *a;                 // [TF*nd]
    __device__ __forceinline__ int n() const { return nc; }
    __device__ __forceinline__ Idx at(int t) const {
        Idx x;
        x.c = t;
        x.i = x.j = 0;
        return x;
    }
    // Passes over the list in order: in pass p the blocks take consecutive
    // runs of KBLOCK cells from p * gridDim.x * KBLOCK on, consecutive blocks
    // on one XCD (xcd_block), so the cells a row above and below (a few
    // thousand entries away) are read by that XCD's blocks in the same pass.
    template <class F>
    __device__ __forceinline__ void each(F &&f) const {
        const int b = xcd_block(blockIdx.x, gridDim.x);
        for (int t = b * KBLOCK + (int)threadIdx.x; t < nc; t += (int)gridDim.x * KBLOCK) f(at(t));
    }
    // V: the vectors' type (double; float in the inner solves of the FP32 form)
    template <class V>
    __device__ __forceinline__ V apply(const V *xv, const V *dinv, const Idx &x) const {
        const int c = x.c, nf = m_ncf(meta[c]);
        V sum = V(0);
        for (int k = 0; k < nf; k++) {
            const int nb = c2c[k * nd + c];
            if (nb < nc) sum += (V)a[k * nd + c] * (xv[nb] * dinv[nb]);
        }
        return xv[c] - sum;
    }
    template <class C>
    __device__ __forceinline__ C apply_z(const float *z, const C *ap, const Idx &x) const {
        const int c = x.c, nf = m_ncf(meta[c]);
        C sum = C(0);
        for (int k = 0; k < nf; k++) {
            const int nb = c2c[k * nd + c];
            if (nb < nc) sum += (C)a[k * nd + c] * (C)z[nb];
        }
        return ap[c] * (C)z[c] - sum;
    }
};
using TelOp = TelOpT<double>;
using TelOp32 = TelOpT<float>;   // the inner solve of the FP32 form

// solve prologue for u (n=2) and v (n=3), relax 0.8 (solvers.f90:232-240) and
// the BiCGStab start (pbicgstab.f90:21-31).
__global__ void __launch_bounds__(KBLOCK) k_tl_uv_prologue(TelG T, TelS s, KrySys<2> K, TelOp A) {
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    A.each([&](const Idx &x_) {
        const int c = x_.c;
        const uint64_t m = T.meta[c];
        const int nf = m_ncf(m);
        double asum = 0.0;
        for (int k = 0; k < nf; k++) asum += s.am[k * T.nd + c];
        const double rinv = 1.0 / F4(0.8), relaxm = 1.0 - F4(0.8);
        const int w = s.iwet[c];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            double *phi = q ? s.v : s.u;
            double *ss = q ? s.sv : s.su;
            const double *sp = q ? s.spv : s.spu;
            const double ap = (asum - sp[c]) * rinv;
            const double ph = phi[c];
            const double ssv = ss[c] + relaxm * ph * ap;
            ss[c] = ssv;
            double anb = 0.0;
            for (int k = 0; k < nf; k++) anb += s.am[k * T.nd + c] * phi[tnb(T, c, k)];
            const double res = w * (anb + ssv - ap * ph) / ap;
            const double r = anb + ssv - ap * ph + SMALL;
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

// ------------------------------------------------------- pressure correction
// coefsourcesink_pp, first loop (flow_imp.F90:600-625)
__global__ void k_tl_pp_prep(TelG T, TelS s) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const uint64_t m = T.meta[i];
    double asum = 0.0;
    for (int k = 0; k < m_ncf(m); k++) asum += s.am[k * T.nd + i];
    const int w = s.iwet[i];
    const double apu = w * RELAX_UV / fmax(asum - s.sp[i], F4(1.0e-6));
    const double apa = apu * (T.dx[i] * T.dy[i]);
    s.apuareap[i] = apa;
    s.sumu[i] = asum * apu;
    s.Hu[i] = w * (s.u[i] + apa * s.dpx[i] * s.h[i]);
    s.Hv[i] = w * (s.v[i] + apa * s.dpy[i] * s.h[i]);
}

// pp coefficients (flow_imp.F90:628-647) and Rhie-Chow fluxes
// (flow_imp_flux.f90:2-169 with skewness terms on joint owners, :361-409
// without) of the faces the cell owns, written to both directed entries.
__global__ void k_tl_pp_coef_flux(TelG T, TelS s, const double *dHuy, const double *dHvx, const double *dapux,
                                  const double *dapuy) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const uint64_t m = T.meta[i];
    const bool joint = T.skewcor && m_joint(m);
    for (int k = 0; k < m_ncf(m); k++) {
        if (!m_own(m, k)) continue;
        const int f = k * T.nd + i, nb = tnb(T, i, k), d = m_dir(m, k);
        const FGeo g = fgeo(T, i, nb, d);
        const int b = nb < T.n ? m_llec(m, k) * T.nd + nb : -1;
        const double hk = s.hk[f];
        double a, fl;
        if (s.iwet[i] * s.iwet[nb] == 0) {
            a = 0.0;
            fl = 0.0;
        } else if (nb >= T.n) {
            a = hk * hk * g.ds * s.apuareap[i] / (1.0 - s.sumu[i]) / g.dn;
            fl = hk * g.ds * ((g.fnx * s.Hu[i] + g.fny * s.Hv[i]) - s.apuareap[i] * hk * (s.p[nb] - s.p[i]) / g.dn);
        } else {
            const double fk = g.fintp, fp = 1.0 - fk;
            const double apf0 = fk * s.apuareap[nb] + fp * s.apuareap[i];
            const double suf = fk * s.sumu[nb] + fp * s.sumu[i];
            a = hk * hk * g.ds * apf0 / (1.0 - suf) / g.dn;
            if (joint) {
                const FGeo gb = fgeo(T, nb, i, (d + 2) & 3);
                const bool xd = d == KE || d == KW;
                const double rpi = rp_of(g, d), rpn = rp_of(gb, d);
                const double *dH = xd ? s.Hu : s.Hv;
                const double *ddH = xd ? dHuy : dHvx;
                const double *dap = xd ? dapuy : dapux;
                const double *dp = xd ? s.dpy : s.dpx;
                double apf, Hf, dpf;
                if (fabs(rpi) > F4(1.0e-6)) {
                    apf = fk * s.apuareap[nb] + fp * (s.apuareap[i] + rpi * dap[i]);
                    Hf = fk * dH[nb] + fp * (dH[i] + rpi * ddH[i]);
                    dpf = (s.p[nb] - s.p[i] - rpi * dp[i]) / g.dn;
                } else if (fabs(rpn) > F4(1.0e-6)) {
                    apf = fk * (s.apuareap[nb] + rpn * dap[nb]) + fp * s.apuareap[i];
                    Hf = fk * (dH[nb] + rpn * ddH[nb]) + fp * dH[i];
                    dpf = (s.p[nb] + rpn * dp[nb] - s.p[i]) / g.dn;
                } else {
                    apf = fk * s.apuareap[nb] + fp * s.apuareap[i];
                    Hf = fk * dH[nb] + fp * dH[i];
                    dpf = (s.p[nb] - s.p[i]) / g.dn;
                }
                fl = hk * g.ds * ((xd ? g.fnx : g.fny) * Hf - apf * hk * dpf);
            } else {
                const double Huf = fk * s.Hu[nb] + fp * s.Hu[i];
                const double Hvf = fk * s.Hv[nb] + fp * s.Hv[i];
                fl = hk * g.ds * ((g.fnx * Huf + g.fny * Hvf) - apf0 * hk * (s.p[nb] - s.p[i]) / g.dn);
            }
        }
        if (m_qstr(m, k))      // bndflux (flow_imp.F90:697): the flux string's faces from their ghosts
            fl = (double)s.iwet[i] * g.ds * hk * (g.fnx * s.u[nb] + g.fny * s.v[nb]);
        s.apc[f] = a;
        s.flux[f] = fl;
        if (b >= 0) { s.apc[b] = a; s.flux[b] = -fl; }
        else { s.apc[m_llec(m, k) * T.nd + nb] = a; s.flux[m_llec(m, k) * T.nd + nb] = -fl; }
    }
}

// The pp skewness correction of face k owned by o (flow_imp.F90:733-776).
__device__ __forceinline__ double tl_fcorr(const TelG &T, const TelS &s, const double *dppx, const double *dppy,
                                           int o, int k) {
    const uint64_t m = T.meta[o];
    if (!m_joint(m) || s.iwet[o] == 0) return 0.0;
    const int nb = tnb(T, o, k), d = m_dir(m, k);
    if (s.iwet[nb] == 0 || nb >= T.n) return 0.0;
    const FGeo f = fgeo(T, o, nb, d), b = fgeo(T, nb, o, (d + 2) & 3);
    const bool xd = d == KE || d == KW;
    const double rpi = rp_of(f, d), rpn = rp_of(b, d);
    const double *dpp = xd ? dppy : dppx;
    const double a = s.apc[k * T.nd + o];
    if (fabs(rpi) > F4(1.0e-6)) return -a * rpi * dpp[o];
    if (fabs(rpn) > F4(1.0e-6)) return a * rpn * dpp[nb];
    return 0.0;
}

// pp sources (flow_imp.F90:700-712) + the joint-cell correction (:720-789) +
// bound_pp (boundaries.F90:3493-3646) + solve prologue n=1 + BiCGStab start.
__global__ void __launch_bounds__(KBLOCK) k_tl_pp_src(TelG T, TelS s, Par P, KrySys<1> K, TelOp A,
                                                      const double *dppx, const double *dppy) {
    double red[2] = {0.0, 0.0};
    A.each([&](const Idx &x_) {
        const int c = x_.c;
        const uint64_t m = T.meta[c];
        const int nf = m_ncf(m);
        double sp = s.sppp0[c];
        double fs = 0.0;
        for (int k = 0; k < nf; k++) fs += s.flux[k * T.nd + c];
        double su = s.sspp0[c] - fs + s.p[c] * sp;
        if (T.skewcor && m_joint(m)) {
            double fc = 0.0;
            for (int k = 0; k < nf; k++) {
                if (m_own(m, k)) fc += tl_fcorr(T, s, dppx, dppy, c, k);
                else {
                    const int nb = tnb(T, c, k);
                    fc += nb < T.n ? -tl_fcorr(T, s, dppx, dppy, nb, m_llec(m, k)) : 0.0;
                }
            }
            su = su + F4(0.2) * fc;
        }
        double a[TF];
#pragma unroll
        for (int k = 0; k < TF; k++) a[k] = k < nf ? s.apc[k * T.nd + c] : 0.0;
        const int wet = s.iwet[c];
        if (wet == 0) {
            int sumwet = 0;
            for (int k = 0; k < nf; k++) {
                const int nb = tnb(T, c, k);
                if (nb < T.n) sumwet += s.iwet[nb];
            }
            double aploc = 0.0;
            su = 0.0;
#pragma unroll
            for (int k = 0; k < TF; k++) {
                if (k >= nf) continue;
                const int nb = tnb(T, c, k);
                if (sumwet >= 1) {
                    a[k] = s.iwet[nb];
                    aploc += a[k];
                    su += a[k] * s.p[nb];
                } else if (nb >= T.n) {
                    a[k] = 0.0;
                } else {
                    a[k] = fgeo(T, c, nb, m_dir(m, k)).dsxy;
                    aploc += a[k];
                    su += a[k] * s.p[nb];
                }
            }
            su -= aploc * s.p[c];
            sp = 0.0;
        }
        bool changed = wet == 0;
#pragma unroll
        for (int k = 0; k < TF; k++)      // flux string: closed faces (boundaries.F90:3544-3551)
            if (k < nf && m_qstr(m, k)) { a[k] = 0.0; changed = true; }
#pragma unroll
        for (int k = 0; k < TF; k++) {
            if (k >= nf || !m_str(m, k)) continue;
            const int gc = tnb(T, c, k);
            const double ppg = s.wseadj[gc - T.n] * GRAV - s.p[gc];
            s.pp[gc] = ppg;
            su += a[k] * ppg;
            sp -= a[k];
            a[k] = 0.0;
            changed = true;
        }
        // the coefficients change only on dry cells and cellstring faces
        if (changed) {
#pragma unroll
            for (int k = 0; k < TF; k++)
                if (k < nf) s.apc[k * T.nd + c] = a[k];
        }
        s.sp[c] = sp;
        double asum = 0.0;
#pragma unroll
        for (int k = 0; k < TF; k++)
            if (k < nf) asum += a[k];
        const double ap = (asum - sp) * 1.0;
        const double phi = s.pp[c];
        const double ss = su + 0.0 * phi * ap;
        s.su[c] = ss;
        double anb = 0.0;
#pragma unroll
        for (int k = 0; k < TF; k++)
            if (k < nf) anb += a[k] * s.pp[tnb(T, c, k)];
        const double res = wet * (anb + ss - ap * phi) / ap;
        s.resp[c] = res;
        const double r = anb + ss - ap * phi + SMALL;
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

// bnd_wse_adjust (boundaries.F90:4620-4686) along the string: the set-up from
// the wave and wind forcing integrated outward from the middle cell; the level
// of each string ghost (indexed ghost - n).
__global__ void __launch_bounds__(1024) k_tl_wse_adjust(TelG T, TelS s, Par P, double w, double ramp, int adjust) {
    extern __shared__ double term[];
    const int n = T.nstr, im0 = n / 2 - 1;
    const bool adj = adjust && n >= 3;
    for (int im = threadIdx.x; im < n; im += blockDim.x) {
        double tv = 0.0;
        if (adj && im != im0) {
            const int im1 = im > im0 ? im - 1 : im + 1;
            const int c = T.str_c[im], c1 = T.str_c[im1];
            const int gc = tnb(T, c, T.str_k[im]);
            const FGeo g = fgeo(T, c, gc, m_dir(T.meta[c], T.str_k[im]));
            double fx = 0.0, fx1 = 0.0, fy = 0.0, fy1 = 0.0;
            if (P.noptset >= 3) { fx = s.wavestrx[c]; fx1 = s.wavestrx[c1]; fy = s.wavestry[c]; fy1 = s.wavestry[c1]; }
            if (P.tauwx != 0.0 || P.tauwy != 0.0) {
                fx = fx + P.tauwx; fx1 = fx1 + P.tauwx; fy = fy + P.tauwy; fy1 = fy1 + P.tauwy;
            }
            const double sx = (T.x[c] - T.x[c1]) * (fx + fx1) / (s.h[c] + s.h[c1]) * GRAVINV;
            const double sy = (T.y[c] - T.y[c1]) * (fy + fy1) / (s.h[c] + s.h[c1]) * GRAVINV;
            if (fabs(g.ry) >= F4(0.999)) tv = sx;
            else if (fabs(g.rx) >= F4(0.999)) tv = sy;
            else tv = sx + sy;
        }
        term[im] = tv;
    }
    __syncthreads();
    if (threadIdx.x == 0 && adj) {
        term[im0] = 0.0;
        double acc = 0.0;
        for (int im = im0 + 1; im < n; im++) { acc = acc + term[im]; term[im] = acc; }
        acc = 0.0;
        for (int im = im0 - 1; im >= 0; im--) { acc = acc + term[im]; term[im] = acc; }
    }
    __syncthreads();
    const double base = (1.0 - ramp) * 0.0 + ramp * (w + 0.0);
    for (int im = threadIdx.x; im < n; im += blockDim.x) {
        const int gc = tnb(T, T.str_c[im], T.str_k[im]);
        s.wseadj[gc - T.n] = adj ? base + term[im] : base;
    }
}

// -------------------------------------------------------------------- pvfc
// Wall ghosts take the pp of their cell (flow_imp.F90:859-863).
__global__ void k_tl_ghost_pp(TelG T, TelS s) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= T.ng) return;
    if (T.gwall[g]) s.pp[T.n + g] = s.pp[T.gown[g]];
}
// Velocity and face-flux corrections (flow_imp.F90:885-908) with the joint-cell
// fixes (:912-962) on the faces the cell owns.
__global__ void k_tl_pvfc1(TelG T, TelS s, const double *dppx, const double *dppy) {
    int i;
    if (!tel_thread(T.n, i)) return;
    if (s.iwet[i] == 0) return;
    const double val = s.h[i] * s.apuareap[i] / (1.0 - s.sumu[i] + SMALL);
    s.u[i] = s.u[i] - val * dppx[i];
    s.v[i] = s.v[i] - val * dppy[i];
    const uint64_t m = T.meta[i];
    const bool joint = T.skewcor && m_joint(m);
    const double ppi = s.pp[i];
    for (int k = 0; k < m_ncf(m); k++) {
        if (!m_own(m, k)) continue;
        const int f = k * T.nd + i, nb = tnb(T, i, k);
        if (s.iwet[nb] == 0) continue;
        double fl = s.flux[f] - s.apc[f] * (s.pp[nb] - ppi);
        if (joint && nb < T.n) {
            const int d = m_dir(m, k);
            const FGeo g = fgeo(T, i, nb, d), gb = fgeo(T, nb, i, (d + 2) & 3);
            const bool xd = d == KE || d == KW;
            const double rpi = rp_of(g, d), rpn = rp_of(gb, d);
            const double *dpp = xd ? dppy : dppx;
            if (fabs(rpi) > F4(1.0e-6)) fl = fl + s.apc[f] * rpi * dpp[i];
            else if (fabs(rpn) > F4(1.0e-6)) fl = fl - s.apc[f] * rpn * dpp[nb];
        }
        s.flux[f] = fl;
        s.flux[m_llec(m, k) * T.nd + nb] = -fl;
    }
}
// p, h, pp of every cell (flow_imp.F90:965-969)
__global__ void k_tl_pvfc2(TelG T, TelS s, Par P) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.nd) return;
    const double pp = s.pp[t];
    const double p = s.p[t] + 1.0 * pp;
    s.p[t] = p;
    s.h[t] = fmax(P.hmin, p * GRAVINV - s.zb[t]);
    s.pp[t] = FACPP * pp;
}
// wall ghosts take the pressure of their cell (flow_imp.F90:972-980)
__global__ void k_tl_ghost_p(TelG T, TelS s, Par P) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= T.ng) return;
    if (T.gwall[g]) s.p[T.n + g] = s.p[T.gown[g]];
    if (T.gwall[g] == 2) s.h[T.n + g] = fmax(P.hmin, s.p[T.n + g] * GRAVINV - s.zb[T.n + g]);
}

// -------------------------------------------------- face values, gradients
// interp_scal_cell2face (interp_lib.f90:2727-2906) of the faces the cell owns,
// ibc 0 (zb), 1 (p) or 3 (vis; values given by vf), with the joint-cell
// reconstruction and clip when gradients are given.
template <int IBC>
__device__ __forceinline__ double tl_interp(const TelG &T, const uint8_t *iwet, const double *phi, int i, int k,
                                            int nb, const FGeo &g, bool skew, const double *gx, const double *gy) {
    const double fk = g.fintp;
    const int wi = iwet[i], wn = iwet[nb];
    double val;
    if (IBC == 0) {
        val = fk * phi[nb] + (1.0 - fk) * phi[i];
    } else if (IBC == 1) {
        if (wi * wn == 1) val = fk * phi[nb] + (1.0 - fk) * phi[i];
        else if (wi == 1) val = phi[i];
        else if (wn == 1) val = phi[nb];
        else val = 0.5 * (phi[i] + phi[nb]);
    } else {
        val = wi * wn == 1 ? fk * phi[nb] + (1.0 - fk) * phi[i] : 0.0;
    }
    if (skew) {
        const int d = m_dir(T.meta[i], k);
        // the neighbour's offset; a ghost carries none and no gradient
        double rpn = 0.0, gn = 0.0;
        if (nb < T.n) {
            const FGeo b = fgeo(T, nb, i, (d + 2) & 3);
            rpn = rp_of(b, d);
            gn = (d == KE || d == KW) ? gy[nb] : gx[nb];
        }
        const double gi = (d == KE || d == KW) ? gy[i] : gx[i];
        val = val + g.fintp * rpn * gn + (1.0 - g.fintp) * rp_of(g, d) * gi;
        val = fmin(fmax(val, fmin(phi[i], phi[nb])), fmax(phi[i], phi[nb]));
    }
    return val;
}
// pk of owned faces (flow_grad_interp, flow_imp.F90:1111)
__global__ void k_tl_pk(TelG T, TelS s) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const uint64_t m = T.meta[i];
    const bool skew = T.skewcor && m_joint(m);
    for (int k = 0; k < m_ncf(m); k++) {
        if (!m_own(m, k)) continue;
        const int nb = tnb(T, i, k);
        const FGeo g = fgeo(T, i, nb, m_dir(m, k));
        const double v = tl_interp<1>(T, s.iwet, s.p, i, k, nb, g, skew, s.dpx, s.dpy);
        s.pk[k * T.nd + i] = v;
        s.pk[m_llec(m, k) * T.nd + nb] = v;
    }
}
// zbk of owned faces (prestart.f90:309-329), all-cell operator
__global__ void k_tl_zbk(TelG T, TelS s, const double *gx, const double *gy) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const uint64_t m = T.meta[i];
    const bool skew = T.skewcor && m_joint(m);
    for (int k = 0; k < m_ncf(m); k++) {
        if (!m_own(m, k)) continue;
        const int nb = tnb(T, i, k);
        const FGeo g = fgeo(T, i, nb, m_dir(m, k));
        const double v = tl_interp<0>(T, s.iwet, s.zb, i, k, nb, g, skew, gx, gy);
        s.zbk[k * T.nd + i] = v;
        s.zbk[m_llec(m, k) * T.nd + nb] = v;
    }
}
// dpx, dpy and hk from the face pressures (flow_imp.F90:1112-1125)
__global__ void k_tl_dphk(TelG T, TelS s, Par P) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const uint64_t m = T.meta[i];
    double sx = 0.0, sy = 0.0;
    for (int k = 0; k < m_ncf(m); k++) {
        const int f = k * T.nd + i, nb = tnb(T, i, k);
        const FGeo g = fgeo(T, i, nb, m_dir(m, k));
        const double pk = s.pk[f];
        sx += g.dsx * pk;
        sy += g.dsy * pk;
        s.hk[f] = fmax(P.hmin, pk * GRAVINV - s.zbk[f]);
    }
    const double area = T.dx[i] * T.dy[i];
    s.dpx[i] = sx / area;
    s.dpy[i] = sy / area;
}

// fric_eval (friction.F90:615-797) and the eddy-viscosity cell value
// (flow_eddyvis.f90:23-160) of active cells; water-level ghosts copy the cell
// value (bndcopy2ghost), wall ghosts keep theirs.
__global__ void k_tl_fric_vis(TelG T, TelS s, Par P, const double *dux, const double *duy, const double *dvx,
                              const double *dvy) {
    int c;
    if (!tel_thread(T.n, c)) return;
    const double h = s.h[c], u = s.u[c], v = s.v[c];
    St sv{};
    sv.cfrict = s.cfrict; sv.coefman = s.coefman; sv.z0 = s.z0; sv.cbf = s.cbf; sv.uelwc = s.uelwc;
    const double cd = fric_law(P, sv, c, h);
    double Um, Uwc, taum;
    if (P.noptset >= 3) {
        const double Um2 = u * u + v * v;
        Um = sqrt(Um2);
        Uwc = sqrt(Um2 + CFRICWAV * 0.0 * 0.0);
        taum = RHOW * cd * Um * Uwc;
    } else {
        Um = sqrt(u * u + v * v);
        Uwc = Um;
        taum = RHOW * cd * Um * Um;
    }
    const double bsvel = sqrt(taum / RHOW + F4(1.0e-15));
    s.uv[c] = Um; s.uelwc[c] = Uwc; s.bsvel[c] = bsvel;
    const int wc = s.iwet[c];
    const double gx0 = dux[c], gy0 = duy[c], gx1 = dvx[c], gy1 = dvy[c];
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
            const double ss12 = 0.5 * (gy0 + gx1);
            const double stens = 2.0 * (gx0 * gx0 + 2.0 * ss12 * ss12 + gy1 * gy1);
            const double b = CVISBOT * bsvel * h;
            vis = fmin(waves(CVISCON + sqrt(P.c2a * P.c2a * stens + b * b)), CVISMAX);
        } else {
            vis = F4(1.0e-6);
        }
    } else if (wc == 1) {
        const double stens = sqrt(2.0 * (gx0 * gx0 + gy1 * gy1) + (gy0 + gx1) * (gy0 + gx1));
        vis = fmin(waves(CVISCON + CVISBOT * bsvel * h + P.c2a * stens), CVISMAX);
    } else {
        vis = F4(1.0e-6);
    }
    s.visb[c] = vis;
}
__global__ void k_tl_vis_ghosts(TelG T, TelS s) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= T.ng) return;
    s.visb[T.n + g] = T.gwall[g] == 1 ? s.vis[T.n + g] : s.visb[T.gown[g]];
}

// Eddy-viscosity fix-ups next to dry cells and ghosts (flow_eddyvis.f90:166-186),
// every read from the values before the loop: a wet cell with a dry or ghost
// neighbour takes the mean over its faces on the opposite side, and so does
// that neighbour; a dry cell or ghost written by several wet cells keeps the
// value of the highest-numbered one (the last writer in the reference's order).
__device__ __forceinline__ double tl_vis_from(const TelG &T, const double *vb, int w, int k) {
    const uint64_t m = T.meta[w];
    const int kkdf = (m_dir(m, k) + 2) & 3;
    double cid = 0.0, vissum = 0.0;
    for (int kk = 0; kk < m_ncf(m); kk++)
        if (m_dir(m, kk) == kkdf) {
            cid += 1.0;
            vissum += vb[tnb(T, w, kk)] - CVISCON;
        }
    return CVISCON + vissum / (cid + SMALL);
}
__global__ void k_tl_vis_fix(TelG T, TelS s) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.nd) return;
    const uint8_t *w = s.iwet;
    const double *vb = s.visb;
    double val = vb[t];
    if (t < T.n) {
        const uint64_t m = T.meta[t];
        const int nf = m_ncf(m);
        if (w[t] == 1) {
            for (int k = 0; k < nf; k++) {
                const int nb = tnb(T, t, k);
                if (w[nb] == 0 || nb >= T.n) val = tl_vis_from(T, vb, t, k);
            }
        } else {
            int best = -1, bk = 0;
            for (int k = 0; k < nf; k++) {
                const int nb = tnb(T, t, k);
                if (nb < T.n && w[nb] == 1 && nb > best) { best = nb; bk = m_llec(m, k); }
            }
            if (best >= 0) val = tl_vis_from(T, vb, best, bk);
        }
    } else {
        const int own = T.gown[t - T.n];
        if (w[own] == 1) val = tl_vis_from(T, vb, own, T.gk[t - T.n]);
    }
    s.vis[t] = val;
}
// corner wet cells (flow_eddyvis.f90:194-206), reading the fixed-up values
__global__ void k_tl_vis_corner(TelG T, TelS s) {
    int i;
    if (!tel_thread(T.n, i)) return;
    if (s.icorner[i] != 1) return;
    const uint64_t m = T.meta[i];
    double cid = 0.0, vs = 0.0;
    for (int k = 0; k < m_ncf(m); k++) {
        const int nb = tnb(T, i, k);
        cid += s.iwet[nb];
        vs += (s.vis[nb] - CVISCON) * s.iwet[nb];
    }
    s.visc[i] = CVISCON + vs / (cid + SMALL);
}
// visk = interp(vis, ibc 3) of owned faces with the corner values committed
__device__ __forceinline__ double tl_vis_final(const TelG &T, const TelS &s, int x) {
    return (x < T.n && s.icorner[x] == 1) ? s.visc[x] : s.vis[x];
}
__global__ void k_tl_visk(TelG T, TelS s) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const uint64_t m = T.meta[i];
    const double vc = tl_vis_final(T, s, i);
    for (int k = 0; k < m_ncf(m); k++) {
        if (!m_own(m, k)) continue;
        const int nb = tnb(T, i, k);
        const FGeo g = fgeo(T, i, nb, m_dir(m, k));
        const double v = s.iwet[i] * s.iwet[nb] == 1 ? g.fintp * tl_vis_final(T, s, nb) + (1.0 - g.fintp) * vc : 0.0;
        s.visk[k * T.nd + i] = v;
        s.visk[m_llec(m, k) * T.nd + nb] = v;
    }
}
__global__ void k_tl_vis_commit(TelG T, TelS s) {
    int i;
    if (!tel_thread(T.n, i)) return;
    if (s.icorner[i] == 1) s.vis[i] = s.visc[i];
}

// ------------------------------------------------------------- step end
// velface_eta (flow_imp.F90:1296-1346, Cartesian and telescoping branch) and
// the final fric_eval
__global__ void k_tl_velface(TelG T, TelS s, Par P) {
    int c;
    if (!tel_thread(T.n, c)) return;
    double h = s.h[c], u = s.u[c], v = s.v[c];
    if (h > P.hmin && s.iwet[c] == 1) {
        double umin = 1.0e6, umax = -1.0e6, vmin = 1.0e6, vmax = -1.0e6;
        const uint64_t m = T.meta[c];
        for (int k = 0; k < m_ncf(m); k++) {
            const int f = k * T.nd + c, nb = tnb(T, c, k);
            const FGeo g = fgeo(T, c, nb, m_dir(m, k));
            const double vel = s.flux[f] / (s.hk[f] * g.ds);
            const double uf = g.fnx * vel, vf = g.fny * vel;
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
    St sv{};
    sv.cfrict = s.cfrict; sv.coefman = s.coefman; sv.z0 = s.z0; sv.cbf = s.cbf; sv.uelwc = s.uelwc;
    const double cd = fric_law(P, sv, c, h);
    const double Um = sqrt(u * u + v * v);
    double Uwc = Um, taum;
    if (P.noptset >= 3) { Uwc = sqrt(u * u + v * v + CFRICWAV * 0.0 * 0.0); taum = RHOW * cd * Um * Uwc; }
    else taum = RHOW * cd * Um * Um;
    s.uv[c] = Um; s.uelwc[c] = Uwc;
    s.bsvel[c] = sqrt(taum / RHOW + F4(1.0e-15));
}

// flow_update (CMS_Flow.f90:312-355) on every cell; flux1 of every face slot
__global__ void k_tl_update(TelG T, TelS s) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.nd) return;
    s.iwet1[t] = s.iwet[t];
    s.u1[t] = s.u[t]; s.v1[t] = s.v[t]; s.p1[t] = s.p[t]; s.h1[t] = s.h[t];
    s.pp[t] = 0.0;
    s.maxeta[t] = fmax(s.eta[t], s.maxeta[t]);
#pragma unroll
    for (int k = 0; k < TF; k++) s.flux1[k * T.nd + t] = s.flux[k * T.nd + t];
}

// flow_convflux (flow.F90:1879-1910) on owned faces
__global__ void k_tl_convflux(TelG T, TelS s) {
    int i;
    if (!tel_thread(T.n, i)) return;
    const uint64_t m = T.meta[i];
    for (int k = 0; k < m_ncf(m); k++) {
        if (!m_own(m, k)) continue;
        const int f = k * T.nd + i, nb = tnb(T, i, k);
        const FGeo g = fgeo(T, i, nb, m_dir(m, k));
        const double uni = g.fnx * s.u[i] + g.fny * s.v[i];
        const double unk = g.fnx * s.u[nb] + g.fny * s.v[nb];
        double fl = (double)(s.iwet[i] * s.iwet[nb]) * g.ds * s.hk[f] * (g.fintp * unk + (1.0 - g.fintp) * uni);
        if (m_qstr(m, k)) fl = (double)s.iwet[i] * g.ds * s.hk[f] * (g.fnx * s.u[nb] + g.fny * s.v[nb]);   // bndflux
        s.flux[f] = fl;
        s.flux[m_llec(m, k) * T.nd + nb] = -fl;
    }
}

// bndriver_vel (boundaries.F90:4259-4356) on the flux string of a cell list;
// as k_river: weights on the ghosts, their flux sum in string order in one
// thread, then scaled to the inflow (qdev[0] for a rating curve).
__global__ void __launch_bounds__(1024) k_tl_river(TelG T, TelS s, double qflux, double cmvel, double qcos,
                                                   double qsin, const double *qdev) {
    extern __shared__ double term[];
    __shared__ double qnorm;
    const int n = T.nq;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int c = T.q_c[j], k = T.q_k[j], f = k * T.nd + c, gc = tnb(T, c, k);
        const FGeo g = fgeo(T, c, gc, m_dir(T.meta[c], k));
        const double hk = s.hk[f];
        const double val = (double)(s.iwet[c] * s.iwet[gc]) * pow(hk, cmvel) / fmax(s.coefman[c], F4(1.0e-5));
        const double ug = val * qcos, vg = val * qsin;
        s.u[gc] = ug;
        s.v[gc] = vg;
        term[j] = g.ds * hk * (ug * g.fnx + vg * g.fny);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sumw = 0.0;
        for (int j = 0; j < n; j++) sumw = sumw + term[j];
        qnorm = (qdev ? qdev[0] : qflux) / fmax(fabs(sumw), F4(1.0e-6));
    }
    __syncthreads();
    const double qn = qnorm;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int c = T.q_c[j], k = T.q_k[j], f = k * T.nd + c, gc = tnb(T, c, k);
        const FGeo g = fgeo(T, c, gc, m_dir(T.meta[c], k));
        const double ug = s.u[gc] * qn, vg = s.v[gc] * qn;
        s.u[gc] = ug;
        s.v[gc] = vg;
        const double fl = (double)(s.iwet[c] * s.iwet[gc]) * g.ds * s.hk[f] * (ug * g.fnx + vg * g.fny);
        s.flux[f] = fl;
        s.flux[m_llec(T.meta[c], k) * T.nd + gc] = -fl;
    }
}
// the rating curve's inflow from the flux string's mean depth (as k_river_rating)
__global__ void k_tl_river_rating(TelG T, TelS s, int m, double ramp, double *q) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double stage = 0.0;
    for (int j = 0; j < T.nq; j++) stage = stage + s.h[T.q_c[j]];
    stage = stage / (double)T.nq;
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

__global__ void k_tl_zero(double *p, long n) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

}  // namespace cms
