// adjoint_kernels.hip -- reverse kernels of the implicit step on a uniform
// Cartesian grid (the discrete adjoint of flow_imp, source/flow/flow_imp.F90:
// 29-312, as the forward kernels of flow_kernels.hip compute it).
//
// For every forward kernel K: x -> y of the step there is a kernel here that
// maps the adjoint of y (dL/dy, held in the Ad arrays) to the adjoint of x, at
// the forward values recorded for that point of the step.  Each reverse
// kernel follows its forward kernel's structure line by line: a thread takes
// the outputs its forward thread wrote, reads their adjoints and clears them
// (an overwritten value has no adjoint of its own), and adds the products
// with the partial derivatives to the adjoints of every input that forward
// thread read (atomic adds, since neighbouring threads read the same
// inputs).  Partial derivatives are written out by hand; the few nonlinear
// per-cell formulas (advection scheme, Manning drag, wall coefficient, eddy
// viscosity) are evaluated a second time on dual numbers (value and one
// directional derivative) to obtain their partial derivatives.
//
// Discrete choices are held at the forward run's values: the wet/dry,
// corner and dry-to-wet flags, the number of outer iterations, the branches
// of max/min and clipping.  At an exact tie of a max or min, and at zero for
// an absolute value or a square root of a sum of squares, the derivative is
// the mean of the one-sided ones (what a central difference sees).
#include "flow_state.h"

namespace cms {

// Adjoint arrays, padded like St.  State adjoints (u .. h1) carry between
// kernels and steps; the source adjoints ssu0 .. sspp0 gather one step's
// outer iterations; su .. qsp live within one outer iteration; lu, lv, lp
// are the transposed solves' solutions and tN .. tW their coefficients;
// coefman, zb, zbkE, zbkN gather the parameter gradients.
struct Ad {
    double *u, *v, *p, *h, *pp;
    double *hkE, *hkN, *viskE, *viskN, *fluxE, *fluxN;
    double *dpx, *dpy, *cfrict, *uelwc, *z0;
    double *vis, *visb, *visc, *visf;
    double *u1, *v1, *p1, *h1;
    double *ssu0, *ssv0, *spuv0, *sspp0;
    double *su, *sv, *sp, *spu, *spv;
    double *aN, *aE, *aS, *aW;
    double *apa, *sumu, *Hu, *Hv;
    double *qN, *qE, *qS, *qW, *qsp;
    double *lu, *lv, *lp;
    double *tN, *tE, *tS, *tW;
    double *tu, *tv;
    double *coefman, *zb, *zbkE, *zbkN;
};
constexpr int AD_NARRAYS = sizeof(Ad) / sizeof(double *);

// Values of one outer iteration recorded by the forward run (padded arrays):
// the momentum coefficients and the u, v that k_mom leaves for the solve,
// the solved u, v, the pp-prep outputs, the pp coefficients before and after
// the boundary folding of k_pp_src, its diagonal term, the solved pp and pp
// after pvfc's ghost copies.
struct Mid {
    double *au[4], *sp, *spu, *spv, *u0, *v0, *us, *vs;
    double *apa, *sumu, *Hu, *Hv;
    double *a0[4], *ap[4], *spp, *pps, *ppv;
};
constexpr int MID_NARRAYS = sizeof(Mid) / sizeof(double *);

__device__ __forceinline__ void aadd(double *p, double v) {
    if (v != 0.0) unsafeAtomicAdd(p, v);
}
// face value of cell c in direction k (the array entry fsym/fdir read)
__device__ __forceinline__ double *face_at(double *E, double *N, const Geo &g, int c, int k) {
    return k == KN ? N + c : k == KE ? E + c : k == KS ? N + (c - g.nxp) : E + (c - 1);
}
// sign with which fdir reads that entry
__device__ __forceinline__ double fdir_sign(int k) { return (k == KS || k == KW) ? -1.0 : 1.0; }

// ------------------------------------------------------------ dual numbers
// In a namespace of their own so that they reach the templates below by
// argument-dependent lookup and do not hide the double overloads.
namespace ad {
struct Dl {
    double v, d;
};
__device__ __forceinline__ Dl operator+(Dl a, Dl b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dl operator+(Dl a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dl operator+(double a, Dl b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dl operator-(Dl a, Dl b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dl operator-(Dl a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dl operator-(double a, Dl b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dl operator-(Dl a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dl operator*(Dl a, Dl b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dl operator*(Dl a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dl operator*(double a, Dl b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dl operator/(Dl a, Dl b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ __forceinline__ Dl operator/(Dl a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dl operator/(double a, Dl b) { return {a / b.v, -a * b.d / (b.v * b.v)}; }
__device__ __forceinline__ Dl fmax(Dl a, Dl b) {
    return a.v > b.v ? a : b.v > a.v ? b : Dl{a.v, 0.5 * (a.d + b.d)};
}
__device__ __forceinline__ Dl fmin(Dl a, Dl b) {
    return a.v < b.v ? a : b.v < a.v ? b : Dl{a.v, 0.5 * (a.d + b.d)};
}
__device__ __forceinline__ Dl fmax(Dl a, double b) { return fmax(a, Dl{b, 0.0}); }
__device__ __forceinline__ Dl fmax(double a, Dl b) { return fmax(Dl{a, 0.0}, b); }
__device__ __forceinline__ Dl fmin(Dl a, double b) { return fmin(a, Dl{b, 0.0}); }
__device__ __forceinline__ Dl fmin(double a, Dl b) { return fmin(Dl{a, 0.0}, b); }
__device__ __forceinline__ Dl fabs(Dl a) { return {::fabs(a.v), a.v > 0.0 ? a.d : a.v < 0.0 ? -a.d : 0.0}; }
__device__ __forceinline__ Dl sqrt(Dl a) {
    const double r = ::sqrt(a.v);
    return {r, r > 0.0 ? a.d / (2.0 * r) : 0.0};
}
__device__ __forceinline__ Dl exp(Dl a) {
    const double e = ::exp(a.v);
    return {e, e * a.d};
}
__device__ __forceinline__ Dl log(Dl a) { return {::log(a.v), a.d / a.v}; }
__device__ __forceinline__ Dl pow(Dl a, double e) {
    return {::pow(a.v, e), a.d == 0.0 ? 0.0 : e * ::pow(a.v, e - 1.0) * a.d};
}
}  // namespace ad
using ad::Dl;

// The per-cell formulas of flow_kernels.hip on dual numbers (fric_vis_t and
// wall_coef_t on a generic scalar).
__device__ __forceinline__ Dl schm_dl(int ndsch, Dl dk, Dl fk) {
    Dl absfk = fabs(fk), peck;
    switch (ndsch) {
    case 0:
        return dk * 0.0;
    case 2:
        peck = absfk / fmax(dk, F4(1.0e-6));
        return dk * fmax(0.0, 1.0 - 0.5 * peck) + fmax(0.0, -fk);
    case 3:
        peck = absfk / fmax(fmax(dk, absfk), F4(1.0e-6));
        return dk * (1.0 - pow(peck, 5.0)) + fmax(0.0, -fk);
    case 4: {
        peck = absfk / fmax(fmax(dk, F4(1.e-6)), absfk);
        // P/(e^P - 1); below P = 1e-6, where the guard max(e^P - 1, 1e-6)
        // turns it into a ramp from 1 down to 0 of slope 1e6, its limit 1 - P/2
        const Dl em1 = exp(peck) - 1.0;
        const Dl a = em1.v < F4(1.0e-6) ? 1.0 - 0.5 * peck : peck / em1;
        return dk * a + fmax(0.0, -fk);
    }
    default:
        return dk + fmax(0.0, -fk);
    }
}
// (d/d dk, d/d fk) of schmcoef
__device__ __forceinline__ void schm_grad(int ndsch, double dk, double fk, double &gd, double &gf) {
    gd = schm_dl(ndsch, Dl{dk, 1.0}, Dl{fk, 0.0}).d;
    gf = schm_dl(ndsch, Dl{dk, 0.0}, Dl{fk, 1.0}).d;
}
template <class T>
__device__ __forceinline__ T wall_coef_t(double yp, T y0) {
    T y0lim = fmin(fmax(y0, F4(1.0e-6)), yp * F4(0.9999));
    T t = F4(0.4) / log(yp / y0lim);
    return t * t;
}
// Manning drag and the roughness length it implies (man2drag, drag2length;
// fric_law's default law), then the current speed and the cell eddy
// viscosity of k_grad_fric_vis for TURBULENCE_MODEL mturbul without waves.
template <class T>
__device__ __forceinline__ void fric_vis_t(const Par &P, int wc, T u, T v, T h, T n, T ux, T uy, T vx, T vy, T &cd,
                                           T &z0, T &uelwc, T &vis) {
    T hh = fmax(fmin(10.0, h), F4(0.1) * n);
    cd = GRAV * n * n / pow(hh, F4(0.3333));
    z0 = fmax(h / exp(F4(0.4) / sqrt(cd) + 1.0), F4(1.0e-6));
    T Um = sqrt(u * u + v * v);
    uelwc = Um;
    T taum = RHOW * cd * Um * Um;
    T bsvel = sqrt(taum / RHOW + F4(1.0e-15));
    if (P.mturbul == 0) {
        vis = h * 0.0 + CVISCON;
    } else if (P.mturbul == 1) {
        vis = fmin(CVISCON + CVISBOT * cd * Um * h, CVISMAX);
    } else if (P.mturbul == 2) {
        vis = fmin(CVISCON + CVISBOT * bsvel * h, CVISMAX);
    } else if (P.mturbul == 3) {
        if (wc == 1) {
            const T ss12 = 0.5 * (uy + vx);
            const T stens = 2.0 * (ux * ux + 2.0 * ss12 * ss12 + vy * vy);
            const T b = CVISBOT * bsvel * h;
            vis = fmin(CVISCON + sqrt(P.c2a * P.c2a * stens + b * b), CVISMAX);
        } else {
            vis = h * 0.0 + F4(1.0e-6);
        }
    } else if (wc == 1) {
        T stens = sqrt(2.0 * (ux * ux + vy * vy) + (uy + vx) * (uy + vx));
        vis = fmin(CVISCON + CVISBOT * bsvel * h + P.c2a * stens, CVISMAX);
    } else {
        vis = h * 0.0 + F4(1.0e-6);
    }
}

// Screened cell-based Green-Gauss weights of cbgg_grad: grad var = sum_k
// (cx[k], cy[k]) var[ix[k]] (zero weights for screened-out entries and for
// dry cells).
__device__ __forceinline__ void cbgg_coef(const Geo &g, const uint8_t *iwet, int c, int *ix, double *cx, double *cy) {
    ix[0] = c;
    for (int k = 0; k < 4; k++) ix[k + 1] = nbr(g, c, k);
    for (int k = 0; k < 5; k++) { cx[k] = 0.0; cy[k] = 0.0; }
    if (!iwet[c]) return;
    double wx[5] = {0, 0, 0, 0, 0}, wy[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (iwet[ix[k + 1]] == 1) {
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
    const double mx = nx_ ? sx / (double)nx_ : 0.0, my = ny_ ? sy / (double)ny_ : 0.0;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        if (fabs(wx[k]) > tolw) cx[k] = wx[k] - mx;
        if (fabs(wy[k]) > tolw) cy[k] = wy[k] - my;
    }
}

__global__ void k_fill0(double *p, long n) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

// ------------------------------------------------------ eddy viscosity tail
// k_vis_fix's source of padded cell t: vis[t] = CVISCON + (visb[src] -
// CVISCON) * wt (or visb[t], wt 1); false where k_vis_fix writes nothing.
__device__ __forceinline__ bool vis_fix_src(const Geo &g, const uint8_t *w, int t, int &src, double &wt) {
    int i, j;
    pad_ij(g, t, i, j);
    src = t;
    wt = 1.0;
    const double wf = 1.0 / (1.0 + SMALL);
    if (active_ij(g, i, j)) {
        if (w[t] == 1) {
            for (int k = 0; k < 4; k++) {
                const int nb = nbr(g, t, k);
                if (w[nb] == 0 || nb_ghost(g, i, j, k)) { src = nbr(g, t, c_kopp[k]); wt = wf; }
            }
        } else {
            const int order[4] = {KS, KW, KE, KN};
            for (int q = 0; q < 4; q++) {
                const int k = order[q];
                if (nb_ghost(g, i, j, k)) continue;
                const int nb = nbr(g, t, k);
                if (w[nb] == 1) { src = nbr(g, nb, k); wt = wf; }
            }
        }
        return true;
    }
    const int sd = ring_side(g, i, j);
    if (sd < 0) return false;
    const int own = nbr(g, t, c_kopp[sd]);
    if (w[own] == 1) { src = nbr(g, own, c_kopp[sd]); wt = wf; }
    return true;
}
// end vis -> the corner value (corner cells) or the fixed-up value
__global__ void k_r_tail0(Geo g, St T, Ad A) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    int i, j;
    pad_ij(g, t, i, j);
    const bool act = active_ij(g, i, j);
    if (act && T.icorner[t] == 1) {
        A.visc[t] = A.vis[t];
        A.visf[t] = 0.0;
        A.vis[t] = 0.0;
    } else if (act || ring_side(g, i, j) >= 0) {
        A.visf[t] = A.vis[t];
        A.vis[t] = 0.0;
    }
}
// k_visk: face viscosities from the final cell values
__global__ void k_r_tail1(Geo g, St T, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const uint8_t *w = T.iwet;
    auto vf_add = [&](int x, double b) {
        int xi, xj;
        pad_ij(g, x, xi, xj);
        if (active_ij(g, xi, xj) && T.icorner[x] == 1) aadd(A.visc + x, b);
        else aadd(A.visf + x, b);
    };
    auto face = [&](double *bk, int nb, int k) {
        const double b = *bk;
        *bk = 0.0;
        if (w[c] * w[nb] == 1) {
            vf_add(nb, g.fintp[k] * b);
            vf_add(c, (1.0 - g.fintp[k]) * b);
        }
    };
    face(A.viskN + c, c + g.nxp, KN);
    face(A.viskE + c, c + 1, KE);
    if (j == 0) face(A.viskN + c - g.nxp, c - g.nxp, KS);
    if (i == 0) face(A.viskE + c - 1, c - 1, KW);
}
// k_vis_corner
__global__ void k_r_tail2(Geo g, St T, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (T.icorner[c] != 1) return;
    const double b = A.visc[c];
    A.visc[c] = 0.0;
    double cid = 0.0;
    for (int k = 0; k < 4; k++) cid += T.iwet[nbr(g, c, k)];
    for (int k = 0; k < 4; k++) {
        const int nb = nbr(g, c, k);
        if (T.iwet[nb]) aadd(A.visf + nb, b / (cid + SMALL));
    }
}
// k_vis_fix
__global__ void k_r_tail3(Geo g, St T, Ad A) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    const double b = A.visf[t];
    A.visf[t] = 0.0;
    if (b == 0.0) return;
    int src;
    double wt;
    if (vis_fix_src(g, T.iwet, t, src, wt)) aadd(A.visb + src, wt * b);
}

// ------------------------------------------------------- k_grad_fric_vis
__global__ void k_r_gfv(Geo g, St E, Par P, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    // ghost values of visb
    double bvis = A.visb[c];
    A.visb[c] = 0.0;
    for (int k = 0; k < 4; k++) {
        if (!nb_ghost(g, i, j, k)) continue;
        const int gc = nbr(g, c, k);
        const double bg = A.visb[gc];
        A.visb[gc] = 0.0;
        if (g.side[k] == 1) bvis += bg;
        else aadd(A.vis + gc, bg);   // wall ghosts keep their vis
    }
    // friction and eddy viscosity (per-cell formulas on dual numbers)
    int ix[5];
    double cx[5], cy[5];
    cbgg_coef(g, E.iwet, c, ix, cx, cy);
    double in[8] = {E.u[c], E.v[c], E.h[c], E.coefman[c], 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < 5; k++) {
        in[4] += cx[k] * E.u[ix[k]];
        in[5] += cy[k] * E.u[ix[k]];
        in[6] += cx[k] * E.v[ix[k]];
        in[7] += cy[k] * E.v[ix[k]];
    }
    const double bcd = A.cfrict[c], bue = A.uelwc[c], bz = A.z0[c];
    A.cfrict[c] = 0.0;
    A.uelwc[c] = 0.0;
    double gin[8];
    bool z0set = true;
    for (int q = 0; q < 8; q++) {
        Dl x[8];
        for (int k = 0; k < 8; k++) x[k] = Dl{in[k], k == q ? 1.0 : 0.0};
        Dl cd, z0, ue, vis;
        fric_vis_t<Dl>(P, E.iwet[c], x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], cd, z0, ue, vis);
        z0set = cd.v != 0.0;
        gin[q] = cd.d * bcd + (z0set ? z0.d * bz : 0.0) + ue.d * bue + vis.d * bvis;
    }
    if (z0set) A.z0[c] = 0.0;
    aadd(A.u + c, gin[0]);
    aadd(A.v + c, gin[1]);
    aadd(A.h + c, gin[2]);
    aadd(A.coefman + c, gin[3]);
    for (int k = 0; k < 5; k++) {
        aadd(A.u + ix[k], cx[k] * gin[4] + cy[k] * gin[5]);
        aadd(A.v + ix[k], cx[k] * gin[6] + cy[k] * gin[7]);
    }
    // face pressures -> dpx, dpy and the face depths hk
    const double bdx = A.dpx[c], bdy = A.dpy[c];
    A.dpx[c] = 0.0;
    A.dpy[c] = 0.0;
    const double pc = E.p[c];
    const int wc = E.iwet[c];
    for (int k = 0; k < 4; k++) {
        const int nb = nbr(g, c, k);
        const int wn = E.iwet[nb];
        double pk, cpc, cpn;
        if (wc * wn == 1) { pk = g.fintp[k] * E.p[nb] + (1.0 - g.fintp[k]) * pc; cpn = g.fintp[k]; cpc = 1.0 - g.fintp[k]; }
        else if (wc == 1) { pk = pc; cpc = 1.0; cpn = 0.0; }
        else if (wn == 1) { pk = E.p[nb]; cpc = 0.0; cpn = 1.0; }
        else { pk = 0.5 * (pc + E.p[nb]); cpc = 0.5; cpn = 0.5; }
        double bpk = (g.dsx[k] * bdx + g.dsy[k] * bdy) / g.area;
        const bool owned = k == KN || k == KE || (k == KS && j == 0) || (k == KW && i == 0);
        if (owned) {
            double *bh = face_at(A.hkE, A.hkN, g, c, k);
            const double b = *bh;
            *bh = 0.0;
            const double zbk = k == KN ? E.zbkN[c] : k == KE ? E.zbkE[c] : k == KS ? E.zbkN[c - g.nxp] : E.zbkE[c - 1];
            if (pk * GRAVINV - zbk > P.hmin) {
                bpk += GRAVINV * b;
                aadd(face_at(A.zbkE, A.zbkN, g, c, k), -b);
            }
        }
        aadd(A.p + c, cpc * bpk);
        aadd(A.p + nb, cpn * bpk);
    }
}

// ------------------------------------------------------- pressure correction
__global__ void k_r_ghost_p(Geo g, Ad A) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = 2 * g.nx + 2 * g.ny;
    if (t >= nr) return;
    int i, j, k;
    if (t < g.nx) { i = t; j = g.ny - 1; k = KN; }
    else if (t < 2 * g.nx) { i = t - g.nx; j = 0; k = KS; }
    else if (t < 2 * g.nx + g.ny) { i = g.nx - 1; j = t - 2 * g.nx; k = KE; }
    else { i = 0; j = t - 2 * g.nx - g.ny; k = KW; }
    if (g.side[k] != 0) return;
    const int c = cidx(g, i, j), gc = nbr(g, c, k);
    const double b = A.p[gc];
    A.p[gc] = 0.0;
    aadd(A.p + c, b);
}
// k_pvfc2: p += pp; h = max(hmin, p/g - zb) on every cell
__global__ void k_r_pvfc2(Geo g, St T, const double *ppv, Par P, Ad A) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    const double pn = T.p[t] + 1.0 * ppv[t];
    const double bh = A.h[t];
    A.h[t] = 0.0;
    double b = A.p[t];
    if (pn * GRAVINV - T.zb[t] > P.hmin) {
        b += GRAVINV * bh;
        A.zb[t] -= bh;
    }
    A.p[t] = b;
    A.pp[t] += b;
}
// k_pvfc1: wall ghosts copy pp, gradient of pp, velocity and flux corrections
__global__ void k_r_pvfc1(Geo g, St T, Mid M, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const uint8_t *w = T.iwet;
    const double *pp = M.ppv;
    const double ppc = pp[c];
    int nb = c + g.nxp;
    if (w[nb]) {
        const double bf = A.fluxN[c], d = pp[nb] - ppc, a = M.ap[KN][c];
        aadd(A.pp + nb, -a * bf);
        aadd(A.pp + c, a * bf);
        A.qN[c] += -d * bf;
    }
    nb = c + 1;
    if (w[nb]) {
        const double bf = A.fluxE[c], d = pp[nb] - ppc, a = M.ap[KE][c];
        aadd(A.pp + nb, -a * bf);
        aadd(A.pp + c, a * bf);
        A.qE[c] += -d * bf;
    }
    if (j == 0) {
        nb = c - g.nxp;
        if (w[nb]) {
            const double bf = A.fluxN[nb], d = pp[nb] - ppc, a = M.ap[KS][c];
            aadd(A.pp + nb, a * bf);
            aadd(A.pp + c, -a * bf);
            A.qS[c] += d * bf;
        }
    }
    if (i == 0) {
        nb = c - 1;
        if (w[nb]) {
            const double bf = A.fluxE[nb], d = pp[nb] - ppc, a = M.ap[KW][c];
            aadd(A.pp + nb, a * bf);
            aadd(A.pp + c, -a * bf);
            A.qW[c] += d * bf;
        }
    }
    if (w[c]) {
        int ix[5];
        double cx[5], cy[5];
        cbgg_coef(g, w, c, ix, cx, cy);
        double gx = 0.0, gy = 0.0;
        for (int k = 0; k < 5; k++) { gx += cx[k] * pp[ix[k]]; gy += cy[k] * pp[ix[k]]; }
        const double den = 1.0 - M.sumu[c] + SMALL;
        const double hc = T.h[c], apa = M.apa[c];
        const double val = hc * apa / den;
        const double bu = A.u[c], bv = A.v[c];
        const double bgx = -val * bu, bgy = -val * bv, bval = -(gx * bu + gy * bv);
        aadd(A.h + c, bval * apa / den);
        A.apa[c] += bval * hc / den;
        A.sumu[c] += bval * hc * apa / (den * den);
        for (int k = 0; k < 5; k++) aadd(A.pp + ix[k], cx[k] * bgx + cy[k] * bgy);
    }
    for (int k = 0; k < 4; k++) {
        if (!(nb_ghost(g, i, j, k) && g.side[k] == 0)) continue;
        const int gc = nbr(g, c, k);
        const double b = A.pp[gc];
        A.pp[gc] = 0.0;
        aadd(A.pp + c, b);
    }
}

// Transposed coefficients: row c of A^T holds, for each neighbour nb, the
// coefficient of c in the row of nb (zero for ghost neighbours).
__global__ void k_r_transpose(Geo g, const double *aN, const double *aE, const double *aS, const double *aW, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    A.tN[c] = j < g.ny - 1 ? aS[c + g.nxp] : 0.0;
    A.tE[c] = i < g.nx - 1 ? aW[c + 1] : 0.0;
    A.tS[c] = j > 0 ? aN[c - g.nxp] : 0.0;
    A.tW[c] = i > 0 ? aE[c - 1] : 0.0;
}

// Start of a transposed solve A^T x = rhs on the fixed Krylov grid: x = 0,
// r = r0 = rhs, the diagonal (asum - sp) * rinv of the forward system and
// the partial sums the first k_bicg_p reads.
template <int NS>
struct AdjStart {
    const double *aN, *aE, *aS, *aW;
    const double *sp[NS], *rhs[NS];
    double rinv;
};
template <int NS>
__global__ void __launch_bounds__(KBLOCK) k_r_start(Geo g, KrySys<NS> K, AdjStart<NS> S) {
    double red[2 * NS];
#pragma unroll
    for (int q = 0; q < 2 * NS; q++) red[q] = 0.0;
    grid_walk(g, [&](const Idx &x_) {
        const int c = x_.c;
        const double asum = ((S.aN[c] + S.aE[c]) + S.aS[c]) + S.aW[c];
#pragma unroll
        for (int q = 0; q < NS; q++) {
            const double ap = (asum - S.sp[q][c]) * S.rinv;
            const double r = S.rhs[q][c];
            K.x[q][c] = 0.0;
            K.ap_w[q][c] = ap;
            K.dinv_w[q][c] = 1.0 / ap;
            K.r[q][c] = r;
            K.r0[q][c] = r;
            red[2 * q] += r * r;
            red[2 * q + 1] += r * r;
        }
    });
    double *dst[2 * NS];
#pragma unroll
    for (int q = 0; q < NS; q++) { dst[2 * q] = K.red.pr[q]; dst[2 * q + 1] = K.red.pr[q] + gridDim.x; }
    put_partials<2 * NS>(red, dst);
#pragma unroll
    for (int q = 0; q < NS; q++) kry_reset(K.sc[q]);
}

// Gradients of the pp solve's inputs from lambda = A^-T (dL/dpp): rhs
// lambda, diagonal term lambda pp, coefficient a_k lambda (pp_nb - pp).
__global__ void k_r_ppsolve(Geo g, Mid M, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const double lam = A.lp[c], ppc = M.pps[c];
    A.pp[c] = 0.0;
    A.qsp[c] = lam * ppc;
    A.qN[c] += lam * (M.pps[c + g.nxp] - ppc);
    A.qE[c] += lam * (M.pps[c + 1] - ppc);
    A.qS[c] += lam * (M.pps[c - g.nxp] - ppc);
    A.qW[c] += lam * (M.pps[c - 1] - ppc);
}
// k_pp_src: pp sources, dry-cell rows, boundary folding
__global__ void k_r_pp_src(Geo g, St T, Mid M, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const uint8_t *w = T.iwet;
    const double bsu = A.lp[c], bsp = A.qsp[c];
    A.qsp[c] = 0.0;
    double *bq[4] = {A.qN + c, A.qE + c, A.qS + c, A.qW + c};
    double bpost[4];
    for (int k = 0; k < 4; k++) { bpost[k] = *bq[k]; *bq[k] = 0.0; }
    const int wet = w[c];
    double a[4], aploc = 0.0;
    bool dry_read[4] = {false, false, false, false};
    if (wet) {
        for (int k = 0; k < 4; k++) a[k] = M.a0[k][c];
    } else {
        int sumwet = 0;
        for (int k = 0; k < 4; k++) if (!nb_ghost(g, i, j, k)) sumwet += w[nbr(g, c, k)];
        for (int k = 0; k < 4; k++) {
            const int nb = nbr(g, c, k);
            if (sumwet >= 1) { a[k] = w[nb]; dry_read[k] = true; }
            else if (nb_ghost(g, i, j, k)) a[k] = 0.0;
            else { a[k] = g.dsxy[k]; dry_read[k] = true; }
            if (dry_read[k]) aploc += a[k];
        }
    }
    double bpre[4];
    for (int k = 0; k < 4; k++) {
        bpre[k] = bpost[k];
        if (!(nb_ghost(g, i, j, k) && g.side[k] == 1)) continue;
        const int gc = nbr(g, c, k);
        const double ppg = M.ppv[gc];
        const double bppg = A.pp[gc] + a[k] * bsu;
        A.pp[gc] = 0.0;
        aadd(A.p + gc, -bppg);
        bpre[k] = bsu * ppg - bsp;
    }
    if (wet) {
        A.sspp0[c] += bsu;
        aadd(A.fluxN + c, -bsu);
        aadd(A.fluxE + c, -bsu);
        aadd(A.fluxN + c - g.nxp, bsu);
        aadd(A.fluxE + c - 1, bsu);
        aadd(A.p + c, T.sppp0[c] * bsu);
        for (int k = 0; k < 4; k++) *bq[k] = bpre[k];
    } else {
        for (int k = 0; k < 4; k++)
            if (dry_read[k]) aadd(A.p + nbr(g, c, k), a[k] * bsu);
        aadd(A.p + c, -aploc * bsu);
    }
}

// pp coefficient and Rhie-Chow flux of the face of `own` in direction k
__device__ __forceinline__ void r_pp_coef(const Geo &g, const St &T, const Mid &M, Ad &A, int own, int oi, int oj,
                                          int k, double b) {
    if (b == 0.0) return;
    const int nck = nbr(g, own, k);
    if (T.iwet[own] * T.iwet[nck] == 0) return;
    const double hk = fsym(T.hkE, T.hkN, g, own, k);
    double *bhk = face_at(A.hkE, A.hkN, g, own, k);
    if (nb_ghost(g, oi, oj, k)) {
        const double den = 1.0 - M.sumu[own], apa = M.apa[own];
        aadd(bhk, b * 2.0 * hk * g.ds[k] * apa / den / g.dn[k]);
        aadd(A.apa + own, b * hk * hk * g.ds[k] / den / g.dn[k]);
        aadd(A.sumu + own, b * hk * hk * g.ds[k] * apa / (den * den) / g.dn[k]);
        return;
    }
    const double fk = g.fintp[k], fp = 1.0 - fk;
    const double apf = fk * M.apa[nck] + fp * M.apa[own];
    const double suf = fk * M.sumu[nck] + fp * M.sumu[own];
    const double den = 1.0 - suf;
    aadd(bhk, b * 2.0 * hk * g.ds[k] * apf / den / g.dn[k]);
    const double bapf = b * hk * hk * g.ds[k] / den / g.dn[k];
    const double bsuf = b * hk * hk * g.ds[k] * apf / (den * den) / g.dn[k];
    aadd(A.apa + nck, fk * bapf);
    aadd(A.apa + own, fp * bapf);
    aadd(A.sumu + nck, fk * bsuf);
    aadd(A.sumu + own, fp * bsuf);
}
__device__ __forceinline__ void r_rc_flux(const Geo &g, const St &T, const Mid &M, Ad &A, int own, int oi, int oj,
                                          int k, double b) {
    if (b == 0.0) return;
    const int nck = nbr(g, own, k);
    if (T.iwet[own] * T.iwet[nck] == 0) return;
    const double hk = fsym(T.hkE, T.hkN, g, own, k);
    const double ds = g.ds[k], dn = g.dn[k];
    const double dp = T.p[nck] - T.p[own];
    double *bhk = face_at(A.hkE, A.hkN, g, own, k);
    if (nb_ghost(g, oi, oj, k)) {
        const double U = g.fnx[k] * M.Hu[own] + g.fny[k] * M.Hv[own], apa = M.apa[own];
        aadd(bhk, b * (ds * U - 2.0 * ds * apa * hk * dp / dn));
        aadd(A.Hu + own, b * hk * ds * g.fnx[k]);
        aadd(A.Hv + own, b * hk * ds * g.fny[k]);
        aadd(A.apa + own, -b * hk * ds * hk * dp / dn);
        aadd(A.p + nck, -b * hk * ds * apa * hk / dn);
        aadd(A.p + own, b * hk * ds * apa * hk / dn);
        return;
    }
    const double fk = g.fintp[k], fp = 1.0 - fk;
    const double apf = fk * M.apa[nck] + fp * M.apa[own];
    const double U = g.fnx[k] * (fk * M.Hu[nck] + fp * M.Hu[own]) + g.fny[k] * (fk * M.Hv[nck] + fp * M.Hv[own]);
    aadd(bhk, b * (ds * U - 2.0 * ds * apf * hk * dp / dn));
    const double bU = b * hk * ds;
    aadd(A.Hu + nck, bU * g.fnx[k] * fk);
    aadd(A.Hu + own, bU * g.fnx[k] * fp);
    aadd(A.Hv + nck, bU * g.fny[k] * fk);
    aadd(A.Hv + own, bU * g.fny[k] * fp);
    const double bapf = -b * hk * ds * hk * dp / dn;
    aadd(A.apa + nck, fk * bapf);
    aadd(A.apa + own, fp * bapf);
    aadd(A.p + nck, -b * hk * ds * apf * hk / dn);
    aadd(A.p + own, b * hk * ds * apf * hk / dn);
}
__global__ void k_r_pp_coef_flux(Geo g, St T, Mid M, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    double ba[4] = {A.qN[c], A.qE[c], A.qS[c], A.qW[c]};
    A.qN[c] = 0.0; A.qE[c] = 0.0; A.qS[c] = 0.0; A.qW[c] = 0.0;
    r_pp_coef(g, T, M, A, c, i, j, KN, ba[KN]);
    r_pp_coef(g, T, M, A, c, i, j, KE, ba[KE]);
    if (j > 0) r_pp_coef(g, T, M, A, c - g.nxp, i, j - 1, KN, ba[KS]);
    else r_pp_coef(g, T, M, A, c, i, j, KS, ba[KS]);
    if (i > 0) r_pp_coef(g, T, M, A, c - 1, i - 1, j, KE, ba[KW]);
    else r_pp_coef(g, T, M, A, c, i, j, KW, ba[KW]);
    double b = A.fluxN[c];
    A.fluxN[c] = 0.0;
    r_rc_flux(g, T, M, A, c, i, j, KN, b);
    b = A.fluxE[c];
    A.fluxE[c] = 0.0;
    r_rc_flux(g, T, M, A, c, i, j, KE, b);
    if (j == 0) {
        b = A.fluxN[c - g.nxp];
        A.fluxN[c - g.nxp] = 0.0;
        r_rc_flux(g, T, M, A, c, i, j, KS, -b);
    }
    if (i == 0) {
        b = A.fluxE[c - 1];
        A.fluxE[c - 1] = 0.0;
        r_rc_flux(g, T, M, A, c, i, j, KW, -b);
    }
}
// k_pp_prep
__global__ void k_r_pp_prep(Geo g, St T, Mid M, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const double asum = ((M.au[0][c] + M.au[1][c]) + M.au[2][c]) + M.au[3][c];
    const int w = T.iwet[c];
    const double D = asum - M.sp[c], Dm = fmax(D, F4(1.0e-6));
    const double apu = w * RELAX_UV / Dm, apa = apu * g.area;
    const double bHu = A.Hu[c], bHv = A.Hv[c], bsumu = A.sumu[c];
    double bapa = A.apa[c];
    A.Hu[c] = 0.0; A.Hv[c] = 0.0; A.apa[c] = 0.0; A.sumu[c] = 0.0;
    const double h = T.h[c], dpx = T.dpx[c], dpy = T.dpy[c];
    aadd(A.u + c, w * bHu);
    aadd(A.v + c, w * bHv);
    bapa += w * (dpx * h * bHu + dpy * h * bHv);
    A.dpx[c] += w * apa * h * bHu;
    A.dpy[c] += w * apa * h * bHv;
    aadd(A.h + c, w * apa * (dpx * bHu + dpy * bHv));
    const double bapu = g.area * bapa + asum * bsumu;
    double basum = apu * bsumu;
    if (D > F4(1.0e-6)) {
        const double bD = -w * RELAX_UV / (Dm * Dm) * bapu;
        basum += bD;
        A.sp[c] -= bD;
    }
    A.aN[c] += basum; A.aE[c] += basum; A.aS[c] += basum; A.aW[c] += basum;
}
// Gradients of the u and v solves' inputs (relaxation 0.8, solvers.f90:232-240):
// the row is R = sum_k a_k phi_nb + su + (1 - w) ap phi_old - ap phi with
// ap = (asum - sp) / w.
__global__ void k_r_uvsolve(Geo g, Mid M, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const double rinv = 1.0 / F4(0.8), relaxm = 1.0 - F4(0.8);
    const double a[4] = {M.au[0][c], M.au[1][c], M.au[2][c], M.au[3][c]};
    const double asum = ((a[0] + a[1]) + a[2]) + a[3];
    const double lu = A.lu[c], lv = A.lv[c];
    const double apu = (asum - M.spu[c]) * rinv, apv = (asum - M.spv[c]) * rinv;
    const double du = rinv * (relaxm * M.u0[c] - M.us[c]), dv = rinv * (relaxm * M.v0[c] - M.vs[c]);
    A.su[c] = lu;
    A.sv[c] = lv;
    A.spu[c] += -lu * du;
    A.spv[c] += -lv * dv;
    double *ba[4] = {A.aN + c, A.aE + c, A.aS + c, A.aW + c};
    for (int k = 0; k < 4; k++) {
        const int nb = nbr(g, c, k);
        *ba[k] += lu * (M.us[nb] + du) + lv * (M.vs[nb] + dv);
        if (nb_ghost(g, i, j, k)) {
            aadd(A.u + nb, a[k] * lu);
            aadd(A.v + nb, a[k] * lv);
        }
    }
    A.u[c] = relaxm * apu * lu;
    A.v[c] = relaxm * apv * lv;
}
// k_mom
__global__ void k_r_mom(Geo g, St T, Par P, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const double Ar = g.area, fc = T.fc[c];
    const double u = T.u[c], v = T.v[c], h = T.h[c];
    const double bsu = A.su[c], bsv = A.sv[c], bsp = A.sp[c], bspu = A.spu[c], bspv = A.spv[c];
    A.su[c] = 0.0; A.sv[c] = 0.0; A.sp[c] = 0.0; A.spu[c] = 0.0; A.spv[c] = 0.0;
    double ba[4] = {A.aN[c], A.aE[c], A.aS[c], A.aW[c]};
    A.aN[c] = 0.0; A.aE[c] = 0.0; A.aS[c] = 0.0; A.aW[c] = 0.0;
    const double bu_out = A.u[c], bv_out = A.v[c];
    double hk[4], vk[4], fl[4], ddk[4];
    for (int k = 0; k < 4; k++) {
        hk[k] = fsym(T.hkE, T.hkN, g, c, k);
        vk[k] = fsym(T.viskE, T.viskN, g, c, k);
        fl[k] = fdir(T.fluxE, T.fluxN, g, c, k);
        ddk[k] = vk[k] * hk[k] * g.dsxy[k];
    }
    const int wet = T.iwet[c];
    double bsp_m, bsu_m, bsv_m, bu_in, bv_in, bz0 = 0.0;
    double bhk[4] = {0, 0, 0, 0}, bvk[4] = {0, 0, 0, 0}, bfl[4] = {0, 0, 0, 0}, bddk[4] = {0, 0, 0, 0};
    if (P.wallfric) {
        if (wet == 1) {
            bsp_m = bsp + bspu + bspv;
            bsu_m = bsu; bsv_m = bsv;
            bu_in = bu_out; bv_in = bv_out;
            const double z0w = T.z0[c] * 1.0;
            for (int k = 0; k < 4; k++) {
                if (T.iwet[nbr(g, c, k)] != 0) continue;
                const bool ns = fabs(g.fny[k]) > F4(0.999);
                const double yp = ns ? fabs(g.ry[k]) : fabs(g.rx[k]);
                const double vel = ns ? u : v, b = ns ? bspu : bspv;
                const Dl wc = wall_coef_t<Dl>(yp, Dl{z0w, 1.0});
                const double av = fabs(vel), sg = vel > 0.0 ? 1.0 : vel < 0.0 ? -1.0 : 0.0;
                // spu -= ds hk wall_coef(z0) |u|
                bhk[k] += -b * g.ds[k] * wc.v * av;
                bz0 += -b * g.ds[k] * hk[k] * av * wc.d;
                if (ns) bu_in += -b * g.ds[k] * hk[k] * wc.v * sg;
                else bv_in += -b * g.ds[k] * hk[k] * wc.v * sg;
            }
        } else {
            bsp_m = 0.0; bsu_m = 0.0; bsv_m = 0.0; bu_in = 0.0; bv_in = 0.0;
        }
    } else {
        bsp_m = wet * (bsp + bspu + bspv);
        bsu_m = wet * bsu; bsv_m = wet * bsv;
        bu_in = wet * bu_out; bv_in = wet * bv_out;
    }
    double bh = 0.0;
    for (int k = 0; k < 4; k++) {
        const bool fold = nb_ghost(g, i, j, k) && g.side[k] == 1;
        if (!fold) {
            double gd, gf;
            schm_grad(P.ndsch, ddk[k], fl[k], gd, gf);
            bddk[k] += gd * ba[k];
            bfl[k] += gf * ba[k];
            continue;
        }
        const int gc = nbr(g, c, k);
        const double ab = schmcoef(1, ddk[k], fl[k]);
        double bab = -bsp_m;
        const double hg = T.h[gc];
        if (fabs(g.fny[k]) > F4(0.9999)) {
            const double vg = g.fny[k] * fl[k] / (g.ds[k] * hk[k]);
            bab += bsv_m * vg;
            const double bvg = bsv_m * ab + A.v[gc];
            A.v[gc] = 0.0;
            const double bug = A.u[gc];
            A.u[gc] = 0.0;
            bu_in += wet * h / hg * bug;
            bh += wet * u / hg * bug;
            aadd(A.h + gc, -wet * u * h / (hg * hg) * bug);
            bfl[k] += g.fny[k] / (g.ds[k] * hk[k]) * bvg;
            bhk[k] += -g.fny[k] * fl[k] / (g.ds[k] * hk[k] * hk[k]) * bvg;
        } else {
            const double ug = g.fnx[k] * fl[k] / (g.ds[k] * hk[k]);
            bab += bsu_m * ug;
            const double bug = bsu_m * ab + A.u[gc];
            A.u[gc] = 0.0;
            const double bvg = A.v[gc];
            A.v[gc] = 0.0;
            bv_in += wet * h / hg * bvg;
            bh += wet * v / hg * bvg;
            aadd(A.h + gc, -wet * v * h / (hg * hg) * bvg);
            bfl[k] += g.fnx[k] / (g.ds[k] * hk[k]) * bug;
            bhk[k] += -g.fnx[k] * fl[k] / (g.ds[k] * hk[k] * hk[k]) * bug;
        }
        double gd, gf;
        schm_grad(1, ddk[k], fl[k], gd, gf);
        bddk[k] += gd * bab;
        bfl[k] += gf * bab;
    }
    // sources and the friction term
    A.ssu0[c] += bsu_m;
    bv_in += fc * h * Ar * bsu_m;
    A.dpx[c] -= h * Ar * bsu_m;
    bh += (fc * v - T.dpx[c]) * Ar * bsu_m;
    A.ssv0[c] += bsv_m;
    bu_in -= fc * h * Ar * bsv_m;
    A.dpy[c] -= h * Ar * bsv_m;
    bh += (-fc * u - T.dpy[c]) * Ar * bsv_m;
    A.spuv0[c] += bsp_m;
    A.cfrict[c] -= T.uelwc[c] * Ar * bsp_m;
    A.uelwc[c] -= T.cfrict[c] * Ar * bsp_m;
    A.z0[c] += bz0;
    for (int k = 0; k < 4; k++) {
        bvk[k] += hk[k] * g.dsxy[k] * bddk[k];
        bhk[k] += vk[k] * g.dsxy[k] * bddk[k];
        aadd(face_at(A.hkE, A.hkN, g, c, k), bhk[k]);
        aadd(face_at(A.viskE, A.viskN, g, c, k), bvk[k]);
        aadd(face_at(A.fluxE, A.fluxN, g, c, k), fdir_sign(k) * bfl[k]);
    }
    A.u[c] = bu_in;
    A.v[c] = bv_in;
    aadd(A.h + c, bh);
}

// ------------------------------------------------------------ step start/end
// flow_srcsnk_static
__global__ void k_r_srcsnk(Geo g, St T, Par P, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const double dti = 1.0 / P.dtime, gdi = 1.0 / (GRAV * P.dtime);
    const double bsp0 = A.spuv0[c], bsu0 = A.ssu0[c], bsv0 = A.ssv0[c], bss = A.sspp0[c];
    A.spuv0[c] = 0.0; A.ssu0[c] = 0.0; A.ssv0[c] = 0.0; A.sspp0[c] = 0.0;
    const double fac1 = g.area * T.h1[c] * dti;
    A.h1[c] += g.area * dti * (-bsp0 + T.u1[c] * bsu0 + T.v1[c] * bsv0);
    A.u1[c] += fac1 * bsu0;
    A.v1[c] += fac1 * bsv0;
    A.p1[c] += g.area * gdi * bss;
}
// one dry-to-wet sweep (k_d2w_sweep + k_d2w_commit): read the new values'
// adjoints, then spread them over the four neighbours
__global__ void k_r_d2w_take(Geo g, St T, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (!(T.iwet[c] == 1 && T.iwet1[c] == 0)) return;
    A.tu[c] = A.u[c]; A.u[c] = 0.0;
    A.tv[c] = A.v[c]; A.v[c] = 0.0;
}
__global__ void k_r_d2w_spread(Geo g, St T, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    if (!(T.iwet[c] == 1 && T.iwet1[c] == 0)) return;
    int ns = 0;
    for (int k = 0; k < 4; k++) ns += T.iwet[nbr(g, c, k)];
    const double den = (double)(float)ns + SMALL;
    const double bu = A.tu[c] / den / F4(1.25), bv = A.tv[c] / den / F4(1.25);
    A.tu[c] = 0.0; A.tv[c] = 0.0;
    for (int k = 0; k < 4; k++) {
        aadd(A.u + nbr(g, c, k), bu);
        aadd(A.v + nbr(g, c, k), bv);
    }
}
// k_wetdry_post: face fluxes and viscosities zeroed next to dry cells
__global__ void k_r_wetdry_post(Geo g, St T, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const uint8_t *w = T.iwet;
    if (w[c] * w[c + 1] == 0) { A.fluxE[c] = 0.0; A.viskE[c] = 0.0; }
    if (w[c] * w[c + g.nxp] == 0) { A.fluxN[c] = 0.0; A.viskN[c] = 0.0; }
    if (i == 0 && w[c] * w[c - 1] == 0) { A.fluxE[c - 1] = 0.0; A.viskE[c - 1] = 0.0; }
    if (j == 0 && w[c] * w[c - g.nxp] == 0) { A.fluxN[c - g.nxp] = 0.0; A.viskN[c - g.nxp] = 0.0; }
}
// k_wetdry_h: h = p/g - zb, held at hmin below hdry
__global__ void k_r_wetdry_h(Geo g, St T, Par P, Ad A) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    const double raw = T.p[t] * GRAVINV - T.zb[t];
    const double bh = A.h[t];
    A.h[t] = 0.0;
    if (!(raw < P.hdry) || raw > P.hmin) {
        A.p[t] += GRAVINV * bh;
        A.zb[t] -= bh;
    }
}
// k_velface: velocity clipping to the face velocities, dry cells, friction
__global__ void k_r_velface(Geo g, St S, Par P, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    const double h = S.h[c], u = S.u[c], v = S.v[c], n = S.coefman[c];
    const double bu = A.u[c], bv = A.v[c], bcd = A.cfrict[c], bue = A.uelwc[c], bz = A.z0[c], bho = A.h[c];
    A.cfrict[c] = 0.0;
    A.uelwc[c] = 0.0;
    const bool wetb = h > P.hmin && S.iwet[c] == 1;
    double uc = 0.0, vc = 0.0, hv = h;
    int kumin = -1, kumax = -1, kvmin = -1, kvmax = -1;
    double vel[4];
    if (wetb) {
        double umin = 1.0e6, umax = -1.0e6, vmin = 1.0e6, vmax = -1.0e6;
        for (int k = 0; k < 4; k++) {
            vel[k] = fdir(S.fluxE, S.fluxN, g, c, k) / (fsym(S.hkE, S.hkN, g, c, k) * g.ds[k]);
            const double uf = g.fnx[k] * vel[k], vf = g.fny[k] * vel[k];
            if (uf < umin) { umin = uf; kumin = k; }
            if (uf > umax) { umax = uf; kumax = k; }
            if (vf < vmin) { vmin = vf; kvmin = k; }
            if (vf > vmax) { vmax = vf; kvmax = k; }
        }
        uc = fmin(fmax(u, umin), umax);
        vc = fmin(fmax(v, vmin), vmax);
        if (!(u < umin)) kumin = -1;
        if (!(u > umax) || u < umin) kumax = -1;
        if (!(v < vmin)) kvmin = -1;
        if (!(v > vmax) || v < vmin) kvmax = -1;
    } else {
        hv = fmax(P.hmin, h);
    }
    double gin[4];
    bool z0set = true;
    for (int q = 0; q < 4; q++) {
        const Dl x[4] = {{uc, q == 0 ? 1.0 : 0.0}, {vc, q == 1 ? 1.0 : 0.0}, {hv, q == 2 ? 1.0 : 0.0},
                         {n, q == 3 ? 1.0 : 0.0}};
        const Dl zero{0.0, 0.0};
        Dl cd, z0, ue, vis;
        fric_vis_t<Dl>(P, 0, x[0], x[1], x[2], x[3], zero, zero, zero, zero, cd, z0, ue, vis);
        z0set = cd.v != 0.0;
        gin[q] = cd.d * bcd + (z0set ? z0.d * bz : 0.0) + ue.d * bue;
    }
    if (z0set) A.z0[c] = 0.0;
    aadd(A.coefman + c, gin[3]);
    if (wetb) {
        const double buc = bu + gin[0], bvc = bv + gin[1];
        A.h[c] = bho + gin[2];
        auto face = [&](int k, double b, bool isu) {
            // uf = fn * flux_dir / (hk ds)
            const double fn = isu ? g.fnx[k] : g.fny[k];
            const double hk = fsym(S.hkE, S.hkN, g, c, k);
            aadd(face_at(A.fluxE, A.fluxN, g, c, k), fdir_sign(k) * b * fn / (hk * g.ds[k]));
            aadd(face_at(A.hkE, A.hkN, g, c, k), -b * fn * vel[k] / hk);
        };
        if (kumin >= 0) { face(kumin, buc, true); A.u[c] = 0.0; }
        else if (kumax >= 0) { face(kumax, buc, true); A.u[c] = 0.0; }
        else A.u[c] = buc;
        if (kvmin >= 0) { face(kvmin, bvc, false); A.v[c] = 0.0; }
        else if (kvmax >= 0) { face(kvmax, bvc, false); A.v[c] = 0.0; }
        else A.v[c] = bvc;
    } else {
        A.h[c] = h > P.hmin ? bho + gin[2] : 0.0;
        A.u[c] = 0.0;
        A.v[c] = 0.0;
    }
}
// flow_update: u1 = u, v1 = v, p1 = p, h1 = h
__global__ void k_r_update(Geo g, Ad A) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * g.nyp) return;
    A.u[t] += A.u1[t]; A.u1[t] = 0.0;
    A.v[t] += A.v1[t]; A.v1[t] = 0.0;
    A.p[t] += A.p1[t]; A.p1[t] = 0.0;
    A.h[t] += A.h1[t]; A.h1[t] = 0.0;
}
// dL/deta of the gauge cells (eta = p/g where velface sets it)
__global__ void k_r_gauge(Geo g, St S, Par P, const int *cells, int ng, const double *geta, Ad A) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ng) return;
    const int a = cells[t], j = a / g.nx, i = a - j * g.nx, c = cidx(g, i, j);
    if (S.h[c] > P.hmin && S.iwet[c] == 1) aadd(A.p + c, GRAVINV * geta[t]);
}
__global__ void k_gauge_eta(Geo g, St s, const int *cells, int ng, double *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ng) return;
    const int a = cells[t], j = a / g.nx, i = a - j * g.nx;
    out[t] = s.eta[cidx(g, i, j)];
}

// ------------------------------------------------------------ initial state
// flow_convflux
__global__ void k_r_convflux(Geo g, St S, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    auto face = [&](double *bf, double sgn, int nb, int k, double hk, double *bhk) {
        const double b = sgn * *bf;
        *bf = 0.0;
        const double iw = (double)(S.iwet[c] * S.iwet[nb]);
        if (iw == 0.0 || b == 0.0) return;
        const double uni = g.fnx[k] * S.u[c] + g.fny[k] * S.v[c];
        const double unk = g.fnx[k] * S.u[nb] + g.fny[k] * S.v[nb];
        aadd(bhk, b * iw * g.ds[k] * (g.fintp[k] * unk + (1.0 - g.fintp[k]) * uni));
        const double q = b * iw * g.ds[k] * hk;
        aadd(A.u + nb, q * g.fintp[k] * g.fnx[k]);
        aadd(A.v + nb, q * g.fintp[k] * g.fny[k]);
        aadd(A.u + c, q * (1.0 - g.fintp[k]) * g.fnx[k]);
        aadd(A.v + c, q * (1.0 - g.fintp[k]) * g.fny[k]);
    };
    face(A.fluxN + c, 1.0, c + g.nxp, KN, S.hkN[c], A.hkN + c);
    face(A.fluxE + c, 1.0, c + 1, KE, S.hkE[c], A.hkE + c);
    if (j == 0) face(A.fluxN + c - g.nxp, -1.0, c - g.nxp, KS, S.hkN[c - g.nxp], A.hkN + c - g.nxp);
    if (i == 0) face(A.fluxE + c - 1, -1.0, c - 1, KW, S.hkE[c - 1], A.hkE + c - 1);
}
// interp_scal_cell2face(zb): the face elevations
__global__ void k_r_zbk(Geo g, Ad A) {
    int i, j, c;
    if (!act_thread(g, i, j, c)) return;
    auto face = [&](double b, int nb, int k) {
        aadd(A.zb + nb, g.fintp[k] * b);
        aadd(A.zb + c, (1.0 - g.fintp[k]) * b);
    };
    face(A.zbkN[c], c + g.nxp, KN);
    face(A.zbkE[c], c + 1, KE);
    if (j == 0) face(A.zbkN[c - g.nxp], c - g.nxp, KS);
    if (i == 0) face(A.zbkE[c - 1], c - 1, KW);
}

}  // namespace cms
