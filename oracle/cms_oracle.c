/*
 * cms_oracle.c -- FP64 CPU restatement of CMS-Flow's implicit hydrodynamic step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path in cms2d_amd/.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every routine cites the reference routine it restates (paths relative to the
 * erdc/cms2d tree).  Constants that the reference writes as default-real
 * literals (e.g. 0.8, 1.0e-6, 0.3333) keep the value of that single-precision
 * literal, F4(x), exactly as an ikind=8 rebuild of the reference evaluates them;
 * constants written with the _ikind suffix (grav=9.81_ikind) are exact doubles.
 *
 * Loops whose OpenMP form in the reference reads values that other iterations
 * of the same loop write (the dry-to-wet initialisation, flow.F90:1477-1491;
 * the eddy-viscosity fix-ups at dry boundaries and corners,
 * flow_eddyvis.f90:168-212) have two modes: serial (the reference run on one
 * thread) and race-free (every read sees the values from before the loop, the
 * semantics of the HIP path).  cms_case.serial selects.
 */
#include "cms_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define F4(x) ((double)(float)(x))
#define NF (o->nf)
#define KN 0
#define KE 1
#define KS 2
#define KW 3

static const int kopp[4] = {KS, KW, KN, KE};      /* kkface, geospatial.F90:58-61 */

/* reference constants (definitions.F90:88-92, flow.F90:21-169, friction.F90:18-65) */
static const double GRAV = 9.81;                  /* 9.81_ikind */
#define SMALL   F4(1.0e-15)                       /* const_def small */
#define RHOW    1025.0
#define CVISCON F4(1.0e-6)
#define CVISMAX 20.0
#define CVISBOT F4(0.0667)
#define CVISHOR F4(0.2)
#define CVISWAV 0.5
#define CVISWAVBRK F4(0.1)
#define RELAX_UV F4(0.8)
#define RELAX_PP 1.0
#define FACPP   F4(0.2)
#define CFRICWAV F4(0.65)
#define HMAXFRIC 10.0
/* check_conv thresholds, flow.F90:135-150 */
#define RMOMMAXP    F4(1.0e-3)
#define RMOMTARGETP F4(1.0e-4)
#define RMOMMINP    F4(1.0e-8)
#define RMOMRATIOP  F4(0.5)
#define RMOMMAXUV   F4(1.0e-2)
#define RMOMTARGETUV F4(1.0e-3)
#define RMOMMINUV   F4(1.0e-7)
#define RMOMRATIOUV 1.0

/* Gradient operator of one cell (der_go_type, der_def.f90): cell-based
 * Green-Gauss weights (wcx, wcy over cells icx, icy) and, on joint cells of a
 * telescoping grid (ider 2), the deferred cell-reconstruction weights applied
 * to the previous perpendicular derivative (wgxy over igx, wgyx over igy). */
#define NSC 7                                     /* nmaxfaces + 1 */
typedef struct {
    int ncx, ncy;
    int icx[NSC], icy[NSC];
    double wcx[NSC], wcy[NSC];
    int ngx, ngy;
    int igx[NSC], igy[NSC];
    double wgxy[NSC], wgyx[NSC];
} gradop;

struct cms_oracle {
    cms_case c;
    int nx, ny, ncells, ncellsD;
    /* connectivity / geometry (geo_def): nf faces per cell slot (nmaxfaces), ncf
     * used by each cell (ncface), idir its directions 0..3 = N,E,S,W (idirface) */
    int nf, *ncf, *idir;
    int *c2c, *llec, *nxy, *kxy;
    int *nxf, *kxf, *nyf, *kyf;       /* no-repeat x (E/W) and y (N/S) face lists */
    int *isjoint, njoint, *joint;     /* joint cells (telescoping), ascending */
    int ider2;                        /* gradients with the joint-cell corrections (nder 2) */
    int skewcor;                      /* skewness corrections on (joint cells present) */
    double *rpx, *rpy;
    double *dHux, *dHuy, *dHvx, *dHvy, *dapux, *dapuy, *dzbx, *dzby;
    gradop *goa;
    double *dx, *dy, *x, *y, *area;
    double *ds, *dn, *fnx, *fny, *rx, *ry, *dsx, *dsy, *dsxy, *fintp;
    double areaavg;
    /* boundary strings (bnd_def) */
    int nW, *Wc, *Wk;
    int nTH, *THc, *THk;
    int nQ, *Qc, *Qk;     /* flux (river) string Q_str, cells and faces in string order */
    /* flow statistics (stat_def): on, window start/end [s], steps counted, the
     * per-cell accumulators in the order of cms_oracle_stat_names */
    int stat_on, stat_n;
    double stat_t1, stat_t2;
    double *st[33];
    double qflux, cmvel, qcos, qsin;   /* its total inflow this step, conveyance, inflow direction */
    int q_inc;            /* window start of the flux curve (Q_str%inc) */
    double *wsebnd, *wseadj;
    int wseadjust;
    int wse_inc;          /* window start of the water-level curve (H_str%inc) */
    /* spatially constant wind (met_def windconst, wndx, wndy) */
    int windconst;
    double wndx, wndy, tauwx, tauwy;   /* the vector and the stress of the current step */
    /* state (flow_def, fric_def, wave_flowgrid_def) */
    double *u, *v, *p, *h, *pp, *dppx, *dppy, *dux, *duy, *dvx, *dvy, *dpx, *dpy;
    double *vis, *zb, *eta, *u1, *v1, *p1, *h1;
    int *iwet, *iwet1, *icorner;
    double *flux, *flux1, *hk, *pk, *visk, *zbk, *acoef;
    double *su, *sv, *sp, *spu, *spv, *ssu0, *ssv0, *spuv0, *sppp0, *sspp0;
    double *apuareap, *sumu, *Hu, *Hv, *rsu, *rsv, *rsp;
    double *coefman, *cfrict, *z0, *uv, *uelwc, *bsxy, *bsvel, *cmb, *cbcfuwcap;
    double *cbf;   /* linear friction coefficient (cbotfric of mbedfric 4) */
    double *maxeta;   /* running maximum of eta over the steps taken */
    double *wavestrx, *wavestry, *fc;
    gradop *gow;
    /* control (comvarbl) */
    double dtime, timesecs, ctime, timehrs, ramp, rampdur, hmin, hdry, hdry1, hdry2;
    int dtvar, jtime, mtime;          /* variable time step (flow.F90:107, 664-666) */
    double dtimebeg, deltime;
    int maxit, niter, iconv, ntime;
    double rmom[3], rmom0[3], rmom1[3];
    int kry[3];
    double bndvol;
    /* solver workspace */
    double *wk;
    int *ia, *ja, *diagpos;
    double *lu;
};

static void *xcalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s); if (!p) abort(); return p; }
#define DALLOC(n) ((double *)xcalloc((size_t)(n), sizeof(double)))
#define IALLOC(n) ((int *)xcalloc((size_t)(n), sizeof(int)))

/* ---------------------------------------------------------------------------
 * Closed-form helpers
 * ------------------------------------------------------------------------- */

/* comp_lib.f90:40-109 zerocoef / upwindcoef / hybridcoef / powerlawcoef /
 * exponentialcoef (ndsch 0-4; every other scheme takes upwind coefficients,
 * flow_imp.F90:120-126) */
double cms_oracle_schmcoef(int ndsch, double dk, double fk)
{
    double absfk = fabs(fk), peck;
    switch (ndsch) {
    case 0:
        return 0.0;
    case 2:
        peck = absfk / fmax(dk, F4(1.0e-6));
        return dk * fmax(0.0, 1.0 - 0.5 * peck) + fmax(0.0, -fk);
    case 3:
        peck = absfk / fmax(fmax(dk, absfk), F4(1.0e-6));
        return dk * (1.0 - pow(peck, 5)) + fmax(0.0, -fk);
    case 4:
        peck = absfk / fmax(fmax(dk, F4(1.e-6)), absfk);
        return dk * peck / fmax(exp(peck) - 1.0, F4(1.0e-6)) + fmax(0.0, -fk);
    default:
        return dk + fmax(0.0, -fk);
    }
}

/* NVD deferred-correction limiters gamma(phi~) of the gamma family,
 * comp_lib.f90:112-208 (ndsch 6 GAMMA, 7 CUBISTA, 8 ALVSMART, 9 HOAB; the
 * reference's 0.1667 is a single-precision literal). */
double cms_oracle_nvd_gamma(int ndsch, double phin)
{
    switch (ndsch) {
    case 6:
        if (phin > 0.0 && phin <= 0.5) return 2.0 * phin;
        if (phin > 0.5 && phin <= 1.0) return 1.0;
        return 0.0;
    case 7:
        if (phin > 0.0 && phin <= 0.375) return 1.5 * phin / (1.0 - phin);
        if (phin > 0.375 && phin <= 0.75) return (3.0 - 2.0 * phin) / (4.0 - 4.0 * phin);
        if (phin > 0.75 && phin <= 1.0) return 1.5;
        return 0.0;
    case 8:
        if (phin > 0.0 && phin <= 0.25) return 2.5 * phin / (1.0 - phin);
        if (phin > 0.25 && phin <= 0.75) return (3.0 - 2.0 * phin) / (4.0 - 4.0 * phin);
        if (phin > 0.75 && phin <= 1.0) return 1.5;
        return 0.0;
    case 9:
        if (phin > 0.0 && phin <= F4(0.1667)) return 5.0 * phin / (1.0 - phin);
        if (phin > F4(0.1667) && phin <= 0.5) return 1.0;
        if (phin > 0.5 && phin <= 0.75) return 0.5 / (1.0 - phin);
        if (phin > 0.75 && phin <= 1.0) return 2.0;
        return 0.0;
    default:
        return 0.0;
    }
}

/* wind_drag_hsu (meteorological/met_lib.f90:33-66): Hsu (1988), reduced
 * above 30 m/s after Powell et al. (2003); wkappa = 0.4 (met_def) */
double cms_oracle_wind_drag_hsu(double w)
{
    if (w <= 30.0) {
        const double t = F4(0.4) / (F4(14.56) - 2.0 * log(fmax(w, F4(0.01))));
        return t * t;
    }
    return fmax(F4(3.86e-3) - F4(4.0e-5) * w, F4(0.0015));
}

/* fric_lib.f90:162-194 wall_coef */
double cms_oracle_wall_coef(double yp, double y0)
{
    double y0lim = fmin(fmax(y0, F4(1.0e-6)), yp * F4(0.9999));
    double t = F4(0.4) / log(yp / y0lim);
    return t * t;
}

/* fric_lib.f90:944-969 fric_conv_man2drag */
double cms_oracle_man2drag(double g, double cman, double h, double hmax)
{
    double hh = fmax(fmin(hmax, h), F4(0.1) * cman);
    return g * cman * cman / pow(hh, F4(0.3333));
}

/* fric_lib.f90:972-996 fric_conv_drag2length */
double cms_oracle_drag2length(double cd, double h)
{
    double z0 = h / exp(F4(0.4) / sqrt(cd) + 1.0);
    return fmax(z0, F4(1.0e-6));
}

/* fric_lib.f90:916-941 fric_conv_drag2man */
double cms_oracle_drag2man(double g, double cd, double h, double hmax)
{
    return sqrt(cd * pow(fmin(hmax, h), F4(0.3333)) / g);
}

/* fric_lib.f90:999-1023 fric_conv_length2drag */
double cms_oracle_length2drag(double z0, double h)
{
    double z0lim = fmax(fmin(z0, 0.5 * h), F4(1.0e-5));
    double t = F4(0.4) / (log(h / z0lim) - 1.0);
    return t * t;
}

/* time_lib.F90:234-256 ramp_func */
static double ramp_func(double timehrs, double rampdur)
{
    return 0.5 - 0.5 * cos(M_PI * fmin(timehrs, rampdur) / rampdur);
}

/* ---------------------------------------------------------------------------
 * Grid.  Topology of a rectangle (geo_cart, geospatial.F90:2709-2930) or of a
 * telescoping grid (read_tel, :403-795); geometry of both (geo_init_cart,
 * :1400-1611, interp_coef_cell2face, interp_lib.f90:2669-2724).
 * ------------------------------------------------------------------------- */
static void alloc_topology(cms_oracle *o, int nf, int nc, int nd)
{
    o->nf = nf;
    o->ncells = nc;
    o->ncellsD = nd;
    o->ncf = IALLOC(nd);
    o->c2c = IALLOC(NF * nd);
    o->llec = IALLOC(NF * nd);
    o->idir = IALLOC(NF * nd);
    for (int f = 0; f < NF * nd; f++) { o->c2c[f] = -1; o->llec[f] = -1; o->idir[f] = -1; }
    o->dx = DALLOC(nd); o->dy = DALLOC(nd); o->x = DALLOC(nd); o->y = DALLOC(nd);
    o->zb = DALLOC(nd);
}

static void build_topology_cart(cms_oracle *o)
{
    const int nx = o->nx, ny = o->ny, nc = nx * ny;
    int nd = nc;
    /* count ghost cells: every missing neighbour of an active cell gets one */
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
            nd += (j == ny - 1) + (i == nx - 1) + (j == 0) + (i == 0);
    alloc_topology(o, 4, nc, nd);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int c = j * nx + i;
            o->ncf[c] = 4;
            for (int k = 0; k < 4; k++) o->idir[NF * c + k] = k;
            o->dx[c] = o->c.dx; o->dy[c] = o->c.dy;
            o->x[c] = (i + 1) * o->c.dx - 0.5 * o->c.dx;
            o->y[c] = (j + 1) * o->c.dy - 0.5 * o->c.dy;
            o->zb[c] = o->c.zb[c];
            if (j < ny - 1) o->c2c[NF * c + KN] = c + nx;
            if (i < nx - 1) o->c2c[NF * c + KE] = c + 1;
            if (j > 0) o->c2c[NF * c + KS] = c - nx;
            if (i > 0) o->c2c[NF * c + KW] = c - 1;
        }
    /* dummy cells, in the reference's order (cell by cell, N,E,S,W) */
    int g = nc;
    for (int c = 0; c < nc; c++) {
        for (int k = 0; k < 4; k++) {
            if (o->c2c[NF * c + k] >= 0) continue;
            o->c2c[NF * c + k] = g;
            o->ncf[g] = 4;
            o->c2c[NF * g + kopp[k]] = c;
            o->dx[g] = o->dx[c]; o->dy[g] = o->dy[c];
            o->x[g] = o->x[c] + (k == KE ? o->dx[c] : k == KW ? -o->dx[c] : 0.0);
            o->y[g] = o->y[c] + (k == KN ? o->dy[c] : k == KS ? -o->dy[c] : 0.0);
            o->zb[g] = o->zb[c];
            g++;
        }
    }
}

/* read_tel (geospatial.F90:403-795) for a grid whose cells are all active:
 * faces in the order of the file's neighbour slots N1 N2 E1 E2 S1 S2 W1 W2,
 * one dummy (ghost) cell for each side without a neighbour, placed one cell
 * size beyond it with the cell's size and bed (:552-690). */
static void build_topology_tel(cms_oracle *o)
{
    const int nc = o->c.tel_n;
    const int *nb = o->c.tel_nb;
    int nd = nc;
    for (int i = 0; i < nc; i++)
        for (int k = 0; k < 8; k += 2)
            if (nb[8 * i + k] == 0 && nb[8 * i + k + 1] == 0) nd++;
    alloc_topology(o, 6, nc, nd);
    int g = nc;
    for (int i = 0; i < nc; i++) {
        const double *gm = &o->c.tel_geom[4 * i];
        o->x[i] = gm[0]; o->y[i] = gm[1]; o->dx[i] = gm[2]; o->dy[i] = gm[3];
        o->zb[i] = o->c.zb[i];
    }
    for (int i = 0; i < nc; i++) {
        int k = 0;
        for (int j = 0; j < 8; j++) {
            const int dir = j / 2;                     /* N, E, S, W */
            const int nck = nb[8 * i + j];
            const int open = (j % 2 == 0) && nck == 0 && nb[8 * i + j + 1] == 0;
            if (open) {
                o->c2c[NF * i + k] = g;
                o->idir[NF * i + k] = dir;
                k++;
                o->ncf[g] = 4;
                o->c2c[NF * g + kopp[dir]] = i;
                o->x[g] = o->x[i] + (dir == KE ? o->dx[i] : dir == KW ? -o->dx[i] : 0.0);
                o->y[g] = o->y[i] + (dir == KN ? o->dy[i] : dir == KS ? -o->dy[i] : 0.0);
                o->dx[g] = o->dx[i];
                o->dy[g] = o->dy[i];
                o->zb[g] = o->zb[i];
                g++;
            } else if (nck > 0) {
                o->c2c[NF * i + k] = nck - 1;
                o->idir[NF * i + k] = dir;
                k++;
            }
            if (k > NF) { fprintf(stderr, "cms_oracle: cell %d has more than 6 faces\n", i + 1); abort(); }
        }
        o->ncf[i] = k;
    }
}

static void geo_init(cms_oracle *o)
{
    const int nc = o->ncells, nd = o->ncellsD;
    o->area = DALLOC(nd);
    double asum = 0.0;
    for (int i = 0; i < nd; i++) o->area[i] = o->dx[i] * o->dy[i];
    for (int i = 0; i < nc; i++) asum += o->area[i];
    o->areaavg = asum / (double)(float)nc;           /* sum(areap)/real(ncells) */
    /* backward connectivity (:1446-1468); ghosts list their cell in the
     * direction-indexed slot */
    for (int i = 0; i < nc; i++)
        for (int k = 0; k < o->ncf[i]; k++) {
            const int f = NF * i + k, nck = o->c2c[f];
            if (nck < nc) {
                for (int kk = 0; kk < o->ncf[nck]; kk++)
                    if (o->c2c[NF * nck + kk] == i) o->llec[f] = kk;
            } else {
                o->llec[f] = kopp[o->idir[f]];
            }
        }
    /* joint cells: a neighbour whose size across the face differs by more
     * than 20 % (:1490-1505) */
    o->isjoint = IALLOC(nd);
    for (int i = 0; i < nc; i++)
        for (int k = 0; k < o->ncf[i]; k++) {
            const int f = NF * i + k, nck = o->c2c[f];
            if (nck >= nc) continue;
            const double r = (o->idir[f] == KE || o->idir[f] == KW) ? o->dy[i] / o->dy[nck] : o->dx[i] / o->dx[nck];
            if (r > F4(1.2) || r < F4(0.8)) { o->isjoint[i] = 1; o->isjoint[nck] = 1; }
        }
    o->njoint = 0;
    for (int i = 0; i < nc; i++) o->njoint += o->isjoint[i];
    o->joint = IALLOC(o->njoint);
    for (int i = 0, n = 0; i < nc; i++)
        if (o->isjoint[i]) o->joint[n++] = i;
    int F = NF * nd;
    o->ds = DALLOC(F); o->dn = DALLOC(F); o->fnx = DALLOC(F); o->fny = DALLOC(F);
    o->rx = DALLOC(F); o->ry = DALLOC(F); o->dsx = DALLOC(F); o->dsy = DALLOC(F);
    o->dsxy = DALLOC(F); o->fintp = DALLOC(F); o->rpx = DALLOC(F); o->rpy = DALLOC(F);
    for (int i = 0; i < nc; i++)
        for (int k = 0; k < o->ncf[i]; k++) {
            const int f = NF * i + k, nck = o->c2c[f], d = o->idir[f];
            if (d == KN || d == KS) o->ds[f] = fmin(o->dx[i], o->dx[nck]);
            else o->ds[f] = fmin(o->dy[i], o->dy[nck]);
            switch (d) {
            case KN: o->fnx[f] = 0; o->fny[f] = 1; o->dn[f] = fabs(o->y[nck] - o->y[i]);
                o->rx[f] = o->x[nck] - o->x[i]; o->ry[f] = o->dy[i] / 2.0; break;
            case KE: o->fnx[f] = 1; o->fny[f] = 0; o->dn[f] = fabs(o->x[nck] - o->x[i]);
                o->rx[f] = o->dx[i] / 2.0; o->ry[f] = o->y[nck] - o->y[i]; break;
            case KS: o->fnx[f] = 0; o->fny[f] = -1; o->dn[f] = fabs(o->y[nck] - o->y[i]);
                o->rx[f] = o->x[nck] - o->x[i]; o->ry[f] = -o->dy[i] / 2.0; break;
            default: o->fnx[f] = -1; o->fny[f] = 0; o->dn[f] = fabs(o->x[nck] - o->x[i]);
                o->rx[f] = -o->dx[i] / 2.0; o->ry[f] = o->y[nck] - o->y[i]; break;
            }
            /* perpendicular centroid-to-face vectors (:1592-1611) */
            o->rpx[f] = o->rx[f] - o->fnx[f] * fabs(o->rx[f]);
            o->rpy[f] = o->ry[f] - o->fny[f] * fabs(o->ry[f]);
            o->dsx[f] = o->fnx[f] * o->ds[f];
            o->dsy[f] = o->fny[f] * o->ds[f];
            o->dsxy[f] = o->ds[f] / o->dn[f];
            /* interp_coef_cell2face (simple and joint cells alike) */
            if (d == KN || d == KS) o->fintp[f] = o->dy[i] / (o->dy[i] + o->dy[nck]);
            else o->fintp[f] = o->dx[i] / (o->dx[i] + o->dx[nck]);
        }
    /* no-repeat face lists kxyface, kxface, kyface (:1528-1552) */
    o->nxy = IALLOC(nd); o->nxf = IALLOC(nd); o->nyf = IALLOC(nd);
    o->kxy = IALLOC(NF * nd); o->kxf = IALLOC(NF * nd); o->kyf = IALLOC(NF * nd);
    char *ind = (char *)xcalloc((size_t)F, 1);
    memset(ind, 1, (size_t)F);
    for (int i = 0; i < nc; i++)
        for (int k = 0; k < o->ncf[i]; k++)
            if (ind[NF * i + k]) {
                const int f = NF * i + k, nck = o->c2c[f];
                ind[f] = 0;
                ind[NF * nck + o->llec[f]] = 0;
                o->kxy[NF * i + o->nxy[i]++] = k;
                if (o->idir[f] == KE || o->idir[f] == KW) o->kxf[NF * i + o->nxf[i]++] = k;
                else o->kyf[NF * i + o->nyf[i]++] = k;
            }
    free(ind);
}

static void build_grid(cms_oracle *o)
{
    if (o->c.tel_n > 0) build_topology_tel(o);
    else build_topology_cart(o);
    geo_init(o);
    /* skewcor is turned off without joint cells (flow.F90:832); gradients then
     * take no corrections (nder 2 -> 1, derivatives.f90:109-112) */
    o->skewcor = !o->c.noskew && o->njoint > 0;
    o->ider2 = o->skewcor;
}

/* Boundary strings: the water-level string (TH_str or H_str) from its cells
 * by cellstr_bnd (bnd_lib.f90:215-316: the edge faces of every inner cell of
 * the string; the end cells take the direction of their neighbour in the
 * string), walls on every other edge face (W_str, boundaries.F90:2380-2435). */
/* The cells of a Cartesian side in string order (the order the string's
 * cellstring file lists them: along x or y from the south-west). */
static int side_cells(const cms_oracle *o, int side, int **cells)
{
    const int nx = o->nx, ny = o->ny;
    const int ns = (side == KE || side == KW) ? ny : nx;
    *cells = IALLOC(ns);
    for (int j = 0; j < ns; j++) {
        if (side == KE) (*cells)[j] = j * nx + (nx - 1);
        else if (side == KW) (*cells)[j] = j * nx;
        else if (side == KN) (*cells)[j] = (ny - 1) * nx + j;
        else (*cells)[j] = j;
    }
    return ns;
}

/* cellstr_bnd: the string's (cell, face) pairs */
static int string_faces(const cms_oracle *o, const int *cells, int ns, int **sc, int **sk)
{
    const int nc = o->ncells;
    int cap = 4 * ns + 1, n = 0;
    *sc = IALLOC(cap); *sk = IALLOC(cap);
    for (int j = 0; j < ns && ns >= 2; j++) {
        const int id = cells[j];
        const int ref = j == 0 ? cells[1] : j == ns - 1 ? cells[ns - 2] : id;
        for (int k = 0; k < o->ncf[ref]; k++) {
            if (o->c2c[NF * ref + k] < nc) continue;
            int kf = k;
            if (ref != id) {
                kf = -1;
                for (int kk = 0; kk < o->ncf[id]; kk++)
                    if (o->idir[NF * id + kk] == o->idir[NF * ref + k]) { kf = kk; break; }
            }
            (*sc)[n] = id; (*sk)[n] = kf; n++;
        }
    }
    return n;
}

static void build_bnd(cms_oracle *o)
{
    const int nc = o->ncells;
    int side = o->c.th_side;
    int ns = 0, *cells = NULL;
    if (side >= 0 && o->c.tel_n > 0) {
        ns = o->c.tel_nbnd;
        cells = IALLOC(ns);
        for (int j = 0; j < ns; j++) cells[j] = o->c.tel_bnd[j];
    } else if (side >= 0) {
        ns = side_cells(o, side, &cells);
    }
    o->nTH = string_faces(o, cells, ns, &o->THc, &o->THk);
    free(cells);
    /* the flux string (Cartesian sides): inflow along the inward normal,
     * local_angle = (4 - idirface) 90 degrees (boundaries.F90:1750-1762) */
    o->nQ = 0;
    if (o->c.q_side >= 0 && (o->c.tel_n == 0 || o->c.tel_nq > 0)) {
        if (o->c.tel_n > 0) {
            ns = o->c.tel_nq;
            cells = IALLOC(ns);
            for (int j = 0; j < ns; j++) cells[j] = o->c.tel_qcells[j];
        } else {
            ns = side_cells(o, o->c.q_side, &cells);
        }
        o->nQ = string_faces(o, cells, ns, &o->Qc, &o->Qk);
        free(cells);
        const double pi = acos(-1.0), deg2rad = pi / 180.0;
        double angle = (double)(float)((float)(4 - (o->c.q_side + 1)) * 90.0f);
        if (o->c.q_angle_on) {   /* INFLOW_ANGLE (boundaries.F90:362-370), grid azimuth 0 */
            angle = 90.0 - o->c.q_angle - 0.0;
            if (angle < 0.0) angle += 360.0;
            else if (angle > 360.0) angle -= 360.0;
        }
        o->qcos = cos(angle * deg2rad);
        o->qsin = sin(angle * deg2rad);
        o->cmvel = o->c.q_cmvel > 0.0 ? fmax(fmin(o->c.q_cmvel, 1.0), 0.5) : F4(0.667);
        o->q_inc = 1;
    } else {
        o->Qc = IALLOC(1); o->Qk = IALLOC(1);
    }
    o->wsebnd = DALLOC(o->nTH > 0 ? o->nTH : 1);
    o->wseadj = DALLOC(o->nTH > 0 ? o->nTH : 1);
    o->wseadjust = o->nTH >= 3;          /* boundaries.F90:1370,1808 */
    if (o->c.wse_ntimes > 0 && o->c.wse_adjust_off) o->wseadjust = 0;   /* WSE_ADJUSTMENT OFF, :552-554 */
    o->wse_inc = 1;
    char *claimed = (char *)xcalloc((size_t)o->ncellsD, 1);
    for (int j = 0; j < o->nTH; j++) claimed[o->c2c[NF * o->THc[j] + o->THk[j]]] = 1;
    for (int j = 0; j < o->nQ; j++) claimed[o->c2c[NF * o->Qc[j] + o->Qk[j]]] = 1;
    int n = 0;
    for (int i = 0; i < nc; i++)
        for (int k = 0; k < o->ncf[i]; k++) {
            int nck = o->c2c[NF * i + k];
            if (nck >= nc && !claimed[nck]) n++;
        }
    o->nW = n; o->Wc = IALLOC(n); o->Wk = IALLOC(n);
    n = 0;
    for (int i = 0; i < nc; i++)
        for (int k = 0; k < o->ncf[i]; k++) {
            int nck = o->c2c[NF * i + k];
            if (nck >= nc && !claimed[nck]) { o->Wc[n] = i; o->Wk[n] = k; n++; }
        }
    free(claimed);
}

/* ---------------------------------------------------------------------------
 * Gradient operator: der_grad_cbgg (der_lib.f90:272-378) + screen_weights
 * (:681-715), gow%ibc=1 (wet cells only) or goa%ibc=0; on joint cells of a
 * telescoping grid also der_grad_cbggcr (:381-504).
 * ------------------------------------------------------------------------- */
static void der_grad_cbgg(cms_oracle *o, gradop *go, int ibc)
{
    const double tolw = F4(1.0e-8);
    for (int i = 0; i < o->ncells; i++) {
        gradop *G = &go[i];
        if (o->iwet[i] == 0 && ibc == 1) { G->ncx = 0; G->ncy = 0; G->ngx = 0; G->ngy = 0; continue; }
        const int nf = o->ncf[i];
        double wx[NSC] = {0}, wy[NSC] = {0};
        int ix[NSC], iy[NSC];
        ix[0] = iy[0] = i;
        for (int k = 0; k < nf; k++) {
            int f = NF * i + k, nck = o->c2c[f];
            ix[k + 1] = iy[k + 1] = nck;
            if (o->iwet[nck] == 1 || ibc == 0) {
                wx[0] += o->dsx[f] * (1.0 - o->fintp[f]);
                wx[k + 1] = o->dsx[f] * o->fintp[f];
                wy[0] += o->dsy[f] * (1.0 - o->fintp[f]);
                wy[k + 1] = o->dsy[f] * o->fintp[f];
            } else {
                wx[0] += o->dsx[f];
                wy[0] += o->dsy[f];
            }
        }
        for (int k = 0; k <= nf; k++) { wx[k] /= o->area[i]; wy[k] /= o->area[i]; }
        G->ncx = G->ncy = 0;
        for (int k = 0; k <= nf; k++) {
            if (fabs(wx[k]) > tolw) { G->icx[G->ncx] = ix[k]; G->wcx[G->ncx] = wx[k]; G->ncx++; }
            if (fabs(wy[k]) > tolw) { G->icy[G->ncy] = iy[k]; G->wcy[G->ncy] = wy[k]; G->ncy++; }
        }
        double sx = 0.0, sy = 0.0;
        for (int k = 0; k < G->ncx; k++) sx += G->wcx[k];
        for (int k = 0; k < G->ncy; k++) sy += G->wcy[k];
        for (int k = 0; k < G->ncx; k++) G->wcx[k] -= sx / (double)G->ncx;
        for (int k = 0; k < G->ncy; k++) G->wcy[k] -= sy / (double)G->ncy;
    }
    if (!o->ider2) return;
    /* der_grad_cbggcr, joint cells: weights of the perpendicular derivative of
     * the previous evaluation, from the linear reconstruction along the faces
     * whose neighbour differs in size (rpx, rpy) */
    for (int jj = 0; jj < o->njoint; jj++) {
        const int i = o->joint[jj];
        gradop *G = &go[i];
        if (o->iwet[i] == 0 && ibc == 1) { G->ngx = 0; G->ngy = 0; continue; }
        double wxy[NSC] = {0}, wyx[NSC] = {0};
        int igx[NSC], igy[NSC];
        for (int k = 0; k < NSC; k++) igx[k] = igy[k] = i;
        igx[0] = igy[0] = i;
        for (int j = 0; j < o->nxf[i]; j++) {
            const int k = o->kxf[NF * i + j], f = NF * i + k, nck = o->c2c[f];
            if (nck >= o->ncells || fabs(o->dy[i] / o->dy[nck] - 1.0) < F4(0.1)) continue;
            if (o->iwet[nck] == 0 && ibc == 1) continue;
            const int jcn = o->llec[f];
            wxy[0] += o->dsx[f] * (1.0 - o->fintp[f]) * o->rpy[f] / o->area[i];
            wxy[k + 1] = o->dsx[f] * o->fintp[f] * o->rpy[NF * nck + jcn] / o->area[i];
            igx[k + 1] = nck;
        }
        for (int j = 0; j < o->nyf[i]; j++) {
            const int k = o->kyf[NF * i + j], f = NF * i + k, nck = o->c2c[f];
            igy[k + 1] = nck;
            if (nck >= o->ncells || fabs(o->dx[i] / o->dx[nck] - 1.0) < F4(0.1)) continue;
            if (o->iwet[nck] == 0 && ibc == 1) continue;
            const int jcn = o->llec[f];
            wyx[0] += o->dsy[f] * (1.0 - o->fintp[f]) * o->rpx[f] / o->area[i];
            wyx[k + 1] = o->dsy[f] * o->fintp[f] * o->rpx[NF * nck + jcn] / o->area[i];
        }
        G->ngx = G->ngy = 0;
        for (int k = 0; k <= o->ncf[i]; k++) {
            if (fabs(wxy[k]) > tolw) { G->igx[G->ngx] = igx[k]; G->wgxy[G->ngx] = wxy[k]; G->ngx++; }
            if (fabs(wyx[k]) > tolw) { G->igy[G->ngy] = igy[k]; G->wgyx[G->ngy] = wyx[k]; G->ngy++; }
        }
    }
}

/* der_grad_eval, der_lib.f90:726-839: first-order Green-Gauss on every cell;
 * on joint cells (nder 2) plus the deferred correction from the previous
 * perpendicular derivative, which is what gx, gy hold on entry. */
static void der_grad_eval(cms_oracle *o, const gradop *go, const double *var, double *gx, double *gy)
{
    double *gxm = NULL, *gym = NULL;
    if (o->ider2) {
        gxm = DALLOC(o->ncellsD); gym = DALLOC(o->ncellsD);
        for (int jj = 0; jj < o->njoint; jj++) {
            const int i = o->joint[jj];
            gxm[i] = gx[i]; gym[i] = gy[i];
        }
    }
    for (int i = 0; i < o->ncells; i++) {
        const gradop *G = &go[i];
        double sx = 0.0, sy = 0.0;
        for (int k = 0; k < G->ncx; k++) sx += G->wcx[k] * var[G->icx[k]];
        for (int k = 0; k < G->ncy; k++) sy += G->wcy[k] * var[G->icy[k]];
        gx[i] = sx; gy[i] = sy;
    }
    if (o->ider2) {
        for (int jj = 0; jj < o->njoint; jj++) {
            const int i = o->joint[jj];
            const gradop *G = &go[i];
            for (int k = 0; k < G->ngx; k++) gx[i] += G->wgxy[k] * gym[G->igx[k]];
            for (int k = 0; k < G->ngy; k++) gy[i] += G->wgyx[k] * gxm[G->igy[k]];
        }
        free(gxm); free(gym);
    }
}

/* interp_scal_cell2face, interp_lib.f90:2727-2906, ibc = 0, 1, 3; with the
 * cell gradients given and skewness corrections on, the faces of joint cells
 * add the linear reconstruction along the face and are clipped to the two
 * cell values (:2853-2873). */
static void interp_cell2face(cms_oracle *o, const double *phi, int ibc, double *phik, const double *dphix,
                             const double *dphiy)
{
    for (int i = 0; i < o->ncells; i++)
        for (int j = 0; j < o->nxy[i]; j++) {
            int k = o->kxy[NF * i + j], f = NF * i + k;
            int nck = o->c2c[f], jcn = o->llec[f];
            double fk = o->fintp[f], val;
            int wi = o->iwet[i], wn = o->iwet[nck];
            if (ibc == 0) {
                val = fk * phi[nck] + (1.0 - fk) * phi[i];
            } else if (ibc == 1) {
                if (wi * wn == 1) val = fk * phi[nck] + (1.0 - fk) * phi[i];
                else if (wi == 1) val = phi[i];
                else if (wn == 1) val = phi[nck];
                else val = 0.5 * (phi[i] + phi[nck]);
            } else { /* 3 */
                if (wi * wn == 1) val = fk * phi[nck] + (1.0 - fk) * phi[i];
                else val = 0.0;
            }
            phik[f] = val;
            phik[NF * nck + jcn] = val;
        }
    if (!o->skewcor || !dphiy) return;
    for (int jj = 0; jj < o->njoint; jj++) {
        const int i = o->joint[jj];
        for (int j = 0; j < o->nxy[i]; j++) {
            const int k = o->kxy[NF * i + j], f = NF * i + k, nck = o->c2c[f], jcn = o->llec[f];
            const int b = NF * nck + jcn;
            double val = phik[f];
            if (fabs(o->fnx[f]) > F4(0.999))
                val = val + o->fintp[f] * o->rpy[b] * dphiy[nck] + (1.0 - o->fintp[f]) * o->rpy[f] * dphiy[i];
            else
                val = val + o->fintp[f] * o->rpx[b] * dphix[nck] + (1.0 - o->fintp[f]) * o->rpx[f] * dphix[i];
            val = fmin(fmax(val, fmin(phi[i], phi[nck])), fmax(phi[i], phi[nck]));
            phik[f] = val;
            phik[b] = val;
        }
    }
}

/* ---------------------------------------------------------------------------
 * Wetting and drying: flow_wetdry(0) (flow.F90:1231-1539), flow_pond (:1807-1876)
 * ------------------------------------------------------------------------- */
static void flow_pond(cms_oracle *o)
{
    const int nc = o->ncells, nd = o->ncellsD;
    int *ipond = IALLOC(nd);
    for (int j = 0; j < o->nQ; j++) {
        int i = o->Qc[j], nck = o->c2c[NF * i + o->Qk[j]];
        ipond[i] = o->iwet[i];
        ipond[nck] = o->iwet[nck];
    }
    for (int j = 0; j < o->nTH; j++) {
        int i = o->THc[j], nck = o->c2c[NF * i + o->THk[j]];
        ipond[i] = o->iwet[i];
        ipond[nck] = o->iwet[nck];
    }
    int npond = 10;
    while (npond > 0) {
        npond = 0;
        for (int i = 0; i < nc; i++)
            if (ipond[i] == 1)
                for (int k = 0; k < o->ncf[i]; k++) {
                    int nck = o->c2c[NF * i + k];
                    int d = o->iwet[nck] - ipond[nck];
                    npond += d > 0 ? d : 0;
                    ipond[nck] = o->iwet[nck];
                }
        for (int i = nc - 1; i >= 0; i--)
            if (ipond[i] == 1)
                for (int k = 0; k < o->ncf[i]; k++) {
                    int nck = o->c2c[NF * i + k];
                    int d = o->iwet[nck] - ipond[nck];
                    npond += d > 0 ? d : 0;
                    ipond[nck] = o->iwet[nck];
                }
    }
    for (int i = 0; i < nd; i++)
        if (ipond[i] == 0 && o->iwet[i] == 1) o->iwet[i] = 0;
    free(ipond);
}

static void flow_wetdry(cms_oracle *o)
{
    const int nc = o->ncells, nd = o->ncellsD;
    for (int i = 0; i < nd; i++) {
        o->h[i] = o->p[i] * (1.0 / GRAV) - o->zb[i];   /* p*gravinv, gravinv = 1/grav */
        if (o->h[i] < o->hdry) { o->h[i] = fmax(o->h[i], o->hmin); o->iwet[i] = 0; }
        else o->iwet[i] = 1;
    }
    for (int j = 0; j < o->nW; j++) o->iwet[o->c2c[NF * o->Wc[j] + o->Wk[j]]] = 0;
    /* forcing strings: the ghost takes its cell's flag (:1606-1613) */
    for (int j = 0; j < o->nQ; j++) o->iwet[o->c2c[NF * o->Qc[j] + o->Qk[j]]] = o->iwet[o->Qc[j]];
    for (int j = 0; j < o->nTH; j++) o->iwet[o->c2c[NF * o->THc[j] + o->THk[j]]] = o->iwet[o->THc[j]];
    /* narrow-channel removal (Cartesian and telescoping branch, :1342-1455),
     * sequential in place; wdir(i, d) sums the wet flags of the neighbours
     * of i in direction d, wdir2 those two steps away through an active cell
     * (a ghost lists no directions) */
#define WDIR(c, d, acc) do { for (int q_ = 0; q_ < o->ncf[c]; q_++) \
        if (o->idir[NF * (c) + q_] == (d)) acc += o->iwet[o->c2c[NF * (c) + q_]]; } while (0)
#define WDIR2(c, d, acc) do { for (int q_ = 0; q_ < o->ncf[c]; q_++) { \
        const int n_ = o->c2c[NF * (c) + q_]; \
        if (o->idir[NF * (c) + q_] == (d) && n_ < nc) WDIR(n_, d, acc); } } while (0)
    int nsw = 1;
    while (nsw > 0) {
        nsw = 0;
        for (int i = 0; i < nc; i++) {
            if (o->iwet[i] == 0) continue;
            int ksum = 0;
            for (int k = 0; k < o->ncf[i]; k++) ksum += o->iwet[o->c2c[NF * i + k]];
            if (ksum <= 1) { o->iwet[i] = 0; nsw++; }
            int kns = 0, kew = 0;
            WDIR(i, KN, kns); WDIR(i, KS, kns);
            WDIR(i, KE, kew); WDIR(i, KW, kew);
            if ((kns == 0 || kew == 0) && o->h[i] < o->hdry1) { o->iwet[i] = 0; nsw++; }
            int knss = 0;
            WDIR(i, KN, knss); WDIR2(i, KS, knss);
            if (knss == 0 && o->h[i] < o->hdry2) { o->iwet[i] = 0; nsw++; }
            /* the reference tests knss a second time here (flow.F90:1403) */
            if (knss == 0 && o->h[i] < o->hdry2) { o->iwet[i] = 0; nsw++; }
            int keww = 0;
            WDIR(i, KE, keww); WDIR2(i, KW, keww);
            if (keww == 0 && o->h[i] < o->hdry2) { o->iwet[i] = 0; nsw++; }
            int keew = 0;
            WDIR(i, KW, keew); WDIR2(i, KE, keew);
            if (keew == 0 && o->h[i] < o->hdry2) { o->iwet[i] = 0; nsw++; }
        }
    }
#undef WDIR
#undef WDIR2
    flow_pond(o);
    for (int i = 0; i < nc; i++) {
        o->icorner[i] = 0;
        if (o->iwet[i] == 1) {
            int nds = 0;
            for (int k = 0; k < o->ncf[i]; k++) nds += o->iwet[o->c2c[NF * i + k]] == 0;
            if (nds > 1) o->icorner[i] = 1;
        }
    }
    /* dry-to-wet initialisation, 5 sweeps (:1477-1491) */
    double *tu = NULL, *tv = NULL, *tvis = NULL;
    if (!o->c.serial) { tu = DALLOC(nd); tv = DALLOC(nd); tvis = DALLOC(nd); }
    for (int sweep = 0; sweep < 5; sweep++) {
        const double *ru = o->u, *rv = o->v, *rvis = o->vis;
        if (!o->c.serial) {
            memcpy(tu, o->u, sizeof(double) * nd); memcpy(tv, o->v, sizeof(double) * nd);
            memcpy(tvis, o->vis, sizeof(double) * nd);
            ru = tu; rv = tv; rvis = tvis;
        }
        for (int i = 0; i < nc; i++)
            if (o->iwet[i] == 1 && o->iwet1[i] == 0) {
                const int *nb = &o->c2c[NF * i];
                double us = 0, vs = 0, viss = 0; int ns = 0;
                for (int k = 0; k < o->ncf[i]; k++) { us += ru[nb[k]]; vs += rv[nb[k]]; viss += rvis[nb[k]]; ns += o->iwet[nb[k]]; }
                o->u[i] = us / ((double)(float)ns + SMALL) / F4(1.25);
                o->v[i] = vs / ((double)(float)ns + SMALL) / F4(1.25);
                o->vis[i] = viss / ((double)(float)ns + SMALL);
            }
    }
    free(tu); free(tv); free(tvis);
    /* zero fluxes and viscosities on faces that are not wet on both sides */
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < o->nxy[i]; j++) {
            int k = o->kxy[NF * i + j], f = NF * i + k, nck = o->c2c[f], jcn = o->llec[f];
            if (o->iwet[i] * o->iwet[nck] == 0) {
                o->flux[f] = 0; o->visk[f] = 0;
                o->flux[NF * nck + jcn] = 0; o->visk[NF * nck + jcn] = 0;
            }
        }
}

/* plagr_fit (plagr_lib.f90:20-138) followed by bnd_eval's weighted sum
 * (boundaries.F90:4129-4130).  Arrays are 1-based through X(i), Y(i). */
double cms_oracle_plagr(int m, const double *x, const double *y, int ni, double xi, int *kk)
{
#define X(i) x[(i) - 1]
#define Y(i) y[(i) - 1]
    int np = ni < m - 1 ? ni : m - 1;
    int k = *kk;
    if (xi < X(1)) { *kk = 1; return Y(1); }
    if (xi > X(m)) { *kk = m - np; return Y(m); }
    if (k >= 1) {                       /* sequential search, forwards then backwards */
        int i = k;
        for (k = i; k <= m - np; k++)
            if (xi <= X(k + np)) break;
        if (k > m - np || xi < X(k) || xi > X(k + np)) {
            for (k = i; k >= 1; k--)
                if (X(k) <= xi) break;
            if (k < 1 || k + np > m || xi < X(k) || xi > X(k + np)) k = 0;
        }
    }
    if (k <= 0) {                       /* bisection */
        int i = 1, j = m - np + 1;
        while (j > i + 1) {
            k = (i + j) / 2;
            if (xi < X(k)) j = k; else i = k;
        }
        k = i;
    }
    double lb[4], sum = 0.0, w = 0.0;
    for (int i = 1; i <= np + 1; i++) {
        lb[i - 1] = 1.0;
        for (int j = 1; j <= np + 1; j++)
            if (i != j) lb[i - 1] = lb[i - 1] * (xi - X(j + k - 1)) / (X(i + k - 1) - X(j + k - 1));
        sum += lb[i - 1];
    }
    for (int i = 1; i <= np + 1; i++) w += lb[i - 1] / sum * Y(i + k - 1);
    *kk = k;
    return w;
#undef X
#undef Y
}

/* ---------------------------------------------------------------------------
 * Boundary values: bnd_eval harmonic branch (boundaries.F90:4098-4124) or the
 * single water level curve (H_str, :4127-4150; nti = 2 by default, limited to
 * 1..ntimes-1, :1417,1869)
 * ------------------------------------------------------------------------- */
static void bnd_eval(cms_oracle *o)
{
    /* flux string (:4078-4094): constant or curve (order nti limited to
     * ntimes - 1, :1783), times the ramp */
    if (o->nQ > 0) {
        double q = o->c.q_const;
        if (o->c.q_rating) {
            /* stage-flow curve (:4084-4093): the mean depth of the string's cells,
             * first-order interpolation from a bisection search */
            double stage = 0.0;
            for (int j = 0; j < o->nQ; j++) stage = stage + o->h[o->Qc[j]];
            stage = stage / (double)o->nQ;
            int k = 0;
            q = cms_oracle_plagr(o->c.q_ntimes, o->c.q_times, o->c.q_values, 1, stage, &k);
        } else if (o->c.q_ntimes > 0) {
            int ni = o->c.q_order ? o->c.q_order : 1;
            if (ni > o->c.q_ntimes - 1) ni = o->c.q_ntimes - 1;
            if (ni < 1) ni = 1;
            q = cms_oracle_plagr(o->c.q_ntimes, o->c.q_times, o->c.q_values, ni, o->timehrs, &o->q_inc);
        }
        o->qflux = o->ramp * q;
    }
    if (o->c.wse_ntimes > 0) {
        int ni = o->c.wse_order ? o->c.wse_order : 2;
        if (ni > o->c.wse_ntimes - 1) ni = o->c.wse_ntimes - 1;
        if (ni < 1) ni = 1;
        const double w = cms_oracle_plagr(o->c.wse_ntimes, o->c.wse_times, o->c.wse_values, ni, o->timehrs,
                                          &o->wse_inc);
        /* wsevar: temporally constant gradients from the string's middle cell (:1892-1901) */
        const int ir = o->THc[o->nTH / 2 - 1 >= 0 ? o->nTH / 2 - 1 : 0];
        for (int j = 0; j < o->nTH; j++) {
            const int ii = o->THc[j];
            const double wv = (o->x[ii] - o->x[ir]) * o->c.wse_gradx + (o->y[ii] - o->y[ir]) * o->c.wse_grady;
            o->wsebnd[j] = (1.0 - o->ramp) * 0.0 + o->ramp * (w + wv);
        }
        return;
    }
    for (int j = 0; j < o->nTH; j++) {
        double w = 0.0;
        for (int k = 0; k < o->c.ntc; k++)
            w += o->c.amp[k] * cos(o->c.speed[k] * o->timehrs - o->c.phase[k]);
        o->wsebnd[j] = (1.0 - o->ramp) * 0.0 + o->ramp * (w + 0.0);
    }
}

/* flow_srcsnk_static, flow_imp.F90:316-406 (ntsch=1) */
static void flow_srcsnk_static(cms_oracle *o)
{
    double gdi = 1.0 / (GRAV * o->dtime), dti = 1.0 / o->dtime;
    for (int i = 0; i < o->ncells; i++) {
        double fac1 = o->area[i] * o->h1[i] * dti;
        o->spuv0[i] = -fac1;
        o->ssu0[i] = fac1 * o->u1[i];
        o->ssv0[i] = fac1 * o->v1[i];
        o->sppp0[i] = -o->area[i] * gdi;
        o->sspp0[i] = -o->p1[i] * o->sppp0[i];
    }
    if (o->c.noptset >= 3)
        for (int i = 0; i < o->ncells; i++) {
            o->ssu0[i] += o->wavestrx[i] * o->area[i];
            o->ssv0[i] += o->wavestry[i] * o->area[i];
        }
    /* spatially constant wind, Eulerian frame (:386-394), its stress from
     * windcurve_eval (meteorological.F90:1020-1027), called each step in
     * flow_imp (:85): tau = ramp Cd rho_air/rho W (wndx, wndy), rho_air/rho
     * = 0.00120 (met_def) */
    if (o->windconst) {
        const double w = sqrt(o->wndx * o->wndx + o->wndy * o->wndy);
        const double val = o->ramp * cms_oracle_wind_drag_hsu(w) * F4(0.00120) * w;
        const double tx = val * o->wndx, ty = val * o->wndy;
        o->tauwx = tx;
        o->tauwy = ty;
        for (int i = 0; i < o->ncells; i++) {
            o->ssu0[i] += tx * o->area[i];
            o->ssv0[i] += ty * o->area[i];
        }
    }
}

/* Spatially constant wind: the 10-m wind vector (wndx, wndy) in grid
 * coordinates [m/s], as windcurve_init/windcurve_eval turn the curve's speed
 * and direction into it (meteorological.F90:700-710,1015-1019). */
int cms_oracle_set_wind(cms_oracle *o, double wndx, double wndy)
{
    o->windconst = 1;
    o->wndx = wndx;
    o->wndy = wndy;
    return 0;
}

/* coeffsourcesink_uv, flow_imp.F90:409-574 */
static void coeffsourcesink_uv(cms_oracle *o)
{
    for (int i = 0; i < o->ncells; i++) {
        o->su[i] = o->ssu0[i] + (o->fc[i] * o->v[i] - o->dpx[i]) * o->h[i] * o->area[i];
        o->sv[i] = o->ssv0[i] + (-o->fc[i] * o->u[i] - o->dpy[i]) * o->h[i] * o->area[i];
        o->cbcfuwcap[i] = o->cmb[i] * o->cfrict[i] * o->uelwc[i] * o->area[i];
        o->sp[i] = o->spuv0[i] - o->cbcfuwcap[i];
        for (int k = 0; k < o->ncf[i]; k++) {
            int f = NF * i + k;
            double ddk = o->visk[f] * o->hk[f] * o->dsxy[f];
            o->acoef[f] = cms_oracle_schmcoef(o->c.ndsch, ddk, o->flux[f]);
        }
    }
}

/* Deferred anti-diffusion corrections of the NVD schemes on a Cartesian grid
 * (flow_imp.F90:140-147): defcorhlpavec (ndsch 5, computational.F90:337-414)
 * and defcorgammavec (ndsch 6-9, :495-581).  Each face is visited once by the
 * cell listing it in kxyface; with C the upwind cell, D the downwind one and U
 * the cell upwind of C, phi~ = (phiC - phiU)/(phiD - phiU) and, for
 * 0 < phi~ < 1, the deferred normal flux is F (phiD - phiC) phi~ (HLPA) or
 * F gamma(phi~) (phif - phiC) with the central phif (gamma family); the
 * neighbour gets its negative.  HLPA skips a face whose U is a ghost or dry,
 * the gamma family reads U whatever it is (ghost values included).  Returns
 * the sums over each cell's faces in dsu, dsv (su -= dsu, sv -= dsv). */
static void defcor_nvd(const cms_oracle *o, double *dsu, double *dsv)
{
    const int nd = o->ncellsD, nc = o->ncells, ndsch = o->c.ndsch;
    double *uf = DALLOC(NF * (size_t)nd), *vf = DALLOC(NF * (size_t)nd);
    const double tiny = F4(1.0e-20);
    for (int i = 0; i < nc; i++) {
        if (o->iwet[i] == 0) continue;
        for (int j = 0; j < o->nxy[i]; j++) {
            const int k = o->kxy[NF * i + j], f = NF * i + k, nck = o->c2c[f];
            if (nck >= nc || o->iwet[nck] == 0) continue;
            const int jcn = o->llec[f];
            int nckk;
            double uc, ud, uu, vc, vd, vu;
            if (o->flux[f] >= 0.0) {
                nckk = o->c2c[NF * i + kopp[k]];
                uc = o->u[i]; ud = o->u[nck];
                vc = o->v[i]; vd = o->v[nck];
            } else {
                nckk = o->c2c[NF * nck + k];
                uc = o->u[nck]; ud = o->u[i];
                vc = o->v[nck]; vd = o->v[i];
            }
            if (ndsch == 5 && (nckk >= nc || o->iwet[nckk] == 0)) continue;
            uu = o->u[nckk]; vu = o->v[nckk];
            if (fabs(ud - uu) > tiny) {
                const double un = (uc - uu) / (ud - uu);
                if (un > 0.0 && un < 1.0) {
                    if (ndsch == 5) uf[f] = o->flux[f] * (ud - uc) * un;
                    else {
                        const double ufc = o->fintp[f] * o->u[nck] + (1.0 - o->fintp[f]) * o->u[i];
                        uf[f] = o->flux[f] * cms_oracle_nvd_gamma(ndsch, un) * (ufc - uc);
                    }
                    uf[NF * nck + jcn] = -uf[f];
                }
            }
            if (fabs(vd - vu) > tiny) {
                const double vn = (vc - vu) / (vd - vu);
                if (vn > 0.0 && vn < 1.0) {
                    if (ndsch == 5) vf[f] = o->flux[f] * (vd - vc) * vn;
                    else {
                        const double vfc = o->fintp[f] * o->v[nck] + (1.0 - o->fintp[f]) * o->v[i];
                        vf[f] = o->flux[f] * cms_oracle_nvd_gamma(ndsch, vn) * (vfc - vc);
                    }
                    vf[NF * nck + jcn] = -vf[f];
                }
            }
        }
    }
    for (int i = 0; i < nc; i++) {
        dsu[i] = ((uf[NF * i] + uf[NF * i + 1]) + uf[NF * i + 2]) + uf[NF * i + 3];
        dsv[i] = ((vf[NF * i] + vf[NF * i + 1]) + vf[NF * i + 2]) + vf[NF * i + 3];
    }
    free(uf); free(vf);
}

/* defcorparagradvec (computational.F90:149-264): deferred skewness
 * correction of the momentum sources on joint cells, called for every scheme
 * without an NVD correction on a grid that is not simple Cartesian
 * (flow_imp.F90:127-139), from the previous velocity gradients. */
static void defcor_skew(cms_oracle *o)
{
    const int nc = o->ncells;
    double *us = DALLOC(NF * (size_t)o->ncellsD), *vs = DALLOC(NF * (size_t)o->ncellsD);
    for (int jj = 0; jj < o->njoint; jj++) {
        const int i = o->joint[jj];
        if (o->iwet[i] == 0) continue;
        for (int j = 0; j < o->nxy[i]; j++) {
            const int k = o->kxy[NF * i + j], f = NF * i + k, nck = o->c2c[f];
            if (nck >= nc || o->iwet[nck] == 0) continue;
            const int b = NF * nck + o->llec[f];
            const int xdir = o->idir[f] == KE || o->idir[f] == KW;
            const double rpi = xdir ? o->rpy[f] : o->rpx[f], rpn = xdir ? o->rpy[b] : o->rpx[b];
            const double *du = xdir ? o->duy : o->dux, *dv = xdir ? o->dvy : o->dvx;
            if (fabs(rpi) > F4(1.0e-6)) {
                const double val = (o->acoef[f] + o->flux[f]) * rpi;
                us[f] = val * du[i]; vs[f] = val * dv[i];
                us[b] = -us[f]; vs[b] = -vs[f];
            } else if (fabs(rpn) > F4(1.0e-6)) {
                const double val = -o->acoef[f] * rpn;
                us[f] = val * du[nck]; vs[f] = val * dv[nck];
                us[b] = -us[f]; vs[b] = -vs[f];
            }
        }
    }
    for (int jj = 0; jj < o->njoint; jj++) {
        const int i = o->joint[jj];
        double su = 0.0, sv = 0.0;
        for (int k = 0; k < o->ncf[i]; k++) { su += us[NF * i + k]; sv += vs[NF * i + k]; }
        o->su[i] = o->su[i] - su;
        o->sv[i] = o->sv[i] - sv;
    }
    free(us); free(vs);
}

int cms_oracle_defcor(cms_oracle *o, double *dsu, double *dsv)
{
    if (o->c.ndsch < 5 || o->c.ndsch > 9) {
        memset(dsu, 0, sizeof(double) * o->ncells);
        memset(dsv, 0, sizeof(double) * o->ncells);
        return 0;
    }
    defcor_nvd(o, dsu, dsv);
    return 0;
}

/* bound_uv, boundaries.F90:3310-3490, with bnd_vel_openwse (bnd_vel.f90:35-95) */
/* bndriver_vel (boundaries.F90:4259-4356): the flux string's inflow split
 * over its faces in proportion to iwet(i) iwet(ghost) hk^cmvel / max(n, 1e-5)
 * times hk (the conveyance h^(1+cmvel)/n), along the inflow direction; the
 * ghost cells take the face velocities, the faces their fluxes. */
static void bndriver_vel(cms_oracle *o)
{
    if (o->nQ == 0) return;
    double sumw = 0.0;
    for (int j = 0; j < o->nQ; j++) {
        const int i = o->Qc[j], f = NF * i + o->Qk[j], nck = o->c2c[f];
        const double val = o->iwet[i] * o->iwet[nck] * pow(o->hk[f], o->cmvel) / fmax(o->coefman[i], F4(1.0e-5));
        o->u[nck] = val * o->qcos;
        o->v[nck] = val * o->qsin;
        sumw = sumw + o->ds[f] * o->hk[f] * (o->u[nck] * o->fnx[f] + o->v[nck] * o->fny[f]);
    }
    const double qnorm = o->qflux / fmax(fabs(sumw), F4(1.0e-6));
    for (int j = 0; j < o->nQ; j++) {
        const int i = o->Qc[j], f = NF * i + o->Qk[j], nck = o->c2c[f];
        o->u[nck] = o->u[nck] * qnorm;
        o->v[nck] = o->v[nck] * qnorm;
        o->flux[f] = o->iwet[i] * o->iwet[nck] * o->ds[f] * o->hk[f] * (o->u[nck] * o->fnx[f] + o->v[nck] * o->fny[f]);
        o->flux[NF * nck + o->llec[f]] = -o->flux[f];
    }
}

/* bndflux / fluxbnd (boundaries.F90:4359-4414,4740-4764): the flux string's
 * face fluxes from its ghost velocities */
static void bndflux(cms_oracle *o)
{
    for (int j = 0; j < o->nQ; j++) {
        const int i = o->Qc[j], f = NF * i + o->Qk[j], nck = o->c2c[f];
        o->flux[f] = o->iwet[i] * o->ds[f] * o->hk[f] * (o->fnx[f] * o->u[nck] + o->fny[f] * o->v[nck]);
        o->flux[NF * nck + o->llec[f]] = -o->flux[f];
    }
}

static void bound_uv(cms_oracle *o)
{
    /* bnd_vel_flux (bnd_vel.f90:2-32): the flux faces' upwind coefficient
     * times the ghost velocity moves into the sources */
    for (int j = 0; j < o->nQ; j++) {
        const int i = o->Qc[j], f = NF * i + o->Qk[j], nck = o->c2c[f];
        const double ddk = o->visk[f] * o->hk[f] * o->dsxy[f];
        o->acoef[f] = cms_oracle_schmcoef(1, ddk, o->flux[f]);
        o->su[i] += o->acoef[f] * o->u[nck];
        o->sv[i] += o->acoef[f] * o->v[nck];
        o->sp[i] -= o->acoef[f];
        o->acoef[f] = 0.0;
    }
    for (int j = 0; j < o->nTH; j++) {
        int i = o->THc[j], k = o->THk[j], f = NF * i + k, nck = o->c2c[f];
        double ddk = o->visk[f] * o->hk[f] * o->dsxy[f];
        o->acoef[f] = cms_oracle_schmcoef(1, ddk, o->flux[f]);
        if (fabs(o->fny[f]) > F4(0.9999)) {
            o->u[nck] = o->iwet[i] * o->u[i] * o->h[i] / o->h[nck];
            o->v[nck] = o->fny[f] * o->flux[f] / (o->ds[f] * o->hk[f]);
            o->sv[i] += o->acoef[f] * o->v[nck];
        } else {
            o->u[nck] = o->fnx[f] * o->flux[f] / (o->ds[f] * o->hk[f]);
            o->v[nck] = o->iwet[i] * o->v[i] * o->h[i] / o->h[nck];
            o->su[i] += o->acoef[f] * o->u[nck];
        }
        o->sp[i] -= o->acoef[f];
        o->acoef[f] = 0.0;
    }
    if (o->c.wallfric) {
        for (int i = 0; i < o->ncells; i++) {
            if (o->iwet[i] == 1) {
                o->spu[i] = o->sp[i];
                o->spv[i] = o->sp[i];
                for (int k = 0; k < o->ncf[i]; k++) {
                    int f = NF * i + k, nck = o->c2c[f];
                    if (o->iwet[nck] == 0) {
                        double z0w = o->z0[i] * 1.0;   /* wallfac = 1.0 */
                        if (fabs(o->fny[f]) > F4(0.999)) {
                            double delta = fabs(o->ry[f]);
                            o->spu[i] -= o->ds[f] * o->hk[f] * cms_oracle_wall_coef(delta, z0w) * fabs(o->u[i]);
                        } else {
                            double delta = fabs(o->rx[f]);
                            o->spv[i] -= o->ds[f] * o->hk[f] * cms_oracle_wall_coef(delta, z0w) * fabs(o->v[i]);
                        }
                    }
                }
            } else {
                o->u[i] = 0.0; o->v[i] = 0.0;
                o->spu[i] = -1.0; o->spv[i] = -1.0; o->sp[i] = -1.0;
                o->su[i] = 0.0; o->sv[i] = 0.0;
            }
        }
    } else {
        for (int i = 0; i < o->ncells; i++) {
            int w = o->iwet[i];
            o->u[i] = w * o->u[i];
            o->v[i] = w * o->v[i];
            o->sp[i] = w * o->sp[i] + (double)(w - 1);
            o->spu[i] = o->sp[i];
            o->spv[i] = o->spu[i];
            o->su[i] = w * o->su[i];
            o->sv[i] = w * o->sv[i];
        }
    }
}

/* ---------------------------------------------------------------------------
 * Linear solver: solve (solvers.f90:205-304) + right-preconditioned BiCGStab
 * (pbicgstab.f90:2-98).  The preconditioner is ILU(0) on the natural ordering;
 * the reference factors with ILUTP (ilutp.f90), which only changes the
 * iterates, not the converged solution.
 * ------------------------------------------------------------------------- */
typedef struct {
    int n, nd, nf;            /* cells, cells + ghosts, faces per cell (nmaxfaces) */
    const double *acoef;
    const int *c2c;
    const double *ap;
    int *ia, *ja, *dpos;
    double *lu;
} sysmat;

static void ilu0_build(sysmat *S)
{
    int n = S->n, nz = 0;
    for (int i = 0; i < n; i++) {
        S->ia[i] = nz;
        int cols[CMS_MAX_FACES + 1]; double vals[CMS_MAX_FACES + 1]; int m = 0;
        cols[m] = i; vals[m] = S->ap[i]; m++;
        for (int k = 0; k < S->nf; k++) {
            int nck = S->c2c[S->nf * i + k];
            if (nck >= 0 && nck < n) { cols[m] = nck; vals[m] = -S->acoef[S->nf * i + k]; m++; }
        }
        for (int a = 1; a < m; a++) {          /* insertion sort by column */
            int cc = cols[a]; double vv = vals[a]; int b = a - 1;
            while (b >= 0 && cols[b] > cc) { cols[b + 1] = cols[b]; vals[b + 1] = vals[b]; b--; }
            cols[b + 1] = cc; vals[b + 1] = vv;
        }
        for (int a = 0; a < m; a++) {
            S->ja[nz] = cols[a]; S->lu[nz] = vals[a];
            if (cols[a] == i) S->dpos[i] = nz;
            nz++;
        }
    }
    S->ia[n] = nz;
    /* IKJ ILU(0) */
    for (int i = 0; i < n; i++) {
        for (int pk = S->ia[i]; pk < S->dpos[i]; pk++) {
            int kc = S->ja[pk];
            double lik = S->lu[pk] / S->lu[S->dpos[kc]];
            S->lu[pk] = lik;
            int pj = S->dpos[kc] + 1, pi = pk + 1;
            while (pj < S->ia[kc + 1] && pi < S->ia[i + 1]) {
                if (S->ja[pj] == S->ja[pi]) { S->lu[pi] -= lik * S->lu[pj]; pj++; pi++; }
                else if (S->ja[pj] < S->ja[pi]) pj++;
                else pi++;
            }
        }
    }
}

static void ilu0_apply(const sysmat *S, const double *r, double *z)
{
    int n = S->n;
    for (int i = 0; i < n; i++) {
        double s = r[i];
        for (int p = S->ia[i]; p < S->dpos[i]; p++) s -= S->lu[p] * z[S->ja[p]];
        z[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = z[i];
        for (int p = S->dpos[i] + 1; p < S->ia[i + 1]; p++) s -= S->lu[p] * z[S->ja[p]];
        z[i] = s / S->lu[S->dpos[i]];
    }
}

/* y = A z with z = 0 on ghost cells (pbicgstab.f90:55,76) */
static void spmv(const sysmat *S, const double *z, double *y)
{
    for (int i = 0; i < S->n; i++) {
        double s = 0.0;
        for (int k = 0; k < S->nf; k++) {
            int nck = S->c2c[S->nf * i + k];
            if (nck >= 0 && nck < S->n) s += S->acoef[S->nf * i + k] * z[nck];
        }
        y[i] = S->ap[i] * z[i] - s;
    }
}

/* Returns Krylov iterations. mode 0: tight; mode 1: reference stopping rule
 * (L1 ratio < 1e-3 or nmax iterations). */
static int bicgstab(sysmat *S, const double *su, double *phi, int mode, int nmax,
                    double rtol, double atol, double *wk)
{
    int n = S->n;
    double *pk = wk, *r = wk + n, *uk = wk + 2 * n, *vk = wk + 3 * n, *r0 = wk + 4 * n, *zz = wk + 5 * n;
    double res0 = 0.0, n20 = 0.0;
    for (int i = 0; i < n; i++) {
        double term = 0.0;
        for (int k = 0; k < S->nf; k++) {
            int nck = S->c2c[S->nf * i + k];
            if (nck >= 0) term += S->acoef[S->nf * i + k] * phi[nck];
        }
        term = term + su[i] - S->ap[i] * phi[i];
        r[i] = term + SMALL;
        r0[i] = r[i];
        res0 += fabs(r0[i]);
        n20 += r0[i] * r0[i];
        pk[i] = uk[i] = vk[i] = 0.0;
    }
    double thr = fmax(rtol * sqrt(n20), atol * sqrt((double)n));
    if (mode == 0 && sqrt(n20) <= thr) return 0;
    ilu0_build(S);
    double alf = 1.0, bet0 = 1.0, gam = 1.0, nr0 = sqrt(n20), nr = nr0;
    int it;
    for (it = 1; it <= nmax; it++) {
        double bet = 0.0;
        for (int i = 0; i < n; i++) bet += r[i] * r0[i];
        double omg;
        if (mode == 1) {
            omg = bet * gam / (alf * bet0 + SMALL);        /* pbicgstab.f90:43 */
        } else if (fabs(bet) <= 1e-12 * nr * nr0 || it == 1) {
            /* start, or rho breakdown: restart with the shadow residual r0 = r */
            if (it > 1) {
                for (int i = 0; i < n; i++) r0[i] = r[i];
                bet = nr * nr; nr0 = nr;
            }
            omg = 0.0;
        } else {
            omg = (bet / bet0) * (gam / alf);
        }
        bet0 = bet;
        for (int i = 0; i < n; i++) pk[i] = r[i] + omg * (pk[i] - alf * uk[i]);
        ilu0_apply(S, pk, zz);
        spmv(S, zz, uk);
        double ukr0 = 0.0;
        for (int i = 0; i < n; i++) ukr0 += uk[i] * r0[i];
        if (ukr0 == 0.0) break;
        gam = bet / ukr0;
        for (int i = 0; i < n; i++) { phi[i] += gam * zz[i]; r[i] -= gam * uk[i]; }
        ilu0_apply(S, r, zz);
        spmv(S, zz, vk);
        double s1 = 0.0, s2 = 0.0;
        for (int i = 0; i < n; i++) { s1 += vk[i] * r[i]; s2 += vk[i] * vk[i]; }
        alf = mode == 1 ? s1 / fmax(s2, SMALL) : (s2 > 0.0 ? s1 / s2 : 0.0);
        double res = 0.0, n2 = 0.0;
        for (int i = 0; i < n; i++) {
            phi[i] += alf * zz[i];
            r[i] -= alf * vk[i];
            res += fabs(r[i]);
            n2 += r[i] * r[i];
        }
        nr = sqrt(n2);
        if (mode == 0 && alf == 0.0) break;
        if (mode == 1) { if (res / res0 < F4(0.001)) break; }
        else if (sqrt(n2) <= thr) break;
    }
    return it > nmax ? nmax : it;
}

/* solve(acoef,ss,sp,res,phi,n), solvers.f90:205-304 */
static int solve_sys(cms_oracle *o, double *ss, const double *sp, double *res, double *phi,
                     int n, double *rmom_out)
{
    const int nc = o->ncells;
    double relax = (n == 1) ? RELAX_PP : RELAX_UV;
    double rinv = 1.0 / relax, relaxm = 1.0 - relax;
    double *ap = o->wk + 6 * nc;
    double adak = 0.0;
    for (int i = 0; i < nc; i++) {
        double asum = 0.0, anb = 0.0;
        for (int k = 0; k < o->ncf[i]; k++) {
            asum += o->acoef[NF * i + k];
            anb += o->acoef[NF * i + k] * phi[o->c2c[NF * i + k]];
        }
        ap[i] = (asum - sp[i]) * rinv;
        ss[i] = ss[i] + relaxm * phi[i] * ap[i];
        res[i] = o->iwet[i] * (anb + ss[i] - ap[i] * phi[i]) / ap[i];
        adak += res[i] * res[i];
    }
    *rmom_out = sqrt(adak / (double)(float)nc);
    sysmat S = {nc, o->ncellsD, NF, o->acoef, o->c2c, ap, o->ia, o->ja, o->diagpos, o->lu};
    int nswp = (n == 1) ? 15 : 5;                       /* flow.F90:779-783 */
    int nmax = o->c.solver_mode == 1 ? nswp : o->c.kmax;
    return bicgstab(&S, ss, phi, o->c.solver_mode, nmax, o->c.rtol, o->c.atol, o->wk);
}

/* ---------------------------------------------------------------------------
 * Pressure correction: coefsourcesink_pp (flow_imp.F90:577-794) with
 * flow_imp_flux_noskew (flow_imp_flux.f90:361-409), bound_pp (boundaries.F90:
 * 3493-3646), bnd_wse_adjust (:4620-4686), bnd_pp (:4767-4795)
 * ------------------------------------------------------------------------- */
static void coefsourcesink_pp(cms_oracle *o)
{
    const int nc = o->ncells;
    for (int i = 0; i < nc; i++) {
        double asum = 0.0;
        for (int k = 0; k < o->ncf[i]; k++) asum += o->acoef[NF * i + k];
        double apu = o->iwet[i] * RELAX_UV / fmax(asum - o->sp[i], F4(1.0e-6));
        o->apuareap[i] = apu * o->area[i];
        o->sumu[i] = asum * apu;
        o->Hu[i] = o->iwet[i] * (o->u[i] + o->apuareap[i] * o->dpx[i] * o->h[i]);
        o->Hv[i] = o->iwet[i] * (o->v[i] + o->apuareap[i] * o->dpy[i] * o->h[i]);
    }
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < o->nxy[i]; j++) {
            int k = o->kxy[NF * i + j], f = NF * i + k, nck = o->c2c[f];
            double a;
            if (o->iwet[i] * o->iwet[nck] == 0) a = 0.0;
            else if (nck >= nc)
                a = o->hk[f] * o->hk[f] * o->ds[f] * o->apuareap[i] / (1.0 - o->sumu[i]) / o->dn[f];
            else {
                double fk = o->fintp[f], fp = 1.0 - fk;
                double apf = fk * o->apuareap[nck] + fp * o->apuareap[i];
                double suf = fk * o->sumu[nck] + fp * o->sumu[i];
                a = o->hk[f] * o->hk[f] * o->ds[f] * apf / (1.0 - suf) / o->dn[f];
            }
            o->acoef[f] = a;
            o->acoef[NF * nck + o->llec[f]] = a;
        }
    /* Rhie-Chow fluxes: flow_imp_flux_noskew (flow_imp_flux.f90:361-409), or
     * with skewness corrections flow_imp_flux_skew (:2-169): on the faces a
     * joint cell lists, the side whose perpendicular offset rp is nonzero is
     * reconstructed linearly from the gradients of H, apuareap and p */
    if (o->skewcor) {
        der_grad_eval(o, o->gow, o->Hu, o->dHux, o->dHuy);
        der_grad_eval(o, o->gow, o->Hv, o->dHvx, o->dHvy);
        der_grad_eval(o, o->gow, o->apuareap, o->dapux, o->dapuy);
    }
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < o->nxy[i]; j++) {
            int k = o->kxy[NF * i + j], f = NF * i + k, nck = o->c2c[f];
            double fl;
            if (o->iwet[i] * o->iwet[nck] == 0) fl = 0.0;
            else if (nck >= nc)
                fl = o->hk[f] * o->ds[f] * ((o->fnx[f] * o->Hu[i] + o->fny[f] * o->Hv[i])
                     - o->apuareap[i] * o->hk[f] * (o->p[nck] - o->p[i]) / o->dn[f]);
            else if (o->skewcor && o->isjoint[i]) {
                const double fk = o->fintp[f], fp = 1.0 - fk;
                const int b = NF * nck + o->llec[f];
                const int xdir = o->idir[f] == KE || o->idir[f] == KW;
                /* x faces reconstruct along y (rpy, d/dy), y faces along x */
                const double rpi = xdir ? o->rpy[f] : o->rpx[f], rpn = xdir ? o->rpy[b] : o->rpx[b];
                const double *dH = xdir ? o->dHuy : o->dHvx, *dap = xdir ? o->dapuy : o->dapux;
                const double *dp = xdir ? o->dpy : o->dpx, *H = xdir ? o->Hu : o->Hv;
                double apf, Hf, dpf;
                if (fabs(rpi) > F4(1.0e-6)) {            /* coarse to fine */
                    apf = fk * o->apuareap[nck] + fp * (o->apuareap[i] + rpi * dap[i]);
                    Hf = fk * H[nck] + fp * (H[i] + rpi * dH[i]);
                    dpf = (o->p[nck] - o->p[i] - rpi * dp[i]) / o->dn[f];
                } else if (fabs(rpn) > F4(1.0e-6)) {     /* fine to coarse */
                    apf = fk * (o->apuareap[nck] + rpn * dap[nck]) + fp * o->apuareap[i];
                    Hf = fk * (H[nck] + rpn * dH[nck]) + fp * H[i];
                    dpf = (o->p[nck] + rpn * dp[nck] - o->p[i]) / o->dn[f];
                } else {
                    apf = fk * o->apuareap[nck] + fp * o->apuareap[i];
                    Hf = fk * H[nck] + fp * H[i];
                    dpf = (o->p[nck] - o->p[i]) / o->dn[f];
                }
                fl = o->hk[f] * o->ds[f] * ((xdir ? o->fnx[f] : o->fny[f]) * Hf - apf * o->hk[f] * dpf);
            } else {
                double fk = o->fintp[f], fp = 1.0 - fk;
                double apf = fk * o->apuareap[nck] + fp * o->apuareap[i];
                double Huf = fk * o->Hu[nck] + fp * o->Hu[i];
                double Hvf = fk * o->Hv[nck] + fp * o->Hv[i];
                fl = o->hk[f] * o->ds[f] * ((o->fnx[f] * Huf + o->fny[f] * Hvf)
                     - apf * o->hk[f] * (o->p[nck] - o->p[i]) / o->dn[f]);
            }
            o->flux[f] = fl;
            o->flux[NF * nck + o->llec[f]] = -fl;
        }
    bndflux(o);
    for (int i = 0; i < nc; i++) {
        o->sp[i] = o->sppp0[i];
        double fs = 0.0;
        for (int k = 0; k < o->ncf[i]; k++) fs += o->flux[NF * i + k];
        o->su[i] = o->sspp0[i] - fs + o->p[i] * o->sp[i];
    }
    if (!o->skewcor) return;
    /* deferred skewness correction of the pp sources on joint cells
     * (flow_imp.F90:718-789), from the previous pp gradient, scaled by 0.2 */
    double *fcorr = DALLOC(NF * (size_t)o->ncellsD);
    for (int jj = 0; jj < o->njoint; jj++) {
        const int i = o->joint[jj];
        if (o->iwet[i] == 0) continue;
        for (int j = 0; j < o->nxy[i]; j++) {
            const int k = o->kxy[NF * i + j], f = NF * i + k, nck = o->c2c[f];
            if (o->iwet[nck] == 0) continue;
            const int b = NF * nck + o->llec[f];
            const int xdir = o->idir[f] == KE || o->idir[f] == KW;
            const double rpi = xdir ? o->rpy[f] : o->rpx[f], rpn = xdir ? o->rpy[b] : o->rpx[b];
            const double *dpp = xdir ? o->dppy : o->dppx;
            if (fabs(rpi) > F4(1.0e-6)) {
                fcorr[f] = -o->acoef[f] * rpi * dpp[i];
                fcorr[b] = -fcorr[f];
            } else if (fabs(rpn) > F4(1.0e-6)) {
                fcorr[f] = o->acoef[f] * rpn * dpp[nck];
                fcorr[b] = -fcorr[f];
            }
        }
    }
    for (int jj = 0; jj < o->njoint; jj++) {
        const int i = o->joint[jj];
        double fs = 0.0;
        for (int k = 0; k < o->ncf[i]; k++) fs += fcorr[NF * i + k];
        o->su[i] = o->su[i] + F4(0.2) * fs;
    }
    free(fcorr);
}

static void bound_pp(cms_oracle *o)
{
    const int nc = o->ncells;
    for (int i = 0; i < nc; i++) {
        if (o->iwet[i] != 0) continue;
        int sumwet = 0;
        for (int k = 0; k < o->ncf[i]; k++) { int nck = o->c2c[NF * i + k]; if (nck < nc) sumwet += o->iwet[nck]; }
        double aploc = 0.0;
        o->su[i] = 0.0;
        if (sumwet >= 1) {
            for (int k = 0; k < o->ncf[i]; k++) {
                int f = NF * i + k, nck = o->c2c[f];
                o->acoef[f] = o->iwet[nck];
                aploc += o->acoef[f];
                o->su[i] += o->acoef[f] * o->p[nck];
            }
        } else {
            for (int k = 0; k < o->ncf[i]; k++) {
                int f = NF * i + k, nck = o->c2c[f];
                if (nck >= nc) o->acoef[f] = 0.0;
                else { o->acoef[f] = o->dsxy[f]; aploc += o->acoef[f]; o->su[i] += o->acoef[f] * o->p[nck]; }
            }
        }
        o->su[i] -= aploc * o->p[i];
        o->sp[i] = 0.0;
    }
    for (int j = 0; j < o->nQ; j++) o->acoef[NF * o->Qc[j] + o->Qk[j]] = 0.0;
    if (o->nTH > 0) {
        const double *wse = o->wsebnd;
        if (o->wseadjust) {
            /* bnd_wse_adjust: setup integrated from the centre cell outward */
            int n = o->nTH, im0 = n / 2 - 1;           /* 1-based nbndcells/2 */
            double *ws = DALLOC(n);
            ws[im0] = 0.0;
            for (int im = im0 + 1; im < n; im++) {
                int i = o->THc[im], i1 = o->THc[im - 1], f = NF * i + o->THk[im];
                double fx = o->c.noptset >= 3 ? o->wavestrx[i] : 0.0, fx1 = o->c.noptset >= 3 ? o->wavestrx[i1] : 0.0;
                double fy = o->c.noptset >= 3 ? o->wavestry[i] : 0.0, fy1 = o->c.noptset >= 3 ? o->wavestry[i1] : 0.0;
                if (o->windconst) {   /* forcex/forcey add the wind stress, boundaries.F90:3978-3982 */
                    fx += o->tauwx; fx1 += o->tauwx; fy += o->tauwy; fy1 += o->tauwy;
                }
                double sx = (o->x[i] - o->x[i1]) * (fx + fx1) / (o->h[i] + o->h[i1]) * (1.0 / GRAV);
                double sy = (o->y[i] - o->y[i1]) * (fy + fy1) / (o->h[i] + o->h[i1]) * (1.0 / GRAV);
                if (fabs(o->ry[f]) >= F4(0.999)) ws[im] = ws[im - 1] + sx;
                else if (fabs(o->rx[f]) >= F4(0.999)) ws[im] = ws[im - 1] + sy;
                else ws[im] = ws[im - 1] + sx + sy;
            }
            for (int im = im0 - 1; im >= 0; im--) {
                int i = o->THc[im], i1 = o->THc[im + 1], f = NF * i + o->THk[im];
                double fx = o->c.noptset >= 3 ? o->wavestrx[i] : 0.0, fx1 = o->c.noptset >= 3 ? o->wavestrx[i1] : 0.0;
                double fy = o->c.noptset >= 3 ? o->wavestry[i] : 0.0, fy1 = o->c.noptset >= 3 ? o->wavestry[i1] : 0.0;
                if (o->windconst) {   /* forcex/forcey add the wind stress, boundaries.F90:3978-3982 */
                    fx += o->tauwx; fx1 += o->tauwx; fy += o->tauwy; fy1 += o->tauwy;
                }
                double sx = (o->x[i] - o->x[i1]) * (fx + fx1) / (o->h[i] + o->h[i1]) * (1.0 / GRAV);
                double sy = (o->y[i] - o->y[i1]) * (fy + fy1) / (o->h[i] + o->h[i1]) * (1.0 / GRAV);
                if (fabs(o->ry[f]) >= F4(0.999)) ws[im] = ws[im + 1] + sx;
                else if (fabs(o->rx[f]) >= F4(0.999)) ws[im] = ws[im + 1] + sy;
                else ws[im] = ws[im + 1] + sx + sy;
            }
            for (int im = 0; im < n; im++) o->wseadj[im] = o->wsebnd[im] + ws[im];
            free(ws);
            wse = o->wseadj;
        }
        for (int j = 0; j < o->nTH; j++) {
            int i = o->THc[j], f = NF * i + o->THk[j], nck = o->c2c[f];
            o->pp[nck] = wse[j] * GRAV - o->p[nck];
            o->su[i] += o->acoef[f] * o->pp[nck];
            o->sp[i] -= o->acoef[f];
            o->acoef[f] = 0.0;
        }
    }
}

/* pvfc, flow_imp.F90:797-1072 (no skewness terms on Cartesian grids) */
static void pvfc(cms_oracle *o)
{
    const int nc = o->ncells, nd = o->ncellsD;
    for (int j = 0; j < o->nW; j++) {
        int i = o->Wc[j];
        o->pp[o->c2c[NF * i + o->Wk[j]]] = o->pp[i];
    }
    for (int j = 0; j < o->nQ; j++) {
        int i = o->Qc[j];
        o->pp[o->c2c[NF * i + o->Qk[j]]] = o->pp[i];
    }
    der_grad_eval(o, o->gow, o->pp, o->dppx, o->dppy);
    for (int i = 0; i < nc; i++) {
        if (o->iwet[i] == 0) continue;
        double val = o->h[i] * o->apuareap[i] / (1.0 - o->sumu[i] + SMALL);
        o->u[i] -= val * o->dppx[i];
        o->v[i] -= val * o->dppy[i];
        for (int j = 0; j < o->nxy[i]; j++) {
            int k = o->kxy[NF * i + j], f = NF * i + k, nck = o->c2c[f];
            if (o->iwet[nck] == 0) continue;
            o->flux[f] -= o->acoef[f] * (o->pp[nck] - o->pp[i]);
            o->flux[NF * nck + o->llec[f]] = -o->flux[f];
        }
    }
    /* skewness fixes of the corrected fluxes on joint cells (flow_imp.F90:912-962) */
    for (int jj = 0; o->skewcor && jj < o->njoint; jj++) {
        const int i = o->joint[jj];
        if (o->iwet[i] == 0) continue;
        for (int j = 0; j < o->nxy[i]; j++) {
            const int k = o->kxy[NF * i + j], f = NF * i + k, nck = o->c2c[f];
            if (nck >= nc || o->iwet[nck] == 0) continue;
            const int b = NF * nck + o->llec[f];
            const int xdir = o->idir[f] == KE || o->idir[f] == KW;
            const double rpi = xdir ? o->rpy[f] : o->rpx[f], rpn = xdir ? o->rpy[b] : o->rpx[b];
            const double *dpp = xdir ? o->dppy : o->dppx;
            if (fabs(rpi) > F4(1.0e-6)) {
                o->flux[f] = o->flux[f] + o->acoef[f] * rpi * dpp[i];
                o->flux[b] = -o->flux[f];
            } else if (fabs(rpn) > F4(1.0e-6)) {
                o->flux[f] = o->flux[f] - o->acoef[f] * rpn * dpp[nck];
                o->flux[b] = -o->flux[f];
            }
        }
    }
    for (int i = 0; i < nd; i++) {
        o->p[i] += RELAX_PP * o->pp[i];
        o->h[i] = fmax(o->hmin, o->p[i] * (1.0 / GRAV) - o->zb[i]);
        o->pp[i] = FACPP * o->pp[i];
    }
    for (int j = 0; j < o->nW; j++) {
        int i = o->Wc[j];
        o->p[o->c2c[NF * i + o->Wk[j]]] = o->p[i];
    }
    for (int j = 0; j < o->nQ; j++) {
        int i = o->Qc[j], nck = o->c2c[NF * i + o->Qk[j]];
        o->p[nck] = o->p[i];
        o->h[nck] = fmax(o->hmin, o->p[nck] * (1.0 / GRAV) - o->zb[nck]);
    }
}

/* flow_grad_interp, flow_imp.F90:1075-1130 */
static void flow_grad_interp(cms_oracle *o)
{
    der_grad_eval(o, o->gow, o->u, o->dux, o->duy);
    der_grad_eval(o, o->gow, o->v, o->dvx, o->dvy);
    interp_cell2face(o, o->p, 1, o->pk, o->dpx, o->dpy);
    for (int i = 0; i < o->ncells; i++) {
        double sx = 0.0, sy = 0.0;
        for (int k = 0; k < o->ncf[i]; k++) {
            int f = NF * i + k;
            sx += o->dsx[f] * o->pk[f];
            sy += o->dsy[f] * o->pk[f];
            o->hk[f] = fmax(o->hmin, o->pk[f] * (1.0 / GRAV) - o->zbk[f]);
        }
        o->dpx[i] = sx / o->area[i];
        o->dpy[i] = sy / o->area[i];
    }
}

/* fric_eval, friction.F90:615-797 (mbedfric 1-4; mwavcurint=1 or no waves).
 * The law's input (cfrict, coefman, z0 or the linear coefficient cbf) is the
 * array fric_init filled (friction.F90:340-390); the others follow from it. */
static void fric_eval(cms_oracle *o)
{
    for (int i = 0; i < o->ncells; i++) {
        const double h = o->h[i];
        switch (o->c.mbedfric) {
        case 1:   /* :639-645 */
            o->coefman[i] = cms_oracle_drag2man(GRAV, o->cfrict[i], h, HMAXFRIC);
            o->z0[i] = cms_oracle_drag2length(o->cfrict[i], h);
            break;
        case 3:   /* :664-670 */
            o->cfrict[i] = cms_oracle_length2drag(o->z0[i], h);
            o->coefman[i] = cms_oracle_drag2man(GRAV, o->cfrict[i], h, HMAXFRIC);
            break;
        case 4:   /* :672-679: uelwc is the previous evaluation's */
            o->cfrict[i] = o->cbf[i] * h / fmax(o->uelwc[i], F4(1.0e-15));
            o->coefman[i] = cms_oracle_drag2man(GRAV, o->cfrict[i], h, h);
            o->z0[i] = cms_oracle_drag2length(o->cfrict[i], h);
            break;
        default:  /* 2, :647-662 */
            o->cfrict[i] = cms_oracle_man2drag(GRAV, o->coefman[i], h, HMAXFRIC);
            if (o->cfrict[i] != 0.0) o->z0[i] = cms_oracle_drag2length(o->cfrict[i], h);
        }
    }
    for (int i = 0; i < o->ncells; i++) {
        double Um, Uwc, taum;
        if (o->c.noptset >= 3) {
            /* fric_wavcurmean_quad with us=vs=Worb=0 (fric_lib.f90:840-874) */
            double Um2 = o->u[i] * o->u[i] + o->v[i] * o->v[i];
            Um = sqrt(Um2);
            Uwc = sqrt(Um2 + CFRICWAV * 0.0 * 0.0);
            taum = RHOW * o->cfrict[i] * Um * Uwc;
        } else {
            Um = sqrt(o->u[i] * o->u[i] + o->v[i] * o->v[i]);
            Uwc = Um;
            taum = RHOW * o->cfrict[i] * Um * Um;
        }
        o->uv[i] = Um; o->uelwc[i] = Uwc; o->bsxy[i] = taum;
        o->bsvel[i] = sqrt(taum / RHOW + F4(1.0e-15));
    }
}

/* flow_eddyvis, flow_eddyvis.f90:2-225 (Cartesian mesh): TURBULENCE_MODEL
 * CONSTANT (mturbul 0, :24-29), FALCONER (1, :31-48), PARABOLIC (2, :50-67),
 * SUBGRID-WU (3, :69-98) and SUBGRID (5, the default, :133-160).  The
 * coefficients keep the values flow_default gives them for mturbul 5
 * (flow.F90:81-92) whichever model a card selects later. */
static double eddy_waves(const cms_oracle *o, int i, double vv)
{
    if (o->c.noptset < 3) return vv;
    vv += CVISWAV * 0.0 * 0.0;              /* cviswav Whgt Worb: no wave height on this path */
    vv += CVISWAVBRK * o->h[i] * pow(fmax(0.0, F4(1.0e-15)) / RHOW, F4(0.33333));
    return vv;
}

static void flow_eddyvis(cms_oracle *o)
{
    const int nc = o->ncells, nd = o->ncellsD, mt = o->c.mturbul;
    const double c2a = CVISHOR * CVISHOR * o->areaavg, c2a2 = c2a * c2a;
    for (int i = 0; i < nc; i++) {
        if (mt == 0) {
            o->vis[i] = CVISCON;
        } else if (mt == 1) {
            const double vv = CVISCON + CVISBOT * o->cfrict[i] * o->uv[i] * o->h[i];
            o->vis[i] = fmin(eddy_waves(o, i, vv), CVISMAX);
        } else if (mt == 2) {
            const double vv = CVISCON + CVISBOT * o->bsvel[i] * o->h[i];
            o->vis[i] = fmin(eddy_waves(o, i, vv), CVISMAX);
        } else if (mt == 3) {
            if (o->iwet[i] == 1) {
                const double ss12 = 0.5 * (o->duy[i] + o->dvx[i]);
                const double stens = 2.0 * (o->dux[i] * o->dux[i] + 2.0 * ss12 * ss12 + o->dvy[i] * o->dvy[i]);
                const double b = CVISBOT * o->bsvel[i] * o->h[i];
                const double vv = CVISCON + sqrt(c2a2 * stens + b * b);
                o->vis[i] = fmin(eddy_waves(o, i, vv), CVISMAX);
            } else {
                o->vis[i] = F4(1.0e-6);
            }
        } else if (o->iwet[i] == 1) {
            double stens = sqrt(2.0 * (o->dux[i] * o->dux[i] + o->dvy[i] * o->dvy[i])
                                + (o->duy[i] + o->dvx[i]) * (o->duy[i] + o->dvx[i]));
            double vv = CVISCON + CVISBOT * o->bsvel[i] * o->h[i] + c2a * stens;
            o->vis[i] = fmin(eddy_waves(o, i, vv), CVISMAX);
        } else {
            o->vis[i] = F4(1.0e-6);
        }
    }
    /* bndcopy2ghost: forcing strings */
    for (int j = 0; j < o->nQ; j++) {
        int i = o->Qc[j];
        o->vis[o->c2c[NF * i + o->Qk[j]]] = o->vis[i];
    }
    for (int j = 0; j < o->nTH; j++) {
        int i = o->THc[j];
        o->vis[o->c2c[NF * i + o->THk[j]]] = o->vis[i];
    }
    /* boundaries near dry cells (Cartesian and telescoping branch, :166-186) */
    double *rv = o->vis;
    double *old = NULL;
    if (!o->c.serial) { old = DALLOC(nd); memcpy(old, o->vis, sizeof(double) * nd); rv = old; }
    for (int i = 0; i < nc; i++) {
        if (o->iwet[i] != 1) continue;
        for (int k = 0; k < o->ncf[i]; k++) {
            int nck = o->c2c[NF * i + k];
            if (o->iwet[nck] == 0 || nck >= nc) {
                /* average over the faces on the opposite side */
                const int kkdf = kopp[o->idir[NF * i + k]];
                double cid = 0.0, vissum = 0.0;
                for (int kk = 0; kk < o->ncf[i]; kk++)
                    if (o->idir[NF * i + kk] == kkdf) {
                        cid += 1.0;
                        vissum += rv[o->c2c[NF * i + kk]] - CVISCON;
                    }
                o->vis[i] = CVISCON + vissum / (cid + SMALL);
                o->vis[nck] = o->vis[i];
            }
        }
    }
    /* corner wet nodes */
    if (!o->c.serial) memcpy(old, o->vis, sizeof(double) * nd);
    for (int i = 0; i < nc; i++) {
        if (o->icorner[i] != 1) continue;
        double cid = 0.0, vs = 0.0;
        for (int k = 0; k < o->ncf[i]; k++) {
            int nck = o->c2c[NF * i + k];
            cid += o->iwet[nck];
            vs += (rv[nck] - CVISCON) * o->iwet[nck];
        }
        o->vis[i] = CVISCON + vs / (cid + SMALL);
    }
    free(old);
    interp_cell2face(o, o->vis, 3, o->visk, NULL, NULL);
}

/* check_conv, flow_imp.F90:1576-1747 */
static void check_conv(cms_oracle *o)
{
    double *rmom = o->rmom, *r0 = o->rmom0, *r1 = o->rmom1;
    int niter = o->niter;
    o->iconv = 1;
    if (niter <= 2) {
        for (int n = 0; n < 3; n++) r0[n] = fmax(rmom[n], F4(1e-15));
        for (int n = 0; n < 3; n++) r1[n] = r0[n];
    }
    double rmax = fmax(rmom[0], fmax(rmom[1], rmom[2]));
    if (rmax > 5.0) { o->iconv = 0; return; }
    if (niter <= 5) return;
    int nhalf = o->maxit / 2;
    double rp = rmom[0] / r0[0];
    double ruv = fmax(rmom[1] / r0[1], rmom[2] / r0[2]);
    if ((rp >= RMOMRATIOP * 10.0 || ruv >= RMOMRATIOUV * 10.0) && rmom[0] > RMOMTARGETP) { o->iconv = 0; return; }
    if (niter == o->maxit && (rp >= RMOMRATIOP || ruv >= RMOMRATIOUV) && rmom[0] > RMOMMAXP) { o->iconv = 0; return; }
    if (niter == o->maxit && (rp >= RMOMRATIOP / 5.0 || ruv >= RMOMRATIOUV / 5.0) && rmom[0] > RMOMTARGETP) {
        o->iconv = o->dtvar ? 3 : 0;                 /* 3: reduce the next time step */
        return;
    }
    double rmomuv = fmax(rmom[1], rmom[2]);
    if (rmom[0] <= RMOMMINP && rmomuv <= RMOMMINUV) { o->iconv = 2; return; }
    if (niter == o->maxit && (rmom[0] > RMOMMAXP || rmomuv > RMOMMAXUV)) { o->iconv = 0; return; }
    if (niter == o->maxit && (rmom[0] > RMOMTARGETP || rmomuv > RMOMTARGETUV)) { o->iconv = o->dtvar ? 3 : 0; return; }
    if (niter <= nhalf || niter % 2 != 0) return;
    double rchp = r1[0] - rmom[0];
    double rchuv = fmax(r1[1], r1[2]) - rmomuv;
    if (rchp < 0.0 || rchuv < 0.0) return;
    if (rchp <= RMOMMINP && rchuv <= RMOMMINUV) { o->iconv = 2; return; }
    for (int n = 0; n < 3; n++) r1[n] = rmom[n];
}

/* velface_eta, flow_imp.F90:1245-1573 (Cartesian branch) */
static void velface_eta(cms_oracle *o)
{
    for (int i = 0; i < o->ncells; i++) {
        if (o->h[i] > o->hmin && o->iwet[i] == 1) {
            double umin = 1.0e6, umax = -1.0e6, vmin = 1.0e6, vmax = -1.0e6;
            for (int k = 0; k < o->ncf[i]; k++) {
                int f = NF * i + k;
                double vel = o->flux[f] / (o->hk[f] * o->ds[f]);
                double uf = o->fnx[f] * vel, vf = o->fny[f] * vel;
                umin = fmin(umin, uf); umax = fmax(umax, uf);
                vmin = fmin(vmin, vf); vmax = fmax(vmax, vf);
            }
            o->u[i] = fmin(fmax(o->u[i], umin), umax);
            o->v[i] = fmin(fmax(o->v[i], vmin), vmax);
            o->eta[i] = o->p[i] * (1.0 / GRAV);
            o->uv[i] = sqrt(o->u[i] * o->u[i] + o->v[i] * o->v[i]);
        } else {
            o->h[i] = fmax(o->hmin, o->h[i]);
            o->u[i] = 0.0; o->v[i] = 0.0; o->uv[i] = 0.0;
            o->eta[i] = -999.0;
        }
    }
}

/* flow_recalculate, flow_imp.F90:1133-1242: back to the previous time level
 * with half the time step (the ramp keeps the value timing gave it).
 * Returns -1 when the step would fall below dtimebeg/2^8 (divergent). */
static int flow_recalculate(cms_oracle *o)
{
    o->mtime = 1;
    o->jtime += 1;
    o->timesecs = fmax(0.0, o->timesecs - o->deltime);
    o->deltime = o->dtimebeg / ldexp(1.0, o->jtime);
    o->timesecs += o->deltime;
    o->dtime = o->deltime;
    o->ctime = o->timesecs;
    o->timehrs = o->ctime / 3600.0;
    if (o->dtime <= o->dtimebeg / 256.0) return -1;
    const int nd = o->ncellsD;
    for (int i = 0; i < nd; i++) {
        o->u[i] = o->u1[i]; o->v[i] = o->v1[i]; o->p[i] = o->p1[i]; o->h[i] = o->h1[i];
        o->pp[i] = 0.0; o->dppx[i] = 0.0; o->dppy[i] = 0.0;
        o->iwet[i] = o->iwet1[i];
    }
    memcpy(o->flux, o->flux1, sizeof(double) * NF * nd);
    flow_wetdry(o);
    bndflux(o);
    der_grad_cbgg(o, o->gow, 1);
    flow_grad_interp(o);
    fric_eval(o);
    flow_eddyvis(o);
    return 0;
}

/* flow_imp, flow_imp.F90:29-312; with dtvar a divergent outer iteration
 * recalculates the step with half the time step (:262-272) */
static int flow_imp(cms_oracle *o)
{
restart:
    bnd_eval(o);
    flow_wetdry(o);
    der_grad_cbgg(o, o->gow, 1);                 /* der_update: weights follow iwet */
    flow_srcsnk_static(o);
    o->kry[0] = o->kry[1] = o->kry[2] = 0;
    for (o->niter = 1; o->niter <= o->maxit; o->niter++) {
        bndriver_vel(o);
        coeffsourcesink_uv(o);
        if (o->c.ndsch >= 5 && o->c.ndsch <= 9) {   /* deferred corrections, flow_imp.F90:140-147 */
            double *dsu = DALLOC(o->ncells), *dsv = DALLOC(o->ncells);
            defcor_nvd(o, dsu, dsv);
            for (int i = 0; i < o->ncells; i++) { o->su[i] -= dsu[i]; o->sv[i] -= dsv[i]; }
            free(dsu); free(dsv);
        } else if (o->njoint > 0 && o->c.ndsch != 0) {
            defcor_skew(o);
        }
        bound_uv(o);
        o->kry[1] += solve_sys(o, o->su, o->spu, o->rsu, o->u, 2, &o->rmom[1]);
        o->kry[2] += solve_sys(o, o->sv, o->spv, o->rsv, o->v, 3, &o->rmom[2]);
        coefsourcesink_pp(o);
        bound_pp(o);
        o->kry[0] += solve_sys(o, o->su, o->sp, o->rsp, o->pp, 1, &o->rmom[0]);
        pvfc(o);
        flow_grad_interp(o);
        bndriver_vel(o);
        bndflux(o);
        fric_eval(o);
        flow_eddyvis(o);
        check_conv(o);
        if (o->iconv == 0) {
            if (!o->dtvar || flow_recalculate(o) != 0) return -1;
            goto restart;
        }
        if (o->iconv == 2) break;
    }
    if (o->niter > o->maxit) o->niter = o->maxit;
    velface_eta(o);
    fric_eval(o);
    return 0;
}

/* flow_update, CMS_Flow.f90:243-411 */
static void flow_update(cms_oracle *o)
{
    const int nd = o->ncellsD;
    for (int i = 0; i < nd; i++) {
        o->iwet1[i] = o->iwet[i];
        o->u1[i] = o->u[i]; o->v1[i] = o->v[i]; o->p1[i] = o->p[i]; o->h1[i] = o->h[i];
        o->pp[i] = 0.0; o->dppx[i] = 0.0; o->dppy[i] = 0.0;
    }
    memcpy(o->flux1, o->flux, sizeof(double) * NF * nd);
    /* maximum water level, :352-355 (maxeta starts at 0, flow.F90:874) */
    for (int i = 0; i < nd; i++) o->maxeta[i] = fmax(o->eta[i], o->maxeta[i]);
}

/* boundary inflow volume over one step from the final boundary fluxes */
static double bnd_inflow(const cms_oracle *o)
{
    double q = 0.0;
    for (int j = 0; j < o->nQ; j++) q -= o->flux[NF * o->Qc[j] + o->Qk[j]];
    for (int j = 0; j < o->nTH; j++) q -= o->flux[NF * o->THc[j] + o->THk[j]];
    return q * o->dtime;
}

/* flow_convflux, flow.F90:1879-1909 */
static void flow_convflux(cms_oracle *o)
{
    for (int i = 0; i < o->ncells; i++)
        for (int j = 0; j < o->nxy[i]; j++) {
            int k = o->kxy[NF * i + j], f = NF * i + k, nck = o->c2c[f];
            double uni = o->fnx[f] * o->u[i] + o->fny[f] * o->v[i];
            double unk = o->fnx[f] * o->u[nck] + o->fny[f] * o->v[nck];
            o->flux[f] = o->iwet[i] * o->iwet[nck] * o->ds[f] * o->hk[f]
                         * (o->fintp[f] * unk + (1.0 - o->fintp[f]) * uni);
            o->flux[NF * nck + o->llec[f]] = -o->flux[f];
        }
    bndflux(o);
}

static void init_state(cms_oracle *o);
static void stat_update_flow(cms_oracle *o);

cms_oracle *cms_oracle_create(const cms_case *c)
{
    cms_oracle *o = (cms_oracle *)xcalloc(1, sizeof(cms_oracle));
    o->c = *c;
    o->nx = c->nx; o->ny = c->ny;
    build_grid(o);
    const int nd = o->ncellsD, nc = o->ncells, F = NF * nd;
#define A(f) o->f = DALLOC(nd)
    A(u); A(v); A(p); A(h); A(pp); A(dppx); A(dppy); A(dux); A(duy); A(dvx); A(dvy); A(dpx); A(dpy);
    A(vis); A(eta); A(u1); A(v1); A(p1); A(h1);
    A(su); A(sv); A(sp); A(spu); A(spv); A(ssu0); A(ssv0); A(spuv0); A(sppp0); A(sspp0);
    A(apuareap); A(sumu); A(Hu); A(Hv); A(rsu); A(rsv); A(rsp);
    A(coefman); A(cfrict); A(z0); A(uv); A(uelwc); A(bsxy); A(bsvel); A(cmb); A(cbcfuwcap);
    A(wavestrx); A(wavestry); A(fc); A(cbf); A(maxeta);
    A(dHux); A(dHuy); A(dHvx); A(dHvy); A(dapux); A(dapuy); A(dzbx); A(dzby);
#undef A
    o->iwet = IALLOC(nd); o->iwet1 = IALLOC(nd); o->icorner = IALLOC(nd);
    o->flux = DALLOC(F); o->flux1 = DALLOC(F); o->hk = DALLOC(F); o->pk = DALLOC(F);
    o->visk = DALLOC(F); o->zbk = DALLOC(F); o->acoef = DALLOC(F);
    o->gow = (gradop *)xcalloc((size_t)nc, sizeof(gradop));
    o->goa = (gradop *)xcalloc((size_t)nc, sizeof(gradop));
    o->wk = DALLOC(7 * (size_t)nc);
    o->ia = IALLOC(nc + 1); o->ja = IALLOC((NF + 1) * nc); o->diagpos = IALLOC(nc);
    o->lu = DALLOC((NF + 1) * (size_t)nc);
    /* fric_init (friction.F90:315-319): coefman = cfrict = 0, z0 = 1e-5, then
     * the law's input array from the dataset */
    for (int i = 0; i < nd; i++) o->z0[i] = F4(1.0e-5);
    for (int i = 0; i < nc; i++) {
        switch (c->mbedfric) {
        case 1: o->cfrict[i] = c->mann[i]; break;
        case 3: o->z0[i] = c->mann[i]; break;
        case 4: o->cbf[i] = c->mann[i]; break;
        default: o->coefman[i] = c->mann[i];
        }
        o->cmb[i] = 1.0;
        o->wavestrx[i] = c->wavestrx ? c->wavestrx[i] : 0.0;
        o->wavestry[i] = c->wavestry ? c->wavestry[i] : 0.0;
    }
    /* fric_init's copy of the friction to the dummy cells (friction.F90:397-407)
     * loops `do i=1,ncells+1,ncellsD`: one pass, i = 1, so the first active
     * cell takes cfrict, coefman and z0 of its first active neighbour in face
     * order, and the dummy cells keep theirs. */
    for (int k = 0; k < o->ncf[0]; k++) {
        const int nck = o->c2c[NF * 0 + k];
        if (nck < nc) { o->cfrict[0] = o->cfrict[nck]; o->coefman[0] = o->coefman[nck]; o->z0[0] = o->z0[nck]; break; }
    }
    /* Coriolis, flow.F90:720-736: f-plane fc = 2 omega sin(lat), beta-plane
     * fc + 2 omega cos(lat)/R (y - mean y); omega = 2 pi/86400, R = 6371 km
     * (definitions.F90:86-96) */
    {
        const double pi = acos(-1.0), omega = 2.0 * pi / 86400.0, deg2rad = pi / 180.0;
        const double f0 = 2.0 * omega * sin(c->latitude * deg2rad);
        if (c->coriolis_plane == 1) {
            const double beta = 2.0 * omega * cos(c->latitude * deg2rad) / 6371000.0;
            double ysum = 0.0;
            for (int i = 0; i < nc; i++) ysum += o->y[i];
            const double yavg = ysum / (double)nc;
            for (int i = 0; i < nd; i++) o->fc[i] = f0 + beta * (o->y[i] - yavg);
        } else {
            for (int i = 0; i < nd; i++) o->fc[i] = f0;
        }
    }
    build_bnd(o);
    /* flow_init (flow.F90:592-835) */
    o->dtime = c->dt;
    o->dtvar = c->dtvar;
    o->dtimebeg = c->dt;
    o->deltime = c->dt;
    o->jtime = 0;
    o->mtime = 0;
    o->hdry = c->hdry > 0 ? c->hdry : F4(0.05);
    o->hmin = o->hdry;
    o->hdry1 = F4(5.0) * o->hdry;
    o->hdry2 = F4(2.0) * o->hdry;
    o->rampdur = fmax(c->rampdur, o->dtime / 3600.0);
    o->maxit = c->maxit > 0 ? c->maxit : 20 + nc / 10000;
    o->timesecs = 0.0; o->ctime = 0.0; o->timehrs = 0.0;
    for (int n = 0; n < 3; n++) { o->rmom0[n] = 1.e10; o->rmom[n] = 0.0; }
    if (o->c.rtol <= 0) o->c.rtol = 1e-10;
    if (o->c.kmax <= 0) o->c.kmax = 2000;
    /* zbk: every face of every active cell, with the all-cell gradient
     * operator goa for the joint-cell reconstruction (prestart.f90:309-329) */
    der_grad_cbgg(o, o->goa, 0);
    der_grad_eval(o, o->goa, o->zb, o->dzbx, o->dzby);
    interp_cell2face(o, o->zb, 0, o->zbk, o->dzbx, o->dzby);
    init_state(o);
    return o;
}

/* Initial state from u, v, p as flow_init and prestart build it
 * (flow.F90:688-707, prestart.f90:309-329): wet/dry, previous time level,
 * gradients, face values, fluxes (flow_convflux), friction, eddy viscosity. */
static void init_state(cms_oracle *o)
{
    const int nd = o->ncellsD;
    flow_wetdry(o);
    memcpy(o->u1, o->u, sizeof(double) * nd); memcpy(o->v1, o->v, sizeof(double) * nd);
    memcpy(o->p1, o->p, sizeof(double) * nd); memcpy(o->h1, o->h, sizeof(double) * nd);
    der_grad_cbgg(o, o->gow, 1);                 /* der_init */
    /* prestart initial state (prestart.f90:309-329) */
    flow_wetdry(o);
    memcpy(o->iwet1, o->iwet, sizeof(int) * nd);
    for (int i = 0; i < nd; i++) o->eta[i] = o->iwet[i] * o->p[i] * (1.0 / GRAV) - 999.0 * (1 - o->iwet[i]);
    der_grad_cbgg(o, o->gow, 1);
    flow_grad_interp(o);
    flow_convflux(o);
    memcpy(o->flux1, o->flux, sizeof(double) * NF * nd);
    fric_eval(o);
    flow_eddyvis(o);
}

void cms_oracle_destroy(cms_oracle *o)
{
    if (!o) return;
    void *ptrs[] = {o->c2c, o->llec, o->nxy, o->kxy, o->dx, o->dy, o->x, o->y, o->area, o->ds, o->dn,
                    o->fnx, o->fny, o->rx, o->ry, o->dsx, o->dsy, o->dsxy, o->fintp, o->Wc, o->Wk, o->THc,
                    o->THk, o->Qc, o->Qk, o->wsebnd, o->wseadj, o->u, o->v, o->p, o->h, o->pp, o->dppx, o->dppy, o->dux,
                    o->duy, o->dvx, o->dvy, o->dpx, o->dpy, o->vis, o->zb, o->eta, o->u1, o->v1, o->p1, o->h1,
                    o->iwet, o->iwet1, o->icorner, o->flux, o->flux1, o->hk, o->pk, o->visk, o->zbk, o->acoef,
                    o->su, o->sv, o->sp, o->spu, o->spv, o->ssu0, o->ssv0, o->spuv0, o->sppp0, o->sspp0,
                    o->apuareap, o->sumu, o->Hu, o->Hv, o->rsu, o->rsv, o->rsp, o->coefman, o->cfrict, o->z0,
                    o->uv, o->uelwc, o->bsxy, o->bsvel, o->cmb, o->cbcfuwcap, o->wavestrx, o->wavestry, o->fc,
                    o->cbf, o->maxeta, o->gow, o->wk, o->ia, o->ja, o->diagpos, o->lu,
                    o->ncf, o->idir, o->nxf, o->kxf, o->nyf, o->kyf, o->isjoint, o->joint, o->rpx, o->rpy,
                    o->dHux, o->dHuy, o->dHvx, o->dHvy, o->dapux, o->dapuy, o->dzbx, o->dzby, o->goa};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
    for (int q = 0; q < 33; q++) free(o->st[q]);
    free(o);
}

int cms_oracle_ncells(const cms_oracle *o) { return o->ncells; }
int cms_oracle_ncellsD(const cms_oracle *o) { return o->ncellsD; }
int cms_oracle_nf(const cms_oracle *o) { return o->nf; }

int cms_oracle_topology(const cms_oracle *o, int *ncface, int *cell2cell, int *idirface, int *isjoint)
{
    if (ncface) memcpy(ncface, o->ncf, sizeof(int) * o->ncells);
    if (cell2cell) memcpy(cell2cell, o->c2c, sizeof(int) * NF * o->ncellsD);
    if (idirface) memcpy(idirface, o->idir, sizeof(int) * NF * o->ncellsD);
    if (isjoint) memcpy(isjoint, o->isjoint, sizeof(int) * o->ncells);
    return 0;
}
double cms_oracle_time(const cms_oracle *o) { return o->timesecs; }

int cms_oracle_step(cms_oracle *o, int nsteps, cms_step_info *info)
{
    for (int s = 0; s < nsteps; s++) {
        /* timing, CMS_Flow.f90:109-240; with dtvar the time step is halved
         * after a step that ended with iconv = 3 and doubled back once the
         * pp residual is small and the time falls on the coarser step
         * (:180-205) */
        o->ntime++;
        if (o->dtvar) {
            if (o->iconv == 3 && o->dtimebeg / o->deltime <= 10.0) {
                o->mtime = 0;
                o->jtime += 1;
                o->deltime = o->dtimebeg / ldexp(1.0, o->jtime);
                o->dtime = o->deltime;
            } else if (o->jtime >= 1) {
                if ((o->mtime >= 2 && o->rmom[0] < RMOMTARGETP / 100.0) ||
                    (o->mtime >= 4 && o->rmom[0] < RMOMTARGETP / 10.0)) {
                    double dtt = o->dtimebeg / ldexp(1.0, o->jtime - 1);
                    double rtime = fmod(o->timesecs + dtt, dtt);
                    double err = fabs(rtime) / dtt;
                    if (rtime < F4(0.01) || err < F4(0.01)) {
                        o->deltime = dtt;
                        o->dtime = o->deltime;
                        o->mtime = 0;
                        o->jtime -= 1;
                    }
                }
            }
            o->mtime += 1;
        }
        o->timesecs += o->deltime;
        o->ctime = o->timesecs;
        o->timehrs = o->ctime / 3600.0;
        o->ramp = ramp_func(o->timehrs, o->rampdur);
        int rc = flow_imp(o);
        if (info) {
            info->niter = o->niter; info->iconv = o->iconv;
            for (int n = 0; n < 3; n++) { info->rmom[n] = o->rmom[n]; info->kry_iters[n] = o->kry[n]; }
            int nw = 0;
            for (int i = 0; i < o->ncells; i++) nw += o->iwet[i];
            info->nwet = nw;
        }
        if (rc != 0) return rc;
        o->bndvol += bnd_inflow(o);
        if (o->stat_on) stat_update_flow(o);       /* CMS_Flow.f90:78-81, before flow_update */
        flow_update(o);
    }
    return 0;
}

/* Flow statistics (stat_update_flow, source/statistics/statistics.F90:476-720):
 * accumulators over the steps of a window, reset when the window closes. */
enum { S_ETAMAX, S_TETAMAX, S_ETAMIN, S_TETAMIN, S_UVMAX, S_UMAX, S_VMAX, S_TUVMAX, S_ETAMEAN, S_UR, S_VR,
       S_UVCURV, S_HYDPER, S_RSUAVG, S_RSVAVG, S_RSPAVG, S_RSUMAX, S_RSVMAX, S_RSPMAX, S_ETAMAXGRAD, S_UVMAXGRAD,
       S_FRXMAX, S_FRYMAX, S_UMAXCURVX, S_UMAXCURVY, S_VMAXCURVX, S_VMAXCURVY, S_CFLXMAX, S_CFLYMAX,
       S_QCFLXMAX, S_QCFLYMAX, S_QFRXMAX, S_QFRYMAX, NSTAT };
static const char *stat_names[NSTAT] = {
    "etamax", "tetamax", "etamin", "tetamin", "uvmax", "umax", "vmax", "tuvmax", "etamean", "ur", "vr", "uvcurv",
    "hydper", "rsuavg", "rsvavg", "rspavg", "rsumax", "rsvmax", "rspmax", "etamaxgrad", "uvmaxgrad", "frxmax",
    "frymax", "umaxcurvx", "umaxcurvy", "vmaxcurvx", "vmaxcurvy", "cflxmax", "cflymax", "qcflxmax", "qcflymax",
    "qfrxmax", "qfrymax"};

/* dx2d / dy2d (der_lib.f90): one cell's gradient component with the wet-cell operator */
static double grad_x(const cms_oracle *o, int i, const double *var)
{
    const gradop *G = &o->gow[i];
    double sx = 0.0;
    for (int k = 0; k < G->ncx; k++) sx += G->wcx[k] * var[G->icx[k]];
    return sx;
}
static double grad_y(const cms_oracle *o, int i, const double *var)
{
    const gradop *G = &o->gow[i];
    double sy = 0.0;
    for (int k = 0; k < G->ncy; k++) sy += G->wcy[k] * var[G->icy[k]];
    return sy;
}

static void stat_update_flow(cms_oracle *o)
{
    if (o->ctime + F4(1.0e-4) < o->stat_t1) return;
    o->stat_n += 1;
    double **S = o->st;
    const int nc = o->ncells;
    for (int i = 0; i < nc; i++) {
        if (o->iwet[i] == 0) continue;
        if (o->eta[i] > S[S_ETAMAX][i]) { S[S_ETAMAX][i] = o->eta[i]; S[S_TETAMAX][i] = o->timehrs; }
        if (o->eta[i] < S[S_ETAMIN][i]) {
            S[S_ETAMIN][i] = o->eta[i];
            for (int q = 0; q < o->ncellsD; q++) S[S_TETAMIN][q] = o->timehrs;   /* tetamin = timehrs: the array */
        }
        if (o->uv[i] > S[S_UVMAX][i]) {
            S[S_UVMAX][i] = o->uv[i]; S[S_UMAX][i] = o->u[i]; S[S_VMAX][i] = o->v[i]; S[S_TUVMAX][i] = o->timehrs;
        }
        S[S_ETAMEAN][i] += o->eta[i];
        S[S_UR][i] += o->u[i];
        S[S_VR][i] += o->v[i];
        /* streamwise_curvature (flow_lib.F90:256-276) */
        const double u = o->u[i], v = o->v[i];
        S[S_UVCURV][i] += (u * u * o->dvx[i] - v * v * o->duy[i] + u * v * (o->dvy[i] - o->dux[i]))
                          / pow(fmax(o->uv[i], F4(1.e-6)), 1.5);
        S[S_HYDPER][i] += o->iwet[i];
        S[S_RSUAVG][i] += o->rsu[i]; S[S_RSVAVG][i] += o->rsv[i]; S[S_RSPAVG][i] += o->rsp[i];
        S[S_RSPMAX][i] = fmax(S[S_RSPMAX][i], o->rsp[i]);
        S[S_RSUMAX][i] = fmax(S[S_RSUMAX][i], o->rsu[i]);
        S[S_RSVMAX][i] = fmax(S[S_RSVMAX][i], o->rsv[i]);
        double val = sqrt(o->dpx[i] * o->dpx[i] + o->dpy[i] * o->dpy[i]) * (1.0 / GRAV);
        S[S_ETAMAXGRAD][i] = fmax(S[S_ETAMAXGRAD][i], val);
        val = sqrt(o->dux[i] * o->dux[i] + o->duy[i] * o->duy[i]);
        S[S_UVMAXGRAD][i] = fmax(S[S_UVMAXGRAD][i], val);
        val = sqrt(o->dvx[i] * o->dvx[i] + o->dvy[i] * o->dvy[i]);
        S[S_UVMAXGRAD][i] = fmax(S[S_UVMAXGRAD][i], val);
        const double c = sqrt(GRAV * o->h[i]);
        S[S_FRXMAX][i] = fmax(S[S_FRXMAX][i], u / c);
        S[S_FRYMAX][i] = fmax(S[S_FRYMAX][i], v / c);
        S[S_UMAXCURVX][i] = fmax(S[S_UMAXCURVX][i], grad_x(o, i, o->dux));
        S[S_UMAXCURVY][i] = fmax(S[S_UMAXCURVY][i], grad_y(o, i, o->duy));
        S[S_VMAXCURVX][i] = fmax(S[S_VMAXCURVX][i], grad_x(o, i, o->dvx));
        S[S_VMAXCURVY][i] = fmax(S[S_VMAXCURVY][i], grad_y(o, i, o->dvy));
    }
    for (int i = 0; i < nc; i++) {
        const double c = sqrt(GRAV * o->h[i]);
        S[S_CFLXMAX][i] = fmax(S[S_CFLXMAX][i], (fabs(o->u[i]) + c) * o->dtime / o->dx[i]);
        S[S_CFLYMAX][i] = fmax(S[S_CFLYMAX][i], (fabs(o->v[i]) + c) * o->dtime / o->dy[i]);
    }
    for (int i = 0; i < nc; i++) {
        S[S_QCFLXMAX][i] = fmax(S[S_QCFLXMAX][i], o->dx[i] / o->dtime / (fabs(o->u[i]) + sqrt(GRAV * o->h[i])));
        S[S_QCFLYMAX][i] = fmax(S[S_QCFLYMAX][i], o->dy[i] / o->dtime / (fabs(o->v[i]) + sqrt(GRAV * o->h[i])));
        S[S_QFRXMAX][i] = fmax(S[S_QFRXMAX][i], o->dx[i] / o->dtime * fabs(o->u[i]) / sqrt(GRAV * o->h[i]));
        S[S_QFRYMAX][i] = fmax(S[S_QFRYMAX][i], o->dy[i] / o->dtime * fabs(o->v[i]) / sqrt(GRAV * o->h[i]));
    }
    if (fmod(o->timesecs - o->stat_t1, o->stat_t2 - o->stat_t1) > F4(1.0e-4)) return;
    /* the window closes: results written (not restated), accumulators reset */
    o->stat_n = 0;
    for (int q = 0; q < NSTAT; q++) {
        if (q == S_TETAMAX || q == S_TETAMIN || q == S_TUVMAX) continue;
        for (int i = 0; i < nc; i++) S[q][i] = q == S_ETAMIN ? F4(1.0e20) : 0.0;
    }
}

/* FLOW_STATISTICS start end [s]: accumulators zero (stat_init_flow, :172-320) */
int cms_oracle_stats(cms_oracle *o, double t1, double t2)
{
    if (!(t2 > t1)) return -1;
    o->stat_on = 1; o->stat_n = 0; o->stat_t1 = t1; o->stat_t2 = t2;
    for (int q = 0; q < NSTAT; q++) { free(o->st[q]); o->st[q] = DALLOC(o->ncellsD); }
    return 0;
}
int cms_oracle_stat_get(const cms_oracle *o, const char *name, double *out, int *count)
{
    if (!o->stat_on) return -1;
    *count = o->stat_n;
    for (int q = 0; q < NSTAT; q++)
        if (!strcmp(name, stat_names[q])) { memcpy(out, o->st[q], sizeof(double) * o->ncells); return 0; }
    return -1;
}

static const double *field(const cms_oracle *o, const char *name)
{
    if (!strcmp(name, "eta")) return o->eta;
    if (!strcmp(name, "u")) return o->u;
    if (!strcmp(name, "v")) return o->v;
    if (!strcmp(name, "p")) return o->p;
    if (!strcmp(name, "h")) return o->h;
    if (!strcmp(name, "vis")) return o->vis;
    if (!strcmp(name, "pp")) return o->pp;
    if (!strcmp(name, "dpx")) return o->dpx;
    if (!strcmp(name, "dpy")) return o->dpy;
    if (!strcmp(name, "cfrict")) return o->cfrict;
    if (!strcmp(name, "coefman")) return o->coefman;
    if (!strcmp(name, "uv")) return o->uv;
    if (!strcmp(name, "fc")) return o->fc;
    if (!strcmp(name, "wavestrx")) return o->wavestrx;
    if (!strcmp(name, "wavestry")) return o->wavestry;
    if (!strcmp(name, "bsvel")) return o->bsvel;
    if (!strcmp(name, "zb")) return o->zb;
    if (!strcmp(name, "z0")) return o->z0;
    if (!strcmp(name, "maxeta")) return o->maxeta;
    return NULL;
}

int cms_oracle_get(const cms_oracle *o, const char *name, double *out)
{
    if (!strcmp(name, "iwet")) {
        for (int i = 0; i < o->ncells; i++) out[i] = o->iwet[i];
        return 0;
    }
    const double *f = field(o, name);
    if (!f) return -1;
    memcpy(out, f, sizeof(double) * o->ncells);
    return 0;
}

/* Initial condition ("u", "v" or "p" of the active cells): the state is then
 * rebuilt as at the start of a run, so both time levels agree. */
int cms_oracle_set(cms_oracle *o, const char *name, const double *in)
{
    if (strcmp(name, "u") && strcmp(name, "v") && strcmp(name, "p")) return -1;
    double *f = (double *)field(o, name);
    memcpy(f, in, sizeof(double) * o->ncells);
    init_state(o);
    return 0;
}

int cms_oracle_get_face(const cms_oracle *o, const char *name, double *out)
{
    const double *f = NULL;
    if (!strcmp(name, "flux")) f = o->flux;
    else if (!strcmp(name, "hk")) f = o->hk;
    else if (!strcmp(name, "visk")) f = o->visk;
    else if (!strcmp(name, "pk")) f = o->pk;
    if (!f) return -1;
    memcpy(out, f, sizeof(double) * NF * o->ncells);
    return 0;
}

void cms_oracle_volume(const cms_oracle *o, double *vol, double *bndvol)
{
    double s = 0.0;
    for (int i = 0; i < o->ncells; i++) s += o->area[i] * o->h[i];
    *vol = s;
    *bndvol = o->bndvol;
}

int cms_oracle_grad(cms_oracle *o, const double *var_act, double *gx, double *gy)
{
    /* gradient of an active-cell field with the current wet-only CBGG weights;
     * ghost values copy their cell (zero-gradient) */
    double *v = DALLOC(o->ncellsD);
    memcpy(v, var_act, sizeof(double) * o->ncells);
    for (int i = 0; i < o->ncells; i++)
        for (int k = 0; k < o->ncf[i]; k++) {
            int nck = o->c2c[NF * i + k];
            if (nck >= o->ncells) v[nck] = var_act[i];
        }
    der_grad_cbgg(o, o->gow, 1);
    der_grad_eval(o, o->gow, v, gx, gy);
    free(v);
    return 0;
}

/* wave_rad_stress_fl (waves/wave_rad_stress_fl.f90:2-59): the kinematic
 * radiation-stress forcing from the wave height, period, length and unit
 * direction on the flow grid (active cells, row-major), with the current depth
 * and the wet-only gradient operator:
 *   kh = min(2 pi h / L, 10), n = 1/2 + kh / sinh(2 kh), E = g H^2 / 16,
 *   Sxx = E (n (1 + ux^2) - 1/2), Syy = E (n (1 + uy^2) - 1/2),
 *   Sxy = E n uy uy   (the reference's uy*uy, :31, kept as it is),
 *   wavestrx = -dSxx/dx - dSxy/dy, wavestry = -dSxy/dx - dSyy/dy.
 * Ghost values copy their cell (the reference reads its uninitialised stack
 * array there).  Returns -1, leaving the forcing unchanged, on a NaN or a
 * component above 10 (the reference's error stop, :42-56). */
int cms_oracle_set_waves(cms_oracle *o, const double *whgt, const double *wper, const double *wlen,
                         const double *ux, const double *uy)
{
    (void)wper;                               /* c = L/T is computed and unused (:26) */
    const int n = o->ncells;
    const double twopi = 2.0 * acos(-1.0);    /* const_def twopi, definitions.F90:86-87 */
    double *sxx = DALLOC(n), *sxy = DALLOC(n), *syy = DALLOC(n);
    double *g1x = DALLOC(n), *g1y = DALLOC(n), *g2x = DALLOC(n), *g2y = DALLOC(n), *g3x = DALLOC(n),
           *g3y = DALLOC(n);
    for (int i = 0; i < n; i++) {
        double kh = twopi / wlen[i] * o->h[i];
        kh = fmin(kh, 10.0);
        const double cn = 0.5 + kh / sinh(2.0 * kh);
        const double e = 0.0625 * GRAV * whgt[i] * whgt[i];
        sxx[i] = e * (cn * (1.0 + ux[i] * ux[i]) - 0.5);
        syy[i] = e * (cn * (1.0 + uy[i] * uy[i]) - 0.5);
        sxy[i] = e * cn * uy[i] * uy[i];
    }
    cms_oracle_grad(o, sxx, g1x, g1y);
    cms_oracle_grad(o, sxy, g2x, g2y);
    cms_oracle_grad(o, syy, g3x, g3y);
    int bad = 0;
    for (int i = 0; i < n; i++) {
        const double fx = -g1x[i] - g2y[i], fy = -g2x[i] - g3y[i];
        if (isnan(fx) || isnan(fy) || fabs(fx) > 10.0 || fabs(fy) > 10.0) bad = 1;
    }
    if (!bad)
        for (int i = 0; i < n; i++) {
            o->wavestrx[i] = -g1x[i] - g2y[i];
            o->wavestry[i] = -g2x[i] - g3y[i];
        }
    free(sxx); free(sxy); free(syy); free(g1x); free(g1y); free(g2x); free(g2y); free(g3x); free(g3y);
    return bad ? -1 : 0;
}

int cms_oracle_solve(int ncells, int ncellsD, int nf, const double *acoef, const int *c2c,
                     const int *iwet, double *ss, const double *sp, double *res,
                     double *phi, double relax, double rtol, double atol, int kmax,
                     double *rmom)
{
    if (nf < 1 || nf > CMS_MAX_FACES) return -1;
    double rinv = 1.0 / relax, relaxm = 1.0 - relax, adak = 0.0;
    double *ap = DALLOC(ncells), *wk = DALLOC(6 * (size_t)ncells);
    for (int i = 0; i < ncells; i++) {
        double asum = 0.0, anb = 0.0;
        for (int k = 0; k < nf; k++) {
            int nck = c2c[nf * i + k];
            if (nck < 0) continue;
            asum += acoef[nf * i + k];
            anb += acoef[nf * i + k] * phi[nck];
        }
        ap[i] = (asum - sp[i]) * rinv;
        ss[i] = ss[i] + relaxm * phi[i] * ap[i];
        res[i] = iwet[i] * (anb + ss[i] - ap[i] * phi[i]) / ap[i];
        adak += res[i] * res[i];
    }
    *rmom = sqrt(adak / (double)(float)ncells);
    int *ia = IALLOC(ncells + 1), *ja = IALLOC((nf + 1) * (size_t)ncells), *dp = IALLOC(ncells);
    double *lu = DALLOC((nf + 1) * (size_t)ncells);
    sysmat S = {ncells, ncellsD, nf, acoef, c2c, ap, ia, ja, dp, lu};
    int it = bicgstab(&S, ss, phi, 0, kmax, rtol, atol, wk);
    free(ap); free(wk); free(ia); free(ja); free(dp); free(lu);
    return it;
}
