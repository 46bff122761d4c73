// flow_state.h -- device state of the implicit step and the solver parameters.
#pragma once
#include "common.h"
#include "krylov.h"

namespace cms {

// Device pointers of every field (padded cell arrays unless noted).
struct St {
    double *u, *v, *p, *h, *pp, *zb, *vis, *visb, *visc, *u1, *v1, *p1, *h1, *eta;
    double *dppx, *dppy, *dux, *duy, *dvx, *dvy, *dpx, *dpy;
    double *coefman, *cfrict, *z0, *uv, *uelwc, *bsvel;
    double *su, *sv, *sp, *spu, *spv, *ssu0, *ssv0, *spuv0, *sppp0, *sspp0;
    double *apuareap, *sumu, *Hu, *Hv, *resu, *resv, *resp;
    double *fluxE, *fluxN, *hkE, *hkN, *viskE, *viskN, *zbkE, *zbkN;   // unique faces
    double *aN, *aE, *aS, *aW;                                         // directed coefficients
    double *wavestrx, *wavestry;
    double *tu, *tv, *tvis;
    double *wseadj;            // [4*lmax] adjusted boundary water level per side
    uint8_t *iwet, *iwet1, *icorner, *ipond;
    int *flags;                // [0] narrow-channel change, [1] ponding change, [2] newly wet cells
    double *partials;          // reduction scratch [nblocks*8]
    unsigned *counter;         // last-block counter (returns to 0 after every reduction)
};

struct Par {
    double hmin, hdry, hdry1, hdry2, dtime, c2a, rtol, atol;
    int kmax, noptset, wallfric, ndsch, lmax;
};

// Stopping threshold and scalar state at the start of a Krylov solve.
__device__ __forceinline__ void kry_start(KryScal &sc, double sumres2, double sumr2, int n, const Par &P) {
    sc.rmom = sqrt(sumres2 / (double)(float)n);       // sqrt(adak/real(ncells))
    double r0n = sqrt(sumr2);
    sc.r0norm = r0n;
    sc.rnorm = r0n;
    sc.thr = fmax(P.rtol * r0n, P.atol * sqrt((double)n));
    sc.rho = sumr2;
    sc.alpha = 1.0;
    sc.omega = 1.0;
    sc.beta = 0.0;
    sc.iters = 0;
    sc.restart = 0;
    sc.first = 1;
    sc.done = (r0n <= sc.thr) ? 1 : 0;
}

}  // namespace cms
