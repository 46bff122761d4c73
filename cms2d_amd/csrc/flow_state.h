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
    double *cbf;               // linear friction coefficient (cbotfric of mbedfric 4)
    double *maxeta;            // running maximum of eta (flow_def maxeta)
    double *su, *sv, *sp, *spu, *spv, *ssu0, *ssv0, *spuv0, *sppp0, *sspp0;
    double *apuareap, *sumu, *Hu, *Hv, *resu, *resv, *resp;
    double *fluxE, *fluxN, *hkE, *hkN, *viskE, *viskN, *zbkE, *zbkN;   // unique faces
    double *aN, *aE, *aS, *aW;                                         // directed coefficients
    double *wavestrx, *wavestry;
    double *fc;                // Coriolis parameter (flow.F90:720-736)
    double *tu, *tv, *tvis;
    double *wseadj;            // [4*lmax] adjusted boundary water level per side
    uint8_t *iwet, *iwet1, *icorner, *ipond;
    double *fluxE1, *fluxN1;   // face fluxes of the previous time level (flux1)
    int *plab;                 // ponding component labels (padded index of the root, -1 dry)
    int *flags;                // [0] narrow-channel change, [1] ponding change, [2] newly wet cells
};

struct Par {
    double hmin, hdry, hdry1, hdry2, dtime, c2a, rtol, atol;
    double tauwx, tauwy;       // constant wind stress / rho of the step (windcurve_eval)
    int kmax, noptset, wallfric, ndsch, lmax;
    int mturbul;               // turbulence model (flow_eddyvis.f90:23-160)
    int mbedfric;              // bottom friction law (fric_eval, friction.F90:638-688)
};

}  // namespace cms
