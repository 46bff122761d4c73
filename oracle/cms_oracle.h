/*
 * cms_oracle.h -- FP64 CPU restatement of CMS-Flow's implicit hydrodynamic
 * time step (reference: erdc/cms2d 5.4, source/flow/flow_imp.F90:29-312).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in cms2d_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls it.
 *
 * Data model: the reference's own general cell list (active cells numbered
 * row-major, ghost cells appended, directed faces k = N,E,S,W per cell, no-repeat
 * face lists), restated in C with 0-based indices.  Face arrays are stored
 * cell-major, a[4*i+k], like the reference's column-major a(k,i).
 */
#ifndef CMS_ORACLE_H
#define CMS_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct cms_oracle cms_oracle;

/* Case description for a rectangular, fully active Cartesian grid. */
typedef struct {
    int nx, ny;               /* active cells in x (columns) and y (rows)        */
    double dx, dy;            /* uniform cell sizes [m]                          */
    const double *zb;         /* bed elevation [nx*ny], row-major (j*nx+i) [m]   */
    const double *mann;       /* Manning n per cell [nx*ny]                      */
    const double *wavestrx;   /* kinematic radiation-stress force [nx*ny] or NULL */
    const double *wavestry;
    int th_side;              /* side carrying the harmonic WSE string: 0=N,1=E,2=S,3=W, -1 none */
    int ntc;                  /* number of harmonic constituents (<=8)           */
    double amp[8];            /* [m]                                             */
    double speed[8];          /* [rad/hr]                                        */
    double phase[8];          /* [rad]                                           */
    double dt;                /* time step [s]                                   */
    double rampdur;           /* ramp duration [hr] (0 -> one time step)         */
    double hdry;              /* drying depth [m] (reference default 0.05)       */
    int noptset;              /* >=3 selects the wave branches (radiation stress) */
    int wallfric;             /* wall friction on/off                            */
    int ndsch;                /* advection scheme: 1 upwind, 2 hybrid, 3 power-law, 4 exponential */
    int maxit;                /* outer iterations cap (0 -> 20+ncells/10000)     */
    int solver_mode;          /* 0 = tight (rtol), 1 = reference stopping rule   */
    double rtol;              /* tight mode: ||r||2 <= rtol*||r0||2 ...          */
    double atol;              /* ... or ||r||2 <= atol*sqrt(ncells)              */
    int kmax;                 /* tight mode iteration cap per solve              */
    int serial;               /* 1: reference serial loop order for the three loops whose
                                 OpenMP form races; 0: race-free (all reads from the values
                                 before the loop), the semantics of the HIP path      */
    int dtvar;                /* VARIABLE_TIMESTEP: a divergent step is redone with half
                                 the time step, which recovers later (0: stop with an error) */
    double latitude;          /* AVERAGE_LATITUDE [deg] (0: no Coriolis on the f-plane)   */
    int coriolis_plane;       /* CORIOLIS_APPROXIMATION: 0 f-plane, 1 beta-plane         */
    int mturbul;              /* TURBULENCE_MODEL: 0 constant, 1 Falconer, 2 parabolic,
                                 3 subgrid (Wu), 5 subgrid (the reference's default)   */
    int mbedfric;             /* bottom friction law of `mann`: 1 drag coefficient, 2 Manning n
                                 (0 is read as 2), 3 roughness length z0, 4 linear coefficient */
    const double *wse_times;  /* wse_ntimes > 0: the th_side string is a single water level  */
    const double *wse_values; /* curve (H_str) [hr], [m] instead of the harmonics          */
    int wse_ntimes;
    int wse_order;            /* TIME_INTERP_ORDER (0 -> 2)                               */
    int wse_adjust_off;       /* WSE_ADJUSTMENT OFF on the curve string                   */
    double wse_gradx, wse_grady;   /* WSE_GRADIENTS of the curve string [m/m]            */
    /* Telescoping grid (read_tel, geospatial.F90:403-795).  tel_n > 0 replaces the
     * rectangle (nx, ny, dx, dy unused): tel_n cells, all active, in the order of the
     * .tel file; tel_geom[4*i..] = centre x, y and size dx, dy [m]; tel_nb[8*i..] = the
     * file's neighbour slots N1 N2 E1 E2 S1 S2 W1 W2 (1-based cell numbers, 0 none).
     * zb, mann, wavestrx/y are per cell in that order.  The water-level string is the
     * cellstring tel_bnd[tel_nbnd] (0-based cells in string order, th_side >= 0); every
     * other edge face is a wall. */
    int tel_n;
    const double *tel_geom;
    const int *tel_nb;
    const int *tel_bnd;
    int tel_nbnd;
    int noskew;               /* SKEWNESS_CORRECTION OFF (the default is on, flow.F90:101) */
    /* Flux (river) string Q_str on a Cartesian side (bnd_block.f90:201-208,429-444;
     * boundaries.F90:1742-1790,4078-4094; bndriver_vel :4259-4356): q_side >= 0 names
     * the side whose cells take the total inflow [m^3/s per boundary, positive into
     * the domain]: q_ntimes = 0 the constant q_const, else the curve q_times [hr] /
     * q_values, Lagrange order q_order (0 -> 1), both times the ramp; distributed over
     * the string in proportion to hk^(1+q_cmvel)/n (q_cmvel <= 0 -> 0.667, the
     * default-real literal) along the inward normal. */
    int q_side;
    double q_const;
    const double *q_times;
    const double *q_values;
    int q_ntimes;
    int q_order;
    double q_cmvel;
    int q_angle_on;           /* INFLOW_ANGLE q_angle [deg clockwise from north] instead of the normal */
    double q_angle;
    int q_rating;             /* q_times/q_values are a stage-flow curve (mean depth of the string) */
    const int *tel_qcells;    /* telescoping grid: the flux string's cells (0-based, string order) */
    int tel_nq;
} cms_case;

typedef struct {
    int niter;                /* outer iterations taken in the last step         */
    int iconv;                /* 0 divergent, 1 continue, 2 converged, 3 reduce dt */
    double rmom[3];           /* pp, u, v normalized residual RMS (last outer iter) */
    int kry_iters[3];         /* Krylov iterations summed over the last step (pp,u,v) */
    int nwet;
} cms_step_info;

cms_oracle *cms_oracle_create(const cms_case *c);
void cms_oracle_destroy(cms_oracle *o);
int cms_oracle_ncells(const cms_oracle *o);
/* nmaxfaces: the face stride of get_face (4 Cartesian, 6 telescoping) */
int cms_oracle_nf(const cms_oracle *o);
/* Cell-list topology: ncface[ncells], cell2cell[nf*ncellsD] (0-based, -1 none),
 * idirface[nf*ncellsD] (0..3 = N,E,S,W, -1 none), isjoint[ncells]. */
int cms_oracle_topology(const cms_oracle *o, int *ncface, int *cell2cell, int *idirface, int *isjoint);
int cms_oracle_ncellsD(const cms_oracle *o);
/* Runs nsteps time steps (timing + flow_imp + flow_update).  Returns 0 on
 * success, -1 if the step diverged (iconv==0 with fixed dt). */
int cms_oracle_step(cms_oracle *o, int nsteps, cms_step_info *info);
/* Copies an active-cell field (row-major nx*ny) out.  name: "eta","u","v","p",
 * "h","iwet","vis","pp","dpx","dpy","cfrict","uv","bsvel","zb". */
int cms_oracle_get(const cms_oracle *o, const char *name, double *out);
/* Directed face field of every active cell: out[nf*c+k], k in the cell's face
 * order (N,E,S,W on Cartesian grids). */
int cms_oracle_get_face(const cms_oracle *o, const char *name, double *out);
int cms_oracle_set(cms_oracle *o, const char *name, const double *in);
double cms_oracle_time(const cms_oracle *o);
/* Total water volume sum(area*h) over wet active cells and the net boundary
 * inflow volume accumulated over the steps run so far. */
void cms_oracle_volume(const cms_oracle *o, double *vol, double *bndvol);

/* --- Unit entry points used by the tests (closed forms) --------------------- */
double cms_oracle_schmcoef(int ndsch, double dk, double fk);
double cms_oracle_wall_coef(double yp, double y0);
double cms_oracle_man2drag(double g, double cman, double h, double hmax);
double cms_oracle_drag2length(double cd, double h);
double cms_oracle_drag2man(double g, double cd, double h, double hmax);
/* plagr_fit + the weighted sum of bnd_eval: curve (x, y) of m points at xi,
 * order ni, window start *k (1-based, updated) */
double cms_oracle_plagr(int m, const double *x, const double *y, int ni, double xi, int *k);
double cms_oracle_length2drag(double z0, double h);

/* Gradient of an active-cell field (row-major) with the current wet-only CBGG
 * operator (der_grad_cbgg + der_grad_eval, der_lib.f90:272-378,726-839). */
int cms_oracle_grad(cms_oracle *o, const double *var_act, double *gx, double *gy);

/* Spatially constant wind (windconst): the 10-m wind vector in grid
 * coordinates [m/s]; the stress of each step is ramp * Cd(W) * 0.0012 * W *
 * (wndx, wndy) with Hsu's drag (met_lib.f90:33-66). */
int cms_oracle_set_wind(cms_oracle *o, double wndx, double wndy);
double cms_oracle_wind_drag_hsu(double w);

/* NVD limiter gamma(phi~) of ndsch 6-9 (comp_lib.f90:112-208) and the
 * deferred-correction source sums of the current state for ndsch 5-9
 * (computational.F90:337-414,495-581; zero for other schemes). */
double cms_oracle_nvd_gamma(int ndsch, double phin);
int cms_oracle_defcor(cms_oracle *o, double *dsu, double *dsv);

/* Radiation-stress forcing from wave fields, wave_rad_stress_fl
 * (waves/wave_rad_stress_fl.f90:2-59); -1 on NaN or |forcing| > 10. */
int cms_oracle_set_waves(cms_oracle *o, const double *whgt, const double *wper, const double *wlen,
                         const double *ux, const double *uy);

/* Stand-alone linear solve on a general cell list, the reference's
 * solve(acoef,ss,sp,res,phi,n) (solvers.f90:205-304) with the tight stopping rule.
 * nf = nmaxfaces (4 on Cartesian grids, 6 on telescoping grids, read_tel,
 * geospatial.F90:433), acoef[nf*ncellsD], c2c[nf*ncellsD] (0-based, -1 none),
 * iwet[ncellsD].  Returns iterations used (-1: nf out of range), rmom via *rmom. */
#define CMS_MAX_FACES 8
int cms_oracle_solve(int ncells, int ncellsD, int nf, const double *acoef, const int *c2c,
                     const int *iwet, double *ss, const double *sp, double *res,
                     double *phi, double relax, double rtol, double atol, int kmax,
                     double *rmom);

/* Flow statistics over windows [t1, t2] [s] (FLOW_STATISTICS, stat_update_flow,
 * source/statistics/statistics.F90:476-720): the accumulators of one field
 * ("etamax", "etamean", ... in the order of stat_def) and the steps counted. */
int cms_oracle_stats(cms_oracle *o, double t1, double t2);
int cms_oracle_stat_get(const cms_oracle *o, const char *name, double *out, int *count);

#ifdef __cplusplus
}
#endif
#endif
