/*
 * cms2d_amd.h -- C ABI of the MI355X implementation of CMS-Flow's implicit
 * hydrodynamic time step (erdc/cms2d 5.4).
 *
 * Plain C types only: pointers, sizes, doubles and ints.  Every entry point
 * returns 0 on success and a negative status on failure; cms2d_last_error()
 * gives the message.  No C++ exception crosses this boundary.
 *
 * Reference interfaces replaced (paths relative to the erdc/cms2d tree):
 *   cms2d_flow_step  <- flow_imp                       source/flow/flow_imp.F90:29-312
 *                       driven as in the time loop     source/main/CMS_Flow.f90:43-87
 *                       (timing :109-240, flow_update :243-411)
 *   cms2d_solve      <- solve(acoef,ss,sp,res,phi,n)   source/solvers/solvers.f90:205-304
 *                       with pbicgstab                 source/solvers/pbicgstab.f90:2-98
 *   cms2d_flow_create <- prestart/flow_init/geo_cart   source/main/prestart.f90:2-329,
 *                       source/flow/flow.F90:592-835, source/geospatial/geospatial.F90:2709-2828
 *   cms2d_flow_create_tel <- the same with read_tel     source/geospatial/geospatial.F90:403-795
 *   cms2d_flow_get/_get_face/_set <- reads/writes of the flow_def module state
 *                       (eta, u, v, p, h, iwet, flux, hk, visk; source/flow/flow_def.f90:32-106)
 *   cms2d_flow_time  <- timesecs of the time loop      source/main/CMS_Flow.f90:109-240
 *   cms2d_flow_set_waves <- wave_rad_stress_fl         source/waves/wave_rad_stress_fl.f90:2-59
 *   cms2d_flow_set_wind  <- windcurve_eval (constant wind) source/meteorological/meteorological.F90:995-1033
 *   cms2d_flow_checkpoint/_restore <- hot start write/read (exact-resume variant)
 *                       source/hot_start/hot_start.F90:693-1122,1125-1344,1512-1560
 *   cms2d_flow_balance <- flow_balance and flow_step_stat source/flow/flow_imp.F90:2260-2372,
 *                       source/flow/flow.F90:1912-2008
 *   cms2d_flow_statistics/_stat <- stat_init_flow, stat_update_flow
 *                       source/statistics/statistics.F90:172-320,476-720
 *   cms2d_last_error <- the diag_print_error messages  source/utilities/diag_lib.f90:177-220
 * Without a counterpart in the reference: cms2d_device_count (device probe),
 * cms2d_flow_kernel_stats (live kernel timing for measurement),
 * cms2d_flow_tuning (the launch configuration picked for a flow),
 * cms2d_ensemble_step (independent domains side by side), cms2d_solve_adjoint
 * and cms2d_flow_adjoint (the discrete adjoint of a solve and of the time
 * loop) and cms2d_solve_noalloc (for callers that cannot allocate).
 */
#ifndef CMS2D_AMD_H
#define CMS2D_AMD_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define CMS2D_OK 0
#define CMS2D_ERR_ARG (-1)        /* invalid argument or size */
#define CMS2D_ERR_HIP (-2)        /* HIP runtime error */
#define CMS2D_ERR_DIVERGED (-3)   /* outer iteration divergent (iconv=0) with fixed dt */
#define CMS2D_ERR_NODEVICE (-4)   /* no HIP device */

/* Side codes, in the reference's face order (geospatial.F90:2841-2875). */
#define CMS2D_SIDE_N 0
#define CMS2D_SIDE_E 1
#define CMS2D_SIDE_S 2
#define CMS2D_SIDE_W 3
/* Boundary types of a side. */
#define CMS2D_BND_WALL 0          /* closed wall (W_str, boundaries.F90:2413-2435)            */
#define CMS2D_BND_WSE_HARMONIC 1  /* harmonic water level (TH_str, boundaries.F90:4098-4124) */
#define CMS2D_BND_FLUX 3          /* flux (river) string, FLUX_VALUE / FLUX_CURVE (Q_str,
                                     boundaries.F90:265-445,4078-4094; bndriver_vel :4259-4356;
                                     bnd_vel_flux, bnd_vel.f90:2-32; bndflux :4359-4414): the total
                                     inflow flux_q, or the curve flux_times/flux_values, split over
                                     the side's faces by the conveyance hk^(1+c)/n along the inward
                                     normal or flux_angle; Manning friction only                */
#define CMS2D_BND_WSE_CURVE 2     /* single water level from a time series, WSE_CURVE / WSE_CONSTANT
                                     (H_str, boundaries.F90:4127-4150): the curve wse_times/
                                     wse_values, Lagrange interpolation of order wse_order
                                     (plagr_fit, plagr_lib.f90:20-138)                       */

/* Preconditioner of the pressure-correction (pp) BiCGStab.  The momentum
 * systems are diagonally dominant and always use Jacobi. */
#define CMS2D_PRECOND_MG 0        /* FP32 geometric multigrid V(1,1) cycle                    */
#define CMS2D_PRECOND_JACOBI 1    /* diagonal                                                 */

/* Krylov method of the pp system (the momentum systems use BiCGStab). */
#define CMS2D_SOLVER_BICGSTAB 0   /* pbicgstab (pbicgstab.f90:2-98)                           */
#define CMS2D_SOLVER_GMRES 1      /* restarted GMRES(20) (pgmres.f90:2-146 runs GMRES(3))     */
/* Turbulence models (flow_eddyvis.f90:23-160); 0 is the reference's default */
#define CMS2D_TURB_SUBGRID 0      /* SUBGRID, mturbul 5                                          */
#define CMS2D_TURB_CONSTANT 1     /* CONSTANT, mturbul 0                                         */
#define CMS2D_TURB_FALCONER 2     /* FALCONER, mturbul 1                                         */
#define CMS2D_TURB_PARABOLIC 3    /* PARABOLIC, mturbul 2                                        */
#define CMS2D_TURB_SUBGRID_WU 4   /* SUBGRID-WU, mturbul 3                                       */
/* Bottom friction laws (fric_eval, friction.F90:638-688); the field `mann` holds
   the law's input per cell.  0 is the reference's default. */
#define CMS2D_FRIC_MANNING 0      /* Manning n, mbedfric 2 (MANNINGS_N_DATASET)                  */
#define CMS2D_FRIC_DRAG 1         /* drag coefficient Cd, mbedfric 1 (BOTTOM_FRICTION_COEF_*)    */
#define CMS2D_FRIC_ROUGHNESS 2    /* hydraulic roughness length z0 [m] = roughness height / 30,
                                     mbedfric 3 (ROUGHNESS_HEIGHT_*, friction.F90:344,385-387) */
#define CMS2D_FRIC_LINEAR 3       /* linear coefficient c [1/s] (bed stress rho c h U), mbedfric 4
                                     (LINEAR_BOTTOM_FRICTION_COEF_CONSTANT)                     */

/* Solver precision forms. */
#define CMS2D_PREC_FP64 0         /* FP64 everywhere                                          */
#define CMS2D_PREC_FP32_REFINE 1  /* FP32 Krylov + SpMV inside FP64 iterative refinement     */

typedef struct {
    int nx, ny;                   /* active cells per row / per column (row-major from SW), >= 3 */
    double dx, dy;                /* uniform cell size [m]                                  */
    const double *zb;             /* bed elevation [nx*ny], host memory [m]                 */
    const double *mann;           /* bottom friction input [nx*ny] of the law `friction`:
                                     Manning n by default (MANNINGS_N_DATASET)              */
    const double *wavestrx;       /* radiation-stress forcing / rho [nx*ny] or NULL         */
    const double *wavestry;       /* (wave_flowgrid_def wavestrx/wavestry)                  */
    int side_type[4];             /* CMS2D_BND_* for N,E,S,W: one water-level side and one flux
                                     side at most (up to 8192 cells long), the others walls  */
    int ntc;                      /* harmonic constituents (<= 8)                            */
    double amp[8];                /* [m]                                                     */
    double speed[8];              /* [rad/hr]                                                */
    double phase[8];              /* [rad]                                                   */
    double dt;                    /* HYDRO_TIMESTEP [s], fixed (VARIABLE_TIMESTEP OFF)       */
    double rampdur;               /* DURATION_RAMP [hr]; 0 -> one time step                  */
    double hdry;                  /* DRYING_DEPTH [m] (reference default 0.05)               */
    int noptset;                  /* >= 3: wave branches of friction/eddy viscosity active   */
    int wallfric;                 /* USE_WALL_FRICTION_TERMS                                 */
    int ndsch;                    /* ADVECTION_SCHEME 0 none, 1 upwind, 2 hybrid, 3 power law,
                                     4 exponential, 5-9 HLPA, GAMMA, CUBISTA, ALVSMART, HOAB
                                     (upwind plus deferred anti-diffusion), 10-18 upwind   */
    int maxit;                    /* HYDRO_MAX_ITERATIONS; 0 -> 20 + ncells/10000            */
    double rtol;                  /* Krylov stop: ||r||2 <= max(rtol*||r0||2, atol*sqrt(n))  */
    double atol;
    int kmax;                     /* Krylov iteration cap per solve                          */
    int precision;                /* CMS2D_PREC_*                                            */
    int device;                   /* HIP device ordinal                                      */
    int precond;                  /* CMS2D_PRECOND_* for the pp system                       */
    int solver;                   /* CMS2D_SOLVER_* for the pp system (GMRES: FP64, multigrid) */
    int dtvar;                    /* VARIABLE_TIMESTEP: a divergent step is redone with half the
                                     time step (flow_recalculate, flow_imp.F90:1133-1242), dt
                                     recovers in timing (CMS_Flow.f90:180-205); 0: divergence
                                     returns CMS2D_ERR_DIVERGED                                   */
    double latitude;              /* AVERAGE_LATITUDE [deg] for the Coriolis parameter          */
    int coriolis_plane;           /* CORIOLIS_APPROXIMATION 0 f-plane, 1 beta-plane
                                     (flow.F90:261-269,720-736)                               */
    int turbulence;               /* TURBULENCE_MODEL, CMS2D_TURB_* (flow.F90:362-384)        */
    int friction;                 /* bottom friction law, CMS2D_FRIC_* (friction.F90:201-256) */
    const double *wse_times;      /* CMS2D_BND_WSE_CURVE sides: curve times [hr], increasing,  */
    const double *wse_values;     /* water levels [m] (offset included), wse_ntimes each;      */
    int wse_ntimes;               /* one point is a constant level (WSE_CONSTANT)             */
    int wse_order;                /* TIME_INTERP_ORDER 1-3 (0 -> 2, the reference's default),
                                     limited to wse_ntimes - 1 (boundaries.F90:1417,1869)     */
    int wse_adjust_off;           /* WSE_ADJUSTMENT OFF on the curve sides: no wind and wave
                                     set-up along the string (boundaries.F90:552-554,3569)    */
    double wse_gradx, wse_grady;  /* WSE_GRADIENTS of the curve sides [m/m], grid axes: the level
                                     varies along the side from its middle cell
                                     (boundaries.F90:556-572,1892-1901)                        */
    int serial_order;             /* 1: the three loops whose iterations read values earlier
                                     iterations write (dry-to-wet start values, flow.F90:1477-1491;
                                     eddy-viscosity fix-ups next to dry cells and at corners,
                                     flow_eddyvis.f90:168-212) give the reference's one-thread
                                     (cell order) result; 0: every read sees the values from
                                     before the loop (independent of any order).  Cartesian grids */
    double flux_q;                /* CMS2D_BND_FLUX side: total inflow [m^3/s], positive into the
                                     domain (FLUX_VALUE ... 'm^3/s'), when flux_ntimes = 0   */
    const double *flux_times;     /* or its curve: times [hr], increasing, and inflows [m^3/s], */
    const double *flux_values;    /* flux_ntimes >= 2 of each (FLUX_CURVE, m^3/s/boundary)     */
    int flux_ntimes;
    int flux_order;               /* TEMPORAL_INTERPOLATION_ORDER 1-3 (0 -> 1, the reference's
                                     default for fluxes), limited to flux_ntimes - 1          */
    double flux_conveyance;       /* CONVEYANCE coefficient, limited to [0.5, 1]; 0 -> 0.667   */
    int flux_angle_on;            /* 1: INFLOW_ANGLE flux_angle [deg clockwise from north, grid
                                     angle 0] instead of the inward normal                    */
    double flux_angle;
    int flux_rating;              /* 1: flux_times/flux_values are a RATING_CURVE: the inflow follows
                                     the mean depth of the string's cells [m] (stage) by linear
                                     interpolation, constant beyond the table (:4084-4093)    */
} cms2d_flow_desc;

typedef struct {
    int niter;                    /* outer (SIMPLEC) iterations of the last step             */
    int iconv;                    /* check_conv state of the last step                       */
    double rmom[3];               /* normalized residual RMS: pp, u, v                       */
    int kry_iters[3];             /* Krylov iterations over the last step: pp, u, v          */
    int nwet;                     /* wet active cells                                        */
    double kernel_ms;             /* device time of the steps run by this call [ms]          */
    double dt;                    /* time step of the last step [s] (varies with dtvar)      */
} cms2d_step_info;

/* Telescoping grid: the content of a .tel file (read_tel,
 * source/geospatial/geospatial.F90:403-795), every cell active.  Cells in file
 * order; geom[4*i..] = centre x, y and size dx, dy [m] in grid coordinates
 * with the origin at the south-west corner of the base grid; nb[8*i..] = the
 * file's neighbour slots N1 N2 E1 E2 S1 S2 W1 W2 (1-based cell numbers, 0
 * none): a side with two finer neighbours lists them clockwise (north: west
 * then east; east: north then south; south: east then west; west: south then
 * north), at most six faces per cell.  bnd_cells: the water-level cellstring
 * (0-based, in string order, 2..8192 cells; its faces follow cellstr_bnd,
 * source/boundaries/bnd_lib.f90:215-316); every other edge face is a wall. */
typedef struct {
    int ncells;
    const double *geom;
    const int *nb;
    const int *bnd_cells;
    int nbnd;
    const int *flux_cells;        /* the flux string (CMS2D_BND_FLUX side), 0-based, string order */
    int nflux;
} cms2d_tel_grid;

typedef struct cms2d_flow cms2d_flow;

const char *cms2d_last_error(void);
int cms2d_device_count(int *count);

/* Builds the grid, uploads the case and sets the cold-start state
 * (prestart.f90:309-329). */
int cms2d_flow_create(const cms2d_flow_desc *desc, cms2d_flow **out);
int cms2d_flow_destroy(cms2d_flow *flow);
/* A flow on a telescoping grid: read_tel and geo_init_cart
 * (geospatial.F90:403-795,1400-1611) on the cells of `grid`, the step of
 * flow_imp with the joint-cell skewness terms (flow_imp_flux.f90:2-169,
 * flow_imp.F90:718-789,912-962, computational.F90:149-264, der_lib.f90:381-504,
 * interp_lib.f90:2853-2873).  desc gives the physics and the base grid
 * (nx, ny, dx, dy: every cell lies in one base cell, which the multigrid
 * preconditioner coarsens); zb, mann and the forcing are per cell of `grid`;
 * one side_type names the water-level string of grid->bnd_cells, and a
 * CMS2D_BND_FLUX side the flux string of grid->flux_cells.  BiCGStab
 * (the FP64 form or CMS2D_PREC_FP32_REFINE) with a fixed time step; the NVD
 * schemes and water-level gradients are built for Cartesian grids only.  Fields of cms2d_flow_get are per cell;
 * cms2d_flow_get_face gives out[6*c+k] in the cell's face order. */
int cms2d_flow_create_tel(const cms2d_flow_desc *desc, const cms2d_tel_grid *grid, cms2d_flow **out);
/* Advances nsteps time steps: timing, flow_imp and flow_update per step. */
int cms2d_flow_step(cms2d_flow *flow, int nsteps, cms2d_step_info *info);

/* Water volume balance of the last step and since creation (flow_balance,
 * source/flow/flow_imp.F90:2260-2372, two-level time scheme, no rain or
 * evaporation) and the step's extremes over the wet cells (flow_step_stat,
 * source/flow/flow.F90:1912-2008), computed on the device after every step.
 * Volumes in m^3: storage = sum area (h - h1); efflux = dt times the outward
 * flux through the forcing strings' faces (water level and flux); flux = dt times the sum of
 * every cell's outward face fluxes; balance = storage + efflux; net =
 * storage + flux; error_pct = 100 balance / max(|storage| + |efflux|, 1e-10);
 * cum_error_pct = 100 cum_balance / max(the largest |cum_storage| so far,
 * 1e-10).  Extremes: the signed water level, u and v of largest magnitude and
 * the largest Courant number dt max((|u| + c)/dx, (|v| + c)/dy), c = sqrt(g h),
 * each with its cell (0-based: row-major on a Cartesian grid, list order on a
 * telescoping one); the first such cell in cell order, cell 0 and 0 when no
 * wet cell exceeds 0.  The running sums are part of the checkpointed state. */
typedef struct {
    double storage, efflux, flux, balance, net, error_pct;
    double cum_storage, cum_efflux, cum_flux, cum_balance, cum_net, cum_error_pct;
    double wse, u, v, courant;
    int wse_cell, u_cell, v_cell, courant_cell;
    int steps;                    /* steps counted since creation */
} cms2d_balance;
int cms2d_flow_balance(cms2d_flow *flow, cms2d_balance *out);

/* Flow statistics over time windows (the FLOW_STATISTICS card; stat_init_flow
 * and stat_update_flow, source/statistics/statistics.F90:172-320,476-720):
 * from t_start [s] every step adds to per-cell accumulators, and the window
 * closes every t_end - t_start after t_start, where the reference writes its
 * statistics and resets them (extremes and sums to 0, etamin to 1e20, the
 * times of the extremes kept).  Cartesian grids. */
int cms2d_flow_statistics(cms2d_flow *flow, double t_start, double t_end);
/* One accumulator per active cell (row-major) and the steps counted in the
 * open window: "etamax", "tetamax" [hr], "etamin", "tetamin" (the reference
 * sets every cell's to the time of the latest new minimum anywhere), "uvmax",
 * "umax", "vmax", "tuvmax", the sums "etamean", "ur", "vr", "uvcurv"
 * (streamwise curvature), "hydper" (wet steps), "rsuavg", "rsvavg", "rspavg",
 * the maxima "rsumax", "rsvmax", "rspmax" (normalized residuals),
 * "etamaxgrad", "uvmaxgrad", "frxmax", "frymax" (u/c, v/c), "umaxcurvx",
 * "umaxcurvy", "vmaxcurvx", "vmaxcurvy" (second derivatives), "cflxmax",
 * "cflymax", "qcflxmax", "qcflymax", "qfrxmax", "qfrymax".  The reference's
 * window results are the sums divided by the count. */
int cms2d_flow_stat(cms2d_flow *flow, const char *name, double *out, int *count);
/* Copies an active-cell field to host memory, row-major [nx*ny]:
 * "eta","u","v","p","h","iwet","vis","cfrict","coefman","z0","uv","bsvel","pp","dpx","dpy",
 * "zb","fc","wavestrx","wavestry", and "maxeta" (running maximum of eta over the steps,
 * CMS_Flow.f90:352-355). */
int cms2d_flow_get(cms2d_flow *flow, const char *name, double *out);
/* Directed face field of every active cell, out[4*c+k], k = N,E,S,W:
 * "flux","hk","visk". */
int cms2d_flow_get_face(cms2d_flow *flow, const char *name, double *out);
/* Initial condition: overwrites an active-cell field ("u","v","p") and
 * rebuilds the state from it as a run starts (wet/dry, previous time level,
 * face fluxes by flow_convflux, friction, eddy viscosity: prestart.f90:309-329,
 * flow.F90:1879-1910), as a hot start does. */
int cms2d_flow_set(cms2d_flow *flow, const char *name, const double *in);
double cms2d_flow_time(const cms2d_flow *flow);
/* Average device time per launch of the named kernel group over the steps run
 * so far, counting the launches that did work, and their number.  Groups:
 * "spmv_pp" (pp-system SpMV v = A z), "mg_down0" and "mg_up0" (multigrid
 * level-0 down and up sweeps).  Measurement only: no reference counterpart. */
int cms2d_flow_kernel_stats(const cms2d_flow *flow, const char *name, double *avg_ms, long *count);
/* Exact checkpoint of a flow: every array of the model state (both time
 * levels, fluxes, wet/dry flags, boundary set-up, forcing) and the time and
 * time-step control, so that cms2d_flow_restore into a flow created from the
 * same description continues bitwise as the original would.  The reference's
 * hot start (hot_write_xmdf, source/hot_start/hot_start.F90:1512-1560; read
 * :693-1122 and hot_init :1125-1344) stores p, eta, u, v (the fluxes too in
 * its DEV_MODE build) and rebuilds the rest, which cms2d_flow_set does here;
 * this is the exact-resume variant.  buf = NULL: *nbytes receives the size. */
int cms2d_flow_checkpoint(cms2d_flow *flow, void *buf, size_t *nbytes);
/* CMS2D_ERR_ARG if buf is not a checkpoint of a flow on the same grid. */
int cms2d_flow_restore(cms2d_flow *flow, const void *buf, size_t nbytes);

/* Spatially constant wind (the reference's windconst: WIND_INPUT_CURVE /
 * WIND_SPEED_CURVE + WIND_DIRECTION_CURVE, meteorological.F90:176-202):
 * the 10-m wind vector in grid coordinates [m/s] as windcurve_init and
 * windcurve_eval make it from the curve (meteorological.F90:700-710,
 * 1015-1019; the host interpolates a curve in time and calls this before each
 * step).  Each step then adds ramp Cd(W) 0.0012 W (wndx, wndy) with Hsu's drag
 * (met_lib.f90:33-66) to the momentum sources (flow_imp.F90:386-394). */
int cms2d_flow_set_wind(cms2d_flow *flow, double wndx, double wndy);

/* Wave forcing from wave fields on the flow grid: wave_rad_stress_fl
 * (source/waves/wave_rad_stress_fl.f90:2-59), called where the reference
 * calls it after interpolating the CMS-Wave fields (wave.F90:125-190).
 * Inputs per active cell (row-major): significant wave height whgt (m),
 * period wper (s, unused as in the reference), wave length wlen (m), unit
 * wave direction (wunitx, wunity).  With the current depth the radiation
 * stresses and their wet-only gradient give the kinematic forcing used by the
 * following steps (readable as "wavestrx"/"wavestry").  Needs a wave mode
 * (noptset >= 3).  CMS2D_ERR_ARG, forcing unchanged, if a component is NaN
 * or above 10 (the reference's error stop). */
int cms2d_flow_set_waves(cms2d_flow *flow, const double *whgt, const double *wper, const double *wlen,
                         const double *wunitx, const double *wunity);
/* Launch configuration picked when the flow was created: the tile width of
 * the level-0 multigrid kernels (64, 128 or 256, the fastest when timed on
 * this grid; 0 without a level 0 cycle) and whether BiCGStab iterations are
 * replayed from captured graphs (grids up to 1M cells).  Neither changes any
 * result.  No reference counterpart. */
int cms2d_flow_tuning(cms2d_flow *flow, int *mg_tile_x, int *graphs);

/* Batched ensemble: advances each of nflows independent flow objects by
 * nsteps time steps.  No counterpart in the reference (which runs one domain
 * per process): it runs that many CMS_Flow time loops (CMS_Flow.f90:43-87)
 * side by side, members taken largest first by nthreads host threads (0 -> 8),
 * each member on its own stream, so small domains share the device.  infos
 * (optional, [nflows]) receives each member's info as cms2d_flow_step gives
 * it.  Every member completes its steps; the status is the first failing
 * member's (in member order) or CMS2D_OK.  A flow object must not appear
 * twice or be used by another thread during the call. */
int cms2d_ensemble_step(cms2d_flow **flows, int nflows, int nsteps, int nthreads, cms2d_step_info *infos);

/* Drop-in for solve(acoef,ss,sp,res,phi,n) on the reference's general cell
 * list (solvers.f90:205-304): nmaxfaces faces per cell (1..8; 4 on Cartesian
 * grids, geospatial.F90:2572, 6 on telescoping grids, :433), 0-based cell2cell
 * with ghost cells numbered ncells..ncellsD-1 (-1 for no face), host arrays in
 * the reference layout (acoef[nmaxfaces*ncellsD], the faces of a cell
 * contiguous).  Any caller of solve (flow n=1..3, sediment n=4, salinity and
 * heat n=5,6: sed_imp.f90:230, salinity.F90:510,599, heat.f90:480,559) passes
 * its own system.  relax = relax(n).  Under-relaxation
 * and the normalized residual follow the reference; the Krylov solve stops at
 * ||r||2 <= max(rtol*||r0||2, atol*sqrt(ncells)).  Outputs: ss, res, phi,
 * *rmom, *iters. */
int cms2d_solve(int ncells, int ncellsD, int nmaxfaces, const double *acoef, const int *cell2cell,
                const int *iwet, double *ss, const double *sp, double *res, double *phi,
                double relax, double rtol, double atol, int kmax, double *rmom, int *iters);

/* The reverse of cms2d_solve, for gradients through the linear solves (the
 * transposed Krylov solve of a discrete adjoint).  No counterpart in the
 * reference, which has no adjoint.  For phi = solve(acoef, ss, sp, phi_old,
 * relax) on the active cells (ghost values fixed; the system of
 * solvers.f90:205-304 solved to convergence) and gphi = dL/dphi [ncells], it
 * solves A^T lambda = gphi (BiCGStab, Jacobi, rtol/atol/kmax as cms2d_solve)
 * and returns dL/dss, dL/dsp [ncellsD], dL/dacoef [nmaxfaces*ncellsD] and
 * dL/dphi_old [ncellsD] (active cells through the relaxation term, ghost cells
 * through their faces); ghost rows of the others are 0.  phi: the solution
 * cms2d_solve returned (ghosts as given).  Faces must be shared: a cell lists
 * each active neighbour that lists it. */
int cms2d_solve_adjoint(int ncells, int ncellsD, int nmaxfaces, const double *acoef, const int *cell2cell,
                        const double *sp, const double *phi_old, const double *phi, double relax,
                        const double *gphi, double rtol, double atol, int kmax, double *gss, double *gsp,
                        double *gacoef, double *gphi_old, int *iters);

/* Discrete adjoint of the implicit time loop.  No counterpart in the
 * reference, which has no adjoint.  From the flow's initial state (a flow
 * object on which no step has been taken), runs nsteps steps as
 * cms2d_flow_step does and returns the gradient of a loss of the water levels
 * at the gauge cells with respect to the Manning n (gmann) and the bed
 * elevation (gzb) of every active cell [ncells, row-major; gmann with respect
 * to the input n of cms2d_flow_desc, so the first cell's entry is 0: the step
 * uses its north neighbour's n there, as the reference's fric_init does], by the reverse of
 * every kernel of the step (transposed u, v and pp solves included) at the
 * recorded forward state.  The state is checkpointed every `every` steps and
 * each interval recomputed in the reverse sweep.
 *   gauge_cells [ngauge]: row-major active-cell indices;
 *   gauge_obs [nsteps*ngauge] (step-major): when given, the loss is
 *     sum (eta - obs)^2 over steps and gauges, returned in *loss;
 *   geta [nsteps*ngauge]: otherwise dL/deta of each gauge after each step;
 *   eta_gauge [nsteps*ngauge]: the gauge water levels of the run (or NULL).
 * Discrete choices (wet/dry flags, outer-iteration counts, clipping branches)
 * are those of the forward run.  Requires a Cartesian grid, the FP64 form with
 * BiCGStab and the multigrid preconditioner, fixed dt, Manning friction,
 * advection schemes 0-4, no flux side and no wave or wind forcing (CMS2D_ERR_ARG
 * otherwise).  Allocates its checkpoints and tape on the device (about
 * 0.6 KB per cell for each stored state); the flow is back at its initial
 * state on return. */
int cms2d_flow_adjoint(cms2d_flow *flow, int nsteps, int every, int ngauge, const int *gauge_cells,
                       const double *gauge_obs, const double *geta, double *eta_gauge, double *loss,
                       double *gmann, double *gzb);

/* Device scratch of cms2d_solve_noalloc per device [bytes]. */
#define CMS2D_SOLVE_SCRATCH_BYTES (64u << 20)

/* cms2d_solve without device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, checks and results as
 * cms2d_solve, which it shares its kernels with; its device buffers (staged
 * system, Krylov vectors, partial sums, scalars) come from one static device
 * array of CMS2D_SOLVE_SCRATCH_BYTES per device, held until the process ends.
 * A system needs about (12 nmaxfaces + 36) ncellsD + 56 ncells bytes of it
 * (nmaxfaces 4: about 470k cells; 6: about 400k); larger systems are refused
 * with CMS2D_ERR_ARG.  Runs on the device's null stream and returns after the
 * solve.  Every call on a device shares that device's array: calls that overlap
 * on one device race. */
int cms2d_solve_noalloc(int ncells, int ncellsD, int nmaxfaces, const double *acoef, const int *cell2cell,
                        const int *iwet, double *ss, const double *sp, double *res, double *phi,
                        double relax, double rtol, double atol, int kmax, double *rmom, int *iters);

#ifdef __cplusplus
}
#endif
#endif
