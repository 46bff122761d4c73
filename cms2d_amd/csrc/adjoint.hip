// adjoint.hip -- cms2d_flow_adjoint: gradients of a loss of gauge water levels
// with respect to the Manning n and the bed elevation of every cell, by the
// discrete adjoint of the implicit time loop (included by cms2d.hip).
//
// Forward sweep: the steps run as cms2d_flow_step runs them; the model state
// is copied to device memory every `every` steps (checkpoints) and the water
// level of the gauge cells is recorded after each step.  Reverse sweep, one
// checkpoint interval at a time from the last: the interval's steps are
// recomputed from its checkpoint, keeping the state at the start of each
// step; then, for each step from the last, the step is recomputed once more
// with the tape on (the state at the start of every outer iteration and the
// intermediate values of each iteration, flow_imp_once) and reversed:
// flow_update, velface_eta, the outer iterations in reverse order (each:
// eddy viscosity, gradients and friction, pvfc, the transposed pp solve, pp
// sources, coefficients and fluxes, the transposed u and v solves, momentum
// coefficients), flow_srcsnk_static, dry-to-wet start values and the wet/dry
// depth.  The initial state's dependence on the bed (depth, face elevations,
// fluxes, friction) closes the sweep.  Each solve's adjoint is the transposed
// system solved by the same BiCGStab (multigrid preconditioner built from
// the transposed pp coefficients; Jacobi for u and v), to the forward's
// relative tolerance.
namespace cms {

struct Snap {
    char *dev = nullptr;
    CkHost host{};
};
struct Tape {
    std::vector<Snap> snap;        // state at the start of outer iteration k (k = niter: at the loop's exit)
    std::vector<double *> mid;     // MID_NARRAYS padded arrays per outer iteration
    int niter = 0;
};

}  // namespace cms

static size_t state_bytes(const cms2d_flow *F) {
    size_t n = 0;
    for (auto &a : F->state) n += a.second;
    return n;
}
static int snap_save(cms2d_flow *F, Snap &S) {
    if (!S.dev) HIPCHK(hipMalloc((void **)&S.dev, state_bytes(F)));
    size_t off = 0;
    for (auto &a : F->state) {
        HIPCHK(hipMemcpyAsync(S.dev + off, a.first, a.second, hipMemcpyDeviceToDevice, F->st));
        off += a.second;
    }
    ck_host_get(F, S.host);
    return 0;
}
static int snap_load(cms2d_flow *F, const Snap &S) {
    size_t off = 0;
    for (auto &a : F->state) {
        HIPCHK(hipMemcpyAsync(a.first, S.dev + off, a.second, hipMemcpyDeviceToDevice, F->st));
        off += a.second;
    }
    ck_host_set(F, S.host);
    return 0;
}
// The state arrays of a snapshot seen as an St (fields outside the state
// list keep the live pointers).
static St snap_view(const cms2d_flow *F, const Snap &S) {
    constexpr int NP = sizeof(St) / sizeof(void *);
    static_assert(sizeof(St) == NP * sizeof(void *), "St holds pointers only");
    void *slot[NP];
    memcpy(slot, &F->s, sizeof(St));
    for (int k = 0; k < NP; k++) {
        size_t off = 0;
        for (auto &a : F->state) {
            if (a.first == slot[k]) { slot[k] = S.dev + off; break; }
            off += a.second;
        }
    }
    St v;
    memcpy(&v, slot, sizeof(St));
    return v;
}
static Mid mid_view(const cms2d_flow *F, double *base) {
    static_assert(sizeof(Mid) == MID_NARRAYS * sizeof(double *), "Mid holds pointers only");
    double *slot[MID_NARRAYS];
    for (int k = 0; k < MID_NARRAYS; k++) slot[k] = base + (size_t)k * F->npad;
    Mid m;
    memcpy(&m, slot, sizeof(Mid));
    return m;
}

static int tape_snap(cms2d_flow *F, int slot) {
    Tape &T = *F->tape;
    if ((int)T.snap.size() <= slot) T.snap.resize(slot + 1);
    if (int rc = snap_save(F, T.snap[slot])) return rc;
    T.niter = slot;
    return 0;
}
static int tape_mid(cms2d_flow *F, int what) {
    Tape &T = *F->tape;
    const int it = F->niter - 1;
    while ((int)T.mid.size() <= it) {
        double *p = nullptr;
        HIPCHK(hipMalloc((void **)&p, (size_t)MID_NARRAYS * F->npad * sizeof(double)));
        T.mid.push_back(p);
    }
    const Mid M = mid_view(F, T.mid[it]);
    const St &s = F->s;
    const size_t nb = (size_t)F->npad * sizeof(double);
    auto cp = [&](double *dst, const double *src) {
        return hipMemcpyAsync(dst, src, nb, hipMemcpyDeviceToDevice, F->st);
    };
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    const double *a4[4] = {s.aN, s.aE, s.aS, s.aW};
    switch (what) {
    case TAPE_MOM:
        for (int k = 0; k < 4; k++) chk(cp(M.au[k], a4[k]));
        chk(cp(M.sp, s.sp)); chk(cp(M.spu, s.spu)); chk(cp(M.spv, s.spv));
        chk(cp(M.u0, s.u)); chk(cp(M.v0, s.v));
        break;
    case TAPE_UV:
        chk(cp(M.us, s.u)); chk(cp(M.vs, s.v));
        break;
    case TAPE_PPCOEF:
        for (int k = 0; k < 4; k++) chk(cp(M.a0[k], a4[k]));
        chk(cp(M.apa, s.apuareap)); chk(cp(M.sumu, s.sumu)); chk(cp(M.Hu, s.Hu)); chk(cp(M.Hv, s.Hv));
        break;
    case TAPE_PPSRC:
        for (int k = 0; k < 4; k++) chk(cp(M.ap[k], a4[k]));
        chk(cp(M.spp, s.sp));
        break;
    case TAPE_PPSOL:
        chk(cp(M.pps, s.pp));
        break;
    default:
        chk(cp(M.ppv, s.pp));
    }
    if (e != hipSuccess) return set_err(CMS2D_ERR_HIP, std::string("tape copy: ") + hipGetErrorString(e));
    return 0;
}

namespace {
// Everything the adjoint call allocates, freed on every exit.
struct AdjMem {
    std::vector<Snap> cp, ss;
    Tape tape;
    double *ad = nullptr, *eta = nullptr, *geta = nullptr;
    int *cells = nullptr;
    ~AdjMem() {
        for (auto &S : cp) if (S.dev) (void)hipFree(S.dev);
        for (auto &S : ss) if (S.dev) (void)hipFree(S.dev);
        for (auto &S : tape.snap) if (S.dev) (void)hipFree(S.dev);
        for (double *p : tape.mid) (void)hipFree(p);
        for (void *p : {(void *)ad, (void *)eta, (void *)geta, (void *)cells}) if (p) (void)hipFree(p);
    }
};
}  // namespace

static Ad ad_view(const cms2d_flow *F, double *base) {
    static_assert(sizeof(Ad) == AD_NARRAYS * sizeof(double *), "Ad holds pointers only");
    double *slot[AD_NARRAYS];
    for (int k = 0; k < AD_NARRAYS; k++) slot[k] = base + (size_t)k * F->npad;
    Ad a;
    memcpy(&a, slot, sizeof(Ad));
    return a;
}

// A transposed solve must end at its tolerance (a BiCGStab run can also stop
// on a breakdown or the iteration cap).
static int adj_solved(cms2d_flow *F, int sc0, int ns, const char *what) {
    for (int q = 0; q < ns; q++) {
        const KryScal &k = F->h_sc[sc0 + q];
        if (!(k.rnorm <= k.thr)) {
            char msg[160];
            snprintf(msg, sizeof msg, "transposed %s solve stopped at ||r|| = %.3e above its threshold %.3e after %d "
                     "iterations", what, k.rnorm, k.thr, k.final_iters);
            return set_err(CMS2D_ERR_DIVERGED, msg);
        }
    }
    return 0;
}

static KryCtx adj_ctx(cms2d_flow *F) {
    KryCtx X = flow_ctx(F, false);
    X.graph = nullptr;
    X.atol = 0.0;    // relative stop only: the adjoint right-hand sides can be small
    return X;
}

// Reverse of one outer iteration: T the state at its start, E at its end, M
// its recorded intermediate values.
static int rev_iter(cms2d_flow *F, Ad &A, const St &T, const St &E, const Mid &M) {
    const Geo &g = F->g;
    const Par &P = F->P;
    const int nbp = F->nblk_pad, nba = F->nblk_act;
    LAUNCH(k_fill0, nbp, A.pp, (long)g.nxp * g.nyp);
    LAUNCH(k_r_tail0, nbp, g, T, A);
    LAUNCH(k_r_tail1, nba, g, T, A);
    LAUNCH(k_r_tail2, nba, g, T, A);
    LAUNCH(k_r_tail3, nbp, g, T, A);
    LAUNCH(k_r_gfv, nba, g, E, P, A);
    LAUNCH(k_r_ghost_p, (2 * g.nx + 2 * g.ny + BLOCK - 1) / BLOCK, g, A);
    LAUNCH(k_r_pvfc2, nbp, g, T, (const double *)M.ppv, P, A);
    LAUNCH(k_r_pvfc1, nba, g, T, M, A);
    // pp: A^T lambda = dL/dpp
    KryCtx X = adj_ctx(F);
    {
        LAUNCH(k_r_transpose, nba, g, (const double *)M.ap[0], (const double *)M.ap[1], (const double *)M.ap[2],
               (const double *)M.ap[3], A);
        KrySys<1> K = F->Kpp;
        K.x[0] = A.lp;
        AdjStart<1> S{M.ap[0], M.ap[1], M.ap[2], M.ap[3], {M.spp}, {A.pp}, 1.0};
        hipLaunchKernelGGL(k_r_start<1>, dim3(F->nblk_kry), dim3(KBLOCK), 0, F->st, g, K, S);
        mg_setup_from(F, A.tN, A.tE, A.tS, A.tW, K.ap_w[0]);
        const StencilOp At{g, A.tN, A.tE, A.tS, A.tW};
        if (int rc = run_bicg<StencilOp, 1, 1, double>(X, K, At, 0, PpPre{F})) return rc;
        if (int rc = adj_solved(F, 0, 1, "pp")) return rc;
    }
    LAUNCH(k_r_ppsolve, nba, g, M, A);
    LAUNCH(k_r_pp_src, nba, g, T, M, A);
    LAUNCH(k_r_pp_coef_flux, nba, g, T, M, A);
    LAUNCH(k_r_pp_prep, nba, g, T, M, A);
    // u, v: A^T lambda = dL/du*, dL/dv*
    {
        LAUNCH(k_r_transpose, nba, g, (const double *)M.au[0], (const double *)M.au[1], (const double *)M.au[2],
               (const double *)M.au[3], A);
        KrySys<2> K = F->Kuv;
        K.x[0] = A.lu;
        K.x[1] = A.lv;
        AdjStart<2> S{M.au[0], M.au[1], M.au[2], M.au[3], {M.spu, M.spv}, {A.u, A.v}, 1.0 / F4(0.8)};
        hipLaunchKernelGGL(k_r_start<2>, dim3(F->nblk_kry), dim3(KBLOCK), 0, F->st, g, K, S);
        const StencilOp At{g, A.tN, A.tE, A.tS, A.tW};
        if (int rc = run_bicg<StencilOp, 2, 0, double>(X, K, At, 1)) return rc;
        if (int rc = adj_solved(F, 1, 2, "u/v")) return rc;
    }
    LAUNCH(k_r_uvsolve, nba, g, M, A);
    LAUNCH(k_r_mom, nba, g, T, P, A);
    return 0;
}

// Reverse of one time step recorded in the tape; geta: dL/deta of the gauges
// after this step (device, ng values).
static int rev_step(cms2d_flow *F, Ad &A, Tape &tp, const int *cells, int ng, const double *geta) {
    const Geo &g = F->g;
    const Par &P = F->P;
    const int nbp = F->nblk_pad, nba = F->nblk_act;
    LAUNCH(k_r_update, nbp, g, A);
    const St S = snap_view(F, tp.snap[tp.niter]);
    if (ng > 0) LAUNCH(k_r_gauge, (ng + BLOCK - 1) / BLOCK, g, S, P, cells, ng, geta, A);
    LAUNCH(k_r_velface, nba, g, S, P, A);
    for (int it = tp.niter - 1; it >= 0; it--) {
        const St T = snap_view(F, tp.snap[it]), E = snap_view(F, tp.snap[it + 1]);
        if (int rc = rev_iter(F, A, T, E, mid_view(F, tp.mid[it]))) return rc;
    }
    const St T0 = snap_view(F, tp.snap[0]);
    LAUNCH(k_r_srcsnk, nba, g, T0, P, A);
    for (int sweep = 0; sweep < 5; sweep++) {
        LAUNCH(k_r_d2w_take, nba, g, T0, A);
        LAUNCH(k_r_d2w_spread, nba, g, T0, A);
    }
    LAUNCH(k_r_wetdry_post, nba, g, T0, A);
    LAUNCH(k_r_wetdry_h, nbp, g, T0, P, A);
    return 0;
}

// Reverse of init_state (the creation state S0): eddy viscosity, fluxes,
// gradients and friction, then the depth from the bed and the face
// elevations.  The initial u, v, p are inputs, so their adjoints end here.
static int rev_init(cms2d_flow *F, Ad &A, const St &S0) {
    const Geo &g = F->g;
    const Par &P = F->P;
    const int nbp = F->nblk_pad, nba = F->nblk_act;
    LAUNCH(k_r_tail0, nbp, g, S0, A);
    LAUNCH(k_r_tail1, nba, g, S0, A);
    LAUNCH(k_r_tail2, nba, g, S0, A);
    LAUNCH(k_r_tail3, nbp, g, S0, A);
    LAUNCH(k_r_convflux, nba, g, S0, A);
    LAUNCH(k_r_gfv, nba, g, S0, P, A);
    LAUNCH(k_r_update, nbp, g, A);          // h1 = h (and u1, v1, p1) of the first wet/dry pass
    LAUNCH(k_r_wetdry_h, nbp, g, S0, P, A);
    LAUNCH(k_r_zbk, nba, g, A);
    return 0;
}

static int adjoint_check(const cms2d_flow *F) {
    if (F->tel) return set_err(CMS2D_ERR_ARG, "the adjoint is built for Cartesian grids");
    if (F->precision != CMS2D_PREC_FP64 || F->solver != CMS2D_SOLVER_BICGSTAB || F->precond != CMS2D_PRECOND_MG)
        return set_err(CMS2D_ERR_ARG, "the adjoint needs the FP64 form, BiCGStab and the multigrid preconditioner");
    if (F->dtvar || F->serial_order) return set_err(CMS2D_ERR_ARG, "the adjoint needs a fixed time step and the default cell order");
    if (F->P.mbedfric != 2) return set_err(CMS2D_ERR_ARG, "the adjoint is built for Manning friction");
    if (F->qside >= 0) return set_err(CMS2D_ERR_ARG, "the adjoint is built without flux sides");
    if (F->P.ndsch > 4) return set_err(CMS2D_ERR_ARG, "the adjoint is built for the schemes without deferred corrections (ndsch 0-4)");
    if (F->P.noptset >= 3 || F->windconst)
        return set_err(CMS2D_ERR_ARG, "the adjoint is built without wave and wind forcing");
    if (F->ntime != 0) return set_err(CMS2D_ERR_ARG, "the adjoint starts from the flow's initial state (no steps taken)");
    return 0;
}

extern "C" int cms2d_flow_adjoint(cms2d_flow *F, int nsteps, int every, int ngauge, const int *gauge_cells,
                                  const double *gauge_obs, const double *geta_in, double *eta_gauge, double *loss,
                                  double *gmann, double *gzb) {
    if (!F || nsteps < 1 || every < 1 || ngauge < 0 || (ngauge > 0 && !gauge_cells) ||
        (ngauge > 0 && !gauge_obs && !geta_in))
        return set_err(CMS2D_ERR_ARG, "invalid argument");
    if (int rc = adjoint_check(F)) return rc;
    const Geo &g = F->g;
    for (int k = 0; k < ngauge; k++)
        if (gauge_cells[k] < 0 || gauge_cells[k] >= g.nact) return set_err(CMS2D_ERR_ARG, "gauge cell out of range");
    HIPCHK(hipSetDevice(F->dev));
    HIPCHK(hipStreamSynchronize(F->st));
    struct StatsOff {
        cms2d_flow *F;
        bool keep;
        ~StatsOff() { F->stats = keep; }
    } stats_off{F, F->stats};
    F->stats = false;
    AdjMem mem;
    const size_t ng = (size_t)ngauge, ns = (size_t)nsteps;
    HIPCHK(hipMalloc((void **)&mem.cells, std::max<size_t>(ng, 1) * sizeof(int)));
    HIPCHK(hipMalloc((void **)&mem.eta, std::max<size_t>(ng * ns, 1) * sizeof(double)));
    HIPCHK(hipMalloc((void **)&mem.geta, std::max<size_t>(ng * ns, 1) * sizeof(double)));
    if (ng) HIPCHK(hipMemcpyAsync(mem.cells, gauge_cells, ng * sizeof(int), hipMemcpyHostToDevice, F->st));
    // forward sweep with checkpoints
    const int ncp = (nsteps + every - 1) / every;
    mem.cp.resize(ncp);
    if (int rc = snap_save(F, mem.cp[0])) return rc;
    auto restore0 = [&]() { return snap_load(F, mem.cp[0]); };
    for (int n = 0; n < nsteps; n++) {
        if (n > 0 && n % every == 0)
            if (int rc = snap_save(F, mem.cp[n / every])) return rc;
        if (int rc = step_once(F)) { restore0(); return rc; }
        if (ng) LAUNCH(k_gauge_eta, (ngauge + BLOCK - 1) / BLOCK, g, F->s, (const int *)mem.cells, ngauge,
                       mem.eta + (size_t)n * ng);
    }
    std::vector<double> eta(ng * ns), geta(ng * ns);
    if (ng) HIPCHK(hipMemcpyAsync(eta.data(), mem.eta, ng * ns * sizeof(double), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    if (eta_gauge) std::copy(eta.begin(), eta.end(), eta_gauge);
    double L = 0.0;
    for (size_t k = 0; k < ng * ns; k++) {
        if (gauge_obs) {
            const double d = eta[k] - gauge_obs[k];
            L += d * d;
            geta[k] = 2.0 * d;
        } else {
            geta[k] = geta_in[k];
        }
    }
    if (loss) *loss = L;
    if (ng) HIPCHK(hipMemcpyAsync(mem.geta, geta.data(), ng * ns * sizeof(double), hipMemcpyHostToDevice, F->st));
    // reverse sweep
    HIPCHK(hipMalloc((void **)&mem.ad, (size_t)AD_NARRAYS * F->npad * sizeof(double)));
    HIPCHK(hipMemsetAsync(mem.ad, 0, (size_t)AD_NARRAYS * F->npad * sizeof(double), F->st));
    Ad A = ad_view(F, mem.ad);
    mem.ss.resize(std::min(every, nsteps));
    for (int s = ncp - 1; s >= 0; s--) {
        const int n0 = s * every, n1 = std::min(nsteps, n0 + every);
        if (int rc = snap_load(F, mem.cp[s])) return rc;
        for (int n = n0; n < n1; n++) {
            if (int rc = snap_save(F, mem.ss[n - n0])) return rc;
            if (n + 1 < n1)
                if (int rc = step_once(F)) { restore0(); return rc; }
        }
        for (int n = n1 - 1; n >= n0; n--) {
            if (int rc = snap_load(F, mem.ss[n - n0])) return rc;
            F->tape = &mem.tape;
            const int rc = step_once(F);
            F->tape = nullptr;
            if (rc) { restore0(); return rc; }
            if (int r2 = rev_step(F, A, mem.tape, mem.cells, ngauge, mem.geta + (size_t)n * ng)) {
                restore0();
                return r2;
            }
        }
    }
    if (int rc = rev_init(F, A, snap_view(F, mem.cp[0]))) return rc;
    // gradients of the active cells; bed ghosts are copies of their cells
    std::vector<double> gm(F->npad), gb(F->npad);
    HIPCHK(hipMemcpyAsync(gm.data(), A.coefman, F->npad * sizeof(double), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipMemcpyAsync(gb.data(), A.zb, F->npad * sizeof(double), hipMemcpyDeviceToHost, F->st));
    if (int rc = restore0()) return rc;
    HIPCHK(hipStreamSynchronize(F->st));
    auto at = [&](int i, int j) -> double & { return gb[(size_t)(j + 1) * g.nxp + i + 1]; };
    for (int i = 0; i < g.nx; i++) {
        at(i, 0) += at(i, -1);
        at(i, g.ny - 1) += at(i, g.ny);
    }
    for (int j = 0; j < g.ny; j++) {
        at(0, j) += at(-1, j);
        at(g.nx - 1, j) += at(g.nx, j);
    }
    for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) {
            const size_t a = (size_t)j * g.nx + i, c = (size_t)(j + 1) * g.nxp + i + 1;
            if (gmann) gmann[a] = gm[c];
            if (gzb) gzb[a] = gb[c];
        }
    // the first cell's n is its north neighbour's (fric_init's copy at creation):
    // its adjoint belongs to that input
    if (gmann) {
        gmann[g.nx] += gmann[0];
        gmann[0] = 0.0;
    }
    return CMS2D_OK;
}
