// cms2d.hip -- host driver and C ABI of the MI355X implicit CMS-Flow step.
//
// One HIP stream per flow object.  The host keeps the reference's scalar
// control (timing, the SIMPLEC outer loop, check_conv) and everything else
// stays on the device for the whole run; only scalars cross per outer
// iteration (Krylov flags and residual norms, wet/dry change flags).
#include "flow_kernels.hip"
#include "krylov_kernels.hip"
#include "mg_kernels.hip"
#include "tel_kernels.hip"
#include "balance_kernels.hip"
#include "adjoint_kernels.hip"
#include "../../include/cms2d_amd.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace cms;

// Fixed grid of the Krylov kernels (4 workgroups per CU on 256 CUs).
static constexpr int KRY_BLOCKS = 1024;

static thread_local std::string g_err;
static int set_err(int code, const std::string &m) {
    g_err = m;
    return code;
}
#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return set_err(CMS2D_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Live timing of selected kernel groups: a timed launch carries start/stop
// events stamped by its own dispatch; after the host synchronises, the
// launches of iterations before the solve stopped (those that did work) are
// added up.
// mg_coarse is the span of a V-cycle's levels below level 0 (events recorded
// on the stream before the first and after the last of those launches).
// Launches inside a captured BiCGStab pair are not timed (an event recorded by
// a graph node gives no elapsed time here); the capture notes each timed
// group's launches per pair, so every replayed launch of an iteration that
// did work is counted, and the average duration comes from the timed direct
// launches (each solve's first two iterations run direct).
enum {
    TG_SPMV_PP = 0, TG_MG_DOWN0, TG_MG_UP0, TG_MG_COARSE, TG_TEL_RESTRICT, TG_TEL_FJAC, TG_TEL_CS_RESTRICT,
    TG_N
};
static const char *const TG_NAME[TG_N] = {"spmv_pp", "mg_down0", "mg_up0", "mg_coarse",
                                          "tel_restrict", "tel_fjac", "tel_cs_restrict"};
constexpr int KT_SLOTS = 4;       // graph slots (the Krylov systems of KryScal)
struct KTimer {
    struct Rec {
        int group, it;
    };
    std::vector<hipEvent_t> pool;   // event pairs of direct launches
    std::vector<Rec> rec;           // one per pair in use
    int it = 0;                     // Krylov iteration of the launches being enqueued
    double ms[TG_N]{};              // timed launches: summed duration and count
    long n[TG_N]{};
    long counted[TG_N]{};           // replayed launches of iterations that did work (not timed)
    int capture_slot = -1;          // >= 0 while a pair is captured into that graph slot
    std::vector<Rec> grec[KT_SLOTS];   // the slot's timed-group launches: group, iteration in the pair (0, 1)
    // A timed launch: events for a direct launch (false if none are left);
    // while capturing, the launch is noted and runs untimed (false).
    bool take(int group, hipEvent_t &a, hipEvent_t &b) {
        if (capture_slot >= 0) {
            grec[capture_slot].push_back({group, it - 2});
            return false;
        }
        const size_t k = rec.size();
        if (2 * k + 2 > pool.size()) return false;
        a = pool[2 * k];
        b = pool[2 * k + 1];
        rec.push_back({group, it});
        return true;
    }
    int harvest(int upto) {
        for (size_t k = 0; k < rec.size(); k++) {
            if (rec[k].it >= upto) continue;
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, pool[2 * k], pool[2 * k + 1]));
            ms[rec[k].group] += t;
            n[rec[k].group] += 1;
        }
        rec.clear();
        return 0;
    }
    // After a batch that replayed the pairs starting at iterations first,
    // first + 2, ..., last of graph slot `slot`: count the launches of
    // iterations below upto (those that did work).
    void count_graph(int slot, int first, int last, int upto) {
        for (int p = first; p <= last; p += 2)
            for (const Rec &g : grec[slot])
                if (p + g.it < upto) counted[g.group] += 1;
    }
};

template <class Kern, class... Args>
static void tlaunch(KTimer *tm, int group, Kern kern, dim3 grid, dim3 blk, size_t lds, hipStream_t st,
                    Args... args) {
    hipEvent_t a, b;
    if (tm && tm->take(group, a, b)) hipExtLaunchKernelGGL(kern, grid, blk, lds, st, a, b, 0, args...);
    else hipLaunchKernelGGL(kern, grid, blk, lds, st, args...);
}

namespace cms {
struct TelHost;
}

namespace cms {
struct Tape;
}

struct cms2d_flow {
    cms::TelHost *tel = nullptr;   // telescoping grid (cell-list step), or null
    int dev = 0;
    hipStream_t st = nullptr;
    Geo g{};
    Par P{};
    St s{};
    KrySys<2> Kuv{};
    KrySys<1> Kpp{};
    int precond = CMS2D_PRECOND_MG;
    int precision = CMS2D_PREC_FP64;
    int solver = CMS2D_SOLVER_BICGSTAB;
    GmSys gm{};                // GMRES basis and scalars (pp system)
    GmScal *h_gm = nullptr;    // pinned mirror
    int prev_gm = 8;
    Mg mg{};                   // pp preconditioner hierarchy (level 0 z is Kpp.zp / Kpp.zs)
    KrySys<1, float> Kin{};    // inner FP32 solve of the refinement form (pp)
    KrySys<2, float> Kin2{};   // inner FP32 solves of the refinement form (u, v)
    float *a32[4]{};           // their FP32 coefficients
    KryScal *d_sc = nullptr;   // [6]: pp, u, v, inner FP32 pp, inner FP32 u, v
    KryScal *h_sc = nullptr;   // pinned mirror
    int *h_flags = nullptr;    // pinned
    std::vector<void *> allocs;
    std::vector<std::pair<void *, size_t>> state;   // device arrays of the model state (checkpointed)
    int npad = 0, nblk_act = 0, nblk_pad = 0, nblk_kry = 0;
    int ntc = 0;
    double amp[8]{}, speed[8]{}, phase[8]{};
    double timesecs = 0.0, rampdur = 0.0, ramp = 0.0;
    int dtvar = 0, jtime = 0, mtime = 0;       // variable time step (flow.F90:107, 664-666)
    double dtimebeg = 0.0, deltime = 0.0;
    int ntime = 0, maxit = 0, niter = 0, iconv = 1;
    double rmom[3]{}, rmom0[3]{}, rmom1[3]{};
    int kry[3]{}, prev_iters[6]{4, 2, 2, 4, 2, 2};
    // timing of the pp-system SpMV launches (events bracket each launch)
    bool stats = true;
    KTimer tm;                 // live timing of the pp solve's main kernels
    KTimer *tm_active = nullptr;
    // launch-bound grids: BiCGStab iteration pairs replayed from captured
    // graphs, one per Krylov system (indexed like d_sc)
    bool use_graphs = false;
    hipGraphExec_t kgraph[6]{};
    int mg_tile = 128;         // level-0 tile width of the multigrid cycle (timed at creation)
    // the reference's cell order for the three order-dependent loops: the
    // buffers of consecutive passes (dry-to-wet u, v, vis; fix-up face values
    // [4*npad]; corner values)
    int serial_order = 0;
    double *ser_d2w[2][3]{};
    double *ser_face[2]{};
    double *ser_corner[2]{};
    int windconst = 0;         // spatially constant wind (met_def windconst) and its 10-m vector
    double wndx = 0.0, wndy = 0.0;
    int side_type[4]{};        // CMS2D_BND_* per side (the kernels see 1 for every water-level side,
                               // 2 for the flux side)
    // the flux (river) side, Q_str: its side (-1 none), total inflow or curve, interpolation
    // order and window start (Q_str%nti, %inc), conveyance, inflow direction, and this
    // step's inflow (bnd_eval)
    int qside = -1;
    double q_const = 0.0;
    std::vector<double> q_t, q_v;
    int q_order = 1, q_inc = 1;
    double cmvel = 0.0, qcos = 0.0, qsin = 0.0, qflux = 0.0;
    int q_rating = 0;          // RATING_CURVE: the table and the step's inflow on the device
    double *q_dev = nullptr;   // [1 + 2m]: inflow, stages, flows
    std::vector<double> wse_t, wse_v;   // water-level curve of CMS2D_BND_WSE_CURVE sides
    int wse_order = 1, wse_inc = 1;     // interpolation order, window start (H_str%inc, 1-based)
    int wse_adjust_off = 0;             // WSE_ADJUSTMENT OFF on the curve sides
    double wse_gx = 0.0, wse_gy = 0.0;  // WSE_GRADIENTS of the curve sides
    hipEvent_t step_ev[2]{};
    cms::Tape *tape = nullptr;   // records the outer iterations for the reverse sweep (or null)
    // volume balance and extremes of the last step (flow_balance, flow_step_stat):
    // per-block partials on the fixed Krylov grid and the running totals
    double *bal_part = nullptr;
    cms::BalAcc *bal = nullptr;
    // flow statistics (FLOW_STATISTICS): window [stat_t1, stat_t2] s, steps counted,
    // accumulators [NSTAT][npad], the new-minimum flag
    int stat_on = 0, stat_n = 0;
    double stat_t1 = 0.0, stat_t2 = 0.0;
    double *stat_acc = nullptr;
    int *stat_flag = nullptr;
};

// Lagrange interpolation of the curve (t, v) at time xi with the window search
// of plagr_fit (plagr_lib.f90:20-138): order np = min(ni, m - 1); constant
// extension outside the curve; the window start k (1-based, kept between calls
// like H_str%inc) moves by a forward search, else backwards, else bisection,
// so that x(k) <= xi <= x(k+np).  Weights normalised by their sum.
static double wse_curve_eval(const std::vector<double> &t, const std::vector<double> &v, int ni, double xi, int &k) {
    const int m = (int)t.size();
    const int np = std::min(ni, m - 1);
    auto x = [&](int i) { return t[i - 1]; };
    if (xi < x(1)) { k = 1; return v[0]; }
    if (xi > x(m)) { k = m - np; return v[m - 1]; }
    if (k >= 1) {
        const int i0 = k;
        for (k = i0; k <= m - np; k++)
            if (xi <= x(k + np)) break;
        if (k > m - np || xi < x(k) || xi > x(k + np)) {
            for (k = std::min(i0, m); k >= 1; k--)
                if (x(k) <= xi) break;
            if (k < 1 || k + np > m || xi < x(k) || xi > x(k + np)) k = 0;
        }
    }
    if (k <= 0) {
        int i = 1, j = m - np + 1;
        while (j > i + 1) {
            k = (i + j) / 2;
            if (xi < x(k)) j = k;
            else i = k;
        }
        k = i;
    }
    double lb[4], sum = 0.0;
    for (int i = 1; i <= np + 1; i++) {
        lb[i - 1] = 1.0;
        for (int j = 1; j <= np + 1; j++)
            if (i != j) lb[i - 1] = lb[i - 1] * (xi - x(j + k - 1)) / (x(i + k - 1) - x(j + k - 1));
        sum += lb[i - 1];
    }
    double w = 0.0;
    for (int i = 1; i <= np + 1; i++) w += lb[i - 1] / sum * v[i + k - 2];
    return w;
}

extern "C" const char *cms2d_last_error(void) { return g_err.c_str(); }

extern "C" int cms2d_device_count(int *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return CMS2D_OK;
}

template <typename T>
static int dalloc(cms2d_flow *F, T **p, size_t n) {
    void *q = nullptr;
    HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    HIPCHK(hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), F->st));
    F->allocs.push_back(q);
    *p = (T *)q;
    return 0;
}

#define LAUNCH(kern, nb, ...) hipLaunchKernelGGL(kern, dim3(nb), dim3(BLOCK), 0, F->st, __VA_ARGS__)

static int sync_flags(cms2d_flow *F) {
    HIPCHK(hipMemcpyAsync(F->h_flags, F->s.flags, 4 * sizeof(int), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    return 0;
}

// Passes of one of the serial-order recurrences until a pass changes nothing:
// pass(first, from, to) evaluates into buffer `to` from buffer `from` (the
// first pass from the values before the loop); `res` is the buffer that holds
// the settled values.  A pass fixes at least the next cell of every chain, so
// the count is bounded by the number of cells.
template <class Pass>
static int serial_settle(cms2d_flow *F, Pass pass, int &res) {
    int *flag = F->s.flags + 3;
    pass(1, 0, 0);
    int cur = 0;
    const long cap = (long)F->g.nx * F->g.ny + 2;
    for (long q = 0; q < cap; q++) {
        LAUNCH(k_set_int, 1, flag, 1, 0);
        pass(0, cur, cur ^ 1);
        cur ^= 1;
        HIPCHK(hipMemcpyAsync(F->h_flags + 3, flag, sizeof(int), hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
        if (!F->h_flags[3]) {
            res = cur;
            return 0;
        }
    }
    return set_err(CMS2D_ERR_HIP, "serial-order passes did not settle");
}

// flow_wetdry(0), flow.F90:1231-1539
static int wetdry(cms2d_flow *F) {
    Geo &g = F->g;
    St &s = F->s;
    LAUNCH(k_wetdry_h, F->nblk_pad, g, s, F->P);
    for (int it = 0;; it++) {
        LAUNCH(k_set_int, 1, s.flags, 3, 0);
        LAUNCH(k_narrow, F->nblk_act, g, s, F->P);
        if (int rc = sync_flags(F)) return rc;
        if (F->h_flags[0] == 0) break;
        if (it > 100000) return set_err(CMS2D_ERR_HIP, "narrow-channel pass did not settle");
    }
    {
        const int ntx = (g.nxp + PT - 1) / PT, nty = (g.nyp + PT - 1) / PT;
        const int nedge = (ntx - 1) * g.nyp + (nty - 1) * g.nxp;
        hipLaunchKernelGGL(k_pond_local, dim3(ntx, nty), dim3(256), 0, F->st, g, s);
        if (nedge > 0) LAUNCH(k_pond_merge, (nedge + BLOCK - 1) / BLOCK, g, s);
        LAUNCH(k_pond_compress, F->nblk_pad, g, s);
        LAUNCH(k_pond_mark, F->nblk_pad, g, s);
        LAUNCH(k_pond_final, F->nblk_pad, g, s);
    }
    LAUNCH(k_set_int, 1, s.flags, 3, 0);
    LAUNCH(k_wetdry_post, F->nblk_act, g, s);
    if (int rc = sync_flags(F)) return rc;
    if (F->h_flags[2] > 0) {
        for (int sweep = 0; sweep < 5; sweep++) {
            if (F->serial_order) {
                int res = 0;
                auto pass = [&](int first, int from, int to) {
                    double *const *a = F->ser_d2w[from], *const *b = F->ser_d2w[to];
                    LAUNCH(k_d2w_serial, F->nblk_act, g, s, (const double *)a[0], (const double *)a[1],
                           (const double *)a[2], b[0], b[1], b[2], first, s.flags + 3);
                };
                if (int rc = serial_settle(F, pass, res)) return rc;
                LAUNCH(k_d2w_serial_commit, F->nblk_act, g, s, (const double *)F->ser_d2w[res][0],
                       (const double *)F->ser_d2w[res][1], (const double *)F->ser_d2w[res][2]);
                continue;
            }
            LAUNCH(k_d2w_sweep, F->nblk_act, g, s);
            LAUNCH(k_d2w_commit, F->nblk_act, g, s);
        }
    }
    return 0;
}

// Krylov launch context: stream, fixed grid and scalar mirrors.
struct KryCtx {
    hipStream_t st;
    int nblk;                      // fixed Krylov grid
    int kmax;
    double rtol, atol;
    KryScal *d_sc, *h_sc;
    int *prev_iters;
    KTimer *tm;                    // timed launches (or null)
    hipGraphExec_t *graph = nullptr;   // per-system graph slots (or null: direct launches)
};

// BiCGStab iterations from the third on replay pairs from a graph captured
// at creation.  On small grids a launch lasts about as long as the host takes
// to issue it; on large ones the graph still shortens the gaps between
// dependent launches (the 2000 x 1500 step: 41.2 -> 37.7 ms).  The kernels
// and their arguments are those of direct launches, so results are bitwise
// the same either way.  CMS2D_GRAPHS=0 in the environment turns it off.

template <class Body>
static int capture_graph(hipStream_t st, hipGraphExec_t &ge, Body body) {
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    body();
    const hipError_t e = hipStreamEndCapture(st, &gr);
    if (e != hipSuccess) return set_err(CMS2D_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e2 != hipSuccess) {
        ge = nullptr;
        return set_err(CMS2D_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e2));
    }
    return 0;
}

struct NoPrec {
    void operator()(int) const {}
};

// Runs BiCGStab iterations in batches until every system has stopped.  A
// batch is enqueued without host synchronisation; kernels launched after a
// system has stopped return at once (they read the device-side done flag).
// PREC 1: pre(0) enqueues zp = M^-1 p, pre(1) enqueues zs = M^-1 s.
// One BiCGStab iteration (index i) of the systems in K.
template <class Op, int NS, int PREC, class T, class Pre>
static void bicg_iter(const KryCtx &X, KrySys<NS, T> &K, Op A, Pre &pre, int i) {
    const dim3 grid(X.nblk), blk(KBLOCK);
    if (X.tm) X.tm->it = i;
    hipLaunchKernelGGL((k_bicg_p<Op, NS, T>), grid, blk, 0, X.st, A, K, i, X.nblk, X.kmax, X.rtol, X.atol, 0);
    if (PREC) pre(0);
    tlaunch(X.tm, TG_SPMV_PP, k_bicg_spmv<Op, NS, 1, PREC, T>, grid, blk, 0, X.st, A, K);
    hipLaunchKernelGGL((k_bicg_s<Op, NS, PREC, T>), grid, blk, 0, X.st, A, K, i, X.nblk);
    if (PREC) pre(1);
    tlaunch(X.tm, TG_SPMV_PP, k_bicg_spmv<Op, NS, 2, PREC, T>, grid, blk, 0, X.st, A, K);
    hipLaunchKernelGGL((k_bicg_r<Op, NS, PREC, T>), grid, blk, 0, X.st, A, K, i, X.nblk);
}

// Captures iterations 2 and 3 of the systems in K into ge.  Iterations from 2
// on depend on their index only through its parity (k_bicg_p/s/r use i & 1
// and i == 0), so the pair replays as any later pair (2m, 2m+1).
template <class Op, int NS, int PREC = 0, class T = double, class Pre = NoPrec>
static int capture_bicg_pair(const KryCtx &X, KrySys<NS, T> &K, Op A, hipGraphExec_t &ge, Pre pre = Pre()) {
    return capture_graph(X.st, ge, [&]() {
        bicg_iter<Op, NS, PREC, T>(X, K, A, pre, 2);
        bicg_iter<Op, NS, PREC, T>(X, K, A, pre, 3);
    });
}

template <class Op, int NS, int PREC = 0, class T = double, class Pre = NoPrec>
static int run_bicg(KryCtx &X, KrySys<NS, T> &K, Op A, int sc0, Pre pre = Pre()) {
    const dim3 blk(KBLOCK);
    // start: thresholds, rmom and the initial stop test (read with the first
    // batch: a solve that needs no iteration only costs no-op launches)
    hipLaunchKernelGGL((k_bicg_p<Op, NS, T>), dim3(1), blk, 0, X.st, A, K, 0, X.nblk, X.kmax, X.rtol, X.atol, 1);
    auto all_done = [&]() {
        int d = 1;
        for (int q = 0; q < NS; q++) d &= X.h_sc[sc0 + q].done;
        return d;
    };
    // batches: the last solve's count, then 2, 2, 4, 8, ... so that few
    // launches follow convergence (each costs a few microseconds even when it
    // returns at once) and few host round trips are needed
    int it = 0, nbatch = 0, grow = 2;
    int batch = std::max(1, X.prev_iters[sc0]);
    const hipGraphExec_t ge = X.graph ? X.graph[sc0] : nullptr;
    for (bool first = true; first || (!all_done() && it < X.kmax); first = false) {
        const int nb = std::min(batch, X.kmax - it);
        int rfirst = -1, rlast = -1;     // first and last replayed pair of the batch
        for (int b = 0; b < nb;) {
            if (ge && it >= 2 && (it & 1) == 0 && nb - b >= 2) {
                HIPCHK(hipGraphLaunch(ge, X.st));
                if (rfirst < 0) rfirst = it;
                rlast = it;
                b += 2;
                it += 2;
            } else {
                bicg_iter<Op, NS, PREC, T>(X, K, A, pre, it);
                b++;
                it++;
            }
        }
        hipLaunchKernelGGL((k_bicg_p<Op, NS, T>), dim3(1), blk, 0, X.st, A, K, it, X.nblk, X.kmax, X.rtol, X.atol, 1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(X.h_sc + sc0, X.d_sc + sc0, NS * sizeof(KryScal), hipMemcpyDeviceToHost, X.st));
        HIPCHK(hipStreamSynchronize(X.st));
        if (X.tm) {
            // launches that did work: those of iterations the first system completed
            const KryScal &s0 = X.h_sc[sc0];
            const int upto = s0.done ? s0.final_iters : it;
            if (int rc = X.tm->harvest(upto)) return rc;
            if (rfirst >= 0) X.tm->count_graph(sc0, rfirst, rlast, upto);
        }
        batch = grow;
        if (++nbatch >= 2) grow = std::min(2 * grow, 64);
    }
    for (int q = 0; q < NS; q++) X.prev_iters[sc0 + q] = std::max(1, X.h_sc[sc0 + q].final_iters);
    return 0;
}

static int mg_nblk(const MgLevel &L) { return (L.nx * L.ny + BLOCK - 1) / BLOCK; }

// Builds the hierarchy from the assembled pp system (after the solve prologue
// has written ap).
static void mg_setup_from(cms2d_flow *F, const double *aN, const double *aE, const double *aS, const double *aW,
                          const double *ap) {
    Mg &M = F->mg;
    LAUNCH(k_mg_setup0, mg_nblk(M.L[0]), F->g, aN, aE, aS, aW, ap, M.L[0]);
    for (int l = 0; l + 1 < M.nlev; l++) LAUNCH(k_mg_coarsen, mg_nblk(M.L[l + 1]), M.L[l], M.L[l + 1]);
}
static void mg_setup(cms2d_flow *F) {
    const St &s = F->s;
    mg_setup_from(F, s.aN, s.aE, s.aS, s.aW, F->Kpp.ap_w[0]);
}

// The level-0 launch of the cycle (down or up) with tile width TX.
template <int TX, class TB>
static void mg_level0_tx(cms2d_flow *F, KTimer *tm, bool down, const MgLevel &L0, const TB *b0, const KryScal *sc) {
    using T = MTile<TX>;
    const dim3 tiles((L0.nx + T::X - 1) / T::X, (L0.ny + T::Y - 1) / T::Y), tb(MT_THREADS);
    if (down) tlaunch(tm, TG_MG_DOWN0, k_mg_down<TX, TB>, tiles, tb, 0, F->st, L0, b0, F->mg.L[1], sc);
    else tlaunch(tm, TG_MG_UP0, k_mg_up<TX, TB>, tiles, tb, 0, F->st, L0, b0, F->mg.L[1], sc);
}
template <class TB>
static void mg_level0(cms2d_flow *F, KTimer *tm, bool down, const MgLevel &L0, const TB *b0, const KryScal *sc) {
    if (F->mg_tile == 64) mg_level0_tx<64>(F, tm, down, L0, b0, sc);
    else if (F->mg_tile == 256) mg_level0_tx<256>(F, tm, down, L0, b0, sc);
    else mg_level0_tx<128>(F, tm, down, L0, b0, sc);
}

// z0 = V(b0): one V(1,1) cycle (mg.h) on the stream of F; every launch
// returns at once when the solve owning sc has stopped.
template <class TB>
static void mg_vcycle(cms2d_flow *F, const TB *b0, float *z0, const KryScal *sc) {
    const Mg &M = F->mg;
    MgLevel L0 = M.L[0];
    L0.zo = z0;
    const MgLevel &Lc = M.nlev == 1 ? L0 : M.L[M.nlev - 1];
    const size_t lds = sizeof(float) * (size_t)Lc.nxp * (Lc.ny + 2);
    KTimer *tm = F->tm_active;
    if (M.nlev == 1) {
        hipLaunchKernelGGL(k_mg_coarsest<TB>, dim3(1), dim3(MG_CT), lds, F->st, L0, b0, sc);
        return;
    }
    mg_level0(F, tm, true, L0, b0, sc);
    hipEvent_t ea, eb;
    const bool span = tm && tm->take(TG_MG_COARSE, ea, eb);
    if (span) (void)hipEventRecord(ea, F->st);
    for (int l = 1; l + 1 < M.nlev; l++)
        LAUNCH(k_mg_down_pc<float>, mg_nblk(M.L[l + 1]), M.L[l], (const float *)M.L[l].b, M.L[l + 1], sc);
    hipLaunchKernelGGL(k_mg_coarsest<float>, dim3(1), dim3(MG_CT), lds, F->st, Lc, (const float *)Lc.b, sc);
    for (int l = M.nlev - 2; l >= 1; l--)
        LAUNCH(k_mg_up_pc<float>, mg_nblk(M.L[l + 1]), M.L[l], (const float *)M.L[l].b, M.L[l + 1], sc);
    if (span) (void)hipEventRecord(eb, F->st);
    mg_level0(F, tm, false, L0, b0, sc);
}

// Picks the level-0 tile width of the cycle by timing the down and up launches
// of each candidate on this grid; every width gives the same values.
static int tune_mg_tile(cms2d_flow *F) {
    static constexpr int cand[3] = {64, 128, 256};
    if (F->mg.nlev < 2) return 0;
    if (const char *e = getenv("CMS2D_MG_TILE")) {   // fixed width (tests)
        const int t = atoi(e);
        if (t == 64 || t == 128 || t == 256) { F->mg_tile = t; return 0; }
    }
    MgLevel L0 = F->mg.L[0];
    L0.zo = F->Kpp.zp;
    const double *b0 = F->Kpp.p[0];
    HIPCHK(hipMemsetAsync(F->d_sc, 0, sizeof(KryScal), F->st));   // done = 0: the launches work
    hipEvent_t ev[2];
    HIPCHK(hipEventCreate(&ev[0]));
    HIPCHK(hipEventCreate(&ev[1]));
    float best = 0.0f;
    int pick = F->mg_tile;
    for (int t : cand) {
        F->mg_tile = t;
        mg_level0(F, nullptr, true, L0, b0, F->d_sc);          // warm-up
        mg_level0(F, nullptr, false, L0, b0, F->d_sc);
        HIPCHK(hipEventRecord(ev[0], F->st));
        for (int r = 0; r < 4; r++) {
            mg_level0(F, nullptr, true, L0, b0, F->d_sc);
            mg_level0(F, nullptr, false, L0, b0, F->d_sc);
        }
        HIPCHK(hipEventRecord(ev[1], F->st));
        HIPCHK(hipEventSynchronize(ev[1]));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        if (t == cand[0] || ms < best) { best = ms; pick = t; }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    F->mg_tile = pick;
    return 0;
}

// Inner FP32 solve: relative stop and iteration cap per refinement sweep.  A
// loose inner stop is the fastest on the tidal workload at 2000 x 1500 (pp
// iterations per step: 60 at 1e-2, 64 at 1e-3, 78 at 1e-5, where FP32
// BiCGStab starts to stall); the FP64 sweeps still reach the FP64 solve's
// own threshold.
#ifndef CMS2D_REF_RTOL
#define CMS2D_REF_RTOL 1e-2
#endif
static constexpr double REF_RTOL = CMS2D_REF_RTOL;
static constexpr int REF_KMAX = 200;
static constexpr int REF_SWEEPS = 40;
// The u and v systems (Jacobi, diagonally dominant) converge at a steady rate
// to FP32 BiCGStab's floor near 1e-6: a 1e-5 inner stop needs two sweeps for
// the FP64 1e-10 where 1e-2 needs five, each with its own host round trip.
static constexpr double REF_RTOL_UV = 1e-5;

// Multigrid preconditioners of the pp solve (FP64 Krylov vectors) and of the
// inner FP32 solve.
struct PpPre {
    cms2d_flow *F;
    void operator()(int which) const {
        if (which == 0) mg_vcycle<double>(F, F->Kpp.p[0], F->Kpp.zp, F->d_sc);
        else mg_vcycle<double>(F, F->Kpp.r[0], F->Kpp.zs, F->d_sc);
    }
};
struct InnerPre {
    cms2d_flow *F;
    void operator()(int which) const {
        KrySys<1, float> &Ki = F->Kin;
        if (which == 0) mg_vcycle<float>(F, Ki.p[0], Ki.zp, Ki.sc);
        else mg_vcycle<float>(F, Ki.r[0], Ki.zs, Ki.sc);
    }
};

// Context and operator of the inner FP32 solve.
static KryCtx inner_ctx(const KryCtx &X) {
    KryCtx Xi = X;
    Xi.kmax = REF_KMAX;
    Xi.rtol = REF_RTOL;
    Xi.atol = 0.0;
    Xi.tm = nullptr;
    return Xi;
}
static StencilOp32 inner_op(const cms2d_flow *F) {
    const MgLevel &L0 = F->mg.L[0];
    return StencilOp32{F->g, L0.aN, L0.aE, L0.aS, L0.aW};
}

// The pp solve of the FP32 form: FP32 BiCGStab (FP32 operator and vectors,
// multigrid) for the correction, FP64 x and residual, FP64 stop test on the
// FP64 solve's own threshold.  Krylov iterations are summed over the sweeps.
static int pp_solve_refine(cms2d_flow *F, KryCtx &X, StencilOp A) {
    KrySys<1> &K = F->Kpp;
    KrySys<1, float> &Ki = F->Kin;
    const int nb = F->nblk_kry;
    hipLaunchKernelGGL((k_bicg_p<StencilOp, 1, double>), dim3(1), dim3(KBLOCK), 0, F->st, A, K, 0, nb, X.kmax,
                       X.rtol, X.atol, 1);
    HIPCHK(hipMemcpyAsync(F->h_sc, F->d_sc, sizeof(KryScal), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    const StencilOp32 A32 = inner_op(F);
    KryCtx Xi = inner_ctx(X);
    InnerPre pre{F};
    int total = 0;
    for (int sw = 0; sw < REF_SWEEPS && !F->h_sc[0].done; sw++) {
        hipLaunchKernelGGL(k_ref_start<StencilOp32>, dim3(nb), dim3(KBLOCK), 0, F->st, A32, Ki, (const double *)K.r[0]);
        if (int rc = run_bicg<StencilOp32, 1, 1, float>(Xi, Ki, A32, 3, pre)) return rc;
        total += F->h_sc[3].final_iters;
        hipLaunchKernelGGL(k_ref_update<StencilOp>, dim3(nb), dim3(KBLOCK), 0, F->st, A, K, (const float *)Ki.x[0]);
        hipLaunchKernelGGL(k_ref_check, dim3(1), dim3(KBLOCK), 0, F->st, K, nb);
        HIPCHK(hipMemcpyAsync(F->h_sc, F->d_sc, sizeof(KryScal), hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
    }
    F->h_sc[0].final_iters = total;
    return 0;
}

// The u and v solves of the FP32 form: FP32 Jacobi BiCGStab of both systems
// (FP32 coefficients and vectors) for each correction, x and the residuals in
// FP64, each system stopped on its FP64 solve's own threshold.  Krylov
// iterations are summed over the sweeps.
static int uv_solve_refine(cms2d_flow *F, KryCtx &X, StencilOp A) {
    KrySys<2> &K = F->Kuv;
    KrySys<2, float> &Ki = F->Kin2;
    const int nb = F->nblk_kry;
    hipLaunchKernelGGL((k_bicg_p<StencilOp, 2, double>), dim3(1), dim3(KBLOCK), 0, F->st, A, K, 0, nb, X.kmax,
                       X.rtol, X.atol, 1);
    hipLaunchKernelGGL(k_uv_to32, dim3(nb), dim3(KBLOCK), 0, F->st, A, K, F->a32[0], F->a32[1], F->a32[2],
                       F->a32[3], Ki);
    HIPCHK(hipMemcpyAsync(F->h_sc + 1, F->d_sc + 1, 2 * sizeof(KryScal), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    const StencilOp32 A32{F->g, F->a32[0], F->a32[1], F->a32[2], F->a32[3]};
    KryCtx Xi = inner_ctx(X);
    Xi.rtol = REF_RTOL_UV;
    int total[2] = {0, 0};
    for (int sw = 0; sw < REF_SWEEPS && !(F->h_sc[1].done && F->h_sc[2].done); sw++) {
        hipLaunchKernelGGL(k_ref_start2<StencilOp32>, dim3(nb), dim3(KBLOCK), 0, F->st, A32, Ki, K);
        if (int rc = run_bicg<StencilOp32, 2, 0, float>(Xi, Ki, A32, 4)) return rc;
        for (int q = 0; q < 2; q++) total[q] += F->h_sc[4 + q].final_iters;
        hipLaunchKernelGGL(k_ref_update2<StencilOp>, dim3(nb), dim3(KBLOCK), 0, F->st, A, K, Ki);
        hipLaunchKernelGGL(k_ref_check2, dim3(1), dim3(KBLOCK), 0, F->st, K, nb);
        HIPCHK(hipMemcpyAsync(F->h_sc + 1, F->d_sc + 1, 2 * sizeof(KryScal), hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
    }
    F->h_sc[1].final_iters = total[0];
    F->h_sc[2].final_iters = total[1];
    return 0;
}

// The pp solve by restarted GMRES with the multigrid preconditioner: cycles
// of Arnoldi steps enqueued in batches (launches after the cycle ended return
// at once), each cycle closed by x += M^-1 V y and the true residual.
static int pp_solve_gmres(cms2d_flow *F, KryCtx &X, StencilOp A) {
    GmSys &G = F->gm;
    const int nb = F->nblk_kry, n = F->g.nact;
    const dim3 grid(nb), blk(KBLOCK);
    auto sync = [&]() -> int {
        HIPCHK(hipMemcpyAsync(F->h_gm, G.sc, sizeof(GmScal), hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
        return 0;
    };
    hipLaunchKernelGGL(k_gm_begin, dim3(1), blk, 0, F->st, G, (const double *)F->Kpp.red.pr[0],
                       (const double *)(F->Kpp.red.pr[0] + nb), nb, 1, n, X.rtol, X.atol);
    if (int rc = sync()) return rc;
    int batch = std::max(1, F->prev_gm - 1);
    while (!F->h_gm->ks.done && F->h_gm->iters < X.kmax) {
        hipLaunchKernelGGL(k_gm_v0<StencilOp>, grid, blk, 0, F->st, A, G);
        int j = 0, grow = 2, nbatch = 0;
        for (;;) {
            const int cnt = std::min(batch, GM_M - j);
            for (int b = 0; b < cnt; b++, j++) {
                mg_vcycle<double>(F, G.V[j], G.z, &G.sc->ks);
                hipLaunchKernelGGL(k_gm_spmv_dot<StencilOp>, grid, blk, 0, F->st, A, G, j);
                hipLaunchKernelGGL(k_gm_hred, dim3(1), blk, 0, F->st, G, j, nb);
                hipLaunchKernelGGL(k_gm_orth<StencilOp>, grid, blk, 0, F->st, A, G, j);
                hipLaunchKernelGGL(k_gm_finish, dim3(1), blk, 0, F->st, G, j, nb);
                hipLaunchKernelGGL(k_gm_scale<StencilOp>, grid, blk, 0, F->st, A, G, j);
            }
            HIPCHK(hipGetLastError());
            if (int rc = sync()) return rc;
            if (F->h_gm->cycle_end || j >= GM_M) break;
            batch = grow;
            if (++nbatch >= 2) grow = std::min(2 * grow, 64);
        }
        const int k = F->h_gm->j;
        if (k > 0) {
            hipLaunchKernelGGL(k_gm_solve, dim3(1), dim3(64), 0, F->st, G);
            hipLaunchKernelGGL(k_gm_combine<StencilOp>, grid, blk, 0, F->st, A, G, k);
            mg_vcycle<double>(F, G.V[k], G.z, &G.sc->ks_final);
            hipLaunchKernelGGL(k_gm_update<StencilOp>, grid, blk, 0, F->st, A, G);
        }
        // next cycle (or the stop) from the true residual
        hipLaunchKernelGGL(k_gm_begin, dim3(1), blk, 0, F->st, G, (const double *)G.part, (const double *)nullptr,
                           nb, 0, n, X.rtol, X.atol);
        if (int rc = sync()) return rc;
        batch = std::max(1, std::min(F->prev_gm, GM_M) - 1);
        if (k == 0) break;
    }
    F->prev_gm = std::max(1, F->h_gm->iters);
    F->h_sc[0].rmom = F->h_gm->ks.rmom;
    F->h_sc[0].final_iters = F->h_gm->iters;
    return 0;
}

static KryCtx flow_ctx(cms2d_flow *F, bool timed) {
    KryCtx X;
    X.st = F->st; X.nblk = F->nblk_kry; X.kmax = F->P.kmax;
    X.rtol = F->P.rtol; X.atol = F->P.atol;
    X.d_sc = F->d_sc; X.h_sc = F->h_sc; X.prev_iters = F->prev_iters;
    X.tm = (timed && F->stats) ? &F->tm : nullptr;
    X.graph = F->use_graphs ? F->kgraph : nullptr;
    return X;
}

// Captures the BiCGStab iteration pairs of the flow's systems (u/v, pp or the
// inner FP32 pp solve) once, at creation, in the creating thread; solves only
// replay them.
static int capture_graphs(cms2d_flow *F) {
    if (!F->use_graphs) return 0;
    const St &s = F->s;
    const StencilOp A{F->g, s.aN, s.aE, s.aS, s.aW};
    const KryCtx X = flow_ctx(F, false);
    if (F->precision == CMS2D_PREC_FP32_REFINE) {
        // the inner FP32 pairs of the refinement form (u and v; pp)
        const StencilOp32 A32{F->g, F->a32[0], F->a32[1], F->a32[2], F->a32[3]};
        if (int rc = capture_bicg_pair<StencilOp32, 2, 0, float>(inner_ctx(X), F->Kin2, A32, F->kgraph[4])) return rc;
        if (int rc = capture_bicg_pair<StencilOp32, 1, 1, float>(inner_ctx(X), F->Kin, inner_op(F), F->kgraph[3],
                                                                 InnerPre{F}))
            return rc;
        return 0;
    }
    if (int rc = capture_bicg_pair<StencilOp, 2, 0, double>(X, F->Kuv, A, F->kgraph[1])) return rc;
    if (F->solver == CMS2D_SOLVER_GMRES) {
        // the pp solve is GMRES
    } else if (F->precond == CMS2D_PRECOND_MG) {
        // the pp pair, its timed-group launches noted for the launch counts
        KryCtx Xt = X;
        if (F->stats) {
            Xt.tm = &F->tm;
            F->tm.capture_slot = 0;
            F->tm_active = &F->tm;
        }
        const int rc = capture_bicg_pair<StencilOp, 1, 1, double>(Xt, F->Kpp, A, F->kgraph[0], PpPre{F});
        F->tm.capture_slot = -1;
        F->tm_active = nullptr;
        if (rc) return rc;
    } else {
        if (int rc = capture_bicg_pair<StencilOp, 1, 0, double>(X, F->Kpp, A, F->kgraph[0])) return rc;
    }
    return 0;
}

// wind_drag_hsu (meteorological/met_lib.f90:33-66): Hsu (1988), reduced above
// 30 m/s after Powell et al. (2003); wkappa = 0.4 (met_def)
static double wind_drag_hsu(double w) {
    if (w <= 30.0) {
        const double t = F4(0.4) / (F4(14.56) - 2.0 * std::log(std::max(w, F4(0.01))));
        return t * t;
    }
    return std::max(F4(3.86e-3) - F4(4.0e-5) * w, F4(0.0015));
}

static double ramp_func(double timehrs, double rampdur) {
    return 0.5 - 0.5 * std::cos(M_PI * std::min(timehrs, rampdur) / rampdur);
}

// check_conv, flow_imp.F90:1576-1747
static void check_conv(cms2d_flow *F) {
    double *rmom = F->rmom, *r0 = F->rmom0, *r1 = F->rmom1;
    const int niter = F->niter, maxit = F->maxit;
    F->iconv = 1;
    if (niter <= 2) {
        for (int n = 0; n < 3; n++) r0[n] = std::max(rmom[n], F4(1e-15));
        for (int n = 0; n < 3; n++) r1[n] = r0[n];
    }
    double rmax = std::max(rmom[0], std::max(rmom[1], rmom[2]));
    if (rmax > 5.0) { F->iconv = 0; return; }
    if (niter <= 5) return;
    int nhalf = maxit / 2;
    double rp = rmom[0] / r0[0];
    double ruv = std::max(rmom[1] / r0[1], rmom[2] / r0[2]);
    const double RP = F4(0.5), RUV = 1.0;
    if ((rp >= RP * 10.0 || ruv >= RUV * 10.0) && rmom[0] > F4(1.0e-4)) { F->iconv = 0; return; }
    if (niter == maxit && (rp >= RP || ruv >= RUV) && rmom[0] > F4(1.0e-3)) { F->iconv = 0; return; }
    if (niter == maxit && (rp >= RP / 5.0 || ruv >= RUV / 5.0) && rmom[0] > F4(1.0e-4)) {
        F->iconv = F->dtvar ? 3 : 0;                 // 3: reduce the next time step
        return;
    }
    double rmomuv = std::max(rmom[1], rmom[2]);
    if (rmom[0] <= F4(1.0e-8) && rmomuv <= F4(1.0e-7)) { F->iconv = 2; return; }
    if (niter == maxit && (rmom[0] > F4(1.0e-3) || rmomuv > F4(1.0e-2))) { F->iconv = 0; return; }
    if (niter == maxit && (rmom[0] > F4(1.0e-4) || rmomuv > F4(1.0e-3))) { F->iconv = F->dtvar ? 3 : 0; return; }
    if (niter <= nhalf || niter % 2 != 0) return;
    double rchp = r1[0] - rmom[0];
    double rchuv = std::max(r1[1], r1[2]) - rmomuv;
    if (rchp < 0.0 || rchuv < 0.0) return;
    if (rchp <= F4(1.0e-8) && rchuv <= F4(1.0e-7)) { F->iconv = 2; return; }
    for (int n = 0; n < 3; n++) r1[n] = rmom[n];
}

static int eddyvis_tail(cms2d_flow *F) {
    if (F->serial_order) {
        int res = 0;
        auto fix = [&](int first, int from, int to) {
            LAUNCH(k_vis_fix_serial, F->nblk_act, F->g, F->s, (const double *)F->ser_face[from], F->ser_face[to],
                   first, F->s.flags + 3);
        };
        if (int rc = serial_settle(F, fix, res)) return rc;
        LAUNCH(k_vis_fix_serial_commit, F->nblk_pad, F->g, F->s, (const double *)F->ser_face[res]);
        auto corner = [&](int first, int from, int to) {
            LAUNCH(k_vis_corner_serial, F->nblk_act, F->g, F->s, (const double *)F->ser_corner[from],
                   F->ser_corner[to], first, F->s.flags + 3);
        };
        if (int rc = serial_settle(F, corner, res)) return rc;
        HIPCHK(hipMemcpyAsync(F->s.visc, F->ser_corner[res], (size_t)F->npad * sizeof(double),
                              hipMemcpyDeviceToDevice, F->st));
    } else {
        LAUNCH(k_vis_fix, F->nblk_pad, F->g, F->s);
        LAUNCH(k_vis_corner, F->nblk_act, F->g, F->s);
    }
    LAUNCH(k_visk, F->nblk_act, F->g, F->s);
    return 0;
}

// flow_recalculate, flow_imp.F90:1133-1242: back to the previous time level
// with half the time step (the ramp keeps the value timing gave it).
// bndriver_vel (boundaries.F90:4259-4356) on the flux side, if any
static void river_vel(cms2d_flow *F) {
    if (F->qside < 0) return;
    const int n = (F->qside == KN || F->qside == KS) ? F->g.nx : F->g.ny;
    hipLaunchKernelGGL(k_river, dim3(1), dim3(1024), (size_t)n * sizeof(double), F->st, F->g, F->s, F->qside,
                       F->qflux, F->cmvel, F->qcos, F->qsin, (const double *)(F->q_rating ? F->q_dev : nullptr));
}

static int flow_recalculate(cms2d_flow *F) {
    F->mtime = 1;
    F->jtime += 1;
    F->timesecs = std::max(0.0, F->timesecs - F->deltime);
    F->deltime = F->dtimebeg / std::ldexp(1.0, F->jtime);
    F->timesecs += F->deltime;
    F->P.dtime = F->deltime;
    if (F->P.dtime <= F->dtimebeg / 256.0)
        return set_err(CMS2D_ERR_DIVERGED, "Model is divergent (time step reduced below 1/256 of its start value)");
    LAUNCH(k_recalc, F->nblk_pad, F->g, F->s);
    if (int rc = wetdry(F)) return rc;
    if (F->qside >= 0) LAUNCH(k_bndflux, F->nblk_act, F->g, F->s);   // bndflux
    LAUNCH(k_grad_fric_vis, F->nblk_act, F->g, F->s, F->P);
    return eddyvis_tail(F);
}

static int flow_imp_once(cms2d_flow *F);

// Recording for the reverse sweep (adjoint.hip): the model state at the start
// of outer iteration `slot` (or at the loop's exit), and the intermediate
// values of the current outer iteration after stage `what`.
enum { TAPE_MOM, TAPE_UV, TAPE_PPCOEF, TAPE_PPSRC, TAPE_PPSOL, TAPE_PVFC1 };
static int tape_snap(cms2d_flow *F, int slot);
static int tape_mid(cms2d_flow *F, int what);
#define TAPE(what)                                              \
    do {                                                        \
        if (F->tape)                                            \
            if (int rc_ = tape_mid(F, what)) return rc_;        \
    } while (0)

// flow_imp, flow_imp.F90:29-312; with dtvar a divergent outer iteration
// recalculates the step with half the time step (:157-161, 262-272)
static int flow_imp(cms2d_flow *F) {
    for (;;) {
        const int rc = flow_imp_once(F);
        if (rc != CMS2D_ERR_DIVERGED || !F->dtvar) return rc;
        if (int r2 = flow_recalculate(F)) return r2;
    }
}

static int flow_imp_once(cms2d_flow *F) {
    Geo &g = F->g;
    St &s = F->s;
    const double timehrs = F->timesecs / 3600.0;
    const double ramp = F->ramp;
    double w = 0.0;
    for (int k = 0; k < F->ntc; k++) w += F->amp[k] * std::cos(F->speed[k] * timehrs - F->phase[k]);
    // bnd_eval (boundaries.F90:4098-4150): the harmonic sum, the curve's level;
    // ramp and gradient terms are applied per cell in k_wse_adjust
    double wsecurve = 0.0;
    if (!F->wse_t.empty()) wsecurve = wse_curve_eval(F->wse_t, F->wse_v, F->wse_order, timehrs, F->wse_inc);
    // the flux string's inflow (:4078-4094): constant or the curve, times the ramp
    if (F->qside >= 0 && F->q_rating) {
        // the stage is the string's mean depth at the step's start: evaluated on the device
        LAUNCH(k_river_rating, 1, g, s, F->qside, (int)F->q_t.size(), F->ramp, F->q_dev);
    } else if (F->qside >= 0) {
        const double q = F->q_t.empty() ? F->q_const : wse_curve_eval(F->q_t, F->q_v, F->q_order, timehrs, F->q_inc);
        F->qflux = F->ramp * q;
    }
    if (int rc = wetdry(F)) return rc;
    F->P.tauwx = F->P.tauwy = 0.0;
    if (F->windconst) {
        // windcurve_eval (meteorological.F90:1020-1027), called each step in flow_imp (:85)
        const double w = std::sqrt(F->wndx * F->wndx + F->wndy * F->wndy);
        const double val = F->ramp * wind_drag_hsu(w) * F4(0.00120) * w;
        F->P.tauwx = val * F->wndx;
        F->P.tauwy = val * F->wndy;
    }
    LAUNCH(k_srcsnk, F->nblk_act, g, s, F->P);
    StencilOp A{g, s.aN, s.aE, s.aS, s.aW};
    for (int n = 0; n < 3; n++) F->kry[n] = 0;
    for (F->niter = 1; F->niter <= F->maxit; F->niter++) {
        if (F->tape)
            if (int rc = tape_snap(F, F->niter - 1)) return rc;
        river_vel(F);
        if (F->P.ndsch >= 5 && F->P.ndsch <= 9) LAUNCH(k_defcor, F->nblk_act, g, s, F->P);
        LAUNCH(k_mom, F->nblk_act, g, s, F->P);
        TAPE(TAPE_MOM);
        hipLaunchKernelGGL(k_uv_prologue, dim3(F->nblk_kry), dim3(KBLOCK), 0, F->st, g, s, F->Kuv);
        {
            KryCtx X = flow_ctx(F, false);
            if (F->precision == CMS2D_PREC_FP32_REFINE) {
                if (int rc = uv_solve_refine(F, X, A)) return rc;
            } else if (int rc = run_bicg(X, F->Kuv, A, 1)) {
                return rc;
            }
        }
        TAPE(TAPE_UV);
        LAUNCH(k_pp_prep, F->nblk_act, g, s);
        LAUNCH(k_pp_coef_flux, F->nblk_act, g, s);
        TAPE(TAPE_PPCOEF);
        for (int sd = 0; sd < 4; sd++) {
            if (g.side[sd] != 1) continue;
            int n = (sd == KE || sd == KW) ? g.ny : g.nx;
            const bool curve = F->side_type[sd] == CMS2D_BND_WSE_CURVE;
            hipLaunchKernelGGL(k_wse_adjust, dim3(1), dim3(1024), n * sizeof(double), F->st, g, s, F->P, sd,
                               curve ? wsecurve : w, ramp, curve ? F->wse_gx : 0.0, curve ? F->wse_gy : 0.0,
                               (int)(curve && F->wse_adjust_off));
        }
        hipLaunchKernelGGL(k_pp_src, dim3(F->nblk_kry), dim3(KBLOCK), 0, F->st, g, s, F->P, F->Kpp);
        TAPE(TAPE_PPSRC);
        {
            KryCtx X = flow_ctx(F, true);
            if (F->precision == CMS2D_PREC_FP32_REFINE) {
                mg_setup(F);
                if (int rc = pp_solve_refine(F, X, A)) return rc;
            } else if (F->solver == CMS2D_SOLVER_GMRES) {
                mg_setup(F);
                if (int rc = pp_solve_gmres(F, X, A)) return rc;
            } else if (F->precond == CMS2D_PRECOND_MG) {
                mg_setup(F);
                F->tm_active = X.tm;
                int rc = run_bicg<StencilOp, 1, 1, double>(X, F->Kpp, A, 0, PpPre{F});
                F->tm_active = nullptr;
                if (rc) return rc;
            } else {
                if (int rc = run_bicg(X, F->Kpp, A, 0)) return rc;
            }
        }
        TAPE(TAPE_PPSOL);
        LAUNCH(k_pvfc1, F->nblk_act, g, s);
        TAPE(TAPE_PVFC1);
        LAUNCH(k_pvfc2, F->nblk_pad, g, s, F->P);
        LAUNCH(k_ghost_p, (2 * g.nx + 2 * g.ny + BLOCK - 1) / BLOCK, g, s, F->P);
        LAUNCH(k_grad_fric_vis, F->nblk_act, g, s, F->P);
        // bndriver_vel and bndflux after flow_grad_interp (flow_imp.F90:236-237): the
        // weights take the new hk; bndflux then writes the fluxes bndriver_vel wrote
        river_vel(F);
        if (int rc = eddyvis_tail(F)) return rc;
        for (int n = 0; n < 3; n++) {
            F->rmom[n] = F->h_sc[n].rmom;
            F->kry[n] += F->h_sc[n].final_iters;
        }
        check_conv(F);
        if (F->iconv == 0) return set_err(CMS2D_ERR_DIVERGED, "Model is divergent (check_conv iconv=0)");
        if (F->iconv == 2) break;
    }
    if (F->niter > F->maxit) F->niter = F->maxit;
    if (F->tape)
        if (int rc = tape_snap(F, F->niter)) return rc;
    LAUNCH(k_velface, F->nblk_act, g, s, F->P);
    return 0;
}

static void fill_geo(Geo &g, int nx, int ny, double dx, double dy, const int *side) {
    g.nx = nx; g.ny = ny; g.nxp = nx + 2; g.nyp = ny + 2; g.nact = nx * ny;
    g.dx = dx; g.dy = dy; g.area = dx * dy;
    for (int k = 0; k < 4; k++) {
        bool ns = (k == KN || k == KS);
        g.ds[k] = ns ? dx : dy;                       // min(dx,dx_nb) on a uniform grid
        g.dn[k] = ns ? dy : dx;                       // |y_nb - y|, |x_nb - x|
        g.dsxy[k] = g.ds[k] / g.dn[k];
        g.fnx[k] = k == KE ? 1.0 : k == KW ? -1.0 : 0.0;
        g.fny[k] = k == KN ? 1.0 : k == KS ? -1.0 : 0.0;
        g.fintp[k] = ns ? dy / (dy + dy) : dx / (dx + dx);
        g.rx[k] = k == KE ? dx / 2.0 : k == KW ? -dx / 2.0 : 0.0;
        g.ry[k] = k == KN ? dy / 2.0 : k == KS ? -dy / 2.0 : 0.0;
        g.dsx[k] = g.fnx[k] * g.ds[k];
        g.dsy[k] = g.fny[k] * g.ds[k];
        g.side[k] = side[k];
    }
}

static int upload_padded(cms2d_flow *F, double *dst, const double *src, bool ghosts_copy) {
    const Geo &g = F->g;
    std::vector<double> h((size_t)g.nxp * g.nyp, 0.0);
    for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) h[(size_t)(j + 1) * g.nxp + (i + 1)] = src ? src[(size_t)j * g.nx + i] : 0.0;
    if (ghosts_copy) {
        for (int i = 0; i < g.nx; i++) {
            h[(size_t)0 * g.nxp + i + 1] = h[(size_t)1 * g.nxp + i + 1];
            h[(size_t)(g.ny + 1) * g.nxp + i + 1] = h[(size_t)g.ny * g.nxp + i + 1];
        }
        for (int j = 0; j < g.ny; j++) {
            h[(size_t)(j + 1) * g.nxp + 0] = h[(size_t)(j + 1) * g.nxp + 1];
            h[(size_t)(j + 1) * g.nxp + g.nx + 1] = h[(size_t)(j + 1) * g.nxp + g.nx];
        }
    }
    HIPCHK(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    return 0;
}

static void tel_destroy(cms2d_flow *F);
static int tel_step(cms2d_flow *F, int nsteps, cms2d_step_info *info);
static int tel_get(cms2d_flow *F, const char *name, double *out);
static int tel_get_face(cms2d_flow *F, const char *name, double *out);
static int tel_set(cms2d_flow *F, const char *name, const double *in);

extern "C" int cms2d_flow_destroy(cms2d_flow *F) {
    if (!F) return CMS2D_OK;
    (void)hipSetDevice(F->dev);
    if (F->st) (void)hipStreamSynchronize(F->st);
    tel_destroy(F);
    for (auto &ge : F->kgraph)
        if (ge) (void)hipGraphExecDestroy(ge);
    for (void *p : F->allocs) (void)hipFree(p);
    if (F->h_sc) (void)hipHostFree(F->h_sc);
    if (F->h_flags) (void)hipHostFree(F->h_flags);
    if (F->h_gm) (void)hipHostFree(F->h_gm);
    for (auto e : F->tm.pool) (void)hipEventDestroy(e);
    for (auto e : F->step_ev) if (e) (void)hipEventDestroy(e);
    if (F->st) (void)hipStreamDestroy(F->st);
    delete F;
    return CMS2D_OK;
}

// Initial state from u, v, p as flow_init and prestart build it
// (flow.F90:688-707, prestart.f90:309-329): wet/dry, previous time level,
// gradients and face values, fluxes (flow_convflux), friction, eddy viscosity.
static int init_state(cms2d_flow *F) {
    const Geo &g = F->g;
    St &s = F->s;
    const size_t N = F->npad;
    if (int rc = wetdry(F)) return rc;
    HIPCHK(hipMemcpyAsync(s.u1, s.u, N * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    HIPCHK(hipMemcpyAsync(s.v1, s.v, N * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    HIPCHK(hipMemcpyAsync(s.p1, s.p, N * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    HIPCHK(hipMemcpyAsync(s.h1, s.h, N * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    if (int rc = wetdry(F)) return rc;
    HIPCHK(hipMemcpyAsync(s.iwet1, s.iwet, N, hipMemcpyDeviceToDevice, F->st));
    LAUNCH(k_grad_fric_vis, F->nblk_act, g, s, F->P);
    LAUNCH(k_convflux, F->nblk_act, g, s);
    HIPCHK(hipMemcpyAsync(s.fluxE1, s.fluxE, N * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    HIPCHK(hipMemcpyAsync(s.fluxN1, s.fluxN, N * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    if (int rc = eddyvis_tail(F)) return rc;
    HIPCHK(hipStreamSynchronize(F->st));
    return 0;
}

// Multigrid hierarchy on an nx x ny grid (level 0 in the padded layout of the
// flow arrays): 2x2 coarsening until the coarsest fits one workgroup.
static int mg_alloc(cms2d_flow *F, int nx, int ny) {
    Mg &M = F->mg;
    M.nlev = 0;
    for (;;) {
        MgLevel &L = M.L[M.nlev];
        L.nx = nx; L.ny = ny; L.nxp = nx + 2;
        const size_t n = (size_t)(nx + 2) * (ny + 2);
        float **arrs[] = {&L.aN, &L.aE, &L.aS, &L.aW, &L.dinv, &L.e};
        for (float **a : arrs)
            if (int rc = dalloc(F, a, n)) return rc;
        if (int rc = dalloc(F, &L.z, n)) return rc;
        if (M.nlev > 0) {
            if (int rc = dalloc(F, &L.zo, n)) return rc;
            if (int rc = dalloc(F, &L.b, n)) return rc;
        } else {
            if (int rc = dalloc(F, &L.ap, n)) return rc;
        }
        M.nlev++;
        const bool fits = (long)nx * ny <= MG_COARSE_MAX && (long)(nx + 2) * (ny + 2) <= MG_COARSE_LDS;
        if (fits && (M.nlev >= MG_MIN_LEVELS || (nx == 1 && ny == 1))) break;
        if (M.nlev == MG_MAX_LEVELS) return set_err(CMS2D_ERR_ARG, "grid too large for the multigrid hierarchy");
        nx = (nx + 1) / 2;
        ny = (ny + 1) / 2;
    }
    return 0;
}

static int create_impl(const cms2d_flow_desc *d, cms2d_flow *F) {
    if (d->nx < 3 || d->ny < 3 || !(d->dx > 0) || !(d->dy > 0) || !d->zb || !d->mann || d->ntc < 0 ||
        d->ntc > 8 || !(d->dt > 0) || !(std::fabs(d->latitude) <= 90.0) || d->coriolis_plane < 0 ||
        d->coriolis_plane > 1 || d->turbulence < 0 || d->turbulence > 4 || d->friction < 0 || d->friction > 3)
        return set_err(CMS2D_ERR_ARG, "invalid flow description");
    if ((long)(d->nx + 2) * (d->ny + 2) > (1L << 30)) return set_err(CMS2D_ERR_ARG, "grid too large");
    bool curve = false;
    for (int k = 0; k < 4; k++) {
        if (d->side_type[k] != CMS2D_BND_WALL && d->side_type[k] != CMS2D_BND_WSE_HARMONIC &&
            d->side_type[k] != CMS2D_BND_WSE_CURVE && d->side_type[k] != CMS2D_BND_FLUX)
            return set_err(CMS2D_ERR_ARG, "unknown boundary type");
        curve = curve || d->side_type[k] == CMS2D_BND_WSE_CURVE;
    }
    int nopen = 0, nflux = 0;
    for (int k = 0; k < 4; k++) {
        if (d->side_type[k] == CMS2D_BND_WALL) continue;
        if (d->side_type[k] == CMS2D_BND_FLUX) nflux++;
        else nopen++;
        // the set-up adjustment of a water-level side keeps the side in LDS
        const long len = (k == CMS2D_SIDE_N || k == CMS2D_SIDE_S) ? d->nx : d->ny;
        if (len * (long)sizeof(double) > 65536)
            return set_err(CMS2D_ERR_ARG, "water-level side longer than 8192 cells");
    }
    if (nopen > 1)
        return set_err(CMS2D_ERR_ARG, "one water-level side at most (the others are walls)");
    if (nflux > 1) return set_err(CMS2D_ERR_ARG, "one flux side at most");
    if (nflux) {
        // bndriver_vel reads the Manning n of the previous friction evaluation; the laws
        // that rewrite n each evaluation are not built with a flux side
        if (d->friction != CMS2D_FRIC_MANNING) return set_err(CMS2D_ERR_ARG, "a flux side needs Manning friction");
        if (d->flux_ntimes == 1 || d->flux_ntimes < 0 || d->flux_order < 0 || d->flux_order > 3 ||
            !(d->flux_conveyance >= 0.0) || !std::isfinite(d->flux_conveyance))
            return set_err(CMS2D_ERR_ARG, "flux curve or conveyance invalid");
        if (d->flux_ntimes == 0 && !std::isfinite(d->flux_q)) return set_err(CMS2D_ERR_ARG, "flux not finite");
        if (d->flux_angle_on && !std::isfinite(d->flux_angle)) return set_err(CMS2D_ERR_ARG, "inflow angle not finite");
        if (d->flux_rating && d->flux_ntimes < 2) return set_err(CMS2D_ERR_ARG, "a rating curve needs 2 or more stages");
        if (d->flux_ntimes > 0) {
            if (!d->flux_times || !d->flux_values) return set_err(CMS2D_ERR_ARG, "flux curve missing");
            for (int i = 0; i < d->flux_ntimes; i++)
                if (!std::isfinite(d->flux_times[i]) || !std::isfinite(d->flux_values[i]) ||
                    (i > 0 && !(d->flux_times[i] > d->flux_times[i - 1])))
                    return set_err(CMS2D_ERR_ARG, "flux curve times must increase, values be finite");
        }
    }
    if (curve) {
        if (d->wse_ntimes < 1 || !d->wse_times || !d->wse_values || d->wse_order < 0 || d->wse_order > 3)
            return set_err(CMS2D_ERR_ARG, "water-level curve missing or invalid");
        if (!std::isfinite(d->wse_gradx) || !std::isfinite(d->wse_grady))
            return set_err(CMS2D_ERR_ARG, "water-level gradients must be finite");
        for (int i = 0; i < d->wse_ntimes; i++)
            if (!std::isfinite(d->wse_times[i]) || !std::isfinite(d->wse_values[i]) ||
                (i > 0 && !(d->wse_times[i] > d->wse_times[i - 1])))
                return set_err(CMS2D_ERR_ARG, "water-level curve times must increase, values be finite");
    }
    if (d->precision != CMS2D_PREC_FP64 && d->precision != CMS2D_PREC_FP32_REFINE)
        return set_err(CMS2D_ERR_ARG, "unknown precision form");
    if (d->precond != CMS2D_PRECOND_MG && d->precond != CMS2D_PRECOND_JACOBI)
        return set_err(CMS2D_ERR_ARG, "unknown preconditioner");
    if (d->precision == CMS2D_PREC_FP32_REFINE && d->precond != CMS2D_PRECOND_MG)
        return set_err(CMS2D_ERR_ARG, "the FP32 refinement form uses the multigrid preconditioner");
    if (d->solver != CMS2D_SOLVER_BICGSTAB && d->solver != CMS2D_SOLVER_GMRES)
        return set_err(CMS2D_ERR_ARG, "unknown solver");
    if (d->solver == CMS2D_SOLVER_GMRES && (d->precond != CMS2D_PRECOND_MG || d->precision != CMS2D_PREC_FP64))
        return set_err(CMS2D_ERR_ARG, "GMRES is built for the FP64 form with the multigrid preconditioner");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= d->device)
        return set_err(CMS2D_ERR_NODEVICE, "no HIP device");
    F->dev = d->device;
    HIPCHK(hipSetDevice(F->dev));
    HIPCHK(hipStreamCreateWithFlags(&F->st, hipStreamNonBlocking));
    {
        int open[4];
        for (int k = 0; k < 4; k++) {
            F->side_type[k] = d->side_type[k];
            open[k] = d->side_type[k] == CMS2D_BND_WALL ? 0 : d->side_type[k] == CMS2D_BND_FLUX ? 2 : 1;
            if (d->side_type[k] == CMS2D_BND_FLUX) {
                // Q_str: inflow along the inward normal, local_angle = (4 - idirface) 90
                // degrees (boundaries.F90:1750-1762); conveyance 0.667 (a default-real
                // literal, bnd_block.f90:201) or the card's value limited to [0.5, 1]
                F->qside = k;
                F->q_const = d->flux_q;
                if (d->flux_ntimes > 0) {
                    F->q_t.assign(d->flux_times, d->flux_times + d->flux_ntimes);
                    F->q_v.assign(d->flux_values, d->flux_values + d->flux_ntimes);
                }
                F->q_order = std::max(std::min(d->flux_order ? d->flux_order : 1, std::max(d->flux_ntimes - 1, 1)), 1);
                F->q_rating = d->flux_rating != 0;
                F->q_inc = 1;
                F->cmvel = d->flux_conveyance > 0.0 ? std::max(std::min(d->flux_conveyance, 1.0), 0.5) : F4(0.667);
                const double pi = std::acos(-1.0), deg2rad = pi / 180.0;
                double angle = (double)((float)(4 - (k + 1)) * 90.0f);
                if (d->flux_angle_on) {
                    // INFLOW_ANGLE (boundaries.F90:362-370): 90 - angle - grid azimuth (0),
                    // within [0, 360]
                    angle = 90.0 - d->flux_angle - 0.0;
                    if (angle < 0.0) angle += 360.0;
                    else if (angle > 360.0) angle -= 360.0;
                }
                F->qcos = std::cos(angle * deg2rad);
                F->qsin = std::sin(angle * deg2rad);
            }
        }
        fill_geo(F->g, d->nx, d->ny, d->dx, d->dy, open);
        bool curve = false;
        for (int k = 0; k < 4; k++) curve = curve || d->side_type[k] == CMS2D_BND_WSE_CURVE;
        if (curve) {
            F->wse_t.assign(d->wse_times, d->wse_times + d->wse_ntimes);
            F->wse_v.assign(d->wse_values, d->wse_values + d->wse_ntimes);
            // H_str%nti = 2 by default, limited to 1..ntimes-1 (boundaries.F90:1417,1869)
            F->wse_order = std::max(std::min(d->wse_order ? d->wse_order : 2, d->wse_ntimes - 1), 1);
            F->wse_inc = 1;
            F->wse_adjust_off = d->wse_adjust_off != 0;
            F->wse_gx = d->wse_gradx;
            F->wse_gy = d->wse_grady;
        }
    }
    Geo &g = F->g;
    F->npad = g.nxp * g.nyp;
    F->nblk_act = (g.nact + BLOCK - 1) / BLOCK;
    F->nblk_pad = (F->npad + BLOCK - 1) / BLOCK;
    F->nblk_kry = std::min(KRY_BLOCKS, F->nblk_act);
    F->use_graphs = true;
    if (const char *e = getenv("CMS2D_GRAPHS")) F->use_graphs = atoi(e) != 0;   // 0: direct launches, 1: graphs
    Par &P = F->P;
    P.hdry = d->hdry > 0 ? d->hdry : F4(0.05);
    P.hmin = P.hdry;                                  // flow.F90:690-694
    P.hdry1 = F4(5.0) * P.hdry;
    P.hdry2 = F4(2.0) * P.hdry;
    P.dtime = d->dt;
    double asum = 0.0;                                // areaavg, geospatial.F90:1441
    for (int c = 0; c < g.nact; c++) asum += g.area;
    double areaavg = asum / (double)(float)g.nact;
    P.c2a = F4(0.2) * F4(0.2) * areaavg;              // cvishor2areaavg, flow.F90:749
    P.rtol = d->rtol > 0 ? d->rtol : 1e-10;
    P.atol = d->atol >= 0 ? d->atol : 1e-12;
    P.kmax = d->kmax > 0 ? d->kmax : 4000;
    P.noptset = d->noptset;
    P.wallfric = d->wallfric;
    P.ndsch = d->ndsch;
    {
        static const int turb[5] = {5, 0, 1, 2, 3};   // CMS2D_TURB_* -> mturbul
        P.mturbul = turb[d->turbulence];
        static const int fric[4] = {2, 1, 3, 4};      // CMS2D_FRIC_* -> mbedfric
        P.mbedfric = fric[d->friction];
    }
    P.lmax = std::max(g.nx, g.ny);
    F->maxit = d->maxit > 0 ? d->maxit : 20 + g.nact / 10000;   // flow.F90:784
    F->rampdur = std::max(d->rampdur, d->dt / 3600.0);
    F->dtvar = d->dtvar != 0;
    F->dtimebeg = F->deltime = d->dt;
    F->ntc = d->ntc;
    for (int k = 0; k < d->ntc; k++) { F->amp[k] = d->amp[k]; F->speed[k] = d->speed[k]; F->phase[k] = d->phase[k]; }
    for (int n = 0; n < 3; n++) { F->rmom0[n] = 1e10; F->rmom[n] = 0.0; }

    St &s = F->s;
    const size_t N = F->npad;
    double **dfields[] = {&s.u, &s.v, &s.p, &s.h, &s.pp, &s.zb, &s.vis, &s.visb, &s.visc, &s.u1, &s.v1, &s.p1,
                          &s.h1, &s.eta, &s.dppx, &s.dppy, &s.dux, &s.duy, &s.dvx, &s.dvy, &s.dpx, &s.dpy,
                          &s.coefman, &s.cfrict, &s.z0, &s.uv, &s.uelwc, &s.bsvel, &s.su, &s.sv, &s.sp, &s.spu,
                          &s.spv, &s.ssu0, &s.ssv0, &s.spuv0, &s.sppp0, &s.sspp0, &s.apuareap, &s.sumu, &s.Hu,
                          &s.Hv, &s.resu, &s.resv, &s.resp, &s.fluxE, &s.fluxN, &s.hkE, &s.hkN, &s.viskE,
                          &s.viskN, &s.zbkE, &s.zbkN, &s.aN, &s.aE, &s.aS, &s.aW, &s.wavestrx, &s.wavestry,
                          &s.tu, &s.tv, &s.tvis, &s.fluxE1, &s.fluxN1, &s.fc, &s.cbf, &s.maxeta};
    for (double **f : dfields) {
        if (int rc = dalloc(F, f, N)) return rc;
        F->state.push_back({*f, N * sizeof(double)});
    }
    if (int rc = dalloc(F, &s.wseadj, 4 * (size_t)P.lmax)) return rc;
    F->state.push_back({s.wseadj, 4 * (size_t)P.lmax * sizeof(double)});
    uint8_t **bfields[] = {&s.iwet, &s.iwet1, &s.icorner, &s.ipond};
    for (uint8_t **f : bfields) {
        if (int rc = dalloc(F, f, N)) return rc;
        F->state.push_back({*f, N});
    }
    if (int rc = dalloc(F, &s.flags, 4)) return rc;
    if (int rc = dalloc(F, &s.plab, N)) return rc;
    F->state.push_back({s.plab, N * sizeof(int)});
    F->serial_order = d->serial_order != 0;
    if (F->serial_order) {
        for (int b = 0; b < 2; b++) {
            for (int q = 0; q < 3; q++)
                if (int rc = dalloc(F, &F->ser_d2w[b][q], N)) return rc;
            if (int rc = dalloc(F, &F->ser_face[b], 4 * N)) return rc;
            if (int rc = dalloc(F, &F->ser_corner[b], N)) return rc;
        }
    }
    {
        // Krylov partial sums: 5 slots per system (pr: 2, pv: 1, pt: 2)
        double *red;
        if (int rc = dalloc(F, &red, 4 * 5 * (size_t)F->nblk_kry)) return rc;
        const size_t nb = F->nblk_kry;
        F->Kpp.red.pr[0] = red; F->Kpp.red.pv[0] = red + 2 * nb; F->Kpp.red.pt[0] = red + 3 * nb;
        for (int q = 0; q < 2; q++) {
            double *b = red + (q + 1) * 5 * nb;
            F->Kuv.red.pr[q] = b; F->Kuv.red.pv[q] = b + 2 * nb; F->Kuv.red.pt[q] = b + 3 * nb;
        }
        double *bi = red + 3 * 5 * nb;
        F->Kin.red.pr[0] = bi; F->Kin.red.pv[0] = bi + 2 * nb; F->Kin.red.pt[0] = bi + 3 * nb;
    }
    if (int rc = dalloc(F, &F->d_sc, 6)) return rc;
    if (F->q_rating) {
        // the rating table beside the step's inflow
        const int m = (int)F->q_t.size();
        if (int rc = dalloc(F, &F->q_dev, 1 + 2 * (size_t)m)) return rc;
        std::vector<double> tab(1 + 2 * (size_t)m, 0.0);
        for (int i = 0; i < m; i++) { tab[1 + i] = F->q_t[i]; tab[1 + m + i] = F->q_v[i]; }
        HIPCHK(hipMemcpyAsync(F->q_dev, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
    }
    if (int rc = dalloc(F, &F->bal_part, BAL_PART * (size_t)F->nblk_kry)) return rc;
    if (int rc = dalloc(F, &F->bal, 1)) return rc;
    F->state.push_back({F->bal, sizeof(BalAcc)});
    // Krylov vectors
    {
        double *kv[3][7];
        for (int q = 0; q < 3; q++)
            for (int m = 0; m < 7; m++)
                if (int rc = dalloc(F, &kv[q][m], N)) return rc;
        F->Kpp.x[0] = s.pp;
        F->Kpp.r[0] = kv[0][0]; F->Kpp.r0[0] = kv[0][1]; F->Kpp.p[0] = kv[0][2];
        F->Kpp.v[0] = kv[0][3]; F->Kpp.t[0] = kv[0][4]; F->Kpp.ap_w[0] = kv[0][5];
        F->Kpp.dinv_w[0] = kv[0][6];
        F->Kpp.sc = F->d_sc;
        for (int q = 0; q < 2; q++) {
            F->Kuv.x[q] = q ? s.v : s.u;
            F->Kuv.r[q] = kv[q + 1][0]; F->Kuv.r0[q] = kv[q + 1][1]; F->Kuv.p[q] = kv[q + 1][2];
            F->Kuv.v[q] = kv[q + 1][3]; F->Kuv.t[q] = kv[q + 1][4]; F->Kuv.ap_w[q] = kv[q + 1][5];
            F->Kuv.dinv_w[q] = kv[q + 1][6];
        }
        F->Kuv.sc = F->d_sc + 1;
    }
    // multigrid hierarchy: level 0 shares the padded layout of the flow arrays
    F->precond = d->precond;
    F->precision = d->precision;
    F->solver = d->solver;
    {
        Mg &M = F->mg;
        if (int rc = mg_alloc(F, g.nx, g.ny)) return rc;
        if (int rc = dalloc(F, &F->Kpp.zp, N)) return rc;
        if (int rc = dalloc(F, &F->Kpp.zs, N)) return rc;
        KrySys<1, float> &Ki = F->Kin;
        if (F->precision == CMS2D_PREC_FP32_REFINE) {
            float **vec[] = {&Ki.x[0], &Ki.r[0], &Ki.r0[0], &Ki.p[0], &Ki.v[0], &Ki.t[0]};
            for (float **v : vec)
                if (int rc = dalloc(F, v, N)) return rc;
        }
        if (F->solver == CMS2D_SOLVER_GMRES) {
            GmSys &G = F->gm;
            for (int i = 0; i <= GM_M; i++)
                if (int rc = dalloc(F, &G.V[i], N)) return rc;
            if (int rc = dalloc(F, &G.z, N)) return rc;
            if (int rc = dalloc(F, &G.part, (size_t)(GM_M + 2) * F->nblk_kry)) return rc;
            if (int rc = dalloc(F, &G.sc, 1)) return rc;
            G.x = s.pp;
            G.r = F->Kpp.r[0];
            G.ap = F->Kpp.ap_w[0];
            HIPCHK(hipHostMalloc((void **)&F->h_gm, sizeof(GmScal), hipHostMallocDefault));
            memset(F->h_gm, 0, sizeof(GmScal));
        }
        Ki.ap_w[0] = M.L[0].ap;
        Ki.zp = F->Kpp.zp;
        Ki.zs = F->Kpp.zs;
        Ki.sc = F->d_sc + 3;
        if (F->precision == CMS2D_PREC_FP32_REFINE) {
            KrySys<2, float> &K2 = F->Kin2;
            for (int q = 0; q < 2; q++) {
                float **vec[] = {&K2.x[q], &K2.r[q], &K2.r0[q], &K2.p[q], &K2.v[q], &K2.t[q], &K2.ap_w[q], &K2.dinv_w[q]};
                for (float **v : vec)
                    if (int rc = dalloc(F, v, N)) return rc;
            }
            for (int k = 0; k < 4; k++)
                if (int rc = dalloc(F, &F->a32[k], N)) return rc;
            double *ri;
            const size_t nbk = F->nblk_kry;
            if (int rc = dalloc(F, &ri, 2 * 5 * nbk)) return rc;
            for (int q = 0; q < 2; q++) {
                K2.red.pr[q] = ri + q * 5 * nbk;
                K2.red.pv[q] = K2.red.pr[q] + 2 * nbk;
                K2.red.pt[q] = K2.red.pr[q] + 3 * nbk;
            }
            K2.sc = F->d_sc + 4;
        }
    }
    HIPCHK(hipHostMalloc((void **)&F->h_sc, 6 * sizeof(KryScal), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&F->h_flags, 4 * sizeof(int), hipHostMallocDefault));
    memset(F->h_sc, 0, 6 * sizeof(KryScal));
    F->tm.pool.resize(2 * 2048);
    for (auto &e : F->tm.pool) HIPCHK(hipEventCreate(&e));
    for (auto &e : F->step_ev) HIPCHK(hipEventCreate(&e));
    // inputs
    if (int rc = upload_padded(F, s.zb, d->zb, true)) return rc;
    {
        // fric_init (friction.F90:315-319,340-390): z0 = 1e-5 until the first
        // evaluation, the law's input array from the dataset
        std::vector<double> z0(N, (double)1.0e-5f);
        HIPCHK(hipMemcpyAsync(s.z0, z0.data(), N * sizeof(double), hipMemcpyHostToDevice, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
        double *in = P.mbedfric == 1 ? s.cfrict : P.mbedfric == 3 ? s.z0 : P.mbedfric == 4 ? s.cbf : s.coefman;
        // fric_init's copy to the dummy cells (friction.F90:397-407) loops
        // `do i=1,ncells+1,ncellsD`, one pass for i = 1: the first cell takes
        // cfrict, coefman and z0 of its first active neighbour in face order
        // (north); the other arrays hold their defaults there
        std::vector<double> fr(d->mann, d->mann + g.nact);
        if (P.mbedfric != 4) fr[0] = fr[g.nx];
        if (int rc = upload_padded(F, in, fr.data(), false)) return rc;
    }
    if (d->wavestrx && int(upload_padded(F, s.wavestrx, d->wavestrx, false))) return CMS2D_ERR_HIP;
    if (d->wavestry && int(upload_padded(F, s.wavestry, d->wavestry, false))) return CMS2D_ERR_HIP;
    {
        // Coriolis, flow.F90:720-736: f-plane 2 omega sin(lat); beta-plane adds
        // 2 omega cos(lat)/R (y - mean y), y the cell centres from the south
        // edge; omega = 2 pi/86400, R = 6371 km (definitions.F90:86-96)
        const double pi = std::acos(-1.0), omega = 2.0 * pi / 86400.0, deg2rad = pi / 180.0;
        const double f0 = 2.0 * omega * std::sin(d->latitude * deg2rad);
        std::vector<double> fc((size_t)g.nx * g.ny, f0);
        if (d->coriolis_plane == 1) {
            const double beta = 2.0 * omega * std::cos(d->latitude * deg2rad) / 6371000.0;
            double ysum = 0.0;
            for (int j = 0; j < g.ny; j++)
                for (int i = 0; i < g.nx; i++) ysum += (j + 1) * g.dy - 0.5 * g.dy;
            const double yavg = ysum / (double)((long)g.nx * g.ny);
            for (int j = 0; j < g.ny; j++)
                for (int i = 0; i < g.nx; i++) fc[(size_t)j * g.nx + i] = f0 + beta * (((j + 1) * g.dy - 0.5 * g.dy) - yavg);
        }
        if (int rc = upload_padded(F, s.fc, fc.data(), false)) return rc;
    }
    LAUNCH(k_zbk, F->nblk_act, g, s);
    if (int rc = init_state(F)) return rc;
    if (int rc = tune_mg_tile(F)) return rc;
    return capture_graphs(F);
}

#include "tel.hip"

extern "C" int cms2d_flow_create(const cms2d_flow_desc *d, cms2d_flow **out) {
    if (!d || !out) return set_err(CMS2D_ERR_ARG, "null argument");
    *out = nullptr;
    cms2d_flow *F = new (std::nothrow) cms2d_flow();
    if (!F) return set_err(CMS2D_ERR_ARG, "out of host memory");
    int rc = create_impl(d, F);
    if (rc != 0) {
        std::string keep = g_err;
        cms2d_flow_destroy(F);
        g_err = keep;
        return rc;
    }
    *out = F;
    return CMS2D_OK;
}

extern "C" int cms2d_flow_create_tel(const cms2d_flow_desc *d, const cms2d_tel_grid *grid, cms2d_flow **out) {
    if (!d || !grid || !out) return set_err(CMS2D_ERR_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (d->device < 0) return set_err(CMS2D_ERR_ARG, "invalid device");
    cms2d_flow *F = new (std::nothrow) cms2d_flow();
    if (!F) return set_err(CMS2D_ERR_ARG, "out of host memory");
    int rc = tel_check(d, grid);
    if (rc == 0 && (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= d->device))
        rc = set_err(CMS2D_ERR_NODEVICE, "no HIP device");
    if (rc == 0) rc = tel_create(d, grid, F);
    if (rc != 0) {
        std::string keep = g_err;
        cms2d_flow_destroy(F);
        g_err = keep;
        return rc;
    }
    *out = F;
    return CMS2D_OK;
}

// One iteration of the CMS_Flow time loop: timing, CMS_Flow.f90:109-240 (with
// dtvar the time step is halved after a step that ended with iconv = 3 and
// doubled back once the pp residual is small and the time falls on the
// coarser step, :180-205), flow_imp, flow_update.
static int step_once(cms2d_flow *F) {
    F->ntime++;
    if (F->dtvar) {
        if (F->iconv == 3 && F->dtimebeg / F->deltime <= 10.0) {
            F->mtime = 0;
            F->jtime += 1;
            F->deltime = F->dtimebeg / std::ldexp(1.0, F->jtime);
        } else if (F->jtime >= 1) {
            if ((F->mtime >= 2 && F->rmom[0] < F4(1.0e-4) / 100.0) ||
                (F->mtime >= 4 && F->rmom[0] < F4(1.0e-4) / 10.0)) {
                const double dtt = F->dtimebeg / std::ldexp(1.0, F->jtime - 1);
                const double rtime = std::fmod(F->timesecs + dtt, dtt);
                const double err = std::fabs(rtime) / dtt;
                if (rtime < F4(0.01) || err < F4(0.01)) {
                    F->deltime = dtt;
                    F->mtime = 0;
                    F->jtime -= 1;
                }
            }
        }
        F->mtime += 1;
        F->P.dtime = F->deltime;
    }
    F->timesecs += F->deltime;
    F->ramp = ramp_func(F->timesecs / 3600.0, F->rampdur);
    if (int rc = flow_imp(F)) return rc;
    hipLaunchKernelGGL(k_bal_part, dim3(F->nblk_kry), dim3(KBLOCK), 0, F->st, F->g, F->s, F->P, F->bal_part);
    hipLaunchKernelGGL(k_bal_fin, dim3(1), dim3(KBLOCK), 0, F->st, F->g, F->s, F->P, (const double *)F->bal_part,
                       F->nblk_kry, F->bal);
    // stat_update (CMS_Flow.f90:78-81, before flow_update): from the window's start,
    // the window closing every stat_t2 - stat_t1 after it (statistics.F90:503,632)
    if (F->stat_on && F->timesecs + F4(1.0e-4) >= F->stat_t1) {
        F->stat_n += 1;
        const double th = F->timesecs / 3600.0;
        const int close = std::fmod(F->timesecs - F->stat_t1, F->stat_t2 - F->stat_t1) <= F4(1.0e-4);
        LAUNCH(k_set_int, 1, F->stat_flag, 1, 0);
        LAUNCH(k_stat, F->nblk_act, F->g, F->s, F->P, F->stat_acc, (size_t)F->npad, th, F->stat_flag);
        LAUNCH(k_stat_fix, F->nblk_pad, F->g, F->stat_acc, (size_t)F->npad, th, (const int *)F->stat_flag, close);
        if (close) F->stat_n = 0;
    }
    LAUNCH(k_update, F->nblk_pad, F->g, F->s);
    return 0;
}

extern "C" int cms2d_flow_step(cms2d_flow *F, int nsteps, cms2d_step_info *info) {
    if (!F || nsteps < 0) return set_err(CMS2D_ERR_ARG, "invalid argument");
    HIPCHK(hipSetDevice(F->dev));
    if (F->tel) return tel_step(F, nsteps, info);
    HIPCHK(hipEventRecord(F->step_ev[0], F->st));
    for (int n = 0; n < nsteps; n++)
        if (int rc = step_once(F)) return rc;
    HIPCHK(hipEventRecord(F->step_ev[1], F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    if (info) {
        info->niter = F->niter;
        info->iconv = F->iconv;
        for (int n = 0; n < 3; n++) { info->rmom[n] = F->rmom[n]; info->kry_iters[n] = F->kry[n]; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, F->step_ev[0], F->step_ev[1]);
        info->kernel_ms = ms;
        info->dt = F->deltime;
        std::vector<uint8_t> w(F->npad);
        HIPCHK(hipMemcpyAsync(w.data(), F->s.iwet, F->npad, hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
        int nw = 0;
        for (int j = 0; j < F->g.ny; j++)
            for (int i = 0; i < F->g.nx; i++) nw += w[(size_t)(j + 1) * F->g.nxp + i + 1];
        info->nwet = nw;
    }
    return CMS2D_OK;
}

// flow_balance (flow_imp.F90:2260-2372) and flow_step_stat (flow.F90:1912-2008)
// of the last step, from the totals the device keeps (k_bal_part, k_bal_fin).
extern "C" int cms2d_flow_balance(cms2d_flow *F, cms2d_balance *out) {
    if (!F || !out) return set_err(CMS2D_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(F->dev));
    BalAcc a{};
    HIPCHK(hipMemcpyAsync(&a, F->bal, sizeof a, hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    out->storage = a.stg; out->efflux = a.bnd; out->flux = a.flux;
    out->balance = a.bal; out->net = a.net; out->error_pct = a.err;
    out->cum_storage = a.cstg; out->cum_efflux = a.cbnd; out->cum_flux = a.cflux;
    out->cum_balance = a.cbal; out->cum_net = a.cnet; out->cum_error_pct = a.cerr;
    out->wse = a.xv[0] * GRAVINV;   // pxtrm*gravinv (flow.F90:2001-2005)
    out->u = a.xv[1]; out->v = a.xv[2]; out->courant = a.xv[3];
    out->wse_cell = (int)a.xc[0]; out->u_cell = (int)a.xc[1]; out->v_cell = (int)a.xc[2];
    out->courant_cell = (int)a.xc[3];
    out->steps = (int)a.steps;
    return CMS2D_OK;
}

// FLOW_STATISTICS (stat_init_flow, statistics.F90:172-320): accumulators zero.
extern "C" int cms2d_flow_statistics(cms2d_flow *F, double t_start, double t_end) {
    if (!F) return set_err(CMS2D_ERR_ARG, "null argument");
    if (F->tel) return set_err(CMS2D_ERR_ARG, "flow statistics are built for Cartesian grids");
    if (!(t_end > t_start) || !std::isfinite(t_start) || !std::isfinite(t_end))
        return set_err(CMS2D_ERR_ARG, "statistics window must have t_end > t_start");
    HIPCHK(hipSetDevice(F->dev));
    if (!F->stat_acc) {
        if (int rc = dalloc(F, &F->stat_acc, (size_t)NSTAT * F->npad)) return rc;
        if (int rc = dalloc(F, &F->stat_flag, 1)) return rc;
    } else {
        HIPCHK(hipMemsetAsync(F->stat_acc, 0, (size_t)NSTAT * F->npad * sizeof(double), F->st));
    }
    F->stat_on = 1;
    F->stat_n = 0;
    F->stat_t1 = t_start;
    F->stat_t2 = t_end;
    return CMS2D_OK;
}

extern "C" int cms2d_flow_stat(cms2d_flow *F, const char *name, double *out, int *count) {
    static const char *names[NSTAT] = {
        "etamax", "tetamax", "etamin", "tetamin", "uvmax", "umax", "vmax", "tuvmax", "etamean", "ur", "vr", "uvcurv",
        "hydper", "rsuavg", "rsvavg", "rspavg", "rsumax", "rsvmax", "rspmax", "etamaxgrad", "uvmaxgrad", "frxmax",
        "frymax", "umaxcurvx", "umaxcurvy", "vmaxcurvx", "vmaxcurvy", "cflxmax", "cflymax", "qcflxmax", "qcflymax",
        "qfrxmax", "qfrymax"};
    if (!F || !name || !out) return set_err(CMS2D_ERR_ARG, "null argument");
    if (!F->stat_on) return set_err(CMS2D_ERR_ARG, "flow statistics not enabled (cms2d_flow_statistics)");
    for (int q = 0; q < NSTAT; q++) {
        if (strcmp(name, names[q])) continue;
        const Geo &g = F->g;
        std::vector<double> buf(F->npad);
        HIPCHK(hipSetDevice(F->dev));
        HIPCHK(hipMemcpyAsync(buf.data(), F->stat_acc + (size_t)q * F->npad, F->npad * sizeof(double),
                              hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
        for (int j = 0; j < g.ny; j++)
            for (int i = 0; i < g.nx; i++) out[(size_t)j * g.nx + i] = buf[(size_t)(j + 1) * g.nxp + i + 1];
        if (count) *count = F->stat_n;
        return CMS2D_OK;
    }
    return set_err(CMS2D_ERR_ARG, std::string("unknown statistic ") + name);
}

extern "C" int cms2d_ensemble_step(cms2d_flow **flows, int nflows, int nsteps, int nthreads,
                                   cms2d_step_info *infos) {
    if (!flows || nflows < 0 || nsteps < 0) return set_err(CMS2D_ERR_ARG, "invalid argument");
    for (int m = 0; m < nflows; m++)
        if (!flows[m]) return set_err(CMS2D_ERR_ARG, "null member");
    std::vector<int> order(nflows);
    for (int m = 0; m < nflows; m++) order[m] = m;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return flows[a]->g.nact > flows[b]->g.nact; });
    std::vector<int> rc(nflows, CMS2D_OK);
    std::vector<std::string> msg(nflows);
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int k; (k = next.fetch_add(1)) < nflows;) {
            const int m = order[k];
            rc[m] = cms2d_flow_step(flows[m], nsteps, infos ? &infos[m] : nullptr);
            if (rc[m] != CMS2D_OK) msg[m] = g_err;
        }
    };
    int nt = nthreads > 0 ? nthreads : 8;
    nt = std::max(1, std::min(nt, nflows));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    for (int m = 0; m < nflows; m++)
        if (rc[m] != CMS2D_OK) return set_err(rc[m], "member " + std::to_string(m) + ": " + msg[m]);
    return CMS2D_OK;
}

static double *field_ptr(cms2d_flow *F, const char *name) {
    St &s = F->s;
    struct { const char *n; double *p; } tab[] = {
        {"eta", s.eta}, {"u", s.u}, {"v", s.v}, {"p", s.p}, {"h", s.h}, {"vis", s.vis}, {"cfrict", s.cfrict},
        {"coefman", s.coefman},
        {"uv", s.uv}, {"bsvel", s.bsvel}, {"pp", s.pp}, {"dpx", s.dpx}, {"dpy", s.dpy}, {"z0", s.z0},
        {"zb", s.zb}, {"maxeta", s.maxeta}, {"wavestrx", s.wavestrx}, {"wavestry", s.wavestry}, {"fc", s.fc}};
    for (auto &t : tab)
        if (!strcmp(t.n, name)) return t.p;
    return nullptr;
}

// wave_rad_stress_fl (waves/wave_rad_stress_fl.f90:2-59): the wave fields of
// the active cells in, the kinematic radiation-stress forcing of the next
// steps out (k_wave_stress, k_wave_force).
extern "C" int cms2d_flow_set_waves(cms2d_flow *F, const double *whgt, const double *wper, const double *wlen,
                                    const double *wunitx, const double *wunity) {
    if (!F || !whgt || !wper || !wlen || !wunitx || !wunity) return set_err(CMS2D_ERR_ARG, "null argument");
    if (F->P.noptset < 3) return set_err(CMS2D_ERR_ARG, "wave forcing needs a wave mode (noptset >= 3)");
    HIPCHK(hipSetDevice(F->dev));
    const Geo &g = F->g;
    const size_t N = F->npad;
    std::vector<void *> tmp;
    auto cleanup = [&]() { for (void *p : tmp) (void)hipFree(p); };
    auto alloc = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        tmp.push_back(p);
        return p;
    };
    double *in[4], *S[3], *out[2];
    for (auto &p : in) p = (double *)alloc(N * sizeof(double));
    for (auto &p : S) p = (double *)alloc(N * sizeof(double));
    for (auto &p : out) p = (double *)alloc(N * sizeof(double));
    int *bad = (int *)alloc(sizeof(int));
    for (double *p : {in[0], in[1], in[2], in[3], S[0], S[1], S[2], out[0], out[1]})
        if (!p) { cleanup(); return set_err(CMS2D_ERR_HIP, "device allocation failed"); }
    if (!bad) { cleanup(); return set_err(CMS2D_ERR_HIP, "device allocation failed"); }
    const double *src[4] = {whgt, wlen, wunitx, wunity};   // the period only gives c = L/T, unused (:26)
    int rc = 0;
    for (int q = 0; q < 4 && !rc; q++) rc = upload_padded(F, in[q], src[q], false);
    if (!rc) {
        bool ok = true;
        for (double *p : {S[0], S[1], S[2], out[0], out[1]})
            ok = ok && hipMemsetAsync(p, 0, N * sizeof(double), F->st) == hipSuccess;
        ok = ok && hipMemsetAsync(bad, 0, sizeof(int), F->st) == hipSuccess;
        if (!ok) rc = set_err(CMS2D_ERR_HIP, "memset failed");
    }
    int h_bad = 0;
    if (!rc) {
        LAUNCH(k_wave_stress, F->nblk_act, g, F->s, (const double *)in[0], (const double *)in[1],
               (const double *)in[2], (const double *)in[3], S[0], S[1], S[2]);
        LAUNCH(k_wave_force, F->nblk_act, g, F->s, (const double *)S[0], (const double *)S[1],
               (const double *)S[2], out[0], out[1], bad);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, F->st) != hipSuccess ||
            hipStreamSynchronize(F->st) != hipSuccess)
            rc = set_err(CMS2D_ERR_HIP, "radiation-stress kernels failed");
    }
    if (!rc && h_bad)
        rc = set_err(CMS2D_ERR_ARG, "radiation-stress forcing is NaN or above 10 at " + std::to_string(h_bad) +
                                        " cells (wave_rad_stress_fl)");
    if (!rc) {
        if (hipMemcpyAsync(F->s.wavestrx, out[0], N * sizeof(double), hipMemcpyDeviceToDevice, F->st) != hipSuccess ||
            hipMemcpyAsync(F->s.wavestry, out[1], N * sizeof(double), hipMemcpyDeviceToDevice, F->st) != hipSuccess ||
            hipStreamSynchronize(F->st) != hipSuccess)
            rc = set_err(CMS2D_ERR_HIP, "copy of the forcing failed");
    }
    cleanup();
    return rc;
}

// Exact checkpoint: a header, the host-side time and solver state, then every
// state array in allocation order.  Scratch (Krylov vectors, multigrid levels,
// partial sums) is rebuilt by each step and not saved.
namespace {
struct CkHeader {
    char magic[8];
    int version, nx, ny, npad, lmax, narrays;
    long long bytes;
};
struct CkHost {
    Par P;
    double timesecs, rampdur, ramp, dtimebeg, deltime;
    double rmom[3], rmom0[3], rmom1[3];
    int dtvar, jtime, mtime, ntime, maxit, niter, iconv;
    int kry[3], prev_iters[6], prev_gm;
    int windconst;
    double wndx, wndy;
    int wse_inc, q_inc;
};
const char CK_MAGIC[8] = {'C', 'M', 'S', '2', 'D', 'C', 'K', '1'};
}  // namespace

static size_t ck_size(const cms2d_flow *F) {
    size_t n = sizeof(CkHeader) + sizeof(CkHost);
    for (auto &a : F->state) n += a.second;
    return n;
}

// The host-side state of the time loop (timing, convergence control, solver
// batch sizes, boundary-curve window) beside the device arrays.
static void ck_host_get(const cms2d_flow *F, CkHost &c) {
    c.P = F->P;
    c.timesecs = F->timesecs; c.rampdur = F->rampdur; c.ramp = F->ramp; c.dtimebeg = F->dtimebeg; c.deltime = F->deltime;
    for (int n = 0; n < 3; n++) { c.rmom[n] = F->rmom[n]; c.rmom0[n] = F->rmom0[n]; c.rmom1[n] = F->rmom1[n]; c.kry[n] = F->kry[n]; }
    c.dtvar = F->dtvar; c.jtime = F->jtime; c.mtime = F->mtime; c.ntime = F->ntime; c.maxit = F->maxit;
    c.niter = F->niter; c.iconv = F->iconv; c.prev_gm = F->prev_gm;
    for (int n = 0; n < 6; n++) c.prev_iters[n] = F->prev_iters[n];
    c.windconst = F->windconst; c.wndx = F->wndx; c.wndy = F->wndy;
    c.wse_inc = F->wse_inc;
    c.q_inc = F->q_inc;
}
static void ck_host_set(cms2d_flow *F, const CkHost &c) {
    F->P = c.P;
    F->timesecs = c.timesecs; F->rampdur = c.rampdur; F->ramp = c.ramp; F->dtimebeg = c.dtimebeg; F->deltime = c.deltime;
    for (int n = 0; n < 3; n++) { F->rmom[n] = c.rmom[n]; F->rmom0[n] = c.rmom0[n]; F->rmom1[n] = c.rmom1[n]; F->kry[n] = c.kry[n]; }
    F->dtvar = c.dtvar; F->jtime = c.jtime; F->mtime = c.mtime; F->ntime = c.ntime; F->maxit = c.maxit;
    F->niter = c.niter; F->iconv = c.iconv; F->prev_gm = c.prev_gm;
    for (int n = 0; n < 6; n++) F->prev_iters[n] = c.prev_iters[n];
    F->windconst = c.windconst; F->wndx = c.wndx; F->wndy = c.wndy;
    F->wse_inc = c.wse_inc;
    F->q_inc = c.q_inc;
}

extern "C" int cms2d_flow_checkpoint(cms2d_flow *F, void *buf, size_t *nbytes) {
    if (!F || !nbytes) return set_err(CMS2D_ERR_ARG, "null argument");
    if (F->tel) return set_err(CMS2D_ERR_ARG, "checkpoints are built for Cartesian grids only");
    const size_t need = ck_size(F);
    if (!buf) { *nbytes = need; return CMS2D_OK; }
    if (*nbytes < need) return set_err(CMS2D_ERR_ARG, "checkpoint buffer too small: " + std::to_string(need) + " bytes needed");
    HIPCHK(hipSetDevice(F->dev));
    HIPCHK(hipStreamSynchronize(F->st));
    char *q = (char *)buf;
    CkHeader h{};
    memcpy(h.magic, CK_MAGIC, 8);
    h.version = 1; h.nx = F->g.nx; h.ny = F->g.ny; h.npad = F->npad; h.lmax = F->P.lmax;
    h.narrays = (int)F->state.size(); h.bytes = (long long)need;
    memcpy(q, &h, sizeof h);
    q += sizeof h;
    CkHost c{};
    ck_host_get(F, c);
    memcpy(q, &c, sizeof c);
    q += sizeof c;
    for (auto &a : F->state) {
        HIPCHK(hipMemcpyAsync(q, a.first, a.second, hipMemcpyDeviceToHost, F->st));
        q += a.second;
    }
    HIPCHK(hipStreamSynchronize(F->st));
    *nbytes = need;
    return CMS2D_OK;
}

extern "C" int cms2d_flow_restore(cms2d_flow *F, const void *buf, size_t nbytes) {
    if (!F || !buf) return set_err(CMS2D_ERR_ARG, "null argument");
    if (F->tel) return set_err(CMS2D_ERR_ARG, "checkpoints are built for Cartesian grids only");
    const size_t need = ck_size(F);
    CkHeader h{};
    if (nbytes < sizeof h) return set_err(CMS2D_ERR_ARG, "not a checkpoint");
    memcpy(&h, buf, sizeof h);
    if (memcmp(h.magic, CK_MAGIC, 8) || h.version != 1) return set_err(CMS2D_ERR_ARG, "not a checkpoint of this library");
    if (h.nx != F->g.nx || h.ny != F->g.ny || h.npad != F->npad || h.lmax != F->P.lmax ||
        h.narrays != (int)F->state.size() || (size_t)h.bytes != need || nbytes < need)
        return set_err(CMS2D_ERR_ARG, "checkpoint of a different grid");
    const char *q = (const char *)buf + sizeof h;
    CkHost c{};
    memcpy(&c, q, sizeof c);
    q += sizeof c;
    HIPCHK(hipSetDevice(F->dev));
    HIPCHK(hipStreamSynchronize(F->st));
    for (auto &a : F->state) {
        HIPCHK(hipMemcpyAsync(a.first, q, a.second, hipMemcpyHostToDevice, F->st));
        q += a.second;
    }
    HIPCHK(hipStreamSynchronize(F->st));
    ck_host_set(F, c);
    return CMS2D_OK;
}

// Spatially constant wind: windcurve_eval's 10-m vector in grid coordinates;
// each following step adds ramp Cd(W) rho_air/rho W (wndx, wndy) to the
// momentum sources (flow_srcsnk_static, flow_imp.F90:386-394).
extern "C" int cms2d_flow_set_wind(cms2d_flow *F, double wndx, double wndy) {
    if (!F) return set_err(CMS2D_ERR_ARG, "null argument");
    if (!std::isfinite(wndx) || !std::isfinite(wndy)) return set_err(CMS2D_ERR_ARG, "wind not finite");
    F->windconst = 1;
    F->wndx = wndx;
    F->wndy = wndy;
    return CMS2D_OK;
}

extern "C" int cms2d_flow_tuning(cms2d_flow *F, int *mg_tile_x, int *graphs) {
    if (!F || !mg_tile_x || !graphs) return set_err(CMS2D_ERR_ARG, "null argument");
    *mg_tile_x = F->mg.nlev > 1 ? F->mg_tile : 0;
    *graphs = F->use_graphs ? 1 : 0;
    return CMS2D_OK;
}

extern "C" int cms2d_flow_get(cms2d_flow *F, const char *name, double *out) {
    if (!F || !name || !out) return set_err(CMS2D_ERR_ARG, "null argument");
    if (F->tel) return tel_get(F, name, out);
    HIPCHK(hipSetDevice(F->dev));
    HIPCHK(hipStreamSynchronize(F->st));
    const Geo &g = F->g;
    std::vector<double> h(F->npad);
    if (!strcmp(name, "iwet")) {
        std::vector<uint8_t> w(F->npad);
        HIPCHK(hipMemcpyAsync(w.data(), F->s.iwet, F->npad, hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
        for (int j = 0; j < g.ny; j++)
            for (int i = 0; i < g.nx; i++) out[(size_t)j * g.nx + i] = w[(size_t)(j + 1) * g.nxp + i + 1];
        return CMS2D_OK;
    }
    double *p = field_ptr(F, name);
    if (!p) return set_err(CMS2D_ERR_ARG, std::string("unknown field ") + name);
    HIPCHK(hipMemcpyAsync(h.data(), p, F->npad * sizeof(double), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) out[(size_t)j * g.nx + i] = h[(size_t)(j + 1) * g.nxp + i + 1];
    return CMS2D_OK;
}

extern "C" int cms2d_flow_set(cms2d_flow *F, const char *name, const double *in) {
    if (!F || !name || !in) return set_err(CMS2D_ERR_ARG, "null argument");
    if (strcmp(name, "u") && strcmp(name, "v") && strcmp(name, "p"))
        return set_err(CMS2D_ERR_ARG, std::string("field not settable: ") + name);
    if (F->tel) return tel_set(F, name, in);
    HIPCHK(hipSetDevice(F->dev));
    HIPCHK(hipStreamSynchronize(F->st));
    const Geo &g = F->g;
    double *p = field_ptr(F, name);
    std::vector<double> h(F->npad);
    HIPCHK(hipMemcpyAsync(h.data(), p, F->npad * sizeof(double), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) h[(size_t)(j + 1) * g.nxp + i + 1] = in[(size_t)j * g.nx + i];
    HIPCHK(hipMemcpyAsync(p, h.data(), F->npad * sizeof(double), hipMemcpyHostToDevice, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    // an initial condition: rebuild the state as at the start of a run
    return init_state(F);
}

extern "C" int cms2d_flow_get_face(cms2d_flow *F, const char *name, double *out) {
    if (!F || !name || !out) return set_err(CMS2D_ERR_ARG, "null argument");
    if (F->tel) return tel_get_face(F, name, out);
    HIPCHK(hipSetDevice(F->dev));
    HIPCHK(hipStreamSynchronize(F->st));
    const Geo &g = F->g;
    const double *E, *Nn;
    bool anti;
    if (!strcmp(name, "flux")) { E = F->s.fluxE; Nn = F->s.fluxN; anti = true; }
    else if (!strcmp(name, "hk")) { E = F->s.hkE; Nn = F->s.hkN; anti = false; }
    else if (!strcmp(name, "visk")) { E = F->s.viskE; Nn = F->s.viskN; anti = false; }
    else return set_err(CMS2D_ERR_ARG, std::string("unknown face field ") + name);
    std::vector<double> he(F->npad), hn(F->npad);
    HIPCHK(hipMemcpyAsync(he.data(), E, F->npad * sizeof(double), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipMemcpyAsync(hn.data(), Nn, F->npad * sizeof(double), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    const double sg = anti ? -1.0 : 1.0;
    for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) {
            size_t c = (size_t)(j + 1) * g.nxp + i + 1, o = 4 * ((size_t)j * g.nx + i);
            out[o + KN] = hn[c];
            out[o + KE] = he[c];
            out[o + KS] = sg * hn[c - g.nxp];
            out[o + KW] = sg * he[c - 1];
        }
    return CMS2D_OK;
}

extern "C" double cms2d_flow_time(const cms2d_flow *F) { return F ? F->timesecs : 0.0; }

extern "C" int cms2d_flow_kernel_stats(const cms2d_flow *F, const char *name, double *avg_ms, long *count) {
    if (!F || !name || !avg_ms || !count) return set_err(CMS2D_ERR_ARG, "null argument");
    for (int k = 0; k < TG_N; k++)
        if (!strcmp(name, TG_NAME[k])) {
            // average over the timed launches; count of every launch that did work
            *count = F->tm.n[k] + F->tm.counted[k];
            *avg_ms = F->tm.n[k] ? F->tm.ms[k] / F->tm.n[k] : 0.0;
            return CMS2D_OK;
        }
    return set_err(CMS2D_ERR_ARG, std::string("unknown kernel group ") + name);
}

// ----------------------------------------------------------------------------
// cms2d_solve: solve(acoef,ss,sp,res,phi,n) on the reference's cell list.

// Device buffers of one cms2d_solve call: the staged inputs/outputs and the
// Krylov scratch (7 vectors of ncells, per-block partials, the scalars).
struct SolveBufs {
    double *a, *ss, *sp, *res, *phi, *vec, *partials;
    int *c2c, *w;
    KryScal *sc;
};

static int solve_check_args(int ncells, int ncellsD, int nmaxfaces, const double *acoef, const int *cell2cell,
                            const int *iwet, const double *ss, const double *sp, const double *res,
                            const double *phi, double relax) {
    if (ncells <= 0 || ncellsD < ncells || !acoef || !cell2cell || !iwet || !ss || !sp || !res || !phi ||
        !(relax > 0.0))
        return set_err(CMS2D_ERR_ARG, "invalid argument");
    if (nmaxfaces < 1 || nmaxfaces > ELL_MAX_FACES)
        return set_err(CMS2D_ERR_ARG, "nmaxfaces must be 1.." + std::to_string(ELL_MAX_FACES));
    const long nf = nmaxfaces;
    for (long f = 0; f < nf * ncells; f++)
        if (cell2cell[f] < -1 || cell2cell[f] >= ncellsD) return set_err(CMS2D_ERR_ARG, "cell2cell out of range");
    return 0;
}

static int solve_device(const SolveBufs &B, hipStream_t st, int ncells, int nmaxfaces, double relax, double rtol,
                        double atol, int kmax, KryScal &h_sc, double bias = SMALL);

// Uploads the system, runs the prologue and BiCGStab, downloads ss, res, phi.
// The scalars are cleared by a kernel (no memset node), so the same sequence
// serves both entry points.
static int solve_run(const SolveBufs &B, hipStream_t st, int ncells, int ncellsD, int nmaxfaces,
                     const double *acoef, const int *cell2cell, const int *iwet, double *ss, const double *sp,
                     double *res, double *phi, double relax, double rtol, double atol, int kmax, double *rmom,
                     int *iters) {
    const size_t nD = ncellsD, n = ncells, nf = nmaxfaces;
    HIPCHK(hipMemcpyAsync(B.a, acoef, nf * nD * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.c2c, cell2cell, nf * nD * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.w, iwet, nD * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.ss, ss, nD * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.sp, sp, nD * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.phi, phi, nD * sizeof(double), hipMemcpyHostToDevice, st));
    KryScal h_sc{};
    if (int rc = solve_device(B, st, ncells, nmaxfaces, relax, rtol, atol, kmax, h_sc)) return rc;
    HIPCHK(hipMemcpyAsync(ss, B.ss, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(res, B.res, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(phi, B.phi, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (rmom) *rmom = h_sc.rmom;
    if (iters) *iters = h_sc.final_iters;
    return 0;
}

// The prologue and BiCGStab on a system already in B (device memory).
static int solve_device(const SolveBufs &B, hipStream_t st, int ncells, int nmaxfaces, double relax, double rtol,
                        double atol, int kmax, KryScal &h_sc, double bias) {
    const size_t n = ncells;
    const int nblk = std::min(KRY_BLOCKS, (ncells + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, (unsigned int *)B.sc,
                       (int)(sizeof(KryScal) / sizeof(unsigned int)));
    KrySys<1> K{};
    K.x[0] = B.phi;
    K.r[0] = B.vec; K.r0[0] = B.vec + n; K.p[0] = B.vec + 2 * n; K.v[0] = B.vec + 3 * n; K.t[0] = B.vec + 4 * n;
    K.ap_w[0] = B.vec + 5 * n;
    K.dinv_w[0] = B.vec + 6 * n;
    K.sc = B.sc;
    K.red.pr[0] = B.partials; K.red.pv[0] = B.partials + 2 * nblk; K.red.pt[0] = B.partials + 3 * nblk;
    EllOp A{ncells, nmaxfaces, B.c2c, B.a};
    hipLaunchKernelGGL(k_ell_prologue, dim3(nblk), dim3(KBLOCK), 0, st, A, (const int *)B.w, B.ss,
                       (const double *)B.sp, B.res, (const double *)B.phi, relax, K, bias);
    int prev = 8;
    KryCtx X{st, nblk, kmax > 0 ? kmax : 4000, rtol > 0 ? rtol : 1e-10, atol >= 0 ? atol : 1e-12,
             B.sc, &h_sc, &prev, nullptr};
    if (int rc = run_bicg(X, K, A, 0)) return rc;
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

extern "C" int cms2d_solve(int ncells, int ncellsD, int nmaxfaces, const double *acoef, const int *cell2cell,
                           const int *iwet, double *ss, const double *sp, double *res, double *phi, double relax,
                           double rtol, double atol, int kmax, double *rmom, int *iters) {
    if (int rc = solve_check_args(ncells, ncellsD, nmaxfaces, acoef, cell2cell, iwet, ss, sp, res, phi, relax))
        return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(CMS2D_ERR_NODEVICE, "no HIP device");
    std::vector<void *> mem;
    auto cleanup = [&]() { for (void *p : mem) (void)hipFree(p); };
    auto alloc = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
        mem.push_back(p);
        return p;
    };
    const size_t nD = ncellsD, n = ncells, nf = nmaxfaces;
    const int nblk = std::min(KRY_BLOCKS, (ncells + BLOCK - 1) / BLOCK);
    SolveBufs B{};
    B.a = (double *)alloc(nf * nD * sizeof(double));
    B.c2c = (int *)alloc(nf * nD * sizeof(int));
    B.w = (int *)alloc(nD * sizeof(int));
    B.ss = (double *)alloc(nD * sizeof(double));
    B.sp = (double *)alloc(nD * sizeof(double));
    B.res = (double *)alloc(nD * sizeof(double));
    B.phi = (double *)alloc(nD * sizeof(double));
    B.vec = (double *)alloc(7 * n * sizeof(double));
    B.partials = (double *)alloc((size_t)nblk * 5 * sizeof(double));
    B.sc = (KryScal *)alloc(sizeof(KryScal));
    if (!B.a || !B.c2c || !B.w || !B.ss || !B.sp || !B.res || !B.phi || !B.vec || !B.partials || !B.sc) {
        cleanup();
        return set_err(CMS2D_ERR_HIP, "device allocation failed");
    }
    hipStream_t st = nullptr;
    int rc = 0;
    if (hipStreamCreate(&st) != hipSuccess) rc = set_err(CMS2D_ERR_HIP, "stream");
    else rc = solve_run(B, st, ncells, ncellsD, nmaxfaces, acoef, cell2cell, iwet, ss, sp, res, phi, relax, rtol,
                        atol, kmax, rmom, iters);
    if (st) { (void)hipStreamSynchronize(st); (void)hipStreamDestroy(st); }
    cleanup();
    return rc;
}

// ----------------------------------------------------------------------------
// cms2d_solve_adjoint: the reverse of solve.  The solved system on the active
// cells is A phi = b with A_ii = ap_i = (sum_k a_ik - sp_i) / relax,
// A_ij = -a_ik for an active neighbour j, and b_i = ss_i + (1 - relax) ap_i
// phi_old_i + sum over ghost neighbours of a_ik phi_g.  For a loss L(phi) the
// adjoint lambda = A^-T dL/dphi gives every input's gradient in closed form.
// A^T has the structure of A (faces are shared by the two cells): its row i
// holds a_jb of each active neighbour j at the slot of j that points back to
// i, so the transposed solve runs the same kernels on a system written by
// k_adj_setup (relax 1, sp chosen so that its diagonal is ap).

__global__ void k_adj_setup(EllOp A, const int *back, const double *sp, double relax, const double *gphi,
                            double *at, double *spt, double *sst, double *lam) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.ncells) return;
    const int nf = A.nf;
    double asum = 0.0, tsum = 0.0;
    for (int k = 0; k < nf; k++) {
        const int j = A.c2c[nf * i + k];
        if (j < 0) continue;
        asum += A.a[nf * i + k];
        const double v = j < A.ncells ? A.a[nf * j + back[nf * i + k]] : 0.0;
        at[nf * i + k] = v;
        tsum += v;
    }
    const double ap = (asum - sp[i]) / relax;
    spt[i] = tsum - ap;
    sst[i] = gphi[i];
    lam[i] = 0.0;
}
// Gradients of the active cells' inputs from lambda.
__global__ void k_adj_grads(EllOp A, const double *sp, const double *phi_old, const double *phi, double relax,
                            const double *lam, double *gss, double *gsp, double *ga, double *gpo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.ncells) return;
    const int nf = A.nf;
    double asum = 0.0;
    for (int k = 0; k < nf; k++)
        if (A.c2c[nf * i + k] >= 0) asum += A.a[nf * i + k];
    const double ap = (asum - sp[i]) / relax, l = lam[i], rm = 1.0 - relax;
    gss[i] = l;
    gpo[i] = l * rm * ap;
    gsp[i] = l * (phi[i] - rm * phi_old[i]) / relax;
    const double base = (rm * phi_old[i] - phi[i]) / relax;
    for (int k = 0; k < nf; k++) {
        const int j = A.c2c[nf * i + k];
        ga[nf * i + k] = j < 0 ? 0.0 : l * (phi[j] + base);
    }
}
// Gradient of each ghost value: the sum of lambda_i a_ik over the faces that
// point to it, in the order of the list gref (flat face indices i*nf+k).
__global__ void k_adj_ghost(EllOp A, int nghost, const int *gstart, const int *gref, const double *lam,
                            double *gpo) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nghost) return;
    double s = 0.0;
    for (int q = gstart[g]; q < gstart[g + 1]; q++) s += lam[gref[q] / A.nf] * A.a[gref[q]];
    gpo[A.ncells + g] = s;
}

extern "C" int cms2d_solve_adjoint(int ncells, int ncellsD, int nmaxfaces, const double *acoef,
                                   const int *cell2cell, const double *sp, const double *phi_old, const double *phi,
                                   double relax, const double *gphi, double rtol, double atol, int kmax,
                                   double *gss, double *gsp, double *gacoef, double *gphi_old, int *iters) {
    if (ncells <= 0 || ncellsD < ncells || !acoef || !cell2cell || !sp || !phi_old || !phi || !gphi || !gss ||
        !gsp || !gacoef || !gphi_old || !(relax > 0.0))
        return set_err(CMS2D_ERR_ARG, "invalid argument");
    if (nmaxfaces < 1 || nmaxfaces > ELL_MAX_FACES)
        return set_err(CMS2D_ERR_ARG, "nmaxfaces must be 1.." + std::to_string(ELL_MAX_FACES));
    const size_t nD = ncellsD, n = ncells, nf = nmaxfaces, ng = nD - n;
    // the slot of each active neighbour that points back, and the faces into each ghost
    std::vector<int> back(nf * n, 0), gcnt(ng + 1, 0);
    for (size_t i = 0; i < n; i++)
        for (size_t k = 0; k < nf; k++) {
            const int j = cell2cell[nf * i + k];
            if (j < -1 || j >= ncellsD) return set_err(CMS2D_ERR_ARG, "cell2cell out of range");
            if (j < 0) continue;
            if ((size_t)j >= n) { gcnt[j - n + 1]++; continue; }
            int b = -1;
            for (size_t kk = 0; kk < nf; kk++)
                if (cell2cell[nf * j + kk] == (int)i) { b = (int)kk; break; }
            if (b < 0)
                return set_err(CMS2D_ERR_ARG, "cell " + std::to_string(j + 1) + " does not list its neighbour " +
                                                  std::to_string(i + 1) + " (faces must be shared)");
            back[nf * i + k] = b;
        }
    for (size_t g = 0; g < ng; g++) gcnt[g + 1] += gcnt[g];
    std::vector<int> gref(std::max<int>(gcnt[ng], 1)), fill(gcnt.begin(), gcnt.end() - 1);
    for (size_t i = 0; i < n; i++)
        for (size_t k = 0; k < nf; k++) {
            const int j = cell2cell[nf * i + k];
            if (j >= (int)n) gref[fill[j - n]++] = (int)(nf * i + k);
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(CMS2D_ERR_NODEVICE, "no HIP device");
    std::vector<void *> mem;
    auto cleanup = [&]() { for (void *p : mem) (void)hipFree(p); };
    auto alloc = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
        mem.push_back(p);
        return p;
    };
    const int nblk = std::min(KRY_BLOCKS, (ncells + BLOCK - 1) / BLOCK);
    double *a = (double *)alloc(nf * nD * sizeof(double)), *spd = (double *)alloc(nD * sizeof(double));
    double *po = (double *)alloc(nD * sizeof(double)), *ph = (double *)alloc(nD * sizeof(double));
    double *gp = (double *)alloc(nD * sizeof(double));
    double *o_gss = (double *)alloc(nD * sizeof(double)), *o_gsp = (double *)alloc(nD * sizeof(double));
    double *o_ga = (double *)alloc(nf * nD * sizeof(double)), *o_gpo = (double *)alloc(nD * sizeof(double));
    int *c2c = (int *)alloc(nf * nD * sizeof(int)), *bk = (int *)alloc(nf * n * sizeof(int));
    int *gs = (int *)alloc((ng + 1) * sizeof(int)), *gr = (int *)alloc(gref.size() * sizeof(int));
    SolveBufs B{};
    B.a = (double *)alloc(nf * nD * sizeof(double));
    B.c2c = c2c;
    B.w = (int *)alloc(nD * sizeof(int));
    B.ss = (double *)alloc(nD * sizeof(double));
    B.sp = (double *)alloc(nD * sizeof(double));
    B.res = (double *)alloc(nD * sizeof(double));
    B.phi = (double *)alloc(nD * sizeof(double));
    B.vec = (double *)alloc(7 * n * sizeof(double));
    B.partials = (double *)alloc((size_t)nblk * 5 * sizeof(double));
    B.sc = (KryScal *)alloc(sizeof(KryScal));
    for (void *p : mem)
        if (!p) { cleanup(); return set_err(CMS2D_ERR_HIP, "device allocation failed"); }
    hipStream_t st = nullptr;
    int rc = 0;
    auto run = [&]() -> int {
        HIPCHK(hipStreamCreate(&st));
        HIPCHK(hipMemcpyAsync(a, acoef, nf * nD * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c2c, cell2cell, nf * nD * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(bk, back.data(), nf * n * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(gs, gcnt.data(), (ng + 1) * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(gr, gref.data(), gref.size() * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(spd, sp, nD * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(po, phi_old, nD * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(ph, phi, nD * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(gp, gphi, n * sizeof(double), hipMemcpyHostToDevice, st));
        const EllOp A{ncells, nmaxfaces, c2c, a};
        const unsigned nb = (unsigned)((n + BLOCK - 1) / BLOCK);
        hipLaunchKernelGGL(k_adj_setup, dim3(nb), dim3(BLOCK), 0, st, A, (const int *)bk, (const double *)spd, relax,
                           (const double *)gp, B.a, B.sp, B.ss, B.phi);
        hipLaunchKernelGGL(k_set_int, dim3(nb), dim3(BLOCK), 0, st, B.w, ncells, 1);
        KryScal h_sc{};
        if (int r = solve_device(B, st, ncells, nmaxfaces, 1.0, rtol, atol, kmax, h_sc, 0.0)) return r;
        hipLaunchKernelGGL(k_adj_grads, dim3(nb), dim3(BLOCK), 0, st, A, (const double *)spd, (const double *)po,
                           (const double *)ph, relax, (const double *)B.phi, o_gss, o_gsp, o_ga, o_gpo);
        if (ng)
            hipLaunchKernelGGL(k_adj_ghost, dim3((unsigned)((ng + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, A, (int)ng,
                               (const int *)gs, (const int *)gr, (const double *)B.phi, o_gpo);
        HIPCHK(hipMemcpyAsync(gss, o_gss, n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(gsp, o_gsp, n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(gacoef, o_ga, nf * n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(gphi_old, o_gpo, nD * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (size_t i = n; i < nD; i++) {        // ghost rows take no part in the system
            gss[i] = gsp[i] = 0.0;
            for (size_t k = 0; k < nf; k++) gacoef[nf * i + k] = 0.0;
        }
        if (iters) *iters = h_sc.final_iters;
        return 0;
    };
    rc = run();
    if (st) { (void)hipStreamSynchronize(st); (void)hipStreamDestroy(st); }
    cleanup();
    return rc;
}

// ----------------------------------------------------------------------------
// cms2d_solve_noalloc: cms2d_solve without device allocation.  Its staging
// and Krylov scratch live in one static device array of CMS2D_SOLVE_SCRATCH_BYTES
// per device (held until the process ends), carved up per call like the
// buffers cms2d_solve allocates.  The host-side name of the array is not a
// device address: every host use goes through the address that
// hipGetSymbolAddress returns, looked up once per device.  Calls run on the
// device's null stream; calls that overlap on one device share the array and
// race.
static __device__ __align__(256) unsigned char scratch[CMS2D_SOLVE_SCRATCH_BYTES];

static constexpr int SCRATCH_MAX_DEV = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[SCRATCH_MAX_DEV];

static int scratch_addr(void **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEV) return set_err(CMS2D_ERR_ARG, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(CMS2D_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// Bytes of scratch a call needs (each part 256-byte aligned).
static size_t solve_scratch_bytes(size_t n, size_t nD, size_t nf, size_t nblk) {
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    return al(nf * nD * sizeof(double)) + al(nf * nD * sizeof(int)) + al(nD * sizeof(int)) +
           4 * al(nD * sizeof(double)) + al(7 * n * sizeof(double)) + al(nblk * 5 * sizeof(double)) +
           al(sizeof(KryScal));
}

extern "C" int cms2d_solve_noalloc(int ncells, int ncellsD, int nmaxfaces, const double *acoef,
                                   const int *cell2cell, const int *iwet, double *ss, const double *sp,
                                   double *res, double *phi, double relax, double rtol, double atol, int kmax,
                                   double *rmom, int *iters) {
    if (int rc = solve_check_args(ncells, ncellsD, nmaxfaces, acoef, cell2cell, iwet, ss, sp, res, phi, relax))
        return rc;
    const size_t nD = ncellsD, n = ncells, nf = nmaxfaces;
    const int nblk = std::min(KRY_BLOCKS, (ncells + BLOCK - 1) / BLOCK);
    const size_t need = solve_scratch_bytes(n, nD, nf, nblk);
    if (need > (size_t)CMS2D_SOLVE_SCRATCH_BYTES)
        return set_err(CMS2D_ERR_ARG, "system needs " + std::to_string(need) + " bytes of scratch, beyond the " +
                                          std::to_string((size_t)CMS2D_SOLVE_SCRATCH_BYTES) +
                                          " of cms2d_solve_noalloc: use cms2d_solve");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(CMS2D_ERR_NODEVICE, "no HIP device");
    void *base = nullptr;
    if (int rc = scratch_addr(&base)) return rc;
    unsigned char *q = (unsigned char *)base;
    auto take = [&](size_t bytes) {
        void *p = q;
        q += (bytes + 255) / 256 * 256;
        return p;
    };
    SolveBufs B{};
    B.a = (double *)take(nf * nD * sizeof(double));
    B.c2c = (int *)take(nf * nD * sizeof(int));
    B.w = (int *)take(nD * sizeof(int));
    B.ss = (double *)take(nD * sizeof(double));
    B.sp = (double *)take(nD * sizeof(double));
    B.res = (double *)take(nD * sizeof(double));
    B.phi = (double *)take(nD * sizeof(double));
    B.vec = (double *)take(7 * n * sizeof(double));
    B.partials = (double *)take((size_t)nblk * 5 * sizeof(double));
    B.sc = (KryScal *)take(sizeof(KryScal));
    const int rc = solve_run(B, nullptr, ncells, ncellsD, nmaxfaces, acoef, cell2cell, iwet, ss, sp, res, phi,
                             relax, rtol, atol, kmax, rmom, iters);
    (void)hipStreamSynchronize(nullptr);
    return rc;
}

#include "adjoint.hip"
