// tel.hip -- host driver of the implicit step on telescoping grids, and the
// pressure-correction preconditioner that couples the cell list to the
// multigrid hierarchy of its base grid.  Included by cms2d.hip after the
// Cartesian driver, whose Krylov, multigrid and control routines it reuses.
//
// Preconditioner (z = M^-1 r, FP32): a V-cycle over the levels of the
// quadtree and then the base grid.  Level k (1 <= k < finest) holds squares of
// 2^-k base cells: the real cells of level k and virtual cells that stand for
// the finer cells inside them.  Going down, the finest cells take damped Jacobi
// sweeps, the fine residual r - A z is summed into the level-k squares (the
// equations are cell integrals), each level takes its sweeps and adds its
// correction to the cells inside its squares; the residual summed into the
// base cells is solved by one V-cycle of the base grid's multigrid, whose
// correction is added to every cell of the base cell; going up, each level
// and then the finest cells take their sweeps again.  The operators of the
// coarser levels are re-discretised from the cell list: the coefficient
// across a square's or base cell's edge is the sum of the cell-face
// coefficients it covers, each scaled by its centre distance over the coarse
// one (a coefficient ~ ds/dn, so the sum keeps the coarse face's ds/dn), and
// the row excess ap - sum(a) (storage term, folded boundary coefficients) is
// summed over the cells.  On unrefined base cells the base operator is the
// cell operator itself.

namespace cms {

constexpr float TEL_OMEGA = 0.8f;     // damping of the Jacobi sweeps
constexpr int TEL_NU = 2;             // sweeps per level before and after the coarse correction
static_assert(TEL_NU == 2, "k_tl_cs_pre2 and k_tl_cs_jacp each run the two sweeps of a level");
constexpr int TEL_MAX_LEVELS = 5;     // refinement levels below the base grid (16 x 16 cells per base cell)

// Level k of the refined region (1 <= k < finest level): squares of 2^-k base
// cells.  Lists are ordered by the squares' first cells.
struct TelLev {
    int nk;                       // squares
    const int *kstart, *kcell;    // cells inside each square, ascending: [nk+1], [kstart[nk]]
    const int *kseg;              // square of each entry of kcell
    const int *blk;               // entry ranges of the restriction's blocks (whole squares) [nblk+1]
    int nblk;
    const int *nbr;               // [4*nk] (side-major) level-k square across N, E, S, W, or -1
    const int *cstart, *cf;       // faces of the cells crossing each square's edge: flat index k*nd+i
    const float *cw;              // their weight dn / dn of the square
    const uint8_t *cs;            // their side 0..3, or 4 toward a coarser cell (diagonal only)
    const int *ge;                // cells of level >= k [nge], ascending
    int nge;
    const int *anc;               // [n] square of each cell of level >= k
    float *a;                     // [4*nk] (side-major) off-diagonal coefficients
    float *ap, *dinv;             // [nk]
    float *r;                     // [nk] summed residual
    float *z;                     // [nk] the first sweep's values
};

// Base-grid level 0 of the multigrid hierarchy from the cell-list pp system
// (once per outer iteration).  One thread per entry of the base-cell-ordered
// list on blocks of whole base cells; the thread of a base cell's first entry
// sums the contributions of its cells in list order from LDS.
__global__ void __launch_bounds__(BLOCK) k_tl_mg_setup0(TelG T, const int *blk, const double *apc, const double *ap,
                                                        const int *bcells, MgLevel L) {
    __shared__ double sv[5][BLOCK];
    const int q0 = blk[blockIdx.x], q1 = blk[blockIdx.x + 1], q = q0 + threadIdx.x;
    double c4[4] = {0.0, 0.0, 0.0, 0.0}, e = 0.0;
    int b = 0;
    if (q < q1) {
        const int i = bcells[q];
        b = T.base[i];
        const uint64_t m = T.meta[i];
        double sa = 0.0;
        for (int k = 0; k < m_ncf(m); k++) {
            const int nb = tnb(T, i, k);
            if (nb >= T.n) continue;
            const double a = apc[k * T.nd + i];
            sa += a;
            if (T.base[nb] == b) continue;
            const int d = m_dir(m, k);
            const double dn = (d == KE || d == KW) ? fabs(T.x[nb] - T.x[i]) : fabs(T.y[nb] - T.y[i]);
            c4[d] += a * dn / ((d == KE || d == KW) ? T.dxb : T.dyb);
        }
        e = fmax(ap[i] - sa, 0.0);
    }
    for (int d = 0; d < 4; d++) sv[d][threadIdx.x] = c4[d];
    sv[4][threadIdx.x] = e;
    __syncthreads();
    if (q >= q1 || T.bstart[b] != q) return;
    double a4[4] = {0.0, 0.0, 0.0, 0.0}, es = 0.0;
    for (int p = q; p < T.bstart[b + 1]; p++) {
        for (int d = 0; d < 4; d++) a4[d] += sv[d][p - q0];
        es += sv[4][p - q0];
    }
    const int c = L.idx(b % T.nbx, b / T.nbx);
    const double apb = es + ((a4[KN] + a4[KE]) + (a4[KS] + a4[KW]));
    L.aN[c] = (float)a4[KN];
    L.aE[c] = (float)a4[KE];
    L.aS[c] = (float)a4[KS];
    L.aW[c] = (float)a4[KW];
    L.e[c] = (float)es;
    L.ap[c] = (float)apb;
    L.dinv[c] = apb > 0.0 ? (float)(1.0 / apb) : 0.0f;
}

// The residual r - A Z of cell i for a correction Z that is nonzero on
// refined cells only (so only cells next to them see it); the face loads
// are issued together (every slot, predicated) and summed in face order.
template <class TB>
__device__ __forceinline__ double tl_res(const TelG &T, const TB *r, const double *ap, const double *apc,
                                         const float *Z, int i, uint64_t m) {
    double ri = (double)r[i];
    if (m_near(m)) {
        const int nf = m_ncf(m);
        int nbk[TF];
        float zk[TF];
#pragma unroll
        for (int k = 0; k < TF; k++) nbk[k] = k < nf ? tnb(T, i, k) : T.n;
#pragma unroll
        for (int k = 0; k < TF; k++) zk[k] = nbk[k] < T.n ? Z[nbk[k]] : 0.0f;
        double az = ap[i] * (double)Z[i];
#pragma unroll
        for (int k = 0; k < TF; k++)
            if (nbk[k] < T.n) az -= apc[k * T.nd + i] * (double)zk[k];
        ri -= az;
    }
    return ri;
}

// Coefficients of level k from the cell list's pp system (once per outer iteration).
__global__ void k_tl_cs_setup(TelG T, TelLev V, const double *apc, const double *ap) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= V.nk) return;
    double e = 0.0;
    for (int q = V.kstart[c]; q < V.kstart[c + 1]; q++) {
        const int i = V.kcell[q];
        const uint64_t m = T.meta[i];
        double sa = 0.0;
        for (int k = 0; k < m_ncf(m); k++)
            if (tnb(T, i, k) < T.n) sa += apc[k * T.nd + i];
        e += fmax(ap[i] - sa, 0.0);
    }
    double a4[4] = {0.0, 0.0, 0.0, 0.0}, diag = 0.0;
    for (int q = V.cstart[c]; q < V.cstart[c + 1]; q++) {
        const double v = (double)V.cw[q] * apc[V.cf[q]];
        diag += v;
        if (V.cs[q] < 4) a4[V.cs[q]] += v;
    }
    const double apk = e + diag;
    for (int d = 0; d < 4; d++) V.a[d * V.nk + c] = (float)a4[d];
    V.ap[c] = (float)apk;
    V.dinv[c] = apk > 0.0 ? (float)(1.0 / apk) : 0.0f;
}

// The two down sweeps on the finest cells from zero in one pass: the first
// sweep's values omega D^-1 r are recomputed where the second reads them;
// the other refined cells get 0.
template <class TB>
__global__ void k_tl_cs_pre2(TelG T, const int *ref, int nref, const uint8_t *lev, int top, const TB *r,
                             const double *ap, const double *apc, float *Z, const KryScal *sc) {
    if (sc->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nref) return;
    const int i = ref[t];
    if (lev[i] != top) {
        Z[i] = 0.0f;
        return;
    }
    auto z1 = [&](int x) { return lev[x] == top ? TEL_OMEGA * (float)((double)r[x] / ap[x]) : 0.0f; };
    const uint64_t m = T.meta[i];
    const int nf = m_ncf(m);
    int nbk[TF];
    float zk[TF];
#pragma unroll
    for (int k = 0; k < TF; k++) nbk[k] = k < nf ? tnb(T, i, k) : T.n;
#pragma unroll
    for (int k = 0; k < TF; k++) zk[k] = nbk[k] < T.n ? z1(nbk[k]) : 0.0f;
    const float zi = z1(i);
    double az = ap[i] * (double)zi;
#pragma unroll
    for (int k = 0; k < TF; k++)
        if (nbk[k] < T.n) az -= apc[k * T.nd + i] * (double)zk[k];
    const double res = (double)r[i] - az;
    Z[i] = zi + TEL_OMEGA * (float)(res / ap[i]);
}
// A damped Jacobi sweep on the finest cells: zout = zin + omega D^-1 (r - A zin)
template <class TB>
__global__ void k_tl_cs_fjac(TelG T, const int *top, int ntop, const TB *r, const double *ap, const double *apc,
                             const float *zin, float *zout, const KryScal *sc) {
    if (sc->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntop) return;
    const int i = top[t];
    const double res = tl_res(T, r, ap, apc, zin, i, T.meta[i]);
    zout[i] = zin[i] + TEL_OMEGA * (float)(res / ap[i]);
}
// Residual r - A Z summed into the level-k squares, and the first sweep from
// zero there.  One thread per entry of kcell; blocks hold whole squares, and
// the thread of a square's first entry sums its entries in order from LDS.
template <class TB>
__global__ void __launch_bounds__(BLOCK) k_tl_cs_restrict(TelG T, TelLev V, const TB *r, const double *ap,
                                                          const double *apc, const float *Z, const KryScal *sc) {
    if (sc->done) return;
    __shared__ double sr[BLOCK];
    const int q0 = V.blk[blockIdx.x], q1 = V.blk[blockIdx.x + 1], q = q0 + threadIdx.x;
    double ri = 0.0;
    int c = 0;
    if (q < q1) {
        const int i = V.kcell[q];
        ri = tl_res(T, r, ap, apc, Z, i, T.meta[i]);
        c = V.kseg[q];
    }
    sr[threadIdx.x] = ri;
    __syncthreads();
    if (q >= q1 || V.kstart[c] != q) return;
    double acc = 0.0;
    for (int p = q; p < V.kstart[c + 1]; p++) acc += sr[p - q0];
    V.r[c] = (float)acc;
    V.z[c] = TEL_OMEGA * (float)acc * V.dinv[c];
}
// The second sweep on level k (from the first, z0) and its correction added
// to the cells inside each square (Z1, and Z2 if given).
__global__ void k_tl_cs_jacp(TelLev V, const float *z0, float *Z1, float *Z2, const KryScal *sc) {
    if (sc->done) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= V.nk) return;
    int nb[4];
    float zn[4];
#pragma unroll
    for (int d = 0; d < 4; d++) nb[d] = V.nbr[d * V.nk + c];
#pragma unroll
    for (int d = 0; d < 4; d++) zn[d] = nb[d] >= 0 ? z0[nb[d]] : 0.0f;
    float az = V.ap[c] * z0[c];
#pragma unroll
    for (int d = 0; d < 4; d++)
        if (nb[d] >= 0) az -= V.a[d * V.nk + c] * zn[d];
    const float zc = z0[c] + TEL_OMEGA * (V.r[c] - az) * V.dinv[c];
    for (int q = V.kstart[c]; q < V.kstart[c + 1]; q++) {
        const int i = V.kcell[q];
        const float v = Z1[i] + zc;
        Z1[i] = v;
        if (Z2) Z2[i] = v;
    }
}
// b of every base cell: the summed residual r - A Z of its cells.  One
// thread per entry of the base-cell-ordered list bcells; blocks hold whole
// base cells, and the thread of a base cell's first entry sums its entries in
// order from LDS.  A block whose entries are all unrefined base cells away
// from refined ones, numbered like their base cells (plain[block] = its
// first base cell, else -1), copies r.
template <class TB>
__global__ void __launch_bounds__(BLOCK) k_tl_mg_restrict(TelG T, const int *blk, const int *plain, const TB *r,
                                                          const double *ap, const double *apc, const float *Z,
                                                          const int *bcells, MgLevel L, float *b0,
                                                          const KryScal *sc) {
    if (sc->done) return;
    __shared__ double sr[BLOCK];
    const int q0 = blk[blockIdx.x], q1 = blk[blockIdx.x + 1], q = q0 + threadIdx.x;
    const int pb = plain[blockIdx.x];
    if (pb >= 0) {
        if (q < q1) {
            const int b = pb + (q - q0);
            b0[L.idx(b % T.nbx, b / T.nbx)] = (float)r[q];
        }
        return;
    }
    int i = 0;
    double ri = 0.0;
    if (q < q1) {
        i = bcells[q];
        ri = tl_res(T, r, ap, apc, Z, i, T.meta[i]);
    }
    sr[threadIdx.x] = ri;
    __syncthreads();
    if (q >= q1) return;
    const int b = T.base[i];
    if (T.bstart[b] != q) return;
    double acc = 0.0;
    for (int p = q; p < T.bstart[b + 1]; p++) acc += sr[p - q0];
    b0[L.idx(b % T.nbx, b / T.nbx)] = (float)acc;
}
// z = Z + P zb on every cell, into both buffers of the up sweeps (a block of
// cells as in k_tl_mg_restrict, plain[block] >= 0, has Z = 0 and is read by
// no sweep: z = zb only).
__global__ void k_tl_cs_post0(TelG T, const int *plain, const float *Z, const float *zb, MgLevel L, float *za,
                              float *zq, const KryScal *sc) {
    if (sc->done) return;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int pb = plain[lb];
    if (pb >= 0) {
        const int i = lb * BLOCK + threadIdx.x;
        if (i < T.n) {
            const int b = pb + threadIdx.x;
            za[i] = zb[L.idx(b % T.nbx, b / T.nbx)];
        }
        return;
    }
    int i;
    if (!tel_thread(T.n, i)) return;
    const int b = T.base[i];
    const float v = Z[i] + zb[L.idx(b % T.nbx, b / T.nbx)];
    za[i] = v;
    zq[i] = v;
}

// Host-side state of a telescoping flow.
struct TelHost {
    TelG T{};
    TelS s{};
    double *gu[2][2], *gv[2][2], *gpp[2][2], *gHu[2][2], *gHv[2][2], *gap[2][2], *gzb[2][2];
    int cu = 0, cv = 0, cpp = 0, cHu = 0, cHv = 0, cap = 0;   // current buffer of each pair
    KrySys<2> Kuv{};
    KrySys<1> Kpp{};
    float *b0 = nullptr;                  // base right-hand side
    float *zp = nullptr;                  // correction of the down sweeps (0 on unrefined cells)
    float *zq = nullptr;                  // second buffer of the up sweeps
    float *apc32 = nullptr;               // FP32 coefficients of the FP32 form's current inner solve (u/v, then pp)
    int *bcells = nullptr;
    int *blkb = nullptr, nblkb = 0;       // blocks of the base restriction (whole base cells)
    int *plainb = nullptr, *plainc = nullptr;   // first base cell of a plain block (see k_tl_mg_restrict), or -1
    TelLev lv[TEL_MAX_LEVELS]{};          // levels 1 .. top-1
    int top = 0;                          // finest level
    int *ref = nullptr, nref = 0;         // refined cells (level >= 1)
    int *topc = nullptr, ntop = 0;        // cells of the finest level
    uint8_t *lev = nullptr;               // [n] level of each cell
    TelTiles tiles{};                     // ponding tiles
    int ntile = 0, tile_max = 0;
    int nblk_n = 0, nblk_nd = 0, nblk_base = 0;
    std::vector<int> gown_h;              // owner of each ghost (host copy)
    int njoint = 0;
};

}  // namespace cms

// GradIO of a ping-pong pair: read buffer cur, write the other; the caller flips cur.
static GradIO tel_io(double *(&g)[2][2], int cur) { return GradIO{g[0][cur], g[1][cur], g[0][cur ^ 1], g[1][cur ^ 1]}; }

#define TLAUNCH(kern, nb, ...) hipLaunchKernelGGL(kern, dim3(nb), dim3(BLOCK), 0, F->st, __VA_ARGS__)

static int tel_sync_flags(cms2d_flow *F) {
    TelHost &H = *F->tel;
    HIPCHK(hipMemcpyAsync(F->h_flags, H.s.flags, 4 * sizeof(int), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    return 0;
}

// flow_wetdry(0) (flow.F90:1231-1539)
static int tel_wetdry(cms2d_flow *F) {
    TelHost &H = *F->tel;
    const TelG &T = H.T;
    TelS &s = H.s;
    TLAUNCH(k_tl_wetdry_h, H.nblk_nd, T, s, F->P);
    for (int it = 0;; it++) {
        TLAUNCH(k_set_int, 1, s.flags, 3, 0);
        TLAUNCH(k_tl_narrow, H.nblk_n, T, s, F->P);
        if (int rc = tel_sync_flags(F)) return rc;
        if (F->h_flags[0] == 0) break;
        if (it > 100000) return set_err(CMS2D_ERR_HIP, "narrow-channel pass did not settle");
    }
    hipLaunchKernelGGL(k_tl_pond_local, dim3(H.ntile), dim3(1024), (size_t)H.tile_max * sizeof(int), F->st, T, s,
                       H.tiles);
    TLAUNCH(k_tl_pond_merge, H.nblk_n, T, s, H.tiles);
    TLAUNCH(k_tl_pond_compress, H.nblk_nd, T, s);
    TLAUNCH(k_tl_pond_mark, (T.nstr + BLOCK - 1) / BLOCK, T, s, T.str_c, T.str_k, T.nstr);
    if (T.nq) TLAUNCH(k_tl_pond_mark, (T.nq + BLOCK - 1) / BLOCK, T, s, T.q_c, T.q_k, T.nq);
    TLAUNCH(k_tl_pond_final, H.nblk_nd, T, s);
    TLAUNCH(k_set_int, 1, s.flags, 3, 0);
    TLAUNCH(k_tl_wetdry_post, H.nblk_n, T, s);
    if (int rc = tel_sync_flags(F)) return rc;
    if (F->h_flags[2] > 0)
        for (int sweep = 0; sweep < 5; sweep++) {
            TLAUNCH(k_tl_d2w_sweep, H.nblk_n, T, s);
            TLAUNCH(k_tl_d2w_commit, H.nblk_n, T, s);
        }
    return 0;
}

// gradients of u and v (der_gradvec_eval), the joint cells' previous values kept
static void tel_grad_uv(cms2d_flow *F) {
    TelHost &H = *F->tel;
    TLAUNCH(k_tl_grad<2>, H.nblk_n, H.T, (const uint8_t *)H.s.iwet, 1, (const double *)H.s.u,
            (const double *)H.s.v, (const double *)nullptr, tel_io(H.gu, H.cu), tel_io(H.gv, H.cv), GradIO{});
    H.cu ^= 1;
    H.cv ^= 1;
}

// flow_grad_interp (flow_imp.F90:1075-1130), fric_eval and flow_eddyvis
static int tel_grad_fric_vis(cms2d_flow *F) {
    TelHost &H = *F->tel;
    const TelG &T = H.T;
    TelS &s = H.s;
    tel_grad_uv(F);
    TLAUNCH(k_tl_pk, H.nblk_n, T, s);
    TLAUNCH(k_tl_dphk, H.nblk_n, T, s, F->P);
    TLAUNCH(k_tl_fric_vis, H.nblk_n, T, s, F->P, (const double *)H.gu[0][H.cu], (const double *)H.gu[1][H.cu],
            (const double *)H.gv[0][H.cv], (const double *)H.gv[1][H.cv]);
    if (T.ng) TLAUNCH(k_tl_vis_ghosts, (T.ng + BLOCK - 1) / BLOCK, T, s);
    TLAUNCH(k_tl_vis_fix, H.nblk_nd, T, s);
    TLAUNCH(k_tl_vis_corner, H.nblk_n, T, s);
    TLAUNCH(k_tl_visk, H.nblk_n, T, s);
    TLAUNCH(k_tl_vis_commit, H.nblk_n, T, s);
    return 0;
}

static unsigned tl_nblk(int n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

// z = M^-1 r of the pp system (r: a Krylov vector of the FP64 solve, or of
// the inner FP32 solve of the refinement form; sc: that solve's scalars)
template <class TB>
static void tel_mg_apply(cms2d_flow *F, const TB *r, float *z, const KryScal *sc) {
    TelHost &H = *F->tel;
    const TelG &T = H.T;
    const double *ap = H.Kpp.ap_w[0], *apc = H.s.apc;
    const MgLevel &L0 = F->mg.L[0];
    hipStream_t st = F->st;
    KTimer *tm = F->tm_active;
    float *Z = H.zp;
    if (H.nref)
        hipLaunchKernelGGL(k_tl_cs_pre2<TB>, dim3(tl_nblk(H.nref)), dim3(BLOCK), 0, st, T, (const int *)H.ref, H.nref,
                           (const uint8_t *)H.lev, H.top, r, ap, apc, Z, sc);
    for (int k = H.top - 1; k >= 1; k--) {
        const TelLev &V = H.lv[k];
        tlaunch(tm, TG_TEL_CS_RESTRICT, k_tl_cs_restrict<TB>, dim3(V.nblk), dim3(BLOCK), 0, st, T, V, r, ap, apc,
                (const float *)Z, sc);
        hipLaunchKernelGGL(k_tl_cs_jacp, dim3(tl_nblk(V.nk)), dim3(BLOCK), 0, st, V, (const float *)V.z, Z,
                           (float *)nullptr, sc);
    }
    tlaunch(tm, TG_TEL_RESTRICT, k_tl_mg_restrict<TB>, dim3(H.nblkb), dim3(BLOCK), 0, st, T, (const int *)H.blkb,
            (const int *)H.plainb, r, ap, apc, (const float *)Z, (const int *)H.bcells, L0, H.b0, sc);
    float *zb = L0.zo;
    mg_vcycle<float>(F, H.b0, zb, sc);
    TLAUNCH(k_tl_cs_post0, H.nblk_n, T, (const int *)H.plainc, (const float *)Z, (const float *)zb, L0, z, H.zq, sc);
    for (int k = 1; k < H.top; k++) {
        const TelLev &V = H.lv[k];
        tlaunch(tm, TG_TEL_CS_RESTRICT, k_tl_cs_restrict<TB>, dim3(V.nblk), dim3(BLOCK), 0, st, T, V, r, ap, apc,
                (const float *)z, sc);
        hipLaunchKernelGGL(k_tl_cs_jacp, dim3(tl_nblk(V.nk)), dim3(BLOCK), 0, st, V, (const float *)V.z, z, H.zq,
                           sc);
    }
    if (H.ntop) {
        tlaunch(tm, TG_TEL_FJAC, k_tl_cs_fjac<TB>, dim3(tl_nblk(H.ntop)), dim3(BLOCK), 0, st, T, (const int *)H.topc,
                H.ntop, r, ap, apc, (const float *)z, H.zq, sc);
        tlaunch(tm, TG_TEL_FJAC, k_tl_cs_fjac<TB>, dim3(tl_nblk(H.ntop)), dim3(BLOCK), 0, st, T, (const int *)H.topc,
                H.ntop, r, ap, apc, (const float *)H.zq, z, sc);
    }
}
struct TelPre {
    cms2d_flow *F;
    void operator()(int which) const {
        TelHost &H = *F->tel;
        if (which == 0) tel_mg_apply<double>(F, H.Kpp.p[0], H.Kpp.zp, H.Kpp.sc);
        else tel_mg_apply<double>(F, H.Kpp.r[0], H.Kpp.zs, H.Kpp.sc);
    }
};

struct TelInnerPre {
    cms2d_flow *F;
    void operator()(int which) const {
        KrySys<1, float> &Ki = F->Kin;
        if (which == 0) tel_mg_apply<float>(F, Ki.p[0], Ki.zp, Ki.sc);
        else tel_mg_apply<float>(F, Ki.r[0], Ki.zs, Ki.sc);
    }
};

static TelOp tel_op(const TelHost &H, const double *a) { return TelOp{H.T.n, H.T.nd, H.T.c2c, H.T.meta, a}; }

// The BiCGStab pairs of the u/v and pp systems, captured once at creation
// (the pp pair's timed-group launches noted for the launch counts).
static int tel_capture_graphs(cms2d_flow *F) {
    if (!F->use_graphs) return 0;
    TelHost &H = *F->tel;
    const KryCtx X = flow_ctx(F, false);
    if (F->precision == CMS2D_PREC_FP32_REFINE) {
        // the inner FP32 pairs of the refinement form (u and v; pp), on the
        // FP32 coefficients converted before each of their solves
        const TelOp32 A32{H.T.n, H.T.nd, H.T.c2c, H.T.meta, H.apc32};
        if (int rc = capture_bicg_pair<TelOp32, 2, 0, float>(inner_ctx(X), F->Kin2, A32, F->kgraph[4])) return rc;
        return capture_bicg_pair<TelOp32, 1, 1, float>(inner_ctx(X), F->Kin, A32, F->kgraph[3], TelInnerPre{F});
    }
    if (int rc = capture_bicg_pair<TelOp, 2, 0, double>(X, H.Kuv, tel_op(H, H.s.am), F->kgraph[1])) return rc;
    const TelOp Ap = tel_op(H, H.s.apc);
    if (F->precond != CMS2D_PRECOND_MG) return capture_bicg_pair<TelOp, 1, 0, double>(X, H.Kpp, Ap, F->kgraph[0]);
    KryCtx Xt = X;
    if (F->stats) {
        Xt.tm = &F->tm;
        F->tm.capture_slot = 0;
        F->tm_active = &F->tm;
    }
    const int rc = capture_bicg_pair<TelOp, 1, 1, double>(Xt, H.Kpp, Ap, F->kgraph[0], TelPre{F});
    F->tm.capture_slot = -1;
    F->tm_active = nullptr;
    return rc;
}

// FP32 copies of the pp coefficients and diagonal for the inner solve
__global__ void k_tl_to32(TelG T, const double *apc, const double *ap, float *apc32, float *ap32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T.nd) return;
    for (int k = 0; k < TF; k++) apc32[k * T.nd + i] = (float)apc[k * T.nd + i];
    ap32[i] = (float)ap[i];
}
// FP32 copies of the momentum coefficients and of each system's diagonal
__global__ void k_tl_uv_to32(TelG T, const double *am, KrySys<2> K, float *a32, KrySys<2, float> Ki) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T.nd) return;
    for (int k = 0; k < TF; k++) a32[k * T.nd + i] = (float)am[k * T.nd + i];
    for (int q = 0; q < 2; q++) {
        Ki.ap_w[q][i] = (float)K.ap_w[q][i];
        Ki.dinv_w[q][i] = (float)K.dinv_w[q][i];
    }
}
// The u and v solves of the FP32 form on the cell list (as uv_solve_refine).
static int tel_uv_solve_refine(cms2d_flow *F, KryCtx &X, const TelOp &A) {
    TelHost &H = *F->tel;
    KrySys<2> &K = H.Kuv;
    KrySys<2, float> &Ki = F->Kin2;
    const int nb = F->nblk_kry;
    hipLaunchKernelGGL((k_bicg_p<TelOp, 2, double>), dim3(1), dim3(KBLOCK), 0, F->st, A, K, 0, nb, X.kmax, X.rtol,
                       X.atol, 1);
    TLAUNCH(k_tl_uv_to32, H.nblk_nd, H.T, (const double *)H.s.am, K, H.apc32, Ki);
    HIPCHK(hipMemcpyAsync(F->h_sc + 1, F->d_sc + 1, 2 * sizeof(KryScal), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    const TelOp32 A32{H.T.n, H.T.nd, H.T.c2c, H.T.meta, H.apc32};
    KryCtx Xi = inner_ctx(X);
    Xi.rtol = REF_RTOL_UV;
    int total[2] = {0, 0};
    for (int sw = 0; sw < REF_SWEEPS && !(F->h_sc[1].done && F->h_sc[2].done); sw++) {
        hipLaunchKernelGGL(k_ref_start2<TelOp32>, dim3(nb), dim3(KBLOCK), 0, F->st, A32, Ki, K);
        if (int rc = run_bicg<TelOp32, 2, 0, float>(Xi, Ki, A32, 4)) return rc;
        for (int q = 0; q < 2; q++) total[q] += F->h_sc[4 + q].final_iters;
        hipLaunchKernelGGL(k_ref_update2<TelOp>, dim3(nb), dim3(KBLOCK), 0, F->st, A, K, Ki);
        hipLaunchKernelGGL(k_ref_check2, dim3(1), dim3(KBLOCK), 0, F->st, K, nb);
        HIPCHK(hipMemcpyAsync(F->h_sc + 1, F->d_sc + 1, 2 * sizeof(KryScal), hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
    }
    F->h_sc[1].final_iters = total[0];
    F->h_sc[2].final_iters = total[1];
    return 0;
}
// The pp solve of the FP32 form on the cell list: FP32 BiCGStab (FP32
// coefficients and vectors, the quadtree V-cycle on FP32 right-hand sides) for
// each correction, x and the residual in FP64, stopped on the FP64 solve's
// own threshold.  Krylov iterations are summed over the sweeps.
static int tel_pp_solve_refine(cms2d_flow *F, KryCtx &X, const TelOp &A) {
    TelHost &H = *F->tel;
    KrySys<1> &K = H.Kpp;
    KrySys<1, float> &Ki = F->Kin;
    const int nb = F->nblk_kry;
    hipLaunchKernelGGL((k_bicg_p<TelOp, 1, double>), dim3(1), dim3(KBLOCK), 0, F->st, A, K, 0, nb, X.kmax, X.rtol,
                       X.atol, 1);
    TLAUNCH(k_tl_to32, H.nblk_nd, H.T, (const double *)H.s.apc, (const double *)K.ap_w[0], H.apc32, Ki.ap_w[0]);
    HIPCHK(hipMemcpyAsync(F->h_sc, F->d_sc, sizeof(KryScal), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    const TelOp32 A32{H.T.n, H.T.nd, H.T.c2c, H.T.meta, H.apc32};
    KryCtx Xi = inner_ctx(X);
    TelInnerPre pre{F};
    int total = 0;
    for (int sw = 0; sw < REF_SWEEPS && !F->h_sc[0].done; sw++) {
        hipLaunchKernelGGL(k_ref_start<TelOp32>, dim3(nb), dim3(KBLOCK), 0, F->st, A32, Ki, (const double *)K.r[0]);
        if (int rc = run_bicg<TelOp32, 1, 1, float>(Xi, Ki, A32, 3, pre)) return rc;
        total += F->h_sc[3].final_iters;
        hipLaunchKernelGGL(k_ref_update<TelOp>, dim3(nb), dim3(KBLOCK), 0, F->st, A, K, (const float *)Ki.x[0]);
        hipLaunchKernelGGL(k_ref_check, dim3(1), dim3(KBLOCK), 0, F->st, K, nb);
        HIPCHK(hipMemcpyAsync(F->h_sc, F->d_sc, sizeof(KryScal), hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
    }
    F->h_sc[0].final_iters = total;
    return 0;
}

// bndriver_vel on the flux string, if any
static void tel_river(cms2d_flow *F) {
    TelHost &H = *F->tel;
    if (F->qside < 0 || H.T.nq == 0) return;
    hipLaunchKernelGGL(k_tl_river, dim3(1), dim3(1024), (size_t)H.T.nq * sizeof(double), F->st, H.T, H.s, F->qflux,
                       F->cmvel, F->qcos, F->qsin, (const double *)(F->q_rating ? F->q_dev : nullptr));
}

static int tel_flow_imp(cms2d_flow *F) {
    TelHost &H = *F->tel;
    const TelG &T = H.T;
    TelS &s = H.s;
    const double timehrs = F->timesecs / 3600.0;
    double w = 0.0;
    for (int k = 0; k < F->ntc; k++) w += F->amp[k] * std::cos(F->speed[k] * timehrs - F->phase[k]);
    if (!F->wse_t.empty()) w = wse_curve_eval(F->wse_t, F->wse_v, F->wse_order, timehrs, F->wse_inc);
    // the flux string's inflow (bnd_eval, boundaries.F90:4078-4094), before flow_wetdry
    if (F->qside >= 0 && F->q_rating) {
        TLAUNCH(k_tl_river_rating, 1, T, s, (int)F->q_t.size(), F->ramp, F->q_dev);
    } else if (F->qside >= 0) {
        const double q = F->q_t.empty() ? F->q_const : wse_curve_eval(F->q_t, F->q_v, F->q_order, timehrs, F->q_inc);
        F->qflux = F->ramp * q;
    }
    if (int rc = tel_wetdry(F)) return rc;
    F->P.tauwx = F->P.tauwy = 0.0;
    if (F->windconst) {
        const double ws = std::sqrt(F->wndx * F->wndx + F->wndy * F->wndy);
        const double val = F->ramp * wind_drag_hsu(ws) * F4(0.00120) * ws;
        F->P.tauwx = val * F->wndx;
        F->P.tauwy = val * F->wndy;
    }
    TLAUNCH(k_tl_srcsnk, H.nblk_n, T, s, F->P);
    const TelOp Am = tel_op(H, s.am), Ap = tel_op(H, s.apc);
    for (int n = 0; n < 3; n++) F->kry[n] = 0;
    for (F->niter = 1; F->niter <= F->maxit; F->niter++) {
        tel_river(F);
        TLAUNCH(k_tl_mom, H.nblk_n, T, s, F->P);
        if (H.njoint && F->P.ndsch != 0)
            TLAUNCH(k_tl_defcor_skew, (H.njoint + BLOCK - 1) / BLOCK, T, s, (const double *)H.gu[0][H.cu],
                    (const double *)H.gu[1][H.cu], (const double *)H.gv[0][H.cv], (const double *)H.gv[1][H.cv]);
        TLAUNCH(k_tl_bound_uv, H.nblk_n, T, s, F->P);
        hipLaunchKernelGGL(k_tl_uv_prologue, dim3(F->nblk_kry), dim3(KBLOCK), 0, F->st, T, s, H.Kuv, Am);
        {
            KryCtx X = flow_ctx(F, false);
            if (F->precision == CMS2D_PREC_FP32_REFINE) {
                if (int rc = tel_uv_solve_refine(F, X, Am)) return rc;
            } else if (int rc = run_bicg(X, H.Kuv, Am, 1)) {
                return rc;
            }
        }
        TLAUNCH(k_tl_pp_prep, H.nblk_n, T, s);
        if (T.skewcor) {
            TLAUNCH(k_tl_grad<3>, H.nblk_n, T, (const uint8_t *)s.iwet, 1, (const double *)s.Hu, (const double *)s.Hv,
                    (const double *)s.apuareap, tel_io(H.gHu, H.cHu), tel_io(H.gHv, H.cHv), tel_io(H.gap, H.cap));
            H.cHu ^= 1;
            H.cHv ^= 1;
            H.cap ^= 1;
        }
        TLAUNCH(k_tl_pp_coef_flux, H.nblk_n, T, s, (const double *)H.gHu[1][H.cHu], (const double *)H.gHv[0][H.cHv],
                (const double *)H.gap[0][H.cap], (const double *)H.gap[1][H.cap]);
        if (T.nstr > 0)
            hipLaunchKernelGGL(k_tl_wse_adjust, dim3(1), dim3(1024), T.nstr * sizeof(double), F->st, T, s, F->P, w,
                               F->ramp, (int)!(F->wse_adjust_off && !F->wse_t.empty()));
        hipLaunchKernelGGL(k_tl_pp_src, dim3(F->nblk_kry), dim3(KBLOCK), 0, F->st, T, s, F->P, H.Kpp, Ap,
                           (const double *)H.gpp[0][H.cpp], (const double *)H.gpp[1][H.cpp]);
        {
            KryCtx X = flow_ctx(F, true);
            if (F->precond == CMS2D_PRECOND_MG) {
                TLAUNCH(k_tl_mg_setup0, H.nblkb, T, (const int *)H.blkb, (const double *)s.apc,
                        (const double *)H.Kpp.ap_w[0], (const int *)H.bcells, F->mg.L[0]);
                for (int k = 1; k < H.top; k++)
                    TLAUNCH(k_tl_cs_setup, tl_nblk(H.lv[k].nk), T, H.lv[k], (const double *)s.apc,
                            (const double *)H.Kpp.ap_w[0]);
                for (int l = 0; l + 1 < F->mg.nlev; l++)
                    LAUNCH(k_mg_coarsen, mg_nblk(F->mg.L[l + 1]), F->mg.L[l], F->mg.L[l + 1]);
                if (F->precision == CMS2D_PREC_FP32_REFINE) {
                    if (int rc = tel_pp_solve_refine(F, X, Ap)) return rc;
                } else {
                    F->tm_active = X.tm;
                    int rc = run_bicg<TelOp, 1, 1, double>(X, H.Kpp, Ap, 0, TelPre{F});
                    F->tm_active = nullptr;
                    if (rc) return rc;
                }
            } else {
                if (int rc = run_bicg(X, H.Kpp, Ap, 0)) return rc;
            }
        }
        if (T.ng) TLAUNCH(k_tl_ghost_pp, (T.ng + BLOCK - 1) / BLOCK, T, s);
        TLAUNCH(k_tl_grad<1>, H.nblk_n, T, (const uint8_t *)s.iwet, 1, (const double *)s.pp, (const double *)nullptr,
                (const double *)nullptr, tel_io(H.gpp, H.cpp), GradIO{}, GradIO{});
        H.cpp ^= 1;
        TLAUNCH(k_tl_pvfc1, H.nblk_n, T, s, (const double *)H.gpp[0][H.cpp], (const double *)H.gpp[1][H.cpp]);
        TLAUNCH(k_tl_pvfc2, H.nblk_nd, T, s, F->P);
        if (T.ng) TLAUNCH(k_tl_ghost_p, (T.ng + BLOCK - 1) / BLOCK, T, s, F->P);
        if (int rc = tel_grad_fric_vis(F)) return rc;
        tel_river(F);       // bndriver_vel after flow_grad_interp (flow_imp.F90:236-237)
        for (int n = 0; n < 3; n++) {
            F->rmom[n] = F->h_sc[n].rmom;
            F->kry[n] += F->h_sc[n].final_iters;
        }
        check_conv(F);
        if (F->iconv == 0) return set_err(CMS2D_ERR_DIVERGED, "Model is divergent (check_conv iconv=0)");
        if (F->iconv == 2) break;
    }
    if (F->niter > F->maxit) F->niter = F->maxit;
    TLAUNCH(k_tl_velface, H.nblk_n, T, s, F->P);
    return 0;
}

// Initial state (flow.F90:688-707, prestart.f90:309-329)
static int tel_init_state(cms2d_flow *F) {
    TelHost &H = *F->tel;
    TelS &s = H.s;
    const size_t nd = H.T.nd;
    if (int rc = tel_wetdry(F)) return rc;
    HIPCHK(hipMemcpyAsync(s.u1, s.u, nd * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    HIPCHK(hipMemcpyAsync(s.v1, s.v, nd * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    HIPCHK(hipMemcpyAsync(s.p1, s.p, nd * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    HIPCHK(hipMemcpyAsync(s.h1, s.h, nd * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    if (int rc = tel_wetdry(F)) return rc;
    HIPCHK(hipMemcpyAsync(s.iwet1, s.iwet, nd, hipMemcpyDeviceToDevice, F->st));
    tel_grad_uv(F);
    TLAUNCH(k_tl_pk, H.nblk_n, H.T, s);
    TLAUNCH(k_tl_dphk, H.nblk_n, H.T, s, F->P);
    TLAUNCH(k_tl_convflux, H.nblk_n, H.T, s);
    HIPCHK(hipMemcpyAsync(s.flux1, s.flux, TF * nd * sizeof(double), hipMemcpyDeviceToDevice, F->st));
    // fric_eval and flow_eddyvis without a new gradient
    TLAUNCH(k_tl_fric_vis, H.nblk_n, H.T, s, F->P, (const double *)H.gu[0][H.cu], (const double *)H.gu[1][H.cu],
            (const double *)H.gv[0][H.cv], (const double *)H.gv[1][H.cv]);
    if (H.T.ng) TLAUNCH(k_tl_vis_ghosts, (H.T.ng + BLOCK - 1) / BLOCK, H.T, s);
    TLAUNCH(k_tl_vis_fix, H.nblk_nd, H.T, s);
    TLAUNCH(k_tl_vis_corner, H.nblk_n, H.T, s);
    TLAUNCH(k_tl_visk, H.nblk_n, H.T, s);
    TLAUNCH(k_tl_vis_commit, H.nblk_n, H.T, s);
    HIPCHK(hipStreamSynchronize(F->st));
    return 0;
}

// The topology of read_tel (geospatial.F90:403-795) and the connectivity part
// of geo_init_cart (:1446-1552) from the .tel arrays, on the host: faces in the
// order of the file's neighbour slots, one ghost per side without a
// neighbour (placed one cell size beyond it, the cell's bed), llec2llec, the
// no-repeat owner of each face, joint cells.
struct TelTopo {
    int n = 0, nd = 0;
    std::vector<int> ncf, c2c, dir, llec, own, joint;
    std::vector<double> x, y, dx, dy;
    std::vector<int> gown, gk;
};
static int tel_topology(const cms2d_tel_grid *tg, TelTopo &t) {
    const int n = tg->ncells;
    t.n = n;
    int nd = n;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 8; k += 2) {
            const int a = tg->nb[8 * i + k], b = tg->nb[8 * i + k + 1];
            if (a < 0 || a > n || b < 0 || b > n) return set_err(CMS2D_ERR_ARG, "neighbour slot out of range");
            if (a == 0 && b == 0) nd++;
        }
    t.nd = nd;
    t.ncf.assign(nd, 0);
    t.c2c.assign((size_t)TF * nd, -1);
    t.dir.assign((size_t)TF * nd, -1);
    t.llec.assign((size_t)TF * nd, -1);
    t.own.assign((size_t)TF * nd, 0);
    t.x.assign(nd, 0.0); t.y.assign(nd, 0.0); t.dx.assign(nd, 0.0); t.dy.assign(nd, 0.0);
    for (int i = 0; i < n; i++) {
        t.x[i] = tg->geom[4 * i]; t.y[i] = tg->geom[4 * i + 1];
        t.dx[i] = tg->geom[4 * i + 2]; t.dy[i] = tg->geom[4 * i + 3];
        if (!(t.dx[i] > 0) || !(t.dy[i] > 0)) return set_err(CMS2D_ERR_ARG, "cell size must be positive");
    }
    static const int kopp[4] = {KS, KW, KN, KE};
    int g = n;
    for (int i = 0; i < n; i++) {
        int k = 0;
        for (int j = 0; j < 8; j++) {
            const int d = j / 2, nck = tg->nb[8 * i + j];
            const bool open = (j % 2 == 0) && nck == 0 && tg->nb[8 * i + j + 1] == 0;
            if (!open && nck == 0) continue;
            if (k == TF) return set_err(CMS2D_ERR_ARG, "a cell with more than six faces");
            if (open) {
                t.c2c[(size_t)k * nd + i] = g;
                t.ncf[g] = 4;
                t.c2c[(size_t)kopp[d] * nd + g] = i;
                t.x[g] = t.x[i] + (d == KE ? t.dx[i] : d == KW ? -t.dx[i] : 0.0);
                t.y[g] = t.y[i] + (d == KN ? t.dy[i] : d == KS ? -t.dy[i] : 0.0);
                t.dx[g] = t.dx[i]; t.dy[g] = t.dy[i];
                t.gown.push_back(i);
                t.gk.push_back(k);
                g++;
            } else {
                t.c2c[(size_t)k * nd + i] = nck - 1;
            }
            t.dir[(size_t)k * nd + i] = d;
            k++;
        }
        t.ncf[i] = k;
    }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < t.ncf[i]; k++) {
            const int nck = t.c2c[(size_t)k * nd + i];
            int l = -1;
            if (nck < n) {
                for (int kk = 0; kk < t.ncf[nck]; kk++)
                    if (t.c2c[(size_t)kk * nd + nck] == i) l = kk;
            } else {
                l = kopp[t.dir[(size_t)k * nd + i]];
            }
            if (l < 0) return set_err(CMS2D_ERR_ARG, "neighbour slots are not reciprocal");
            t.llec[(size_t)k * nd + i] = l;
        }
    std::vector<char> seen((size_t)TF * nd, 0);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < t.ncf[i]; k++) {
            const size_t f = (size_t)k * nd + i;
            if (seen[f]) continue;
            const int nck = t.c2c[f];
            seen[f] = 1;
            seen[(size_t)t.llec[f] * nd + nck] = 1;
            t.own[f] = 1;
        }
    t.joint.assign(n, 0);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < t.ncf[i]; k++) {
            const size_t f = (size_t)k * nd + i;
            const int nck = t.c2c[f];
            if (nck >= n) continue;
            const int d = t.dir[f];
            const double r = (d == KE || d == KW) ? t.dy[i] / t.dy[nck] : t.dx[i] / t.dx[nck];
            if (r > F4(1.2) || r < F4(0.8)) { t.joint[i] = 1; t.joint[nck] = 1; }
        }
    return 0;
}

// Device copy of n elements (an empty array still gets one element).
template <typename T>
static int tel_upload(cms2d_flow *F, T **dst, const T *src, size_t n) {
    if (int rc = dalloc(F, dst, std::max<size_t>(n, 1))) return rc;
    if (n) HIPCHK(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, F->st));
    return 0;
}

static void tel_destroy(cms2d_flow *F) {
    delete F->tel;
    F->tel = nullptr;
}

// Blocks of at most BLOCK entries over consecutive segments [st[s], st[s+1]),
// each block holding whole segments (no segment is longer than BLOCK).
static std::vector<int> tel_seg_blocks(const std::vector<int> &st, int nseg) {
    std::vector<int> blk{0};
    for (int s = 0; s < nseg; s++)
        if (st[s + 1] - blk.back() > BLOCK) blk.push_back(st[s]);
    blk.push_back(st[nseg]);
    return blk;
}

// Host copy of a level of the refined region (see TelLev).
struct TelLevHost {
    int nk = 0;
    std::vector<int> kstart, kcell, kseg, blk, nbr, cstart, cf, ge, anc;
    std::vector<float> cw;
    std::vector<uint8_t> cs;
};

// Levels of the cells (base cells split into 2^L x 2^L squares) and the
// squares of levels 1 .. top-1.
static int tel_levels(const TelTopo &t, const cms2d_flow_desc *d, std::vector<uint8_t> &lev, int &top,
                      std::vector<TelLevHost> &lv) {
    const int n = t.n, nd = t.nd;
    lev.assign(n, 0);
    top = 0;
    for (int i = 0; i < n; i++) {
        const int L = (int)std::lround(std::log2(d->dx / t.dx[i]));
        if (L < 0 || L >= TEL_MAX_LEVELS || std::fabs(t.dx[i] * (1 << L) - d->dx) > 1e-6 * d->dx ||
            std::fabs(t.dy[i] * (1 << L) - d->dy) > 1e-6 * d->dy)
            return set_err(CMS2D_ERR_ARG, "cell " + std::to_string(i + 1) + " is not a 2^-L part (L < " +
                                              std::to_string(TEL_MAX_LEVELS) + ") of its base cell");
        lev[i] = (uint8_t)L;
        top = std::max(top, L);
    }
    lv.assign(std::max(top, 1), TelLevHost{});
    for (int k = 1; k < top; k++) {
        TelLevHost &V = lv[k];
        const double hx = d->dx / (1 << k), hy = d->dy / (1 << k);
        const long long W = (long long)d->nx << k;
        std::unordered_map<long long, int> id;
        std::vector<int> ci, cj;
        V.anc.assign(n, -1);
        for (int i = 0; i < n; i++) {
            if (lev[i] < k) continue;
            const long long I = (long long)std::floor(t.x[i] / hx), J = (long long)std::floor(t.y[i] / hy);
            auto it = id.find(J * W + I);
            if (it == id.end()) {
                it = id.emplace(J * W + I, V.nk++).first;
                ci.push_back((int)I);
                cj.push_back((int)J);
            }
            V.anc[i] = it->second;
        }
        const int nk = V.nk;
        V.kstart.assign(nk + 1, 0);
        for (int i = 0; i < n; i++)
            if (V.anc[i] >= 0) V.kstart[V.anc[i] + 1]++;
        for (int c = 0; c < nk; c++) V.kstart[c + 1] += V.kstart[c];
        V.kcell.assign(V.kstart[nk], 0);
        V.kseg.assign(V.kstart[nk], 0);
        {
            std::vector<int> fill(V.kstart.begin(), V.kstart.end() - 1);
            for (int i = 0; i < n; i++)
                if (V.anc[i] >= 0) {
                    V.kseg[fill[V.anc[i]]] = V.anc[i];
                    V.kcell[fill[V.anc[i]]++] = i;
                    V.ge.push_back(i);
                }
        }
        V.blk = tel_seg_blocks(V.kstart, nk);
        V.nbr.assign((size_t)4 * nk, -1);
        static const int di[4] = {0, 1, 0, -1}, dj[4] = {1, 0, -1, 0};
        for (int c = 0; c < nk; c++)
            for (int dd = 0; dd < 4; dd++) {
                auto it = id.find((long long)(cj[c] + dj[dd]) * W + (ci[c] + di[dd]));
                if (ci[c] + di[dd] >= 0 && ci[c] + di[dd] < W && it != id.end()) V.nbr[(size_t)dd * nk + c] = it->second;
            }
        V.cstart.assign(nk + 1, 0);
        for (int c = 0; c < nk; c++) {
            for (int q = V.kstart[c]; q < V.kstart[c + 1]; q++) {
                const int i = V.kcell[q];
                for (int kf = 0; kf < t.ncf[i]; kf++) {
                    const size_t f = (size_t)kf * nd + i;
                    const int nb = t.c2c[f];
                    if (nb >= n || (lev[nb] >= k && V.anc[nb] == c)) continue;
                    const int dd = t.dir[f];
                    const bool ew = dd == KE || dd == KW;
                    const double dnf = ew ? std::fabs(t.x[nb] - t.x[i]) : std::fabs(t.y[nb] - t.y[i]);
                    double dnk;
                    uint8_t side;
                    if (lev[nb] >= k) {
                        if (V.nbr[(size_t)dd * nk + c] != V.anc[nb])
                            return set_err(CMS2D_ERR_ARG, "cell " + std::to_string(i + 1) +
                                                              ": a neighbour outside the adjacent square");
                        dnk = ew ? hx : hy;
                        side = (uint8_t)dd;
                    } else {
                        const double xc = (ci[c] + 0.5) * hx, yc = (cj[c] + 0.5) * hy;
                        dnk = ew ? std::fabs(t.x[nb] - xc) : std::fabs(t.y[nb] - yc);
                        side = 4;
                    }
                    V.cf.push_back((int)f);
                    V.cw.push_back((float)(dnf / dnk));
                    V.cs.push_back(side);
                }
            }
            V.cstart[c + 1] = (int)V.cf.size();
        }
    }
    return 0;
}

// Checks of the description that need no device.
static int tel_check(const cms2d_flow_desc *d, const cms2d_tel_grid *tg) {
    if (!tg->geom || !tg->nb || tg->ncells < 1 || (tg->nbnd > 0 && !tg->bnd_cells))
        return set_err(CMS2D_ERR_ARG, "invalid telescoping grid");
    if (d->nx < 1 || d->ny < 1 || !(d->dx > 0) || !(d->dy > 0) || !d->zb || !d->mann)
        return set_err(CMS2D_ERR_ARG, "invalid flow description");
    if (d->ndsch >= 5 && d->ndsch <= 9)
        return set_err(CMS2D_ERR_ARG, "the NVD schemes are built for Cartesian grids only");
    if ((d->precision != CMS2D_PREC_FP64 && d->precision != CMS2D_PREC_FP32_REFINE) ||
        d->solver != CMS2D_SOLVER_BICGSTAB || d->dtvar)
        return set_err(CMS2D_ERR_ARG, "telescoping grids run BiCGStab (FP64 or the FP32 form) with a fixed time step");
    if (d->precision == CMS2D_PREC_FP32_REFINE && d->precond != CMS2D_PRECOND_MG)
        return set_err(CMS2D_ERR_ARG, "the FP32 refinement form uses the multigrid preconditioner");
    if (d->precond != CMS2D_PRECOND_MG && d->precond != CMS2D_PRECOND_JACOBI)
        return set_err(CMS2D_ERR_ARG, "unknown preconditioner");
    if (d->serial_order) return set_err(CMS2D_ERR_ARG, "serial_order is built for Cartesian grids only");
    int nopen = 0, curve = 0, nflux = 0;
    for (int k = 0; k < 4; k++) {
        if (d->side_type[k] != CMS2D_BND_WALL && d->side_type[k] != CMS2D_BND_WSE_HARMONIC &&
            d->side_type[k] != CMS2D_BND_WSE_CURVE && d->side_type[k] != CMS2D_BND_FLUX)
            return set_err(CMS2D_ERR_ARG, "unknown boundary type");
        nflux += d->side_type[k] == CMS2D_BND_FLUX;
        nopen += d->side_type[k] != CMS2D_BND_WALL && d->side_type[k] != CMS2D_BND_FLUX;
        curve += d->side_type[k] == CMS2D_BND_WSE_CURVE;
    }
    if (nflux > 1) return set_err(CMS2D_ERR_ARG, "one flux side at most");
    if ((nflux == 1) != (tg->nflux > 0) || (tg->nflux > 0 && !tg->flux_cells))
        return set_err(CMS2D_ERR_ARG, "the flux side needs its cellstring");
    if (tg->nflux == 1 || tg->nflux > 8192) return set_err(CMS2D_ERR_ARG, "the flux cellstring needs 2..8192 cells");
    if (nflux) {
        if (d->friction != CMS2D_FRIC_MANNING) return set_err(CMS2D_ERR_ARG, "a flux side needs Manning friction");
        if (d->flux_ntimes == 1 || d->flux_ntimes < 0 || d->flux_order < 0 || d->flux_order > 3 ||
            !(d->flux_conveyance >= 0.0) || !std::isfinite(d->flux_conveyance) ||
            (d->flux_ntimes == 0 && !std::isfinite(d->flux_q)) || (d->flux_ntimes > 0 && (!d->flux_times || !d->flux_values)))
            return set_err(CMS2D_ERR_ARG, "flux curve or conveyance invalid");
        for (int i = 0; i < d->flux_ntimes; i++)
            if (!std::isfinite(d->flux_times[i]) || !std::isfinite(d->flux_values[i]) ||
                (i > 0 && !(d->flux_times[i] > d->flux_times[i - 1])))
                return set_err(CMS2D_ERR_ARG, "flux curve times must increase, values be finite");
    }
    if (nopen > 1) return set_err(CMS2D_ERR_ARG, "one water-level side at most (the others are walls)");
    if ((nopen == 1) != (tg->nbnd > 0)) return set_err(CMS2D_ERR_ARG, "the water-level side needs its cellstring");
    if (tg->nbnd == 1 || tg->nbnd > 8192) return set_err(CMS2D_ERR_ARG, "the cellstring needs 2..8192 cells");
    if (curve) {
        if (d->wse_ntimes < 1 || !d->wse_times || !d->wse_values || d->wse_order < 0 || d->wse_order > 3)
            return set_err(CMS2D_ERR_ARG, "water-level curve missing or invalid");
        for (int i = 0; i < d->wse_ntimes; i++)
            if (!std::isfinite(d->wse_times[i]) || !std::isfinite(d->wse_values[i]) ||
                (i > 0 && !(d->wse_times[i] > d->wse_times[i - 1])))
                return set_err(CMS2D_ERR_ARG, "water-level curve times must increase, values be finite");
        if (d->wse_gradx != 0.0 || d->wse_grady != 0.0)
            return set_err(CMS2D_ERR_ARG, "water-level gradients are built for Cartesian grids only");
    }
    TelTopo t;
    return tel_topology(tg, t);
}

static int tel_create(const cms2d_flow_desc *d, const cms2d_tel_grid *tg, cms2d_flow *F) {
    int curve = 0;
    for (int k = 0; k < 4; k++) curve += d->side_type[k] == CMS2D_BND_WSE_CURVE;
    TelTopo t;
    if (int rc = tel_topology(tg, t)) return rc;
    const int n = t.n, nd = t.nd, ng = nd - n;
    // the water-level and flux strings: cellstr_bnd (bnd_lib.f90:215-316)
    auto string_faces = [&](const int *cells, int ncs, std::vector<int> &sc, std::vector<int> &sk) -> int {
        for (int j = 0; j < ncs; j++) {
            const int id = cells[j];
            if (id < 0 || id >= n) return set_err(CMS2D_ERR_ARG, "cellstring cell out of range");
            const int ref = j == 0 ? cells[1] : j == ncs - 1 ? cells[ncs - 2] : id;
            if (ref < 0 || ref >= n) return set_err(CMS2D_ERR_ARG, "cellstring cell out of range");
            for (int k = 0; k < t.ncf[ref]; k++) {
                if (t.c2c[(size_t)k * nd + ref] < n) continue;
                int kf = k;
                if (ref != id) {
                    kf = -1;
                    for (int kk = 0; kk < t.ncf[id]; kk++)
                        if (t.dir[(size_t)kk * nd + id] == t.dir[(size_t)k * nd + ref]) { kf = kk; break; }
                }
                if (kf < 0 || t.c2c[(size_t)kf * nd + id] < n)
                    return set_err(CMS2D_ERR_ARG, "cellstring cell without an edge face in the string's direction");
                sc.push_back(id);
                sk.push_back(kf);
            }
        }
        if (sc.size() > 8192) return set_err(CMS2D_ERR_ARG, "the cellstring has more than 8192 faces");
        return 0;
    };
    std::vector<int> str_c, str_k, q_c, q_k;
    if (int rc = string_faces(tg->bnd_cells, tg->nbnd, str_c, str_k)) return rc;
    if (int rc = string_faces(tg->flux_cells, tg->nflux, q_c, q_k)) return rc;
    std::vector<uint8_t> gwall(ng, 1), strmask(n, 0), qmask(n, 0);
    for (size_t j = 0; j < str_c.size(); j++) {
        const int gc = t.c2c[(size_t)str_k[j] * nd + str_c[j]];
        if (!gwall[gc - n]) return set_err(CMS2D_ERR_ARG, "a cellstring face listed twice");
        gwall[gc - n] = 0;
        strmask[str_c[j]] |= (uint8_t)(1u << str_k[j]);
    }
    for (size_t j = 0; j < q_c.size(); j++) {
        const int gc = t.c2c[(size_t)q_k[j] * nd + q_c[j]];
        if (gwall[gc - n] != 1) return set_err(CMS2D_ERR_ARG, "a cellstring face listed twice");
        gwall[gc - n] = 2;
        qmask[q_c[j]] |= (uint8_t)(1u << q_k[j]);
    }
    for (int k = 0; k < 4; k++) {
        if (d->side_type[k] != CMS2D_BND_FLUX) continue;
        // Q_str as on Cartesian grids: inflow along the side's inward normal or
        // INFLOW_ANGLE, conveyance, curve or rating table
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
            angle = 90.0 - d->flux_angle - 0.0;
            if (angle < 0.0) angle += 360.0;
            else if (angle > 360.0) angle -= 360.0;
        }
        F->qcos = std::cos(angle * deg2rad);
        F->qsin = std::sin(angle * deg2rad);
    }
    // base cells
    std::vector<int> base(n), cnt((size_t)d->nx * d->ny + 1, 0);
    for (int i = 0; i < n; i++) {
        const int bi = (int)std::floor(t.x[i] / d->dx), bj = (int)std::floor(t.y[i] / d->dy);
        if (bi < 0 || bi >= d->nx || bj < 0 || bj >= d->ny) return set_err(CMS2D_ERR_ARG, "cell outside the base grid");
        base[i] = bj * d->nx + bi;
        cnt[base[i] + 1]++;
    }
    for (size_t b = 0; b < (size_t)d->nx * d->ny; b++) {
        if (cnt[b + 1] == 0) return set_err(CMS2D_ERR_ARG, "a base cell without cells");
        cnt[b + 1] += cnt[b];
    }
    std::vector<int> bcells(n), fill(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; i++) bcells[fill[base[i]]++] = i;
    for (size_t b = 0; b < (size_t)d->nx * d->ny; b++)
        if (cnt[b + 1] - cnt[b] > BLOCK) return set_err(CMS2D_ERR_ARG, "a base cell with more than 256 cells");
    const std::vector<int> blkb = tel_seg_blocks(cnt, d->nx * d->ny);
    std::vector<uint8_t> lev;
    int top = 0;
    std::vector<TelLevHost> lvh;
    if (int rc = tel_levels(t, d, lev, top, lvh)) return rc;
    std::vector<int> refc, topc;
    for (int i = 0; i < n; i++) {
        if (lev[i] >= 1) refc.push_back(i);
        if (top >= 1 && lev[i] == top) topc.push_back(i);
    }
    std::vector<uint64_t> meta(n, 0);
    std::vector<int> jl;
    for (int i = 0; i < n; i++) {
        uint64_t m = (uint64_t)t.ncf[i];
        for (int k = 0; k < t.ncf[i]; k++) {
            const size_t f = (size_t)k * nd + i;
            m |= (uint64_t)t.dir[f] << (3 + 2 * k);
            m |= (uint64_t)t.llec[f] << (15 + 3 * k);
            m |= (uint64_t)t.own[f] << (33 + k);
        }
        if (t.joint[i]) { m |= 1ull << 39; jl.push_back(i); }
        const bool sm = cnt[base[i] + 1] - cnt[base[i]] > 1;
        if (sm) m |= 1ull << 40;
        bool near = sm;
        for (int k = 0; k < t.ncf[i]; k++) {
            const int nb = t.c2c[(size_t)k * nd + i];
            if (nb < n && cnt[base[nb] + 1] - cnt[base[nb]] > 1) near = true;
        }
        if (near) m |= 1ull << 41;
        m |= (uint64_t)strmask[i] << 42;
        m |= (uint64_t)qmask[i] << 48;
        meta[i] = m;
    }

    F->dev = d->device;
    HIPCHK(hipSetDevice(F->dev));
    HIPCHK(hipStreamCreateWithFlags(&F->st, hipStreamNonBlocking));
    F->tel = new (std::nothrow) TelHost();
    if (!F->tel) return set_err(CMS2D_ERR_ARG, "out of host memory");
    TelHost &H = *F->tel;
    TelG &T = H.T;
    T.n = n; T.nd = nd; T.ng = ng;
    T.nbx = d->nx; T.nby = d->ny; T.dxb = d->dx; T.dyb = d->dy;
    T.njoint = (int)jl.size();
    H.njoint = T.njoint;
    T.skewcor = T.njoint > 0;
    T.nstr = (int)str_c.size();
    T.nq = (int)q_c.size();
    {
        double asum = 0.0;
        for (int i = 0; i < n; i++) asum += t.dx[i] * t.dy[i];
        T.areaavg = asum / (double)(float)n;
    }
    int *c2c, *gown, *gk, *bas, *bst, *jnt, *sc_, *sk_, *qc_ = nullptr, *qk_ = nullptr;
    uint64_t *met;
    double *xx, *yy, *ddx, *ddy;
    uint8_t *gw;
    if (int rc = tel_upload(F, &c2c, t.c2c.data(), t.c2c.size())) return rc;
    if (int rc = tel_upload(F, &met, meta.data(), meta.size())) return rc;
    if (int rc = tel_upload(F, &xx, t.x.data(), (size_t)nd)) return rc;
    if (int rc = tel_upload(F, &yy, t.y.data(), (size_t)nd)) return rc;
    if (int rc = tel_upload(F, &ddx, t.dx.data(), (size_t)nd)) return rc;
    if (int rc = tel_upload(F, &ddy, t.dy.data(), (size_t)nd)) return rc;
    if (int rc = tel_upload(F, &gown, t.gown.data(), t.gown.size())) return rc;
    if (int rc = tel_upload(F, &gk, t.gk.data(), t.gk.size())) return rc;
    if (int rc = tel_upload(F, &gw, gwall.data(), gwall.size())) return rc;
    if (int rc = tel_upload(F, &bas, base.data(), (size_t)n)) return rc;
    if (int rc = tel_upload(F, &bst, cnt.data(), cnt.size())) return rc;
    if (int rc = tel_upload(F, &H.bcells, bcells.data(), (size_t)n)) return rc;
    if (int rc = tel_upload(F, &H.blkb, blkb.data(), blkb.size())) return rc;
    H.nblkb = (int)blkb.size() - 1;
    {
        // plain blocks: every entry an unrefined base cell away from refined
        // cells, the cells numbered like their base cells
        auto plain_cell = [&](int q) {
            const int b = base[q];
            return bcells[q] == q && cnt[b + 1] - cnt[b] == 1 && !((meta[q] >> 41) & 1u);
        };
        auto plain_run = [&](int a, int e) -> int {
            if (a >= e) return -1;
            for (int q = a; q < e; q++)
                if (!plain_cell(q) || base[q] != base[a] + (q - a)) return -1;
            return base[a];
        };
        const int nbc = (n + BLOCK - 1) / BLOCK;
        std::vector<int> pb(H.nblkb), pc(nbc);
        for (int k = 0; k < H.nblkb; k++) pb[k] = plain_run(blkb[k], blkb[k + 1]);
        for (int k = 0; k < nbc; k++) pc[k] = plain_run(k * BLOCK, std::min(n, (k + 1) * BLOCK));
        if (int rc = tel_upload(F, &H.plainb, pb.data(), pb.size())) return rc;
        if (int rc = tel_upload(F, &H.plainc, pc.data(), pc.size())) return rc;
    }
    if (int rc = tel_upload(F, &H.lev, lev.data(), lev.size())) return rc;
    {
        // ponding tiles: squares of base cells, the side halved until a tile
        // holds at most 16384 cells (its labels fit in LDS)
        int side = 32, ntx = 0, nty = 0, mx = 0;
        std::vector<int> tid(nd), tcnt;
        for (;; side /= 2) {
            ntx = (d->nx + side - 1) / side;
            nty = (d->ny + side - 1) / side;
            tcnt.assign((size_t)ntx * nty + 1, 0);
            for (int i = 0; i < nd; i++) {
                const int b = base[i < n ? i : t.gown[i - n]];
                tid[i] = (b / d->nx / side) * ntx + (b % d->nx) / side;
                tcnt[tid[i] + 1]++;
            }
            mx = *std::max_element(tcnt.begin(), tcnt.end());
            if (mx <= 16384 || side == 1) break;
        }
        if (mx > 16384) return set_err(CMS2D_ERR_ARG, "a base cell with more than 16384 cells");
        const int nt = ntx * nty;
        for (int q = 0; q < nt; q++) tcnt[q + 1] += tcnt[q];
        std::vector<int> tcell(nd), loc(nd), fill(tcnt.begin(), tcnt.end() - 1);
        for (int i = 0; i < nd; i++) {
            loc[i] = fill[tid[i]] - tcnt[tid[i]];
            tcell[fill[tid[i]]++] = i;
        }
        int *a1, *a2, *a3, *a4;
        if (int rc = tel_upload(F, &a1, tcnt.data(), tcnt.size())) return rc;
        if (int rc = tel_upload(F, &a2, tcell.data(), tcell.size())) return rc;
        if (int rc = tel_upload(F, &a3, tid.data(), tid.size())) return rc;
        if (int rc = tel_upload(F, &a4, loc.data(), loc.size())) return rc;
        H.tiles = TelTiles{a1, a2, a3, a4};
        H.ntile = nt;
        H.tile_max = std::max(mx, 1);
    }
    if (int rc = tel_upload(F, &H.ref, refc.data(), refc.size())) return rc;
    if (int rc = tel_upload(F, &H.topc, topc.data(), topc.size())) return rc;
    H.nref = (int)refc.size();
    H.ntop = (int)topc.size();
    H.top = top;
    for (int k = 1; k < top; k++) {
        const TelLevHost &Vh = lvh[k];
        TelLev &V = H.lv[k];
        V.nk = Vh.nk;
        V.nblk = (int)Vh.blk.size() - 1;
        V.nge = (int)Vh.ge.size();
        int *kst, *kce, *ksg, *bl, *nbr, *cst, *cff, *ge, *anc;
        float *cww;
        uint8_t *css;
        if (int rc = tel_upload(F, &kst, Vh.kstart.data(), Vh.kstart.size())) return rc;
        if (int rc = tel_upload(F, &kce, Vh.kcell.data(), Vh.kcell.size())) return rc;
        if (int rc = tel_upload(F, &ksg, Vh.kseg.data(), Vh.kseg.size())) return rc;
        if (int rc = tel_upload(F, &bl, Vh.blk.data(), Vh.blk.size())) return rc;
        if (int rc = tel_upload(F, &nbr, Vh.nbr.data(), Vh.nbr.size())) return rc;
        if (int rc = tel_upload(F, &cst, Vh.cstart.data(), Vh.cstart.size())) return rc;
        if (int rc = tel_upload(F, &cff, Vh.cf.data(), Vh.cf.size())) return rc;
        if (int rc = tel_upload(F, &cww, Vh.cw.data(), Vh.cw.size())) return rc;
        if (int rc = tel_upload(F, &css, Vh.cs.data(), Vh.cs.size())) return rc;
        if (int rc = tel_upload(F, &ge, Vh.ge.data(), Vh.ge.size())) return rc;
        if (int rc = tel_upload(F, &anc, Vh.anc.data(), Vh.anc.size())) return rc;
        V.kstart = kst; V.kcell = kce; V.kseg = ksg; V.blk = bl; V.nbr = nbr; V.cstart = cst; V.cf = cff;
        V.cw = cww; V.cs = css; V.ge = ge; V.anc = anc;
        if (int rc = dalloc(F, &V.a, (size_t)4 * V.nk)) return rc;
        float **vf[] = {&V.ap, &V.dinv, &V.r, &V.z};
        for (float **f : vf)
            if (int rc = dalloc(F, f, (size_t)V.nk)) return rc;
    }
    if (int rc = tel_upload(F, &jnt, jl.data(), jl.size())) return rc;
    if (int rc = tel_upload(F, &sc_, str_c.data(), str_c.size())) return rc;
    if (int rc = tel_upload(F, &sk_, str_k.data(), str_k.size())) return rc;
    T.c2c = c2c; T.meta = met; T.x = xx; T.y = yy; T.dx = ddx; T.dy = ddy;
    T.gown = gown; T.gk = gk; T.gwall = gw; T.base = bas; T.bstart = bst; T.joint = jnt;
    T.str_c = sc_; T.str_k = sk_;
    if (!q_c.empty()) {
        if (int rc = tel_upload(F, &qc_, q_c.data(), q_c.size())) return rc;
        if (int rc = tel_upload(F, &qk_, q_k.data(), q_k.size())) return rc;
    }
    T.q_c = qc_; T.q_k = qk_;
    H.gown_h = t.gown;
    H.nblk_n = (n + BLOCK - 1) / BLOCK;
    H.nblk_nd = (nd + BLOCK - 1) / BLOCK;
    H.nblk_base = (d->nx * d->ny + BLOCK - 1) / BLOCK;
    F->nblk_kry = std::min(KRY_BLOCKS, H.nblk_n);
    F->use_graphs = true;
    if (const char *e = getenv("CMS2D_GRAPHS")) F->use_graphs = atoi(e) != 0;   // 0: direct launches

    Par &P = F->P;
    P.hdry = d->hdry > 0 ? d->hdry : F4(0.05);
    P.hmin = P.hdry;
    P.hdry1 = F4(5.0) * P.hdry;
    P.hdry2 = F4(2.0) * P.hdry;
    P.dtime = d->dt;
    P.c2a = F4(0.2) * F4(0.2) * T.areaavg;
    P.rtol = d->rtol > 0 ? d->rtol : 1e-10;
    P.atol = d->atol >= 0 ? d->atol : 1e-12;
    P.kmax = d->kmax > 0 ? d->kmax : 4000;
    P.noptset = d->noptset;
    P.wallfric = d->wallfric;
    P.ndsch = d->ndsch;
    {
        static const int turb[5] = {5, 0, 1, 2, 3};
        P.mturbul = turb[d->turbulence];
        static const int fric[4] = {2, 1, 3, 4};
        P.mbedfric = fric[d->friction];
    }
    F->maxit = d->maxit > 0 ? d->maxit : 20 + n / 10000;
    F->rampdur = std::max(d->rampdur, d->dt / 3600.0);
    F->dtvar = 0;
    F->dtimebeg = F->deltime = d->dt;
    F->ntc = d->ntc;
    for (int k = 0; k < d->ntc; k++) { F->amp[k] = d->amp[k]; F->speed[k] = d->speed[k]; F->phase[k] = d->phase[k]; }
    for (int q = 0; q < 3; q++) { F->rmom0[q] = 1e10; F->rmom[q] = 0.0; }
    for (int k = 0; k < 4; k++) F->side_type[k] = d->side_type[k];
    if (curve) {
        F->wse_t.assign(d->wse_times, d->wse_times + d->wse_ntimes);
        F->wse_v.assign(d->wse_values, d->wse_values + d->wse_ntimes);
        F->wse_order = std::max(std::min(d->wse_order ? d->wse_order : 2, d->wse_ntimes - 1), 1);
        F->wse_inc = 1;
        F->wse_adjust_off = d->wse_adjust_off != 0;
    }
    F->precond = d->precond;
    F->precision = d->precision;
    F->solver = d->solver;

    TelS &s = H.s;
    double **cf[] = {&s.u, &s.v, &s.p, &s.h, &s.pp, &s.zb, &s.vis, &s.visb, &s.visc, &s.u1, &s.v1, &s.p1, &s.h1,
                     &s.eta, &s.maxeta, &s.coefman, &s.cfrict, &s.z0, &s.uv, &s.uelwc, &s.bsvel, &s.cbf, &s.fc,
                     &s.wavestrx, &s.wavestry, &s.su, &s.sv, &s.sp, &s.spu, &s.spv, &s.ssu0, &s.ssv0, &s.spuv0,
                     &s.sppp0, &s.sspp0, &s.apuareap, &s.sumu, &s.Hu, &s.Hv, &s.resu, &s.resv, &s.resp, &s.dpx,
                     &s.dpy, &s.tu, &s.tv, &s.tvis};
    for (double **f : cf)
        if (int rc = dalloc(F, f, (size_t)nd)) return rc;
    double **ff[] = {&s.flux, &s.flux1, &s.hk, &s.pk, &s.visk, &s.zbk, &s.am, &s.apc};
    for (double **f : ff)
        if (int rc = dalloc(F, f, (size_t)TF * nd)) return rc;
    if (int rc = dalloc(F, &s.wseadj, std::max<size_t>(ng, 1))) return rc;
    uint8_t **bf[] = {&s.iwet, &s.iwet1, &s.icorner, &s.ipond};
    for (uint8_t **f : bf)
        if (int rc = dalloc(F, f, (size_t)nd)) return rc;
    if (int rc = dalloc(F, &s.plab, (size_t)nd)) return rc;
    if (int rc = dalloc(F, &s.flags, 4)) return rc;
    double *(*pairs[])[2][2] = {&H.gu, &H.gv, &H.gpp, &H.gHu, &H.gHv, &H.gap, &H.gzb};
    for (auto pr : pairs)
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                if (int rc = dalloc(F, &(*pr)[a][b], (size_t)nd)) return rc;
    // Krylov systems: partial sums and vectors
    {
        const size_t nb = F->nblk_kry;
        double *red;
        if (int rc = dalloc(F, &red, 3 * 5 * nb)) return rc;
        H.Kpp.red.pr[0] = red; H.Kpp.red.pv[0] = red + 2 * nb; H.Kpp.red.pt[0] = red + 3 * nb;
        for (int q = 0; q < 2; q++) {
            double *b = red + (q + 1) * 5 * nb;
            H.Kuv.red.pr[q] = b; H.Kuv.red.pv[q] = b + 2 * nb; H.Kuv.red.pt[q] = b + 3 * nb;
        }
        if (int rc = dalloc(F, &F->d_sc, 6)) return rc;
        if (int rc = dalloc(F, &F->bal_part, BAL_PART * nb)) return rc;
        if (F->q_rating) {
            const int m = (int)F->q_t.size();
            if (int rc = dalloc(F, &F->q_dev, 1 + 2 * (size_t)m)) return rc;
            std::vector<double> tab(1 + 2 * (size_t)m, 0.0);
            for (int i = 0; i < m; i++) { tab[1 + i] = F->q_t[i]; tab[1 + m + i] = F->q_v[i]; }
            HIPCHK(hipMemcpyAsync(F->q_dev, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, F->st));
            HIPCHK(hipStreamSynchronize(F->st));
        }
        if (int rc = dalloc(F, &F->bal, 1)) return rc;
        double *kv[3][7];
        for (int q = 0; q < 3; q++)
            for (int mm = 0; mm < 7; mm++)
                if (int rc = dalloc(F, &kv[q][mm], (size_t)nd)) return rc;
        H.Kpp.x[0] = s.pp;
        H.Kpp.r[0] = kv[0][0]; H.Kpp.r0[0] = kv[0][1]; H.Kpp.p[0] = kv[0][2];
        H.Kpp.v[0] = kv[0][3]; H.Kpp.t[0] = kv[0][4]; H.Kpp.ap_w[0] = kv[0][5]; H.Kpp.dinv_w[0] = kv[0][6];
        H.Kpp.sc = F->d_sc;
        for (int q = 0; q < 2; q++) {
            H.Kuv.x[q] = q ? s.v : s.u;
            H.Kuv.r[q] = kv[q + 1][0]; H.Kuv.r0[q] = kv[q + 1][1]; H.Kuv.p[q] = kv[q + 1][2];
            H.Kuv.v[q] = kv[q + 1][3]; H.Kuv.t[q] = kv[q + 1][4]; H.Kuv.ap_w[q] = kv[q + 1][5];
            H.Kuv.dinv_w[q] = kv[q + 1][6];
        }
        H.Kuv.sc = F->d_sc + 1;
        if (int rc = dalloc(F, &H.Kpp.zp, (size_t)nd)) return rc;
        if (int rc = dalloc(F, &H.Kpp.zs, (size_t)nd)) return rc;
        for (float **f : {&H.zp, &H.zq})
            if (int rc = dalloc(F, f, (size_t)nd)) return rc;
        if (d->precision == CMS2D_PREC_FP32_REFINE) {
            // the inner FP32 solve: vectors, diagonal, coefficients, partial sums
            KrySys<1, float> &Ki = F->Kin;
            float **vec[] = {&Ki.x[0], &Ki.r[0], &Ki.r0[0], &Ki.p[0], &Ki.v[0], &Ki.t[0], &Ki.ap_w[0], &Ki.dinv_w[0]};
            for (float **v : vec)
                if (int rc = dalloc(F, v, (size_t)nd)) return rc;
            if (int rc = dalloc(F, &H.apc32, (size_t)TF * nd)) return rc;
            double *ri;
            if (int rc = dalloc(F, &ri, 5 * nb)) return rc;
            Ki.red.pr[0] = ri; Ki.red.pv[0] = ri + 2 * nb; Ki.red.pt[0] = ri + 3 * nb;
            Ki.zp = H.Kpp.zp;
            Ki.zs = H.Kpp.zs;
            Ki.sc = F->d_sc + 3;
            KrySys<2, float> &K2 = F->Kin2;
            for (int q = 0; q < 2; q++) {
                float **v2[] = {&K2.x[q], &K2.r[q], &K2.r0[q], &K2.p[q], &K2.v[q], &K2.t[q], &K2.ap_w[q], &K2.dinv_w[q]};
                for (float **v : v2)
                    if (int rc = dalloc(F, v, (size_t)nd)) return rc;
            }
            double *r2;
            if (int rc = dalloc(F, &r2, 2 * 5 * nb)) return rc;
            for (int q = 0; q < 2; q++) {
                K2.red.pr[q] = r2 + q * 5 * nb;
                K2.red.pv[q] = K2.red.pr[q] + 2 * nb;
                K2.red.pt[q] = K2.red.pr[q] + 3 * nb;
            }
            K2.sc = F->d_sc + 4;
        }
    }
    if (int rc = mg_alloc(F, d->nx, d->ny)) return rc;
    if (int rc = dalloc(F, &F->mg.L[0].zo, (size_t)(d->nx + 2) * (d->ny + 2))) return rc;
    if (int rc = dalloc(F, &H.b0, (size_t)(d->nx + 2) * (d->ny + 2))) return rc;
    F->mg_tile = 128;
    if (const char *e = getenv("CMS2D_MG_TILE")) {
        const int tt = atoi(e);
        if (tt == 64 || tt == 128 || tt == 256) F->mg_tile = tt;
    }
    HIPCHK(hipHostMalloc((void **)&F->h_sc, 6 * sizeof(KryScal), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&F->h_flags, 4 * sizeof(int), hipHostMallocDefault));
    memset(F->h_sc, 0, 6 * sizeof(KryScal));
    F->tm.pool.resize(2 * 2048);
    for (auto &e : F->tm.pool) HIPCHK(hipEventCreate(&e));
    for (auto &e : F->step_ev) HIPCHK(hipEventCreate(&e));
    // inputs: bed (ghosts take their cell's), friction input, forcing, Coriolis
    {
        std::vector<double> zb(nd), z0(nd, (double)1.0e-5f), buf(nd, 0.0);
        for (int i = 0; i < n; i++) zb[i] = d->zb[i];
        for (int q = 0; q < ng; q++) zb[n + q] = zb[t.gown[q]];
        HIPCHK(hipMemcpyAsync(s.zb, zb.data(), nd * sizeof(double), hipMemcpyHostToDevice, F->st));
        HIPCHK(hipMemcpyAsync(s.z0, z0.data(), nd * sizeof(double), hipMemcpyHostToDevice, F->st));
        for (int i = 0; i < n; i++) buf[i] = d->mann[i];
        // fric_init's copy to the dummy cells (friction.F90:397-407) runs for the
        // first cell only (its loop stride is ncellsD): it takes the friction of its
        // first active neighbour in face order
        for (int k = 0; k < t.ncf[0] && P.mbedfric != 4; k++) {
            const int nb = t.c2c[(size_t)k * nd + 0];
            if (nb >= 0 && nb < n) { buf[0] = buf[nb]; break; }
        }
        double *in = P.mbedfric == 1 ? s.cfrict : P.mbedfric == 3 ? s.z0 : P.mbedfric == 4 ? s.cbf : s.coefman;
        HIPCHK(hipMemcpyAsync(in, buf.data(), n * sizeof(double), hipMemcpyHostToDevice, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
        if (d->wavestrx) HIPCHK(hipMemcpyAsync(s.wavestrx, d->wavestrx, n * sizeof(double), hipMemcpyHostToDevice, F->st));
        if (d->wavestry) HIPCHK(hipMemcpyAsync(s.wavestry, d->wavestry, n * sizeof(double), hipMemcpyHostToDevice, F->st));
        const double pi = std::acos(-1.0), omega = 2.0 * pi / 86400.0, deg2rad = pi / 180.0;
        const double f0 = 2.0 * omega * std::sin(d->latitude * deg2rad);
        std::vector<double> fc(nd, f0);
        if (d->coriolis_plane == 1) {
            const double beta = 2.0 * omega * std::cos(d->latitude * deg2rad) / 6371000.0;
            double ysum = 0.0;
            for (int i = 0; i < n; i++) ysum += t.y[i];
            const double yavg = ysum / (double)n;
            for (int i = 0; i < nd; i++) fc[i] = f0 + beta * (t.y[i] - yavg);
        }
        HIPCHK(hipMemcpyAsync(s.fc, fc.data(), nd * sizeof(double), hipMemcpyHostToDevice, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
    }
    // zbk with the all-cell gradient of the bed (its previous value zero)
    TLAUNCH(k_tl_grad<1>, H.nblk_n, T, (const uint8_t *)s.iwet, 0, (const double *)s.zb, (const double *)nullptr,
            (const double *)nullptr, tel_io(H.gzb, 0), GradIO{}, GradIO{});
    TLAUNCH(k_tl_zbk, H.nblk_n, T, s, (const double *)H.gzb[0][1], (const double *)H.gzb[1][1]);
    if (int rc = tel_init_state(F)) return rc;
    return tel_capture_graphs(F);
}

static int tel_step(cms2d_flow *F, int nsteps, cms2d_step_info *info) {
    TelHost &H = *F->tel;
    HIPCHK(hipEventRecord(F->step_ev[0], F->st));
    for (int q = 0; q < nsteps; q++) {
        F->ntime++;
        F->timesecs += F->deltime;
        F->ramp = ramp_func(F->timesecs / 3600.0, F->rampdur);
        if (int rc = tel_flow_imp(F)) return rc;
        hipLaunchKernelGGL(k_tl_bal_part, dim3(F->nblk_kry), dim3(KBLOCK), 0, F->st, H.T, H.s, F->P, F->bal_part);
        hipLaunchKernelGGL(k_tl_bal_fin, dim3(1), dim3(KBLOCK), 0, F->st, H.T, H.s, F->P,
                           (const double *)F->bal_part, F->nblk_kry, F->bal);
        TLAUNCH(k_tl_update, H.nblk_nd, H.T, H.s);
        // flow_update clears dppx, dppy: the next pp gradient starts from zero
        TLAUNCH(k_tl_zero, H.nblk_nd, H.gpp[0][H.cpp], (long)H.T.nd);
        TLAUNCH(k_tl_zero, H.nblk_nd, H.gpp[1][H.cpp], (long)H.T.nd);
    }
    HIPCHK(hipEventRecord(F->step_ev[1], F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    if (info) {
        info->niter = F->niter;
        info->iconv = F->iconv;
        for (int q = 0; q < 3; q++) { info->rmom[q] = F->rmom[q]; info->kry_iters[q] = F->kry[q]; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, F->step_ev[0], F->step_ev[1]);
        info->kernel_ms = ms;
        info->dt = F->deltime;
        std::vector<uint8_t> w(H.T.n);
        HIPCHK(hipMemcpyAsync(w.data(), H.s.iwet, H.T.n, hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
        int nw = 0;
        for (uint8_t x : w) nw += x;
        info->nwet = nw;
    }
    return CMS2D_OK;
}

static double *tel_field(TelHost &H, const char *name) {
    TelS &s = H.s;
    struct { const char *n; double *p; } tab[] = {
        {"eta", s.eta}, {"u", s.u}, {"v", s.v}, {"p", s.p}, {"h", s.h}, {"vis", s.vis}, {"cfrict", s.cfrict},
        {"coefman", s.coefman}, {"uv", s.uv}, {"bsvel", s.bsvel}, {"pp", s.pp}, {"dpx", s.dpx}, {"dpy", s.dpy},
        {"z0", s.z0}, {"zb", s.zb}, {"maxeta", s.maxeta}, {"wavestrx", s.wavestrx}, {"wavestry", s.wavestry},
        {"fc", s.fc}};
    for (auto &t : tab)
        if (!strcmp(t.n, name)) return t.p;
    return nullptr;
}

static int tel_get(cms2d_flow *F, const char *name, double *out) {
    TelHost &H = *F->tel;
    const int n = H.T.n;
    HIPCHK(hipSetDevice(F->dev));
    if (!strcmp(name, "iwet")) {
        std::vector<uint8_t> w(n);
        HIPCHK(hipMemcpyAsync(w.data(), H.s.iwet, n, hipMemcpyDeviceToHost, F->st));
        HIPCHK(hipStreamSynchronize(F->st));
        for (int i = 0; i < n; i++) out[i] = w[i];
        return CMS2D_OK;
    }
    double *p = tel_field(H, name);
    if (!p) return set_err(CMS2D_ERR_ARG, std::string("unknown field ") + name);
    HIPCHK(hipMemcpyAsync(out, p, n * sizeof(double), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    return CMS2D_OK;
}

static int tel_get_face(cms2d_flow *F, const char *name, double *out) {
    TelHost &H = *F->tel;
    const int n = H.T.n, nd = H.T.nd;
    double *p = !strcmp(name, "flux") ? H.s.flux : !strcmp(name, "hk") ? H.s.hk : !strcmp(name, "visk") ? H.s.visk
                : !strcmp(name, "pk") ? H.s.pk : nullptr;
    if (!p) return set_err(CMS2D_ERR_ARG, std::string("unknown face field ") + name);
    HIPCHK(hipSetDevice(F->dev));
    std::vector<double> buf((size_t)TF * nd);
    HIPCHK(hipMemcpyAsync(buf.data(), p, buf.size() * sizeof(double), hipMemcpyDeviceToHost, F->st));
    HIPCHK(hipStreamSynchronize(F->st));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < TF; k++) out[(size_t)TF * i + k] = buf[(size_t)k * nd + i];
    return CMS2D_OK;
}

static int tel_set(cms2d_flow *F, const char *name, const double *in) {
    TelHost &H = *F->tel;
    double *p = nullptr;
    if (!strcmp(name, "u")) p = H.s.u;
    else if (!strcmp(name, "v")) p = H.s.v;
    else if (!strcmp(name, "p")) p = H.s.p;
    if (!p) return set_err(CMS2D_ERR_ARG, "set takes u, v or p");
    HIPCHK(hipSetDevice(F->dev));
    HIPCHK(hipMemcpyAsync(p, in, H.T.n * sizeof(double), hipMemcpyHostToDevice, F->st));
    return tel_init_state(F);
}

