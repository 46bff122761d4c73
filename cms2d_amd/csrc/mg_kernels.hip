// mg_kernels.hip -- the V(1,1) cycle of mg.h.
//
// Per level above the coarsest, one cycle is two launches (over LDS tiles on
// level 0, one thread per coarse cell on the coarser levels):
//   k_mg_down  red-black sweep from z = 0 (red z = b/ap, black
//              z = (b + sum a_k b_nb/ap_nb)/ap) and the coarse right-hand
//              side: the sum over the red children of their residual, which
//              after that sweep is sum a_k z_nb exactly (red z = b/ap cancels
//              b - ap z) and zero on black children
//   (coarser levels, then k_mg_coarsest in one workgroup)
//   k_mg_up    red cells: z = (b + sum a_k (z_nb + zc(parent(nb))))/ap -- the
//              coarse correction of the red cell itself is replaced by this
//              update, so only its neighbours' is needed; then black cells:
//              z = (b + sum a_k z_nb)/ap with the new red values
// The down sweep writes z and reads only b; the up sweep reads z and the
// coarse zo and writes zo, so no launch reads values it also writes.
#include "mg.h"
#include "krylov.h"

namespace cms {

__device__ __forceinline__ bool mg_thread(const MgLevel &L, int &i, int &j, int &c) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= L.nx * L.ny) return false;
    j = t / L.nx;
    i = t - j * L.nx;
    c = L.idx(i, j);
    return true;
}

// Level 0 from the FP64 pressure-correction system (directed coefficients on
// the padded grid, ap from the solve prologue).
__global__ void __launch_bounds__(BLOCK) k_mg_setup0(Geo g, const double *aN, const double *aE, const double *aS,
                                                     const double *aW, const double *ap, MgLevel L) {
    int i, j, c;
    if (!mg_thread(L, i, j, c)) return;
    const double n_ = j < g.ny - 1 ? aN[c] : 0.0, e_ = i < g.nx - 1 ? aE[c] : 0.0;
    const double s_ = j > 0 ? aS[c] : 0.0, w_ = i > 0 ? aW[c] : 0.0;
    const double d = ap[c];
    L.aN[c] = (float)n_;
    L.aE[c] = (float)e_;
    L.aS[c] = (float)s_;
    L.aW[c] = (float)w_;
    L.dinv[c] = d > 0.0 ? (float)(1.0 / d) : 0.0f;
    L.ap[c] = (float)d;
    L.e[c] = (float)fmax(d - (((n_ + e_) + s_) + w_), 0.0);
}

// Coarse level from its fine level (children (2I+a, 2J+b), a,b in {0,1}).
__global__ void __launch_bounds__(BLOCK) k_mg_coarsen(MgLevel f, MgLevel C) {
    int I, J, c;
    if (!mg_thread(C, I, J, c)) return;
    const int i0 = 2 * I, j0 = 2 * J;
    const bool hi = i0 + 1 < f.nx, hj = j0 + 1 < f.ny;
    const int i1 = hi ? i0 + 1 : i0, j1 = hj ? j0 + 1 : j0;
    const int c00 = f.idx(i0, j0), c10 = f.idx(i1, j0), c01 = f.idx(i0, j1), c11 = f.idx(i1, j1);
    float e = f.e[c00];
    if (hi) e += f.e[c10];
    if (hj) e += f.e[c01];
    if (hi && hj) e += f.e[c11];
    const float aE = 0.5f * (f.aE[c10] + (hj ? f.aE[c11] : 0.0f));
    const float aW = 0.5f * (f.aW[c00] + (hj ? f.aW[c01] : 0.0f));
    const float aN = 0.5f * (f.aN[c01] + (hi ? f.aN[c11] : 0.0f));
    const float aS = 0.5f * (f.aS[c00] + (hi ? f.aS[c10] : 0.0f));
    const float ap = e + ((aN + aE) + (aS + aW));
    C.aN[c] = aN;
    C.aE[c] = aE;
    C.aS[c] = aS;
    C.aW[c] = aW;
    C.e[c] = e;
    C.dinv[c] = ap > 0.0f ? 1.0f / ap : 0.0f;
}

// --- fused cycle: one launch down and one up per level ------------------------
// A workgroup owns a tile of TX x 2048/TX fine cells; each thread owns two
// coarse cells of it and their eight children, whose coefficients it loads
// once into registers and keeps for both phases of the sweep.  Neighbour
// values come from LDS over the tile and a halo; halo values are recomputed
// by the same expressions as in the neighbouring tile (mg_black, mg_red), so
// they are bitwise the values stored there and the result does not depend on
// the tile shape.  The shape (TX = 64, 128 or 256) is picked per grid by
// timing the variants when the flow is created.  At 2000 x 1500 the 1504
// tiles of 128 x 16 fit the chip's workgroup slots (seven down or eight up
// workgroups per CU) in one round.
constexpr int MT_THREADS = 256, MT_OWN = 8;
template <int TX>
struct MTile {
    static constexpr int X = TX, Y = 2048 / TX, RING = 2 * (X + 2) + 2 * Y;
    static_assert(MT_THREADS * 2 == (X / 2) * (Y / 2) && Y % 4 == 0, "two coarse cells per thread");
};

__device__ __forceinline__ bool mg_inside(const MgLevel &L, int i, int j) {
    return i >= 0 && j >= 0 && i < L.nx && j < L.ny;
}
// Fine cell k (0..7) of thread t in tile coordinates: coarse cells (cx, cy)
// and (cx, cy + Y/4), children (sub & 1, sub >> 1).
template <int TX>
__device__ __forceinline__ void mg_own(int t, int k, int &fx, int &fy) {
    using T = MTile<TX>;
    const int cx = t % (T::X / 2), cy = t / (T::X / 2) + (k >> 2) * (T::Y / 4), sub = k & 3;
    fx = 2 * cx + (sub & 1);
    fy = 2 * cy + (sub >> 1);
}
// Origin of this workgroup's tile.  Tiles are numbered row-major and dealt
// to the XCDs in contiguous runs (xcd_block), so a tile and the neighbours
// whose halo lines it shares mostly run on one XCD and meet in its L2; the
// hardware's own round-robin would put horizontal neighbours on different
// XCDs, each fetching the shared lines from HBM.
template <int TX>
__device__ __forceinline__ void mg_tile_origin(int &X0, int &Y0) {
    const int nt = (int)(gridDim.x * gridDim.y);
    const int t = xcd_block((int)(blockIdx.y * gridDim.x + blockIdx.x), nt);
    const int ty = t / (int)gridDim.x, tx = t - ty * (int)gridDim.x;
    X0 = tx * MTile<TX>::X;
    Y0 = ty * MTile<TX>::Y;
}
// Cells of the one-cell ring around the tile, in tile coordinates.
template <int TX>
__device__ __forceinline__ void mg_ring(int t, int &fx, int &fy) {
    using T = MTile<TX>;
    if (t < T::X + 2) { fx = t - 1; fy = -1; }
    else if (t < 2 * (T::X + 2)) { fx = t - (T::X + 2) - 1; fy = T::Y; }
    else { const int u = t - 2 * (T::X + 2); fx = u < T::Y ? -1 : T::X; fy = u % T::Y; }
}

// Black z of the sweep from zero: (b + sum a_k q_nb)/ap = q + (sum a_k q_nb)/ap,
// q = b/ap (red z).
__device__ __forceinline__ float mg_black(float aN, float aE, float aS, float aW, float qN, float qE, float qS,
                                          float qW, float qc, float d) {
    const float s = (aN * qN + aE * qE) + (aS * qS + aW * qW);
    return qc + s * d;
}
// Red z after the coarse correction: (b + sum a_k w_nb)/ap, w = z + zc.
__device__ __forceinline__ float mg_red(float aN, float aE, float aS, float aW, float wN, float wE, float wS,
                                        float wW, float bc, float d) {
    const float s = (aN * wN + aE * wE) + (aS * wS + aW * wW);
    return (bc + s) * d;
}

// Down: pre-smoothing sweep of level f (z of its cells) and coarse b.
//   q = b/ap over the tile + 2, z over the tile + 1 in LDS; coarse b = sum over
//   the red children of sum a_k z_nb (their residual after the sweep).
template <int TX, class TB>
__global__ void __launch_bounds__(MT_THREADS) k_mg_down(MgLevel f, const TB *b, MgLevel C, const KryScal *sc) {
    constexpr int MT_X = MTile<TX>::X, MT_Y = MTile<TX>::Y, MT_RING = MTile<TX>::RING;
    if (sc->done) return;
    // 1/ap of the black cells is read into registers with the coefficients
    // (the staging reads the same lines) rather than kept in LDS: without that second
    // array the workgroup fits seven times on a CU, and the 1504 tiles of a
    // 2000 x 1500 grid run in one round instead of two
    __shared__ float q[MT_Y + 4][MT_X + 4];
    __shared__ float zt[MT_Y + 2][MT_X + 2];
    int X0, Y0;
    mg_tile_origin<TX>(X0, Y0);
    // the owned cells' coefficients are requested before the staging loads,
    // so both are in flight together instead of one after the barrier
    float a[MT_OWN][4], dk[MT_OWN];
#pragma unroll
    for (int k = 0; k < MT_OWN; k++) {
        int fx, fy;
        mg_own<TX>(threadIdx.x, k, fx, fy);
        const int i = X0 + fx, j = Y0 + fy;
        a[k][0] = a[k][1] = a[k][2] = a[k][3] = 0.0f;
        dk[k] = 0.0f;
        if (mg_inside(f, i, j)) {
            const int c = f.idx(i, j);
            a[k][0] = f.aN[c]; a[k][1] = f.aE[c]; a[k][2] = f.aS[c]; a[k][3] = f.aW[c];
            if ((i + j) & 1) dk[k] = f.dinv[c];
        }
    }
    for (int t = threadIdx.x; t < (MT_X + 4) * (MT_Y + 4); t += MT_THREADS) {
        const int ly = t / (MT_X + 4), lx = t - ly * (MT_X + 4), i = X0 - 2 + lx, j = Y0 - 2 + ly;
        float v = 0.0f;
        if (mg_inside(f, i, j)) {
            const int c = f.idx(i, j);
            v = (float)b[c] * f.dinv[c];
        }
        q[ly][lx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MT_OWN; k++) {
        int fx, fy;
        mg_own<TX>(threadIdx.x, k, fx, fy);
        const int i = X0 + fx, j = Y0 + fy, lx = fx + 2, ly = fy + 2;
        float z = 0.0f;
        if (mg_inside(f, i, j)) {
            const int c = f.idx(i, j);
            if (((i + j) & 1) == 0) z = q[ly][lx];
            else z = mg_black(a[k][0], a[k][1], a[k][2], a[k][3], q[ly + 1][lx], q[ly][lx + 1], q[ly - 1][lx],
                              q[ly][lx - 1], q[ly][lx], dk[k]);
            f.z[c] = z;
        }
        zt[fy + 1][fx + 1] = z;
    }
    for (int t = threadIdx.x; t < MT_RING; t += MT_THREADS) {
        int fx, fy;
        mg_ring<TX>(t, fx, fy);
        const int i = X0 + fx, j = Y0 + fy, lx = fx + 2, ly = fy + 2;
        float z = 0.0f;
        if (mg_inside(f, i, j)) {
            if (((i + j) & 1) == 0) {
                z = q[ly][lx];
            } else {
                const int c = f.idx(i, j);
                z = mg_black(f.aN[c], f.aE[c], f.aS[c], f.aW[c], q[ly + 1][lx], q[ly][lx + 1], q[ly - 1][lx],
                             q[ly][lx - 1], q[ly][lx], f.dinv[c]);
            }
        }
        zt[fy + 1][fx + 1] = z;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
        int fx, fy;
        mg_own<TX>(threadIdx.x, 4 * h, fx, fy);                 // child (0,0): red
        const int I = (X0 + fx) / 2, J = (Y0 + fy) / 2;
        if (I >= C.nx || J >= C.ny) continue;
        const int x = fx + 1, y = fy + 1;                    // zt frame
        const float *a0 = a[4 * h], *a3 = a[4 * h + 3];
        float r = (a0[0] * zt[y + 1][x] + a0[1] * zt[y][x + 1]) + (a0[2] * zt[y - 1][x] + a0[3] * zt[y][x - 1]);
        // child (1,1), red; its coefficients are zero outside the grid
        r += (a3[0] * zt[y + 2][x + 1] + a3[1] * zt[y + 1][x + 2]) + (a3[2] * zt[y][x + 1] + a3[3] * zt[y + 1][x]);
        C.b[C.idx(I, J)] = r;
    }
}

// Up: coarse correction and post-smoothing sweep (red, then black) of level
// L into L.zo.  w = z + zc(parent) over the tile + 2 and the new red z over
// the tile + 1 in LDS; the coarse correction of a red cell itself is
// replaced by its update.
template <int TX, class TB>
__global__ void __launch_bounds__(MT_THREADS) k_mg_up(MgLevel L, const TB *b, MgLevel C, const KryScal *sc) {
    constexpr int MT_X = MTile<TX>::X, MT_Y = MTile<TX>::Y, MT_RING = MTile<TX>::RING;
    if (sc->done) return;
    __shared__ float w[MT_Y + 4][MT_X + 4];
    __shared__ float rt[MT_Y + 2][MT_X + 2];
    int X0, Y0;
    mg_tile_origin<TX>(X0, Y0);
    // the owned cells' operands are requested before the staging loads (see k_mg_down)
    float a[MT_OWN][4], bb[MT_OWN], dd[MT_OWN];
#pragma unroll
    for (int k = 0; k < MT_OWN; k++) {
        int fx, fy;
        mg_own<TX>(threadIdx.x, k, fx, fy);
        const int i = X0 + fx, j = Y0 + fy;
        a[k][0] = a[k][1] = a[k][2] = a[k][3] = 0.0f;
        bb[k] = dd[k] = 0.0f;
        if (mg_inside(L, i, j)) {
            const int c = L.idx(i, j);
            a[k][0] = L.aN[c]; a[k][1] = L.aE[c]; a[k][2] = L.aS[c]; a[k][3] = L.aW[c];
            bb[k] = (float)b[c];
            dd[k] = L.dinv[c];
        }
    }
    for (int t = threadIdx.x; t < (MT_X + 4) * (MT_Y + 4); t += MT_THREADS) {
        const int ly = t / (MT_X + 4), lx = t - ly * (MT_X + 4), i = X0 - 2 + lx, j = Y0 - 2 + ly;
        float v = 0.0f;
        if (mg_inside(L, i, j)) v = L.z[L.idx(i, j)] + C.zo[C.idx(i >> 1, j >> 1)];
        w[ly][lx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MT_OWN; k++) {
        int fx, fy;
        mg_own<TX>(threadIdx.x, k, fx, fy);
        const int i = X0 + fx, j = Y0 + fy, lx = fx + 2, ly = fy + 2;
        float z = 0.0f;
        if (mg_inside(L, i, j)) {
            const int c = L.idx(i, j);
            if (((i + j) & 1) == 0) {
                z = mg_red(a[k][0], a[k][1], a[k][2], a[k][3], w[ly + 1][lx], w[ly][lx + 1], w[ly - 1][lx],
                           w[ly][lx - 1], bb[k], dd[k]);
                L.zo[c] = z;
            }
        }
        rt[fy + 1][fx + 1] = z;
    }
    for (int t = threadIdx.x; t < MT_RING; t += MT_THREADS) {
        int fx, fy;
        mg_ring<TX>(t, fx, fy);
        const int i = X0 + fx, j = Y0 + fy, lx = fx + 2, ly = fy + 2;
        float z = 0.0f;
        if (mg_inside(L, i, j) && ((i + j) & 1) == 0) {
            const int c = L.idx(i, j);
            z = mg_red(L.aN[c], L.aE[c], L.aS[c], L.aW[c], w[ly + 1][lx], w[ly][lx + 1], w[ly - 1][lx],
                       w[ly][lx - 1], (float)b[c], L.dinv[c]);
        }
        rt[fy + 1][fx + 1] = z;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MT_OWN; k++) {
        int fx, fy;
        mg_own<TX>(threadIdx.x, k, fx, fy);
        const int i = X0 + fx, j = Y0 + fy, x = fx + 1, y = fy + 1;
        if (!mg_inside(L, i, j) || ((i + j) & 1) == 0) continue;
        const float s = (a[k][0] * rt[y + 1][x] + a[k][1] * rt[y][x + 1]) + (a[k][2] * rt[y - 1][x] + a[k][3] * rt[y][x - 1]);
        L.zo[L.idx(i, j)] = (bb[k] + s) * dd[k];
    }
}

// --- the same cycle on coarse levels: one thread per coarse cell ----------------
// On coarse levels a tile kernel is latency-bound (few tiles, three dependent
// phases); here every thread issues all its loads at once and recomputes the
// values it needs from neighbouring children by the same expressions
// (mg_black0, mg_red1), so they are bitwise the values stored by their owners.

// Black z after the red-black sweep from zero (red z = b/ap), or 0 outside
// the grid.
template <class TB>
__device__ __forceinline__ float mg_black0(const MgLevel &L, const TB *b, int i, int j) {
    if (i < 0 || j < 0 || i >= L.nx || j >= L.ny) return 0.0f;
    const int c = L.idx(i, j);
    const int cN = c + L.nxp, cE = c + 1, cS = c - L.nxp, cW = c - 1;
    const float s = (L.aN[c] * ((float)b[cN] * L.dinv[cN]) + L.aE[c] * ((float)b[cE] * L.dinv[cE])) +
                    (L.aS[c] * ((float)b[cS] * L.dinv[cS]) + L.aW[c] * ((float)b[cW] * L.dinv[cW]));
    return ((float)b[c] + s) * L.dinv[c];
}

// Residual of red cell (i,j) after that sweep: sum a_k z_nb (b - ap z = 0).
template <class TB>
__device__ __forceinline__ float mg_red_res(const MgLevel &L, const TB *b, int i, int j, float zN, float zE,
                                            float zS, float zW) {
    const int c = L.idx(i, j);
    return (L.aN[c] * zN + L.aE[c] * zE) + (L.aS[c] * zS + L.aW[c] * zW);
}

// Down, coarse cell (I, J): pre-smoothing sweep of its children on level f
// (stored to f.z when store) and their restricted residual, returned.
template <class TB>
__device__ __forceinline__ float mg_down_cell(const MgLevel &f, const TB *b, int I, int J, bool store) {
    const int i0 = 2 * I, j0 = 2 * J, i1 = i0 + 1, j1 = j0 + 1;
    const bool hi = i1 < f.nx, hj = j1 < f.ny;
    // black children (i1,j0), (i0,j1) and the black neighbours of the red ones
    const float zb10 = mg_black0(f, b, i1, j0), zb01 = mg_black0(f, b, i0, j1);
    const float zbW = mg_black0(f, b, i0 - 1, j0), zbS = mg_black0(f, b, i0, j0 - 1);
    const int c00 = f.idx(i0, j0);
    const float z00 = (float)b[c00] * f.dinv[c00];
    float r = mg_red_res(f, b, i0, j0, zb01, zb10, zbS, zbW);
    float z11 = 0.0f;
    if (hi && hj) {
        const int c11 = f.idx(i1, j1);
        z11 = (float)b[c11] * f.dinv[c11];
        const float zbE = mg_black0(f, b, i1 + 1, j1), zbN = mg_black0(f, b, i1, j1 + 1);
        r += mg_red_res(f, b, i1, j1, zbN, zbE, zb10, zb01);
    }
    if (store) {
        f.z[c00] = z00;
        if (hi) f.z[f.idx(i1, j0)] = zb10;
        if (hj) f.z[f.idx(i0, j1)] = zb01;
        if (hi && hj) f.z[f.idx(i1, j1)] = z11;
    }
    return r;
}

// Down: pre-smoothing sweep of level f (z of its cells) and coarse b.
template <class TB>
__global__ void __launch_bounds__(BLOCK) k_mg_down_pc(MgLevel f, const TB *b, MgLevel C, const KryScal *sc) {
    const bool done = sc->done;   // tested before the stores: its load overlaps the others
    int I, J, cc;
    if (!mg_thread(C, I, J, cc)) return;
    const float r = mg_down_cell(f, b, I, J, !done);
    if (!done) C.b[cc] = r;
}

// Red z after the coarse correction and the red half-sweep, or 0 outside.
// zd: z of the down sweep, zc: coarse correction (padded layout of C, in
// global memory or LDS).
template <class TB>
__device__ __forceinline__ float mg_red1(const MgLevel &L, const TB *b, const MgLevel &C, const float *zc, int i,
                                         int j) {
    if (i < 0 || j < 0 || i >= L.nx || j >= L.ny) return 0.0f;
    const int c = L.idx(i, j);
    const int pN = C.idx(i >> 1, min(j + 1, L.ny - 1) >> 1), pS = C.idx(i >> 1, max(j - 1, 0) >> 1);
    const int pE = C.idx(min(i + 1, L.nx - 1) >> 1, j >> 1), pW = C.idx(max(i - 1, 0) >> 1, j >> 1);
    const float s = (L.aN[c] * (L.z[c + L.nxp] + zc[pN]) + L.aE[c] * (L.z[c + 1] + zc[pE])) +
                    (L.aS[c] * (L.z[c - L.nxp] + zc[pS]) + L.aW[c] * (L.z[c - 1] + zc[pW]));
    return ((float)b[c] + s) * L.dinv[c];
}

// Up, coarse cell (I, J): coarse correction and post-smoothing sweep (red,
// then black) of its children on level L into L.zo (when store).
template <class TB>
__device__ __forceinline__ void mg_up_cell(const MgLevel &L, const TB *b, const MgLevel &C, const float *zc, int I,
                                           int J, bool store) {
    const int i0 = 2 * I, j0 = 2 * J, i1 = i0 + 1, j1 = j0 + 1;
    const bool hi = i1 < L.nx, hj = j1 < L.ny;
    const float r00 = mg_red1(L, b, C, zc, i0, j0);
    const float r11 = (hi && hj) ? mg_red1(L, b, C, zc, i1, j1) : 0.0f;
    float z10 = 0.0f, z01 = 0.0f;
    if (hi) {
        // black (i1,j0): red neighbours (i1,j1) N, (i1+1,j0) E, (i1,j0-1) S, (i0,j0) W
        const int c = L.idx(i1, j0);
        const float zN = r11, zE = mg_red1(L, b, C, zc, i1 + 1, j0), zS = mg_red1(L, b, C, zc, i1, j0 - 1), zW = r00;
        z10 = ((float)b[c] + ((L.aN[c] * zN + L.aE[c] * zE) + (L.aS[c] * zS + L.aW[c] * zW))) * L.dinv[c];
    }
    if (hj) {
        // black (i0,j1): red neighbours (i0,j1+1) N, (i1,j1) E, (i0,j0) S, (i0-1,j1) W
        const int c = L.idx(i0, j1);
        const float zN = mg_red1(L, b, C, zc, i0, j1 + 1), zE = r11, zS = r00, zW = mg_red1(L, b, C, zc, i0 - 1, j1);
        z01 = ((float)b[c] + ((L.aN[c] * zN + L.aE[c] * zE) + (L.aS[c] * zS + L.aW[c] * zW))) * L.dinv[c];
    }
    if (!store) return;
    L.zo[L.idx(i0, j0)] = r00;
    if (hi && hj) L.zo[L.idx(i1, j1)] = r11;
    if (hi) L.zo[L.idx(i1, j0)] = z10;
    if (hj) L.zo[L.idx(i0, j1)] = z01;
}

// Up: coarse correction and post-smoothing sweep (red, then black) of level
// L into L.zo.
template <class TB>
__global__ void __launch_bounds__(BLOCK) k_mg_up_pc(MgLevel L, const TB *b, MgLevel C, const KryScal *sc) {
    const bool done = sc->done;   // tested before the stores: its load overlaps the others
    int I, J, cc;
    if (!mg_thread(C, I, J, cc)) return;
    mg_up_cell(L, b, C, C.zo, I, J, !done);
}

// Coarsest grid: MG_COARSE_SWEEPS red-black sweeps from z = 0 in one block
// with z in LDS (zs, padded layout of the level, ghost ring zero) and each
// thread's cells (t = thread + k * MG_CT) in registers; bsrc(c) gives the
// right-hand side of padded cell c.  The cell coordinates are stepped, not
// stored, to keep the registers for the coefficients.  Ends with a barrier.
constexpr int MG_CT = 1024;
constexpr int MG_CPT = MG_COARSE_MAX / MG_CT;   // cells per thread
template <class BSrc>
__device__ __forceinline__ void mg_coarse_sweeps(const MgLevel &L, BSrc bsrc, float *zs) {
    const int n = L.nx * L.ny;
    const int j0 = (int)threadIdx.x / L.nx, i0 = (int)threadIdx.x - j0 * L.nx;
    const int dj = MG_CT / L.nx, di_ = MG_CT - dj * L.nx;
    float aN[MG_CPT], aE[MG_CPT], aS[MG_CPT], aW[MG_CPT], di[MG_CPT], bb[MG_CPT];
    {
        int i = i0, j = j0;
#pragma unroll
        for (int k = 0; k < MG_CPT; k++) {
            aN[k] = aE[k] = aS[k] = aW[k] = di[k] = bb[k] = 0.0f;
            if ((int)threadIdx.x + k * MG_CT < n) {
                const int c = L.idx(i, j);
                aN[k] = L.aN[c]; aE[k] = L.aE[c]; aS[k] = L.aS[c]; aW[k] = L.aW[c];
                di[k] = L.dinv[c];
                bb[k] = bsrc(c);
            }
            i += di_; j += dj;
            if (i >= L.nx) { i -= L.nx; j++; }
        }
    }
    __syncthreads();
    for (int sw = 0; sw < MG_COARSE_SWEEPS; sw++) {
        for (int cl = 0; cl < 2; cl++) {
            int i = i0, j = j0;
#pragma unroll
            for (int k = 0; k < MG_CPT; k++) {
                if ((int)threadIdx.x + k * MG_CT < n && ((i + j) & 1) == cl) {
                    const int c = L.idx(i, j);
                    const float s = (aN[k] * zs[c + L.nxp] + aE[k] * zs[c + 1]) + (aS[k] * zs[c - L.nxp] + aW[k] * zs[c - 1]);
                    zs[c] = (bb[k] + s) * di[k];
                }
                i += di_; j += dj;
                if (i >= L.nx) { i -= L.nx; j++; }
            }
            __syncthreads();
        }
    }
}

// The coarsest grid: result to L.zo.
template <class TB>
__global__ void __launch_bounds__(MG_CT) k_mg_coarsest(MgLevel L, const TB *b, const KryScal *sc) {
    if (sc->done) return;
    extern __shared__ float zs[];
    const int n = L.nx * L.ny, npad = L.nxp * (L.ny + 2);
    for (int q = threadIdx.x; q < npad; q += MG_CT) zs[q] = 0.0f;
    mg_coarse_sweeps(L, [&](int c) { return (float)b[c]; }, zs);
    for (int t = threadIdx.x; t < n; t += MG_CT) {
        const int j = t / L.nx, i = t - j * L.nx, c = L.idx(i, j);
        L.zo[c] = zs[c];
    }
}

}  // namespace cms
