// mg_kernels.hip -- the V(1,1) cycle of mg.h.
//
// Per level above the coarsest, one cycle is two launches over tiles of the
// level (LDS holds the sweep's intermediate values over a halo):
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
// A workgroup owns a tile of MT_X x MT_Y fine cells (MT_X/2 x MT_Y/2 coarse
// cells, two per thread).  Intermediate values of the sweep are built in LDS
// over the tile and a halo; halo values are recomputed by the same expressions
// as in the neighbouring tile, so they are bitwise the values stored there.
// At 2000 x 1500 the 1504 tiles of level 0 fit the chip's workgroup slots in
// one round.
constexpr int MT_X = 64, MT_Y = 32, MT_THREADS = 256;

__device__ __forceinline__ bool mg_inside(const MgLevel &L, int i, int j) {
    return i >= 0 && j >= 0 && i < L.nx && j < L.ny;
}

// Down: pre-smoothing sweep of level f (z of its cells) and coarse b.
//   q  = b/ap over the tile + 2   (red z of the sweep from zero)
//   zt = z over the tile + 1      (black z = (b + sum a_k q_nb)/ap)
//   coarse b = sum over the red children of sum a_k zt_nb (their residual)
template <class TB>
__global__ void __launch_bounds__(MT_THREADS) k_mg_down(MgLevel f, const TB *b, MgLevel C, const KryScal *sc) {
    if (sc->done) return;
    __shared__ float q[MT_Y + 4][MT_X + 4];
    __shared__ float zt[MT_Y + 2][MT_X + 2];
    const int X0 = blockIdx.x * MT_X, Y0 = blockIdx.y * MT_Y;
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
    for (int t = threadIdx.x; t < (MT_X + 2) * (MT_Y + 2); t += MT_THREADS) {
        const int ly = t / (MT_X + 2), lx = t - ly * (MT_X + 2), i = X0 - 1 + lx, j = Y0 - 1 + ly;
        float z = 0.0f;
        if (mg_inside(f, i, j)) {
            const int c = f.idx(i, j);
            if (((i + j) & 1) == 0) {
                z = q[ly + 1][lx + 1];
            } else {
                const float s = (f.aN[c] * q[ly + 2][lx + 1] + f.aE[c] * q[ly + 1][lx + 2]) +
                                (f.aS[c] * q[ly][lx + 1] + f.aW[c] * q[ly + 1][lx]);
                z = ((float)b[c] + s) * f.dinv[c];
            }
            if (lx >= 1 && lx <= MT_X && ly >= 1 && ly <= MT_Y) f.z[c] = z;
        }
        zt[ly][lx] = z;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < (MT_X / 2) * (MT_Y / 2); t += MT_THREADS) {
        const int cx = t % (MT_X / 2), cy = t / (MT_X / 2);
        const int I = X0 / 2 + cx, J = Y0 / 2 + cy;
        if (I >= C.nx || J >= C.ny) continue;
        const int i0 = 2 * I, j0 = 2 * J, lx = i0 - X0 + 1, ly = j0 - Y0 + 1;
        const int c00 = f.idx(i0, j0);
        float r = (f.aN[c00] * zt[ly + 1][lx] + f.aE[c00] * zt[ly][lx + 1]) +
                  (f.aS[c00] * zt[ly - 1][lx] + f.aW[c00] * zt[ly][lx - 1]);
        if (i0 + 1 < f.nx && j0 + 1 < f.ny) {
            const int c11 = f.idx(i0 + 1, j0 + 1);
            r += (f.aN[c11] * zt[ly + 2][lx + 1] + f.aE[c11] * zt[ly + 1][lx + 2]) +
                 (f.aS[c11] * zt[ly][lx + 1] + f.aW[c11] * zt[ly + 1][lx]);
        }
        C.b[C.idx(I, J)] = r;
    }
}

// Up: coarse correction and post-smoothing sweep (red, then black) of level
// L into L.zo.
//   w  = z + zc(parent) over the tile + 2
//   rt = new red z = (b + sum a_k w_nb)/ap over the tile + 1
//   black zo = (b + sum a_k rt_nb)/ap on the tile
// (the coarse correction of a red cell itself is replaced by its update).
template <class TB>
__global__ void __launch_bounds__(MT_THREADS) k_mg_up(MgLevel L, const TB *b, MgLevel C, const KryScal *sc) {
    if (sc->done) return;
    __shared__ float w[MT_Y + 4][MT_X + 4];
    __shared__ float rt[MT_Y + 2][MT_X + 2];
    const int X0 = blockIdx.x * MT_X, Y0 = blockIdx.y * MT_Y;
    for (int t = threadIdx.x; t < (MT_X + 4) * (MT_Y + 4); t += MT_THREADS) {
        const int ly = t / (MT_X + 4), lx = t - ly * (MT_X + 4), i = X0 - 2 + lx, j = Y0 - 2 + ly;
        float v = 0.0f;
        if (mg_inside(L, i, j)) v = L.z[L.idx(i, j)] + C.zo[C.idx(i >> 1, j >> 1)];
        w[ly][lx] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < (MT_X + 2) * (MT_Y + 2); t += MT_THREADS) {
        const int ly = t / (MT_X + 2), lx = t - ly * (MT_X + 2), i = X0 - 1 + lx, j = Y0 - 1 + ly;
        float z = 0.0f;
        if (mg_inside(L, i, j) && ((i + j) & 1) == 0) {
            const int c = L.idx(i, j);
            const float s = (L.aN[c] * w[ly + 2][lx + 1] + L.aE[c] * w[ly + 1][lx + 2]) +
                            (L.aS[c] * w[ly][lx + 1] + L.aW[c] * w[ly + 1][lx]);
            z = ((float)b[c] + s) * L.dinv[c];
            if (lx >= 1 && lx <= MT_X && ly >= 1 && ly <= MT_Y) L.zo[c] = z;
        }
        rt[ly][lx] = z;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < MT_X * MT_Y; t += MT_THREADS) {
        const int ly = t / MT_X + 1, lx = t % MT_X + 1, i = X0 - 1 + lx, j = Y0 - 1 + ly;
        if (!mg_inside(L, i, j) || ((i + j) & 1) == 0) continue;
        const int c = L.idx(i, j);
        const float s = (L.aN[c] * rt[ly + 1][lx] + L.aE[c] * rt[ly][lx + 1]) +
                        (L.aS[c] * rt[ly - 1][lx] + L.aW[c] * rt[ly][lx - 1]);
        L.zo[c] = ((float)b[c] + s) * L.dinv[c];
    }
}

// Coarsest grid: MG_COARSE_SWEEPS red-black sweeps from z = 0 in one block
// with z in LDS (padded layout of the level, ghost ring zero) and each
// thread's cells in registers; result to L.zo.
constexpr int MG_CT = 1024;
constexpr int MG_CPT = MG_COARSE_MAX / MG_CT;   // cells per thread
template <class TB>
__global__ void __launch_bounds__(MG_CT) k_mg_coarsest(MgLevel L, const TB *b, const KryScal *sc) {
    if (sc->done) return;
    extern __shared__ float zs[];
    const int n = L.nx * L.ny, npad = L.nxp * (L.ny + 2);
    for (int q = threadIdx.x; q < npad; q += MG_CT) zs[q] = 0.0f;
    float aN[MG_CPT], aE[MG_CPT], aS[MG_CPT], aW[MG_CPT], di[MG_CPT], bb[MG_CPT];
    int cix[MG_CPT], col[MG_CPT];
#pragma unroll
    for (int k = 0; k < MG_CPT; k++) {
        const int t = threadIdx.x + k * MG_CT;
        cix[k] = -1;
        col[k] = -1;
        aN[k] = aE[k] = aS[k] = aW[k] = di[k] = bb[k] = 0.0f;
        if (t < n) {
            const int j = t / L.nx, i = t - j * L.nx, c = L.idx(i, j);
            cix[k] = c;
            col[k] = (i + j) & 1;
            aN[k] = L.aN[c]; aE[k] = L.aE[c]; aS[k] = L.aS[c]; aW[k] = L.aW[c];
            di[k] = L.dinv[c];
            bb[k] = (float)b[c];
        }
    }
    __syncthreads();
    for (int sw = 0; sw < MG_COARSE_SWEEPS; sw++) {
        for (int cl = 0; cl < 2; cl++) {
#pragma unroll
            for (int k = 0; k < MG_CPT; k++) {
                if (col[k] != cl) continue;
                const int c = cix[k];
                const float s = (aN[k] * zs[c + L.nxp] + aE[k] * zs[c + 1]) + (aS[k] * zs[c - L.nxp] + aW[k] * zs[c - 1]);
                zs[c] = (bb[k] + s) * di[k];
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int k = 0; k < MG_CPT; k++)
        if (cix[k] >= 0) L.zo[cix[k]] = zs[cix[k]];
}

}  // namespace cms
