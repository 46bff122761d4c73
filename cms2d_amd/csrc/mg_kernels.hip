// mg_kernels.hip -- the V(1,1) cycle of mg.h.
//
// Per level above the coarsest, one cycle is two launches:
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
// A thread owns one coarse cell and its 2x2 children: (i0,j0) and (i0+1,j0+1)
// are red, the other two black.  Values needed from children of other coarse
// cells are recomputed by the same expressions, so they are bitwise the values
// their owners store.

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

// Down: pre-smoothing sweep of level f (z of its cells) and coarse b.
template <class TB>
__global__ void __launch_bounds__(BLOCK) k_mg_down(MgLevel f, const TB *b, MgLevel C, const KryScal *sc) {
    if (sc->done) return;
    int I, J, cc;
    if (!mg_thread(C, I, J, cc)) return;
    const int i0 = 2 * I, j0 = 2 * J, i1 = i0 + 1, j1 = j0 + 1;
    const bool hi = i1 < f.nx, hj = j1 < f.ny;
    // black children (i1,j0), (i0,j1) and the black neighbours of the red ones
    const float zb10 = mg_black0(f, b, i1, j0), zb01 = mg_black0(f, b, i0, j1);
    const float zbW = mg_black0(f, b, i0 - 1, j0), zbS = mg_black0(f, b, i0, j0 - 1);
    const int c00 = f.idx(i0, j0);
    f.z[c00] = (float)b[c00] * f.dinv[c00];
    float r = mg_red_res(f, b, i0, j0, zb01, zb10, zbS, zbW);
    if (hi) f.z[f.idx(i1, j0)] = zb10;
    if (hj) f.z[f.idx(i0, j1)] = zb01;
    if (hi && hj) {
        const int c11 = f.idx(i1, j1);
        f.z[c11] = (float)b[c11] * f.dinv[c11];
        const float zbE = mg_black0(f, b, i1 + 1, j1), zbN = mg_black0(f, b, i1, j1 + 1);
        r += mg_red_res(f, b, i1, j1, zbN, zbE, zb10, zb01);
    }
    C.b[cc] = r;
}

// Red z after the coarse correction and the red half-sweep, or 0 outside.
// zd: z of the down sweep, zc: coarse correction (on C).
template <class TB>
__device__ __forceinline__ float mg_red1(const MgLevel &L, const TB *b, const MgLevel &C, int i, int j) {
    if (i < 0 || j < 0 || i >= L.nx || j >= L.ny) return 0.0f;
    const int c = L.idx(i, j);
    const int pN = C.idx(i >> 1, min(j + 1, L.ny - 1) >> 1), pS = C.idx(i >> 1, max(j - 1, 0) >> 1);
    const int pE = C.idx(min(i + 1, L.nx - 1) >> 1, j >> 1), pW = C.idx(max(i - 1, 0) >> 1, j >> 1);
    const float s = (L.aN[c] * (L.z[c + L.nxp] + C.zo[pN]) + L.aE[c] * (L.z[c + 1] + C.zo[pE])) +
                    (L.aS[c] * (L.z[c - L.nxp] + C.zo[pS]) + L.aW[c] * (L.z[c - 1] + C.zo[pW]));
    return ((float)b[c] + s) * L.dinv[c];
}

// Up: coarse correction and post-smoothing sweep (red, then black) of level
// L into L.zo.
template <class TB>
__global__ void __launch_bounds__(BLOCK) k_mg_up(MgLevel L, const TB *b, MgLevel C, const KryScal *sc) {
    if (sc->done) return;
    int I, J, cc;
    if (!mg_thread(C, I, J, cc)) return;
    const int i0 = 2 * I, j0 = 2 * J, i1 = i0 + 1, j1 = j0 + 1;
    const bool hi = i1 < L.nx, hj = j1 < L.ny;
    const float r00 = mg_red1(L, b, C, i0, j0);
    const float r11 = (hi && hj) ? mg_red1(L, b, C, i1, j1) : 0.0f;
    L.zo[L.idx(i0, j0)] = r00;
    if (hi && hj) L.zo[L.idx(i1, j1)] = r11;
    if (hi) {
        // black (i1,j0): red neighbours (i1,j1) N, (i1+1,j0) E, (i1,j0-1) S, (i0,j0) W
        const int c = L.idx(i1, j0);
        const float zN = r11, zE = mg_red1(L, b, C, i1 + 1, j0), zS = mg_red1(L, b, C, i1, j0 - 1), zW = r00;
        L.zo[c] = ((float)b[c] + ((L.aN[c] * zN + L.aE[c] * zE) + (L.aS[c] * zS + L.aW[c] * zW))) * L.dinv[c];
    }
    if (hj) {
        // black (i0,j1): red neighbours (i0,j1+1) N, (i1,j1) E, (i0,j0) S, (i0-1,j1) W
        const int c = L.idx(i0, j1);
        const float zN = mg_red1(L, b, C, i0, j1 + 1), zE = r11, zS = r00, zW = mg_red1(L, b, C, i0 - 1, j1);
        L.zo[c] = ((float)b[c] + ((L.aN[c] * zN + L.aE[c] * zE) + (L.aS[c] * zS + L.aW[c] * zW))) * L.dinv[c];
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
