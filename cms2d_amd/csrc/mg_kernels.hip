// mg_kernels.hip -- the V(1,1) cycle of mg.h.
//
// Per level above the coarsest, one cycle is four launches:
//   k_mg_pre        red-black sweep from z = 0: red z = b/ap, black
//                   z = (b + sum a_k b_nb/ap_nb)/ap (the red values it needs
//                   are recomputed, so one launch does both colours)
//   k_mg_restrict   coarse b = sum over the red children of their residual,
//                   which after that sweep is sum a_k z_nb exactly (red z =
//                   b/ap cancels b - ap z) and zero on black children
//   (coarser levels, then k_mg_coarsest in one workgroup)
//   k_mg_post<RED>  red cells: z = (b + sum a_k (z_nb + zc(parent(nb))))/ap;
//                   the coarse correction of the red cell itself is replaced
//                   by this update, so only the neighbours' is needed
//   k_mg_post<BLACK> z = (b + sum a_k z_nb)/ap on black cells
// Every kernel writes only cells of one colour or reads only what an earlier
// launch wrote, so no launch reads values it also writes.
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

template <class TB>
__global__ void __launch_bounds__(BLOCK) k_mg_pre(MgLevel L, const TB *b, const KryScal *sc) {
    if (sc->done) return;
    int i, j, c;
    if (!mg_thread(L, i, j, c)) return;
    const float bc = (float)b[c];
    float z;
    if (((i + j) & 1) == 0) {
        z = bc * L.dinv[c];
    } else {
        const int cN = c + L.nxp, cE = c + 1, cS = c - L.nxp, cW = c - 1;
        const float s = (L.aN[c] * ((float)b[cN] * L.dinv[cN]) + L.aE[c] * ((float)b[cE] * L.dinv[cE])) +
                        (L.aS[c] * ((float)b[cS] * L.dinv[cS]) + L.aW[c] * ((float)b[cW] * L.dinv[cW]));
        z = (bc + s) * L.dinv[c];
    }
    L.z[c] = z;
}

__device__ __forceinline__ float mg_nbsum(const MgLevel &L, int c) {
    return (L.aN[c] * L.z[c + L.nxp] + L.aE[c] * L.z[c + 1]) + (L.aS[c] * L.z[c - L.nxp] + L.aW[c] * L.z[c - 1]);
}

__global__ void __launch_bounds__(BLOCK) k_mg_restrict(MgLevel f, MgLevel C, const KryScal *sc) {
    if (sc->done) return;
    int I, J, c;
    if (!mg_thread(C, I, J, c)) return;
    const int i0 = 2 * I, j0 = 2 * J;
    float r = mg_nbsum(f, f.idx(i0, j0));
    if (i0 + 1 < f.nx && j0 + 1 < f.ny) r += mg_nbsum(f, f.idx(i0 + 1, j0 + 1));
    C.b[c] = r;
}

// RED = 1: red cells with the coarse correction of their neighbours (zc on
// the coarse level C); RED = 0: black cells.
template <class TB, int RED>
__global__ void __launch_bounds__(BLOCK) k_mg_post(MgLevel L, const TB *b, MgLevel C, const KryScal *sc) {
    if (sc->done) return;
    int i, j, c;
    if (!mg_thread(L, i, j, c)) return;
    if ((((i + j) & 1) == 0) != (RED == 1)) return;
    float s;
    if (RED) {
        // parents of the four neighbours; a neighbour outside the grid has a
        // zero coefficient, its parent index is clamped to stay in the array
        const int pN = C.idx(i >> 1, min(j + 1, L.ny - 1) >> 1), pS = C.idx(i >> 1, max(j - 1, 0) >> 1);
        const int pE = C.idx(min(i + 1, L.nx - 1) >> 1, j >> 1), pW = C.idx(max(i - 1, 0) >> 1, j >> 1);
        s = (L.aN[c] * (L.z[c + L.nxp] + C.z[pN]) + L.aE[c] * (L.z[c + 1] + C.z[pE])) +
            (L.aS[c] * (L.z[c - L.nxp] + C.z[pS]) + L.aW[c] * (L.z[c - 1] + C.z[pW]));
    } else {
        s = mg_nbsum(L, c);
    }
    L.z[c] = ((float)b[c] + s) * L.dinv[c];
}

// Coarsest grid: MG_COARSE_SWEEPS red-black sweeps from z = 0 in one block.
template <class TB>
__global__ void __launch_bounds__(1024) k_mg_coarsest(MgLevel L, const TB *b, const KryScal *sc) {
    if (sc->done) return;
    const int n = L.nx * L.ny;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int j = t / L.nx, i = t - j * L.nx;
        L.z[L.idx(i, j)] = 0.0f;
    }
    __syncthreads();
    for (int sw = 0; sw < MG_COARSE_SWEEPS; sw++) {
        for (int col = 0; col < 2; col++) {
            for (int t = threadIdx.x; t < n; t += blockDim.x) {
                const int j = t / L.nx, i = t - j * L.nx;
                if (((i + j) & 1) != col) continue;
                const int c = L.idx(i, j);
                L.z[c] = ((float)b[c] + mg_nbsum(L, c)) * L.dinv[c];
            }
            __syncthreads();
        }
    }
}

}  // namespace cms
