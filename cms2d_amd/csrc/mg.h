// mg.h -- geometric multigrid preconditioner for the pressure-correction
// system on the Cartesian cell grid.
//
// The reference preconditions its Krylov solvers with ILUT (ilutp.f90,
// solvers.f90:185-199).  Here M^-1 is one V(1,1) cycle:
//   * cell-centred 2x2 coarsening, levels until the grid has at most
//     MG_COARSE_MAX cells, its padded layout fits MG_COARSE_LDS floats and
//     there are at least MG_MIN_LEVELS levels (or the grid is one cell);
//   * smoother: one red-black Gauss-Seidel sweep before (red first, from a zero
//     guess) and after (red first) the coarse correction;
//   * per level two launches (down: sweep + restriction; up: correction +
//     sweep): LDS tiles on level 0, one thread per coarse cell and its four
//     children on the coarser levels;
//   * restriction: sum of the children's residuals (the equations are cell
//     integrals); prolongation: the coarse value injected into its children;
//   * coarse operators re-discretised from the fine coefficients: the
//     coefficient across a coarse face is half the sum of the fine face
//     coefficients it covers (a 2-D face of twice the length at twice the
//     distance keeps ds/dn), the row excess ap - sum(a) (storage term and the
//     folded open-boundary coefficients) is summed over the children;
//   * coarsest grid: MG_COARSE_SWEEPS red-black sweeps in one workgroup.
// The hierarchy is FP32 throughout; the Krylov vectors, the SpMV and the
// stopping test stay FP64, so the converged solution does not depend on the
// preconditioner's precision.  z is FP32 and the operator is applied to the
// same z that updates x, so the recurrence stays consistent.
//
// Layout: level l has its own ghost ring (pitch nx_l+2), coefficients toward
// a neighbour outside the grid are zero, dinv = 1/ap (zero on the ghost ring
// and on rows without a diagonal), and z is zero on the ghost ring.
#pragma once
#include "common.h"

namespace cms {

constexpr int MG_MAX_LEVELS = 16;
#ifndef CMS2D_MG_COARSE_MAX
#define CMS2D_MG_COARSE_MAX 4096
#endif
#ifndef CMS2D_MG_COARSE_SWEEPS
#define CMS2D_MG_COARSE_SWEEPS 6
#endif
constexpr int MG_COARSE_MAX = CMS2D_MG_COARSE_MAX;   // cells of the coarsest grid (a multiple of 1024)
constexpr int MG_COARSE_LDS = 16384;                 // its padded size in floats (64 KB of LDS)
constexpr int MG_COARSE_SWEEPS = CMS2D_MG_COARSE_SWEEPS;
// Each coarsening quadruples the storage term of a row against its face
// coefficients, so three levels down the coarsest operator is diagonally
// dominant enough for MG_COARSE_SWEEPS sweeps to act as its solve.  One level
// down (the 64 x 64 coarsest of a 128 x 128 grid) the sweeps barely touch its
// smooth error and the pp BiCGStab needs about three times the iterations.
constexpr int MG_MIN_LEVELS = 4;

struct MgLevel {
    int nx, ny, nxp;
    float *aN, *aE, *aS, *aW;   // directed off-diagonal coefficients
    float *dinv;                // 1/ap
    float *ap;                  // ap (level 0 only: the FP32 operator of the inner solve)
    float *e;                   // row excess ap - sum(a), summed when coarsening
    float *z, *zo, *b;          // down-sweep z, cycle output z, right-hand side (zo of level 0 is the
                                // caller's M^-1 x and level 0 takes its b from the caller)
    __device__ __forceinline__ int idx(int i, int j) const { return (j + 1) * nxp + (i + 1); }
    __device__ __forceinline__ int n() const { return nx * ny; }
};

struct Mg {
    int nlev;
    MgLevel L[MG_MAX_LEVELS];
};

}  // namespace cms
