// Device-side views of a batch of coupled block systems (shared by the
// kernels and the host driver).  All arrays are indexed by batch-global
// slice / position; column indices are batch-global positions.
#pragma once

#include <cstdint>

namespace mrfg {

// Per-system Krylov state, resident on the device.  Scalars are FP64 in
// every precision mode.
struct SysState {
    double rho, rhoOld, alpha, omega, beta;
    double normFactor, xSum, initRes, finalRes;
    double tol, relTol;
    double red[4];          // last finalised reduction values (diagnostics)
    int nIter, minIter, maxIter, status;
    int nCells, pad0, pad1, pad2;
};

enum SysStatus { kRunning = 0, kConverged = 1, kMaxIter = 2, kSingular = 3, kIdle = 4 };

// Block of up to 4 slices of one system (256 threads).
struct BlockEntry {
    int sys, slice0, slice1, pad;
};

// Per-system metadata for reductions and level synchronisation.
struct SysInfo {
    int blkBegin, nBlk;        // range in the block table
    int sliceBegin, nSlices;
    int levelBase, nLevels;    // range in the level arrays
    int nCells, pad;
};

template <typename T>
struct MatView {
    const int32_t *ellOff;
    const int32_t *ellCol;
    const T *ellVal;
    const T *diag;
    const unsigned long long *ifMask;
    const int32_t *ifBase;
    const int32_t *ifRowFace;
    const T *ifCoeff;
    const int32_t *ifPairOff;
    const int32_t *ifPairCol;
    const T *ifPairW;
    const int32_t *lOff;
    const int32_t *lCol;
    const T *lVal;
    const T *lProd;
    const int32_t *uOff;
    const int32_t *uCol;
    const T *uVal;
};

struct LevelView {
    const int32_t *sliceSys;      // per slice
    const int32_t *sliceLevel;    // per slice, level index within its system
    const int32_t *levelCount;    // per global level: slices in that level
    const int32_t *order;         // DILU slice order (level-major, systems interleaved)
    int *cntA;                    // per global level counters (two sets)
    int *cntB;
    int *err;                     // set to 1 if a bounded level wait expired
    int nLevelsTotal;
    int nOrder;
};

}  // namespace mrfg
