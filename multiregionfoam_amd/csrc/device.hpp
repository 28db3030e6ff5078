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
    double red0, red1, red2, red3;   // last finalised reduction values (diagnostics)
    int nIter, minIter, maxIter, status;
    int nCells;
    int sZero;   // BiCGStab: s = r - alpha*v vanished in the last iteration (omega taken as 0)
    int pad1, pad2;
};

// kDiverged: a residual (or the normFactor it is scaled by) is not finite.
enum SysStatus { kRunning = 0, kConverged = 1, kMaxIter = 2, kSingular = 3, kIdle = 4, kDiverged = 5 };

// Block of up to 4 slices of one system (256 threads).
struct BlockEntry {
    int sys, slice0, slice1, pad;
};

// Per-system metadata for the reductions.
struct SysInfo {
    int blkBegin, nBlk;        // range in the block table
    int sliceBegin, nSlices;
    int waveBegin, nWaves;     // range of the system's DILU waves
    int nCells, pad;
};

template <typename T>
struct MatView {
    const int32_t *ellOff;
    const int32_t *ellCol;
    const T *ellVal;
    const T *diag;
    const unsigned long long *ifMask;
    const unsigned long long *rowMask;   // per slice: the lanes that hold a cell (the others are padding, y = 0)
    const int32_t *ifBase;
    const int32_t *ifRowFace;
    const T *ifCoeff;
    const int32_t *ifPairOff;
    const int32_t *ifPairCol;
    const T *ifPairW;
    int ellChunk;   // ELL entries loaded together per row (2, 4, 8 or 16; the launchers' template choice)
};

}  // namespace mrfg
