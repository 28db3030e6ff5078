// Host-side launchers of the HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"
#include "dilu.hpp"

namespace mrfg {

// Krylov workspace of one precision over a whole batch.
template <typename T>
struct Krylov {
    const BlockEntry *blocks = nullptr;
    int nBlk = 0;
    int threads = 256;   // 64 per slice of a block (the block table's slices per block)
    int nSys = 0;
    const SysInfo *info = nullptr;
    SysState *state = nullptr;
    double *partials = nullptr;
    const double *sumA = nullptr;
    T *x = nullptr, *b = nullptr, *ax = nullptr, *r = nullptr, *rw = nullptr;
    T *pv = nullptr, *v = nullptr, *sv = nullptr, *t = nullptr, *ph = nullptr, *sh = nullptr;
    T *w = nullptr, *pA = nullptr;     // PCG
    T *rD = nullptr;
};

template <typename T>
hipError_t launchSpmvPlain(const MatView<T> &A, const BlockEntry *blocks, int nBlk, int threads, const T *x, T *y,
                           hipStream_t st);
template <typename T>
// gate (optional, one byte per system): only systems with gate[q] != 0 are set up
hipError_t launchSetup(const MatView<T> &A, const Krylov<T> &K, bool pcg, hipStream_t st,
                       const int8_t *gate = nullptr);
template <typename T>
hipError_t launchPrecondition(const DiluView<T> &D, const Krylov<T> &K, const T *in, T *out, hipStream_t st);
// Events recorded around one launch of a BiCGStab iteration (timing only).
struct IterProbe {
    int which;              // 1 = first forward sweep, 2 = first backward sweep, 3 = SpMV with the alpha sum
    hipEvent_t begin, end;
};

template <typename T>
hipError_t launchBicgIteration(const MatView<T> &A, const DiluView<T> &D, const Krylov<T> &K, hipStream_t st,
                               const IterProbe *probe = nullptr);
template <typename T>
hipError_t launchPcgIteration(const MatView<T> &A, const DiluView<T> &D, const Krylov<T> &K, hipStream_t st);


// Heat flow through one side of a monolithic interface (FP64, double-double
// sum) into out[0]; partials needs 8 doubles per 256 faces.
hipError_t launchInterfaceFlux(int nFaces, const int32_t *fcPos, const double *coeff, const int32_t *off,
                               const int32_t *pairPos, const double *w, const double *x, double *partials,
                               double *out, hipStream_t st);

// Patch values of one interface side (neighbour field and monolithicTemperature
// wall value), one per face.
hipError_t launchInterfacePatch(int nFaces, const int32_t *fcPos, const int32_t *off, const int32_t *pairPos,
                                const double *w, const double *kOwn, const double *kNei, const double *x,
                                double *tn, double *tw, hipStream_t st);

// Average time (ms) of a plain device-to-device copy of `bytes` (reads and
// writes `bytes` per launch), for calibrating HBM traffic counters.
hipError_t copyProbe(int device, size_t bytes, int reps, double *ms);

// The refinement steps act on the systems whose outer solve runs and, when
// a gate is given, whose gate byte is set (one byte per system).
hipError_t launchRefinePrepare(int nSys, const SysState *outer, SysState *inner, double innerRelTol,
                               int innerMaxIter, const int8_t *gate, hipStream_t st);
hipError_t launchRefineResidual(const MatView<double> &A, const Krylov<double> &K64, const int8_t *gate,
                                hipStream_t st);
hipError_t launchRefineRhs(const Krylov<double> &K64, const Krylov<float> &K32, const int8_t *gate, hipStream_t st);
hipError_t launchRefineUpdate(const Krylov<double> &K64, const Krylov<float> &K32, const int8_t *gate,
                              hipStream_t st);

// Time stepping of transient systems (rows of steady systems have a = 0):
// diag = diagBase + coefft*a, written to the SpMV diagonal and to the factor
// records' diagonal slot (record width K), and the row sums sumA recomputed
// in the order of lduMatrix::sumA; source = sourceBase + a*x (Euler) or
// sourceBase + a*(2 x - 0.5 xPrev) (backward, constant time step).
hipError_t launchDdtDiag(const MatView<double> &A, const BlockEntry *blocks, int nBlk, int K, double coefft,
                         const double *diagBase, const double *a, double *diag, double2 *prodPack, double *sumA,
                         hipStream_t st);
hipError_t launchDdtSource(const BlockEntry *blocks, int nBlk, bool backward, const double *sourceBase,
                           const double *a, const double *x, const double *xPrev, double *b, hipStream_t st);

}  // namespace mrfg
