// Host-side launchers of the HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace mrfg {

// Krylov workspace of one precision over a whole batch.
template <typename T>
struct Krylov {
    const BlockEntry *blocks = nullptr;
    int nBlk = 0;
    const SysInfo *info = nullptr;
    SysState *state = nullptr;
    double *partials = nullptr;
    int *tickets = nullptr;
    const double *sumA = nullptr;
    T *x = nullptr, *b = nullptr, *ax = nullptr, *r = nullptr, *rw = nullptr;
    T *pv = nullptr, *v = nullptr, *sv = nullptr, *t = nullptr, *ph = nullptr, *sh = nullptr;
    T *w = nullptr, *pA = nullptr;     // PCG
    T *rD = nullptr;
};

template <typename T>
hipError_t launchSpmvPlain(const MatView<T> &A, const BlockEntry *blocks, int nBlk, const T *x, T *y, hipStream_t st);
template <typename T>
hipError_t launchSetup(const MatView<T> &A, const Krylov<T> &K, bool pcg, hipStream_t st);
template <typename T>
hipError_t launchFactor(const MatView<T> &A, const LevelView &LV, const Krylov<T> &K, const MatView<double> &A64,
                        double *dwork, const int32_t *cellOfPos, int *cntC, hipStream_t st);
template <typename T>
hipError_t launchPrecondition(const MatView<T> &A, const LevelView &LV, const Krylov<T> &K, const T *in, T *out,
                              hipStream_t st);
// Forward then backward sweep with optional events around each sweep.
template <typename T>
hipError_t launchDiluSweeps(const MatView<T> &A, const LevelView &LV, const Krylov<T> &K, const T *in, T *out,
                            hipStream_t st, hipEvent_t f0, hipEvent_t f1, hipEvent_t b0, hipEvent_t b1);
template <typename T>
hipError_t launchBicgIteration(const MatView<T> &A, const LevelView &LV, const Krylov<T> &K, hipStream_t st);
template <typename T>
hipError_t launchPcgIteration(const MatView<T> &A, const LevelView &LV, const Krylov<T> &K, hipStream_t st);

hipError_t launchRefinePrepare(int nSys, const SysState *outer, SysState *inner, double innerRelTol,
                               int innerMaxIter, hipStream_t st);
hipError_t launchRefineResidual(const MatView<double> &A, const Krylov<double> &K64, hipStream_t st);
hipError_t launchRefineRhs(const Krylov<double> &K64, const Krylov<float> &K32, hipStream_t st);
hipError_t launchRefineUpdate(const Krylov<double> &K64, const Krylov<float> &K32, hipStream_t st);
hipError_t launchFillInt(int *a, int n, int v, hipStream_t st);

}  // namespace mrfg
