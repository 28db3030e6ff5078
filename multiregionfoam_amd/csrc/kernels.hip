// HIP kernels for gfx950: coupled SpMV with the GGI interface gather,
// level-synchronised DILU factor and sweeps, and the BiCGStab / PCG vector
// updates with deterministic per-system reductions.
//
// Every kernel is memory-bound (about 0.1-0.2 flop per byte); none uses
// MFMA.  Row-per-lane inside 64-row slices keeps every ELL / vector access a
// contiguous 64-element wavefront access.  FP contraction is disabled at
// build level so that each row is evaluated with the same roundings as the
// LDU face loop it restates.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"

namespace mrfg {

namespace {

constexpr int kBlock = 256;
constexpr double kGreat = 1e20;   // lduMatrix::great_
constexpr double kSmall = 1e-20;  // lduMatrix::small_

__device__ inline double waveSum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int NV>
struct Red {
    double v[NV];
};

// Block-reduce NV values, write the block partial, and let the last block of
// the system reduce all partials in block order (deterministic).  Returns
// true in every thread of the last block; `out` is valid there.
template <int NV>
__device__ bool reduceToSystem(double (&v)[NV], int blk, int sys, const SysInfo *info, double *partials,
                               int *tickets, double (&out)[NV])
{
    __shared__ double sh[kBlock / 64][NV];
    __shared__ int last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = waveSum(v[j]);
    if (lane == 0)
        for (int j = 0; j < NV; ++j) sh[wid][j] = v[j];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 0; j < NV; ++j) {
            double s = 0.0;
            for (int w = 0; w < kBlock / 64; ++w) s += sh[w][j];
            partials[(size_t)blk * 4 + j] = s;
        }
        __threadfence();
        int t = atomicAdd(&tickets[sys], 1);
        last = (t == info[sys].nBlk - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    const int b0 = info[sys].blkBegin, b1 = b0 + info[sys].nBlk;
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += kBlock)
        for (int j = 0; j < NV; ++j) acc[j] += __builtin_nontemporal_load(&partials[(size_t)b * 4 + j]);
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = waveSum(acc[j]);
    __syncthreads();
    if (lane == 0)
        for (int j = 0; j < NV; ++j) sh[wid][j] = acc[j];
    __syncthreads();
    for (int j = 0; j < NV; ++j) {
        double s = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) s += sh[w][j];
        out[j] = s;
    }
    if (threadIdx.x == 0) tickets[sys] = 0;
    return true;
}

// lduSolver::stop with SolverPerformance::checkConvergence
__device__ inline int stopStatus(const SysState &s)
{
    if (s.nIter < s.minIter) return kRunning;
    bool conv = s.finalRes < s.tol || (s.relTol > kSmall && s.finalRes < s.relTol * s.initRes);
    if (conv) return kConverged;
    if (s.nIter >= s.maxIter) return kMaxIter;
    return kRunning;
}

template <typename T>
__device__ inline T spmvRow(const MatView<T> &A, int s, int lane, int p, const T *__restrict__ x)
{
    T y = A.diag[p] * x[p];
    const int e0 = A.ellOff[s], e1 = A.ellOff[s + 1];
    for (int k = e0; k < e1; ++k) {
        const size_t e = (size_t)k * 64 + lane;
        y += A.ellVal[e] * x[A.ellCol[e]];
    }
    const unsigned long long m = A.ifMask[s];
    if ((m >> lane) & 1ull) {
        const int row = A.ifBase[s] + __popcll(m & ((1ull << lane) - 1ull));
        for (int f = A.ifRowFace[row]; f < A.ifRowFace[row + 1]; ++f) {
            T pnf = 0;
            for (int q = A.ifPairOff[f]; q < A.ifPairOff[f + 1]; ++q) pnf += A.ifPairW[q] * x[A.ifPairCol[q]];
            y -= A.ifCoeff[f] * pnf;
        }
    }
    return y;
}

__device__ inline bool blockRow(const BlockEntry *blocks, int &sys, int &s, int &lane, int &p)
{
    const BlockEntry b = blocks[blockIdx.x];
    sys = b.sys;
    s = b.slice0 + (threadIdx.x >> 6);
    lane = threadIdx.x & 63;
    p = s * 64 + lane;
    return s < b.slice1;
}

}  // namespace

// ---------------------------------------------------------------------------
// SpMV family
// ---------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv_plain(MatView<T> A, const BlockEntry *blocks, const T *x, T *y)
{
    int sys, s, lane, p;
    if (blockRow(blocks, sys, s, lane, p)) y[p] = spmvRow(A, s, lane, p, x);
}

// Setup: Ax = A x0, r = b - Ax, rw = r, p = v = 0; reduce sum(x0).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_setup_residual(MatView<T> A, const BlockEntry *blocks, const SysInfo *info,
                                                         SysState *st, double *partials, int *tickets, const T *x,
                                                         const T *b, T *Ax, T *r)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double v[1] = {0.0};
    if (ok) {
        T ax = spmvRow(A, s, lane, p, x);
        Ax[p] = ax;
        r[p] = b[p] - ax;
        v[0] = (double)x[p];
    }
    double out[1];
    if (reduceToSystem<1>(v, blockIdx.x, sys, info, partials, tickets, out) && threadIdx.x == 0)
        st[sys].xSum = out[0];
}

// normFactor (lduMatrix::solver::normFactor) and the initial residual; also
// the first iteration's rho and beta (BiCGStab) or rho seed (PCG).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_setup_norm(const BlockEntry *blocks, const SysInfo *info, SysState *st,
                                                     double *partials, int *tickets, const double *sumA, const T *Ax,
                                                     const T *b, T *r, T *rw, T *pv, T *vv, int isPCG)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    const double xRef = st[sys].xSum / (double)st[sys].nCells;
    double v[3] = {0.0, 0.0, 0.0};
    if (ok) {
        double tmp = sumA[p] * xRef;
        double rp = (double)r[p];
        v[0] = fabs((double)Ax[p] - tmp) + fabs((double)b[p] - tmp);
        v[1] = fabs(rp);
        v[2] = rp * rp;
        if (rw) rw[p] = r[p];
        if (pv) pv[p] = 0;
        if (vv) vv[p] = 0;
    }
    double out[3];
    if (reduceToSystem<3>(v, blockIdx.x, sys, info, partials, tickets, out) && threadIdx.x == 0) {
        SysState &S = st[sys];
        S.normFactor = out[0] + kSmall;
        S.initRes = out[1] / S.normFactor;
        S.finalRes = S.initRes;
        S.nIter = 0;
        S.status = stopStatus(S);
        if (isPCG) {
            S.rho = kGreat;
        } else {
            S.rho = kGreat;
            S.alpha = 0.0;
            S.omega = kGreat;
            S.rhoOld = S.rho;
            S.rho = out[2];
            S.beta = S.rho / S.rhoOld * (S.alpha / S.omega);
        }
    }
}

// v = A ph, alpha = rho / (rw . v)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv_alpha(MatView<T> A, const BlockEntry *blocks, const SysInfo *info,
                                                     SysState *st, double *partials, int *tickets, const T *ph,
                                                     const T *rw, T *v)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[1] = {0.0};
    if (ok) {
        T y = spmvRow(A, s, lane, p, ph);
        v[p] = y;
        acc[0] = (double)rw[p] * (double)y;
    }
    double out[1];
    if (reduceToSystem<1>(acc, blockIdx.x, sys, info, partials, tickets, out) && threadIdx.x == 0) {
        SysState &S = st[sys];
        S.red[0] = out[0];
        if (S.rho == 0.0 || out[0] == 0.0) S.status = kSingular;
        else S.alpha = S.rho / out[0];
    }
}

// t = A sh, omega = (t . s) / (t . t)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv_omega(MatView<T> A, const BlockEntry *blocks, const SysInfo *info,
                                                     SysState *st, double *partials, int *tickets, const T *sh,
                                                     const T *sv, T *t)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[2] = {0.0, 0.0};
    if (ok) {
        T y = spmvRow(A, s, lane, p, sh);
        t[p] = y;
        double yd = (double)y;
        acc[0] = yd * (double)sv[p];
        acc[1] = yd * yd;
    }
    double out[2];
    if (reduceToSystem<2>(acc, blockIdx.x, sys, info, partials, tickets, out) && threadIdx.x == 0) {
        SysState &S = st[sys];
        S.red[0] = out[0];
        S.red[1] = out[1];
        if (out[1] == 0.0) S.status = kSingular;
        else S.omega = out[0] / out[1];
    }
}

// p = r + beta*p - beta*omega*v
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_p(const BlockEntry *blocks, const SysState *st, const T *r, T *pv,
                                                 const T *v)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (!ok || st[sys].status != kRunning) return;
    const T beta = (T)st[sys].beta, bo = (T)(st[sys].beta * st[sys].omega);
    pv[p] = r[p] + beta * pv[p] - bo * v[p];
}

// s = r - alpha*v
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_s(const BlockEntry *blocks, const SysState *st, const T *r,
                                                 const T *v, T *sv)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (!ok || st[sys].status != kRunning) return;
    sv[p] = r[p] - (T)st[sys].alpha * v[p];
}

// x = x + alpha*ph + omega*sh; r = s - omega*t; residual and next rho
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_xr(const BlockEntry *blocks, const SysInfo *info, SysState *st,
                                                  double *partials, int *tickets, T *x, T *r, const T *ph,
                                                  const T *sh, const T *sv, const T *t, const T *rw)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[2] = {0.0, 0.0};
    if (ok) {
        const T alpha = (T)st[sys].alpha, omega = (T)st[sys].omega;
        x[p] = x[p] + alpha * ph[p] + omega * sh[p];
        T rn = sv[p] - omega * t[p];
        r[p] = rn;
        acc[0] = fabs((double)rn);
        acc[1] = (double)rw[p] * (double)rn;
    }
    double out[2];
    if (reduceToSystem<2>(acc, blockIdx.x, sys, info, partials, tickets, out) && threadIdx.x == 0) {
        SysState &S = st[sys];
        S.finalRes = out[0] / S.normFactor;
        S.nIter++;
        S.status = stopStatus(S);
        S.rhoOld = S.rho;
        S.rho = out[1];
        S.beta = S.rho / S.rhoOld * (S.alpha / S.omega);
    }
}

// PCG: rho = w . r, beta
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_rho(const BlockEntry *blocks, const SysInfo *info, SysState *st,
                                                  double *partials, int *tickets, const T *w, const T *r)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[1] = {0.0};
    if (ok) acc[0] = (double)w[p] * (double)r[p];
    double out[1];
    if (reduceToSystem<1>(acc, blockIdx.x, sys, info, partials, tickets, out) && threadIdx.x == 0) {
        SysState &S = st[sys];
        S.rhoOld = S.rho;
        S.rho = out[0];
        S.beta = S.rho / S.rhoOld;
    }
}

// PCG: pA = wA (first iteration) or wA + beta*pA
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_p(const BlockEntry *blocks, const SysState *st, const T *w, T *pA)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (!ok || st[sys].status != kRunning) return;
    if (st[sys].nIter == 0) pA[p] = w[p];
    else pA[p] = w[p] + (T)st[sys].beta * pA[p];
}

// PCG: wA = A pA, alpha = rho / (wA . pA)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv_pcg(MatView<T> A, const BlockEntry *blocks, const SysInfo *info,
                                                   SysState *st, double *partials, int *tickets, const T *pA, T *w)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[1] = {0.0};
    if (ok) {
        T y = spmvRow(A, s, lane, p, pA);
        w[p] = y;
        acc[0] = (double)y * (double)pA[p];
    }
    double out[1];
    if (reduceToSystem<1>(acc, blockIdx.x, sys, info, partials, tickets, out) && threadIdx.x == 0) {
        SysState &S = st[sys];
        S.red[0] = out[0];
        if (out[0] == 0.0 || fabs(out[0]) / S.normFactor < 1e-300) S.status = kSingular;
        else S.alpha = S.rho / out[0];
    }
}

// PCG: x += alpha*pA; r -= alpha*wA; residual
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_xr(const BlockEntry *blocks, const SysInfo *info, SysState *st,
                                                 double *partials, int *tickets, T *x, T *r, const T *pA,
                                                 const T *w)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[1] = {0.0};
    if (ok) {
        const T alpha = (T)st[sys].alpha;
        x[p] += alpha * pA[p];
        T rn = r[p] - alpha * w[p];
        r[p] = rn;
        acc[0] = fabs((double)rn);
    }
    double out[1];
    if (reduceToSystem<1>(acc, blockIdx.x, sys, info, partials, tickets, out) && threadIdx.x == 0) {
        SysState &S = st[sys];
        S.finalRes = out[0] / S.normFactor;
        S.nIter++;
        S.status = stopStatus(S);
    }
}

// ---------------------------------------------------------------------------
// DILU (Cholesky preconditioner) -- one 64-row slice per block, levels
// released through per-level counters.  Blocks are launched in level order,
// so every slice a block waits for was dispatched before it.
// ---------------------------------------------------------------------------

enum { kDiluFwd = 0, kDiluBwd = 1, kDiluFactor = 2 };

// Bounded wait: a level that never completes (which the launch order rules
// out) raises the error flag instead of hanging the device.
constexpr int kSpinLimit = 1 << 22;

__device__ inline void waitCount(int *ctr, int target, int *err)
{
    int spins = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++spins > kSpinLimit) {
            atomicOr(err, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

template <typename T, int MODE>
__global__ __launch_bounds__(64) void k_dilu(MatView<T> A, LevelView LV, const SysInfo *info, const SysState *st,
                                           const T *rD, const T *in, T *out, int *cntUse, int *cntReset)
{
    const int i = blockIdx.x;
    if (cntReset)
        for (int q = i * 64 + (int)threadIdx.x; q < LV.nLevelsTotal; q += gridDim.x * 64) cntReset[q] = 0;
    const int s = MODE == kDiluBwd ? LV.order[LV.nOrder - 1 - i] : LV.order[i];
    const int sys = LV.sliceSys[s];
    if (MODE != kDiluFactor && st[sys].status != kRunning) return;
    const SysInfo I = info[sys];
    const int lv = LV.sliceLevel[s];
    if (MODE != kDiluBwd) {
        if (lv > 0) waitCount(&cntUse[I.levelBase + lv - 1], LV.levelCount[I.levelBase + lv - 1], LV.err);
    } else {
        if (lv < I.nLevels - 1) waitCount(&cntUse[I.levelBase + lv + 1], LV.levelCount[I.levelBase + lv + 1], LV.err);
    }
    const int lane = threadIdx.x;
    const int p = s * 64 + lane;
    if (MODE == kDiluFwd) {
        const T d = rD[p];
        T y = in[p] * d;
        for (int k = A.lOff[s]; k < A.lOff[s + 1]; ++k) {
            const size_t e = (size_t)k * 64 + lane;
            const int c = A.lCol[e];
            if (c >= 0) y -= d * A.lVal[e] * out[c];
        }
        out[p] = y;
    } else if (MODE == kDiluBwd) {
        const T d = rD[p];
        T y = in[p];
        for (int k = A.uOff[s]; k < A.uOff[s + 1]; ++k) {
            const size_t e = (size_t)k * 64 + lane;
            const int c = A.uCol[e];
            if (c >= 0) y -= d * A.uVal[e] * out[c];
        }
        out[p] = y;
    } else {
        T d = A.diag[p];
        for (int k = A.lOff[s]; k < A.lOff[s + 1]; ++k) {
            const size_t e = (size_t)k * 64 + lane;
            const int c = A.lCol[e];
            if (c >= 0) d -= A.lProd[e] / out[c];
        }
        out[p] = d;
    }
    __threadfence();
    if (lane == 0) __hip_atomic_fetch_add(&cntUse[I.levelBase + lv], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// rD = 1/d on real rows, 0 on padding rows
template <typename T>
__global__ __launch_bounds__(kBlock) void k_invert(const BlockEntry *blocks, const int32_t *cellOfPos, const double *d,
                                                 T *rD)
{
    int sys, s, lane, p;
    if (!blockRow(blocks, sys, s, lane, p)) return;
    rD[p] = cellOfPos[p] >= 0 ? (T)(1.0 / d[p]) : (T)0;
}

__global__ void k_fill_int(int *a, int n, int v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 inner BiCGStab inside FP64 iterative refinement.
// The outer residual, the solution update and the convergence test stay in
// FP64 on the device.
// ---------------------------------------------------------------------------

// Fresh inner state for every system whose outer solve is still running.
__global__ void k_refine_prepare(int nSys, const SysState *outer, SysState *inner, double innerRelTol,
                                 int innerMaxIter)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nSys) return;
    SysState s = {};
    s.tol = 0.0;
    s.relTol = innerRelTol;
    s.minIter = 0;
    s.maxIter = innerMaxIter;
    s.nCells = outer[q].nCells;
    s.status = outer[q].status == kRunning ? kRunning : kIdle;
    inner[q] = s;
}

// r = b - A x in FP64 and the outer convergence test against the FP64
// normFactor of the initial guess.
__global__ __launch_bounds__(kBlock) void k_refine_residual(MatView<double> A, const BlockEntry *blocks,
                                                            const SysInfo *info, SysState *st, double *partials,
                                                            int *tickets, const double *x, const double *b, double *r)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[1] = {0.0};
    if (ok) {
        double rp = b[p] - spmvRow(A, s, lane, p, x);
        r[p] = rp;
        acc[0] = fabs(rp);
    }
    double out[1];
    if (reduceToSystem<1>(acc, blockIdx.x, sys, info, partials, tickets, out) && threadIdx.x == 0) {
        SysState &S = st[sys];
        S.finalRes = out[0] / S.normFactor;
        S.nIter++;
        S.status = stopStatus(S);
    }
}

// b32 = (float) r64 for systems whose outer solve is running; x32 = 0
__global__ __launch_bounds__(kBlock) void k_refine_rhs(const BlockEntry *blocks, const SysState *outer,
                                                       const double *r, float *b32, float *x32)
{
    int sys, s, lane, p;
    if (!blockRow(blocks, sys, s, lane, p)) return;
    if (outer[sys].status != kRunning) return;
    b32[p] = (float)r[p];
    x32[p] = 0.0f;
}

// x64 += (double) d32
__global__ __launch_bounds__(kBlock) void k_refine_update(const BlockEntry *blocks, const SysState *outer, double *x,
                                                          const float *d)
{
    int sys, s, lane, p;
    if (!blockRow(blocks, sys, s, lane, p)) return;
    if (outer[sys].status != kRunning) return;
    x[p] += (double)d[p];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

#define MRFG_CHECK_LAUNCH() return hipGetLastError()

template <typename T>
hipError_t launchSpmvPlain(const MatView<T> &A, const BlockEntry *blocks, int nBlk, const T *x, T *y, hipStream_t st)
{
    k_spmv_plain<T><<<nBlk, kBlock, 0, st>>>(A, blocks, x, y);
    MRFG_CHECK_LAUNCH();
}

template <typename T>
hipError_t launchSetup(const MatView<T> &A, const Krylov<T> &K, bool pcg, hipStream_t st)
{
    k_setup_residual<T><<<K.nBlk, kBlock, 0, st>>>(A, K.blocks, K.info, K.state, K.partials, K.tickets, K.x, K.b,
                                                   K.ax, K.r);
    k_setup_norm<T><<<K.nBlk, kBlock, 0, st>>>(K.blocks, K.info, K.state, K.partials, K.tickets, K.sumA, K.ax, K.b,
                                               K.r, pcg ? nullptr : K.rw, pcg ? K.pA : K.pv, pcg ? nullptr : K.v,
                                               pcg ? 1 : 0);
    MRFG_CHECK_LAUNCH();
}

template <typename T>
hipError_t launchFactor(const MatView<T> &A, const LevelView &LV, const Krylov<T> &K, const MatView<double> &A64,
                        double *dwork, const int32_t *cellOfPos, int *cntC, hipStream_t st)
{
    (void)A;
    k_fill_int<<<64, 256, 0, st>>>(cntC, LV.nLevelsTotal, 0);
    k_dilu<double, kDiluFactor><<<LV.nOrder, 64, 0, st>>>(A64, LV, K.info, K.state, nullptr, nullptr, dwork, cntC,
                                                           nullptr);
    k_invert<T><<<K.nBlk, kBlock, 0, st>>>(K.blocks, cellOfPos, dwork, K.rD);
    MRFG_CHECK_LAUNCH();
}

template <typename T>
hipError_t launchPrecondition(const MatView<T> &A, const LevelView &LV, const Krylov<T> &K, const T *in, T *out,
                              hipStream_t st)
{
    k_dilu<T, kDiluFwd><<<LV.nOrder, 64, 0, st>>>(A, LV, K.info, K.state, K.rD, in, out, LV.cntA, LV.cntB);
    k_dilu<T, kDiluBwd><<<LV.nOrder, 64, 0, st>>>(A, LV, K.info, K.state, K.rD, out, out, LV.cntB, LV.cntA);
    MRFG_CHECK_LAUNCH();
}

template <typename T>
hipError_t launchDiluSweeps(const MatView<T> &A, const LevelView &LV, const Krylov<T> &K, const T *in, T *out,
                            hipStream_t st, hipEvent_t f0, hipEvent_t f1, hipEvent_t b0, hipEvent_t b1)
{
    if (f0) (void)hipEventRecord(f0, st);
    k_dilu<T, kDiluFwd><<<LV.nOrder, 64, 0, st>>>(A, LV, K.info, K.state, K.rD, in, out, LV.cntA, LV.cntB);
    if (f1) (void)hipEventRecord(f1, st);
    if (b0) (void)hipEventRecord(b0, st);
    k_dilu<T, kDiluBwd><<<LV.nOrder, 64, 0, st>>>(A, LV, K.info, K.state, K.rD, out, out, LV.cntB, LV.cntA);
    if (b1) (void)hipEventRecord(b1, st);
    MRFG_CHECK_LAUNCH();
}

template <typename T>
hipError_t launchBicgIteration(const MatView<T> &A, const LevelView &LV, const Krylov<T> &K, hipStream_t st)
{
    k_bicg_p<T><<<K.nBlk, kBlock, 0, st>>>(K.blocks, K.state, K.r, K.pv, K.v);
    {
        hipError_t e = launchPrecondition(A, LV, K, K.pv, K.ph, st);
        if (e != hipSuccess) return e;
    }
    k_spmv_alpha<T><<<K.nBlk, kBlock, 0, st>>>(A, K.blocks, K.info, K.state, K.partials, K.tickets, K.ph, K.rw, K.v);
    k_bicg_s<T><<<K.nBlk, kBlock, 0, st>>>(K.blocks, K.state, K.r, K.v, K.sv);
    {
        hipError_t e = launchPrecondition(A, LV, K, K.sv, K.sh, st);
        if (e != hipSuccess) return e;
    }
    k_spmv_omega<T><<<K.nBlk, kBlock, 0, st>>>(A, K.blocks, K.info, K.state, K.partials, K.tickets, K.sh, K.sv, K.t);
    k_bicg_xr<T><<<K.nBlk, kBlock, 0, st>>>(K.blocks, K.info, K.state, K.partials, K.tickets, K.x, K.r, K.ph, K.sh,
                                            K.sv, K.t, K.rw);
    MRFG_CHECK_LAUNCH();
}

template <typename T>
hipError_t launchPcgIteration(const MatView<T> &A, const LevelView &LV, const Krylov<T> &K, hipStream_t st)
{
    {
        hipError_t e = launchPrecondition(A, LV, K, K.r, K.w, st);
        if (e != hipSuccess) return e;
    }
    k_pcg_rho<T><<<K.nBlk, kBlock, 0, st>>>(K.blocks, K.info, K.state, K.partials, K.tickets, K.w, K.r);
    k_pcg_p<T><<<K.nBlk, kBlock, 0, st>>>(K.blocks, K.state, K.w, K.pA);
    k_spmv_pcg<T><<<K.nBlk, kBlock, 0, st>>>(A, K.blocks, K.info, K.state, K.partials, K.tickets, K.pA, K.w);
    k_pcg_xr<T><<<K.nBlk, kBlock, 0, st>>>(K.blocks, K.info, K.state, K.partials, K.tickets, K.x, K.r, K.pA, K.w);
    MRFG_CHECK_LAUNCH();
}

hipError_t launchRefinePrepare(int nSys, const SysState *outer, SysState *inner, double innerRelTol,
                               int innerMaxIter, hipStream_t st)
{
    k_refine_prepare<<<(nSys + 63) / 64, 64, 0, st>>>(nSys, outer, inner, innerRelTol, innerMaxIter);
    MRFG_CHECK_LAUNCH();
}

hipError_t launchRefineResidual(const MatView<double> &A, const Krylov<double> &K64, hipStream_t st)
{
    k_refine_residual<<<K64.nBlk, kBlock, 0, st>>>(A, K64.blocks, K64.info, K64.state, K64.partials, K64.tickets,
                                                   K64.x, K64.b, K64.r);
    MRFG_CHECK_LAUNCH();
}

hipError_t launchRefineRhs(const Krylov<double> &K64, const Krylov<float> &K32, hipStream_t st)
{
    k_refine_rhs<<<K64.nBlk, kBlock, 0, st>>>(K64.blocks, K64.state, K64.r, K32.b, K32.x);
    MRFG_CHECK_LAUNCH();
}

hipError_t launchRefineUpdate(const Krylov<double> &K64, const Krylov<float> &K32, hipStream_t st)
{
    k_refine_update<<<K64.nBlk, kBlock, 0, st>>>(K64.blocks, K64.state, K64.x, K32.x);
    MRFG_CHECK_LAUNCH();
}

hipError_t launchFillInt(int *a, int n, int v, hipStream_t st)
{
    k_fill_int<<<64, 256, 0, st>>>(a, n, v);
    MRFG_CHECK_LAUNCH();
}

#define MRFG_INSTANTIATE(T)                                                                                          \
    template hipError_t launchSpmvPlain<T>(const MatView<T> &, const BlockEntry *, int, const T *, T *, hipStream_t); \
    template hipError_t launchSetup<T>(const MatView<T> &, const Krylov<T> &, bool, hipStream_t);                    \
    template hipError_t launchFactor<T>(const MatView<T> &, const LevelView &, const Krylov<T> &,                    \
                                        const MatView<double> &, double *, const int32_t *, int *, hipStream_t);     \
    template hipError_t launchPrecondition<T>(const MatView<T> &, const LevelView &, const Krylov<T> &, const T *,   \
                                              T *, hipStream_t);                                                     \
    template hipError_t launchDiluSweeps<T>(const MatView<T> &, const LevelView &, const Krylov<T> &, const T *, T *, \
                                            hipStream_t, hipEvent_t, hipEvent_t, hipEvent_t, hipEvent_t);            \
    template hipError_t launchBicgIteration<T>(const MatView<T> &, const LevelView &, const Krylov<T> &,             \
                                               hipStream_t);                                                         \
    template hipError_t launchPcgIteration<T>(const MatView<T> &, const LevelView &, const Krylov<T> &, hipStream_t);

MRFG_INSTANTIATE(double)
MRFG_INSTANTIATE(float)

}  // namespace mrfg
