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

#include <type_traits>

#include "dd.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace mrfg {

namespace {

constexpr int kBlock = 256;       // interface kernels
constexpr int kMaxBlock = 1024;   // batch kernels: blocks of 4, 8 or 16 slices (autotuned)
constexpr double kGreat = 1e20;   // lduMatrix::great_
constexpr double kSmall = 1e-20;  // lduMatrix::small_

// Reductions carry double-double sums (hi + lo, built from the error-free
// TwoSum), so a reduced value is the sum of its terms to about 2^-100
// relative and its rounding to double does not depend on the order of
// summation: every system reduces to the same scalars alone or in a batch,
// and the same as a sequential compensated sum over its cells.
// Block-reduce NV values and store the block's partial sums (slot j as the
// double-double pair partials[(blockIdx.x * 4 + j) * 2 + {0, 1}]).  No device-scope fence is needed: the
// partials are read by the finalise kernel launched after this one.
//
// EXACT = false (the FP32 inner solves of mixed precision, whose terms carry
// 24-bit values): plain double sums in the same fixed order, about a third of
// the work; the partial's low part is zero.
template <int NV, bool EXACT = true>
__device__ void blockPartial(double (&v)[NV], double *partials)
{
    __shared__ DD sh[kMaxBlock / 64][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    DD d[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        if (EXACT) {
            d[j] = waveSum(DD{v[j], 0.0});
        } else {
            double t = v[j];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
            d[j] = DD{t, 0.0};
        }
    }
    if (lane == 0)
        for (int j = 0; j < NV; ++j) sh[wid][j] = d[j];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < NV; ++j) {
            DD t = sh[0][j];
            for (int w = 1; w < nw; ++w) {
                if (EXACT) t = ddAdd(t, sh[w][j]);
                else t.hi += sh[w][j].hi;
            }
            partials[((size_t)blockIdx.x * 4 + j) * 2] = t.hi;
            partials[((size_t)blockIdx.x * 4 + j) * 2 + 1] = t.lo;
        }
}

// the block sums of a kernel over values T: exact (double-double) for the
// FP64 path, plain double for the FP32 inner solves
template <typename T>
constexpr bool kExactSums = sizeof(T) == sizeof(double);

// lduSolver::stop with SolverPerformance::checkConvergence.  A residual
// that is not finite (NaN or Inf in the system, the field or the Krylov
// scalars) stops the system at once as diverged, so it releases its batch.
__device__ inline int stopStatus(const SysState &s)
{
    if (!isfinite(s.finalRes)) return kDiverged;
    if (s.nIter < s.minIter) return kRunning;
    bool conv = s.finalRes < s.tol || (s.relTol > kSmall && s.finalRes < s.relTol * s.initRes);
    if (conv) return kConverged;
    if (s.nIter >= s.maxIter) return kMaxIter;
    return kRunning;
}

enum FinStage {
    kFinXSum, kFinInitBicg, kFinInitPcg, kFinAlpha, kFinOmega, kFinResBicg,
    kFinPcgRho, kFinPcgAlpha, kFinResPcg, kFinRefine
};

// Scalar update of one BiCGStab / PCG stage from the system's reduced sums
// (the gSumProd / gSumMag results of the lduSolver loop).
__device__ inline void applyStage(int stage, SysState &S, const double *out)
{
    switch (stage) {
    case kFinXSum:
        S.xSum = out[0];
        break;
    case kFinInitBicg:
    case kFinInitPcg:
        // normFactor, initial residual, and the first loop pass's rho/beta
        S.normFactor = out[0] + kSmall;
        S.initRes = out[1] / S.normFactor;
        S.finalRes = S.initRes;
        S.nIter = 0;
        S.status = stopStatus(S);
        S.rho = kGreat;
        if (stage == kFinInitBicg) {
            S.alpha = 0.0;
            S.omega = kGreat;
            S.rhoOld = S.rho;
            S.rho = out[2];
            S.beta = S.rho / S.rhoOld * (S.alpha / S.omega);
        }
        break;
    case kFinAlpha:
        S.red0 = out[0];
        if (S.rho == 0.0 || out[0] == 0.0) S.status = kSingular;
        else S.alpha = S.rho / out[0];
        break;
    case kFinOmega:
        S.red0 = out[0];
        S.red1 = out[1];
        // t . t == 0: t = A M^-1 s vanished, so s = r - alpha*v did and
        // x + alpha*ph solves the system.  The step is taken with omega = 0
        // (x += alpha*ph, r = s) and the stop test decides; if it does not
        // stop, the system is singular (omega = 0 cannot continue).
        S.sZero = out[1] == 0.0;
        S.omega = S.sZero ? 0.0 : out[0] / out[1];
        break;
    case kFinResBicg:
        S.finalRes = out[0] / S.normFactor;
        S.nIter++;
        S.status = stopStatus(S);
        if (S.status == kRunning && S.sZero) S.status = kSingular;
        S.rhoOld = S.rho;
        S.rho = out[1];
        S.beta = S.rho / S.rhoOld * (S.alpha / S.omega);
        break;
    case kFinPcgRho:
        S.rhoOld = S.rho;
        S.rho = out[0];
        S.beta = S.rho / S.rhoOld;
        break;
    case kFinPcgAlpha:
        S.red0 = out[0];
        if (out[0] == 0.0 || fabs(out[0]) / S.normFactor < 1e-300) S.status = kSingular;
        else S.alpha = S.rho / out[0];
        break;
    case kFinResPcg:
    case kFinRefine:
        S.finalRes = out[0] / S.normFactor;
        S.nIter++;
        S.status = stopStatus(S);
        break;
    }
}

// One block per system: reduce the system's block partials and apply the
// stage's scalar update.  The kernel sits between two vector kernels, so it
// is written for latency: every partial load, and the state load, is issued
// before the first use (one memory round trip for systems of up to
// kFinBlock * kFinUnroll blocks), and the per-wave sums are combined by a
// wave-level tree instead of one thread.  The order of every sum is fixed, so
// the result does not depend on the other systems of the batch.
constexpr int kFinBlock = 1024;
constexpr int kFinUnroll = 8;

template <int NV>
__global__ __launch_bounds__(kFinBlock) void k_finalize(int stage, const SysInfo *info, SysState *st,
                                                        const double *partials, const int8_t *gate)
{
    const int q = blockIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const SysInfo I = info[q];
    SysState S = st[q];   // uniform: scalar loads, in flight with the partials
    DD acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = DD{0.0, 0.0};
    for (int k0 = threadIdx.x; k0 < I.nBlk; k0 += kFinBlock * kFinUnroll) {
        double2 h[kFinUnroll][NV];
#pragma unroll
        for (int u = 0; u < kFinUnroll; ++u) {
            const int k = k0 + u * kFinBlock;
            const double2 *rec = reinterpret_cast<const double2 *>(partials + (size_t)(I.blkBegin + min(k, I.nBlk - 1)) * 8);
#pragma unroll
            for (int j = 0; j < NV; ++j) h[u][j] = rec[j];
        }
#pragma unroll
        for (int u = 0; u < kFinUnroll; ++u)
            if (k0 + u * kFinBlock < I.nBlk)
#pragma unroll
                for (int j = 0; j < NV; ++j) acc[j] = ddAdd(acc[j], DD{h[u][j].x, h[u][j].y});
    }
    __shared__ DD sh[kFinBlock / 64][NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = waveSum(acc[j]);
    if (lane == 0)
        for (int j = 0; j < NV; ++j) sh[wid][j] = acc[j];
    __syncthreads();
    if (wid != 0) return;
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = waveSum(lane < kFinBlock / 64 ? sh[lane][j] : DD{0.0, 0.0});
    if (lane == 0 && S.status == kRunning && (!gate || gate[q])) {
        double out[4] = {0.0, 0.0, 0.0, 0.0};
        for (int j = 0; j < NV; ++j) out[j] = acc[j].hi + acc[j].lo;
        applyStage(stage, S, out);
        // field by field: a whole-struct copy would keep S in memory
        SysState &G = st[q];
        G.rho = S.rho;
        G.rhoOld = S.rhoOld;
        G.alpha = S.alpha;
        G.omega = S.omega;
        G.beta = S.beta;
        G.normFactor = S.normFactor;
        G.xSum = S.xSum;
        G.initRes = S.initRes;
        G.finalRes = S.finalRes;
        G.red0 = S.red0;
        G.red1 = S.red1;
        G.nIter = S.nIter;
        G.status = S.status;
        G.sZero = S.sZero;
    }
}

template <typename T, int CH>
__device__ inline T spmvRow(const MatView<T> &A, int s, int lane, int p, const T *__restrict__ x)
{
    // a padding position (no cell) has no coefficients: its row is zero, and
    // its lanes load nothing, so slices cut by a tile's skew fetch less
    if (!((A.rowMask[s] >> lane) & 1ull)) return T(0);
    T y = A.diag[p] * x[p];
    const int e0 = A.ellOff[s], e1 = A.ellOff[s + 1];
    // chunks of CH entries (CH autotuned per batch): every column, value and
    // gathered x of a chunk is loaded before the first product (loads at
    // clamped entries, so they are unconditional), then the products are
    // added in entry order -- the row sum does not depend on CH
    for (int k0 = e0; k0 < e1; k0 += CH) {
        int col[CH];
        T val[CH], xv[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const size_t e = (size_t)min(k0 + j, e1 - 1) * 64 + lane;
            col[j] = A.ellCol[e];
            val[j] = A.ellVal[e];
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) xv[j] = x[col[j]];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const T t = y + val[j] * xv[j];
            y = k0 + j < e1 ? t : y;
        }
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

template <typename T, int CH>
__global__ __launch_bounds__(kMaxBlock) void k_spmv_plain(MatView<T> A, const BlockEntry *blocks, const T *x, T *y)
{
    int sys, s, lane, p;
    if (blockRow(blocks, sys, s, lane, p)) y[p] = spmvRow<T, CH>(A, s, lane, p, x);
}

// Setup: Ax = A x0, r = b - Ax, rw = r, p = v = 0; reduce sum(x0).
template <typename T, int CH>
__global__ __launch_bounds__(kMaxBlock) void k_setup_residual(MatView<T> A, const BlockEntry *blocks, const SysInfo *info,
                                                         SysState *st, double *partials, const T *x,
                                                         const T *b, T *Ax, T *r, const int8_t *gate)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning || (gate && !gate[sys])) return;
    double v[1] = {0.0};
    if (ok) {
        T ax = spmvRow<T, CH>(A, s, lane, p, x);
        Ax[p] = ax;
        r[p] = b[p] - ax;
        v[0] = (double)x[p];
    }
    blockPartial<1, kExactSums<T>>(v, partials);
}

// normFactor (lduMatrix::solver::normFactor) and the initial residual; also
// the first iteration's rho and beta (BiCGStab) or rho seed (PCG).
template <typename T>
__global__ __launch_bounds__(kMaxBlock) void k_setup_norm(const BlockEntry *blocks, const SysInfo *info, SysState *st,
                                                     double *partials, const double *sumA, const T *Ax,
                                                     const T *b, T *r, T *rw, T *pv, T *vv, int isPCG,
                                                     const int8_t *gate)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning || (gate && !gate[sys])) return;
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
    blockPartial<3, kExactSums<T>>(v, partials);
}

// v = A ph, alpha = rho / (rw . v)
template <typename T, int CH>
__global__ __launch_bounds__(kMaxBlock) void k_spmv_alpha(MatView<T> A, const BlockEntry *blocks, const SysInfo *info,
                                                     SysState *st, double *partials, const T *ph,
                                                     const T *rw, T *v)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[1] = {0.0};
    if (ok) {
        T y = spmvRow<T, CH>(A, s, lane, p, ph);
        v[p] = y;
        acc[0] = (double)rw[p] * (double)y;
    }
    blockPartial<1, kExactSums<T>>(acc, partials);
}

// t = A sh, omega = (t . s) / (t . t)
template <typename T, int CH>
__global__ __launch_bounds__(kMaxBlock) void k_spmv_omega(MatView<T> A, const BlockEntry *blocks, const SysInfo *info,
                                                     SysState *st, double *partials, const T *sh,
                                                     const T *sv, T *t)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[2] = {0.0, 0.0};
    if (ok) {
        T y = spmvRow<T, CH>(A, s, lane, p, sh);
        t[p] = y;
        double yd = (double)y;
        acc[0] = yd * (double)sv[p];
        acc[1] = yd * yd;
    }
    blockPartial<2, kExactSums<T>>(acc, partials);
}

// p = r + beta*p - beta*omega*v
template <typename T>
__global__ __launch_bounds__(kMaxBlock) void k_bicg_p(const BlockEntry *blocks, const SysState *st, const T *r, T *pv,
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
__global__ __launch_bounds__(kMaxBlock) void k_bicg_s(const BlockEntry *blocks, const SysState *st, const T *r,
                                                 const T *v, T *sv)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (!ok || st[sys].status != kRunning) return;
    sv[p] = r[p] - (T)st[sys].alpha * v[p];
}

// x = x + alpha*ph + omega*sh; r = s - omega*t; residual and next rho
template <typename T>
__global__ __launch_bounds__(kMaxBlock) void k_bicg_xr(const BlockEntry *blocks, const SysInfo *info, SysState *st,
                                                  double *partials, T *x, T *r, const T *ph,
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
    blockPartial<2, kExactSums<T>>(acc, partials);
}

// PCG: rho = w . r, beta
template <typename T>
__global__ __launch_bounds__(kMaxBlock) void k_pcg_rho(const BlockEntry *blocks, const SysInfo *info, SysState *st,
                                                  double *partials, const T *w, const T *r)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[1] = {0.0};
    if (ok) acc[0] = (double)w[p] * (double)r[p];
    blockPartial<1, kExactSums<T>>(acc, partials);
}

// PCG: pA = wA (first iteration) or wA + beta*pA
template <typename T>
__global__ __launch_bounds__(kMaxBlock) void k_pcg_p(const BlockEntry *blocks, const SysState *st, const T *w, T *pA)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (!ok || st[sys].status != kRunning) return;
    if (st[sys].nIter == 0) pA[p] = w[p];
    else pA[p] = w[p] + (T)st[sys].beta * pA[p];
}

// PCG: wA = A pA, alpha = rho / (wA . pA)
template <typename T, int CH>
__global__ __launch_bounds__(kMaxBlock) void k_spmv_pcg(MatView<T> A, const BlockEntry *blocks, const SysInfo *info,
                                                   SysState *st, double *partials, const T *pA, T *w)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning) return;
    double acc[1] = {0.0};
    if (ok) {
        T y = spmvRow<T, CH>(A, s, lane, p, pA);
        w[p] = y;
        acc[0] = (double)y * (double)pA[p];
    }
    blockPartial<1, kExactSums<T>>(acc, partials);
}

// PCG: x += alpha*pA; r -= alpha*wA; residual
template <typename T>
__global__ __launch_bounds__(kMaxBlock) void k_pcg_xr(const BlockEntry *blocks, const SysInfo *info, SysState *st,
                                                 double *partials, T *x, T *r, const T *pA,
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
    blockPartial<1, kExactSums<T>>(acc, partials);
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 inner BiCGStab inside FP64 iterative refinement.
// The outer residual, the solution update and the convergence test stay in
// FP64 on the device.
// ---------------------------------------------------------------------------

// Fresh inner state for every system whose outer solve is still running.
__global__ void k_refine_prepare(int nSys, const SysState *outer, SysState *inner, double innerRelTol,
                                 int innerMaxIter, const int8_t *gate)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nSys || (gate && !gate[q])) return;
    SysState s = {};
    s.tol = 0.0;
    // reduce the correction's residual by innerRelTol, or only by as much as
    // the outer tolerance still needs (with a margin of 4): the last outer
    // step does not overshoot
    const double need = outer[q].finalRes > 0.0 ? 0.25 * outer[q].tol / outer[q].finalRes : 1.0;
    s.relTol = fmin(fmax(innerRelTol, need), 0.5);
    s.minIter = 0;
    s.maxIter = innerMaxIter;
    s.nCells = outer[q].nCells;
    s.status = outer[q].status == kRunning ? kRunning : kIdle;
    inner[q] = s;
}

// r = b - A x in FP64 and the outer convergence test against the FP64
// normFactor of the initial guess.
template <int CH>
__global__ __launch_bounds__(kMaxBlock) void k_refine_residual(MatView<double> A, const BlockEntry *blocks,
                                                            const SysInfo *info, SysState *st, double *partials,
                                                            const double *x, const double *b, double *r,
                                                            const int8_t *gate)
{
    int sys, s, lane, p;
    bool ok = blockRow(blocks, sys, s, lane, p);
    if (st[sys].status != kRunning || (gate && !gate[sys])) return;
    double acc[1] = {0.0};
    if (ok) {
        double rp = b[p] - spmvRow<double, CH>(A, s, lane, p, x);
        r[p] = rp;
        acc[0] = fabs(rp);
    }
    blockPartial<1>(acc, partials);
}

// b32 = (float) r64 for systems whose outer solve is running; x32 = 0
__global__ __launch_bounds__(kMaxBlock) void k_refine_rhs(const BlockEntry *blocks, const SysState *outer,
                                                       const double *r, float *b32, float *x32, const int8_t *gate)
{
    int sys, s, lane, p;
    if (!blockRow(blocks, sys, s, lane, p)) return;
    if (outer[sys].status != kRunning || (gate && !gate[sys])) return;
    b32[p] = (float)r[p];
    x32[p] = 0.0f;
}

// x64 += (double) d32
__global__ __launch_bounds__(kMaxBlock) void k_refine_update(const BlockEntry *blocks, const SysState *outer, double *x,
                                                          const float *d, const int8_t *gate)
{
    int sys, s, lane, p;
    if (!blockRow(blocks, sys, s, lane, p)) return;
    if (outer[sys].status != kRunning || (gate && !gate[sys])) return;
    x[p] += (double)d[p];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

#define MRFG_CHECK_LAUNCH() return hipGetLastError()

// number of double-double sums each stage reduces
constexpr int finSums(int stage)
{
    return stage == kFinInitBicg || stage == kFinInitPcg ? 3 : stage == kFinOmega || stage == kFinResBicg ? 2 : 1;
}

template <int Stage, typename T>
static void fin(const Krylov<T> &K, hipStream_t st, const int8_t *gate = nullptr)
{
    k_finalize<finSums(Stage)><<<K.nSys, kFinBlock, 0, st>>>(Stage, K.info, K.state, K.partials, gate);
}

// the SpMV kernels instantiated for the ELL chunk the batch was tuned to
template <typename F>
static void withChunk(int ch, F &&f)
{
    switch (ch) {
    case 2: f(std::integral_constant<int, 2>()); break;
    case 4: f(std::integral_constant<int, 4>()); break;
    case 16: f(std::integral_constant<int, 16>()); break;
    default: f(std::integral_constant<int, 8>()); break;
    }
}

template <typename T>
hipError_t launchSpmvPlain(const MatView<T> &A, const BlockEntry *blocks, int nBlk, int threads, const T *x, T *y,
                           hipStream_t st)
{
    withChunk(A.ellChunk, [&](auto c) { k_spmv_plain<T, decltype(c)::value><<<nBlk, threads, 0, st>>>(A, blocks, x, y); });
    MRFG_CHECK_LAUNCH();
}

template <typename T>
hipError_t launchSetup(const MatView<T> &A, const Krylov<T> &K, bool pcg, hipStream_t st, const int8_t *gate)
{
    withChunk(A.ellChunk, [&](auto c) {
        k_setup_residual<T, decltype(c)::value><<<K.nBlk, K.threads, 0, st>>>(A, K.blocks, K.info, K.state,
                                                                              K.partials, K.x, K.b, K.ax, K.r, gate);
    });
    fin<kFinXSum>(K, st, gate);
    k_setup_norm<T><<<K.nBlk, K.threads, 0, st>>>(K.blocks, K.info, K.state, K.partials, K.sumA, K.ax, K.b,
                                               K.r, pcg ? nullptr : K.rw, pcg ? K.pA : K.pv, pcg ? nullptr : K.v,
                                               pcg ? 1 : 0, gate);
    if (pcg) fin<kFinInitPcg>(K, st, gate);
    else fin<kFinInitBicg>(K, st, gate);
    MRFG_CHECK_LAUNCH();
}

template <typename T>
hipError_t launchPrecondition(const DiluView<T> &D, const Krylov<T> &K, const T *in, T *out, hipStream_t st)
{
    if (D.boxLone && D.nWaves == 0 && D.nBoxWaves > 0) return diluApplyLone<T>(D, K.state, in, out, st);
    hipError_t e = diluForward<T>(D, K.state, in, out, st);
    if (e != hipSuccess) return e;
    return diluBackward<T>(D, K.state, out, st);
}


template <typename T>
hipError_t launchBicgIteration(const MatView<T> &A, const DiluView<T> &D, const Krylov<T> &K, hipStream_t st,
                               const IterProbe *probe)
{
    auto mark = [&](int which, bool begin) {
        if (probe && probe->which == which) (void)hipEventRecord(begin ? probe->begin : probe->end, st);
    };
    // box-only batches: the p and s updates are computed by the forward
    // sweeps that precondition them (no update kernels)
    const bool probed = probe && (probe->which == 1 || probe->which == 2);
    const bool fuse = diluCanFuse(D) && !probed;
    if (fuse) {
        hipError_t e = diluApplyFused<T>(D, K.state, kInpP, K.r, K.pv, K.v, K.ph, st);
        if (e != hipSuccess) return e;
    } else if (!probed) {
        k_bicg_p<T><<<K.nBlk, K.threads, 0, st>>>(K.blocks, K.state, K.r, K.pv, K.v);
        hipError_t e = launchPrecondition(D, K, K.pv, K.ph, st);
        if (e != hipSuccess) return e;
    } else {
        k_bicg_p<T><<<K.nBlk, K.threads, 0, st>>>(K.blocks, K.state, K.r, K.pv, K.v);
        // the two sweeps timed one by one
        mark(1, true);
        hipError_t e = diluForward<T>(D, K.state, K.pv, K.ph, st);
        if (e != hipSuccess) return e;
        mark(1, false);
        mark(2, true);
        e = diluBackward<T>(D, K.state, K.ph, st);
        if (e != hipSuccess) return e;
        mark(2, false);
    }
    mark(3, true);
    withChunk(A.ellChunk, [&](auto c) {
        k_spmv_alpha<T, decltype(c)::value><<<K.nBlk, K.threads, 0, st>>>(A, K.blocks, K.info, K.state, K.partials,
                                                                          K.ph, K.rw, K.v);
    });
    mark(3, false);
    fin<kFinAlpha>(K, st);
    if (fuse) {
        hipError_t e = diluApplyFused<T>(D, K.state, kInpS, K.r, K.sv, K.v, K.sh, st);
        if (e != hipSuccess) return e;
    } else {
        k_bicg_s<T><<<K.nBlk, K.threads, 0, st>>>(K.blocks, K.state, K.r, K.v, K.sv);
        hipError_t e = launchPrecondition(D, K, K.sv, K.sh, st);
        if (e != hipSuccess) return e;
    }
    withChunk(A.ellChunk, [&](auto c) {
        k_spmv_omega<T, decltype(c)::value><<<K.nBlk, K.threads, 0, st>>>(A, K.blocks, K.info, K.state, K.partials,
                                                                          K.sh, K.sv, K.t);
    });
    fin<kFinOmega>(K, st);
    k_bicg_xr<T><<<K.nBlk, K.threads, 0, st>>>(K.blocks, K.info, K.state, K.partials, K.x, K.r, K.ph, K.sh,
                                            K.sv, K.t, K.rw);
    fin<kFinResBicg>(K, st);
    MRFG_CHECK_LAUNCH();
}

template <typename T>
hipError_t launchPcgIteration(const MatView<T> &A, const DiluView<T> &D, const Krylov<T> &K, hipStream_t st)
{
    {
        hipError_t e = launchPrecondition(D, K, K.r, K.w, st);
        if (e != hipSuccess) return e;
    }
    k_pcg_rho<T><<<K.nBlk, K.threads, 0, st>>>(K.blocks, K.info, K.state, K.partials, K.w, K.r);
    fin<kFinPcgRho>(K, st);
    k_pcg_p<T><<<K.nBlk, K.threads, 0, st>>>(K.blocks, K.state, K.w, K.pA);
    withChunk(A.ellChunk, [&](auto c) {
        k_spmv_pcg<T, decltype(c)::value><<<K.nBlk, K.threads, 0, st>>>(A, K.blocks, K.info, K.state, K.partials,
                                                                        K.pA, K.w);
    });
    fin<kFinPcgAlpha>(K, st);
    k_pcg_xr<T><<<K.nBlk, K.threads, 0, st>>>(K.blocks, K.info, K.state, K.partials, K.x, K.r, K.pA, K.w);
    fin<kFinResPcg>(K, st);
    MRFG_CHECK_LAUNCH();
}

hipError_t launchRefinePrepare(int nSys, const SysState *outer, SysState *inner, double innerRelTol,
                               int innerMaxIter, const int8_t *gate, hipStream_t st)
{
    k_refine_prepare<<<(nSys + 63) / 64, 64, 0, st>>>(nSys, outer, inner, innerRelTol, innerMaxIter, gate);
    MRFG_CHECK_LAUNCH();
}

hipError_t launchRefineResidual(const MatView<double> &A, const Krylov<double> &K64, const int8_t *gate,
                                hipStream_t st)
{
    withChunk(A.ellChunk, [&](auto c) {
        k_refine_residual<decltype(c)::value><<<K64.nBlk, K64.threads, 0, st>>>(A, K64.blocks, K64.info, K64.state,
                                                                               K64.partials, K64.x, K64.b, K64.r,
                                                                               gate);
    });
    fin<kFinRefine>(K64, st, gate);
    MRFG_CHECK_LAUNCH();
}

hipError_t launchRefineRhs(const Krylov<double> &K64, const Krylov<float> &K32, const int8_t *gate, hipStream_t st)
{
    k_refine_rhs<<<K64.nBlk, K64.threads, 0, st>>>(K64.blocks, K64.state, K64.r, K32.b, K32.x, gate);
    MRFG_CHECK_LAUNCH();
}

hipError_t launchRefineUpdate(const Krylov<double> &K64, const Krylov<float> &K32, const int8_t *gate,
                              hipStream_t st)
{
    k_refine_update<<<K64.nBlk, K64.threads, 0, st>>>(K64.blocks, K64.state, K64.x, K32.x, gate);
    MRFG_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// Time stepping: the ddt terms of transient systems on the device
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kMaxBlock) void k_ddt_diag(MatView<double> A, const BlockEntry *blocks, int K,
                                                     double coefft, const double *diagBase, const double *a,
                                                     double *diag, double2 *prodPack, double *sumA)
{
    int sys, s, lane, p;
    if (!blockRow(blocks, sys, s, lane, p)) return;
    const double d = diagBase[p] + coefft * a[p];
    diag[p] = d;
    const int G = (K + 2) / 2;
    ((double *)&prodPack[((size_t)s * G + K / 2) * 64 + lane])[K % 2] = d;
    // lduMatrix::sumA: the diagonal, the face coefficients in face order (the
    // row's ELL order; padding entries are zeros at the end), minus the
    // interface coefficients in interface order
    double r = d;
    for (int k = A.ellOff[s]; k < A.ellOff[s + 1]; ++k) r += A.ellVal[(size_t)k * 64 + lane];
    const unsigned long long m = A.ifMask[s];
    if ((m >> lane) & 1ull) {
        const int row = A.ifBase[s] + __popcll(m & ((1ull << lane) - 1ull));
        for (int f = A.ifRowFace[row]; f < A.ifRowFace[row + 1]; ++f) r -= A.ifCoeff[f];
    }
    sumA[p] = r;
}

__global__ __launch_bounds__(kMaxBlock) void k_ddt_source(const BlockEntry *blocks, int backward,
                                                       const double *sourceBase, const double *a, const double *x,
                                                       const double *xPrev, double *b)
{
    int sys, s, lane, p;
    if (!blockRow(blocks, sys, s, lane, p)) return;
    const double t = backward ? 2.0 * x[p] - 0.5 * xPrev[p] : x[p];
    b[p] = sourceBase[p] + a[p] * t;
}

hipError_t launchDdtDiag(const MatView<double> &A, const BlockEntry *blocks, int nBlk, int K, double coefft,
                         const double *diagBase, const double *a, double *diag, double2 *prodPack, double *sumA,
                         hipStream_t st)
{
    k_ddt_diag<<<nBlk, kBlock, 0, st>>>(A, blocks, K, coefft, diagBase, a, diag, prodPack, sumA);
    MRFG_CHECK_LAUNCH();
}

hipError_t launchDdtSource(const BlockEntry *blocks, int nBlk, bool backward, const double *sourceBase,
                           const double *a, const double *x, const double *xPrev, double *b, hipStream_t st)
{
    k_ddt_source<<<nBlk, kBlock, 0, st>>>(blocks, backward ? 1 : 0, sourceBase, a, x, xPrev, b);
    MRFG_CHECK_LAUNCH();
}

// Interface heat flow of one side of a monolithic interface,
// Q = sum_f coeff_f * (Tn_f - To_f) with Tn_f = sum_p w_p x[pos_p] (the
// patch-neighbour field through the GGI weights), as a double-double sum:
// one term per face, block partials, then one block adds the partials.
__global__ __launch_bounds__(kBlock) void k_iface_flux(int nFaces, const int32_t *fcPos, const double *coeff,
                                                     const int32_t *off, const int32_t *pairPos,
                                                     const double *w, const double *x, double *partials)
{
    const int f = blockIdx.x * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (f < nFaces) {
        double tn = 0.0;
        for (int q = off[f]; q < off[f + 1]; ++q) tn += w[q] * x[pairPos[q]];
        v[0] = coeff[f] * (tn - x[fcPos[f]]);
    }
    blockPartial<1>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_iface_flux_sum(int nBlk, const double *partials, double *out)
{
    DD acc{0.0, 0.0};
    for (int b = threadIdx.x; b < nBlk; b += kBlock)
        acc = ddAdd(acc, DD{partials[(size_t)b * 8], partials[(size_t)b * 8 + 1]});
    __shared__ DD sh[kBlock / 64];
    acc = waveSum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        DD t = sh[0];
        for (int q = 1; q < kBlock / 64; ++q) t = ddAdd(t, sh[q]);
        out[0] = t.hi + t.lo;
    }
}

hipError_t launchInterfaceFlux(int nFaces, const int32_t *fcPos, const double *coeff, const int32_t *off,
                               const int32_t *pairPos, const double *w, const double *x, double *partials,
                               double *out, hipStream_t st)
{
    const int nBlk = (nFaces + kBlock - 1) / kBlock;
    if (nBlk > 0)
        k_iface_flux<<<nBlk, kBlock, 0, st>>>(nFaces, fcPos, coeff, off, pairPos, w, x, partials);
    k_iface_flux_sum<<<1, kBlock, 0, st>>>(nBlk, partials, out);
    MRFG_CHECK_LAUNCH();
}

// monolithicTemperature's patch values on one side of an interface: the
// neighbour field tn = sum_p w_p x[pos_p] (patchNeighbourField: the GGI
// interpolate of the other side's face-cell values,
// monolithicCouplingFvPatchField.C:208-236) and the wall value
// tw = (kOwn To + kNei tn) / (kOwn + kNei) (calcTemperature with Qr = 0,
// monolithicThermalDiffusivityFvPatchScalarField.C:300-380), one thread per
// face.
__global__ __launch_bounds__(kBlock) void k_iface_patch(int nFaces, const int32_t *fcPos, const int32_t *off,
                                                      const int32_t *pairPos, const double *w, const double *kOwn,
                                                      const double *kNei, const double *x, double *tn, double *tw)
{
    const int f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= nFaces) return;
    double t = 0.0;
    for (int q = off[f]; q < off[f + 1]; ++q) t += w[q] * x[pairPos[q]];
    const double to = x[fcPos[f]];
    tn[f] = t;
    tw[f] = (kOwn[f] * to + kNei[f] * t) / (kOwn[f] + kNei[f]);
}

hipError_t launchInterfacePatch(int nFaces, const int32_t *fcPos, const int32_t *off, const int32_t *pairPos,
                                const double *w, const double *kOwn, const double *kNei, const double *x,
                                double *tn, double *tw, hipStream_t st)
{
    if (nFaces > 0)
        k_iface_patch<<<(nFaces + kBlock - 1) / kBlock, kBlock, 0, st>>>(nFaces, fcPos, off, pairPos, w, kOwn, kNei,
                                                                         x, tn, tw);
    MRFG_CHECK_LAUNCH();
}

// Plain device copy (16 bytes per thread per iteration, grid-stride): the
// reference point of the HBM traffic counters.
__global__ void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t copyProbe(int device, size_t bytes, int reps, double *ms)
{
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return e;
    const size_t n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if ((e = hipMalloc(&a, n * sizeof(uint4))) != hipSuccess) return e;
    if ((e = hipMalloc(&b, n * sizeof(uint4))) != hipSuccess) {
        (void)hipFree(a);
        return e;
    }
    (void)hipMemset(a, 0, n * sizeof(uint4));
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    k_copy_probe<<<2048, 256>>>(a, b, n);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) k_copy_probe<<<2048, 256>>>(a, b, n);
    (void)hipEventRecord(e1, nullptr);
    e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    *ms = (double)t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return e;
}

#define MRFG_INSTANTIATE(T)                                                                                          \
    template hipError_t launchSpmvPlain<T>(const MatView<T> &, const BlockEntry *, int, int, const T *, T *,         \
                                           hipStream_t);                                                             \
    template hipError_t launchSetup<T>(const MatView<T> &, const Krylov<T> &, bool, hipStream_t, const int8_t *);                    \
    template hipError_t launchPrecondition<T>(const DiluView<T> &, const Krylov<T> &, const T *, T *, hipStream_t);   \
    template hipError_t launchBicgIteration<T>(const MatView<T> &, const DiluView<T> &, const Krylov<T> &,           \
                                               hipStream_t, const IterProbe *);                                      \
    template hipError_t launchPcgIteration<T>(const MatView<T> &, const DiluView<T> &, const Krylov<T> &, hipStream_t);

MRFG_INSTANTIATE(double)
MRFG_INSTANTIATE(float)

}  // namespace mrfg
