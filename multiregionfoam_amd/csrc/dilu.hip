// DILU (coupled Cholesky preconditioner) on gfx950: factorisation and the
// forward / backward sweeps of the preconditioner, executed by a schedule of
// waves (one wavefront per workgroup).
//
// Restated arithmetic (per region; interface coefficients are not part of
// the factor, the coupled preconditioner works on each region's lduMatrix):
//   factor    d[u] = diag[u] - sum_{faces f, u(f)=u, face order} upper_f*lower_f / d[l(f)],  rD = 1/d
//   forward   w[u] = rD[u]*r[u] - sum_{lower faces, face order} rD[u]*lower_f*w[l(f)]
//   backward  w[l] -= sum_{upper faces, reverse face order} rD[l]*upper_f*w[u(f)]
//
// Each lane holds one row per step.  A dependency produced by the same wave
// at most kRing-1 steps earlier is read from the wave's LDS ring; any other
// is read from a halo slot that its producer stores with a relaxed
// agent-scope atomic.  Halo buffers are reset to a sentinel bit pattern
// before a sweep (by the previous sweep), so a consumer simply polls the
// slot until the value differs from the sentinel: no fences, no counters.
// Every wait is bounded; an expired wait raises the error flag.
//
// A step is latency-bound (one LDS round trip plus a short FMA chain), so
// the kernel keeps memory out of the critical path: each row's record is a
// few 16-byte chunks, records are prefetched kPrefetch steps ahead and halo
// values kHalf steps ahead, every prefetch load is unconditional (the
// compiler then counts outstanding loads exactly instead of draining them),
// and the fast path has no divergent branch.
#include <hip/hip_runtime.h>

#include "dilu.hpp"

namespace mrfg {

namespace {

constexpr int kPrefetch = 8;
constexpr int kHalf = kPrefetch / 2;
constexpr int kPollLimit = 1 << 20;

template <typename T>
struct Bits;
template <>
struct Bits<double> {
    using U = unsigned long long;
    static constexpr U kSent = kHaloSentinel64;
    __device__ static U to(double v) { return (U)__double_as_longlong(v); }
    __device__ static double from(U u) { return __longlong_as_double((long long)u); }
};
template <>
struct Bits<float> {
    using U = unsigned int;
    static constexpr U kSent = kHaloSentinel32;
    __device__ static U to(float v) { return (U)__float_as_uint(v); }
    __device__ static float from(U u) { return __uint_as_float(u); }
};

template <typename U>
__device__ inline U haloLoad(const U *h, int slot)
{
    return __hip_atomic_load(&h[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename U>
__device__ inline U haloWait(const U *h, int slot, U sent, int *err)
{
    U v = haloLoad(h, slot);
    int spins = 0;
    while (v == sent) {
        // once any wait has expired every later one gives up at once, so a
        // broken schedule ends the kernel quickly instead of hanging it
        if (++spins > kPollLimit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            atomicOr(err, 2);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = haloLoad(h, slot);
    }
    return v;
}

__device__ inline double elem(const double2 &v, int i) { return i == 0 ? v.x : v.y; }
__device__ inline float elem(const float4 &v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; }
__device__ inline int elem(const int4 &v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; }

// Value type of a record chunk for sweep mode MODE over values T.
template <typename T, int MODE>
struct RecordOf {
    using V = typename Pack<T>::V;
    static constexpr int N = Pack<T>::N;
};
template <typename T>
struct RecordOf<T, kDiluFactor> {
    using V = double2;
    static constexpr int N = 2;
};

}  // namespace

// MODE: kDiluFwd / kDiluBwd (T = value type) or kDiluFactor (T = double).
// HMASK: dependency slots that have an inbox (compile-time, so all inbox
// prefetch loads are unconditional code).
template <typename T, int K, int MODE, int HMASK>
__global__ __launch_bounds__(64) void k_dilu_wave(DiluView<T> V, const SysState *st, const T *__restrict__ in,
                                                  T *out, typename Bits<T>::U *inbox,
                                                  typename Bits<T>::U *inboxClear, int nClear)
{
    using U = typename Bits<T>::U;
    using R = typename RecordOf<T, MODE>::V;
    constexpr int N = RecordOf<T, MODE>::N;
    constexpr int G = packChunks(K + 1, N);      // value chunks per row
    constexpr int NC = packChunks(K, 4);         // ints holding the byte codes
    constexpr int H = packChunks(codeInts(K), 4);
    constexpr int KI = __builtin_popcount(HMASK);
    constexpr U kSent = Bits<T>::kSent;
    __shared__ T ring[kRing * 64];
    const int lane = threadIdx.x;
    const int w = MODE == kDiluBwd ? V.order[V.nWaves - 1 - blockIdx.x] : V.order[blockIdx.x];
    const int sys = V.waveSys[w];
    const bool active = MODE == kDiluFactor || st[sys].status == kRunning;
    const int s0 = V.waveSlice[w], nSteps = V.waveSteps[w];
    const int4 *codes = MODE == kDiluBwd ? V.bCodes : V.fCodes;
    const R *recs = MODE == kDiluFactor ? (const R *)V.prodPack : (const R *)(MODE == kDiluBwd ? V.bPack : V.fPack);
    const unsigned long long tStart = V.trace ? wall_clock64() : 0ull;
    unsigned long long nSlow = 0;

    struct Step {
        R v[G];
        int4 c[H];
        U hv[KI > 0 ? KI : 1];
        T a;
    };
    auto sliceOf = [&](int i) {
        i = i < nSteps ? i : nSteps - 1;
        return s0 + (MODE == kDiluBwd ? nSteps - 1 - i : i);
    };
    // the row's record, input and inbox values, kPrefetch steps ahead.  Every
    // address is computable without another load, and steps past the end
    // re-read the last step, so no load waits for another or is conditional.
    auto load = [&](int i, Step &S) {
        const int s = sliceOf(i);
#pragma unroll
        for (int g = 0; g < G; ++g) S.v[g] = recs[((size_t)s * G + g) * 64 + lane];
#pragma unroll
        for (int h = 0; h < H; ++h) S.c[h] = codes[((size_t)s * H + h) * 64 + lane];
#pragma unroll
        for (int kk = 0; kk < KI; ++kk) S.hv[kk] = haloLoad(inbox, ((s * KI + kk) << 6) + lane);
        if (MODE != kDiluFactor) S.a = in[s * 64 + lane];
    };

    if (active && nSteps > 0) {
        Step buf[kPrefetch];
#pragma unroll
        for (int u = 0; u < kPrefetch; ++u) load(u, buf[u]);
        for (int i0 = 0; i0 < nSteps; i0 += kPrefetch) {
#pragma unroll
            for (int u = 0; u < kPrefetch; ++u) {
                const int i = i0 + u;
                if (i < nSteps) {
                    Step &S = buf[u];
                    const int s = sliceOf(i);
                    int code[K];
                    T c[K];
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        code[k] = (elem(S.c[k / 16], (k / 4) % 4) >> (8 * (k % 4))) & 0xFF;
                        c[k] = (T)elem(S.v[k / N], k % N);
                    }
                    const T last = (T)elem(S.v[K / N], K % N);   // rD, or diag for the factor
                    // rare slow path (wave-uniform branch): an inbox value not yet
                    // delivered when it was prefetched is polled here, and the
                    // inbox prefetches of the following steps are refreshed so
                    // that one late producer costs one wait, not one per step
                    bool miss = false;
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        if (HMASK & (1 << k))
                            miss |= code[k] == kCodeInbox && S.hv[__builtin_popcount(HMASK & ((1 << k) - 1))] == kSent;
                    if (__builtin_expect(__any(miss), 0)) {
                        ++nSlow;
#pragma unroll
                        for (int k = 0; k < K; ++k)
                            if (HMASK & (1 << k)) {
                                const int kk = __builtin_popcount(HMASK & ((1 << k) - 1));
                                if (code[k] == kCodeInbox && S.hv[kk] == kSent)
                                    S.hv[kk] = haloWait(inbox, ((s * KI + kk) << 6) + lane, kSent, V.err);
                            }
#pragma unroll
                        for (int d = 1; d < kPrefetch; ++d) {
                            Step &F = buf[(u + d) % kPrefetch];
                            const int sf = sliceOf(i + d);
#pragma unroll
                            for (int kk = 0; kk < KI; ++kk)
                                if (F.hv[kk] == kSent) F.hv[kk] = haloLoad(inbox, ((sf * KI + kk) << 6) + lane);
                        }
                        __builtin_amdgcn_s_waitcnt(0);
                    }
                    // branch-free fast path: ring or inbox value per dependency;
                    // a missing dependency contributes an exact zero
                    T y;
                    if (MODE == kDiluFwd) y = S.a * last;
                    else if (MODE == kDiluBwd) y = S.a;
                    else y = last;
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const int slot = code[k] < 256 - 2 ? code[k] : 0;
                        T v = ring[slot];
                        if (HMASK & (1 << k)) {
                            const U h = S.hv[__builtin_popcount(HMASK & ((1 << k) - 1))];
                            v = code[k] == kCodeInbox ? Bits<T>::from(h) : v;
                        }
                        const bool has = code[k] != kCodeNone;
                        if (MODE == kDiluFactor) y = has ? y - c[k] / v : y;
                        else y = has ? y - last * c[k] * v : y;
                    }
                    const int p = s * 64 + lane;
                    // the ring is private to this wavefront and LDS operations of
                    // a wavefront execute in order: no barrier (which would also
                    // drain the prefetch loads) is needed between steps
                    ring[(i % kRing) * 64 + lane] = y;
                    __builtin_amdgcn_wave_barrier();
                    if (MODE == kDiluFactor) {
                        const double rd = V.cellOfPos[p] >= 0 ? 1.0 / (double)y : 0.0;
                        V.rDout[p] = rd;
                        constexpr int NP = Pack<double>::N;
                        constexpr int GP = packChunks(K + 1, NP);
                        ((double *)&V.fPack[((size_t)s * GP + K / NP) * 64 + lane])[K % NP] = rd;
                        ((double *)&V.bPack[((size_t)s * GP + K / NP) * 64 + lane])[K % NP] = rd;
                    } else {
                        out[p] = y;
                    }
                    // deliver the value to the inboxes of its consumers in other waves
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const int o = elem(S.c[(NC + j) / 4], (NC + j) % 4);
                        if (o >= 0)
                            __hip_atomic_store(&inbox[o], Bits<T>::to(y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    load(i + kPrefetch, S);
                }
            }
        }
    }
    if (V.trace && lane == 0) {
        unsigned long long *tr = V.trace + 4 * (size_t)blockIdx.x;
        tr[0] = tStart;
        tr[1] = wall_clock64();
        tr[2] = nSlow;
        tr[3] = (unsigned long long)nSteps;
    }
    // reset the inbox of the next sweep (nobody reads it during this one)
    if (inboxClear)
        for (int q = blockIdx.x * 64 + lane; q < nClear; q += gridDim.x * 64) inboxClear[q] = kSent;
}

// rD of the FP32 records from the FP64 factor (one thread per row)
template <int K>
__global__ void k_convert_rd(DiluView<float> V, const double *rD64, int nPos)
{
    constexpr int N = Pack<float>::N;
    constexpr int G = packChunks(K + 1, N);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nPos) return;
    const int s = p / 64, lane = p % 64;
    const float r = (float)rD64[p];
    ((float *)&V.fPack[((size_t)s * G + K / N) * 64 + lane])[K % N] = r;
    ((float *)&V.bPack[((size_t)s * G + K / N) * 64 + lane])[K % N] = r;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

template <typename T, int MODE>
static hipError_t launchK(const DiluView<T> &V, const SysState *st, const T *in, T *out,
                          typename Bits<T>::U *use, typename Bits<T>::U *clr, int nClr, hipStream_t s)
{
    const int mask = effectiveHaloMask(V.K, MODE == kDiluBwd ? V.haloMaskB : V.haloMaskF);
    switch (V.K) {
    case 3:
        if (mask == 3) k_dilu_wave<T, 3, MODE, 3><<<V.nWaves, 64, 0, s>>>(V, st, in, out, use, clr, nClr);
        else k_dilu_wave<T, 3, MODE, 7><<<V.nWaves, 64, 0, s>>>(V, st, in, out, use, clr, nClr);
        break;
    case 6: k_dilu_wave<T, 6, MODE, 63><<<V.nWaves, 64, 0, s>>>(V, st, in, out, use, clr, nClr); break;
    case 12: k_dilu_wave<T, 12, MODE, 4095><<<V.nWaves, 64, 0, s>>>(V, st, in, out, use, clr, nClr); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t diluForward(const DiluView<T> &V, const SysState *st, const T *in, T *out, hipStream_t s)
{
    using U = typename Bits<T>::U;
    return launchK<T, kDiluFwd>(V, st, in, out, (U *)V.haloF, (U *)V.haloB, V.nHaloB, s);
}

template <typename T>
hipError_t diluBackward(const DiluView<T> &V, const SysState *st, T *inout, hipStream_t s)
{
    using U = typename Bits<T>::U;
    return launchK<T, kDiluBwd>(V, st, inout, inout, (U *)V.haloB, (U *)V.haloF, V.nHaloF, s);
}

__global__ void k_fill_u64(unsigned long long *a, int n, unsigned long long v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

__global__ void k_fill_u32(unsigned int *a, int n, unsigned int v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

hipError_t diluFactor(const DiluView<double> &V, hipStream_t s)
{
    // the factor publishes d through its own halo buffer, reset first
    k_fill_u64<<<256, 256, 0, s>>>((unsigned long long *)V.haloC, V.nHaloF > 0 ? V.nHaloF : 1, kHaloSentinel64);
    return launchK<double, kDiluFactor>(V, nullptr, nullptr, nullptr, (unsigned long long *)V.haloC, nullptr, 0, s);
}

hipError_t diluConvertRD(const DiluView<float> &V32, const double *rD64, int nSlices, hipStream_t s)
{
    const int nPos = nSlices * 64, blocks = (nPos + 255) / 256;
    switch (V32.K) {
    case 3: k_convert_rd<3><<<blocks, 256, 0, s>>>(V32, rD64, nPos); break;
    case 6: k_convert_rd<6><<<blocks, 256, 0, s>>>(V32, rD64, nPos); break;
    case 12: k_convert_rd<12><<<blocks, 256, 0, s>>>(V32, rD64, nPos); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t fillHalo64(void *h, int n, hipStream_t s)
{
    k_fill_u64<<<256, 256, 0, s>>>((unsigned long long *)h, n, kHaloSentinel64);
    return hipGetLastError();
}

hipError_t fillHalo32(void *h, int n, hipStream_t s)
{
    k_fill_u32<<<256, 256, 0, s>>>((unsigned int *)h, n, kHaloSentinel32);
    return hipGetLastError();
}

template hipError_t diluForward<double>(const DiluView<double> &, const SysState *, const double *, double *,
                                        hipStream_t);
template hipError_t diluForward<float>(const DiluView<float> &, const SysState *, const float *, float *,
                                       hipStream_t);
template hipError_t diluBackward<double>(const DiluView<double> &, const SysState *, double *, hipStream_t);
template hipError_t diluBackward<float>(const DiluView<float> &, const SysState *, float *, hipStream_t);

}  // namespace mrfg
