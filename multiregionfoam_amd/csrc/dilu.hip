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
// few 16-byte chunks, records and inbox values are prefetched kPrefetch
// steps ahead, every prefetch load is unconditional (the
// compiler then counts outstanding loads exactly instead of draining them),
// and the fast path has no divergent branch.
#include <hip/hip_runtime.h>
#include <type_traits>

#include "dilu.hpp"

namespace mrfg {

namespace {

constexpr int kPrefetch = 8;
constexpr int kWaveRounds = 3;   // prefetch-ring rounds per unrolled group of steps
constexpr int kWaveGroup = kWaveRounds * kPrefetch;
static_assert(kWaveGroup % kRing == 0, "a group of steps spans whole rounds of the LDS ring");
// LDS values of a wave's ring: kRing steps, then zeros up to the largest
// code byte, so that every code byte indexes the ring
constexpr int kRingSlots = 256;
static_assert(kRing * 64 <= kCodeInbox, "ring slots lie below the reserved code bytes");
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

// Raw buffer access (gfx9 buffer resource: base, byte size, 32-bit
// offsets).  Every address of the sweeps' fast paths is a wave-uniform
// resource plus a scalar step offset plus a per-lane constant, so a step
// spends no instructions on 64-bit address arithmetic.  A lane offset of kNoLane lies
// beyond every resource: the hardware drops such a store, which replaces the
// per-step exec-mask juggling of conditional stores.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kRawBufferWord3 = 0x00020000;   // 32-bit data format, no swizzle, no index stride
constexpr int kSysCoherent = 16;              // sc1: the policy of the agent-scope relaxed atomics
constexpr unsigned kNoLane = 0x80000000u;

// (the arguments are wave-uniform; readfirstlane states it, so the resource
// lives in scalar registers)
__device__ inline __amdgpu_buffer_rsrc_t rawBuffer(const void *base, size_t bytes)
{
    const unsigned long long a = (unsigned long long)base;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    int n = __builtin_amdgcn_readfirstlane((int)bytes);
    int w3 = kRawBufferWord3;
    // every resource keeps its own size and flag registers: two resources
    // sharing one would be reassembled before each use
    asm volatile("" : "+s"(n), "+s"(w3));
    return __builtin_amdgcn_make_buffer_rsrc((void *)(((unsigned long long)hi << 32) | lo), (short)0, n, w3);
}

template <typename X, int aux>
__device__ inline X bufLoad(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(X) == 16) return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, aux));
    else if constexpr (sizeof(X) == 8) return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, aux));
    else return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, aux));
}

template <typename X, int aux>
__device__ inline void bufStore(X v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(X) == 8) __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, aux);
    else __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, aux);
}

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
    __shared__ T ring[kRingSlots];
    const int lane = threadIdx.x;
    const int w = MODE == kDiluBwd ? V.order[V.nWaves - 1 - blockIdx.x] : V.order[blockIdx.x];
    const int sys = V.waveSys[w];
    const bool active = MODE == kDiluFactor || st[sys].status == kRunning;
    const int s0 = V.waveSlice[w], nSteps = V.waveSteps[w];
    const int4 *codes = MODE == kDiluBwd ? V.bCodes : V.fCodes;
    // whether any row of this wave delivers to another wave (wave-uniform)
    const bool delivers = (V.waveOut[w] & (MODE == kDiluBwd ? 2 : 1)) != 0;
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

    // FULL groups (below) address a step through per-wave resources: the
    // wave's rows of the records, codes, input and output, a scalar offset of
    // the step's slice in the wave and a per-lane constant
    constexpr bool kCanFull = MODE != kDiluFactor && HMASK == 0;
    const size_t recBytes = (size_t)nSteps * G * 64 * sizeof(R);
    const bool fullWave = kCanFull && !delivers && recBytes <= (size_t)INT_MAX;
    const auto rRec = rawBuffer(recs + (size_t)s0 * G * 64, fullWave ? recBytes : 0);
    const auto rCode = rawBuffer(codes + (size_t)s0 * H * 64, fullWave ? (size_t)nSteps * H * 64 * sizeof(int4) : 0);
    const auto rIn = rawBuffer(kCanFull ? in + (size_t)s0 * 64 : nullptr, fullWave ? (size_t)nSteps * 64 * sizeof(T) : 0);
    const auto rOut = rawBuffer(kCanFull ? out + (size_t)s0 * 64 : nullptr, fullWave ? (size_t)nSteps * 64 * sizeof(T) : 0);
    const unsigned vRec = lane * (unsigned)sizeof(R), vCode = lane * (unsigned)sizeof(int4),
                   vRow = lane * (unsigned)sizeof(T);
    auto loadF = [&](int i, Step &S) {
        const unsigned j = (unsigned)(MODE == kDiluBwd ? nSteps - 1 - i : i);
#pragma unroll
        for (int g = 0; g < G; ++g) S.v[g] = bufLoad<R, 0>(rRec, vRec + g * 64 * sizeof(R), j * G * 64 * sizeof(R));
#pragma unroll
        for (int h = 0; h < H; ++h) S.c[h] = bufLoad<int4, 0>(rCode, vCode + h * 64 * sizeof(int4), j * H * 64 * sizeof(int4));
        S.a = bufLoad<T, 0>(rIn, vRow, j * 64 * sizeof(T));
    };

    if (active && nSteps > 0) {
        // ring slots above the wave's kRing steps hold zeros: a code without
        // a ring slot (kCodeInbox, kCodeNone) reads a finite zero there
        for (int q = kRing * 64 + lane; q < kRingSlots; q += 64) ring[q] = T(0);
        __builtin_amdgcn_wave_barrier();
        Step buf[kPrefetch];
        // left to the scheduler, which reorders these loads so the loop's
        // first use waits for all of them; issuing them in consumption order
        // (as the box kernel does) made the waits exact here but the
        // tutorial's solve 1.3 % slower
#pragma unroll
        for (int u = 0; u < kPrefetch; ++u) load(u, buf[u]);
        // ring values of a step's dependencies, read as soon as the previous
        // step has written its slot (a dependency is at most kRing - 1 steps
        // back, and the slot the step overwrites is never among them), so the
        // LDS round trip overlaps the rest of the previous step.  The code
        // byte is the ring slot.
        auto ringSlot = [&](const Step &S, int k) {
            return (elem(S.c[k / 16], (k / 4) % 4) >> (8 * (k % 4))) & 0xFF;
        };
        T rv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) rv[k] = ring[ringSlot(buf[0], k)];
        // Whole groups of steps: a step past the end reads the clamped last
        // slice and stores nothing (its LDS ring slot is never read), so the
        // group has no early exits and the record waits stay exact within it.
        // A group spans kWaveRounds rounds of the prefetch ring and a whole
        // number of ring rounds, so every step's ring slot is a constant.
        //
        // A group is FULL when its steps and their prefetches all lie inside
        // the wave's steps, in a sweep without inbox values or deliveries:
        // there nothing is clamped, the store is unconditional, every address
        // is a scalar offset plus a lane constant, and each dependency term is
        // taken without a select (a missing dependency has a zero coefficient
        // and reads a zero slot, so for rD >= 0 its term is +0 and leaves the
        // value unchanged bit for bit).
        auto runGroup = [&](int i0, auto fullTag) {
            constexpr bool FULL = decltype(fullTag)::value;
#pragma unroll
            for (int uu = 0; uu < kWaveGroup; ++uu) {
                const int u = uu % kPrefetch;
                const int i = i0 + uu;
                const bool inRange = FULL || i < nSteps;
                Step &S = buf[u];
                const int s = FULL ? 0 : sliceOf(i);
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
                if constexpr (!FULL && KI > 0) {
                    bool miss = false;
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        if (HMASK & (1 << k))
                            miss |= inRange && code[k] == kCodeInbox &&
                                    S.hv[__builtin_popcount(HMASK & ((1 << k) - 1))] == kSent;
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
                }
                // branch-free: ring or inbox value per dependency; a missing
                // dependency contributes an exact zero
                T y;
                if (MODE == kDiluFwd) y = S.a * last;
                else if (MODE == kDiluBwd) y = S.a;
                else y = last;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    T v = rv[k];
                    if (HMASK & (1 << k)) {
                        const U h = S.hv[__builtin_popcount(HMASK & ((1 << k) - 1))];
                        v = code[k] == kCodeInbox ? Bits<T>::from(h) : v;
                    }
                    const bool has = code[k] != kCodeNone;
                    if (MODE == kDiluFactor) y = has ? y - c[k] / v : y;
                    else if (FULL) y = y - last * c[k] * v;
                    else y = has ? y - last * c[k] * v : y;
                }
                // the ring is private to this wavefront and LDS operations of
                // a wavefront execute in order: no barrier (which would also
                // drain the prefetch loads) is needed between steps
                ring[(uu % kRing) * 64 + lane] = y;
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int k = 0; k < K; ++k) rv[k] = ring[ringSlot(buf[(u + 1) % kPrefetch], k)];
                if (FULL) {
                    bufStore<T, 0>(y, rOut, vRow, (unsigned)(MODE == kDiluBwd ? nSteps - 1 - i : i) * 64 * sizeof(T));
                    loadF(i + kPrefetch, S);
                    continue;
                }
                const int p = s * 64 + lane;
                if (inRange) {
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
                    if (delivers) {
#pragma unroll
                        for (int j = 0; j < K; ++j) {
                            const int o = elem(S.c[(NC + j) / 4], (NC + j) % 4);
                            if (o >= 0)
                                __hip_atomic_store(&inbox[o], Bits<T>::to(y), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                        }
                    }
                }
                load(i + kPrefetch, S);
            }
        };
        for (int i0 = 0; i0 < nSteps; i0 += kWaveGroup) {
            if (kCanFull && fullWave && i0 + kWaveGroup - 1 + kPrefetch < nSteps) runGroup(i0, std::true_type{});
            else runGroup(i0, std::false_type{});
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

// ---------------------------------------------------------------------------
// Structured-box tile waves.
//
// Lane l = bl + TB*cl of a tile wave owns grid line (bl, cl) of the tile's
// TB x TC lines along the sequential axis; at step t it computes cell
// ia = t - bl - cl of its line.  Each of the cell's three dependencies was
// computed one step earlier (in processing order): along the line by the
// lane itself (kept in a register), across the line by lane l -/+ 1 and
// lane l -/+ TB (exchanged through 64 LDS values), and across the tile's
// edge by the neighbouring tile's wave, which delivers it to this wave's
// inbox: (TB + TC) values per step, [cl] for the b edge and [TC + bl] for
// the c edge.  Records hold the three coefficients in fixed slots (slot
// 2 - axis of the neighbour, which is also the reference's face order), so a
// step loads its record, its input and two inbox values, all addressed
// without another load.
// ---------------------------------------------------------------------------

// Prefetch distances.  A step is bound by its own instruction issue and LDS
// round trip, not by memory latency, and a deeper inbox prefetch makes every
// consumer tile lag its producer by that many more steps.
constexpr int kBoxAhead = 4;     // steps between a record load and its use
constexpr int kBoxInAhead = 3;   // steps between an inbox load and its use (<= kBoxAhead)
// A lone sweep has no inbox: a deeper prefetch costs it only registers, and
// keeps more of its HBM loads in flight (a wave's steps wait on HBM latency
// when the records are not in cache).
constexpr int kBoxAheadLone = 8;
constexpr int kBoxGroup = 16;    // steps per unrolled group (measured best)

// KT: the batch's record width K when it is 3 (records of two FP64 / one
// FP32 chunk, rD in the last), or 0 for a wider batch (K read from the view,
// rD in a chunk of its own).
// LONE: the tile has no neighbouring tile (alone in its region), so it has
// no inbox to read and no values to deliver.
// seg, segs: a lone sweep of segment seg of segs of the tile's rows (rows
// of the other segments are not this wave's: it neither uses nor writes them)
// INP (forward sweeps): the input is BiCGStab's update, computed and stored
// by the sweep (see SweepInput) instead of read from `in`.
template <typename T, int KT, int MODE, int AX, bool LONE, int INP = kInpPlain>
__device__ __forceinline__ void boxSweep(const DiluView<T> &V, const int32_t *m, const T *__restrict__ in, T *out,
                                         typename Bits<T>::U *inbox, T *xch, unsigned long long &nSlow, int seg,
                                         int segs, const SweepInput<T> &inp = SweepInput<T>{})
{
    using U = typename Bits<T>::U;
    using R = typename RecordOf<T, MODE>::V;
    constexpr int N = RecordOf<T, MODE>::N;
    constexpr int G2 = 2 / N + 1;                // chunks holding slots 0..2
    constexpr bool kExtra = KT == 0;             // rD (diag) in a chunk of its own
    constexpr int NL = G2 + (kExtra ? 1 : 0);
    constexpr U kSent = Bits<T>::kSent;
    constexpr bool kFwd = MODE != kDiluBwd;
    constexpr int kAhead = LONE ? kBoxAheadLone : kBoxAhead;   // record prefetch distance (steps)
    // record slot of the dependency along the sequential axis and along b, c
    constexpr int kAxB = AX == 0 ? 1 : 0;
    constexpr int kSlotA = 2 - AX, kSlotB = 2 - kAxB;   // the c neighbour holds the third slot
    const int K = KT ? KT : V.K;
    const int G = packChunks(K + 1, N);          // chunks per row of the record array
    const int gLast = K / N, eLast = K % N;      // chunk / element of rD (diag)
    const R *recs = MODE == kDiluFactor ? (const R *)V.prodPack : (const R *)(MODE == kDiluBwd ? V.bPack : V.fPack);

    const unsigned lane = threadIdx.x;
    const int TB = m[kBmTB], lg = __builtin_ctz(TB), TC = 64 >> lg, TBC = TB + TC;
    const int nA = m[kBmNA], actB = m[kBmActB], actC = m[kBmActC];
    const int s0 = m[kBmSlice], nSteps = m[kBmSteps], base = m[kBmInbox];
    // a LONE sweep treats every tile as alone: a tile that has neighbours
    // then applies the DILU of its own rows only (the tile-local
    // preconditioner, whose records hold no coefficient across tile edges)
    const int prevB = LONE ? -1 : m[kBmPrevB], prevC = LONE ? -1 : m[kBmPrevC];
    const int nextB = LONE ? -1 : m[kBmNextB], nextC = LONE ? -1 : m[kBmNextC];
    const int bl = lane & (TB - 1), cl = lane >> lg;
    const bool inTile = bl < actB && cl < actC;
    // this wave's rows [iaLo, iaHi) along the sequential axis and its steps
    // tFirst..tLast (the whole tile: 0..nSteps-1); a row at step t is
    // ia = t - bl - cl
    int iaLo = 0, iaHi = nA;
    if (LONE && segs > 1) boxSegment(nA, seg, segs, iaLo, iaHi);
    const int tLo = (actB - 1) + (actC - 1);   // rows valid from step iaLo + tLo on for every lane
    const int tFirst = iaLo, tLast = iaHi - 1 + tLo, nLocal = tLast - tFirst + 1;
    const bool cut = LONE && segs > 1;
    if (nLocal <= 0) return;   // a segment without rows (nA < segs)
    // a lane's row at step t that belongs to no segment (padding) or to this
    // one: the only rows a segment wave writes
    // (bitwise, so that it compiles to a few compares and no exec-mask branches)
    auto mine = [&](int ia) {
        return !inTile | ((unsigned)ia >= (unsigned)nA) | ((unsigned)(ia - iaLo) < (unsigned)(iaHi - iaLo));
    };
    // where each dependency comes from: the lane itself (sequential axis),
    // a neighbouring lane, or the inbox on the tile's edge
    const bool edgeB = kFwd ? bl == 0 : bl == TB - 1;
    const bool edgeC = kFwd ? cl == 0 : cl == TC - 1;
    const bool hasB = kFwd ? (bl > 0 || prevB >= 0) : (bl + 1 < actB || (edgeB && nextB >= 0));
    const bool hasC = kFwd ? (cl > 0 || prevC >= 0) : (cl + 1 < actC || (edgeC && nextC >= 0));
    const bool useB = edgeB && hasB, useC = edgeC && hasC;
    const unsigned srcB = (kFwd ? lane - 1 : lane + 1) & 63;
    const unsigned srcC = (kFwd ? lane - TB : lane + TB) & 63;
    // where this lane's values go: the consumer's step is t -/+ (TB - 1)
    // (resp. TC - 1), into the consumer's inbox at [cl] (resp. [TC + bl])
    const int toB = kFwd ? nextB : prevB, toC = kFwd ? nextC : prevC;
    const bool sendB = (kFwd ? bl == TB - 1 : bl == 0) && toB >= 0;
    const bool sendC = (kFwd ? cl == TC - 1 : cl == 0) && toC >= 0;
    const int dB = kFwd ? 1 - TB : TB - 1, dC = kFwd ? 1 - TC : TC - 1;

    // the wave's resources: its slices of the records, of in / out (factor:
    // of rD and of the two record arrays), its own inbox and its consumers'
    const int rowBytes = G * 64 * (int)sizeof(R);
    const int inRowBytes = TBC * (int)sizeof(U);
    const R *myRecs = recs + (size_t)s0 * G * 64;
    const auto rRec = rawBuffer(myRecs, (size_t)nSteps * rowBytes);
    const size_t rowsBytes = (size_t)nSteps * 64 * sizeof(T);
    const auto rIn = MODE == kDiluFactor ? rawBuffer(nullptr, 0)
                                          : rawBuffer((INP ? inp.r : in) + (size_t)s0 * 64, rowsBytes);
    // the fused input's other operands: the vector it is stored to (p, read
    // first for its old value; s) and v
    const auto rDst = INP ? rawBuffer(inp.dst + (size_t)s0 * 64, rowsBytes) : rawBuffer(nullptr, 0);
    const auto rV = INP ? rawBuffer(inp.v + (size_t)s0 * 64, rowsBytes) : rawBuffer(nullptr, 0);
    const auto rOut = MODE == kDiluFactor ? rawBuffer(nullptr, 0) : rawBuffer(out + (size_t)s0 * 64, rowsBytes);
    const U *myIn = inbox + base;
    const auto rInbox = rawBuffer(myIn, LONE ? 0 : (size_t)nSteps * inRowBytes);
    auto consumer = [&](int to) {
        if (LONE || to < 0) return rawBuffer(inbox, 0);
        const int32_t *mt = V.boxMeta + (size_t)to * kBoxMeta;
        return rawBuffer(inbox + mt[kBmInbox], (size_t)mt[kBmSteps] * inRowBytes);
    };
    const auto rToB = consumer(toB), rToC = consumer(toC);
    const unsigned vRec = lane * (unsigned)sizeof(R), vRow = lane * (unsigned)sizeof(T);
    const unsigned vInB = (unsigned)cl * sizeof(U), vInC = (unsigned)(TC + bl) * sizeof(U);
    const unsigned vToB = sendB ? (unsigned)cl * sizeof(U) : kNoLane;
    const unsigned vToC = sendC ? (unsigned)(TC + bl) * sizeof(U) : kNoLane;

    struct Step {
        R v[NL];
        T a;
        T pOld, vv;   // INP: the update's other operands
    };
    // step i of the sweep (clamped: loads past the end re-read the last step)
    auto stepOf = [&](int i) {
        i = i < nLocal ? i : nLocal - 1;
        return kFwd ? tFirst + i : tLast - i;
    };
    auto load = [&](int i, Step &S) {
        const int j = stepOf(i);
        const unsigned so = (unsigned)(j * rowBytes);
#pragma unroll
        for (int g = 0; g < G2; ++g) S.v[g] = bufLoad<R, 0>(rRec, vRec + g * 64 * sizeof(R), so);
        if (kExtra) S.v[NL - 1] = bufLoad<R, 0>(rRec, vRec + gLast * 64 * sizeof(R), so);
        if (MODE != kDiluFactor) S.a = bufLoad<T, 0>(rIn, vRow, (unsigned)(j * 64 * sizeof(T)));
        if (INP == kInpP) S.pOld = bufLoad<T, 0>(rDst, vRow, (unsigned)(j * 64 * sizeof(T)));
        if (INP) S.vv = bufLoad<T, 0>(rV, vRow, (unsigned)(j * 64 * sizeof(T)));
    };
    auto loadIn = [&](int i, U &hb, U &hc) {
        const unsigned so = (unsigned)(stepOf(i) * inRowBytes);
        hb = bufLoad<U, kSysCoherent>(rInbox, vInB, so);
        hc = bufLoad<U, kSysCoherent>(rInbox, vInC, so);
    };
    // A FULL group (below) addresses every stream of step t through one scalar
    // offset per array and per-lane constants that hold the rest: the records
    // and input of step t + dA and the output of step t share the offset of
    // step min(t, t + dA); the inbox row of step t + dI and the consumers'
    // rows of steps t + dB and t + dC share that of step t + mOff, the lowest
    // of them.  The lane constants are non-negative and every address stays
    // inside its array (a FULL group only addresses steps inside the sweep).
    const int dA = kFwd ? kAhead : -kAhead, dI = kFwd ? kBoxInAhead : -kBoxInAhead;
    const int dBe = toB >= 0 ? dB : 0, dCe = toC >= 0 ? dC : 0;
    const int mOff = min(min(dI, 0), min(dBe, dCe));
    const unsigned vRecF = vRec + (kFwd ? (unsigned)(dA * rowBytes) : 0u);
    const unsigned vInF = vRow + (kFwd ? (unsigned)(dA * 64 * sizeof(T)) : 0u);
    const unsigned vOutF = vRow + (kFwd ? 0u : (unsigned)(-dA * 64 * sizeof(T)));
    const unsigned vInBF = vInB + (unsigned)((dI - mOff) * inRowBytes);
    const unsigned vInCF = vInC + (unsigned)((dI - mOff) * inRowBytes);
    const unsigned vToBF = vToB == kNoLane || !inTile ? kNoLane : vToB + (unsigned)((dB - mOff) * inRowBytes);
    const unsigned vToCF = vToC == kNoLane || !inTile ? kNoLane : vToC + (unsigned)((dC - mOff) * inRowBytes);
    auto loadF = [&](int t, Step &S) {
        const int tm = kFwd ? t : t + dA;
        const unsigned so = (unsigned)(tm * rowBytes);
#pragma unroll
        for (int g = 0; g < G2; ++g) S.v[g] = bufLoad<R, 0>(rRec, vRecF + g * 64 * sizeof(R), so);
        if (kExtra) S.v[NL - 1] = bufLoad<R, 0>(rRec, vRecF + gLast * 64 * sizeof(R), so);
        if (MODE != kDiluFactor) S.a = bufLoad<T, 0>(rIn, vInF, (unsigned)(tm * 64 * sizeof(T)));
        if (INP == kInpP) S.pOld = bufLoad<T, 0>(rDst, vInF, (unsigned)(tm * 64 * sizeof(T)));
        if (INP) S.vv = bufLoad<T, 0>(rV, vInF, (unsigned)(tm * 64 * sizeof(T)));
    };
    auto loadInF = [&](int t, U &hb, U &hc) {
        const unsigned so = (unsigned)((t + mOff) * inRowBytes);
        hb = bufLoad<U, kSysCoherent>(rInbox, vInBF, so);
        hc = bufLoad<U, kSysCoherent>(rInbox, vInCF, so);
    };

    // the exchange values start finite: a missing neighbour's zero
    // coefficient multiplies whatever its lane holds
    xch[lane] = T(0);
    __builtin_amdgcn_wave_barrier();
    // the neighbours' values for the next step, read as soon as they are
    // written (below), so the LDS round trip overlaps the rest of the step
    T xbN = xch[srcB], xcN = xch[srcC];
    Step buf[kAhead];
    U hbuf[kAhead], cbuf[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) hbuf[u] = cbuf[u] = kSent;   // a lone tile never loads them
    // the first steps' loads in the order they are consumed (not reordered):
    // the loop's first use then waits only for the loads it needs
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
        if (!LONE && u < kBoxInAhead) loadIn(u, hbuf[u], cbuf[u]);
        load(u, buf[u]);
        __builtin_amdgcn_sched_barrier(0);
    }
    T yprev = T(0);
    // The loop runs whole groups of steps: a step past the end is invalid for
    // every lane, reads clamped addresses and stores nothing.  The body is
    // unrolled over whole rounds of the prefetch ring, every record wait
    // exact (the initial loads above are issued in order for that); 16 steps
    // measured fastest (8 and 32 were slower).
    //
    // A group is FULL when every in-tile row of its steps is valid and its
    // prefetches stay inside the sweep: there the validity mask is the
    // loop-invariant tile mask, the output store is unconditional, no
    // prefetch index is clamped, and every delivery offset lies inside its
    // consumer's inbox, so a step spends no instructions on those tests.
    // For the fluid block of the duct that is 9 of every 11 groups.
    constexpr int kGroup = kBoxGroup;
    static_assert(kGroup % kAhead == 0, "a group spans whole rounds of the prefetch ring");
    // (offTag: the prefetch-ring slot of the group's first step; cutTag: the
    // sweep is a segment of its tile, so rows of other segments are not
    // stored)
    auto runGroup = [&](int i0, auto fullTag, auto stepsTag, auto offTag, auto cutTag) {
        constexpr bool FULL = decltype(fullTag)::value;
        constexpr bool CUT = decltype(cutTag)::value;
#pragma unroll
        for (int uu = 0; uu < decltype(stepsTag)::value; ++uu) {
            const int u = (decltype(offTag)::value + uu) % kAhead;
            const int i = i0 + uu;
            const bool inRange = FULL || i < nLocal;
            Step &S = buf[u];
            U &hb = hbuf[u];
            U &hc = cbuf[u];
            const int t = kFwd ? tFirst + i : tLast - i;
            const int ia = t - bl - cl;
            const bool valid = FULL ? inTile : inTile && (unsigned)(ia - iaLo) < (unsigned)(iaHi - iaLo);
            const bool hasA = kFwd ? ia > iaLo : ia + 1 < iaHi;
            const bool needB = valid & useB, needC = valid & useC;
            T xb = xbN, xc = xcN;
            const T vA = yprev;
            auto slotVal = [&](int q) -> T { return (T)elem(S.v[q / N], q % N); };
            T last;
            if (kExtra) last = (T)elem(S.v[NL - 1], eLast);
            else last = (T)elem(S.v[3 / N], 3 % N);
            // the row's input: BiCGStab's p or s update where fused (the
            // expressions of k_bicg_p / k_bicg_s), else the loaded value
            T ain = S.a;
            if (INP == kInpP) ain = S.a + inp.c1 * S.pOld - inp.c2 * S.vv;
            else if (INP == kInpS) ain = S.a - inp.c1 * S.vv;
            // the parts of the row value that need no neighbour value (the
            // sweeps' first term and coefficient products)
            T y0 = T(0), lc[3] = {T(0), T(0), T(0)}, pA = T(0);
            if (MODE != kDiluFactor) {
                y0 = MODE == kDiluFwd ? ain * last : ain;
#pragma unroll
                for (int q = 0; q < 3; ++q) lc[q] = last * slotVal(q);
                pA = lc[kSlotA] * vA;   // the term along the sequential axis
            }
            // both exchange reads are issued by every lane (not sunk into the
            // select), so the step has no divergent code; the products above
            // are operands too, so they are computed before the wait for the
            // exchange instead of after it, on the step's dependency chain
            asm volatile("" : "+v"(xb), "+v"(xc), "+v"(y0), "+v"(lc[0]), "+v"(lc[1]), "+v"(lc[2]), "+v"(pA));
            // the row's value from its neighbours' values
            auto rowValue = [&](U hbv, U hcv) -> T {
                const T vB = useB ? Bits<T>::from(hbv) : xb;
                const T vC = useC ? Bits<T>::from(hcv) : xc;
                T y = MODE == kDiluFactor ? last : y0;
                // the three terms in slot order, which is the face order:
                // y - (last*c)*v each (the products first, so that the second
                // is not issued behind the first subtraction).  In the sweeps
                // a missing neighbour has a zero coefficient and a finite
                // value (exchanged values of invalid rows are zero), so its
                // term is an exact zero; the factor divides and selects
                // instead.
                T pr[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const T v = q == kSlotA ? vA : q == kSlotB ? vB : vC;
                    if (MODE == kDiluFactor) {
                        const bool has = q == kSlotA ? hasA : q == kSlotB ? hasB : hasC;
                        y = has ? y - slotVal(q) / v : y;
                    } else {
                        pr[q] = q == kSlotA ? pA : lc[q] * v;
                    }
                }
                if (MODE != kDiluFactor) {
                    // (kept apart: otherwise the scheduler interleaves them again)
                    asm volatile("" : "+v"(pr[kSlotB]), "+v"(pr[3 - kSlotA - kSlotB]));
#pragma unroll
                    for (int q = 0; q < 3; ++q) y = y - pr[q];
                }
                return y;
            };
            // the value is computed with the prefetched inbox values first and
            // the exchange started; whether one of them had not arrived is
            // tested after that, off the step's dependency chain.  The inbox
            // sentinel is a NaN, and a NaN operand makes the row's value NaN,
            // so one test of the value covers both inbox values.  A value that
            // is NaN for another reason (a diverged system) only takes the
            // slow path, which then waits for nothing.
            // The step runs twice at most: a second pass follows a wait for
            // the late values (wave-uniform), which also reloads the inbox
            // values already requested for later steps.  The second pass
            // redoes the row value and the exchange (read again, as it may
            // have taken the provisional values).
            T y, yz;
            for (int pass = 0;; ++pass) {
                y = rowValue(hb, hc);
                yz = valid ? y : T(0);
                // all lanes have read the previous step's values (LDS
                // operations of a wavefront execute in order) before any
                // overwrites them
                xch[lane] = yz;
                __builtin_amdgcn_wave_barrier();
                xbN = xch[srcB];
                xcN = xch[srcC];
                // the reads stay here, so their latency overlaps the step's
                // stores and prefetches
                asm volatile("" ::: "memory");
                // (invalid rows hold an exact zero, so the test needs no mask)
                if (LONE || pass || __builtin_expect(__builtin_amdgcn_ballot_w64(yz != yz) == 0, 1)) break;
                ++nSlow;
                // (the slow path's addresses are computed here, not hoisted
                // into every step)
                int is = i;
                asm volatile("" : "+s"(is));
                // Poll rounds: each reloads this step's inbox values and those
                // already requested for the next steps, all in flight together,
                // until this step's have arrived.  The later steps' values then
                // come with the last round instead of a round trip of their own.
                // Every lane loads and selects (no divergent code), so the ring
                // keeps its registers.
                const U *ib[kBoxInAhead];
#pragma unroll
                for (int d = 0; d < kBoxInAhead; ++d) ib[d] = myIn + (size_t)stepOf(is + d) * TBC;
                for (int spins = 0;; ++spins) {
                    U nb[kBoxInAhead], nc[kBoxInAhead];
#pragma unroll
                    for (int d = 0; d < kBoxInAhead; ++d) {
                        nb[d] = haloLoad(ib[d], cl);
                        nc[d] = haloLoad(ib[d], TC + bl);
                    }
                    hb = hb == kSent ? nb[0] : hb;
                    hc = hc == kSent ? nc[0] : hc;
#pragma unroll
                    for (int d = 1; d < kBoxInAhead; ++d) {
                        U &vb = hbuf[(u + d) % kAhead];
                        U &vc = cbuf[(u + d) % kAhead];
                        vb = vb == kSent ? nb[d] : vb;
                        vc = vc == kSent ? nc[d] : vc;
                    }
                    const bool late = (needB && hb == kSent) || (needC && hc == kSent);
                    if (__builtin_amdgcn_ballot_w64(late) == 0) break;
                    // an expired wait ends every later one at once, so a
                    // broken schedule ends the kernel instead of hanging it
                    if (spins >= kPollLimit ||
                        __hip_atomic_load(V.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        if (threadIdx.x == 0) atomicOr(V.err, 2);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            yprev = yz;
            if (inRange) {
                const unsigned so = (unsigned)t * 64 * sizeof(T);
                if (MODE == kDiluFactor) {
                    const double rd = valid ? 1.0 / (double)y : 0.0;
                    if (FULL || !CUT || mine(ia)) (V.rDout + (size_t)(s0 + t) * 64)[lane] = rd;
                    if (V.fPack) {   // (a tile-local factor writes rD only)
                        constexpr int NP = Pack<double>::N;
                        const int GP = packChunks(K + 1, NP);
                        const int s = s0 + t;
                        ((double *)&V.fPack[((size_t)s * GP + K / NP) * 64 + lane])[K % NP] = rd;
                        ((double *)&V.bPack[((size_t)s * GP + K / NP) * 64 + lane])[K % NP] = rd;
                    }
                } else if (FULL) {
                    bufStore<T, 0>(yz, rOut, vOutF, (unsigned)((kFwd ? t : t + dA) * 64 * sizeof(T)));
                    if (INP) bufStore<T, 0>(ain, rDst, vOutF, (unsigned)((kFwd ? t : t + dA) * 64 * sizeof(T)));
                } else {
                    // (arithmetic select: no branch, no exec-mask moves)
                    const unsigned vSt = CUT ? vRow | (unsigned)!mine(ia) * kNoLane : vRow;
                    bufStore<T, 0>(yz, rOut, vSt, so);
                    if (INP) bufStore<T, 0>(ain, rDst, vSt, so);
                }
            }
            if (!LONE) {
                // the deliveries: every lane stores, those without a value for
                // a consumer (not on the edge, or an invalid row) at kNoLane.
                // A valid row's consumer step is valid too, so its offset lies
                // inside the consumer's inbox; an invalid row's step offset is
                // only kept non-negative.  In a FULL group every in-tile row is
                // valid and every step offset lies inside the consumer's inbox.
                const U yb = Bits<T>::to(y);
                if (FULL) {
                    const unsigned so = (unsigned)((t + mOff) * inRowBytes);
                    bufStore<U, kSysCoherent>(yb, rToB, vToBF, so);
                    bufStore<U, kSysCoherent>(yb, rToC, vToCF, so);
                } else {
                    const int tb = t + dB, tc = t + dC;
                    bufStore<U, kSysCoherent>(yb, rToB, valid ? vToB : kNoLane,
                                              (unsigned)((tb > 0 ? tb : 0) * inRowBytes));
                    bufStore<U, kSysCoherent>(yb, rToC, valid ? vToC : kNoLane,
                                              (unsigned)((tc > 0 ? tc : 0) * inRowBytes));
                }
            }
            // the inbox values of a later step first, so that waiting for them
            // leaves the record loads issued after them in flight
            U &hbF = hbuf[(u + kBoxInAhead) % kAhead], &hcF = cbuf[(u + kBoxInAhead) % kAhead];
            if (FULL) {
                if (!LONE) loadInF(t, hbF, hcF);
                loadF(t, S);
            } else {
                if (!LONE) loadIn(i + kBoxInAhead, hbF, hcF);
                load(i + kAhead, S);
            }
            // keep each step's prefetches where they are (not sunk towards
            // their use)
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // the last steps (at most kGroup - kTail) run in groups of kTail, so a
    // sweep runs at most kTail - 1 steps past its end
    constexpr int kTail = 4;
    static_assert(kAhead % kTail == 0, "tail groups start on a prefetch-ring slot that is a multiple of kTail");
    using GroupSteps = std::integral_constant<int, kGroup>;
    using TailSteps = std::integral_constant<int, kTail>;
    using Off0 = std::integral_constant<int, 0>;
    auto sweep = [&](auto cutTag) {
        for (int i0 = 0; i0 < nLocal;) {
            if (nLocal - i0 > kGroup - kTail) {
                const int tMin = kFwd ? tFirst + i0 : tLast - (i0 + kGroup - 1);
                const int tMax = kFwd ? tFirst + i0 + kGroup - 1 : tLast - i0;
                const bool full = tMin >= iaLo + tLo && tMax <= iaHi - 1 && i0 + kGroup - 1 + kAhead < nLocal;
                if (full) runGroup(i0, std::true_type{}, GroupSteps{}, Off0{}, cutTag);
                else runGroup(i0, std::false_type{}, GroupSteps{}, Off0{}, cutTag);
                i0 += kGroup;
            } else {
                if constexpr (kAhead == kTail) {
                    runGroup(i0, std::false_type{}, TailSteps{}, Off0{}, cutTag);
                } else {
                    static_assert(kAhead == 2 * kTail, "tail groups start at ring slot 0 or kTail");
                    if (i0 % kAhead == 0) runGroup(i0, std::false_type{}, TailSteps{}, Off0{}, cutTag);
                    else runGroup(i0, std::false_type{}, TailSteps{}, std::integral_constant<int, kTail>{}, cutTag);
                }
                i0 += kTail;
            }
        }
    };
    // (only lone sweeps are ever cut)
    if constexpr (LONE) {
        if (cut) sweep(std::true_type{});
        else sweep(std::false_type{});
    } else {
        sweep(std::false_type{});
    }
}

// LONE: every box tile of the launch is alone in its region (chosen by the
// launcher, so a launch with neighbouring tiles runs no lone-tile code).
// INP: a forward sweep computing BiCGStab's update as its input (SweepInput;
// the scalars are filled in from each wave's system).
template <typename T>
__device__ inline SweepInput<T> sweepInputOf(int inpKind, SweepInput<T> in, const SysState &S)
{
    if (inpKind == kInpP) {
        in.c1 = (T)S.beta;
        in.c2 = (T)(S.beta * S.omega);
    } else {
        in.c1 = (T)S.alpha;
    }
    return in;
}

template <typename T, int KT, int MODE, bool LONE, int INP = kInpPlain>
__global__ __launch_bounds__(64) void k_dilu_box(DiluView<T> V, const SysState *st, const T *__restrict__ in, T *out,
                                                 typename Bits<T>::U *inbox, typename Bits<T>::U *inboxClear,
                                                 int nClear, SweepInput<T> inpBase = SweepInput<T>{})
{
    __shared__ T xch[64];
    // a lone launch may cut every tile into segs segments: segs blocks per tile
    const int segs = LONE ? V.boxSegs : 1;
    const int j = (int)blockIdx.x / segs, seg = (int)blockIdx.x - j * segs;
    const int w = V.boxOrder[MODE == kDiluBwd ? V.nBoxWaves - 1 - j : j];
    const int32_t *m = V.boxMeta + (size_t)w * kBoxMeta;
    const bool active = MODE == kDiluFactor || st[m[kBmSys]].status == kRunning;
    const unsigned long long tStart = V.trace ? wall_clock64() : 0ull;
    unsigned long long nSlow = 0;
    if (active && m[kBmSteps] > 0) {
        const SweepInput<T> inp = INP ? sweepInputOf(INP, inpBase, st[m[kBmSys]]) : inpBase;
        switch (m[kBmAxis]) {
        case 0: boxSweep<T, KT, MODE, 0, LONE, INP>(V, m, in, out, inbox, xch, nSlow, seg, segs, inp); break;
        case 1: boxSweep<T, KT, MODE, 1, LONE, INP>(V, m, in, out, inbox, xch, nSlow, seg, segs, inp); break;
        default: boxSweep<T, KT, MODE, 2, LONE, INP>(V, m, in, out, inbox, xch, nSlow, seg, segs, inp); break;
        }
    }
    if (V.trace && threadIdx.x == 0 && segs == 1) {
        unsigned long long *tr = V.trace + 4 * ((size_t)V.nWaves + blockIdx.x);
        tr[0] = tStart;
        tr[1] = wall_clock64();
        tr[2] = nSlow;
        tr[3] = (unsigned long long)m[kBmSteps];
    }
    if (inboxClear)
        for (int q = blockIdx.x * 64 + threadIdx.x; q < nClear; q += gridDim.x * 64) inboxClear[q] = Bits<T>::kSent;
}

// Both sweeps of lone tiles in one launch: each wave sweeps its tile (or tile
// segment) forward and then backward, since its backward rows depend only on
// its own forward rows.  The same per-row operations as the two launches.
template <typename T, int KT, int INP = kInpPlain>
__global__ __launch_bounds__(64) void k_dilu_box_pair(DiluView<T> V, const SysState *st, const T *__restrict__ in,
                                                      T *out, SweepInput<T> inpBase = SweepInput<T>{})
{
    __shared__ T xch[64];
    const int segs = V.boxSegs;
    const int j = (int)blockIdx.x / segs, seg = (int)blockIdx.x - j * segs;
    const int32_t *m = V.boxMeta + (size_t)V.boxOrder[j] * kBoxMeta;
    if (st[m[kBmSys]].status != kRunning || m[kBmSteps] <= 0) return;
    unsigned long long nSlow = 0;
    using U = typename Bits<T>::U;
    const SweepInput<T> inp = INP ? sweepInputOf(INP, inpBase, st[m[kBmSys]]) : inpBase;
    switch (m[kBmAxis]) {
    case 0: boxSweep<T, KT, kDiluFwd, 0, true, INP>(V, m, in, out, (U *)nullptr, xch, nSlow, seg, segs, inp); break;
    case 1: boxSweep<T, KT, kDiluFwd, 1, true, INP>(V, m, in, out, (U *)nullptr, xch, nSlow, seg, segs, inp); break;
    default: boxSweep<T, KT, kDiluFwd, 2, true, INP>(V, m, in, out, (U *)nullptr, xch, nSlow, seg, segs, inp); break;
    }
    // the forward rows this wave stored are read back below (workgroup scope:
    // the same wave; an agent-scope fence would also write the L2 back)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    switch (m[kBmAxis]) {
    case 0: boxSweep<T, KT, kDiluBwd, 0, true>(V, m, out, out, (U *)nullptr, xch, nSlow, seg, segs); break;
    case 1: boxSweep<T, KT, kDiluBwd, 1, true>(V, m, out, out, (U *)nullptr, xch, nSlow, seg, segs); break;
    default: boxSweep<T, KT, kDiluBwd, 2, true>(V, m, out, out, (U *)nullptr, xch, nSlow, seg, segs); break;
    }
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
    if (V.nWaves <= 0) return hipSuccess;
    if (!V.waveOut) return hipErrorInvalidValue;
    const int mask = effectiveHaloMask(V.K, MODE == kDiluBwd ? V.haloMaskB : V.haloMaskF);
    switch (V.K) {
    case 3:
        if (mask == 0) k_dilu_wave<T, 3, MODE, 0><<<V.nWaves, 64, 0, s>>>(V, st, in, out, use, clr, nClr);
        else if (mask == 3) k_dilu_wave<T, 3, MODE, 3><<<V.nWaves, 64, 0, s>>>(V, st, in, out, use, clr, nClr);
        else k_dilu_wave<T, 3, MODE, 7><<<V.nWaves, 64, 0, s>>>(V, st, in, out, use, clr, nClr);
        break;
    case 6: k_dilu_wave<T, 6, MODE, 63><<<V.nWaves, 64, 0, s>>>(V, st, in, out, use, clr, nClr); break;
    case 12: k_dilu_wave<T, 12, MODE, 4095><<<V.nWaves, 64, 0, s>>>(V, st, in, out, use, clr, nClr); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, int MODE>
static hipError_t launchBox(const DiluView<T> &V, const SysState *st, const T *in, T *out,
                            typename Bits<T>::U *use, typename Bits<T>::U *clr, int nClr, hipStream_t s)
{
    if (V.nBoxWaves <= 0) return hipSuccess;
    const int n = V.nBoxWaves;
    if (V.boxLone) {
        const int nb = n * (V.boxSegs > 1 ? V.boxSegs : 1);
        if (V.K == 3) k_dilu_box<T, 3, MODE, true><<<nb, 64, 0, s>>>(V, st, in, out, use, clr, nClr);
        else k_dilu_box<T, 0, MODE, true><<<nb, 64, 0, s>>>(V, st, in, out, use, clr, nClr);
        return hipGetLastError();
    }
    if (V.boxSegs > 1) return hipErrorInvalidValue;   // segments only in lone sweeps
    if (V.K == 3) k_dilu_box<T, 3, MODE, false><<<n, 64, 0, s>>>(V, st, in, out, use, clr, nClr);
    else k_dilu_box<T, 0, MODE, false><<<n, 64, 0, s>>>(V, st, in, out, use, clr, nClr);
    return hipGetLastError();
}

// The general sweep on s and the box sweep on V.side when both exist and a
// side stream is set (the box sweep forked after the work already on s and
// joined before the next), else both on s.
template <typename T, typename GenF, typename BoxF>
static hipError_t bothSweeps(const DiluView<T> &V, hipStream_t s, GenF &&gen, BoxF &&box)
{
    const bool fork = V.side && V.fork && V.join && V.nWaves > 0 && V.nBoxWaves > 0;
    if (!fork) {
        hipError_t e = gen(s);
        return e != hipSuccess ? e : box(s);
    }
    hipError_t e = hipEventRecord(V.fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(V.side, V.fork, 0);
    if (e == hipSuccess) e = box(V.side);
    if (e == hipSuccess) e = hipEventRecord(V.join, V.side);
    if (e == hipSuccess) e = gen(s);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, V.join, 0);
    return e;
}

template <typename T>
hipError_t diluForward(const DiluView<T> &V, const SysState *st, const T *in, T *out, hipStream_t s)
{
    using U = typename Bits<T>::U;
    return bothSweeps(
        V, s, [&](hipStream_t q) { return launchK<T, kDiluFwd>(V, st, in, out, (U *)V.haloF, (U *)V.haloB, V.nHaloB, q); },
        [&](hipStream_t q) {
            return launchBox<T, kDiluFwd>(V, st, in, out, (U *)V.boxInF, (U *)V.boxInB, V.nBoxIn, q);
        });
}

template <typename T>
hipError_t diluBackward(const DiluView<T> &V, const SysState *st, T *inout, hipStream_t s)
{
    using U = typename Bits<T>::U;
    return bothSweeps(
        V, s,
        [&](hipStream_t q) {
            return launchK<T, kDiluBwd>(V, st, inout, inout, (U *)V.haloB, (U *)V.haloF, V.nHaloF, q);
        },
        [&](hipStream_t q) {
            return launchBox<T, kDiluBwd>(V, st, inout, inout, (U *)V.boxInB, (U *)V.boxInF, V.nBoxIn, q);
        });
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
    // the factor publishes d through its own inbox buffer, reset first
    const int n = V.nHaloF + V.nBoxIn;
    k_fill_u64<<<256, 256, 0, s>>>((unsigned long long *)V.haloC, n > 0 ? n : 1, kHaloSentinel64);
    hipError_t e = launchK<double, kDiluFactor>(V, nullptr, nullptr, nullptr, (unsigned long long *)V.haloC,
                                                nullptr, 0, s);
    if (e != hipSuccess) return e;
    return launchBox<double, kDiluFactor>(V, nullptr, nullptr, nullptr, (unsigned long long *)V.boxInC, nullptr, 0,
                                          s);
}

hipError_t diluFactorTileLocal(const DiluView<double> &V, double *rD, int segs, hipStream_t s)
{
    if (V.nBoxWaves <= 0) return hipSuccess;
    if (segs < 1) return hipErrorInvalidValue;
    DiluView<double> L = V;
    L.rDout = rD;
    L.fPack = nullptr;
    L.bPack = nullptr;
    L.boxLone = true;
    L.boxSegs = segs;
    L.trace = nullptr;
    return launchBox<double, kDiluFactor>(L, nullptr, nullptr, nullptr, (unsigned long long *)V.boxInC, nullptr, 0,
                                          s);
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

template <typename T>
hipError_t diluApplyLone(const DiluView<T> &V, const SysState *st, const T *in, T *out, hipStream_t s)
{
    if (V.nBoxWaves <= 0 || V.nWaves > 0 || !V.boxLone) return hipErrorInvalidValue;
    const int nb = V.nBoxWaves * (V.boxSegs > 1 ? V.boxSegs : 1);
    if (V.K == 3) k_dilu_box_pair<T, 3><<<nb, 64, 0, s>>>(V, st, in, out);
    else k_dilu_box_pair<T, 0><<<nb, 64, 0, s>>>(V, st, in, out);
    return hipGetLastError();
}

template <typename T>
hipError_t diluApplyFused(const DiluView<T> &V, const SysState *st, int inp, const T *r, T *dst, const T *v, T *out,
                          hipStream_t s)
{
    if (!diluCanFuse(V) || (inp != kInpP && inp != kInpS)) return hipErrorInvalidValue;
    SweepInput<T> in;
    in.r = r;
    in.dst = dst;
    in.v = v;
    if (V.boxLone) {
        const int nb = V.nBoxWaves * (V.boxSegs > 1 ? V.boxSegs : 1);
        if (inp == kInpP) k_dilu_box_pair<T, 3, kInpP><<<nb, 64, 0, s>>>(V, st, dst, out, in);
        else k_dilu_box_pair<T, 3, kInpS><<<nb, 64, 0, s>>>(V, st, dst, out, in);
        return hipGetLastError();
    }
    using U = typename Bits<T>::U;
    const int n = V.nBoxWaves;
    U *use = (U *)V.boxInF, *clr = (U *)V.boxInB;
    if (inp == kInpP) k_dilu_box<T, 3, kDiluFwd, false, kInpP><<<n, 64, 0, s>>>(V, st, dst, out, use, clr, V.nBoxIn, in);
    else k_dilu_box<T, 3, kDiluFwd, false, kInpS><<<n, 64, 0, s>>>(V, st, dst, out, use, clr, V.nBoxIn, in);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : diluBackward<T>(V, st, out, s);
}

template hipError_t diluApplyFused<double>(const DiluView<double> &, const SysState *, int, const double *, double *,
                                           const double *, double *, hipStream_t);
template hipError_t diluApplyFused<float>(const DiluView<float> &, const SysState *, int, const float *, float *,
                                          const float *, float *, hipStream_t);
template hipError_t diluApplyLone<double>(const DiluView<double> &, const SysState *, const double *, double *,
                                          hipStream_t);
template hipError_t diluApplyLone<float>(const DiluView<float> &, const SysState *, const float *, float *,
                                         hipStream_t);
template hipError_t diluForward<double>(const DiluView<double> &, const SysState *, const double *, double *,
                                        hipStream_t);
template hipError_t diluForward<float>(const DiluView<float> &, const SysState *, const float *, float *,
                                       hipStream_t);
template hipError_t diluBackward<double>(const DiluView<double> &, const SysState *, double *, hipStream_t);
template hipError_t diluBackward<float>(const DiluView<float> &, const SysState *, float *, hipStream_t);

}  // namespace mrfg
