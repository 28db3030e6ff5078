// DILU (coupled Cholesky preconditioner) executed by a wave schedule
// (dilu.hip).  See layout.hpp for how the schedule and dependency codes are
// built.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"
#include "layout.hpp"

namespace mrfg {

enum { kDiluFwd = 0, kDiluBwd = 1, kDiluFactor = 2 };

// Halo slots hold the raw bits of a value; a slot equal to the sentinel has
// not been published yet in the current sweep (signalling-NaN payloads that
// arithmetic never produces).
constexpr unsigned long long kHaloSentinel64 = 0xFFF4A5A5A5A5A5A5ull;
constexpr unsigned int kHaloSentinel32 = 0xFFA5A5A5u;

// Per-row DILU records, packed in 16-byte chunks so that a step of a wave
// costs few vector-memory instructions.  Chunk g of lane l of slice s sits
// at (s*G + g)*64 + l.
//   values  [c_0..c_{K-1}, rD]: forward = lower coefficients, backward =
//           upper coefficients, factor = upper*lower products with the
//           diagonal in place of rD;
//   ints    [codes, out_0..out_{K-1}]: one byte per dependency (kCodeNone,
//           kCodeInbox, or an LDS ring slot 0..255), four per int, then the
//           inbox addresses this row's value is delivered to (-1: none).
// A dependency on another wave is read from the consumer's inbox slot
// (slice, kk, lane), kk = rank of the dependency slot among the slots that
// can cross waves, so its prefetch needs no other load first.
template <typename T>
struct Pack;
template <>
struct Pack<double> {
    using V = double2;
    static constexpr int N = 2;
};
template <>
struct Pack<float> {
    using V = float4;
    static constexpr int N = 4;
};
constexpr int packChunks(int n, int per) { return (n + per - 1) / per; }
constexpr int kCodeNone = 0xFF, kCodeInbox = 0xFE;
// ints per row: packed codes then K outs
constexpr int codeInts(int K) { return packChunks(K, 4) + K; }
// Dependency slots that are given an inbox for a K-slot schedule whose
// cross-wave dependencies occur in `mask`: structured-box schedules (K = 3)
// only cross waves in the first two slots; anything else gets all slots.
inline int effectiveHaloMask(int K, int mask)
{
    if (K == 3 && mask == 0) return 0;   // no cross-wave dependency: no inbox at all
    return (K == 3 && (mask & ~3) == 0) ? 3 : (1 << K) - 1;
}

template <typename T>
struct DiluView {
    const int32_t *order = nullptr;      // general (code-driven) waves in forward dispatch order
    const int32_t *waveSlice = nullptr, *waveSteps = nullptr, *waveSys = nullptr;
    int nWaves = 0, K = 3;
    const int4 *fCodes = nullptr, *bCodes = nullptr;
    const int32_t *waveOut = nullptr;    // per wave: bit 0 / 1 when it delivers forward / backward values
    typename Pack<T>::V *fPack = nullptr, *bPack = nullptr;
    const double2 *prodPack = nullptr;   // factor records (FP64)
    double *rDout = nullptr;             // factor: plain rD per row
    const int32_t *cellOfPos = nullptr;  // factor: padding rows get rD = 0
    int haloMaskF = 0, haloMaskB = 0;    // dependency slots that may cross waves
    void *haloF = nullptr, *haloB = nullptr, *haloC = nullptr;   // inboxes
    int nHaloF = 0, nHaloB = 0;          // inbox sizes (values)
    int *err = nullptr;                  // bit 2: a halo wait expired
    // structured-box tile waves (dilu.hip, k_dilu_box): their own dispatch
    // order, kBoxMeta ints of tile geometry per wave, and compact inboxes of
    // (TB + TC) values per step per wave, one per direction
    const int32_t *boxOrder = nullptr;
    int nBoxWaves = 0;
    bool boxLone = false;                // every box tile is alone in its region
    // lone sweeps only (the tile-local preconditioner): each tile's rows are
    // cut along the sequential axis into boxSegs segments of nA / boxSegs
    // rows, each swept by a wave of its own with the coupling across the cut
    // left out
    int boxSegs = 1;
    const int32_t *boxMeta = nullptr;
    void *boxInF = nullptr, *boxInB = nullptr, *boxInC = nullptr;
    int nBoxIn = 0;
    // optional trace: per wave start / end realtime clock ticks, slow-path
    // hits and steps (4 entries each); general waves by dispatch index, then
    // box waves by dispatch index
    unsigned long long *trace = nullptr;
    // host side: when both kinds of waves exist, the box sweep runs on `side`
    // concurrently with the general sweep (regions never depend on each
    // other), forked and joined with the two events
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

// A forward sweep of box tiles may compute its input itself, BiCGStab's
// update of the vector it preconditions, and store that vector (the
// expressions of k_bicg_p / k_bicg_s, so the results are the same):
//   kInpP: in = r + c1*dst - c2*v (c1 = beta, c2 = beta*omega), stored to dst (p);
//   kInpS: in = r - c1*v (c1 = alpha), stored to dst (s).
enum { kInpPlain = 0, kInpP = 1, kInpS = 2 };
template <typename T>
struct SweepInput {
    const T *r = nullptr;
    T *dst = nullptr;
    const T *v = nullptr;
    T c1 = T(0), c2 = T(0);
};

// Forward sweep out = (D+L)^-1 in (with the D^-1 scaling of DILU) and the
// backward sweep in place; systems whose status is not running are skipped.
template <typename T>
hipError_t diluForward(const DiluView<T> &V, const SysState *st, const T *in, T *out, hipStream_t s);
template <typename T>
hipError_t diluBackward(const DiluView<T> &V, const SysState *st, T *inout, hipStream_t s);
// Both sweeps in one launch when every wave is a lone box tile (no general
// waves): out = M^-1 in, each wave forward then backward over its own rows.
template <typename T>
hipError_t diluApplyLone(const DiluView<T> &V, const SysState *st, const T *in, T *out, hipStream_t s);
// The preconditioner applied to BiCGStab's p (inp = kInpP: p = r + beta*p -
// beta*omega*v) or s (kInpS: s = r - alpha*v) of every system, with the
// update computed and stored by the forward sweeps (batches of K = 3 box
// tiles only, no general waves): out = M^-1 dst.  The scalars come from each
// system's state.
template <typename T>
hipError_t diluApplyFused(const DiluView<T> &V, const SysState *st, int inp, const T *r, T *dst, const T *v, T *out,
                          hipStream_t s);
template <typename T>
bool diluCanFuse(const DiluView<T> &V)
{
    return V.nWaves == 0 && V.nBoxWaves > 0 && V.K == 3;
}
// Factorisation: rD = 1/d for every row of the batch (FP64), written to
// rDout and into the forward and backward records.
hipError_t diluFactor(const DiluView<double> &V, hipStream_t s);
// The tile-local factor: rD of every box tile's rows with the couplings across
// tile edges (and across the cuts between its segs segments) left out, each
// tile segment factorised alone, written to rD only (rows of general waves
// are left as they are).  FP64, from the FP64 product records.
hipError_t diluFactorTileLocal(const DiluView<double> &V, double *rD, int segs, hipStream_t s);
// Segment s of segs of a box tile's nA rows along its sequential axis: rows
// [lo, hi) (shared by the host's record derivation and the device sweeps)
__host__ __device__ inline void boxSegment(int nA, int s, int segs, int &lo, int &hi)
{
    lo = (int)((long long)s * nA / segs);
    hi = (int)((long long)(s + 1) * nA / segs);
}
// rD of the FP32 records from the FP64 factor
hipError_t diluConvertRD(const DiluView<float> &V32, const double *rD64, int nSlices, hipStream_t s);
hipError_t fillHalo64(void *h, int n, hipStream_t s);
hipError_t fillHalo32(void *h, int n, hipStream_t s);

}  // namespace mrfg
