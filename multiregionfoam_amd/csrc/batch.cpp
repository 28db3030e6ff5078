// Device-resident batch of coupled systems and the Krylov solve drivers.
//
// Plays the role of coupledFvMatrix<scalar>::solve(dict) and the coupled
// solver selected from it (call site multiRegionSystem.C:150-153; solver
// dictionary flowOverHeatedPlate/system/fluid/fvSolution:43-57): BiCGStab or
// PCG with the Cholesky (DILU) preconditioner, lduSolver stop rules and
// normFactor, on every system of the batch at once.
#include "batch.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "dna.hpp"
#include "kernels.hpp"

namespace mrfg {

#define HIP_OK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
    } while (0)

// Buffers keep their capacity: a batch re-filled with systems of at most
// the same size allocates (and frees, which synchronises the device) nothing.
std::recursive_mutex &deviceMutex()
{
    static std::recursive_mutex m;
    return m;
}

template <typename T>
void DevBuf<T>::alloc(size_t count)
{
    if (p && count <= cap) {
        n = count;
        return;
    }
    std::lock_guard<std::recursive_mutex> lk(deviceMutex());
    release();
    n = count;
    cap = std::max<size_t>(count, 1);
    HIP_OK(hipMalloc((void **)&p, cap * sizeof(T)));
}

template <typename T>
void DevBuf<T>::upload(const std::vector<T> &h, hipStream_t st)
{
    alloc(h.size());
    if (!h.empty()) HIP_OK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

template <typename T>
void DevBuf<T>::zero(hipStream_t st)
{
    HIP_OK(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), st));
}

template struct DevBuf<double>;
template struct DevBuf<float>;
template struct DevBuf<int32_t>;
template struct DevBuf<unsigned long long>;
template struct DevBuf<int8_t>;
template struct DevBuf<FvRegionParams>;
template struct DevBuf<DnaSide>;
template struct DevBuf<DnaCell>;
template struct DevBuf<DnaPair>;

// ---------------------------------------------------------------------------
// fvSolution solver entry
// ---------------------------------------------------------------------------

SolverControls parseControls(const Dict &d)
{
    SolverControls c;
    std::string s = d.word("solver");
    if (s == "BiCGStab" || s == "hipBiCGStab" || s == "PBiCGStab") c.solver = SolverControls::BiCGStab;
    else if (s == "PCG" || s == "CG" || s == "hipPCG") c.solver = SolverControls::PCG;
    else throw DictError("Unknown coupled solver '" + s + "'; valid: BiCGStab hipBiCGStab PCG CG hipPCG");
    std::string pre = "Cholesky";
    if (d.isDict("preconditioner")) pre = d.subDict("preconditioner").word("preconditioner");
    else if (d.found("preconditioner")) pre = d.word("preconditioner");
    if (pre != "Cholesky" && pre != "DILU" && pre != "DIC")
        throw DictError("Unknown coupled preconditioner '" + pre + "'; valid: Cholesky DILU DIC");
    c.tolerance = d.numberOrDefault("tolerance", 1e-6);
    c.relTol = d.numberOrDefault("relTol", 0.0);
    c.minIter = (int)d.numberOrDefault("minIter", 0);
    c.maxIter = (int)d.numberOrDefault("maxIter", 1000);
    std::string prec = d.wordOrDefault("precision", "fp64");
    if (prec == "fp64" || prec == "double") c.precision = SolverControls::FP64;
    else if (prec == "mixed") c.precision = SolverControls::Mixed;
    else throw DictError("Unknown precision '" + prec + "'; valid: fp64 mixed");
    if (d.isDict("refinement")) {
        const Dict &r = d.subDict("refinement");
        c.innerRelTol = r.numberOrDefault("innerRelTol", c.innerRelTol);
        c.innerMaxIter = (int)r.numberOrDefault("innerMaxIter", c.innerMaxIter);
        c.maxOuter = (int)r.numberOrDefault("maxOuter", c.maxOuter);
        if (r.found("innerPreconditioner")) {
            const std::string ip = r.word("innerPreconditioner");
            if (ip == "DILU") c.innerPre = SolverControls::InnerDILU;
            else if (ip == "tileDILU") c.innerPre = SolverControls::InnerTileDILU;
            else throw DictError("Unknown innerPreconditioner '" + ip + "'; valid: DILU tileDILU");
        }
    }
    c.checkInterval = std::max(1, (int)d.numberOrDefault("checkInterval", c.checkInterval));
    if (d.found("autotune")) {
        const std::string g = d.word("autotune");
        if (g == "on" || g == "yes" || g == "true") c.autotune = true;
        else if (g == "off" || g == "no" || g == "false") c.autotune = false;
        else throw DictError("autotune must be on or off, not '" + g + "'");
    }
    if (d.found("graphs")) {
        const std::string g = d.word("graphs");
        if (g == "on" || g == "yes" || g == "true") c.graphs = true;
        else if (g == "off" || g == "no" || g == "false") c.graphs = false;
        else throw DictError("graphs must be on or off, not '" + g + "'");
    }
    return c;
}

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------

Batch::Batch(int device) : device_(device)
{
    std::lock_guard<std::recursive_mutex> lk(deviceMutex());
    HIP_OK(hipSetDevice(device_));
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&join_, hipEventDisableTiming));
}

Batch::~Batch()
{
    std::lock_guard<std::recursive_mutex> lk(deviceMutex());
    (void)hipSetDevice(device_);
    dropGraphs();
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
    if (side_) {
        (void)hipStreamSynchronize(side_);
        (void)hipStreamDestroy(side_);
    }
    if (fork_) (void)hipEventDestroy(fork_);
    if (join_) (void)hipEventDestroy(join_);
}

namespace {

template <typename T>
void appendShift(std::vector<T> &dst, const std::vector<T> &src, T shift)
{
    dst.reserve(dst.size() + src.size());
    for (T v : src) dst.push_back((T)(v + shift));
}

int roundK(int k)
{
    if (k <= 3) return 3;
    if (k <= 6) return 6;
    if (k <= 12) return 12;
    throw std::runtime_error("a cell has more than 12 lower or upper neighbours; not supported by the DILU kernels");
}

}  // namespace

void Batch::dropGraphs()
{
    std::lock_guard<std::recursive_mutex> lk(deviceMutex());
    for (auto &g : graphs_) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    graphs_.clear();
}

template <typename T>
void Batch::runChunk(bool pcg, int chunk, bool useGraph, const MatView<T> &A, const DiluView<T> &D,
                     const Krylov<T> &K)
{
    auto launch = [&]() -> hipError_t {
        for (int it = 0; it < chunk; ++it) {
            hipError_t e = pcg ? launchPcgIteration<T>(A, D, K, stream_) : launchBicgIteration<T>(A, D, K, stream_);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    if (!useGraph) {
        HIP_OK(launch());
        return;
    }
    const int kind = (pcg ? 1 : 0) + (sizeof(T) == 4 ? 2 : 0);
    IterGraph *g = nullptr;
    for (auto &x : graphs_)
        if (x.kind == kind && x.chunk == chunk) g = &x;
    if (!g) {
        std::lock_guard<std::recursive_mutex> lk(deviceMutex());
        IterGraph n;
        n.kind = kind;
        n.chunk = chunk;
        HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        hipError_t e = launch();
        hipError_t e2 = hipStreamEndCapture(stream_, &n.graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (n.graph) (void)hipGraphDestroy(n.graph);
            HIP_OK(e != hipSuccess ? e : e2);
        }
        e = hipGraphInstantiate(&n.exec, n.graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(n.graph);
            HIP_OK(e);
        }
        graphs_.push_back(n);
        g = &graphs_.back();
    }
    HIP_OK(hipGraphLaunch(g->exec, stream_));
}

void Batch::setSystems(const std::vector<const CoupledSystem *> &systems,
                       std::vector<std::unique_ptr<SystemLayout>> *layouts)
{
    HIP_OK(hipSetDevice(device_));
    HIP_OK(hipStreamSynchronize(stream_));
    dropGraphs();   // the captured launches name the buffers about to change
    sys_ = systems;
    lay_.clear();
    ifDev_.clear();
    // layouts of several systems are built concurrently (each depends only
    // on its own system)
    lay_.resize(systems.size());
    if (layouts) {
        if (layouts->size() != systems.size()) throw std::runtime_error("one layout per system is needed");
        for (size_t q = 0; q < systems.size(); ++q) {
            if (!(*layouts)[q]) throw std::runtime_error("missing layout");
            lay_[q] = std::move((*layouts)[q]);
        }
    } else {
        std::vector<std::thread> th;
        std::vector<std::string> errs(systems.size());
        for (size_t q = 0; q < systems.size(); ++q) {
            auto job = [&, q] {
                try {
                    lay_[q].reset(new SystemLayout(buildLayout(*systems[q], preferLevels_)));
                } catch (const std::exception &e) {
                    errs[q] = e.what();
                }
            };
            if (systems.size() == 1) job();
            else th.emplace_back(job);
        }
        for (auto &t : th) t.join();
        for (auto &e : errs)
            if (!e.empty()) throw std::runtime_error(e);
    }
    const int nS = (int)systems.size();
    sliceBegin_.assign(nS, 0);
    K_ = 3;
    for (auto &L : lay_) K_ = std::max(K_, roundK(L->K));

    std::vector<int32_t> ellOff, ellCol, ifBase, ifRowFace, ifPairOff, ifPairCol;
    std::vector<unsigned long long> ifMask;
    std::vector<int32_t> cellOfPos, fCode, bCode, fHalo, bHalo, waveSlice, waveSteps, waveSys, waveOut;
    std::vector<double> diag, source, sumA, psi0, fProd;
    std::vector<std::array<int, 3>> waveKey;   // (key, system, global wave index)
    std::vector<uint8_t> waveBox;
    std::vector<int32_t> boxMeta;
    int boxIn = 0;
    hEllVal_.clear(); hDiag_.clear(); hIfCoeff_.clear(); hIfPairW_.clear(); hFVal_.clear(); hBVal_.clear();
    std::vector<BlockEntry> blocks;
    hInfo_.assign(nS, SysInfo{});
    std::vector<double> ddtA, diagBase, srcBase;
    bool anyTransient = false;

    int slice = 0, ellBase = 0, ifRowBase = 0, ifFaceBase = 0, ifPairBase = 0, hfBase = 0, hbBase = 0;
    for (int q = 0; q < nS; ++q) {
        const SystemLayout &L = *lay_[q];
        const int pos0 = slice * kSlice;
        sliceBegin_[q] = slice;
        for (int s = 0; s < L.nSlices; ++s) {
            ellOff.push_back(L.ellOff[s] + ellBase);
            ifBase.push_back(L.ifBase[s] + ifRowBase);
            ifMask.push_back(L.ifMask[s]);
        }
        appendShift(ellCol, L.ellCol, pos0);
        appendShift(ifPairCol, L.ifPairCol, pos0);
        for (size_t i = 0; i + 1 < L.ifRowFace.size(); ++i) ifRowFace.push_back(L.ifRowFace[i] + ifFaceBase);
        for (size_t i = 0; i + 1 < L.ifPairOff.size(); ++i) ifPairOff.push_back(L.ifPairOff[i] + ifPairBase);
        cellOfPos.insert(cellOfPos.end(), L.cellOfPos.begin(), L.cellOfPos.end());
        diag.insert(diag.end(), L.diag.begin(), L.diag.end());
        source.insert(source.end(), L.source.begin(), L.source.end());
        sumA.insert(sumA.end(), L.sumA.begin(), L.sumA.end());
        psi0.insert(psi0.end(), L.psi0.begin(), L.psi0.end());
        // time stepping: steady regions have a = 0 and their own diag / source
        ddtA.resize(diag.size(), 0.0);
        diagBase.insert(diagBase.end(), L.diag.begin(), L.diag.end());
        srcBase.insert(srcBase.end(), L.source.begin(), L.source.end());
        for (int r = 0; r < (int)systems[q]->regions.size(); ++r) {
            const LduRegion &R = systems[q]->regions[r];
            if (!R.transient()) continue;
            anyTransient = true;
            for (int c = 0; c < R.nCells; ++c) {
                const size_t p = pos0 + L.posOfCell[L.regionOffset[r] + c];
                ddtA[p] = R.ddtCoeff[c];
                diagBase[p] = R.diagBase[c];
                srcBase[p] = R.sourceBase[c];
            }
        }
        hEllVal_.insert(hEllVal_.end(), L.ellVal.begin(), L.ellVal.end());
        hIfCoeff_.insert(hIfCoeff_.end(), L.ifCoeff.begin(), L.ifCoeff.end());
        hIfPairW_.insert(hIfPairW_.end(), L.ifPairW.begin(), L.ifPairW.end());
        // DILU dependency slots, re-padded to the batch width K_; halo codes
        // and halo slots are shifted to the system's halo range
        for (int s = 0; s < L.nSlices; ++s)
            for (int k = 0; k < K_; ++k)
                for (int l = 0; l < kSlice; ++l) {
                    if (k < L.K) {
                        size_t e = ((size_t)s * L.K + k) * kSlice + l;
                        int fc = L.fCode[e], bc = L.bCode[e];
                        fCode.push_back(fc >= 0 ? fc + hfBase : fc);
                        bCode.push_back(bc >= 0 ? bc + hbBase : bc);
                        hFVal_.push_back(L.fVal[e]);
                        hBVal_.push_back(L.bVal[e]);
                        fProd.push_back(L.fProd[e]);
                    } else {
                        fCode.push_back(kNoDep);
                        bCode.push_back(kNoDep);
                        hFVal_.push_back(0.0);
                        hBVal_.push_back(0.0);
                        fProd.push_back(0.0);
                    }
                }
        for (int p = 0; p < L.nPos; ++p) {
            fHalo.push_back(L.fHalo[p] >= 0 ? L.fHalo[p] + hfBase : -1);
            bHalo.push_back(L.bHalo[p] >= 0 ? L.bHalo[p] + hbBase : -1);
        }
        SysInfo &I = hInfo_[q];
        I.waveBegin = (int)waveSlice.size();
        for (int w = 0; w < L.nWaves; ++w) {
            waveKey.push_back({L.waveKey[w], q, (int)waveSlice.size()});
            waveSlice.push_back(L.waveSlice[w] + slice);
            waveSteps.push_back(L.waveSteps[w]);
            waveSys.push_back(q);
            waveBox.push_back(L.waveBox[w]);
            // box tile geometry with batch-wide wave, slice and inbox indices
            const int32_t *lm = &L.boxMeta[(size_t)w * kBoxMeta];
            std::array<int32_t, kBoxMeta> gm;
            for (int k = 0; k < kBoxMeta; ++k) gm[k] = lm[k];
            if (L.waveBox[w]) {
                for (int k : {kBmPrevB, kBmPrevC, kBmNextB, kBmNextC})
                    gm[k] = lm[k] >= 0 ? lm[k] + I.waveBegin : -1;
                gm[kBmSlice] = L.waveSlice[w] + slice;
                gm[kBmSteps] = L.waveSteps[w];
                gm[kBmInbox] = boxIn;
                gm[kBmSys] = q;
                boxIn += L.waveSteps[w] * (lm[kBmTB] + kSlice / lm[kBmTB]);
            }
            boxMeta.insert(boxMeta.end(), gm.begin(), gm.end());
        }
        I.nWaves = L.nWaves;
        I.blkBegin = (int)blocks.size();
        for (int s = 0; s < L.nSlices; s += kSlicesPerBlock)
            blocks.push_back(BlockEntry{q, slice + s, slice + std::min(L.nSlices, s + kSlicesPerBlock), 0});
        I.nBlk = (int)blocks.size() - I.blkBegin;
        I.sliceBegin = slice;
        I.nSlices = L.nSlices;
        I.nCells = (int)L.nCells;

        slice += L.nSlices;
        ellBase += L.ellOff[L.nSlices];
        ifRowBase += (int)L.ifRowFace.size() - 1;
        ifFaceBase += (int)L.ifCoeff.size();
        ifPairBase += (int)L.ifPairW.size();
        hfBase += L.nHaloF;
        hbBase += L.nHaloB;
    }
    ellOff.push_back(ellBase);
    ifRowFace.push_back(ifFaceBase);
    // pack the DILU records: values [c_0..c_{K-1}, rD | diag] in chunks of
    // Pack<T>::N; ints [byte codes, out_0..out_{K-1}] in chunks of 4.  A
    // cross-wave dependency is delivered to the consumer's inbox slot
    // (slice, kk, lane); its producer's outs list those inbox addresses.
    {
        const int nSl = slice;
        const int G = packChunks(K_ + 1, 2);
        const int NC = packChunks(K_, 4), H = packChunks(codeInts(K_), 4);
        int rawF = 0, rawB = 0;
        for (size_t e = 0; e < fCode.size(); ++e) {
            if (fCode[e] >= 0) rawF |= 1 << (int)((e / kSlice) % K_);
            if (bCode[e] >= 0) rawB |= 1 << (int)((e / kSlice) % K_);
        }
        haloMaskF_ = rawF;
        haloMaskB_ = rawB;
        const int mF = effectiveHaloMask(K_, rawF), mB = effectiveHaloMask(K_, rawB);
        const int KIF = __builtin_popcount(mF), KIB = __builtin_popcount(mB);
        auto kkOf = [](int m, int k) { return __builtin_popcount(m & ((1 << k) - 1)); };
        hFPack_.assign((size_t)nSl * G * 64 * 2, 0.0);
        hBPack_.assign((size_t)nSl * G * 64 * 2, 0.0);
        hFPack32_.clear();   // derived from the FP64 records when mixed precision first needs them
        hBPack32_.clear();
        std::vector<double> prod((size_t)nSl * G * 64 * 2, 0.0);
        std::vector<int32_t> fc((size_t)nSl * H * 64 * 4, -1), bc((size_t)nSl * H * 64 * 4, -1);
        // consumers of every published halo value (slot id -> inbox addresses)
        std::vector<std::vector<int32_t>> fOuts(hfBase), bOuts(hbBase);
        auto intAt = [&](int s, int j, int l) { return (((size_t)s * H + j / 4) * 64 + l) * 4 + j % 4; };
        for (int s = 0; s < nSl; ++s)
            for (int l = 0; l < kSlice; ++l) {
                const int p = s * kSlice + l;
                uint32_t fw[4] = {0, 0, 0, 0}, bw[4] = {0, 0, 0, 0};
                for (int k = 0; k < K_; ++k) {
                    size_t e = ((size_t)s * K_ + k) * kSlice + l;
                    auto byteOf = [&](int code, int mask, int KI, std::vector<std::vector<int32_t>> &outs) {
                        if (code == kNoDep) return kCodeNone;
                        if (code < 0) return -code - 1;            // LDS ring slot
                        if (!(mask & (1 << k))) throw std::runtime_error("DILU inbox mask misses a dependency slot");
                        outs[code].push_back(((s * KI + kkOf(mask, k)) << 6) + l);
                        return kCodeInbox;
                    };
                    uint32_t fb = (uint32_t)byteOf(fCode[e], mF, KIF, fOuts);
                    uint32_t bb = (uint32_t)byteOf(bCode[e], mB, KIB, bOuts);
                    fw[k / 4] |= fb << (8 * (k % 4));
                    bw[k / 4] |= bb << (8 * (k % 4));
                    size_t e64 = (((size_t)s * G + k / 2) * 64 + l) * 2 + k % 2;
                    hFPack_[e64] = hFVal_[e];
                    hBPack_[e64] = hBVal_[e];
                    prod[e64] = fProd[e];
                }
                // padding slots of the last code int are "none"
                for (int k = K_; k < NC * 4; ++k) {
                    fw[k / 4] |= (uint32_t)kCodeNone << (8 * (k % 4));
                    bw[k / 4] |= (uint32_t)kCodeNone << (8 * (k % 4));
                }
                for (int j = 0; j < NC; ++j) {
                    fc[intAt(s, j, l)] = (int32_t)fw[j];
                    bc[intAt(s, j, l)] = (int32_t)bw[j];
                }
                prod[(((size_t)s * G + K_ / 2) * 64 + l) * 2 + K_ % 2] = diag[p];
            }
        for (int s = 0; s < nSl; ++s)
            for (int l = 0; l < kSlice; ++l) {
                const int p = s * kSlice + l;
                auto putOuts = [&](int h, std::vector<std::vector<int32_t>> &outs, std::vector<int32_t> &dst) {
                    if (h < 0) return;
                    if ((int)outs[h].size() > K_) throw std::runtime_error("DILU value has more consumers than slots");
                    for (size_t j = 0; j < outs[h].size(); ++j) dst[intAt(s, NC + (int)j, l)] = outs[h][j];
                };
                putOuts(fHalo[p], fOuts, fc);
                putOuts(bHalo[p], bOuts, bc);
            }
        // per wave: bit 0 (1) when a row of it delivers a forward (backward)
        // value to another wave's inbox; a wave without skips the deliveries
        waveOut.assign(waveSlice.size(), 0);
        for (size_t w = 0; w < waveSlice.size(); ++w)
            for (int s = waveSlice[w]; s < waveSlice[w] + waveSteps[w]; ++s)
                for (int l = 0; l < kSlice; ++l)
                    for (int j = 0; j < K_; ++j) {
                        if (fc[intAt(s, NC + j, l)] >= 0) waveOut[w] |= 1;
                        if (bc[intAt(s, NC + j, l)] >= 0) waveOut[w] |= 2;
                    }
        nHaloF_ = nSl * KIF * 64;
        nHaloB_ = nSl * KIB * 64;
        auto asVec = [](const std::vector<int32_t> &a) {
            std::vector<int4> v(a.size() / 4);
            std::memcpy(v.data(), a.data(), a.size() * sizeof(int32_t));
            return v;
        };
        auto asD2 = [](const std::vector<double> &a) {
            std::vector<double2> v(a.size() / 2);
            std::memcpy(v.data(), a.data(), a.size() * sizeof(double));
            return v;
        };
        fCodes_.upload(asVec(fc), stream_);
        bCodes_.upload(asVec(bc), stream_);
        prodPack_.upload(asD2(prod), stream_);
        m64_.fPack.upload(asD2(hFPack_), stream_);
        m64_.bPack.upload(asD2(hBPack_), stream_);
    }
    ifPairOff.push_back(ifPairBase);
    nSlices_ = slice;
    nPos_ = slice * kSlice;
    nBlk_ = (int)blocks.size();
    nWaves_ = (int)waveSlice.size();

    // forward dispatch order: by key (tile anti-diagonal / level), systems
    // interleaved; every wave's producers precede it.  Box tile waves and
    // general waves never depend on each other (DILU dependencies stay inside
    // a region) and run in kernels of their own.
    std::sort(waveKey.begin(), waveKey.end());
    std::vector<int32_t> order, boxOrder;
    for (int i = 0; i < nWaves_; ++i) (waveBox[waveKey[i][2]] ? boxOrder : order).push_back(waveKey[i][2]);
    nGenWaves_ = (int)order.size();
    nBoxWaves_ = (int)boxOrder.size();
    nBoxIn_ = boxIn;
    hBoxMeta_ = boxMeta;
    hBoxWaves_.assign(boxOrder.begin(), boxOrder.end());
    boxLone_ = nBoxWaves_ > 0;
    for (int w : boxOrder)
        for (int k : {kBmPrevB, kBmPrevC, kBmNextB, kBmNextC})
            if (boxMeta[(size_t)w * kBoxMeta + k] >= 0) boxLone_ = false;
    if (nGenWaves_ == 0) nHaloF_ = nHaloB_ = 0;
    if (order.empty()) order.push_back(0);
    if (boxOrder.empty()) boxOrder.push_back(0);

    hipStream_t st = stream_;
    ellOff_.upload(ellOff, st);
    ellCol_.upload(ellCol, st);
    ifMask_.upload(ifMask, st);
    ifBase_.upload(ifBase, st);
    ifRowFace_.upload(ifRowFace, st);
    ifPairOff_.upload(ifPairOff, st);
    ifPairCol_.upload(ifPairCol, st);
    cellOfPos_.upload(cellOfPos, st);
    {
        std::vector<unsigned long long> rm(std::max(1, slice), 0ull);
        for (size_t p = 0; p < cellOfPos.size(); ++p)
            if (cellOfPos[p] >= 0) rm[p / kSlice] |= 1ull << (p % kSlice);
        rowMask_.upload(rm, st);
    }
    blocks_.upload(blocks, st);
    info_.upload(hInfo_, st);
    for (int k = 0; k < 2; ++k) {
        const int S = 8 << k;
        std::vector<BlockEntry> bl;
        std::vector<SysInfo> inf = hInfo_;
        for (int q = 0; q < nS; ++q) {
            inf[q].blkBegin = (int)bl.size();
            for (int s0 = 0; s0 < hInfo_[q].nSlices; s0 += S)
                bl.push_back(BlockEntry{q, hInfo_[q].sliceBegin + s0,
                                        hInfo_[q].sliceBegin + std::min(hInfo_[q].nSlices, s0 + S), 0});
            inf[q].nBlk = (int)bl.size() - inf[q].blkBegin;
        }
        bcfg_[k].nBlk = (int)bl.size();
        bcfg_[k].blocks.upload(bl.empty() ? std::vector<BlockEntry>{BlockEntry{0, 0, 0, 0}} : bl, st);
        bcfg_[k].info.upload(inf, st);
    }
    tune_ = Tuning{};
    order_.upload(order, st);
    boxOrder_.upload(boxOrder, st);
    waveOut_.upload(waveOut, st);
    boxMeta_.upload(boxMeta, st);
    waveSlice_.upload(waveSlice, st);
    waveSteps_.upload(waveSteps, st);
    waveSys_.upload(waveSys, st);
    // each inbox buffer: general part, then the box part
    haloF64_.alloc(std::max(1, nHaloF_ + nBoxIn_));
    haloB64_.alloc(std::max(1, nHaloB_ + nBoxIn_));
    haloC64_.alloc(std::max(1, nHaloF_ + nBoxIn_));
    HIP_OK(fillHalo64(haloF64_.p, (int)haloF64_.n, st));
    HIP_OK(fillHalo64(haloB64_.p, (int)haloB64_.n, st));
    HIP_OK(fillHalo64(haloC64_.p, (int)haloC64_.n, st));
    err_.alloc(1);
    err_.zero(st);

    m64_.ellVal.upload(hEllVal_, st);
    m64_.diag.upload(diag, st);
    hDiag_ = diag;
    m64_.ifCoeff.upload(hIfCoeff_, st);
    m64_.ifPairW.upload(hIfPairW_, st);
    sumA_.upload(sumA, st);
    b64_.upload(source, st);
    transient_ = anyTransient;
    stepsDone_ = 0;
    diagCoefft_ = 1.0;
    if (transient_) {
        ddtA_.upload(ddtA, st);
        diagBase_.upload(diagBase, st);
        srcBase_.upload(srcBase, st);
        xPrev_.alloc(diag.size());
        xPrev_.zero(st);
    }
    x64_.upload(psi0, st);
    x0_.upload(psi0, st);
    for (auto &v : v64_) {
        v.alloc(nPos_);
        v.zero(st);
    }
    st64_.alloc(nS);
    st32_.alloc(nS);
    partials_.alloc((size_t)std::max(1, nBlk_) * 8);   // 4 double-double sums per block
    factorised_ = false;
    haveFloat_ = false;   // FP32 copies are rebuilt (in the buffers kept) on demand
    hostStale_ = false;
    buildFv();
    rf_.reset();
    // the interface face lists (flux check, patch values)
    ifDev_.clear();
    for (int q = 0; q < nS; ++q)
        for (int i = 0; i < (int)systems[q]->interfaces.size(); ++i)
            for (int sd = 0; sd < 2; ++sd) (void)ifaceSlot(q, i, sd);
    HIP_OK(hipStreamSynchronize(st));
}

void Batch::ensureFloat(bool tileLocal)
{
    tileLocal = tileLocal && nBoxWaves_ > 0;
    if (haveFloat_ && tileLocal == tileLocal32_) return;
    tileLocal32_ = tileLocal;
    if (hostStale_) refreshHostCopies();
    else derivePack32(tileLocal);
    dropGraphs();   // FP32 buffers may be allocated anew
    auto toF = [](const std::vector<double> &a) {
        std::vector<float> b(a.size());
        for (size_t i = 0; i < a.size(); ++i) b[i] = (float)a[i];
        return b;
    };
    hipStream_t st = stream_;
    m32_.ellVal.upload(toF(hEllVal_), st);
    m32_.diag.upload(toF(hDiag_), st);
    m32_.ifCoeff.upload(toF(hIfCoeff_), st);
    m32_.ifPairW.upload(toF(hIfPairW_), st);
    {
        std::vector<float4> f(hFPack32_.size() / 4), b(hBPack32_.size() / 4);
        std::memcpy(f.data(), hFPack32_.data(), hFPack32_.size() * sizeof(float));
        std::memcpy(b.data(), hBPack32_.data(), hBPack32_.size() * sizeof(float));
        m32_.fPack.upload(f, st);
        m32_.bPack.upload(b, st);
    }
    for (auto &v : v32_) {
        v.alloc(nPos_);
        v.zero(st);
    }
    haloF32_.alloc(std::max(1, nHaloF_ + nBoxIn_));
    haloB32_.alloc(std::max(1, nHaloB_ + nBoxIn_));
    HIP_OK(fillHalo32(haloF32_.p, (int)haloF32_.n, st));
    HIP_OK(fillHalo32(haloB32_.p, (int)haloB32_.n, st));
    haveFloat_ = true;
    factorised_ = false;   // rD32 is derived from the FP64 factor
    HIP_OK(hipStreamSynchronize(st));
}

template <typename T>
MatView<T> Batch::view(const DevMatrix<T> &m) const
{
    MatView<T> A;
    A.ellOff = ellOff_.p;
    A.ellCol = ellCol_.p;
    A.ellVal = m.ellVal.p;
    A.diag = m.diag.p;
    A.ifMask = ifMask_.p;
    A.rowMask = rowMask_.p;
    A.ifBase = ifBase_.p;
    A.ifRowFace = ifRowFace_.p;
    A.ifCoeff = m.ifCoeff.p;
    A.ifPairOff = ifPairOff_.p;
    A.ifPairCol = ifPairCol_.p;
    A.ifPairW = m.ifPairW.p;
    A.ellChunk = tune_.ellChunk;
    return A;
}

template <typename T>
DiluView<T> Batch::diluView(DevMatrix<T> &m, void *hF, void *hB)
{
    DiluView<T> D;
    D.order = order_.p;
    D.waveSlice = waveSlice_.p;
    D.waveSteps = waveSteps_.p;
    D.waveSys = waveSys_.p;
    D.nWaves = nGenWaves_;
    D.K = K_;
    D.boxOrder = boxOrder_.p;
    D.nBoxWaves = nBoxWaves_;
    // the FP32 records of the tile-local DILU are swept tile by tile
    const bool local = (const void *)&m == (const void *)&m32_ && tileLocal32_;
    D.boxLone = boxLone_ || local;
    D.boxSegs = local ? tileSegments() : 1;
    D.boxMeta = boxMeta_.p;
    D.nBoxIn = nBoxIn_;
    D.fCodes = fCodes_.p;
    D.waveOut = waveOut_.p;
    D.bCodes = bCodes_.p;
    D.fPack = m.fPack.p;
    D.bPack = m.bPack.p;
    D.prodPack = prodPack_.p;
    D.rDout = v64_[11].p;
    D.cellOfPos = cellOfPos_.p;
    D.haloMaskF = haloMaskF_;
    D.haloMaskB = haloMaskB_;
    D.haloF = hF;
    D.haloB = hB;
    D.haloC = haloC64_.p;
    D.nHaloF = nHaloF_;
    D.nHaloB = nHaloB_;
    D.boxInF = (char *)hF + (size_t)nHaloF_ * sizeof(T);
    D.boxInB = (char *)hB + (size_t)nHaloB_ * sizeof(T);
    D.boxInC = (char *)haloC64_.p + (size_t)nHaloF_ * sizeof(double);
    D.err = err_.p;
    D.side = side_;
    D.fork = fork_;
    D.join = join_;
    return D;
}

void Batch::launchCfg(int slices, const BlockEntry *&blocks, int &nBlk, const SysInfo *&info, int &threads) const
{
    const int k = slices == 16 ? 1 : 0;
    blocks = slices == 4 ? blocks_.p : bcfg_[k].blocks.p;
    nBlk = slices == 4 ? nBlk_ : bcfg_[k].nBlk;
    info = slices == 4 ? info_.p : bcfg_[k].info.p;
    threads = 64 * slices;
}

void Batch::setTuning(int slices, int ellChunk)
{
    if (tune_.slices == slices && tune_.ellChunk == ellChunk) return;
    tune_.slices = slices;
    tune_.ellChunk = ellChunk;
    dropGraphs();   // the captured launches use the previous configuration
}

void Batch::autotune()
{
    HIP_OK(hipSetDevice(device_));
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    const int reps = 5;
    double best = 1e300, def = 0.0;
    int bs = 4, bc = 8;
    for (int slices : {4, 8, 16})
        for (int ch : {2, 4, 8, 16}) {
            MatView<double> A = view(m64_);
            A.ellChunk = ch;
            const BlockEntry *bl;
            const SysInfo *inf;
            int nb, th;
            launchCfg(slices, bl, nb, inf, th);
            HIP_OK(launchSpmvPlain<double>(A, bl, nb, th, x64_.p, v64_[0].p, stream_));   // warm-up
            HIP_OK(hipEventRecord(e0, stream_));
            for (int r = 0; r < reps; ++r) HIP_OK(launchSpmvPlain<double>(A, bl, nb, th, x64_.p, v64_[0].p, stream_));
            HIP_OK(hipEventRecord(e1, stream_));
            HIP_OK(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            const double t = (double)ms / reps;
            if (slices == 4 && ch == 8) def = t;
            if (t < best) {
                best = t;
                bs = slices;
                bc = ch;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    setTuning(bs, bc);
    tune_.tuned = true;
    tune_.bestMs = best;
    tune_.defaultMs = def;
}

template <>
void Batch::bindKrylov<double>(Krylov<double> &K)
{
    launchCfg(tune_.slices, K.blocks, K.nBlk, K.info, K.threads);
    K.nSys = (int)sys_.size();
    K.state = st64_.p;
    K.partials = partials_.p;
    K.sumA = sumA_.p;
    K.x = x64_.p;
    K.b = b64_.p;
    K.ax = v64_[0].p; K.r = v64_[1].p; K.rw = v64_[2].p; K.pv = v64_[3].p; K.v = v64_[4].p;
    K.sv = v64_[5].p; K.t = v64_[6].p; K.ph = v64_[7].p; K.sh = v64_[8].p; K.w = v64_[9].p;
    K.pA = v64_[10].p; K.rD = v64_[11].p;
}

template <>
void Batch::bindKrylov<float>(Krylov<float> &K)
{
    launchCfg(tune_.slices, K.blocks, K.nBlk, K.info, K.threads);
    K.nSys = (int)sys_.size();
    K.state = st32_.p;
    K.partials = partials_.p;
    K.sumA = sumA_.p;
    K.x = v32_[0].p; K.b = v32_[1].p;
    K.ax = v32_[2].p; K.r = v32_[3].p; K.rw = v32_[4].p; K.pv = v32_[5].p; K.v = v32_[6].p;
    K.sv = v32_[7].p; K.t = v32_[8].p; K.ph = v32_[9].p; K.sh = v32_[10].p; K.w = v32_[11].p;
    K.pA = v32_[12].p; K.rD = v32_[13].p;
}

void Batch::factorise()
{
    if (factorised_) return;
    DiluView<double> D = diluView(m64_, haloF64_.p, haloB64_.p);
    HIP_OK(mrfg::diluFactor(D, stream_));
    if (haveFloat_) {
        DiluView<float> D32 = diluView(m32_, haloF32_.p, haloB32_.p);
        const double *rd = v64_[11].p;
        if (tileLocal32_) {
            // the general waves' rows keep the full factor's rD
            rdLocal_.alloc(nPos_);
            HIP_OK(hipMemcpyAsync(rdLocal_.p, v64_[11].p, (size_t)nPos_ * sizeof(double), hipMemcpyDeviceToDevice,
                                  stream_));
            HIP_OK(diluFactorTileLocal(D, rdLocal_.p, tileSegments(), stream_));
            rd = rdLocal_.p;
        }
        HIP_OK(diluConvertRD(D32, rd, nSlices_, stream_));
    }
    factorised_ = true;
}

void Batch::checkDeviceError()
{
    int e = 0;
    HIP_OK(hipMemcpyAsync(&e, err_.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    if (e) {
        err_.zero(stream_);
        throw std::runtime_error("HIP error: DILU dependency wait expired (device error flag " + std::to_string(e) +
                                 ")");
    }
}

void Batch::initState(const SolverControls &c, const std::vector<char> *active)
{
    std::vector<SysState> h(sys_.size());
    for (size_t q = 0; q < sys_.size(); ++q) {
        SysState s = {};
        s.tol = c.tolerance;
        s.relTol = c.relTol;
        s.minIter = c.minIter;
        s.maxIter = c.precision == SolverControls::Mixed ? c.maxOuter : c.maxIter;
        s.status = (!active || (*active)[q]) ? kRunning : kIdle;
        s.nCells = hInfo_[q].nCells;
        h[q] = s;
    }
    HIP_OK(hipMemcpyAsync(st64_.p, h.data(), h.size() * sizeof(SysState), hipMemcpyHostToDevice, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
}

void Batch::markAllRunning()
{
    std::vector<SysState> hs(sys_.size());
    for (auto &x : hs) x.status = kRunning;
    HIP_OK(hipMemcpyAsync(st64_.p, hs.data(), hs.size() * sizeof(SysState), hipMemcpyHostToDevice, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
}

void Batch::readState(std::vector<SysState> &out)
{
    out.resize(sys_.size());
    HIP_OK(hipMemcpyAsync(out.data(), st64_.p, out.size() * sizeof(SysState), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
}

static bool anyRunning(const std::vector<SysState> &s)
{
    for (auto &x : s)
        if (x.status == kRunning) return true;
    return false;
}

SolvePerf Batch::perfFp64(int q, const SysState &s) const
{
    SolvePerf p;
    p.initialResidual = s.initRes;
    p.finalResidual = s.finalRes;
    p.nIterations = s.nIter;
    p.converged = s.status == kConverged;
    p.singular = s.status == kSingular ? 1 : s.status == kDiverged ? 2 : 0;
    p.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - solveStart_).count();
    p.bytes = p.nIterations * bytesIter(q, 8);
    return p;
}

std::vector<int> Batch::newlyFinished(const std::vector<SysState> &hs)
{
    std::vector<int> out;
    if (!retire_) return out;
    for (size_t q = 0; q < hs.size(); ++q)
        if (!retired_[q] && hs[q].status != kRunning && hs[q].status != kIdle) {
            retired_[q] = 1;
            out.push_back((int)q);
        }
    return out;
}

void Batch::runFp64(const SolverControls &c)
{
    Krylov<double> K;
    bindKrylov(K);
    DiluView<double> D = diluView(m64_, haloF64_.p, haloB64_.p);
    MatView<double> A = view(m64_);
    const bool pcg = c.solver == SolverControls::PCG;
    HIP_OK(launchSetup<double>(A, K, pcg, stream_));
    std::vector<SysState> hs;
    readState(hs);
    // systems that stopped at the last check are handed to the retire
    // callback after the next chunk is launched, so their results are
    // copied out while the others iterate
    std::vector<int> fin = newlyFinished(hs);
    std::vector<SysState> finState = hs;
    auto retireAll = [&]() {
        for (int q : fin) (*retire_)(q, perfFp64(q, finState[q]));
        fin.clear();
    };
    int done = 0;
    while (anyRunning(hs) && done <= c.maxIter) {
        int chunk = std::min(c.checkInterval, c.maxIter - done + 1);
        runChunk<double>(pcg, chunk, c.graphs, A, D, K);
        retireAll();
        done += chunk;
        readState(hs);
        fin = newlyFinished(hs);
        finState = hs;
    }
    retireAll();
}

void Batch::copyToHost(void *dst, const void *src, size_t bytes) const
{
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
}

void Batch::copyToDevice(void *dst, const void *src, size_t bytes) const
{
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
}

void Batch::copyPsiNow(int sys, int region, double *psi)
{
    const SystemLayout &L = *lay_.at(sys);
    const CoupledSystem &S = *sys_.at(sys);
    const int n = S.regions.at(region).nCells, o = L.regionOffset[region];
    std::vector<double> x(L.nPos);
    copyToHost(x.data(), x64_.p + posBegin(sys), L.nPos * sizeof(double));
    for (int c = 0; c < n; ++c) psi[c] = x[L.posOfCell[o + c]];
}

void Batch::runMixed(const SolverControls &c, std::vector<int> &innerIters, std::vector<int> &fp64Iters)
{
    Krylov<double> K64;
    bindKrylov(K64);
    Krylov<float> K32;
    bindKrylov(K32);
    DiluView<float> D32 = diluView(m32_, haloF32_.p, haloB32_.p);
    MatView<double> A64 = view(m64_);
    MatView<float> A32 = view(m32_);
    const bool pcg = c.solver == SolverControls::PCG;
    const int nS = (int)sys_.size();
    innerIters.assign(nS, 0);
    fp64Iters.assign(nS, 0);
    gate_.alloc(std::max(1, nS));
    std::vector<int8_t> gate(nS, 0);
    auto putGate = [&]() {
        HIP_OK(hipMemcpyAsync(gate_.p, gate.data(), nS * sizeof(int8_t), hipMemcpyHostToDevice, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    };
    auto putOuter = [&](const std::vector<SysState> &h) {
        HIP_OK(hipMemcpyAsync(st64_.p, h.data(), nS * sizeof(SysState), hipMemcpyHostToDevice, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    };
    HIP_OK(launchSetup<double>(A64, K64, false, stream_));
    std::vector<SysState> hs, hi(nS);
    readState(hs);
    // a system whose outer residual stops falling (the FP32 correction no
    // longer resolves its error) continues in FP64 from its current field
    std::vector<double> prevRes(nS);
    std::vector<char> fallback(nS, 0);
    for (int q = 0; q < nS; ++q) prevRes[q] = hs[q].finalRes;
    // residual history of each running inner solve (stagnation test)
    std::vector<std::vector<double>> hist(nS);
    // a fresh inner solve for the gated systems: FP32 state, right-hand side
    // the FP64 residual, zero correction
    auto startInner = [&]() {
        putGate();
        HIP_OK(launchRefinePrepare(nS, st64_.p, st32_.p, c.innerRelTol, c.innerMaxIter, gate_.p, stream_));
        HIP_OK(launchRefineRhs(K64, K32, gate_.p, stream_));
        HIP_OK(launchSetup<float>(A32, K32, pcg, stream_, gate_.p));
        // the history starts with the inner solve's initial residual
        HIP_OK(hipMemcpyAsync(hi.data(), st32_.p, nS * sizeof(SysState), hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        for (int q = 0; q < nS; ++q)
            if (gate[q]) hist[q].assign(1, hi[q].finalRes);
    };
    // Each system takes its outer step as soon as its own inner solve has
    // ended (at the check that finds it), and restarts its inner solve,
    // while the others keep iterating: no system waits for the slowest
    // inner solve of the batch, and the results are the same as alone.
    for (int q = 0; q < nS; ++q) gate[q] = hs[q].status == kRunning;
    startInner();
    // before the first chunk the histories already hold the current residuals
    bool firstCheck = true;
    long long done = 0;
    const long long budget = (long long)(c.maxOuter + 1) * (c.innerMaxIter + c.checkInterval);
    for (;;) {
        HIP_OK(hipMemcpyAsync(hi.data(), st32_.p, nS * sizeof(SysState), hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        // an inner solve whose best residual over the last kStallWindow
        // checks is not 10 % below its best before them has stagnated (the
        // FP32 system resolves no further): it stops there, and the outer
        // step takes the correction it has
        bool stopped = false;
        for (int q = 0; q < nS; ++q) {
            if (hi[q].status != kRunning) continue;
            std::vector<double> &h = hist[q];
            if (firstCheck) continue;
            h.push_back(hi[q].finalRes);
            if ((int)h.size() <= kStallWindow) continue;
            const double recent = *std::min_element(h.end() - kStallWindow, h.end());
            const double prior = *std::min_element(h.begin(), h.end() - kStallWindow);
            if (!(recent < 0.9 * prior)) {
                hi[q].status = kMaxIter;
                stopped = true;
            }
        }
        firstCheck = false;
        if (stopped) {
            HIP_OK(hipMemcpyAsync(st32_.p, hi.data(), nS * sizeof(SysState), hipMemcpyHostToDevice, stream_));
            HIP_OK(hipStreamSynchronize(stream_));
        }
        // the systems whose inner solve has ended; one that broke down or
        // left the FP32 range (singular, diverged) gives no correction and
        // continues in FP64 from its current field
        bool due = false, outerChanged = false;
        for (int q = 0; q < nS; ++q) {
            gate[q] = 0;
            if (hs[q].status != kRunning || hi[q].status == kRunning || hi[q].status == kIdle) continue;
            innerIters[q] += hi[q].nIter;
            if (hi[q].status == kSingular || hi[q].status == kDiverged) {
                fallback[q] = 1;
                hs[q].status = kIdle;
                outerChanged = true;
            } else {
                gate[q] = 1;
                due = true;
            }
        }
        if (outerChanged) putOuter(hs);
        if (due) {
            putGate();
            HIP_OK(launchRefineUpdate(K64, K32, gate_.p, stream_));
            HIP_OK(launchRefineResidual(A64, K64, gate_.p, stream_));
            readState(hs);
            bool changed = false;
            for (int q = 0; q < nS; ++q) {
                if (!gate[q] || hs[q].status != kRunning) continue;
                if (!(hs[q].finalRes < kStallFactor * prevRes[q])) {
                    fallback[q] = 1;
                    hs[q].status = kIdle;
                    changed = true;
                }
                prevRes[q] = hs[q].finalRes;
            }
            if (changed) putOuter(hs);
            // systems whose refinement has ended (not those continuing in FP64)
            for (int q : newlyFinished(hs)) {
                SolvePerf p = perfFp64(q, hs[q]);
                p.nIterations = innerIters[q];
                p.nOuter = hs[q].nIter;
                p.bytes = p.nIterations * bytesIter(q, 4) + p.nOuter * bytesSpmv(q, 8);
                (*retire_)(q, p);
            }
            bool restart = false;
            for (int q = 0; q < nS; ++q) {
                gate[q] = gate[q] && hs[q].status == kRunning;
                restart = restart || gate[q];
            }
            if (restart) startInner();   // the next check follows a chunk
        }
        bool running = false;
        for (int q = 0; q < nS; ++q) running = running || hs[q].status == kRunning;
        if (!running || done > budget) break;
        runChunk<float>(pcg, c.checkInterval, c.graphs, A32, D32, K32);
        done += c.checkInterval;
    }
    for (int q = 0; q < nS; ++q)
        if (hs[q].status == kRunning) fallback[q] = 1;   // outer steps exhausted
    bool any = false;
    for (int q = 0; q < nS; ++q) any |= fallback[q] != 0;
    if (!any) return;
    // FP64 continuation of the fallback systems only
    std::vector<SysState> mixedState = hs, h(nS);
    for (int q = 0; q < nS; ++q) {
        SysState s = {};
        s.tol = c.tolerance;
        s.relTol = c.relTol;
        s.minIter = 0;
        s.maxIter = c.maxIter;
        s.nCells = hInfo_[q].nCells;
        s.status = fallback[q] ? kRunning : kIdle;
        h[q] = s;
    }
    HIP_OK(hipMemcpyAsync(st64_.p, h.data(), nS * sizeof(SysState), hipMemcpyHostToDevice, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    SolverControls c64 = c;
    c64.precision = SolverControls::FP64;
    const RetireFn *keep = retire_;
    retire_ = nullptr;   // the continuation's records are composed below
    runFp64(c64);
    retire_ = keep;
    readState(h);
    for (int q = 0; q < nS; ++q) {
        if (fallback[q]) {
            fp64Iters[q] = h[q].nIter;
            SysState m = h[q];
            m.initRes = mixedState[q].initRes;
            m.nIter = mixedState[q].nIter;
            mixedState[q] = m;
        }
    }
    HIP_OK(hipMemcpyAsync(st64_.p, mixedState.data(), nS * sizeof(SysState), hipMemcpyHostToDevice, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
}

void Batch::solve(const SolverControls &c, std::vector<SolvePerf> &perf, const std::vector<char> *active,
                  const RetireFn *retire)
{
    if (active && active->size() != sys_.size()) throw std::runtime_error("active mask does not match the batch");
    HIP_OK(hipSetDevice(device_));
    struct ClearRetire {
        const RetireFn *&r;
        ~ClearRetire() { r = nullptr; }
    } clearRetire{retire_};
    retire_ = retire;
    retired_.assign(sys_.size(), 0);
    if (active)
        for (size_t q = 0; q < sys_.size(); ++q) retired_[q] = (*active)[q] ? 0 : 1;
    applyRefresh();
    if (c.precision == SolverControls::Mixed) ensureFloat(c.innerPre == SolverControls::InnerTileDILU);
    if (c.autotune && !tune_.tuned) autotune();
    if (!c.autotune) setTuning(4, 8);
    HIP_OK(hipStreamSynchronize(stream_));
    auto t0 = std::chrono::steady_clock::now();
    solveStart_ = t0;
    // the preconditioner is built per solve call, as the coupled solver
    // object is constructed per coupledFvMatrix::solve call
    factorised_ = false;
    factorise();
    initState(c, active);
    std::vector<int> inner, fp64More;
    if (c.precision == SolverControls::Mixed) runMixed(c, inner, fp64More);
    else runFp64(c);
    HIP_OK(hipStreamSynchronize(stream_));
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    checkDeviceError();
    std::vector<SysState> hs;
    readState(hs);
    perf.assign(sys_.size(), SolvePerf{});
    for (size_t q = 0; q < sys_.size(); ++q) {
        SolvePerf &p = perf[q];
        p.initialResidual = hs[q].initRes;
        p.finalResidual = hs[q].finalRes;
        p.converged = hs[q].status == kConverged;
        p.singular = hs[q].status == kSingular ? 1 : hs[q].status == kDiverged ? 2 : 0;
        p.seconds = secs;
        if (c.precision == SolverControls::Mixed) {
            p.nIterations = inner[q] + fp64More[q];
            p.nOuter = hs[q].nIter;
            p.bytes = p.nIterations * bytesIter((int)q, 4) + p.nOuter * bytesSpmv((int)q, 8);
        } else {
            p.nIterations = hs[q].nIter;
            p.bytes = p.nIterations * bytesIter((int)q, 8);
        }
    }
    if (retire_)
        for (size_t q = 0; q < sys_.size(); ++q)
            if (!retired_[q]) {
                retired_[q] = 1;
                (*retire_)((int)q, perf[q]);
            }
}

void Batch::resetPsi()
{
    HIP_OK(hipMemcpyAsync(x64_.p, x0_.p, (size_t)nPos_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    stepsDone_ = 0;
}

void Batch::setDdtDiag(double coefft)
{
    if (coefft == diagCoefft_) return;
    HIP_OK(launchDdtDiag(view(m64_), blocks_.p, nBlk_, K_, coefft, diagBase_.p, ddtA_.p, m64_.diag.p,
                         prodPack_.p, sumA_.p, stream_));
    // the host copy the FP32 matrix is derived from
    HIP_OK(hipMemcpyAsync(hDiag_.data(), m64_.diag.p, hDiag_.size() * sizeof(double), hipMemcpyDeviceToHost,
                          stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    haveFloat_ = false;
    diagCoefft_ = coefft;
}

void Batch::timeSteps(const SolverControls &c, bool backward, int nSteps, std::vector<SolvePerf> &perf)
{
    HIP_OK(hipSetDevice(device_));
    if (!transient_) throw std::runtime_error("time stepping needs a transient system (ddt coefficients)");
    if (c.precision != SolverControls::FP64) throw std::runtime_error("time stepping runs in precision fp64");
    if (nSteps < 1) throw std::runtime_error("time stepping needs nSteps >= 1");
    perf.clear();
    for (int n = 0; n < nSteps; ++n) {
        beginTimeStep(backward);
        std::vector<SolvePerf> p;
        solve(c, p);
        perf.insert(perf.end(), p.begin(), p.end());
    }
}

void Batch::beginTimeStep(bool backward)
{
    HIP_OK(hipSetDevice(device_));
    if (!transient_) throw std::runtime_error("time stepping needs a transient system (ddt coefficients)");
    const bool bwd = backward && stepsDone_ >= 1;
    setDdtDiag(bwd ? 1.5 : 1.0);
    HIP_OK(launchDdtSource(blocks_.p, nBlk_, bwd, srcBase_.p, ddtA_.p, x64_.p, xPrev_.p, b64_.p, stream_));
    HIP_OK(hipMemcpyAsync(xPrev_.p, x64_.p, (size_t)nPos_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    ++stepsDone_;
}

void Batch::setPsi(int sys, int region, const double *psi)
{
    const SystemLayout &L = *lay_.at(sys);
    const CoupledSystem &S = *sys_.at(sys);
    int n = S.regions.at(region).nCells, o = L.regionOffset[region];
    std::vector<double> x(L.nPos);
    HIP_OK(hipStreamSynchronize(stream_));
    copyToHost(x.data(), x64_.p + posBegin(sys), L.nPos * sizeof(double));
    for (int c = 0; c < n; ++c) x[L.posOfCell[o + c]] = psi[c];
    copyToDevice(x64_.p + posBegin(sys), x.data(), L.nPos * sizeof(double));
}

void Batch::getPsi(int sys, int region, double *psi)
{
    const SystemLayout &L = *lay_.at(sys);
    const CoupledSystem &S = *sys_.at(sys);
    int n = S.regions.at(region).nCells, o = L.regionOffset[region];
    std::vector<double> x(L.nPos);
    HIP_OK(hipStreamSynchronize(stream_));
    copyToHost(x.data(), x64_.p + posBegin(sys), L.nPos * sizeof(double));
    for (int c = 0; c < n; ++c) psi[c] = x[L.posOfCell[o + c]];
}

void Batch::spmv(const double *xHost, double *yHost, int sys)
{
    applyRefresh();
    const SystemLayout &L = *lay_.at(sys);
    std::vector<double> x(nPos_, 0.0), y(nPos_);
    for (long long c = 0; c < L.nCells; ++c) x[posBegin(sys) + L.posOfCell[c]] = xHost[c];
    DevBuf<double> dx, dy;
    dx.upload(x, stream_);
    dy.alloc(nPos_);
    HIP_OK(launchSpmvPlain<double>(view(m64_), blocks_.p, nBlk_, 256, dx.p, dy.p, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    copyToHost(y.data(), dy.p, nPos_ * sizeof(double));
    for (long long c = 0; c < L.nCells; ++c) yHost[c] = y[posBegin(sys) + L.posOfCell[c]];
}

void Batch::diluFactor(int sys, double *rDHost)
{
    applyRefresh();
    factorised_ = false;
    factorise();
    const SystemLayout &L = *lay_.at(sys);
    std::vector<double> r(nPos_);
    HIP_OK(hipStreamSynchronize(stream_));
    checkDeviceError();
    copyToHost(r.data(), v64_[11].p, nPos_ * sizeof(double));
    for (long long c = 0; c < L.nCells; ++c) rDHost[c] = r[posBegin(sys) + L.posOfCell[c]];
}

void Batch::precondition(const double *rHost, double *wHost, int sys)
{
    applyRefresh();
    factorise();
    const SystemLayout &L = *lay_.at(sys);
    std::vector<double> r(nPos_, 0.0), w(nPos_);
    for (long long c = 0; c < L.nCells; ++c) r[posBegin(sys) + L.posOfCell[c]] = rHost[c];
    markAllRunning();
    DevBuf<double> dr, dw;
    dr.upload(r, stream_);
    dw.alloc(nPos_);
    dw.zero(stream_);
    Krylov<double> K;
    bindKrylov(K);
    DiluView<double> D = diluView(m64_, haloF64_.p, haloB64_.p);
    HIP_OK(launchPrecondition<double>(D, K, dr.p, dw.p, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    checkDeviceError();
    copyToHost(w.data(), dw.p, nPos_ * sizeof(double));
    for (long long c = 0; c < L.nCells; ++c) wHost[c] = w[posBegin(sys) + L.posOfCell[c]];
}

Batch::IfaceDev &Batch::ifaceSlot(int sys, int iface, int side)
{
    const CoupledSystem &S = *sys_.at(sys);
    const SystemLayout &L = *lay_.at(sys);
    const CoupledInterface &I = S.interfaces.at(iface);
    if (ifDev_.size() != sys_.size()) ifDev_.resize(sys_.size());
    if (ifDev_[sys].size() != S.interfaces.size()) ifDev_[sys].resize(S.interfaces.size());
    auto &slot = ifDev_[sys][iface][side];
    if (!slot) {
        const InterfaceSide &O = side ? I.b : I.a, &Nb = side ? I.a : I.b;
        slot.reset(new IfaceDev);
        const int p0 = posBegin(sys);
        const int oO = L.regionOffset[O.region], oN = L.regionOffset[Nb.region];
        std::vector<int32_t> fcPos(O.nFaces()), pairPos(O.idx.size());
        for (int f = 0; f < O.nFaces(); ++f) fcPos[f] = p0 + L.posOfCell[oO + O.fc[f]];
        for (size_t k = 0; k < O.idx.size(); ++k) pairPos[k] = p0 + L.posOfCell[oN + Nb.fc[O.idx[k]]];
        slot->nFaces = O.nFaces();
        slot->fcPos.upload(fcPos.empty() ? std::vector<int32_t>{0} : fcPos, stream_);
        slot->pairPos.upload(pairPos.empty() ? std::vector<int32_t>{0} : pairPos, stream_);
        slot->off.upload(O.off, stream_);
        slot->coeff.upload(O.coeff.empty() ? std::vector<double>{0.0} : O.coeff, stream_);
        slot->w.upload(O.w.empty() ? std::vector<double>{0.0} : O.w, stream_);
        slot->wall = !O.kOwn.empty() || O.nFaces() == 0;
        slot->kOwn.upload(O.kOwn.empty() ? std::vector<double>{0.0} : O.kOwn, stream_);
        slot->kNei.upload(O.kNei.empty() ? std::vector<double>{0.0} : O.kNei, stream_);
        HIP_OK(hipStreamSynchronize(stream_));
    }
    return *slot;
}

void Batch::ifaceSync(int sys, int iface, int side, IfaceDev &slot)
{
    if (slot.nFaces == 0) return;
    const CoupledSystem &S = *sys_.at(sys);
    const bool fvNew = hasModel(sys) && fv_->assembled;
    const bool rfNew = !fvNew && rf_ && rf_->changed[sys];
    if (!fvNew && !rfNew) return;
    size_t off = fvNew ? fv_->sideBase[sys] : rf_->sideBase[sys];
    for (int i = 0; i < iface; ++i) off += S.interfaces[i].a.nFaces() + S.interfaces[i].b.nFaces();
    if (side) off += S.interfaces[iface].a.nFaces();
    const size_t n = slot.nFaces * sizeof(double);
    HIP_OK(hipMemcpyAsync(slot.coeff.p, (fvNew ? fv_->ifCoeff.p : rf_->sideCoeff.p) + off, n,
                          hipMemcpyDeviceToDevice, stream_));
    if (fvNew) {   // the wall conductances follow k
        HIP_OK(hipMemcpyAsync(slot.kOwn.p, fv_->sKOwn.p + off, n, hipMemcpyDeviceToDevice, stream_));
        HIP_OK(hipMemcpyAsync(slot.kNei.p, fv_->sKNei.p + off, n, hipMemcpyDeviceToDevice, stream_));
    }
}

void Batch::interfaceFlux(int sys, int iface, double *qa, double *qb)
{
    HIP_OK(hipSetDevice(device_));
    (void)ifaceSlot(sys, iface, 0);
    (void)ifaceSlot(sys, iface, 1);
    // reduction partials of up to one per 256 positions of the batch, then
    // the two results
    fluxBuf_.alloc(partials_.n + 2);
    interfaceFluxScratch(sys, iface, qa, qb, fluxBuf_.p, fluxBuf_.n);
}

void Batch::interfaceFluxScratch(int sys, int iface, double *qa, double *qb, double *scratch, size_t nScratch)
{
    HIP_OK(hipSetDevice(device_));
    applyRefresh();
    const CoupledSystem &S = *sys_.at(sys);
    if (iface < 0 || iface >= (int)S.interfaces.size()) throw std::out_of_range("interface index");
    if (ifDev_.size() != sys_.size() || ifDev_[sys].size() != S.interfaces.size() || !ifDev_[sys][iface][0] ||
        !ifDev_[sys][iface][1])
        throw std::runtime_error("interface face lists are not resident");
    // the two results sit after the partials
    if (nScratch < 2) throw std::runtime_error("flux scratch too small");
    double *out = scratch + nScratch - 2;
    for (int q = 0; q < 2; ++q) {
        IfaceDev &slot = *ifDev_[sys][iface][q];
        ifaceSync(sys, iface, q, slot);
        const size_t nBlk = ((size_t)slot.nFaces + 255) / 256;
        if (nBlk * 8 > nScratch - 2)
            throw std::runtime_error("interface side of " + std::to_string(slot.nFaces) +
                                     " faces exceeds the reduction scratch (at most " +
                                     std::to_string((nScratch - 2) / 8 * 256) + " faces)");
        HIP_OK(launchInterfaceFlux(slot.nFaces, slot.fcPos.p, slot.coeff.p, slot.off.p, slot.pairPos.p, slot.w.p,
                                   x64_.p, scratch, out + q, stream_));
    }
    double h[2] = {0.0, 0.0};
    HIP_OK(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    *qa = h[0];
    *qb = h[1];
}

void Batch::interfacePatchValues(int sys, int iface, int side, double *tn, double *tw)
{
    HIP_OK(hipSetDevice(device_));
    applyRefresh();
    if (side < 0 || side > 1) throw std::out_of_range("interface side");
    IfaceDev &slot = ifaceSlot(sys, iface, side);
    if (!slot.wall)
        throw std::runtime_error("interface " + std::to_string(iface) + " has no wall conductances "
                                 "(mrfg_set_interface_wall)");
    ifaceSync(sys, iface, side, slot);
    if (slot.nFaces == 0) return;
    DevBuf<double> out;
    out.alloc(2 * (size_t)slot.nFaces);
    HIP_OK(launchInterfacePatch(slot.nFaces, slot.fcPos.p, slot.off.p, slot.pairPos.p, slot.w.p, slot.kOwn.p,
                                slot.kNei.p, x64_.p, out.p, out.p + slot.nFaces, stream_));
    HIP_OK(hipMemcpyAsync(tn, out.p, slot.nFaces * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipMemcpyAsync(tw, out.p + slot.nFaces, slot.nFaces * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
}

void Batch::setInterfaceWall(int sys, int iface, int side, const double *kOwn, const double *kNei)
{
    HIP_OK(hipSetDevice(device_));
    if (side < 0 || side > 1) throw std::out_of_range("interface side");
    IfaceDev &slot = ifaceSlot(sys, iface, side);
    if (!kOwn || !kNei) throw std::invalid_argument("null wall conductances");
    if (slot.nFaces == 0) return;
    std::vector<double> a(kOwn, kOwn + slot.nFaces), b(kNei, kNei + slot.nFaces);
    slot.kOwn.upload(a, stream_);
    slot.kNei.upload(b, stream_);
    HIP_OK(hipStreamSynchronize(stream_));
    slot.wall = true;
}

void Batch::tilesOf(int sys, int *tileOfCell) const
{
    const SystemLayout &L = *lay_.at(sys);
    const int w0 = hInfo_.at(sys).waveBegin;
    std::vector<int> waveOfSlice(L.nSlices, -1);
    for (int w = 0; w < L.nWaves; ++w)
        for (int t = 0; t < L.waveSteps[w]; ++t) waveOfSlice[L.waveSlice[w] + t] = w;
    const int segs = tileSegments();
    for (long long c = 0; c < L.nCells; ++c) {
        const int p = L.posOfCell[c], w = waveOfSlice[p / kSlice];
        if (!L.waveBox[w]) {
            tileOfCell[c] = -1;
            continue;
        }
        // the segment of the cell's row along the tile's sequential axis
        const int32_t *m = &L.boxMeta[(size_t)w * kBoxMeta];
        const int TB = m[kBmTB], lane = p % kSlice, t = p / kSlice - L.waveSlice[w];
        const int ia = t - lane % TB - lane / TB;
        int seg = 0, lo, hi;
        for (; seg < segs - 1; ++seg) {
            boxSegment(m[kBmNA], seg, segs, lo, hi);
            if (ia < hi) break;
        }
        tileOfCell[c] = (w0 + w) * segs + seg;
    }
}

void Batch::precondition32(const float *rHost, float *wHost, int sys, bool tileLocal)
{
    applyRefresh();
    ensureFloat(tileLocal);
    factorise();
    const SystemLayout &L = *lay_.at(sys);
    std::vector<float> r(nPos_, 0.0f), w(nPos_);
    for (long long c = 0; c < L.nCells; ++c) r[posBegin(sys) + L.posOfCell[c]] = rHost[c];
    markAllRunning();
    DevBuf<float> dr, dw;
    dr.upload(r, stream_);
    dw.alloc(nPos_);
    dw.zero(stream_);
    Krylov<float> K;
    bindKrylov(K);
    K.state = st64_.p;   // running flags of the systems (set above)
    DiluView<float> D = diluView(m32_, haloF32_.p, haloB32_.p);
    HIP_OK(launchPrecondition<float>(D, K, dr.p, dw.p, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    checkDeviceError();
    copyToHost(w.data(), dw.p, nPos_ * sizeof(float));
    for (long long c = 0; c < L.nCells; ++c) wHost[c] = w[posBegin(sys) + L.posOfCell[c]];
}

double Batch::timeKernel(int which, int reps)
{
    HIP_OK(hipSetDevice(device_));
    applyRefresh();
    if (which == 7 || which == 8) return timeAssembly(which, reps);
    factorise();
    markAllRunning();
    Krylov<double> K;
    bindKrylov(K);
    DiluView<double> D = diluView(m64_, haloF64_.p, haloB64_.p);
    MatView<double> A = view(m64_);
    const int nE = 2 * reps + 2;
    std::vector<hipEvent_t> ev(nE);
    for (auto &e : ev) HIP_OK(hipEventCreate(&e));
    // sweeps run as forward/backward pairs (each resets the other's halo);
    // only the requested one sits between its events
    auto one = [&](int i) {
        hipEvent_t a = ev[2 * i], b = ev[2 * i + 1];
        if (which == 0) {
            HIP_OK(hipEventRecord(a, stream_));
            HIP_OK(launchSpmvPlain<double>(A, K.blocks, K.nBlk, K.threads, x64_.p, v64_[0].p, stream_));
            HIP_OK(hipEventRecord(b, stream_));
        } else if (which == 3) {
            HIP_OK(hipEventRecord(a, stream_));
            HIP_OK(mrfg::diluFactor(D, stream_));
            HIP_OK(hipEventRecord(b, stream_));
        } else {
            if (which == 1) HIP_OK(hipEventRecord(a, stream_));
            HIP_OK(diluForward<double>(D, K.state, b64_.p, v64_[9].p, stream_));
            if (which == 1) HIP_OK(hipEventRecord(b, stream_));
            if (which == 2) HIP_OK(hipEventRecord(a, stream_));
            HIP_OK(diluBackward<double>(D, K.state, v64_[9].p, stream_));
            if (which == 2) HIP_OK(hipEventRecord(b, stream_));
        }
    };
    if (which >= 4) {
        // inside BiCGStab iterations of the solve (tolerance 0, from the
        // initial guess): the kernel sees the caches and clocks of a solve
        SolverControls c;
        c.tolerance = 0.0;
        c.relTol = 0.0;
        c.maxIter = reps + 2;
        initState(c);
        resetPsi();
        HIP_OK(launchSetup<double>(A, K, false, stream_));
        for (int i = 0; i <= reps; ++i) {
            const IterProbe probe{which - 3, ev[2 * (i == 0 ? reps : i - 1)], ev[2 * (i == 0 ? reps : i - 1) + 1]};
            HIP_OK(launchBicgIteration<double>(A, D, K, stream_, &probe));
        }
        resetPsi();
    } else {
        one(reps);   // warm-up into the spare event pair
        for (int i = 0; i < reps; ++i) one(i);
    }
    HIP_OK(hipStreamSynchronize(stream_));
    checkDeviceError();
    double tot = 0;
    for (int i = 0; i < reps; ++i) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
        tot += ms;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return tot / reps;
}

int Batch::traceDilu(long long *out, int maxWaves)
{
    HIP_OK(hipSetDevice(device_));
    applyRefresh();
    factorise();
    markAllRunning();
    Krylov<double> K;
    bindKrylov(K);
    DiluView<double> D = diluView(m64_, haloF64_.p, haloB64_.p);
    DevBuf<unsigned long long> tr;
    tr.alloc((size_t)8 * nWaves_);
    tr.zero(stream_);
    int freq = 0;
    HIP_OK(hipDeviceGetAttribute(&freq, hipDeviceAttributeWallClockRate, device_));   // kHz
    D.trace = tr.p;
    HIP_OK(diluForward<double>(D, K.state, b64_.p, v64_[9].p, stream_));
    D.trace = tr.p + 4 * (size_t)nWaves_;
    HIP_OK(diluBackward<double>(D, K.state, v64_[9].p, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    checkDeviceError();
    std::vector<unsigned long long> h((size_t)8 * nWaves_);
    copyToHost(h.data(), tr.p, h.size() * sizeof(unsigned long long));
    const double nsPerTick = freq > 0 ? 1e6 / freq : 10.0;
    for (int sw = 0; sw < 2; ++sw) {
        unsigned long long t0 = ~0ull;
        for (int w = 0; w < nWaves_; ++w) t0 = std::min(t0, h[(size_t)4 * (sw * nWaves_ + w)]);
        for (int w = 0; w < nWaves_ && w < maxWaves; ++w) {
            const unsigned long long *e = &h[(size_t)4 * (sw * nWaves_ + w)];
            long long *o = out + (size_t)4 * (sw * std::min(nWaves_, maxWaves) + w);
            o[0] = (long long)((e[0] - t0) * nsPerTick);
            o[1] = (long long)((e[1] - t0) * nsPerTick);
            o[2] = (long long)e[2];
            o[3] = (long long)e[3];
        }
    }
    return nWaves_;
}

double Batch::bytesSpmv(int sys, int sv) const
{
    const SystemLayout &L = *lay_.at(sys);
    double N = (double)L.nCells, F = (double)L.nFaces, P = (double)L.nPairs;
    return N * (3 * sv + 4) + (F + P) * (2 * sv + 8) + 4;
}

double Batch::bytesDilu(int sys, int sv) const
{
    const SystemLayout &L = *lay_.at(sys);
    double N = (double)L.nCells, F = (double)L.nFaces;
    return N * (6 * sv + 8) + F * (2 * sv + 8) + 8;
}

double Batch::bytesIter(int sys, int sv) const
{
    const SystemLayout &L = *lay_.at(sys);
    return 2 * bytesSpmv(sys, sv) + 2 * bytesDilu(sys, sv) + 21.0 * (double)L.nCells * sv;
}

}  // namespace mrfg
