// Device-resident batch of coupled systems and the Krylov solve drivers.
//
// Plays the role of coupledFvMatrix<scalar>::solve(dict) and the coupled
// solver selected from it (call site multiRegionSystem.C:150-153; solver
// dictionary flowOverHeatedPlate/system/fluid/fvSolution:43-57): BiCGStab or
// PCG with the Cholesky (DILU) preconditioner, lduSolver stop rules and
// normFactor, on every system of the batch at once.
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <stdexcept>

#include "kernels.hpp"

namespace mrfg {

#define HIP_OK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
    } while (0)

template <typename T>
void DevBuf<T>::alloc(size_t count)
{
    release();
    n = count;
    HIP_OK(hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T)));
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
    }
    c.checkInterval = std::max(1, (int)d.numberOrDefault("checkInterval", c.checkInterval));
    return c;
}

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------

Batch::Batch(int device) : device_(device)
{
    HIP_OK(hipSetDevice(device_));
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

Batch::~Batch()
{
    (void)hipSetDevice(device_);
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

namespace {

template <typename T>
void appendShift(std::vector<T> &dst, const std::vector<T> &src, T shift, bool keepNeg = false)
{
    dst.reserve(dst.size() + src.size());
    for (T v : src) dst.push_back(keepNeg && v < 0 ? v : (T)(v + shift));
}

}  // namespace

void Batch::setSystems(const std::vector<const CoupledSystem *> &systems)
{
    HIP_OK(hipSetDevice(device_));
    sys_ = systems;
    lay_.clear();
    for (auto *s : systems) lay_.emplace_back(new SystemLayout(buildLayout(*s)));
    const int nS = (int)systems.size();
    sliceBegin_.assign(nS, 0);
    levelBase_.assign(nS, 0);

    std::vector<int32_t> ellOff, ellCol, ifBase, ifRowFace, ifPairOff, ifPairCol, lOff, lCol, uOff, uCol;
    std::vector<unsigned long long> ifMask;
    std::vector<int32_t> cellOfPos, sliceSys, sliceLevel, levelCount;
    std::vector<double> diag, source, sumA, psi0;
    hEllVal_.clear(); hDiag_.clear(); hIfCoeff_.clear(); hIfPairW_.clear();
    hLVal_.clear(); hLProd_.clear(); hUVal_.clear();
    std::vector<BlockEntry> blocks;
    hInfo_.assign(nS, SysInfo{});

    int slice = 0, ellBase = 0, lBase = 0, uBase = 0, ifRowBase = 0, ifFaceBase = 0, ifPairBase = 0, lvlBase = 0;
    int maxLevels = 0;
    for (int q = 0; q < nS; ++q) {
        const SystemLayout &L = *lay_[q];
        const int pos0 = slice * kSlice;
        sliceBegin_[q] = slice;
        levelBase_[q] = lvlBase;
        for (int s = 0; s < L.nSlices; ++s) {
            ellOff.push_back(L.ellOff[s] + ellBase);
            lOff.push_back(L.lOff[s] + lBase);
            uOff.push_back(L.uOff[s] + uBase);
            ifBase.push_back(L.ifBase[s] + ifRowBase);
            ifMask.push_back(L.ifMask[s]);
            sliceSys.push_back(q);
        }
        for (int lv = 0; lv < L.nLevels; ++lv) {
            levelCount.push_back(L.levelSliceBegin[lv + 1] - L.levelSliceBegin[lv]);
            for (int s = L.levelSliceBegin[lv]; s < L.levelSliceBegin[lv + 1]; ++s) sliceLevel.push_back(lv);
        }
        appendShift(ellCol, L.ellCol, pos0);
        appendShift(lCol, L.lCol, pos0, true);
        appendShift(uCol, L.uCol, pos0, true);
        appendShift(ifPairCol, L.ifPairCol, pos0);
        for (size_t i = 0; i + 1 < L.ifRowFace.size(); ++i) ifRowFace.push_back(L.ifRowFace[i] + ifFaceBase);
        for (size_t i = 0; i + 1 < L.ifPairOff.size(); ++i) ifPairOff.push_back(L.ifPairOff[i] + ifPairBase);
        cellOfPos.insert(cellOfPos.end(), L.cellOfPos.begin(), L.cellOfPos.end());
        diag.insert(diag.end(), L.diag.begin(), L.diag.end());
        source.insert(source.end(), L.source.begin(), L.source.end());
        sumA.insert(sumA.end(), L.sumA.begin(), L.sumA.end());
        psi0.insert(psi0.end(), L.psi0.begin(), L.psi0.end());
        hEllVal_.insert(hEllVal_.end(), L.ellVal.begin(), L.ellVal.end());
        hIfCoeff_.insert(hIfCoeff_.end(), L.ifCoeff.begin(), L.ifCoeff.end());
        hIfPairW_.insert(hIfPairW_.end(), L.ifPairW.begin(), L.ifPairW.end());
        hLVal_.insert(hLVal_.end(), L.lVal.begin(), L.lVal.end());
        hLProd_.insert(hLProd_.end(), L.lProd.begin(), L.lProd.end());
        hUVal_.insert(hUVal_.end(), L.uVal.begin(), L.uVal.end());

        SysInfo &I = hInfo_[q];
        I.blkBegin = (int)blocks.size();
        for (int s = 0; s < L.nSlices; s += kSlicesPerBlock)
            blocks.push_back(BlockEntry{q, slice + s, slice + std::min(L.nSlices, s + kSlicesPerBlock), 0});
        I.nBlk = (int)blocks.size() - I.blkBegin;
        I.sliceBegin = slice;
        I.nSlices = L.nSlices;
        I.levelBase = lvlBase;
        I.nLevels = L.nLevels;
        I.nCells = (int)L.nCells;

        slice += L.nSlices;
        ellBase += L.ellOff[L.nSlices];
        lBase += L.lOff[L.nSlices];
        uBase += L.uOff[L.nSlices];
        ifRowBase += (int)L.ifRowFace.size() - 1;
        ifFaceBase += (int)L.ifCoeff.size();
        ifPairBase += (int)L.ifPairW.size();
        lvlBase += L.nLevels;
        maxLevels = std::max(maxLevels, L.nLevels);
    }
    ellOff.push_back(ellBase);
    lOff.push_back(lBase);
    uOff.push_back(uBase);
    ifRowFace.push_back(ifFaceBase);
    ifPairOff.push_back(ifPairBase);
    nSlices_ = slice;
    nPos_ = slice * kSlice;
    nBlk_ = (int)blocks.size();
    nLevelsTotal_ = lvlBase;

    // DILU launch order: level-major, systems interleaved
    std::vector<int32_t> order;
    order.reserve(nSlices_);
    for (int lv = 0; lv < maxLevels; ++lv)
        for (int q = 0; q < nS; ++q) {
            const SystemLayout &L = *lay_[q];
            if (lv >= L.nLevels) continue;
            for (int s = L.levelSliceBegin[lv]; s < L.levelSliceBegin[lv + 1]; ++s) order.push_back(sliceBegin_[q] + s);
        }
    nOrder_ = (int)order.size();

    hipStream_t st = stream_;
    ellOff_.upload(ellOff, st);
    ellCol_.upload(ellCol, st);
    ifMask_.upload(ifMask, st);
    ifBase_.upload(ifBase, st);
    ifRowFace_.upload(ifRowFace, st);
    ifPairOff_.upload(ifPairOff, st);
    ifPairCol_.upload(ifPairCol, st);
    lOff_.upload(lOff, st);
    lCol_.upload(lCol, st);
    uOff_.upload(uOff, st);
    uCol_.upload(uCol, st);
    cellOfPos_.upload(cellOfPos, st);
    blocks_.upload(blocks, st);
    info_.upload(hInfo_, st);
    sliceSys_.upload(sliceSys, st);
    sliceLevel_.upload(sliceLevel, st);
    levelCount_.upload(levelCount, st);
    order_.upload(order, st);
    cntA_.alloc(std::max(1, nLevelsTotal_));
    cntB_.alloc(std::max(1, nLevelsTotal_));
    cntC_.alloc(std::max(1, nLevelsTotal_));
    cntA_.zero(st);
    cntB_.zero(st);
    cntC_.zero(st);
    err_.alloc(1);
    err_.zero(st);

    m64_.ellVal.upload(hEllVal_, st);
    m64_.diag.upload(diag, st);
    hDiag_ = diag;
    m64_.ifCoeff.upload(hIfCoeff_, st);
    m64_.ifPairW.upload(hIfPairW_, st);
    m64_.lVal.upload(hLVal_, st);
    m64_.lProd.upload(hLProd_, st);
    m64_.uVal.upload(hUVal_, st);
    sumA_.upload(sumA, st);
    b64_.upload(source, st);
    x64_.upload(psi0, st);
    x0_.upload(psi0, st);
    d64_.alloc(nPos_);
    d64_.zero(st);
    for (auto &v : v64_) {
        v.alloc(nPos_);
        v.zero(st);
    }
    st64_.alloc(nS);
    st32_.alloc(nS);
    partials_.alloc((size_t)std::max(1, nBlk_) * 4);
    tickets_.alloc(std::max(1, nS));
    tickets_.zero(st);
    factorised_ = false;
    haveFloat_ = false;
    for (auto &v : v32_) v.release();
    HIP_OK(hipStreamSynchronize(st));
}

void Batch::ensureFloat()
{
    if (haveFloat_) return;
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
    m32_.lVal.upload(toF(hLVal_), st);
    m32_.lProd.upload(toF(hLProd_), st);
    m32_.uVal.upload(toF(hUVal_), st);
    for (auto &v : v32_) {
        v.alloc(nPos_);
        v.zero(st);
    }
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
    A.ifBase = ifBase_.p;
    A.ifRowFace = ifRowFace_.p;
    A.ifCoeff = m.ifCoeff.p;
    A.ifPairOff = ifPairOff_.p;
    A.ifPairCol = ifPairCol_.p;
    A.ifPairW = m.ifPairW.p;
    A.lOff = lOff_.p;
    A.lCol = lCol_.p;
    A.lVal = m.lVal.p;
    A.lProd = m.lProd.p;
    A.uOff = uOff_.p;
    A.uCol = uCol_.p;
    A.uVal = m.uVal.p;
    return A;
}

template <>
void Batch::bindKrylov<double>(Krylov<double> &K, bool)
{
    K.blocks = blocks_.p;
    K.nBlk = nBlk_;
    K.info = info_.p;
    K.state = st64_.p;
    K.partials = partials_.p;
    K.tickets = tickets_.p;
    K.sumA = sumA_.p;
    K.x = x64_.p;
    K.b = b64_.p;
    K.ax = v64_[0].p; K.r = v64_[1].p; K.rw = v64_[2].p; K.pv = v64_[3].p; K.v = v64_[4].p;
    K.sv = v64_[5].p; K.t = v64_[6].p; K.ph = v64_[7].p; K.sh = v64_[8].p; K.w = v64_[9].p;
    K.pA = v64_[10].p; K.rD = v64_[11].p;
}

template <>
void Batch::bindKrylov<float>(Krylov<float> &K, bool)
{
    K.blocks = blocks_.p;
    K.nBlk = nBlk_;
    K.info = info_.p;
    K.state = st32_.p;
    K.partials = partials_.p;
    K.tickets = tickets_.p;
    K.sumA = sumA_.p;
    K.x = v32_[0].p; K.b = v32_[1].p;
    K.ax = v32_[2].p; K.r = v32_[3].p; K.rw = v32_[4].p; K.pv = v32_[5].p; K.v = v32_[6].p;
    K.sv = v32_[7].p; K.t = v32_[8].p; K.ph = v32_[9].p; K.sh = v32_[10].p; K.w = v32_[11].p;
    K.pA = v32_[12].p; K.rD = v32_[13].p;
}

LevelView makeLevelView(const int32_t *sliceSys, const int32_t *sliceLevel, const int32_t *levelCount,
                        const int32_t *order, int *cntA, int *cntB, int *err, int nLevelsTotal, int nOrder)
{
    LevelView LV;
    LV.err = err;
    LV.sliceSys = sliceSys;
    LV.sliceLevel = sliceLevel;
    LV.levelCount = levelCount;
    LV.order = order;
    LV.cntA = cntA;
    LV.cntB = cntB;
    LV.nLevelsTotal = nLevelsTotal;
    LV.nOrder = nOrder;
    return LV;
}

void Batch::factorise()
{
    if (factorised_) return;
    LevelView LV = makeLevelView(sliceSys_.p, sliceLevel_.p, levelCount_.p, order_.p, cntA_.p, cntB_.p, err_.p,
                                 nLevelsTotal_, nOrder_);
    Krylov<double> K;
    bindKrylov(K, false);
    MatView<double> A = view(m64_);
    HIP_OK(launchFactor<double>(A, LV, K, A, d64_.p, cellOfPos_.p, cntC_.p, stream_));
    if (haveFloat_) {
        Krylov<float> K32;
        bindKrylov(K32, true);
        MatView<float> A32 = view(m32_);
        // FP32 rD from the FP64 factor: re-run the inversion into rD32
        HIP_OK(launchFactor<float>(A32, LV, K32, A, d64_.p, cellOfPos_.p, cntC_.p, stream_));
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
        throw std::runtime_error("HIP error: DILU level wait expired (device error flag set)");
    }
}

void Batch::initState(const SolverControls &c)
{
    std::vector<SysState> h(sys_.size());
    for (size_t q = 0; q < sys_.size(); ++q) {
        SysState s = {};
        s.tol = c.tolerance;
        s.relTol = c.relTol;
        s.minIter = c.minIter;
        s.maxIter = c.precision == SolverControls::Mixed ? c.maxOuter : c.maxIter;
        s.status = kRunning;
        s.nCells = hInfo_[q].nCells;
        h[q] = s;
    }
    HIP_OK(hipMemcpyAsync(st64_.p, h.data(), h.size() * sizeof(SysState), hipMemcpyHostToDevice, stream_));
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

void Batch::runFp64(const SolverControls &c)
{
    LevelView LV = makeLevelView(sliceSys_.p, sliceLevel_.p, levelCount_.p, order_.p, cntA_.p, cntB_.p, err_.p,
                                 nLevelsTotal_, nOrder_);
    Krylov<double> K;
    bindKrylov(K, false);
    MatView<double> A = view(m64_);
    const bool pcg = c.solver == SolverControls::PCG;
    HIP_OK(launchSetup<double>(A, K, pcg, stream_));
    std::vector<SysState> hs;
    readState(hs);
    int done = 0;
    while (anyRunning(hs) && done <= c.maxIter) {
        int chunk = std::min(c.checkInterval, c.maxIter - done + 1);
        for (int it = 0; it < chunk; ++it) {
            if (pcg) HIP_OK(launchPcgIteration<double>(A, LV, K, stream_));
            else HIP_OK(launchBicgIteration<double>(A, LV, K, stream_));
        }
        done += chunk;
        readState(hs);
    }
}

void Batch::runMixed(const SolverControls &c, std::vector<int> &innerIters)
{
    ensureFloat();
    factorise();
    LevelView LV = makeLevelView(sliceSys_.p, sliceLevel_.p, levelCount_.p, order_.p, cntA_.p, cntB_.p, err_.p,
                                 nLevelsTotal_, nOrder_);
    Krylov<double> K64;
    bindKrylov(K64, false);
    Krylov<float> K32;
    bindKrylov(K32, true);
    MatView<double> A64 = view(m64_);
    MatView<float> A32 = view(m32_);
    const bool pcg = c.solver == SolverControls::PCG;
    const int nS = (int)sys_.size();
    innerIters.assign(nS, 0);
    HIP_OK(launchSetup<double>(A64, K64, false, stream_));
    std::vector<SysState> hs, hi(nS);
    readState(hs);
    int outer = 0;
    while (anyRunning(hs) && outer <= c.maxOuter) {
        HIP_OK(launchRefinePrepare(nS, st64_.p, st32_.p, c.innerRelTol, c.innerMaxIter, stream_));
        HIP_OK(launchRefineRhs(K64, K32, stream_));
        HIP_OK(launchSetup<float>(A32, K32, pcg, stream_));
        int done = 0;
        for (;;) {
            HIP_OK(hipMemcpyAsync(hi.data(), st32_.p, nS * sizeof(SysState), hipMemcpyDeviceToHost, stream_));
            HIP_OK(hipStreamSynchronize(stream_));
            if (!anyRunning(hi) || done > c.innerMaxIter) break;
            int chunk = std::min(c.checkInterval, c.innerMaxIter - done + 1);
            for (int it = 0; it < chunk; ++it) {
                if (pcg) HIP_OK(launchPcgIteration<float>(A32, LV, K32, stream_));
                else HIP_OK(launchBicgIteration<float>(A32, LV, K32, stream_));
            }
            done += chunk;
        }
        for (int q = 0; q < nS; ++q)
            if (hi[q].status != kIdle) innerIters[q] += hi[q].nIter;
        HIP_OK(launchRefineUpdate(K64, K32, stream_));
        HIP_OK(launchRefineResidual(A64, K64, stream_));
        readState(hs);
        ++outer;
    }
}

void Batch::solve(const SolverControls &c, std::vector<SolvePerf> &perf)
{
    HIP_OK(hipSetDevice(device_));
    if (c.precision == SolverControls::Mixed) ensureFloat();
    HIP_OK(hipStreamSynchronize(stream_));
    auto t0 = std::chrono::steady_clock::now();
    // the preconditioner is built per solve call, as the coupled solver
    // object is constructed per coupledFvMatrix::solve call
    factorised_ = false;
    factorise();
    initState(c);
    std::vector<int> inner;
    if (c.precision == SolverControls::Mixed) runMixed(c, inner);
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
        p.singular = hs[q].status == kSingular;
        p.seconds = secs;
        if (c.precision == SolverControls::Mixed) {
            p.nIterations = inner[q];
            p.nOuter = hs[q].nIter;
            p.bytes = p.nIterations * bytesIter((int)q, 4) + p.nOuter * bytesSpmv((int)q, 8);
        } else {
            p.nIterations = hs[q].nIter;
            p.bytes = p.nIterations * bytesIter((int)q, 8);
        }
    }
}

void Batch::resetPsi()
{
    HIP_OK(hipMemcpyAsync(x64_.p, x0_.p, (size_t)nPos_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
}

void Batch::setPsi(int sys, int region, const double *psi)
{
    const SystemLayout &L = *lay_.at(sys);
    const CoupledSystem &S = *sys_.at(sys);
    int n = S.regions.at(region).nCells, o = L.regionOffset[region];
    std::vector<double> x(L.nPos);
    HIP_OK(hipMemcpy(x.data(), x64_.p + posBegin(sys), L.nPos * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < n; ++c) x[L.posOfCell[o + c]] = psi[c];
    HIP_OK(hipMemcpy(x64_.p + posBegin(sys), x.data(), L.nPos * sizeof(double), hipMemcpyHostToDevice));
}

void Batch::getPsi(int sys, int region, double *psi)
{
    const SystemLayout &L = *lay_.at(sys);
    const CoupledSystem &S = *sys_.at(sys);
    int n = S.regions.at(region).nCells, o = L.regionOffset[region];
    std::vector<double> x(L.nPos);
    HIP_OK(hipMemcpy(x.data(), x64_.p + posBegin(sys), L.nPos * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < n; ++c) psi[c] = x[L.posOfCell[o + c]];
}

void Batch::spmv(const double *xHost, double *yHost, int sys)
{
    const SystemLayout &L = *lay_.at(sys);
    std::vector<double> x(nPos_, 0.0), y(nPos_);
    for (long long c = 0; c < L.nCells; ++c) x[posBegin(sys) + L.posOfCell[c]] = xHost[c];
    DevBuf<double> dx, dy;
    dx.upload(x, stream_);
    dy.alloc(nPos_);
    HIP_OK(launchSpmvPlain<double>(view(m64_), blocks_.p, nBlk_, dx.p, dy.p, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipMemcpy(y.data(), dy.p, nPos_ * sizeof(double), hipMemcpyDeviceToHost));
    for (long long c = 0; c < L.nCells; ++c) yHost[c] = y[posBegin(sys) + L.posOfCell[c]];
}

void Batch::diluFactor(int sys, double *rDHost)
{
    factorise();
    const SystemLayout &L = *lay_.at(sys);
    std::vector<double> r(nPos_);
    HIP_OK(hipStreamSynchronize(stream_));
    checkDeviceError();
    HIP_OK(hipMemcpy(r.data(), v64_[11].p, nPos_ * sizeof(double), hipMemcpyDeviceToHost));
    for (long long c = 0; c < L.nCells; ++c) rDHost[c] = r[posBegin(sys) + L.posOfCell[c]];
}

void Batch::precondition(const double *rHost, double *wHost, int sys)
{
    factorise();
    const SystemLayout &L = *lay_.at(sys);
    std::vector<double> r(nPos_, 0.0), w(nPos_);
    for (long long c = 0; c < L.nCells; ++c) r[posBegin(sys) + L.posOfCell[c]] = rHost[c];
    // every system must be marked running for the sweeps to act
    std::vector<SysState> hs(sys_.size());
    for (auto &s : hs) s.status = kRunning;
    HIP_OK(hipMemcpyAsync(st64_.p, hs.data(), hs.size() * sizeof(SysState), hipMemcpyHostToDevice, stream_));
    DevBuf<double> dr, dw;
    dr.upload(r, stream_);
    dw.alloc(nPos_);
    dw.zero(stream_);
    LevelView LV = makeLevelView(sliceSys_.p, sliceLevel_.p, levelCount_.p, order_.p, cntA_.p, cntB_.p, err_.p,
                                 nLevelsTotal_, nOrder_);
    Krylov<double> K;
    bindKrylov(K, false);
    HIP_OK(launchPrecondition<double>(view(m64_), LV, K, dr.p, dw.p, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    checkDeviceError();
    HIP_OK(hipMemcpy(w.data(), dw.p, nPos_ * sizeof(double), hipMemcpyDeviceToHost));
    for (long long c = 0; c < L.nCells; ++c) wHost[c] = w[posBegin(sys) + L.posOfCell[c]];
}

double Batch::timeKernel(int which, int reps)
{
    HIP_OK(hipSetDevice(device_));
    factorise();
    std::vector<SysState> hs(sys_.size());
    for (auto &x : hs) x.status = kRunning;
    HIP_OK(hipMemcpyAsync(st64_.p, hs.data(), hs.size() * sizeof(SysState), hipMemcpyHostToDevice, stream_));
    LevelView LV = makeLevelView(sliceSys_.p, sliceLevel_.p, levelCount_.p, order_.p, cntA_.p, cntB_.p, err_.p,
                                 nLevelsTotal_, nOrder_);
    Krylov<double> K;
    bindKrylov(K, false);
    MatView<double> A = view(m64_);
    const int nE = 2 * reps + 2;
    std::vector<hipEvent_t> ev(nE);
    for (auto &e : ev) HIP_OK(hipEventCreate(&e));
    // DILU sweeps are timed as forward/backward pairs (the counters of one
    // sweep are reset by the other), each sweep between its own events.
    auto one = [&](int i) {
        if (which == 0) {
            HIP_OK(hipEventRecord(ev[2 * i], stream_));
            HIP_OK(launchSpmvPlain<double>(A, blocks_.p, nBlk_, x64_.p, v64_[0].p, stream_));
            HIP_OK(hipEventRecord(ev[2 * i + 1], stream_));
        } else if (which == 3) {
            HIP_OK(hipEventRecord(ev[2 * i], stream_));
            HIP_OK(launchFactor<double>(A, LV, K, A, d64_.p, cellOfPos_.p, cntC_.p, stream_));
            HIP_OK(hipEventRecord(ev[2 * i + 1], stream_));
        } else {
            HIP_OK(launchDiluSweeps<double>(A, LV, K, b64_.p, v64_[9].p, stream_, which == 1 ? ev[2 * i] : nullptr,
                                             which == 1 ? ev[2 * i + 1] : nullptr,
                                             which == 2 ? ev[2 * i] : nullptr, which == 2 ? ev[2 * i + 1] : nullptr));
        }
    };
    one(reps);   // warm-up into the spare event pair
    for (int i = 0; i < reps; ++i) one(i);
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
