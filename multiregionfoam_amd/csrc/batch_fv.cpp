// Batch: finite-volume models of systems read from disk, re-assembled on
// the device (the role of setCoupledEqns per solve, transportTemperature.C:
// 129-150, with the interface coefficients of monolithicTemperature).
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "batch.hpp"
#include "dilu.hpp"

namespace mrfg {

#define HIP_OK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
    } while (0)

void Batch::buildFv()
{
    fv_.reset();
    bool any = false;
    for (auto *S : sys_) any = any || S->fv;
    if (!any) return;
    std::unique_ptr<FvBatch> F(new FvBatch);
    const int nS = (int)sys_.size();
    F->cellBase.assign(nS, -1);
    F->faceBase.assign(nS, -1);
    F->sideBase.assign(nS, -1);
    F->regionBase.assign(nS, -1);

    std::vector<int32_t> cellRegion, faceRegion, own, nei, cfOff{0}, cf, cbOff{0}, bSide, sCell, sOff, sPairCell;
    std::vector<int8_t> bT, bU, bKFixed;
    std::vector<double> w, dc, magSf, Sf, V, k, U, T, bNDist, bMagSf, bSf, bTb, bUb, bKb, sNDist, sW, sPairNDist;
    std::vector<int32_t> cellOfPos(nPos_, -1), ellSrc, fSrc, bSrc, ifSrc;
    int cellBase = 0, faceBase = 0, sideBase = 0, pairBase = 0;

    for (int q = 0; q < nS; ++q) {
        const CoupledSystem &S = *sys_[q];
        const SystemLayout &L = *lay_[q];
        const size_t nDepQ = (size_t)L.nSlices * K_ * kSlice;
        if (!S.fv) {
            ellSrc.insert(ellSrc.end(), L.ellVal.size(), -1);
            fSrc.insert(fSrc.end(), nDepQ, -1);
            bSrc.insert(bSrc.end(), nDepQ, -1);
            ifSrc.insert(ifSrc.end(), L.ifCoeff.size(), -1);
            continue;
        }
        const FvModel &M = *S.fv;
        if (M.regions.size() != S.regions.size() || M.sides.size() != 2 * S.interfaces.size())
            throw std::runtime_error("finite-volume model does not match its system");
        F->cellBase[q] = cellBase;
        F->faceBase[q] = faceBase;
        F->sideBase[q] = sideBase;
        F->regionBase[q] = (int)F->hRegion.size();
        for (size_t r = 0; r < S.regions.size(); ++r) {
            const LduRegion &R = S.regions[r];
            const FvRegionModel &m = M.regions[r];
            const int rid = (int)F->hRegion.size();
            F->hRegion.push_back(FvRegionParams{m.transport ? 1 : 0, 0, m.rhoCp, m.rDeltaT});
            const int c0 = cellBase + L.regionOffset[r];
            const int f0 = faceBase + L.faceOffset[r];
            const size_t nb = m.bOwner.size();
            if ((int)m.w.size() != R.nFaces || (int)m.V.size() != R.nCells || (int)m.k.size() != R.nCells ||
                (int)m.T.size() != R.nCells || (m.transport && (int)m.U.size() != 3 * R.nCells) ||
                m.bT.size() != nb || m.bSide.size() != nb)
                throw std::runtime_error("finite-volume model of region '" + R.name + "' has inconsistent sizes");
            for (int32_t c : m.bOwner)
                if (c < 0 || c >= R.nCells)
                    throw std::runtime_error("finite-volume model of region '" + R.name + "': boundary owner out of range");
            for (int f = 0; f < R.nFaces; ++f) {
                own.push_back(c0 + R.lAddr[f]);
                nei.push_back(c0 + R.uAddr[f]);
                faceRegion.push_back(rid);
            }
            w.insert(w.end(), m.w.begin(), m.w.end());
            dc.insert(dc.end(), m.deltaCoeffs.begin(), m.deltaCoeffs.end());
            magSf.insert(magSf.end(), m.magSf.begin(), m.magSf.end());
            Sf.insert(Sf.end(), m.Sf.begin(), m.Sf.end());
            V.insert(V.end(), m.V.begin(), m.V.end());
            k.insert(k.end(), m.k.begin(), m.k.end());
            T.insert(T.end(), m.T.begin(), m.T.end());
            if (m.transport) U.insert(U.end(), m.U.begin(), m.U.end());
            else U.insert(U.end(), 3 * (size_t)R.nCells, 0.0);
            cellRegion.insert(cellRegion.end(), R.nCells, rid);
            // incident internal faces per cell, ascending face order
            std::vector<int> cnt(R.nCells + 1, 0);
            for (int f = 0; f < R.nFaces; ++f) {
                ++cnt[R.lAddr[f] + 1];
                ++cnt[R.uAddr[f] + 1];
            }
            for (int c = 0; c < R.nCells; ++c) cnt[c + 1] += cnt[c];
            const int cfBase = (int)cf.size();
            cf.resize(cf.size() + 2 * (size_t)R.nFaces);
            std::vector<int> fill(cnt.begin(), cnt.end() - 1);
            for (int f = 0; f < R.nFaces; ++f) {
                cf[cfBase + fill[R.lAddr[f]]++] = 2 * (f0 + f);
                cf[cfBase + fill[R.uAddr[f]]++] = 2 * (f0 + f) + 1;
            }
            for (int c = 0; c < R.nCells; ++c) cfOff.push_back(cfBase + cnt[c + 1]);
            // boundary faces grouped by owner cell, the reader's order kept
            std::vector<int> bc(R.nCells + 1, 0);
            for (size_t b = 0; b < nb; ++b) ++bc[m.bOwner[b] + 1];
            for (int c = 0; c < R.nCells; ++c) bc[c + 1] += bc[c];
            const int bBase = (int)bT.size();
            std::vector<size_t> order(nb);
            std::vector<int> bfill(bc.begin(), bc.end() - 1);
            for (size_t b = 0; b < nb; ++b) order[bfill[m.bOwner[b]]++] = b;
            for (size_t j = 0; j < nb; ++j) {
                const size_t b = order[j];
                bT.push_back(m.bT[b]);
                bU.push_back(m.bU[b]);
                bKFixed.push_back(m.bKFixed[b]);
                bNDist.push_back(m.bNDist[b]);
                bMagSf.push_back(m.bMagSf[b]);
                for (int a = 0; a < 3; ++a) {
                    bSf.push_back(m.bSf[3 * b + a]);
                    bUb.push_back(m.bUb[3 * b + a]);
                }
                bTb.push_back(m.bTb[b]);
                bKb.push_back(m.bKb[b]);
                bSide.push_back(m.bSide[b] >= 0 ? sideBase + m.bSide[b] : -1);
            }
            for (int c = 0; c < R.nCells; ++c) cbOff.push_back(bBase + bc[c + 1]);
        }
        int nSideQ = 0;
        for (size_t i = 0; i < S.interfaces.size(); ++i)
            for (int s = 0; s < 2; ++s) {
                const InterfaceSide &IS = s ? S.interfaces[i].b : S.interfaces[i].a;
                const InterfaceSide &IN = s ? S.interfaces[i].a : S.interfaces[i].b;
                const FvSideModel &D = M.sides[2 * i + s];
                if ((int)D.nDist.size() != IS.nFaces() || D.pairCell.size() != IS.w.size())
                    throw std::runtime_error("finite-volume model of interface '" + S.interfaces[i].name +
                                             "' has inconsistent sizes");
                const int cO = cellBase + L.regionOffset[IS.region], cN = cellBase + L.regionOffset[IN.region];
                for (int j = 0; j < IS.nFaces(); ++j) {
                    sCell.push_back(cO + IS.fc[j]);
                    sNDist.push_back(D.nDist[j]);
                    sOff.push_back(pairBase + IS.off[j]);
                }
                for (size_t p = 0; p < IS.w.size(); ++p) {
                    sW.push_back(IS.w[p]);
                    sPairNDist.push_back(D.pairNDist[p]);
                    sPairCell.push_back(cN + D.pairCell[p]);
                }
                pairBase += (int)IS.w.size();
                nSideQ += IS.nFaces();
            }
        // placement maps: this system's positions, ELL entries, DILU slots
        // (re-padded to the batch width K_) and interface-row faces
        const int p0 = posBegin(q);
        for (int p = 0; p < L.nPos; ++p)
            if (L.cellOfPos[p] >= 0) cellOfPos[p0 + p] = cellBase + L.cellOfPos[p];
        for (int32_t v : L.ellSrc) ellSrc.push_back(v >= 0 ? v + 2 * faceBase : -1);
        for (int s = 0; s < L.nSlices; ++s)
            for (int kk = 0; kk < K_; ++kk)
                for (int l = 0; l < kSlice; ++l) {
                    int fs = -1, bs = -1;
                    if (kk < L.K) {
                        const size_t e = ((size_t)s * L.K + kk) * kSlice + l;
                        fs = L.fSrc[e] >= 0 ? L.fSrc[e] + 2 * faceBase : -1;
                        bs = L.bSrc[e] >= 0 ? L.bSrc[e] + 2 * faceBase : -1;
                    }
                    fSrc.push_back(fs);
                    bSrc.push_back(bs);
                }
        for (int32_t v : L.ifSrc) ifSrc.push_back(v + sideBase);
        cellBase += (int)S.nCells();
        faceBase += (int)S.nFaces();
        sideBase += nSideQ;
    }
    sOff.push_back(pairBase);
    if (ellSrc.size() != m64_.ellVal.n || fSrc.size() != (size_t)nSlices_ * K_ * kSlice ||
        ifSrc.size() != m64_.ifCoeff.n)
        throw std::runtime_error("finite-volume placement maps do not match the batch layout");
    F->nCells = cellBase;
    F->nFaces = faceBase;
    F->nSideFaces = sideBase;
    F->nBFaces = (long long)bT.size();
    F->nPairs = pairBase;
    for (int q = 0; q < nS; ++q) {
        if (F->cellBase[q] < 0) continue;
        const SystemLayout &L = *lay_[q];
        F->nPosModel += L.nPos;
        F->nEll += (long long)L.ellVal.size();
        F->nSlots += (long long)L.nSlices * K_ * kSlice;
        F->nIfEntries += (long long)L.ifCoeff.size();
    }
    // every index the assembly kernels follow stays inside its array
    {
        auto fail = [](const char *what) {
            throw std::runtime_error(std::string("finite-volume model: index out of range in ") + what);
        };
        auto inRange = [&](const std::vector<int32_t> &v, long long lo, long long hi, const char *what) {
            for (int32_t x : v)
                if (x < lo || x >= hi) fail(what);
        };
        auto offsets = [&](const std::vector<int32_t> &off, size_t n, size_t total, const char *what) {
            if (off.size() != n + 1 || off[0] != 0 || (size_t)off.back() != total) fail(what);
            for (size_t i = 0; i < n; ++i)
                if (off[i] > off[i + 1]) fail(what);
        };
        const long long nC = cellBase, nF = faceBase, nSd = sideBase, nRg = (long long)F->hRegion.size();
        inRange(own, 0, nC, "owner");
        inRange(nei, 0, nC, "neighbour");
        inRange(faceRegion, 0, nRg, "face region");
        inRange(cellRegion, 0, nRg, "cell region");
        if ((long long)cellRegion.size() != nC || (long long)faceRegion.size() != nF ||
            w.size() != (size_t)nF || Sf.size() != 3 * (size_t)nF || V.size() != (size_t)nC ||
            k.size() != (size_t)nC || T.size() != (size_t)nC || U.size() != 3 * (size_t)nC)
            fail("array sizes");
        offsets(cfOff, nC, cf.size(), "cell faces");
        inRange(cf, 0, 2 * nF, "cell faces");
        offsets(cbOff, nC, bT.size(), "cell boundary faces");
        inRange(bSide, -1, nSd, "boundary side");
        for (size_t b = 0; b < bT.size(); ++b)
            if (bT[b] < 0 || bT[b] > 3 || (bT[b] == 3 && bSide[b] < 0) || bU[b] < 0 || bU[b] > 3) fail("boundary codes");
        if (bNDist.size() != bT.size() || bSf.size() != 3 * bT.size() || bUb.size() != 3 * bT.size() ||
            bTb.size() != bT.size() || bKb.size() != bT.size() || bMagSf.size() != bT.size())
            fail("boundary sizes");
        inRange(sCell, 0, nC, "side cell");
        offsets(sOff, (size_t)nSd, sW.size(), "side pairs");
        inRange(sPairCell, 0, nC, "side pair cell");
        if (sNDist.size() != (size_t)nSd || sPairNDist.size() != sW.size()) fail("side sizes");
        inRange(cellOfPos, -1, nC, "position cell");
        inRange(ellSrc, -1, 2 * nF, "ELL source");
        inRange(fSrc, -1, 2 * nF, "forward record source");
        inRange(bSrc, -1, 2 * nF, "backward record source");
        for (int32_t x : ifSrc)
            if (x >= nSd || x < -1) fail("interface source");
    }

    hipStream_t st = stream_;
    F->region.upload(F->hRegion, st);
    F->cellRegion.upload(cellRegion, st);
    F->faceRegion.upload(faceRegion, st);
    F->own.upload(own, st);
    F->nei.upload(nei, st);
    F->cfOff.upload(cfOff, st);
    F->cf.upload(cf, st);
    F->cbOff.upload(cbOff, st);
    F->bSide.upload(bSide, st);
    F->sCell.upload(sCell, st);
    F->sOff.upload(sOff, st);
    F->sPairCell.upload(sPairCell, st);
    F->bT.upload(bT, st);
    F->bU.upload(bU, st);
    F->bKFixed.upload(bKFixed, st);
    for (auto pr : {std::make_pair(&F->w, &w), std::make_pair(&F->dc, &dc), std::make_pair(&F->magSf, &magSf),
                    std::make_pair(&F->Sf, &Sf), std::make_pair(&F->V, &V), std::make_pair(&F->k, &k),
                    std::make_pair(&F->U, &U), std::make_pair(&F->T, &T), std::make_pair(&F->bNDist, &bNDist),
                    std::make_pair(&F->bMagSf, &bMagSf), std::make_pair(&F->bSf, &bSf),
                    std::make_pair(&F->bTb, &bTb), std::make_pair(&F->bUb, &bUb), std::make_pair(&F->bKb, &bKb),
                    std::make_pair(&F->sNDist, &sNDist), std::make_pair(&F->sW, &sW),
                    std::make_pair(&F->sPairNDist, &sPairNDist)})
        pr.first->upload(*pr.second, st);
    for (auto *b : {&F->lapU, &F->aLo, &F->aUp}) b->alloc(std::max(1, F->nFaces));
    F->faceVal.alloc(2 * (size_t)std::max(1, F->nFaces));
    for (auto *b : {&F->sK, &F->sDelta, &F->sKOwn, &F->sKNei, &F->ifCoeff}) b->alloc(std::max(1, F->nSideFaces));
    for (auto *b : {&F->diag, &F->source, &F->ddtA, &F->diagBase, &F->srcBase}) b->alloc(std::max(1, F->nCells));
    F->cellOfPos.upload(cellOfPos, st);
    F->ellSrc.upload(ellSrc.empty() ? std::vector<int32_t>{-1} : ellSrc, st);
    F->fSrc.upload(fSrc, st);
    F->bSrc.upload(bSrc, st);
    F->ifSrc.upload(ifSrc.empty() ? std::vector<int32_t>{-1} : ifSrc, st);
    fv_ = std::move(F);
}

bool Batch::hasModel(int sys) const
{
    return fv_ && sys >= 0 && sys < (int)fv_->cellBase.size() && fv_->cellBase[sys] >= 0;
}

void Batch::setField(int sys, int region, const std::string &name, const double *v, size_t n)
{
    HIP_OK(hipSetDevice(device_));
    if (!hasModel(sys)) throw std::runtime_error("system " + std::to_string(sys) + " has no finite-volume model "
                                                 "(only cases read from disk have one)");
    const CoupledSystem &S = *sys_[sys];
    const SystemLayout &L = *lay_[sys];
    if (region < 0 || region >= (int)S.regions.size()) throw std::runtime_error("region index out of range");
    const LduRegion &R = S.regions[region];
    const int rid = fv_->regionBase[sys] + region;
    const size_t c0 = (size_t)fv_->cellBase[sys] + L.regionOffset[region];
    auto put = [&](DevBuf<double> &dst, size_t per) {
        if (n != per * (size_t)R.nCells)
            throw std::runtime_error("field '" + name + "' of region '" + R.name + "' needs " +
                                     std::to_string(per * (size_t)R.nCells) + " values, got " + std::to_string(n));
        HIP_OK(hipMemcpyAsync(dst.p + per * c0, v, n * sizeof(double), hipMemcpyHostToDevice, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    };
    if (name == "k") {
        for (size_t i = 0; i < n; ++i)
            if (!(v[i] > 0.0)) throw std::runtime_error("region '" + R.name + "': conductivity must be positive");
        put(fv_->k, 1);
    } else if (name == "U") {
        if (!fv_->hRegion[rid].transport)
            throw std::runtime_error("region '" + R.name + "' has no velocity (not a transportTemperature region)");
        put(fv_->U, 3);
    } else if (name == "T") {
        put(fv_->T, 1);
        setPsi(sys, region, v);
        const size_t p0 = posBegin(sys);
        HIP_OK(hipMemcpyAsync(x0_.p + p0, x64_.p + p0, (size_t)L.nPos * sizeof(double), hipMemcpyDeviceToDevice,
                              stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    } else if (name == "rhoCp") {
        if (n != 1 || !(v[0] > 0.0)) throw std::runtime_error("rhoCp takes one positive value");
        fv_->hRegion[rid].rhoCp = v[0];
        fv_->region.upload(fv_->hRegion, stream_);
        HIP_OK(hipStreamSynchronize(stream_));
    } else {
        throw std::runtime_error("unknown field '" + name + "'; valid: k U T rhoCp");
    }
}

FvDev Batch::fvDevView() const
{
    const FvBatch &F = *fv_;
    FvDev m;
    m.nCells = F.nCells;
    m.nFaces = F.nFaces;
    m.nSideFaces = F.nSideFaces;
    m.region = F.region.p;
    m.cellRegion = F.cellRegion.p;
    m.faceRegion = F.faceRegion.p;
    m.own = F.own.p;
    m.nei = F.nei.p;
    m.w = F.w.p;
    m.dc = F.dc.p;
    m.magSf = F.magSf.p;
    m.Sf = F.Sf.p;
    m.V = F.V.p;
    m.k = F.k.p;
    m.U = F.U.p;
    m.T = F.T.p;
    m.cfOff = F.cfOff.p;
    m.cf = F.cf.p;
    m.cbOff = F.cbOff.p;
    m.bT = F.bT.p;
    m.bU = F.bU.p;
    m.bKFixed = F.bKFixed.p;
    m.bNDist = F.bNDist.p;
    m.bMagSf = F.bMagSf.p;
    m.bSf = F.bSf.p;
    m.bTb = F.bTb.p;
    m.bUb = F.bUb.p;
    m.bKb = F.bKb.p;
    m.bSide = F.bSide.p;
    m.sCell = F.sCell.p;
    m.sOff = F.sOff.p;
    m.sPairCell = F.sPairCell.p;
    m.sNDist = F.sNDist.p;
    m.sW = F.sW.p;
    m.sPairNDist = F.sPairNDist.p;
    m.lapU = F.lapU.p;
    m.aLo = F.aLo.p;
    m.aUp = F.aUp.p;
    m.faceVal = F.faceVal.p;
    m.sK = F.sK.p;
    m.sDelta = F.sDelta.p;
    m.sKOwn = F.sKOwn.p;
    m.sKNei = F.sKNei.p;
    m.ifCoeff = F.ifCoeff.p;
    m.diag = F.diag.p;
    m.source = F.source.p;
    m.ddtA = F.ddtA.p;
    m.diagBase = F.diagBase.p;
    m.srcBase = F.srcBase.p;
    return m;
}

FvPlace Batch::fvPlaceView()
{
    const FvBatch &F = *fv_;
    FvPlace q;
    q.nPos = nPos_;
    q.pos0 = 0;
    q.K = K_;
    q.G = packChunks(K_ + 1, 2);
    q.cellOfPos = F.cellOfPos.p;
    q.ellOff = ellOff_.p;
    q.ellSrc = F.ellSrc.p;
    q.fSrc = F.fSrc.p;
    q.bSrc = F.bSrc.p;
    q.ifMask = ifMask_.p;
    q.ifBase = ifBase_.p;
    q.ifRowFace = ifRowFace_.p;
    q.ifSrc = F.ifSrc.p;
    q.faceVal = F.faceVal.p;
    q.sideCoeff = F.ifCoeff.p;
    q.diag = F.diag.p;
    q.source = F.source.p;
    q.ddtA = F.ddtA.p;
    q.diagBase = F.diagBase.p;
    q.srcBase = F.srcBase.p;
    q.ellVal = m64_.ellVal.p;
    q.diagOut = m64_.diag.p;
    q.b = b64_.p;
    q.sumA = sumA_.p;
    q.ifCoeff = m64_.ifCoeff.p;
    q.fPack = (double *)m64_.fPack.p;
    q.bPack = (double *)m64_.bPack.p;
    q.prodPack = (double *)prodPack_.p;
    q.ddtAOut = transient_ ? ddtA_.p : nullptr;
    q.diagBaseOut = transient_ ? diagBase_.p : nullptr;
    q.srcBaseOut = transient_ ? srcBase_.p : nullptr;
    return q;
}

void Batch::assemble()
{
    HIP_OK(hipSetDevice(device_));
    if (!fv_) throw std::runtime_error("no system of the batch has a finite-volume model");
    HIP_OK(launchFvAssemble(fvDevView(), stream_));
    HIP_OK(launchFvPlace(fvPlaceView(), stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    fv_->assembled = true;
    factorised_ = false;
    haveFloat_ = false;
    hostStale_ = true;
    diagCoefft_ = 1.0;   // diag = diagBase + a, as assembled
}

double Batch::timeAssembly(int which, int reps)
{
    HIP_OK(hipSetDevice(device_));
    if (!fv_) throw std::runtime_error("no system of the batch has a finite-volume model");
    const FvDev m = fvDevView();
    const FvPlace q = fvPlaceView();
    // both passes once, so the timed one reads what a real assembly reads
    HIP_OK(launchFvAssemble(m, stream_));
    HIP_OK(launchFvPlace(q, stream_));
    std::vector<hipEvent_t> ev(2 * reps);
    for (auto &e : ev) HIP_OK(hipEventCreate(&e));
    for (int i = 0; i < reps; ++i) {
        HIP_OK(hipEventRecord(ev[2 * i], stream_));
        if (which == 7) HIP_OK(launchFvAssemble(m, stream_));
        else HIP_OK(launchFvPlace(q, stream_));
        HIP_OK(hipEventRecord(ev[2 * i + 1], stream_));
    }
    HIP_OK(hipStreamSynchronize(stream_));
    double tot = 0;
    for (int i = 0; i < reps; ++i) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
        tot += ms;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    fv_->assembled = true;
    factorised_ = false;
    haveFloat_ = false;
    hostStale_ = true;
    diagCoefft_ = 1.0;
    return tot / reps;
}

void Batch::assemblyBytes(double &assemble, double &place) const
{
    if (!fv_) throw std::runtime_error("no system of the batch has a finite-volume model");
    const FvBatch &F = *fv_;
    const double nF = F.nFaces, nC = F.nCells, nB = (double)F.nBFaces, nSd = F.nSideFaces, nP = F.nPairs;
    // each array the passes read or write, once: per internal face the
    // addressing, region, w, deltaCoeff, |S|, S (60), the cell-face lists
    // (8), the two coefficients (16) and the Laplacian / convection parts
    // written by the face pass and read by the cell pass (48); per cell the
    // region, V, list offsets (20), the fields k, U, T (40) and five results
    // (40); per boundary face 87; per interface-side face 40, per GGI pair 20
    assemble = nF * 132.0 + nC * 100.0 + nB * 87.0 + nSd * 40.0 + nP * 20.0;
    // per position of a model system: its cell (4), diag, source (16) and
    // diag, b, sumA written (24), ddt terms (48 when transient); per ELL
    // entry source + value (12); per DILU slot two sources + three values
    // (32) and the diagonal slot (8 per position); per interface-row face
    // source, gathered and written coefficient (20); the face values (16 per
    // face) and side coefficients (8) read once
    place = (double)F.nPosModel * (4 + 16 + 24 + 8 + (transient_ ? 48 : 0)) + (double)F.nEll * 12.0 +
            (double)F.nSlots * 32.0 + (double)F.nIfEntries * 20.0 + nF * 16.0 + nSd * 8.0;
}

void Batch::getRegionMatrix(int sys, int region, double *diag, double *lower, double *upper, double *source)
{
    HIP_OK(hipSetDevice(device_));
    if (!hasModel(sys) || !fv_->assembled) throw std::runtime_error("system has no device-assembled matrix");
    const CoupledSystem &S = *sys_[sys];
    const SystemLayout &L = *lay_[sys];
    const LduRegion &R = S.regions.at(region);
    const size_t c0 = (size_t)fv_->cellBase[sys] + L.regionOffset[region];
    const size_t f0 = (size_t)fv_->faceBase[sys] + L.faceOffset[region];
    std::vector<double> fvals(2 * (size_t)R.nFaces);
    copyToHost(diag, fv_->diag.p + c0, R.nCells * sizeof(double));
    copyToHost(source, fv_->source.p + c0, R.nCells * sizeof(double));
    if (R.nFaces > 0)
        copyToHost(fvals.data(), fv_->faceVal.p + 2 * f0, fvals.size() * sizeof(double));
    for (int f = 0; f < R.nFaces; ++f) {
        lower[f] = fvals[2 * (size_t)f];
        upper[f] = fvals[2 * (size_t)f + 1];
    }
}

void Batch::getInterfaceCoeff(int sys, int iface, int side, double *coeff)
{
    HIP_OK(hipSetDevice(device_));
    if (!hasModel(sys) || !fv_->assembled) throw std::runtime_error("system has no device-assembled matrix");
    const CoupledSystem &S = *sys_[sys];
    if (iface < 0 || iface >= (int)S.interfaces.size() || side < 0 || side > 1)
        throw std::runtime_error("interface or side out of range");
    size_t off = fv_->sideBase[sys];
    for (int i = 0; i < iface; ++i) off += S.interfaces[i].a.nFaces() + S.interfaces[i].b.nFaces();
    if (side) off += S.interfaces[iface].a.nFaces();
    const int n = side ? S.interfaces[iface].b.nFaces() : S.interfaces[iface].a.nFaces();
    if (n > 0) copyToHost(coeff, fv_->ifCoeff.p + off, n * sizeof(double));
}

void Batch::refreshHostCopies()
{
    HIP_OK(hipStreamSynchronize(stream_));
    auto down = [this](std::vector<double> &h, const void *d) {
        if (!h.empty()) copyToHost(h.data(), d, h.size() * sizeof(double));
    };
    down(hEllVal_, m64_.ellVal.p);
    down(hDiag_, m64_.diag.p);
    down(hIfCoeff_, m64_.ifCoeff.p);
    down(hFPack_, m64_.fPack.p);
    down(hBPack_, m64_.bPack.p);
    derivePack32(tileLocal32_);
    hostStale_ = false;
}

void Batch::derivePack32(bool tileLocal)
{
    const int G = packChunks(K_ + 1, 2), G32 = packChunks(K_ + 1, 4);
    hFPack32_.assign((size_t)nSlices_ * G32 * 64 * 4, 0.0f);
    hBPack32_.assign((size_t)nSlices_ * G32 * 64 * 4, 0.0f);
    auto at32 = [G32](int s, int k, int l) { return (((size_t)s * G32 + k / 4) * 64 + l) * 4 + k % 4; };
    for (int s = 0; s < nSlices_; ++s)
        for (int k = 0; k < K_; ++k)
            for (int l = 0; l < kSlice; ++l) {
                const size_t e64 = (((size_t)s * G + k / 2) * 64 + l) * 2 + k % 2;
                hFPack32_[at32(s, k, l)] = (float)hFPack_[e64];
                hBPack32_[at32(s, k, l)] = (float)hBPack_[e64];
            }
    if (!tileLocal) return;
    // The tile-local DILU: a box tile's row whose lower (forward record) or
    // upper (backward record) neighbour lies in the neighbouring tile drops
    // that coefficient.  Lane l = bl + TB*cl of a tile holds line (bl, cl); the
    // neighbour along axis q sits in record slot 2 - q.
    for (int w : hBoxWaves_) {
        const int32_t *m = &hBoxMeta_[(size_t)w * kBoxMeta];
        const int TB = m[kBmTB], TC = kSlice / TB, ax = m[kBmAxis];
        const int b = ax == 0 ? 1 : 0, c = ax == 2 ? 1 : 2;
        const int slotB = 2 - b, slotC = 2 - c;
        for (int t = 0; t < m[kBmSteps]; ++t) {
            const int s = m[kBmSlice] + t;
            for (int l = 0; l < kSlice; ++l) {
                const int bl = l % TB, cl = l / TB;
                if (bl == 0 && m[kBmPrevB] >= 0) hFPack32_[at32(s, slotB, l)] = 0.0f;
                if (cl == 0 && m[kBmPrevC] >= 0) hFPack32_[at32(s, slotC, l)] = 0.0f;
                if (bl == TB - 1 && m[kBmNextB] >= 0) hBPack32_[at32(s, slotB, l)] = 0.0f;
                if (cl == TC - 1 && m[kBmNextC] >= 0) hBPack32_[at32(s, slotC, l)] = 0.0f;
            }
        }
        // the cuts between the tile's segments along its sequential axis:
        // a segment's first row drops its lower neighbour (slot 2 - axis),
        // the previous segment's last row its upper neighbour
        const int segs = tileSegments(), slotA = 2 - ax;
        for (int sg = 1; sg < segs; ++sg) {
            int lo, hi;
            boxSegment(m[kBmNA], sg, segs, lo, hi);
            if (lo <= 0 || lo >= m[kBmNA]) continue;
            for (int cl = 0; cl < m[kBmActC]; ++cl)
                for (int bl = 0; bl < m[kBmActB]; ++bl) {
                    const int l = bl + TB * cl;
                    hFPack32_[at32(m[kBmSlice] + lo + bl + cl, slotA, l)] = 0.0f;
                    hBPack32_[at32(m[kBmSlice] + lo - 1 + bl + cl, slotA, l)] = 0.0f;
                }
        }
    }
}

}  // namespace mrfg
