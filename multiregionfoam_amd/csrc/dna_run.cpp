// Partitioned (Dirichlet-Neumann) coupling driver; see dna_run.hpp.
#include "dna_run.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <stdexcept>

namespace mrfg {

#define HIP_OK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
    } while (0)

DnaRun::DnaRun(int device, const std::vector<const CoupledSystem *> &cases, bool preferLevels)
    : B_(device), cases_(cases)
{
    if (cases.empty()) throw std::runtime_error("no partitioned case");
    std::vector<const CoupledSystem *> ptrs;
    for (size_t c = 0; c < cases.size(); ++c) {
        const CoupledSystem &S = *cases[c];
        if (!S.part) throw std::runtime_error("system " + std::to_string(c) + " is not a partitioned case");
        if (!S.interfaces.empty()) throw std::runtime_error("a partitioned case has monolithic interfaces");
        maxRegions_ = std::max(maxRegions_, (int)S.regions.size());
        sysOf_.emplace_back();
        for (const LduRegion &R : S.regions) {
            std::unique_ptr<CoupledSystem> one(new CoupledSystem);
            one->regions.push_back(R);
            sysOf_.back().push_back((int)regionSys_.size());
            ptrs.push_back(one.get());
            regionSys_.push_back(std::move(one));
        }
    }
    B_.setPreferLevels(preferLevels);
    B_.setSystems(ptrs);

    std::vector<int32_t> dPos, dOff{0}, dIdx, nPos, nOff{0}, nIdx, termVal;
    std::vector<double> dDc, dKb, dW, nDc, nKb, nW, termCoeff;
    dSidesByRegion_.assign(maxRegions_, {});
    nSidesByRegion_.assign(maxRegions_, {});
    cellsByRegion_.assign(maxRegions_, {});
    for (size_t c = 0; c < cases.size(); ++c) {
        const CoupledSystem &S = *cases[c];
        const PartitionedModel &M = *S.part;
        const int nR = (int)S.regions.size();
        if ((int)M.sourceBase.size() != nR) throw std::runtime_error("partitioned model does not match its case");
        std::vector<int> dBase(M.pairs.size()), nBase(M.pairs.size());
        int dCount = (int)dPos.size(), nCount = (int)nPos.size();
        for (size_t q = 0; q < M.pairs.size(); ++q) {
            dBase[q] = dCount;
            nBase[q] = nCount;
            dCount += M.pairs[q].dir.nFaces();
            nCount += M.pairs[q].neu.nFaces();
        }
        for (size_t q = 0; q < M.pairs.size(); ++q) {
            const PartitionedPair &P = M.pairs[q];
            const PartitionedSide &D = P.dir, &N = P.neu;
            auto checkSide = [&](const PartitionedSide &X, int nOther) {
                if (X.region < 0 || X.region >= nR || X.off.size() != (size_t)X.nFaces() + 1 ||
                    X.idx.size() != X.w.size() || (size_t)X.off.back() != X.w.size() ||
                    X.coeff.size() != (size_t)X.nFaces() || X.kb.size() != (size_t)X.nFaces() ||
                    X.dc.size() != (size_t)X.nFaces() || X.value0.size() != (size_t)X.nFaces())
                    throw std::runtime_error("partitioned interface '" + P.name + "' has inconsistent sizes");
                for (int32_t i : X.idx)
                    if (i < 0 || i >= nOther) throw std::runtime_error("partitioned interface '" + P.name +
                                                                       "': gather index out of range");
                for (int32_t f : X.fc)
                    if (f < 0 || f >= S.regions[X.region].nCells)
                        throw std::runtime_error("partitioned interface '" + P.name + "': face cell out of range");
            };
            checkSide(D, N.nFaces());
            checkSide(N, D.nFaces());
            const int sd = sysOf_[c][D.region], sn = sysOf_[c][N.region];
            const int mc = std::max(1, M.maxCoupleIter);
            DnaSide ds{(int)dPos.size(), (int)dPos.size() + D.nFaces(), D.accType, (int)c, (int)dSides_.size(), mc,
                       D.relax, P.K, dHistSize_};
            DnaSide ns{(int)nPos.size(), (int)nPos.size() + N.nFaces(), N.accType, (int)c, (int)nSides_.size(), mc,
                       N.relax, P.K, nHistSize_};
            if (D.accType == 2) dHistSize_ += dnaHistSize(D.nFaces(), mc);
            if (N.accType == 2) nHistSize_ += dnaHistSize(N.nFaces(), mc);
            for (int i = 0; i < D.nFaces(); ++i) {
                dPos.push_back(B_.position(sd, D.fc[i]));
                dDc.push_back(D.dc[i]);
                dKb.push_back(D.kb[i]);
                dValue0_.push_back(D.value0[i]);
                for (int p = D.off[i]; p < D.off[i + 1]; ++p) {
                    dIdx.push_back(nBase[q] + D.idx[p]);
                    dW.push_back(D.w[p]);
                }
                dOff.push_back((int32_t)dW.size());
            }
            for (int j = 0; j < N.nFaces(); ++j) {
                nPos.push_back(B_.position(sn, N.fc[j]));
                nDc.push_back(N.dc[j]);
                nKb.push_back(N.kb[j]);
                nValue0_.push_back(N.value0[j]);
                for (int p = N.off[j]; p < N.off[j + 1]; ++p) {
                    nIdx.push_back(dBase[q] + N.idx[p]);
                    nW.push_back(N.w[p]);
                }
                nOff.push_back((int32_t)nW.size());
            }
            pairs_.push_back(DnaPair{ds.id, ns.id, (int)c, 0});
            pairCase_.push_back((int)c);
            dSides_.push_back(ds);
            nSides_.push_back(ns);
            dSidesByRegion_[D.region].push_back(ds);
            nSidesByRegion_[N.region].push_back(ns);
        }
        // source terms per cell, in pair order, Dirichlet side first, faces
        // in order (a cell's terms keep that order)
        for (int r = 0; r < nR; ++r) {
            std::map<int, std::vector<std::pair<double, int>>> terms;
            std::vector<int> order;
            for (size_t q = 0; q < M.pairs.size(); ++q) {
                const PartitionedPair &P = M.pairs[q];
                for (int side = 0; side < 2; ++side) {
                    const PartitionedSide &X = side ? P.neu : P.dir;
                    if (X.region != r) continue;
                    for (int i = 0; i < X.nFaces(); ++i) {
                        auto &t = terms[X.fc[i]];
                        if (t.empty()) order.push_back(X.fc[i]);
                        t.push_back({X.coeff[i], side ? -(nBase[q] + i) - 1 : dBase[q] + i});
                    }
                }
            }
            for (int cell : order) {
                DnaCell C;
                C.pos = B_.position(sysOf_[c][r], cell);
                C.caseId = (int)c;
                C.tBegin = (int)termVal.size();
                for (auto &t : terms[cell]) {
                    termCoeff.push_back(t.first);
                    termVal.push_back(t.second);
                }
                C.tEnd = (int)termVal.size();
                C.base = M.sourceBase[r].at(cell);
                C.pad = 0.0;
                cellsByRegion_[r].push_back(C);
            }
        }
    }
    hipStream_t st = B_.stream();
    auto up32 = [&](DevBuf<int32_t> &b, const std::vector<int32_t> &v) { b.upload(v.empty() ? std::vector<int32_t>{0} : v, st); };
    auto upD = [&](DevBuf<double> &b, const std::vector<double> &v) { b.upload(v.empty() ? std::vector<double>{0.0} : v, st); };
    up32(dPos_, dPos); up32(dOff_, dOff); up32(dIdx_, dIdx);
    up32(nPos_, nPos); up32(nOff_, nOff); up32(nIdx_, nIdx);
    up32(termVal_, termVal);
    upD(dDc_, dDc); upD(dKb_, dKb); upD(dW_, dW);
    upD(nDc_, nDc); upD(nKb_, nKb); upD(nW_, nW);
    upD(termCoeff_, termCoeff);
    for (auto *b : {&dT_, &dPrev_, &dRes_, &dPrevRes_}) b->alloc(std::max<size_t>(1, dPos.size()));
    for (auto *b : {&nG_, &nPrev_, &nRes_, &nPrevRes_, &nT_}) b->alloc(std::max<size_t>(1, nPos.size()));
    dOmega_.alloc(std::max<size_t>(1, dSides_.size()));
    nOmega_.alloc(std::max<size_t>(1, nSides_.size()));
    dCorr_.alloc(std::max<size_t>(1, dSides_.size()));
    nCorr_.alloc(std::max<size_t>(1, nSides_.size()));
    dCols_.alloc(std::max<size_t>(1, dSides_.size()));
    nCols_.alloc(std::max<size_t>(1, nSides_.size()));
    dHist_.alloc((size_t)std::max<long long>(1, dHistSize_));
    nHist_.alloc((size_t)std::max<long long>(1, nHistSize_));
    caseActive_.alloc(cases.size());
    resOut_.alloc(2 * std::max<size_t>(1, pairs_.size()));
    dSidesDev_.upload(dSides_, st);
    nSidesDev_.upload(nSides_, st);
    pairsDev_.upload(pairs_, st);
    for (int r = 0; r < maxRegions_; ++r) {
        dSidesR_.emplace_back(new DevBuf<DnaSide>);
        nSidesR_.emplace_back(new DevBuf<DnaSide>);
        cellsR_.emplace_back(new DevBuf<DnaCell>);
        if (!dSidesByRegion_[r].empty()) dSidesR_[r]->upload(dSidesByRegion_[r], st);
        if (!nSidesByRegion_[r].empty()) nSidesR_[r]->upload(nSidesByRegion_[r], st);
        if (!cellsByRegion_[r].empty()) cellsR_[r]->upload(cellsByRegion_[r], st);
    }
    // every index the kernels follow stays inside its array
    for (int32_t p : dPos) if (p < 0) throw std::runtime_error("DNA: face position out of range");
    for (int32_t p : nPos) if (p < 0) throw std::runtime_error("DNA: face position out of range");
    for (int32_t v : termVal)
        if (v >= (int32_t)dPos.size() || (v < 0 && -v - 1 >= (int32_t)nPos.size()))
            throw std::runtime_error("DNA: source term names no face");
    HIP_OK(hipStreamSynchronize(st));
    history_.assign(cases.size(), {});
}

DnaDev DnaRun::dev()
{
    DnaDev m;
    m.x = B_.fieldDevice();
    m.base = nullptr;
    m.caseActive = caseActive_.p;
    m.dPos = dPos_.p; m.dOff = dOff_.p; m.dIdx = dIdx_.p;
    m.dDc = dDc_.p; m.dKb = dKb_.p; m.dW = dW_.p;
    m.dT = dT_.p; m.dPrev = dPrev_.p; m.dRes = dRes_.p; m.dPrevRes = dPrevRes_.p;
    m.nPos = nPos_.p; m.nOff = nOff_.p; m.nIdx = nIdx_.p;
    m.nDc = nDc_.p; m.nKb = nKb_.p; m.nW = nW_.p;
    m.nG = nG_.p; m.nPrev = nPrev_.p; m.nRes = nRes_.p; m.nPrevRes = nPrevRes_.p; m.nT = nT_.p;
    m.dOmega = dOmega_.p; m.nOmega = nOmega_.p; m.dCorr = dCorr_.p; m.nCorr = nCorr_.p;
    m.dCols = dCols_.p; m.nCols = nCols_.p; m.dHist = dHist_.p; m.nHist = nHist_.p;
    return m;
}

// the initial state: fields as committed, Dirichlet values (and their
// acceleration's previous field) the patch values, Neumann gradients zero
// with the acceleration's previous field the patch values (initialize(*this)
// of both patch types), relaxation factors at their initial values
void DnaRun::reset()
{
    hipStream_t st = B_.stream();
    B_.resetPsi();
    const size_t nd = dValue0_.size(), nn = nValue0_.size();
    std::vector<double> zd(std::max<size_t>(1, nd), 0.0), zn(std::max<size_t>(1, nn), 0.0);
    auto put = [&](DevBuf<double> &b, const std::vector<double> &v) {
        if (!v.empty()) HIP_OK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, st));
    };
    put(dT_, dValue0_); put(dPrev_, dValue0_); put(dRes_, zd); put(dPrevRes_, zd);
    put(nG_, zn); put(nPrev_, nValue0_); put(nRes_, zn); put(nPrevRes_, zn); put(nT_, nValue0_);
    std::vector<double> wd, wn;
    for (auto &s : dSides_) wd.push_back(s.relax);
    for (auto &s : nSides_) wn.push_back(s.relax);
    put(dOmega_, wd); put(nOmega_, wn);
    std::vector<int32_t> cd(std::max<size_t>(1, dSides_.size()), 0), cn(std::max<size_t>(1, nSides_.size()), 0);
    HIP_OK(hipMemcpyAsync(dCorr_.p, cd.data(), cd.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(nCorr_.p, cn.data(), cn.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dCols_.p, cd.data(), cd.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(nCols_.p, cn.data(), cn.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    for (auto &h : history_) h.clear();
}

void DnaRun::run(const SolverControls &c, std::vector<DnaPerf> &out)
{
    if (c.precision != SolverControls::FP64) throw std::runtime_error("partitioned coupling runs in precision fp64");
    reset();
    loop(c, out, false);
}

void DnaRun::timeSteps(const SolverControls &c, bool backward, int nSteps, std::vector<DnaPerf> &out)
{
    if (c.precision != SolverControls::FP64) throw std::runtime_error("partitioned coupling runs in precision fp64");
    if (!B_.transient()) throw std::runtime_error("time stepping needs a transient partitioned case (ddt Euler or backward)");
    if (nSteps < 1) throw std::runtime_error("time stepping needs nSteps >= 1");
    reset();
    hipStream_t st = B_.stream();
    out.clear();
    const size_t nPosB = B_.nPositions();
    bBase_.alloc(nPosB);
    std::vector<int32_t> zd(std::max<size_t>(1, dSides_.size()), 0), zn(std::max<size_t>(1, nSides_.size()), 0);
    for (int n = 0; n < nSteps; ++n) {
        // the ddt terms of this step; the sources without the partitioned
        // terms are kept for the interface cells
        B_.beginTimeStep(backward);
        HIP_OK(hipMemcpyAsync(bBase_.p, B_.rhsDevice(), nPosB * sizeof(double), hipMemcpyDeviceToDevice, st));
        // a new time value: the acceleration models restart their correction
        // count (IQN-ILS with couplingReuse 0 drops its columns)
        HIP_OK(hipMemcpyAsync(dCorr_.p, zd.data(), zd.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(nCorr_.p, zn.data(), zn.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(dCols_.p, zd.data(), zd.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(nCols_.p, zn.data(), zn.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIP_OK(hipStreamSynchronize(st));
        std::vector<DnaPerf> step;
        loop(c, step, true);
        out.insert(out.end(), step.begin(), step.end());
    }
}

void DnaRun::loop(const SolverControls &c, std::vector<DnaPerf> &out, bool stepBase)
{
    hipStream_t st = B_.stream();
    const int nC = (int)cases_.size();
    out.assign(nC, DnaPerf{});
    for (auto &h : history_) h.clear();
    std::vector<int> corr(nC, 0), active(nC, 1);
    std::vector<double> jump(nC, 0.0), flux(nC, 0.0);
    DnaDev m = dev();
    m.base = stepBase ? bBase_.p : nullptr;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> res(2 * std::max<size_t>(1, pairs_.size()));
    while (true) {
        bool any = false;
        for (int q = 0; q < nC; ++q) {
            if (!active[q]) continue;
            const PartitionedModel &M = *cases_[q]->part;
            ++corr[q];
            // dnaControl::loop: no check on the first pass; residual control,
            // then the iteration limit
            if (corr[q] > 1 && M.maxJumpRes >= 0.0 && jump[q] <= M.maxJumpRes && flux[q] <= M.maxFluxRes) {
                out[q].converged = 1;
                active[q] = 0;
            } else if (corr[q] > M.maxCoupleIter) {
                active[q] = 0;
            }
            if (!active[q]) out[q].iterations = corr[q] - 1;
            any = any || active[q];
        }
        if (!any) break;
        std::vector<int32_t> act(active.begin(), active.end());
        HIP_OK(hipMemcpyAsync(caseActive_.p, act.data(), act.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        for (int r = 0; r < maxRegions_; ++r) {
            HIP_OK(launchDnaDirichlet(m, dSidesR_[r]->p, (int)dSidesByRegion_[r].size(), st));
            HIP_OK(launchDnaNeumann(m, nSidesR_[r]->p, (int)nSidesByRegion_[r].size(), st));
            HIP_OK(launchDnaSource(m, cellsR_[r]->p, (int)cellsByRegion_[r].size(), termCoeff_.p, termVal_.p,
                                   B_.rhsDevice(), st));
            std::vector<char> mask(regionSys_.size(), 0);
            bool anyR = false;
            for (int q = 0; q < nC; ++q)
                if (active[q] && r < (int)sysOf_[q].size()) {
                    mask[sysOf_[q][r]] = 1;
                    anyR = true;
                }
            if (anyR) {
                std::vector<SolvePerf> perf;
                B_.solve(c, perf, &mask);
                for (int q = 0; q < nC; ++q)
                    if (active[q] && r < (int)sysOf_[q].size()) {
                        out[q].regionSolves += 1;
                        out[q].solverIterations += perf[sysOf_[q][r]].nIterations;
                    }
            }
            HIP_OK(launchDnaPatchValues(m, nSidesR_[r]->p, (int)nSidesByRegion_[r].size(), st));
        }
        HIP_OK(launchDnaResiduals(m, dSidesDev_.p, nSidesDev_.p, pairsDev_.p, (int)pairs_.size(), resOut_.p, st));
        HIP_OK(hipMemcpyAsync(res.data(), resOut_.p, res.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        std::vector<double> omega(std::max<size_t>(1, dSides_.size()));
        HIP_OK(hipMemcpyAsync(omega.data(), dOmega_.p, omega.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        for (int q = 0; q < nC; ++q) {
            if (!active[q]) continue;
            jump[q] = flux[q] = 0.0;
            double w0 = -1.0;
            for (size_t p = 0; p < pairs_.size(); ++p)
                if (pairCase_[p] == q) {
                    jump[q] = std::max(jump[q], res[2 * p]);
                    flux[q] = std::max(flux[q], res[2 * p + 1]);
                    if (w0 < 0.0) w0 = omega[pairs_[p].dir];
                }
            history_[q].push_back({jump[q], flux[q], w0});
        }
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int q = 0; q < nC; ++q) {
        out[q].jumpResidual = jump[q];
        out[q].fluxResidual = flux[q];
        out[q].seconds = secs;
    }
}

void DnaRun::getPsi(int caseId, int region, double *psi)
{
    B_.getPsi(sysOf_.at(caseId).at(region), 0, psi);
}

}  // namespace mrfg
