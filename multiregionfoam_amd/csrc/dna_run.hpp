// Partitioned (Dirichlet-Neumann) coupling driver: the DNA loop of
// multiRegionSystem::solve (multiRegionSystem.C:571-610) under dnaControl
// (dnaControl.C:993-1061) for heatTransferInterface pairs, steady.  Every
// region of every case is a system of one device batch; per DNA iteration
// each region in order gets its partitioned patches updated on the device
// and the systems of that region (all cases still iterating) are solved in
// one batched solve.  Each case stops on its own residual control.
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "batch.hpp"
#include "dna.hpp"

namespace mrfg {

struct DnaPerf {
    int iterations = 0, converged = 0, regionSolves = 0, solverIterations = 0;
    double jumpResidual = 0.0, fluxResidual = 0.0, seconds = 0.0;
};

class DnaRun {
public:
    DnaRun(int device, const std::vector<const CoupledSystem *> &cases, bool preferLevels);
    // Run every case's DNA loop from the initial state; one DnaPerf per case.
    void run(const SolverControls &c, std::vector<DnaPerf> &out);
    // nSteps time steps from the initial state (Euler, or backward with a
    // first Euler step), each a DNA loop per case; out gets nSteps x nCases
    // entries (step-major).
    void timeSteps(const SolverControls &c, bool backward, int nSteps, std::vector<DnaPerf> &out);
    void getPsi(int caseId, int region, double *psi);
    // per DNA iteration of a case: jump residual, flux residual, relaxation
    // factor of the case's first Dirichlet side
    const std::vector<std::array<double, 3>> &history(int caseId) const { return history_.at(caseId); }
    int nCases() const { return (int)cases_.size(); }

private:
    void reset();
    void loop(const SolverControls &c, std::vector<DnaPerf> &out, bool stepBase);

    Batch B_;
    std::vector<const CoupledSystem *> cases_;
    std::vector<std::unique_ptr<CoupledSystem>> regionSys_;
    std::vector<std::vector<int>> sysOf_;          // [case][region] -> batch system
    int maxRegions_ = 0;
    // side faces and their initial state
    std::vector<double> dValue0_, nValue0_;
    std::vector<DnaSide> dSides_, nSides_;
    std::vector<DnaPair> pairs_;
    std::vector<int> pairCase_;
    std::vector<std::vector<DnaSide>> dSidesByRegion_, nSidesByRegion_;
    std::vector<std::vector<DnaCell>> cellsByRegion_;
    std::vector<std::vector<std::array<double, 3>>> history_;
    // device
    DevBuf<int32_t> dPos_, dOff_, dIdx_, nPos_, nOff_, nIdx_, termVal_, dCorr_, nCorr_, dCols_, nCols_, caseActive_;
    DevBuf<double> dHist_, nHist_;     // IQN-ILS histories
    DevBuf<double> bBase_;             // per step: sources with the ddt terms, without the partitioned terms
    long long dHistSize_ = 0, nHistSize_ = 0;
    DevBuf<double> dDc_, dKb_, dW_, dT_, dPrev_, dRes_, dPrevRes_, nDc_, nKb_, nW_, nG_, nPrev_, nRes_, nPrevRes_,
        nT_, dOmega_, nOmega_, termCoeff_, resOut_;
    DevBuf<DnaSide> dSidesDev_, nSidesDev_;
    std::vector<std::unique_ptr<DevBuf<DnaSide>>> dSidesR_, nSidesR_;
    std::vector<std::unique_ptr<DevBuf<DnaCell>>> cellsR_;
    DevBuf<DnaPair> pairsDev_;
    DnaDev dev() ;
};

}  // namespace mrfg
