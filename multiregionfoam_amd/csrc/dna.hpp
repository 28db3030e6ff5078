// Partitioned (Dirichlet-Neumann) coupling on the device: the interface
// updates of one DNA iteration between the region solves
// (genericRegionCoupledJump / genericRegionCoupledFlux updateCoeffs with
// their acceleration models, the Neumann patch evaluation, the source terms
// of the partitioned faces and the interface residual norms of dnaControl).
//
// Side faces are numbered batch-wide: Dirichlet faces (d*) and Neumann faces
// (n*) of every pair of every case; a side's GGI gather names faces of the
// other side set.  One block per side (or pair); per-face work is spread
// over the block and the sums are double-double.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrfg {

struct DnaSide {
    int begin, end;      // face range in its side set
    int accType;         // 0 fixed relaxation, 1 Aitken, 2 IQN-ILS
    int caseId;
    int id;              // index of the side's acceleration state
    int maxCols;         // IQN-ILS: history columns held
    double relax;        // initial relaxation factor
    double K;            // jump coefficient of the pair (Dirichlet sides)
    long long histBase;  // IQN-ILS: the side's slice of the history buffer
};

// IQN-ILS history of a side with n faces and m columns, from histBase:
// V (m x n), W (m x n), Q (m x n), the reference field and residual (n
// each), R (m x m, row-major) and the coefficients c (m).
inline long long dnaHistSize(int n, int m)
{
    return 3LL * m * n + 2LL * n + (long long)m * m + m;
}

struct DnaPair {
    int dir, neu;        // side indices
    int caseId, pad;
};

struct DnaDev {
    const double *x;                               // batch field (positions)
    const double *base;                            // per-position source base (time steps), or null
    const int *caseActive;
    // Dirichlet faces: cell position, deltaCoeffs, face k, GGI gather of Neumann faces
    const int32_t *dPos, *dOff, *dIdx;
    const double *dDc, *dKb, *dW;
    double *dT, *dPrev, *dRes, *dPrevRes;
    // Neumann faces: the same for the gradient; nT = patch values
    const int32_t *nPos, *nOff, *nIdx;
    const double *nDc, *nKb, *nW;
    double *nG, *nPrev, *nRes, *nPrevRes, *nT;
    // per side: Aitken factor, correction counter, IQN-ILS columns held
    double *dOmega, *nOmega;
    int *dCorr, *nCorr, *dCols, *nCols;
    double *dHist, *nHist;
};

// Source of one cell with partitioned faces: b[pos] = base (DnaDev::base[pos]
// when set, the step's sources with their ddt terms), then
// b += coeff * value for its terms in order (value: Dirichlet face v >= 0,
// Neumann face -v-1).
struct DnaCell {
    int pos, caseId, tBegin, tEnd;
    double base;
    double pad;
};

hipError_t launchDnaDirichlet(const DnaDev &m, const DnaSide *sides, int nSides, hipStream_t st);
hipError_t launchDnaNeumann(const DnaDev &m, const DnaSide *sides, int nSides, hipStream_t st);
hipError_t launchDnaSource(const DnaDev &m, const DnaCell *cells, int nCells, const double *coeff,
                           const int32_t *val, double *b, hipStream_t st);
hipError_t launchDnaPatchValues(const DnaDev &m, const DnaSide *sides, int nSides, hipStream_t st);
// out[2 * p] = jump residual, out[2 * p + 1] = flux residual of pair p
hipError_t launchDnaResiduals(const DnaDev &m, const DnaSide *dSides, const DnaSide *nSides, const DnaPair *pairs,
                              int nPairs, double *out, hipStream_t st);

}  // namespace mrfg
