// Device layout of coupled block systems.
//
// Rows ("positions") of a system are permuted into DILU level order: a
// cell's level is one more than the largest level of its lower neighbours
// (l(f) < u(f) in the region's own numbering), levels of all regions are
// merged, and each level is padded to a whole number of 64-row slices (one
// wavefront each).  Level order is a topological order of the factor's
// dependency graph, so the Cholesky/DILU factor and sweeps computed in it
// are the reference preconditioner itself, not a reordered variant.
//
// Per slice of 64 rows:
//   SpMV   sliced ELL, column-major inside the slice: entry k of lane l at
//          (ellOff[s] + k) * 64 + l; diagonal kept apart; row entries in
//          the LDU face-loop order (lower ascending, then upper ascending).
//   GGI    rows with interface faces are flagged in a 64-bit mask; their
//          faces hold the boundary coefficient and the weighted gather
//          pairs (position, weight) of the other region's cells.
//   DILU   lower dependencies (ascending l, face order) with lower coeffs
//          and upper*lower products, upper dependencies in descending u
//          (reverse face order) with upper coeffs; -1 marks padding.
#pragma once

#include "coupled_system.hpp"

#include <cstdint>
#include <vector>

namespace mrfg {

constexpr int kSlice = 64;          // rows per slice = wavefront size
constexpr int kSlicesPerBlock = 4;  // 256-thread blocks for SpMV / vector ops

struct SystemLayout {
    long long nCells = 0, nFaces = 0, nPairs = 0;
    int nPos = 0, nSlices = 0, nLevels = 0;
    std::vector<int32_t> regionOffset;       // first global cell of each region
    std::vector<int32_t> cellOfPos;          // global cell or -1 (padding)
    std::vector<int32_t> posOfCell;
    std::vector<int32_t> levelSliceBegin;    // nLevels + 1
    // SpMV
    std::vector<int32_t> ellOff;             // nSlices + 1, in 64-entry columns
    std::vector<int32_t> ellCol;
    std::vector<double> ellVal;
    std::vector<double> diag, source, sumA, psi0;
    // GGI interface rows
    std::vector<uint64_t> ifMask;            // per slice
    std::vector<int32_t> ifBase;             // per slice, first interface row
    std::vector<int32_t> ifRowFace;          // interface rows + 1
    std::vector<double> ifCoeff;             // per interface face
    std::vector<int32_t> ifPairOff;          // faces + 1
    std::vector<int32_t> ifPairCol;
    std::vector<double> ifPairW;
    // DILU
    std::vector<int32_t> lOff, uOff;         // nSlices + 1, in 64-entry columns
    std::vector<int32_t> lCol, uCol;
    std::vector<double> lVal, lProd, uVal;
};

SystemLayout buildLayout(const CoupledSystem &sys);

}  // namespace mrfg
