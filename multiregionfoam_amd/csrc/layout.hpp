// Device layout of coupled block systems.
//
// Rows ("positions") of a system are grouped into 64-row slices; a slice is
// one step of one DILU wave.  A wave is a sequence of steps executed by one
// wavefront, lane l of step t holding one row (or nothing).  The waves of a
// system form its DILU schedule:
//
//   * structured hex regions (detected from the LDU addressing: faces to
//     c+1, c+nx, c+nx*ny in upper-triangular order) are cut into tiles of 64
//     grid lines along the longest axis; lane (b,c) of a tile walks its line
//     with a skew of b+c steps, so each of a cell's three lower neighbours
//     was computed one step earlier by the same or an adjacent lane;
//   * any other region is scheduled by levels (a cell's level is one more
//     than its lower neighbours' largest), 64 rows of one level per wave.
//
// Either way the schedule is a topological order of the factor's dependency
// graph, so the Cholesky/DILU factor and sweeps computed in it are the
// reference preconditioner itself.  A dependency on a row of the same wave
// at most kRing-1 steps back is read from the wave's LDS ring; any other is
// read from a halo slot the producing wave publishes.
//
// Per slice of 64 rows:
//   SpMV   sliced ELL, column-major inside the slice: entry k of lane l at
//          (ellOff[s] + k) * 64 + l; diagonal kept apart; row entries in the
//          LDU face-loop order (lower ascending, then upper ascending).
//   GGI    rows with interface faces are flagged in a 64-bit mask; their
//          faces hold the boundary coefficient and the weighted gather pairs
//          (position, weight) of the other region's cells.
//   DILU   K dependency slots per row at (s*K + k)*64 + l: forward = lower
//          neighbours in face order with lower coefficients and upper*lower
//          products, backward = upper neighbours in reverse face order with
//          upper coefficients; codes: kNoDep, ring slot -(slot+1), or halo.
#pragma once

#include "coupled_system.hpp"

#include <climits>
#include <cstdint>
#include <string>
#include <vector>

namespace mrfg {

constexpr int kSlice = 64;          // rows per slice = wavefront size
constexpr int kSlicesPerBlock = 4;  // 256-thread blocks for SpMV / vector ops
constexpr int kRing = 3;            // LDS ring depth (steps) of a DILU wave (slots < 192)
constexpr int32_t kNoDep = INT32_MIN;

// Per-wave metadata of a structured-box tile wave (kBoxMeta ints each).
enum {
    kBmTB = 0,      // lanes along the tile's first tangential axis (TC = 64/TB)
    kBmNA,          // cells along the sequential axis
    kBmPrevB,       // wave of the tile before / after along each tangential axis (-1: none)
    kBmPrevC,
    kBmNextB,
    kBmNextC,
    kBmAxis,        // sequential axis (0 x, 1 y, 2 z); b, c are the other two, ascending
    kBmActB,        // active lanes along b / c in this tile
    kBmActC,
    kBmSlice,       // first slice (filled by the batch)
    kBmSteps,
    kBmInbox,       // first inbox value of this wave (filled by the batch)
    kBmSys,
    kBoxMeta = 16
};

struct RegionSchedule {
    std::string kind;                // "box" or "levels"
    int n[3] = {0, 0, 0};            // box dimensions
    int seqAxis = -1, tileB = 0, tileC = 0;
    int nWaves = 0, maxSteps = 0;
};

struct SystemLayout {
    long long nCells = 0, nFaces = 0, nPairs = 0;
    int nPos = 0, nSlices = 0, nWaves = 0, K = 0;
    std::vector<RegionSchedule> schedules;
    std::vector<int32_t> regionOffset;       // first global cell of each region
    std::vector<int32_t> cellOfPos;          // global cell or -1 (padding)
    std::vector<int32_t> posOfCell;
    // waves, in dispatch order
    std::vector<int32_t> waveSlice, waveSteps, waveKey;
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
    std::vector<int32_t> fCode, bCode;       // nSlices * K * 64
    std::vector<double> fVal, fProd, bVal;
    std::vector<int32_t> fHalo, bHalo;       // per position: halo slot published or -1
    int nHaloF = 0, nHaloB = 0;
    // structured-box tile waves: their rows keep the DILU values in fixed
    // slots (slot = 2 - axis of the neighbour: z, y, x, which is both the
    // face order of the lower neighbours and the reverse face order of the
    // upper ones; zero where a neighbour is missing) and have no codes: the
    // dependencies follow from the tile geometry in boxMeta
    std::vector<uint8_t> waveBox;            // per wave: 1 if a box tile
    std::vector<int32_t> boxMeta;            // kBoxMeta per wave
    // where each value came from, so re-assembled coefficients can be put in
    // place on the device: face values are numbered 2*(faceOffset[r] + f)
    // (+0 the lower coefficient, +1 the upper; a symmetric region's lower
    // entries name its upper), interface-side faces in interface order, side
    // a then side b
    std::vector<int32_t> faceOffset;         // per region
    std::vector<int32_t> ellSrc;             // per ELL entry, -1 for padding
    std::vector<int32_t> fSrc, bSrc;         // per DILU slot (as fVal / bVal), -1 for none
    std::vector<int32_t> ifSrc;              // per interface-row face (as ifCoeff)
};

// Build the layout.  `preferLevels` forces the level schedule for every
// region (used to test the general path on structured meshes).
SystemLayout buildLayout(const CoupledSystem &sys, bool preferLevels = false);

// Detect a lexicographically numbered structured hex box from LDU
// addressing; returns false if the region is not one.
bool detectBox(const LduRegion &R, int n[3]);

}  // namespace mrfg
