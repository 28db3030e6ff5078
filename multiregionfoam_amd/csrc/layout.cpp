#include "layout.hpp"

#include <algorithm>
#include <numeric>
#include <stdexcept>

namespace mrfg {

namespace {

struct RowDeps {
    // CSR of lower / upper neighbours per global cell, in face order
    std::vector<int32_t> lo_off, lo_nb, lo_face;
    std::vector<int32_t> up_off, up_nb, up_face;
};

// One DILU wave under construction: cells[step * 64 + lane] (global cell or -1)
struct WaveBuild {
    int key = 0;       // dispatch key: tiles on one anti-diagonal / one level share it
    int region = 0;
    int steps = 0;
    std::vector<int32_t> cells;
    // structured-box tile (box = false for level waves)
    bool box = false;
    int tb = 0, tc = 0, TB = 0, nA = 0, seqAxis = 0, actB = 0, actC = 0;
};

void tangential(int a, int &b, int &c)
{
    if (a == 0) { b = 1; c = 2; }
    else if (a == 1) { b = 0; c = 2; }
    else { b = 0; c = 1; }
}

// Model of one sweep's time (us) for a tiling: the dependency chain of a
// tile (its steps plus one cross-tile hop per tile boundary crossed) against
// the bandwidth time of all slots.
double tileCost(const int n[3], int a, int TB, int TC)
{
    int b, c;
    tangential(a, b, c);
    int nTB = (n[b] + TB - 1) / TB, nTC = (n[c] + TC - 1) / TC;
    int steps = n[a] + std::min(TB, n[b]) - 1 + std::min(TC, n[c]) - 1;
    double slots = (double)nTB * nTC * steps * 64.0;
    double chain = steps * 0.06 + (nTB + nTC - 2) * 2.0;
    double bw = slots * 56.0 / 4.0e6;
    return std::max(chain, bw) + 0.05 * bw;
}

void boxWaves(const int n[3], int regionIdx, int cellOffset, RegionSchedule &rs, std::vector<WaveBuild> &out)
{
    int a = 0;
    for (int q = 1; q < 3; ++q)
        if (n[q] > n[a]) a = q;
    int b, c;
    tangential(a, b, c);
    int bestTB = 64;
    double best = 1e300;
    for (int TB = 1; TB <= 64; TB *= 2) {
        int TC = 64 / TB;
        if (TB > 1 && TB / 2 >= n[b]) continue;
        if (TC > 1 && TC / 2 >= n[c]) continue;
        double cst = tileCost(n, a, TB, TC);
        if (cst < best) {
            best = cst;
            bestTB = TB;
        }
    }
    const int TB = bestTB, TC = 64 / TB;
    rs.kind = "box";
    for (int q = 0; q < 3; ++q) rs.n[q] = n[q];
    rs.seqAxis = a;
    rs.tileB = TB;
    rs.tileC = TC;
    const int nTB = (n[b] + TB - 1) / TB, nTC = (n[c] + TC - 1) / TC;
    for (int tc = 0; tc < nTC; ++tc)
        for (int tb = 0; tb < nTB; ++tb) {
            WaveBuild w;
            w.key = tb + tc;
            w.region = regionIdx;
            int maxBl = std::min(TB, n[b] - tb * TB) - 1, maxCl = std::min(TC, n[c] - tc * TC) - 1;
            w.box = true;
            w.tb = tb;
            w.tc = tc;
            w.TB = TB;
            w.nA = n[a];
            w.seqAxis = a;
            w.actB = maxBl + 1;
            w.actC = maxCl + 1;
            w.steps = n[a] + maxBl + maxCl;
            w.cells.assign((size_t)w.steps * kSlice, -1);
            for (int cl = 0; cl <= maxCl; ++cl)
                for (int bl = 0; bl <= maxBl; ++bl) {
                    int lane = bl + TB * cl;
                    int idx[3];
                    idx[b] = tb * TB + bl;
                    idx[c] = tc * TC + cl;
                    for (int ia = 0; ia < n[a]; ++ia) {
                        idx[a] = ia;
                        int cell = idx[0] + n[0] * (idx[1] + n[1] * idx[2]);
                        w.cells[(size_t)(ia + bl + cl) * kSlice + lane] = cellOffset + cell;
                    }
                }
            rs.maxSteps = std::max(rs.maxSteps, w.steps);
            out.push_back(std::move(w));
            rs.nWaves++;
        }
}

void levelWaves(int regionIdx, int c0, int c1, const RowDeps &D, RegionSchedule &rs, std::vector<WaveBuild> &out)
{
    std::vector<int32_t> level(c1 - c0, 0);
    int maxLevel = 0;
    for (int c = c0; c < c1; ++c) {
        int lv = 0;
        for (int q = D.lo_off[c]; q < D.lo_off[c + 1]; ++q) lv = std::max(lv, level[D.lo_nb[q] - c0] + 1);
        level[c - c0] = lv;
        maxLevel = std::max(maxLevel, lv);
    }
    std::vector<std::vector<int32_t>> byLevel(maxLevel + 1);
    for (int c = c0; c < c1; ++c) byLevel[level[c - c0]].push_back(c);
    size_t maxWidth = 0;
    for (const auto &b : byLevel) maxWidth = std::max(maxWidth, b.size());
    if (maxWidth <= (size_t)kSlice) {
        // every level fits one slice (e.g. 2-D and multi-block meshes whose
        // levels are anti-diagonals of at most 64 cells): one wave walks the
        // levels as its steps, so a dependency on the previous levels comes
        // from the wave's own LDS ring instead of another wave's inbox
        rs.kind = "chain";
        WaveBuild w;
        w.key = 0;
        w.region = regionIdx;
        w.steps = maxLevel + 1;
        w.cells.assign((size_t)w.steps * kSlice, -1);
        for (int lv = 0; lv <= maxLevel; ++lv)
            for (size_t l = 0; l < byLevel[lv].size(); ++l) w.cells[(size_t)lv * kSlice + l] = byLevel[lv][l];
        out.push_back(std::move(w));
        rs.nWaves++;
        rs.maxSteps = maxLevel + 1;
        return;
    }
    rs.kind = "levels";
    for (int lv = 0; lv <= maxLevel; ++lv)
        for (size_t i = 0; i < byLevel[lv].size(); i += kSlice) {
            WaveBuild w;
            w.key = lv;
            w.region = regionIdx;
            w.steps = 1;
            w.cells.assign(kSlice, -1);
            for (size_t l = 0; l < (size_t)kSlice && i + l < byLevel[lv].size(); ++l) w.cells[l] = byLevel[lv][i + l];
            out.push_back(std::move(w));
            rs.nWaves++;
        }
    rs.maxSteps = 1;
}

}  // namespace

bool detectBox(const LduRegion &R, int n[3])
{
    const int N = R.nCells;
    if (N < 2 || R.nFaces < 1) return false;
    int d1 = 0, d2 = 0, d3 = 0;   // smallest distinct offsets
    for (int f = 0; f < R.nFaces; ++f) {
        int d = R.uAddr[f] - R.lAddr[f];
        if (d == d1 || d == d2 || d == d3) continue;
        if (!d1 || d < d1) { d3 = d2; d2 = d1; d1 = d; }
        else if (!d2 || d < d2) { d3 = d2; d2 = d; }
        else if (!d3 || d < d3) d3 = d;
    }
    int nx, ny, nz;
    if (d1 != 1) return false;
    if (!d2) { nx = N; ny = 1; nz = 1; }
    else if (!d3) { nx = d2; ny = N / d2; nz = 1; }
    else { nx = d2; if (d3 % d2) return false; ny = d3 / d2; nz = N / d3; }
    if ((long long)nx * ny * nz != N || nx < 1 || ny < 1 || nz < 1) return false;
    // verify the face list is exactly the lexicographic box pattern
    int f = 0;
    for (int c = 0; c < N; ++c) {
        int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
        int nb[3] = {i + 1 < nx ? c + 1 : -1, j + 1 < ny ? c + nx : -1, k + 1 < nz ? c + nx * ny : -1};
        for (int q = 0; q < 3; ++q) {
            if (nb[q] < 0) continue;
            if (f >= R.nFaces || R.lAddr[f] != c || R.uAddr[f] != nb[q]) return false;
            ++f;
        }
    }
    if (f != R.nFaces) return false;
    n[0] = nx;
    n[1] = ny;
    n[2] = nz;
    return true;
}

SystemLayout buildLayout(const CoupledSystem &sys, bool preferLevels)
{
    sys.check();
    SystemLayout L;
    const int nR = (int)sys.regions.size();
    L.regionOffset.resize(nR + 1, 0);
    for (int r = 0; r < nR; ++r) L.regionOffset[r + 1] = L.regionOffset[r] + sys.regions[r].nCells;
    const int N = L.regionOffset[nR];
    L.nCells = N;
    L.nFaces = sys.nFaces();
    L.nPairs = sys.nPairs();

    // ---- lower/upper neighbour lists per cell (face order) ----
    std::vector<int32_t> nLo(N, 0), nUp(N, 0);
    for (int r = 0; r < nR; ++r) {
        const LduRegion &R = sys.regions[r];
        int o = L.regionOffset[r];
        for (int f = 0; f < R.nFaces; ++f) {
            nLo[o + R.uAddr[f]]++;
            nUp[o + R.lAddr[f]]++;
        }
    }
    RowDeps D;
    D.lo_off.assign(N + 1, 0);
    D.up_off.assign(N + 1, 0);
    for (int c = 0; c < N; ++c) {
        D.lo_off[c + 1] = D.lo_off[c] + nLo[c];
        D.up_off[c + 1] = D.up_off[c] + nUp[c];
    }
    D.lo_nb.resize(D.lo_off[N]);
    D.lo_face.resize(D.lo_off[N]);
    D.up_nb.resize(D.up_off[N]);
    D.up_face.resize(D.up_off[N]);
    {
        std::vector<int32_t> fl(D.lo_off.begin(), D.lo_off.end() - 1), fu(D.up_off.begin(), D.up_off.end() - 1);
        for (int r = 0; r < nR; ++r) {
            const LduRegion &R = sys.regions[r];
            int o = L.regionOffset[r];
            for (int f = 0; f < R.nFaces; ++f) {
                int l = o + R.lAddr[f], u = o + R.uAddr[f];
                D.lo_nb[fl[u]] = l;
                D.lo_face[fl[u]++] = f;
                D.up_nb[fu[l]] = u;
                D.up_face[fu[l]++] = f;
            }
        }
    }
    int K = 3;
    for (int c = 0; c < N; ++c) K = std::max(K, std::max(nLo[c], nUp[c]));
    L.K = K;

    // ---- DILU schedule: waves of every region, ordered by dispatch key ----
    std::vector<WaveBuild> waves;
    L.schedules.resize(nR);
    for (int r = 0; r < nR; ++r) {
        int n[3];
        if (!preferLevels && detectBox(sys.regions[r], n))
            boxWaves(n, r, L.regionOffset[r], L.schedules[r], waves);
        else
            levelWaves(r, L.regionOffset[r], L.regionOffset[r + 1], D, L.schedules[r], waves);
    }
    std::vector<int> ord(waves.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return waves[x].key < waves[y].key; });
    L.nWaves = (int)waves.size();
    L.waveSlice.resize(L.nWaves);
    L.waveSteps.resize(L.nWaves);
    L.waveKey.resize(L.nWaves);
    int slice = 0;
    for (int wi = 0; wi < L.nWaves; ++wi) {
        const WaveBuild &w = waves[ord[wi]];
        L.waveSlice[wi] = slice;
        L.waveSteps[wi] = w.steps;
        L.waveKey[wi] = w.key;
        slice += w.steps;
    }
    L.nSlices = slice;
    L.nPos = slice * kSlice;
    L.cellOfPos.assign(L.nPos, -1);
    L.posOfCell.assign(N, -1);
    std::vector<int32_t> waveOfSlice(L.nSlices);
    for (int wi = 0; wi < L.nWaves; ++wi) {
        const WaveBuild &w = waves[ord[wi]];
        for (int t = 0; t < w.steps; ++t) {
            waveOfSlice[L.waveSlice[wi] + t] = wi;
            for (int l = 0; l < kSlice; ++l) {
                int c = w.cells[(size_t)t * kSlice + l];
                if (c < 0) continue;
                int p = (L.waveSlice[wi] + t) * kSlice + l;
                L.cellOfPos[p] = c;
                L.posOfCell[c] = p;
            }
        }
    }
    for (int c = 0; c < N; ++c)
        if (L.posOfCell[c] < 0) throw std::runtime_error("DILU schedule does not cover every cell");

    // structured-box waves: tile neighbours and fixed-slot records (slot =
    // 2 - axis: z, y, x, the face order of a cell's lower neighbours and the
    // reverse face order of its upper neighbours)
    L.waveBox.assign(L.nWaves, 0);
    L.boxMeta.assign((size_t)L.nWaves * kBoxMeta, -1);
    {
        std::vector<std::vector<int>> tileWave(nR);
        std::vector<int> tileNB(nR, 0);
        for (int wi = 0; wi < L.nWaves; ++wi) {
            const WaveBuild &w = waves[ord[wi]];
            if (!w.box) continue;
            const RegionSchedule &rs = L.schedules[w.region];
            int b, cc;
            tangential(rs.seqAxis, b, cc);
            int nTB = (rs.n[b] + rs.tileB - 1) / rs.tileB, nTC = (rs.n[cc] + rs.tileC - 1) / rs.tileC;
            if (tileWave[w.region].empty()) tileWave[w.region].assign((size_t)nTB * nTC, -1);
            tileNB[w.region] = nTB;
            tileWave[w.region][(size_t)w.tc * nTB + w.tb] = wi;
        }
        for (int wi = 0; wi < L.nWaves; ++wi) {
            const WaveBuild &w = waves[ord[wi]];
            if (!w.box) continue;
            L.waveBox[wi] = 1;
            const RegionSchedule &rs = L.schedules[w.region];
            int b, cc;
            tangential(rs.seqAxis, b, cc);
            const int nTB = tileNB[w.region], nTC = (rs.n[cc] + rs.tileC - 1) / rs.tileC;
            auto tw = [&](int tb, int tc) {
                return tb < 0 || tc < 0 || tb >= nTB || tc >= nTC ? -1 : tileWave[w.region][(size_t)tc * nTB + tb];
            };
            int32_t *m = &L.boxMeta[(size_t)wi * kBoxMeta];
            m[kBmTB] = w.TB;
            m[kBmNA] = w.nA;
            m[kBmPrevB] = tw(w.tb - 1, w.tc);
            m[kBmPrevC] = tw(w.tb, w.tc - 1);
            m[kBmNextB] = tw(w.tb + 1, w.tc);
            m[kBmNextC] = tw(w.tb, w.tc + 1);
            m[kBmAxis] = rs.seqAxis;
            m[kBmActB] = w.actB;
            m[kBmActC] = w.actC;
        }
    }

    auto regionOf = [&](int c) {
        return int(std::upper_bound(L.regionOffset.begin(), L.regionOffset.end(), c) - L.regionOffset.begin()) - 1;
    };

    L.diag.assign(L.nPos, 0.0);
    L.source.assign(L.nPos, 0.0);
    L.psi0.assign(L.nPos, 0.0);
    L.sumA.assign(L.nPos, 0.0);
    for (int r = 0; r < nR; ++r) {
        const LduRegion &R = sys.regions[r];
        int o = L.regionOffset[r];
        for (int c = 0; c < R.nCells; ++c) {
            int p = L.posOfCell[o + c];
            L.diag[p] = R.diag[c];
            L.source[p] = R.source[c];
            L.psi0[p] = R.psi[c];
        }
    }
    // sumA: diag, then the face loop, then minus interface coefficients
    // (lduMatrix::sumA restated on the coupled system)
    {
        std::vector<double> s(N);
        for (int r = 0; r < nR; ++r) {
            const LduRegion &R = sys.regions[r];
            int o = L.regionOffset[r];
            for (int c = 0; c < R.nCells; ++c) s[o + c] = R.diag[c];
            for (int f = 0; f < R.nFaces; ++f) {
                s[o + R.uAddr[f]] += R.lower[f];
                s[o + R.lAddr[f]] += R.upper[f];
            }
        }
        for (const auto &I : sys.interfaces) {
            const InterfaceSide *sides[2] = {&I.a, &I.b};
            for (int q = 0; q < 2; ++q) {
                const InterfaceSide &S = *sides[q];
                int o = L.regionOffset[S.region];
                for (int f = 0; f < S.nFaces(); ++f) s[o + S.fc[f]] -= S.coeff[f];
            }
        }
        for (int c = 0; c < N; ++c) L.sumA[L.posOfCell[c]] = s[c];
    }

    // ---- interface faces per row (interface order, then face order) ----
    std::vector<std::vector<std::pair<int, int>>> rowFaces(N);
    for (size_t i = 0; i < sys.interfaces.size(); ++i) {
        const CoupledInterface &I = sys.interfaces[i];
        const InterfaceSide *sides[2] = {&I.a, &I.b};
        for (int q = 0; q < 2; ++q) {
            const InterfaceSide &S = *sides[q];
            int o = L.regionOffset[S.region];
            for (int f = 0; f < S.nFaces(); ++f) rowFaces[o + S.fc[f]].push_back({(int)i * 2 + q, f});
        }
    }

    // ---- per-slice ELL (SpMV) ----
    L.ellOff.assign(L.nSlices + 1, 0);
    L.ifMask.assign(L.nSlices, 0);
    L.ifBase.assign(L.nSlices, 0);
    L.ifRowFace.assign(1, 0);
    L.ifPairOff.assign(1, 0);
    for (int s = 0; s < L.nSlices; ++s) {
        int wE = 0;
        for (int l = 0; l < kSlice; ++l) {
            int c = L.cellOfPos[s * kSlice + l];
            if (c < 0) continue;
            wE = std::max(wE, (D.lo_off[c + 1] - D.lo_off[c]) + (D.up_off[c + 1] - D.up_off[c]));
        }
        L.ellOff[s + 1] = L.ellOff[s] + wE;
    }
    L.ellCol.assign((size_t)L.ellOff[L.nSlices] * kSlice, 0);
    L.ellVal.assign((size_t)L.ellOff[L.nSlices] * kSlice, 0.0);
    L.ellSrc.assign((size_t)L.ellOff[L.nSlices] * kSlice, -1);
    L.faceOffset.assign(nR, 0);
    for (int r = 1; r < nR; ++r) L.faceOffset[r] = L.faceOffset[r - 1] + sys.regions[r - 1].nFaces;
    std::vector<int32_t> sideOffset(2 * sys.interfaces.size() + 1, 0);
    for (size_t i = 0; i < sys.interfaces.size(); ++i) {
        sideOffset[2 * i + 1] = sideOffset[2 * i] + sys.interfaces[i].a.nFaces();
        sideOffset[2 * i + 2] = sideOffset[2 * i + 1] + sys.interfaces[i].b.nFaces();
    }
    const size_t nDep = (size_t)L.nSlices * K * kSlice;
    L.fCode.assign(nDep, kNoDep);
    L.bCode.assign(nDep, kNoDep);
    L.fVal.assign(nDep, 0.0);
    L.fProd.assign(nDep, 0.0);
    L.bVal.assign(nDep, 0.0);
    L.fSrc.assign(nDep, -1);
    L.bSrc.assign(nDep, -1);
    L.fHalo.assign(L.nPos, -1);
    L.bHalo.assign(L.nPos, -1);

    auto waveStep = [&](int p, int &w, int &t, int &lane) {
        int s = p / kSlice;
        lane = p % kSlice;
        w = waveOfSlice[s];
        t = s - L.waveSlice[w];
    };

    for (int s = 0; s < L.nSlices; ++s) {
        L.ifBase[s] = (int32_t)L.ifRowFace.size() - 1;
        const int wE = L.ellOff[s + 1] - L.ellOff[s];
        for (int l = 0; l < kSlice; ++l) {
            int p = s * kSlice + l;
            int c = L.cellOfPos[p];
            // padding entries of the SpMV point at the row itself with value 0
            for (int k = 0; k < wE; ++k) L.ellCol[(size_t)(L.ellOff[s] + k) * kSlice + l] = p;
            if (c < 0) continue;
            int r = regionOf(c);
            const LduRegion &R = sys.regions[r];
            const std::vector<double> &lo = R.symmetric ? R.upper : R.lower;
            auto loSrc = [&](int f) { return 2 * (L.faceOffset[r] + f) + (R.symmetric ? 1 : 0); };
            auto upSrc = [&](int f) { return 2 * (L.faceOffset[r] + f) + 1; };
            int k = 0;
            for (int q = D.lo_off[c]; q < D.lo_off[c + 1]; ++q, ++k) {
                size_t e = (size_t)(L.ellOff[s] + k) * kSlice + l;
                L.ellCol[e] = L.posOfCell[D.lo_nb[q]];
                L.ellVal[e] = lo[D.lo_face[q]];
                L.ellSrc[e] = loSrc(D.lo_face[q]);
            }
            for (int q = D.up_off[c]; q < D.up_off[c + 1]; ++q, ++k) {
                size_t e = (size_t)(L.ellOff[s] + k) * kSlice + l;
                L.ellCol[e] = L.posOfCell[D.up_nb[q]];
                L.ellVal[e] = R.upper[D.up_face[q]];
                L.ellSrc[e] = upSrc(D.up_face[q]);
            }

            int w, t, lane;
            waveStep(p, w, t, lane);
            const int T = L.waveSteps[w];
            if (L.waveBox[w]) {
                // box tile rows: fixed slots by the axis of the neighbour
                // (slot 2 - axis), zero where there is none; the dependencies
                // themselves follow from the tile geometry
                const RegionSchedule &rs = L.schedules[r];
                const int sxy = rs.n[0] * rs.n[1];
                auto slotOf = [&](int d) { return d == 1 ? 2 : d == rs.n[0] ? 1 : d == sxy ? 0 : -1; };
                for (int q = D.lo_off[c]; q < D.lo_off[c + 1]; ++q) {
                    int sl = slotOf(c - D.lo_nb[q]);
                    if (sl < 0) throw std::runtime_error("box region with an unexpected lower neighbour");
                    int f = D.lo_face[q];
                    size_t e = ((size_t)s * K + sl) * kSlice + l;
                    L.fVal[e] = lo[f];
                    L.fProd[e] = R.upper[f] * lo[f];
                    L.fSrc[e] = loSrc(f);
                }
                for (int q = D.up_off[c]; q < D.up_off[c + 1]; ++q) {
                    int sl = slotOf(D.up_nb[q] - c);
                    if (sl < 0) throw std::runtime_error("box region with an unexpected upper neighbour");
                    L.bVal[((size_t)s * K + sl) * kSlice + l] = R.upper[D.up_face[q]];
                    L.bSrc[((size_t)s * K + sl) * kSlice + l] = upSrc(D.up_face[q]);
                }
            } else {
                // forward sweep: lower neighbours, ascending face order
                k = 0;
                for (int q = D.lo_off[c]; q < D.lo_off[c + 1]; ++q, ++k) {
                    int f = D.lo_face[q];
                    int pd = L.posOfCell[D.lo_nb[q]];
                    int wd, td, ld;
                    waveStep(pd, wd, td, ld);
                    size_t e = ((size_t)s * K + k) * kSlice + l;
                    L.fVal[e] = lo[f];
                    L.fProd[e] = R.upper[f] * lo[f];
                    L.fSrc[e] = loSrc(f);
                    if (wd == w && t - td >= 1 && t - td <= kRing - 1) {
                        L.fCode[e] = -((td % kRing) * kSlice + ld + 1);
                    } else {
                        if (wd == w ? td >= t : wd > w)
                            throw std::runtime_error("DILU schedule violates a forward dependency");
                        if (L.fHalo[pd] < 0) L.fHalo[pd] = L.nHaloF++;
                        L.fCode[e] = L.fHalo[pd];
                    }
                }
                // backward sweep: upper neighbours, reverse face order; processing
                // index i = T-1-t, so the ring slot of a dependency uses its own i
                k = 0;
                for (int q = D.up_off[c + 1] - 1; q >= D.up_off[c]; --q, ++k) {
                    int pd = L.posOfCell[D.up_nb[q]];
                    int wd, td, ld;
                    waveStep(pd, wd, td, ld);
                    size_t e = ((size_t)s * K + k) * kSlice + l;
                    L.bVal[e] = R.upper[D.up_face[q]];
                    L.bSrc[e] = upSrc(D.up_face[q]);
                    if (wd == w && td - t >= 1 && td - t <= kRing - 1) {
                        int id = T - 1 - td;
                        L.bCode[e] = -((id % kRing) * kSlice + ld + 1);
                    } else {
                        if (wd == w ? td <= t : wd < w)
                            throw std::runtime_error("DILU schedule violates a backward dependency");
                        if (L.bHalo[pd] < 0) L.bHalo[pd] = L.nHaloB++;
                        L.bCode[e] = L.bHalo[pd];
                    }
                }
            }
            // interface faces of this row
            if (!rowFaces[c].empty()) {
                L.ifMask[s] |= 1ULL << l;
                for (auto &pf : rowFaces[c]) {
                    const CoupledInterface &I = sys.interfaces[pf.first / 2];
                    const InterfaceSide &S = (pf.first % 2) ? I.b : I.a;
                    const InterfaceSide &O = (pf.first % 2) ? I.a : I.b;
                    int oN = L.regionOffset[O.region];
                    int f = pf.second;
                    L.ifCoeff.push_back(S.coeff[f]);
                    L.ifSrc.push_back(sideOffset[pf.first] + f);
                    for (int q = S.off[f]; q < S.off[f + 1]; ++q) {
                        L.ifPairCol.push_back(L.posOfCell[oN + O.fc[S.idx[q]]]);
                        L.ifPairW.push_back(S.w[q]);
                    }
                    L.ifPairOff.push_back((int32_t)L.ifPairW.size());
                }
                L.ifRowFace.push_back((int32_t)L.ifCoeff.size());
            }
        }
    }
    return L;
}

}  // namespace mrfg
