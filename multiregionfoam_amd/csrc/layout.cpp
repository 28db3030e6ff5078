#include "layout.hpp"

#include <algorithm>
#include <stdexcept>

namespace mrfg {

namespace {

struct RowDeps {
    // CSR of lower / upper neighbours per global cell, in face order
    std::vector<int32_t> lo_off, lo_nb, lo_face;
    std::vector<int32_t> up_off, up_nb, up_face;
};

}  // namespace

SystemLayout buildLayout(const CoupledSystem &sys)
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

    // ---- levels (merged over regions) ----
    std::vector<int32_t> level(N, 0);
    int maxLevel = 0;
    for (int c = 0; c < N; ++c) {
        int lv = 0;
        for (int q = D.lo_off[c]; q < D.lo_off[c + 1]; ++q) lv = std::max(lv, level[D.lo_nb[q]] + 1);
        level[c] = lv;
        maxLevel = std::max(maxLevel, lv);
    }
    L.nLevels = N ? maxLevel + 1 : 0;
    std::vector<int32_t> perLevel(L.nLevels + 1, 0);
    for (int c = 0; c < N; ++c) perLevel[level[c] + 1]++;
    L.levelSliceBegin.assign(L.nLevels + 1, 0);
    std::vector<int32_t> levelPosBegin(L.nLevels + 1, 0);
    for (int lv = 0; lv < L.nLevels; ++lv) {
        int slices = (perLevel[lv + 1] + kSlice - 1) / kSlice;
        L.levelSliceBegin[lv + 1] = L.levelSliceBegin[lv] + slices;
        levelPosBegin[lv + 1] = L.levelSliceBegin[lv + 1] * kSlice;
    }
    L.nSlices = L.levelSliceBegin[L.nLevels];
    L.nPos = L.nSlices * kSlice;
    L.cellOfPos.assign(L.nPos, -1);
    L.posOfCell.resize(N);
    {
        std::vector<int32_t> fill(levelPosBegin.begin(), levelPosBegin.end() - 1);
        for (int c = 0; c < N; ++c) {
            int p = fill[level[c]]++;
            L.posOfCell[c] = p;
            L.cellOfPos[p] = c;
        }
    }

    // per-position coefficient accessors
    auto regionOf = [&](int c) {
        int r = int(std::upper_bound(L.regionOffset.begin(), L.regionOffset.end(), c) - L.regionOffset.begin()) - 1;
        return r;
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
    std::vector<std::vector<std::pair<int, int>>> rowFaces(N);   // (iface*2+side, face)
    for (size_t i = 0; i < sys.interfaces.size(); ++i) {
        const CoupledInterface &I = sys.interfaces[i];
        const InterfaceSide *sides[2] = {&I.a, &I.b};
        for (int q = 0; q < 2; ++q) {
            const InterfaceSide &S = *sides[q];
            int o = L.regionOffset[S.region];
            for (int f = 0; f < S.nFaces(); ++f) rowFaces[o + S.fc[f]].push_back({(int)i * 2 + q, f});
        }
    }

    // ---- per-slice ELL (SpMV) and DILU arrays ----
    L.ellOff.assign(L.nSlices + 1, 0);
    L.lOff.assign(L.nSlices + 1, 0);
    L.uOff.assign(L.nSlices + 1, 0);
    L.ifMask.assign(L.nSlices, 0);
    L.ifBase.assign(L.nSlices, 0);
    L.ifRowFace.assign(1, 0);
    L.ifPairOff.assign(1, 0);
    for (int s = 0; s < L.nSlices; ++s) {
        int wE = 0, wL = 0, wU = 0;
        for (int l = 0; l < kSlice; ++l) {
            int c = L.cellOfPos[s * kSlice + l];
            if (c < 0) continue;
            int nl = D.lo_off[c + 1] - D.lo_off[c], nu = D.up_off[c + 1] - D.up_off[c];
            wE = std::max(wE, nl + nu);
            wL = std::max(wL, nl);
            wU = std::max(wU, nu);
        }
        L.ellOff[s + 1] = L.ellOff[s] + wE;
        L.lOff[s + 1] = L.lOff[s] + wL;
        L.uOff[s + 1] = L.uOff[s] + wU;
    }
    L.ellCol.assign((size_t)L.ellOff[L.nSlices] * kSlice, 0);
    L.ellVal.assign((size_t)L.ellOff[L.nSlices] * kSlice, 0.0);
    L.lCol.assign((size_t)L.lOff[L.nSlices] * kSlice, -1);
    L.lVal.assign((size_t)L.lOff[L.nSlices] * kSlice, 0.0);
    L.lProd.assign((size_t)L.lOff[L.nSlices] * kSlice, 0.0);
    L.uCol.assign((size_t)L.uOff[L.nSlices] * kSlice, -1);
    L.uVal.assign((size_t)L.uOff[L.nSlices] * kSlice, 0.0);

    for (int s = 0; s < L.nSlices; ++s) {
        L.ifBase[s] = (int32_t)L.ifRowFace.size() - 1;
        for (int l = 0; l < kSlice; ++l) {
            int p = s * kSlice + l;
            int c = L.cellOfPos[p];
            int wE = L.ellOff[s + 1] - L.ellOff[s];
            int wL = L.lOff[s + 1] - L.lOff[s], wU = L.uOff[s + 1] - L.uOff[s];
            // padding entries of the SpMV point at the row itself with value 0
            for (int k = 0; k < wE; ++k) L.ellCol[(size_t)(L.ellOff[s] + k) * kSlice + l] = p;
            if (c < 0) continue;
            int r = regionOf(c);
            const LduRegion &R = sys.regions[r];
            const std::vector<double> &lo = R.symmetric ? R.upper : R.lower;
            int k = 0;
            for (int q = D.lo_off[c]; q < D.lo_off[c + 1]; ++q, ++k) {
                size_t e = (size_t)(L.ellOff[s] + k) * kSlice + l;
                L.ellCol[e] = L.posOfCell[D.lo_nb[q]];
                L.ellVal[e] = lo[D.lo_face[q]];
            }
            for (int q = D.up_off[c]; q < D.up_off[c + 1]; ++q, ++k) {
                size_t e = (size_t)(L.ellOff[s] + k) * kSlice + l;
                L.ellCol[e] = L.posOfCell[D.up_nb[q]];
                L.ellVal[e] = R.upper[D.up_face[q]];
            }
            (void)wE;
            // DILU lower deps: ascending face order
            k = 0;
            for (int q = D.lo_off[c]; q < D.lo_off[c + 1]; ++q, ++k) {
                size_t e = (size_t)(L.lOff[s] + k) * kSlice + l;
                int f = D.lo_face[q];
                L.lCol[e] = L.posOfCell[D.lo_nb[q]];
                L.lVal[e] = lo[f];
                L.lProd[e] = R.upper[f] * lo[f];
            }
            (void)wL;
            // DILU upper deps: reverse face order
            k = 0;
            for (int q = D.up_off[c + 1] - 1; q >= D.up_off[c]; --q, ++k) {
                size_t e = (size_t)(L.uOff[s] + k) * kSlice + l;
                L.uCol[e] = L.posOfCell[D.up_nb[q]];
                L.uVal[e] = R.upper[D.up_face[q]];
            }
            (void)wU;
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
