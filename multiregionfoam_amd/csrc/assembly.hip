// Finite-volume assembly on the device.  Every expression is evaluated in
// the order and grouping of the host reader's loops (fvm::laplacian,
// upwind fvm::div, negSumDiag, addBoundaryDiag / addBoundarySource,
// calcThermalDiffusivity, the Euler ddt term) with FP contraction off, so
// the device matrices equal the host-assembled ones bit for bit.
//
// Three passes over the model (faces, interface-side faces, cells), each
// one thread per item and independent of the others within a pass; a cell
// reads its faces' coefficients in ascending face order, which is the order
// in which the sequential negSumDiag loop updates that cell.  A fourth pass
// puts the values into the solver's position-ordered arrays.
#include "assembly.hpp"

namespace mrfg {

namespace {

constexpr int kAsmBlock = 256;

inline int gridFor(int n) { return (n + kAsmBlock - 1) / kAsmBlock; }

// internal faces: Laplacian coefficient k_f |S_f| |d|^-1 with k_f linearly
// interpolated; upwind convection split of phi = (w U_P + (1-w) U_N) . S_f;
// the region equation's lower / upper.
__global__ __launch_bounds__(kAsmBlock) void k_fv_faces(FvDev m)
{
    const int f = blockIdx.x * kAsmBlock + threadIdx.x;
    if (f >= m.nFaces) return;
    const FvRegionParams P = m.region[m.faceRegion[f]];
    const int o = m.own[f], n = m.nei[f];
    const double w = m.w[f];
    const double kf = w * m.k[o] + (1.0 - w) * m.k[n];
    const double lap = m.dc[f] * (kf * m.magSf[f]);
    double lower, upper, alo = 0.0, aup = 0.0;
    if (P.transport) {
        const double *Uo = m.U + 3 * (size_t)o, *Un = m.U + 3 * (size_t)n, *S = m.Sf + 3 * (size_t)f;
        const double u0 = w * Uo[0] + (1.0 - w) * Un[0];
        const double u1 = w * Uo[1] + (1.0 - w) * Un[1];
        const double u2 = w * Uo[2] + (1.0 - w) * Un[2];
        const double phi = u0 * S[0] + u1 * S[1] + u2 * S[2];
        const double wu = phi >= 0.0 ? 1.0 : 0.0;
        alo = -wu * phi;
        aup = alo + phi;
        upper = P.rhoCp * aup - lap;
        lower = P.rhoCp * alo - lap;
    } else {
        upper = 0.0 - lap;
        lower = upper;
    }
    m.lapU[f] = lap;
    m.aLo[f] = alo;
    m.aUp[f] = aup;
    m.faceVal[2 * (size_t)f] = lower;
    m.faceVal[2 * (size_t)f + 1] = upper;
}

// interface-side faces: calcThermalDiffusivity (radiation off) with the
// partner side's distance and conductivity gathered through the GGI weights
__global__ __launch_bounds__(kAsmBlock) void k_fv_sides(FvDev m)
{
    const int i = blockIdx.x * kAsmBlock + threadIdx.x;
    if (i >= m.nSideFaces) return;
    const double dOwn = m.sNDist[i];
    double dNei = 0.0, fNei = 0.0;
    for (int p = m.sOff[i]; p < m.sOff[i + 1]; ++p) {
        dNei += m.sW[p] * m.sPairNDist[p];
        fNei += m.sW[p] * m.k[m.sPairCell[p]];
    }
    const double fOwn = m.k[m.sCell[i]];
    const double magDelta = dOwn + dNei;
    const double wP = dNei / magDelta;
    const double delta = 1.0 / magDelta;
    const double kOwn = fOwn / (1.0 - wP) / magDelta;
    const double kNei = fNei / wP / magDelta;
    m.sKOwn[i] = kOwn;
    m.sKNei[i] = kNei;
    double kk = kOwn * kNei;
    kk /= delta * (kOwn + kNei);
    const double kHarm = 1.0 / (wP / fOwn + (1.0 - wP) / fNei);
    m.sK[i] = fmax(fmin(kk, 100.0 * kHarm), 0.01 * kHarm);
    m.sDelta[i] = delta;
}

// cells: negSumDiag of both operators, the region equation's diagonal and
// source, the boundary faces' internal / boundary coefficients (coupled
// faces hand theirs to the interface), then the ddt term
__global__ __launch_bounds__(kAsmBlock) void k_fv_cells(FvDev m)
{
    const int c = blockIdx.x * kAsmBlock + threadIdx.x;
    if (c >= m.nCells) return;
    const FvRegionParams P = m.region[m.cellRegion[c]];
    double lapD = 0.0, aD = 0.0;
    for (int j = m.cfOff[c]; j < m.cfOff[c + 1]; ++j) {
        const int e = m.cf[j], f = e >> 1;
        lapD -= m.lapU[f];
        aD -= (e & 1) ? m.aUp[f] : m.aLo[f];
    }
    double d, src;
    if (P.transport) {
        d = P.rhoCp * aD - lapD;
        src = P.rhoCp * 0.0 - 0.0;
    } else {
        d = 0.0 - lapD;
        src = 0.0 - 0.0;
    }
    const double kc = m.k[c];
    const double *Uc = m.U + 3 * (size_t)c;
    for (int b = m.cbOff[c]; b < m.cbOff[c + 1]; ++b) {
        const int bt = m.bT[b];
        double lic = 0.0, lbc = 0.0;
        if (bt == 0) {            // fixedValue
            const double deltab = 1.0 / m.bNDist[b];
            const double g = (m.bKFixed[b] ? m.bKb[b] : kc) * m.bMagSf[b];
            lic = g * (-deltab);
            lbc = -g * (deltab * m.bTb[b]);
        } else if (bt == 3) {     // monolithic interface
            const int i = m.bSide[b];
            const double g = m.sK[i] * m.bMagSf[b];
            const double delta = m.sDelta[i];
            lic = g * (-delta);
            lbc = -g * delta;
        }
        double eic, ebc;
        if (P.transport) {
            const double *S = m.bSf + 3 * (size_t)b;
            double ub0 = 0.0, ub1 = 0.0, ub2 = 0.0;
            switch (m.bU[b]) {
            case 0:
                ub0 = m.bUb[3 * (size_t)b];
                ub1 = m.bUb[3 * (size_t)b + 1];
                ub2 = m.bUb[3 * (size_t)b + 2];
                break;
            case 1:
                ub0 = Uc[0];
                ub1 = Uc[1];
                ub2 = Uc[2];
                break;
            case 2: {
                const double r = 1.0 / m.bMagSf[b];
                const double n0 = r * S[0], n1 = r * S[1], n2 = r * S[2];
                const double t = Uc[0] * n0 + Uc[1] * n1 + Uc[2] * n2;
                ub0 = Uc[0] - t * n0;
                ub1 = Uc[1] - t * n1;
                ub2 = Uc[2] - t * n2;
                break;
            }
            default:
                break;
            }
            const double ph = ub0 * S[0] + ub1 * S[1] + ub2 * S[2];
            double aic = 0.0, abc = 0.0;
            if (bt == 0) {
                aic = ph * 0.0;
                abc = -ph * m.bTb[b];
            } else if (bt == 3) {
                const double wu = ph >= 0.0 ? 1.0 : 0.0;
                aic = ph * wu;
                abc = -ph * (1.0 - wu);
            } else if (bt == 1) {
                aic = ph * 1.0;
                abc = -ph * 0.0;
            }
            eic = P.rhoCp * aic - lic;
            ebc = P.rhoCp * abc - lbc;
        } else {
            eic = 0.0 - lic;
            ebc = 0.0 - lbc;
        }
        d += eic;
        if (bt == 3) m.ifCoeff[m.bSide[b]] = ebc;
        else src += ebc;
    }
    double a = 0.0;
    m.diagBase[c] = d;
    m.srcBase[c] = src;
    if (P.rDeltaT > 0.0) {
        a = P.rDeltaT * P.rhoCp * m.V[c];
        d = d + a;
        src = src + a * m.T[c];
    }
    m.ddtA[c] = a;
    m.diag[c] = d;
    m.source[c] = src;
}

// one thread per batch position: the row's ELL values, DILU record values
// and products, diagonal, source, interface coefficients and
// lduMatrix::sumA (the diagonal, the face coefficients in face order, minus
// the interface coefficients in interface order)
__global__ __launch_bounds__(kAsmBlock) void k_fv_place(FvPlace q)
{
    const int p = q.pos0 + blockIdx.x * kAsmBlock + threadIdx.x;
    if (p >= q.nPos) return;
    const int c = q.cellOfPos[p];
    if (c < 0) return;
    const int s = p >> 6, lane = p & 63;
    const double d = q.diag[c];
    double r = d;
    for (int k = q.ellOff[s]; k < q.ellOff[s + 1]; ++k) {
        const size_t e = (size_t)k * 64 + lane;
        const int src = q.ellSrc[e];
        const double v = src >= 0 ? q.faceVal[src] : 0.0;
        q.ellVal[e] = v;
        r += v;
    }
    for (int k = 0; k < q.K; ++k) {
        const size_t e = ((size_t)s * q.K + k) * 64 + lane;
        const size_t e2 = (((size_t)s * q.G + k / 2) * 64 + lane) * 2 + k % 2;
        const int fs = q.fSrc[e], bs = q.bSrc[e];
        q.fPack[e2] = fs >= 0 ? q.faceVal[fs] : 0.0;
        q.prodPack[e2] = fs >= 0 ? q.faceVal[fs | 1] * q.faceVal[fs] : 0.0;
        q.bPack[e2] = bs >= 0 ? q.faceVal[bs] : 0.0;
    }
    q.prodPack[(((size_t)s * q.G + q.K / 2) * 64 + lane) * 2 + q.K % 2] = d;
    const unsigned long long msk = q.ifMask[s];
    if ((msk >> lane) & 1ull) {
        const int row = q.ifBase[s] + __popcll(msk & ((1ull << lane) - 1ull));
        for (int f = q.ifRowFace[row]; f < q.ifRowFace[row + 1]; ++f) {
            const double v = q.sideCoeff[q.ifSrc[f]];
            q.ifCoeff[f] = v;
            r -= v;
        }
    }
    q.diagOut[p] = d;
    q.b[p] = q.source[c];
    q.sumA[p] = r;
    if (q.ddtAOut) {
        q.ddtAOut[p] = q.ddtA[c];
        q.diagBaseOut[p] = q.diagBase[c];
        q.srcBaseOut[p] = q.srcBase[c];
    }
}

}  // namespace

hipError_t launchFvAssemble(const FvDev &m, hipStream_t st)
{
    if (m.nFaces > 0) k_fv_faces<<<gridFor(m.nFaces), kAsmBlock, 0, st>>>(m);
    if (m.nSideFaces > 0) k_fv_sides<<<gridFor(m.nSideFaces), kAsmBlock, 0, st>>>(m);
    if (m.nCells > 0) k_fv_cells<<<gridFor(m.nCells), kAsmBlock, 0, st>>>(m);
    return hipGetLastError();
}

hipError_t launchFvPlace(const FvPlace &q, hipStream_t st)
{
    if (q.nPos > q.pos0) k_fv_place<<<gridFor(q.nPos - q.pos0), kAsmBlock, 0, st>>>(q);
    return hipGetLastError();
}

}  // namespace mrfg
