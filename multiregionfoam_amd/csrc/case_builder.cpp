// Native case builder: structured hex box regions, FV assembly of the
// temperature equations and the monolithic interface coefficients.
//
// Mirrors the reference's region / boundary-condition classes:
//   conductTemperature::setCoupledEqns    (conductTemperature.C:135-152)
//   transportTemperature::setCoupledEqns  (transportTemperature.C:129-150)
//   monolithicThermalDiffusivity::calcThermalDiffusivity
//                          (monolithicThermalDiffusivityFvPatchScalarField.C:167-277)
//   monolithicCouplingFvPatchField / regionCouplePatch().interpolate
//                          (monolithicCouplingFvPatchField.C:207-464)
//   GGI weights, rescaled (ggiInterfaceToInterfaceMapping.C:62-77)
// and the fvMatrix arithmetic order of foam-extend's fvm::laplacian,
// fvm::div (Gauss upwind), negSumDiag, operator== and addBoundaryDiag /
// addBoundarySource(couples=false) before the coupled solve
// (multiRegionSystem.C:150-153).
#include "coupled_system.hpp"
#include "dictionary.hpp"

#include <array>
#include <cmath>
#include <map>
#include <stdexcept>

namespace mrfg {

double hashU01(uint64_t seed, uint64_t index)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (index + 1ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

namespace {

const char *kSides[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};

enum class PType { ZeroGradient, FixedValue, Empty, Coupled };
enum class FluxKind { Wall, Inlet, Outlet };

void tangential(int a, int &b, int &c)
{
    if (a == 0) { b = 1; c = 2; }
    else if (a == 1) { b = 0; c = 2; }
    else { b = 0; c = 1; }
}

// ---------------------------------------------------------------------------
// Structured hex box: cells i fastest, internal faces in upper-triangular
// order (owner ascending, neighbours c+1, c+nx, c+nx*ny), six side patches
// whose faces run first-tangential-axis fastest.
// ---------------------------------------------------------------------------
struct BoxFvMesh {
    std::string name;
    int n[3];
    double origin[3], length[3], d[3], area[3];
    int nCells = 0, nFaces = 0;
    std::vector<int32_t> owner, neighbour;
    std::vector<int8_t> faceAxis;
    std::array<PType, 6> ptype;

    void init()
    {
        for (int a = 0; a < 3; ++a) d[a] = length[a] / (double)n[a];
        area[0] = d[1] * d[2];
        area[1] = d[0] * d[2];
        area[2] = d[0] * d[1];
        nCells = n[0] * n[1] * n[2];
        owner.reserve((size_t)nCells * 3);
        for (int c = 0; c < nCells; ++c) {
            int i = c % n[0], j = (c / n[0]) % n[1], k = c / (n[0] * n[1]);
            int nb[3] = {i + 1 < n[0] ? c + 1 : -1, j + 1 < n[1] ? c + n[0] : -1,
                         k + 1 < n[2] ? c + n[0] * n[1] : -1};
            for (int a = 0; a < 3; ++a)
                if (nb[a] >= 0) {
                    owner.push_back(c);
                    neighbour.push_back(nb[a]);
                    faceAxis.push_back((int8_t)a);
                }
        }
        nFaces = (int)owner.size();
    }
    int cellIndex(int i, int j, int k) const { return i + n[0] * (j + n[1] * k); }
    void cellIJK(int c, int &i, int &j, int &k) const
    {
        i = c % n[0];
        j = (c / n[0]) % n[1];
        k = c / (n[0] * n[1]);
    }
    int sideFaces(int s) const
    {
        if (ptype[s] == PType::Empty) return 0;
        int b, c;
        tangential(s / 2, b, c);
        return n[b] * n[c];
    }
    int sideFaceCell(int s, int f, int idx[3]) const
    {
        int a = s / 2, b, c;
        tangential(a, b, c);
        idx[a] = (s % 2) ? n[a] - 1 : 0;
        idx[b] = f % n[b];
        idx[c] = f / n[b];
        return cellIndex(idx[0], idx[1], idx[2]);
    }
};

// ---------------------------------------------------------------------------
// fvMatrix<scalar> stand-in: LDU coefficients plus per-patch internal and
// boundary coefficients.
// ---------------------------------------------------------------------------
struct FvScalarMatrix {
    std::vector<double> diag, upper, lower, source;
    bool asymmetric = false;
    std::array<std::vector<double>, 6> internalCoeffs, boundaryCoeffs;

    void negSumDiag(const BoxFvMesh &m)
    {
        const std::vector<double> &lo = asymmetric ? lower : upper;
        for (int f = 0; f < m.nFaces; ++f) {
            diag[m.owner[f]] -= lo[f];
            diag[m.neighbour[f]] -= upper[f];
        }
    }
};

struct TemperaturePatches {
    std::array<PType, 6> type;
    std::array<double, 6> value;
};

// fvm::laplacian(k, T) with Gauss linear (uniform orthogonal hex, no
// non-orthogonal correction): upper = deltaCoeffs * (interpolate(k)*magSf),
// negSumDiag, internalCoeffs = gammaMagSf*gradientInternalCoeffs,
// boundaryCoeffs = -gammaMagSf*gradientBoundaryCoeffs.
FvScalarMatrix fvmLaplacian(const BoxFvMesh &m, const std::vector<double> &k,
                            const TemperaturePatches &T,
                            const std::array<std::vector<double>, 6> &coupledK,
                            const std::array<double, 6> &coupledDelta)
{
    FvScalarMatrix M;
    M.diag.assign(m.nCells, 0.0);
    M.upper.resize(m.nFaces);
    M.source.assign(m.nCells, 0.0);
    for (int f = 0; f < m.nFaces; ++f) {
        int a = m.faceAxis[f];
        double kf = 0.5 * k[m.owner[f]] + 0.5 * k[m.neighbour[f]];
        double gammaMagSf = kf * m.area[a];
        M.upper[f] = (1.0 / m.d[a]) * gammaMagSf;
    }
    M.negSumDiag(m);
    int idx[3];
    for (int s = 0; s < 6; ++s) {
        int nf = m.sideFaces(s);
        M.internalCoeffs[s].resize(nf);
        M.boundaryCoeffs[s].resize(nf);
        int a = s / 2;
        double magSf = m.area[a];
        for (int f = 0; f < nf; ++f) {
            int c = m.sideFaceCell(s, f, idx);
            double ic, bc;
            switch (T.type[s]) {
            case PType::FixedValue: {
                double deltab = 1.0 / (0.5 * m.d[a]);
                double gammaMagSf = k[c] * magSf;
                ic = gammaMagSf * (-deltab);
                bc = -gammaMagSf * (deltab * T.value[s]);
                break;
            }
            case PType::Coupled: {
                double gammaMagSf = coupledK[s][f] * magSf;
                double delta = coupledDelta[s];
                ic = gammaMagSf * (-delta);
                bc = -gammaMagSf * delta;
                break;
            }
            default: {
                double gammaMagSf = k[c] * magSf;
                ic = gammaMagSf * 0.0;
                bc = -gammaMagSf * 0.0;
            }
            }
            M.internalCoeffs[s][f] = ic;
            M.boundaryCoeffs[s][f] = bc;
        }
    }
    return M;
}

// fvm::div(phi, T) with Gauss upwind: lower = -w*phi, upper = lower + phi,
// negSumDiag, internalCoeffs = phi*valueInternalCoeffs(w),
// boundaryCoeffs = -phi*valueBoundaryCoeffs(w).
FvScalarMatrix fvmDivUpwind(const BoxFvMesh &m, const std::vector<double> &phi,
                            const std::array<std::vector<double>, 6> &phiB,
                            const TemperaturePatches &T)
{
    FvScalarMatrix M;
    M.asymmetric = true;
    M.diag.assign(m.nCells, 0.0);
    M.upper.resize(m.nFaces);
    M.lower.resize(m.nFaces);
    M.source.assign(m.nCells, 0.0);
    for (int f = 0; f < m.nFaces; ++f) {
        double w = phi[f] >= 0.0 ? 1.0 : 0.0;
        M.lower[f] = -w * phi[f];
        M.upper[f] = M.lower[f] + phi[f];
    }
    M.negSumDiag(m);
    for (int s = 0; s < 6; ++s) {
        int nf = m.sideFaces(s);
        M.internalCoeffs[s].resize(nf);
        M.boundaryCoeffs[s].resize(nf);
        for (int f = 0; f < nf; ++f) {
            double p = phiB[s][f];
            double w = p >= 0.0 ? 1.0 : 0.0;
            switch (T.type[s]) {
            case PType::FixedValue:
                M.internalCoeffs[s][f] = p * 0.0;
                M.boundaryCoeffs[s][f] = -p * T.value[s];
                break;
            case PType::Coupled:
                M.internalCoeffs[s][f] = p * w;
                M.boundaryCoeffs[s][f] = -p * (1.0 - w);
                break;
            default:
                M.internalCoeffs[s][f] = p * 1.0;
                M.boundaryCoeffs[s][f] = -p * 0.0;
            }
        }
    }
    return M;
}

// ---------------------------------------------------------------------------
// Region data
// ---------------------------------------------------------------------------
struct RegionCase {
    BoxFvMesh mesh;
    bool transport = false;
    double rhoCp = 0.0;
    std::vector<double> k;
    TemperaturePatches T;
    std::array<FluxKind, 6> flux;
    double T0 = 0.0;
    // velocity
    int flowAxis = 0, profileMask = 0;
    double Umax = 0.0, noise = 0.0;
    uint64_t velSeed = 0;
    // interface data per side (filled once the partner is known)
    std::array<std::vector<double>, 6> coupledK;
    std::array<double, 6> coupledDelta{};
    std::array<std::vector<double>, 6> coupledCoeff;

    double profile(const int idx[3]) const
    {
        double f = 1.0;
        for (int a = 0; a < 3; ++a)
            if (profileMask & (1 << a)) {
                double eta = ((double)idx[a] + 0.5) / (double)mesh.n[a];
                f *= 4.0 * eta * (1.0 - eta);
            }
        return f;
    }

    // Edge vector potential of the divergence-free velocity noise: uniform in
    // +-noise*Umax*h/4 on interior edges, zero on boundary edges.
    double edgePotential(int comp, int I, int J, int K) const
    {
        const int nx = mesh.n[0], ny = mesh.n[1], nz = mesh.n[2];
        uint64_t id;
        if (comp == 0) {
            if (J <= 0 || J >= ny || K <= 0 || K >= nz) return 0.0;
            id = (uint64_t)I + (uint64_t)nx * ((uint64_t)J + (uint64_t)(ny + 1) * (uint64_t)K);
        } else if (comp == 1) {
            if (I <= 0 || I >= nx || K <= 0 || K >= nz) return 0.0;
            id = (uint64_t)I + (uint64_t)(nx + 1) * ((uint64_t)J + (uint64_t)ny * (uint64_t)K);
        } else {
            if (I <= 0 || I >= nx || J <= 0 || J >= ny) return 0.0;
            id = (uint64_t)I + (uint64_t)(nx + 1) * ((uint64_t)J + (uint64_t)(ny + 1) * (uint64_t)K);
        }
        double xi = hashU01(velSeed, 3ULL * id + (uint64_t)comp);
        return 0.25 * noise * Umax * mesh.d[comp] * (2.0 * xi - 1.0);
    }

    // Noise flux through the internal face normal to axis a owned by (i,j,k):
    // the circulation of the edge potential around the face (discrete curl),
    // so the noise adds no net flux to any cell.
    double noiseFlux(int a, int i, int j, int k) const
    {
        if (noise <= 0.0) return 0.0;
        const double *d = mesh.d;
        if (a == 0) {
            int I = i + 1;
            return (edgePotential(2, I, j + 1, k) - edgePotential(2, I, j, k)) * d[2] -
                   (edgePotential(1, I, j, k + 1) - edgePotential(1, I, j, k)) * d[1];
        } else if (a == 1) {
            int J = j + 1;
            return (edgePotential(0, i, J, k + 1) - edgePotential(0, i, J, k)) * d[0] -
                   (edgePotential(2, i + 1, J, k) - edgePotential(2, i, J, k)) * d[2];
        }
        int K = k + 1;
        return (edgePotential(1, i + 1, j, K) - edgePotential(1, i, j, K)) * d[1] -
               (edgePotential(0, i, j + 1, K) - edgePotential(0, i, j, K)) * d[0];
    }
};

std::array<double, 3> vec3(const Dict &d, const std::string &key)
{
    const auto &items = d.list(key);
    if (items.size() != 3) throw DictError("'" + key + "' needs 3 components");
    return {items[0].asNumber(), items[1].asNumber(), items[2].asNumber()};
}

void readRegion(const std::string &name, const Dict &d, RegionCase &R)
{
    BoxFvMesh &m = R.mesh;
    m.name = name;
    auto nc = vec3(d, "nCells");
    auto org = vec3(d, "origin");
    auto len = vec3(d, "length");
    for (int a = 0; a < 3; ++a) {
        m.n[a] = (int)nc[a];
        if (m.n[a] < 1) throw DictError("region '" + name + "': nCells must be >= 1");
        m.origin[a] = org[a];
        m.length[a] = len[a];
    }
    std::string type = d.word("type");
    if (type == "transportTemperature") R.transport = true;
    else if (type == "conductTemperature") R.transport = false;
    else throw DictError("region '" + name + "': unknown regionType '" + type + "'");
    R.rhoCp = d.numberOrDefault("rhoCp", 0.0);
    R.T0 = d.number("T0");

    const Dict &bd = d.subDict("boundary");
    for (int s = 0; s < 6; ++s) {
        const Dict &p = bd.subDict(kSides[s]);
        std::string t = p.word("type");
        if (t == "fixedValue") R.T.type[s] = PType::FixedValue;
        else if (t == "zeroGradient") R.T.type[s] = PType::ZeroGradient;
        else if (t == "empty") R.T.type[s] = PType::Empty;
        else if (t == "monolithicTemperature") R.T.type[s] = PType::Coupled;
        else throw DictError("patch '" + std::string(kSides[s]) + "': unknown type '" + t + "'");
        R.T.value[s] = R.T.type[s] == PType::FixedValue ? p.number("value") : 0.0;
        std::string fk = p.wordOrDefault("flux", "wall");
        R.flux[s] = fk == "inlet" ? FluxKind::Inlet : fk == "outlet" ? FluxKind::Outlet : FluxKind::Wall;
        m.ptype[s] = R.T.type[s];
    }
    m.init();

    // conductivity field
    const Dict &kd = d.subDict("k");
    std::string kt = kd.word("type");
    R.k.resize(m.nCells);
    if (kt == "uniform") {
        double v = kd.number("value");
        for (auto &x : R.k) x = v;
    } else if (kt == "blockLogUniform") {
        int B = (int)kd.number("blockSize");
        double kmin = kd.number("min"), kmax = kd.number("max");
        uint64_t seed = (uint64_t)kd.number("seed");
        int nbx = (m.n[0] + B - 1) / B, nby = (m.n[1] + B - 1) / B;
        double lo = std::log(kmin), hi = std::log(kmax);
        for (int c = 0; c < m.nCells; ++c) {
            int i, j, k;
            m.cellIJK(c, i, j, k);
            uint64_t bidx = (uint64_t)(i / B) +
                            (uint64_t)nbx * ((uint64_t)(j / B) + (uint64_t)nby * (uint64_t)(k / B));
            double u = hashU01(seed, bidx);
            R.k[c] = std::exp(lo + u * (hi - lo));
        }
    } else {
        throw DictError("region '" + name + "': unknown k type '" + kt + "'");
    }

    if (R.transport) {
        const Dict &ud = d.subDict("U");
        R.flowAxis = (int)ud.number("flowAxis");
        R.profileMask = 0;
        for (auto &v : ud.list("profileAxes")) R.profileMask |= 1 << (int)v.asNumber();
        R.Umax = ud.number("Umax");
        R.noise = ud.numberOrDefault("noise", 0.0);
        R.velSeed = (uint64_t)ud.numberOrDefault("seed", 0.0);
    }
}

// ---------------------------------------------------------------------------
// GGI for axis-aligned rectangular faces
// ---------------------------------------------------------------------------
struct Overlap1D {
    int j;
    double ov;
};

std::vector<std::vector<Overlap1D>> overlaps1D(int nA, double oA, double dA, int nB, double oB, double dB)
{
    std::vector<std::vector<Overlap1D>> L(nA);
    for (int i = 0; i < nA; ++i) {
        double lo = oA + (double)i * dA, hi = oA + (double)(i + 1) * dA;
        int j0 = (int)std::floor((lo - oB) / dB) - 1, j1 = (int)std::floor((hi - oB) / dB) + 1;
        if (j0 < 0) j0 = 0;
        if (j1 > nB - 1) j1 = nB - 1;
        for (int j = j0; j <= j1; ++j) {
            double blo = oB + (double)j * dB, bhi = oB + (double)(j + 1) * dB;
            double ov = std::fmin(hi, bhi) - std::fmax(lo, blo);
            if (ov > 1e-10 * std::fmin(hi - lo, bhi - blo)) L[i].push_back({j, ov});
        }
    }
    return L;
}

// Faces of side (A,sA) gather from faces of (B,sB): exact overlap areas
// rescaled to sum 1 per covered face; uncovered faces get no pairs.
void ggiWeights(const BoxFvMesh &A, int sA, const BoxFvMesh &B, InterfaceSide &S)
{
    int a = sA / 2, b, c;
    tangential(a, b, c);
    auto Lb = overlaps1D(A.n[b], A.origin[b], A.d[b], B.n[b], B.origin[b], B.d[b]);
    auto Lc = overlaps1D(A.n[c], A.origin[c], A.d[c], B.n[c], B.origin[c], B.d[c]);
    int nf = A.n[b] * A.n[c];
    S.off.assign(nf + 1, 0);
    S.idx.clear();
    S.w.clear();
    double faceArea = A.area[a];
    for (int f = 0; f < nf; ++f) {
        const auto &lb = Lb[f % A.n[b]];
        const auto &lc = Lc[f / A.n[b]];
        S.off[f] = (int32_t)S.w.size();
        size_t p0 = S.w.size();
        double sum = 0.0;
        for (const auto &qc : lc)
            for (const auto &qb : lb) {
                double ar = qb.ov * qc.ov;
                S.idx.push_back(qb.j + B.n[b] * qc.j);
                S.w.push_back(ar);
                sum += ar;
            }
        if (sum > 1e-12 * faceArea) {
            for (size_t q = p0; q < S.w.size(); ++q) S.w[q] = S.w[q] / sum;
        } else {
            S.idx.resize(p0);
            S.w.resize(p0);
        }
    }
    S.off[nf] = (int32_t)S.w.size();
}

// monolithicThermalDiffusivityFvPatchScalarField::calcThermalDiffusivity with
// radiation off: the owner side's face conductivity from its own k and the
// GGI-interpolated neighbour k (forig), series resistance, harmonic clamp.
void calcThermalDiffusivity(const RegionCase &own, int sOwn, const RegionCase &nei, int sNei,
                            InterfaceSide &S, std::vector<double> &kface, double &delta)
{
    const BoxFvMesh &mo = own.mesh, &mn = nei.mesh;
    int a = sOwn / 2;
    int nf = mo.sideFaces(sOwn);
    double dOwn = 0.5 * mo.d[a], dNei = 0.5 * mn.d[a];
    double magDelta = dOwn + dNei;        // magLongDelta
    double wP = dNei / magDelta;          // patch weights
    delta = 1.0 / magDelta;               // patch deltaCoeffs
    kface.resize(nf);
    S.fc.resize(nf);
    S.magSf.assign(nf, mo.area[a]);
    S.kOwn.resize(nf);
    S.kNei.resize(nf);
    int idx[3];
    for (int f = 0; f < nf; ++f) {
        int cO = mo.sideFaceCell(sOwn, f, idx);
        S.fc[f] = cO;
        double fOwn = own.k[cO];
        double fNei = 0.0;   // regionCouplePatch().interpolate(neighbour.forig())
        for (int p = S.off[f]; p < S.off[f + 1]; ++p)
            fNei += S.w[p] * nei.k[mn.sideFaceCell(sNei, S.idx[p], idx)];
        double kOwn = fOwn / (1.0 - wP) / magDelta;
        double kNei = fNei / wP / magDelta;
        S.kOwn[f] = kOwn;
        S.kNei[f] = kNei;
        double k = kOwn * kNei;
        k /= delta * (kOwn + kNei);
        double kHarm = 1.0 / (wP / fOwn + (1.0 - wP) / fNei);
        kface[f] = std::fmax(std::fmin(k, 100.0 * kHarm), 0.01 * kHarm);
    }
}

// Face velocities -> face fluxes: internal phi = (w U_P + (1-w) U_N) . Sf plus
// the divergence-free noise flux; boundary phi = +-U_n |Sf| with the inlet
// profile (fixedValue U), the face cell's U (outlet, zeroGradient U) or zero
// (walls and coupled patches).
void faceFluxes(const RegionCase &R, std::vector<double> &phi, std::array<std::vector<double>, 6> &phiB)
{
    const BoxFvMesh &m = R.mesh;
    std::vector<double> Ua(m.nCells, 0.0);   // flow-axis velocity per cell
    for (int c = 0; c < m.nCells; ++c) {
        int idx[3];
        m.cellIJK(c, idx[0], idx[1], idx[2]);
        Ua[c] = R.Umax * R.profile(idx);
    }
    phi.resize(m.nFaces);
    for (int f = 0; f < m.nFaces; ++f) {
        int a = m.faceAxis[f], c = m.owner[f], u = m.neighbour[f];
        double UP = a == R.flowAxis ? Ua[c] : 0.0;
        double UN = a == R.flowAxis ? Ua[u] : 0.0;
        int i, j, k;
        m.cellIJK(c, i, j, k);
        phi[f] = (0.5 * UP + 0.5 * UN) * m.area[a] + R.noiseFlux(a, i, j, k);
    }
    for (int s = 0; s < 6; ++s) {
        int nf = m.sideFaces(s), a = s / 2;
        double sign = (s % 2) ? 1.0 : -1.0;
        phiB[s].assign(nf, 0.0);
        for (int f = 0; f < nf; ++f) {
            int idx[3];
            int c = m.sideFaceCell(s, f, idx);
            if (R.flux[s] == FluxKind::Inlet) {
                double Ub = a == R.flowAxis ? R.Umax * R.profile(idx) : 0.0;
                phiB[s][f] = sign * Ub * m.area[a];
            } else if (R.flux[s] == FluxKind::Outlet) {
                double Uc = a == R.flowAxis ? Ua[c] : 0.0;
                phiB[s][f] = sign * Uc * m.area[a];
            }
        }
    }
}

// setCoupledEqns of the region types; returns the TEqn fvMatrix.
FvScalarMatrix setCoupledEqns(const RegionCase &R)
{
    const BoxFvMesh &m = R.mesh;
    FvScalarMatrix lap = fvmLaplacian(m, R.k, R.T, R.coupledK, R.coupledDelta);
    FvScalarMatrix E;
    E.source.assign(m.nCells, 0.0);
    if (R.transport) {
        // rho*cp*(fvm::ddt(T) + fvm::div(phi,T)) == fvm::laplacian(k,T), steady ddt
        std::vector<double> phi;
        std::array<std::vector<double>, 6> phiB;
        faceFluxes(R, phi, phiB);
        FvScalarMatrix div = fvmDivUpwind(m, phi, phiB, R.T);
        E.asymmetric = true;
        E.diag.resize(m.nCells);
        E.upper.resize(m.nFaces);
        E.lower.resize(m.nFaces);
        for (int c = 0; c < m.nCells; ++c) E.diag[c] = R.rhoCp * div.diag[c] - lap.diag[c];
        for (int f = 0; f < m.nFaces; ++f) {
            E.upper[f] = R.rhoCp * div.upper[f] - lap.upper[f];
            E.lower[f] = R.rhoCp * div.lower[f] - lap.upper[f];
        }
        for (int s = 0; s < 6; ++s) {
            size_t nf = lap.internalCoeffs[s].size();
            E.internalCoeffs[s].resize(nf);
            E.boundaryCoeffs[s].resize(nf);
            for (size_t f = 0; f < nf; ++f) {
                E.internalCoeffs[s][f] = R.rhoCp * div.internalCoeffs[s][f] - lap.internalCoeffs[s][f];
                E.boundaryCoeffs[s][f] = R.rhoCp * div.boundaryCoeffs[s][f] - lap.boundaryCoeffs[s][f];
            }
        }
    } else {
        // fvm::ddt(rho*cv,T) == fvm::laplacian(k,T), steady ddt
        E.diag.resize(m.nCells);
        E.upper.resize(m.nFaces);
        for (int c = 0; c < m.nCells; ++c) E.diag[c] = 0.0 - lap.diag[c];
        for (int f = 0; f < m.nFaces; ++f) E.upper[f] = 0.0 - lap.upper[f];
        for (int s = 0; s < 6; ++s) {
            size_t nf = lap.internalCoeffs[s].size();
            E.internalCoeffs[s].resize(nf);
            E.boundaryCoeffs[s].resize(nf);
            for (size_t f = 0; f < nf; ++f) {
                E.internalCoeffs[s][f] = 0.0 - lap.internalCoeffs[s][f];
                E.boundaryCoeffs[s][f] = 0.0 - lap.boundaryCoeffs[s][f];
            }
        }
    }
    return E;
}

}  // namespace

void CoupledSystem::check() const
{
    for (const auto &r : regions) {
        if ((int)r.lAddr.size() != r.nFaces || (int)r.uAddr.size() != r.nFaces ||
            (int)r.diag.size() != r.nCells || (int)r.upper.size() != r.nFaces ||
            (int)r.source.size() != r.nCells || (int)r.psi.size() != r.nCells)
            throw std::runtime_error("region '" + r.name + "': array sizes do not match nCells/nFaces");
        if (!r.symmetric && (int)r.lower.size() != r.nFaces)
            throw std::runtime_error("region '" + r.name + "': asymmetric matrix needs lower coefficients");
        if (r.transient() && ((int)r.ddtCoeff.size() != r.nCells || (int)r.diagBase.size() != r.nCells ||
                              (int)r.sourceBase.size() != r.nCells))
            throw std::runtime_error("region '" + r.name + "': ddt arrays do not match nCells");
        for (int f = 0; f < r.nFaces; ++f) {
            int l = r.lAddr[f], u = r.uAddr[f];
            if (l < 0 || u >= r.nCells || l >= u)
                throw std::runtime_error("region '" + r.name + "': face " + std::to_string(f) +
                                         " is not upper-triangular (l < u)");
            if (f > 0 && (r.lAddr[f - 1] > l || (r.lAddr[f - 1] == l && r.uAddr[f - 1] >= u)))
                throw std::runtime_error("region '" + r.name + "': faces are not in upper-triangular order");
        }
    }
    for (const auto &I : interfaces) {
        const InterfaceSide *sides[2] = {&I.a, &I.b};
        for (int s = 0; s < 2; ++s) {
            const InterfaceSide &S = *sides[s], &O = *sides[1 - s];
            if (S.region < 0 || S.region >= (int)regions.size())
                throw std::runtime_error("interface '" + I.name + "': bad region index");
            int nC = regions[S.region].nCells;
            if ((int)S.coeff.size() != S.nFaces() || (int)S.off.size() != S.nFaces() + 1 ||
                S.idx.size() != S.w.size() || S.off.back() != (int32_t)S.w.size())
                throw std::runtime_error("interface '" + I.name + "': inconsistent GGI addressing");
            if (S.kOwn.size() != S.kNei.size() || (!S.kOwn.empty() && (int)S.kOwn.size() != S.nFaces()))
                throw std::runtime_error("interface '" + I.name + "': wall conductances do not match the faces");
            for (int32_t c : S.fc)
                if (c < 0 || c >= nC) throw std::runtime_error("interface '" + I.name + "': face cell out of range");
            for (int32_t j : S.idx)
                if (j < 0 || j >= O.nFaces()) throw std::runtime_error("interface '" + I.name + "': GGI index out of range");
        }
    }
}

CoupledSystem buildCase(const std::string &text, CaseExtras *extras)
{
    Dict root = parseDict(text);
    const auto &rl = root.list("regions");
    std::vector<RegionCase> R;
    std::map<std::string, int> byName;
    for (size_t i = 0; i < rl.size(); ++i) {
        if (!rl[i].isDict()) continue;
        std::string name = rl[i].dict->name;
        R.emplace_back();
        readRegion(name, *rl[i].dict, R.back());
        byName[name] = (int)R.size() - 1;
    }
    CoupledSystem sys;
    std::vector<std::array<int, 2>> sidePartner;
    if (root.found("monolithicCoupledPatches")) {
        for (const auto &it : root.list("monolithicCoupledPatches")) {
            if (!it.isDict()) continue;
            const Dict &d = *it.dict;
            const auto &pp = d.list("coupledPatchPair");
            if (pp.size() != 2 || !pp[0].isList() || !pp[1].isList())
                throw DictError("coupledPatchPair must be ((regionA patchA) (regionB patchB))");
            int reg[2], side[2];
            for (int s = 0; s < 2; ++s) {
                const std::string &rn = pp[s].items.at(0).asWord();
                const std::string &pn = pp[s].items.at(1).asWord();
                auto f = byName.find(rn);
                if (f == byName.end()) throw DictError("unknown region '" + rn + "'");
                reg[s] = f->second;
                side[s] = -1;
                for (int q = 0; q < 6; ++q)
                    if (pn == kSides[q]) side[s] = q;
                if (side[s] < 0) throw DictError("unknown patch '" + pn + "'");
                if (R[reg[s]].T.type[side[s]] != PType::Coupled)
                    throw DictError("patch '" + pn + "' of '" + rn + "' is not monolithicTemperature");
            }
            if (side[0] / 2 != side[1] / 2) throw DictError("coupled patches must share a normal axis");
            CoupledInterface I;
            I.name = d.name;
            I.a.region = reg[0];
            I.b.region = reg[1];
            ggiWeights(R[reg[0]].mesh, side[0], R[reg[1]].mesh, I.a);
            ggiWeights(R[reg[1]].mesh, side[1], R[reg[0]].mesh, I.b);
            calcThermalDiffusivity(R[reg[0]], side[0], R[reg[1]], side[1], I.a,
                                   R[reg[0]].coupledK[side[0]], R[reg[0]].coupledDelta[side[0]]);
            calcThermalDiffusivity(R[reg[1]], side[1], R[reg[0]], side[0], I.b,
                                   R[reg[1]].coupledK[side[1]], R[reg[1]].coupledDelta[side[1]]);
            sys.interfaces.push_back(std::move(I));
            sidePartner.push_back({6 * reg[0] + side[0], 6 * reg[1] + side[1]});
        }
    }
    for (size_t r = 0; r < R.size(); ++r)
        for (int s = 0; s < 6; ++s)
            if (R[r].T.type[s] == PType::Coupled && R[r].coupledK[s].empty() && R[r].mesh.sideFaces(s) > 0)
                throw DictError("coupled patch '" + std::string(kSides[s]) + "' of region '" + R[r].mesh.name +
                                "' has no monolithicCoupledPatches entry");

    sys.regions.resize(R.size());
    for (size_t r = 0; r < R.size(); ++r) {
        const RegionCase &rc = R[r];
        const BoxFvMesh &m = rc.mesh;
        FvScalarMatrix E = setCoupledEqns(rc);
        LduRegion &L = sys.regions[r];
        L.name = m.name;
        L.nCells = m.nCells;
        L.nFaces = m.nFaces;
        L.lAddr = m.owner;
        L.uAddr = m.neighbour;
        L.symmetric = !E.asymmetric;
        L.diag = E.diag;
        L.upper = E.upper;
        L.lower = E.asymmetric ? E.lower : E.upper;
        L.source = E.source;
        L.psi.assign(m.nCells, rc.T0);
        // coupledFvMatrix::solve preparation: addBoundaryDiag for every patch,
        // addBoundarySource(couples=false) for non-coupled patches; the coupled
        // patches' boundaryCoeffs become the interface coefficients.
        int idx[3];
        for (int s = 0; s < 6; ++s)
            for (size_t f = 0; f < E.internalCoeffs[s].size(); ++f)
                L.diag[m.sideFaceCell(s, (int)f, idx)] += E.internalCoeffs[s][f];
        for (int s = 0; s < 6; ++s) {
            if (rc.T.type[s] == PType::Coupled) continue;
            for (size_t f = 0; f < E.boundaryCoeffs[s].size(); ++f)
                L.source[m.sideFaceCell(s, (int)f, idx)] += E.boundaryCoeffs[s][f];
        }
        for (size_t i = 0; i < sys.interfaces.size(); ++i)
            for (int s2 = 0; s2 < 2; ++s2) {
                int key = sidePartner[i][s2];
                if (key / 6 != (int)r) continue;
                InterfaceSide &S = s2 ? sys.interfaces[i].b : sys.interfaces[i].a;
                S.coeff = E.boundaryCoeffs[key % 6];
            }
    }
    if (extras) {
        extras->k.clear();
        extras->cellSize.clear();
        for (auto &rc : R) {
            extras->k.push_back(rc.k);
            extras->cellSize.push_back({rc.mesh.d[0], rc.mesh.d[1], rc.mesh.d[2]});
        }
        extras->ifaceNormalAxis.clear();
        for (auto &sp : sidePartner) extras->ifaceNormalAxis.push_back((sp[0] % 6) / 2);
    }
    sys.check();
    return sys;
}

}  // namespace mrfg
