// The coupled block system handed to the solver: one LDU matrix per region
// plus monolithic interfaces with GGI gather weights.  This is what
// coupledFvMatrix<scalar>::solve hands to the coupled solver after
// addBoundaryDiag / addBoundarySource (call site multiRegionSystem.C:150-153).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace mrfg {

struct LduRegion {
    std::string name;
    int nCells = 0, nFaces = 0;
    std::vector<int32_t> lAddr, uAddr;      // owner / neighbour, upper-triangular order
    std::vector<double> diag, lower, upper; // diag includes patch internalCoeffs
    std::vector<double> source;             // includes non-coupled boundaryCoeffs
    std::vector<double> psi;                // initial guess / solution
    bool symmetric = false;
    // transient regions (empty when steady): the Euler ddt coefficient
    // a = rho*c*V/dt per cell and the diagonal and source without the ddt
    // term, so that diag = diagBase + a and source = sourceBase + a*psi
    std::vector<double> ddtCoeff, diagBase, sourceBase;
    bool transient() const { return !ddtCoeff.empty(); }
};

// One side of a monolithic interface: faces of this side gather the other
// side's face-cell values, pnf_i = sum_p w[p] * psi_other[fcOther[idx[p]]],
// and Amul does result[fc[i]] -= coeff[i] * pnf_i
// (monolithicCouplingFvPatchField.C:381-464).
struct InterfaceSide {
    int region = -1;
    std::vector<int32_t> fc;
    std::vector<double> coeff;
    std::vector<double> magSf;
    std::vector<int32_t> off, idx;
    std::vector<double> w;
    // per face, the two conductances of monolithicTemperature's wall value
    // (calcTemperature, monolithicThermalDiffusivityFvPatchScalarField.C:
    // 300-380, radiation off): kOwn = forig / (1 - w) / |delta| of this
    // face (monolithicBase::korig) and kNei = the GGI-interpolated other
    // side's forig / w / |delta|; empty when the caller gave none
    std::vector<double> kOwn, kNei;
    int nFaces() const { return (int)fc.size(); }
};

struct CoupledInterface {
    std::string name;
    InterfaceSide a, b;
};

// The finite-volume model a system read from disk was assembled from: mesh
// geometry, boundary conditions and the fields k, U, T, so the system can be
// re-assembled on the device when fields change (transportTemperature.C:
// 129-150 setCoupledEqns, monolithicTemperature's calcThermalDiffusivity).
enum FvTBc : int8_t { kBcFixedValue = 0, kBcZeroGradient = 1, kBcEmpty = 2, kBcCoupled = 3 };
enum FvUBc : int8_t { kUFixedValue = 0, kUZeroGradient = 1, kUSlip = 2, kUZero = 3 };

struct FvRegionModel {
    bool transport = false;             // rho*cp*(ddt + div(phi,T)) - laplacian(k,T), else ddt - laplacian
    double rhoCp = 0.0, rDeltaT = 0.0;  // rDeltaT = 0: steady
    // internal faces, in the region's face order
    std::vector<double> w, deltaCoeffs, magSf, Sf;   // Sf: 3 per face
    std::vector<double> V;                            // cell volumes
    // boundary faces, patch by patch (the reader's loop order)
    std::vector<int32_t> bOwner;
    std::vector<int8_t> bT, bU, bKFixed;              // FvTBc, FvUBc, k given (else the owner cell's)
    std::vector<double> bNDist, bMagSf, bSf;          // bSf: 3 per face
    std::vector<double> bTb, bUb, bKb;                // fixed T, fixed U (3 per face), fixed k
    std::vector<int32_t> bSide;                       // coupled faces: the interface-side face, else -1
    // fields
    std::vector<double> k, U, T;                      // U: 3 per cell (transport regions)
};

// Per face of one interface side (same order as InterfaceSide::fc): the
// face's normal distance and boundary-face index; per GGI pair: the partner
// face's normal distance and owner cell (in the other region).
struct FvSideModel {
    std::vector<double> nDist;
    std::vector<int32_t> bFace;
    std::vector<double> pairNDist;
    std::vector<int32_t> pairCell;
};

struct FvModel {
    std::vector<FvRegionModel> regions;
    std::vector<FvSideModel> sides;     // interface i: side a = 2i, side b = 2i + 1
};

// Partitioned (Dirichlet-Neumann) coupling through heatTransferInterface
// pairs (regionInterfacePropertiesPartitioned, dnaControl): each region is
// solved on its own; the Dirichlet side (regionCoupledTemperatureJump) takes
// a relaxed fixed value from the other side's patch values, the Neumann side
// (regionCoupledHeatFlux) the gradient that matches the other side's heat
// flux.  A side's source term is linear in its boundary value (temperature
// or gradient): source[fc[i]] += coeff[i] * value[i], after every other
// boundary contribution of the cell.
struct PartitionedSide {
    int region = -1;
    std::vector<int32_t> fc;          // face cells
    std::vector<double> coeff;        // source term per unit boundary value
    std::vector<double> kb, dc;       // face conductivity, patch deltaCoeffs 1/|d|
    std::vector<double> value0;       // initial patch values
    std::vector<int32_t> off, idx;    // per face: the other side's faces it gathers (GGI)
    std::vector<double> w;
    int accType = 0;                  // 0 fixed relaxation, 1 Aitken, 2 IQN-ILS
    double relax = 1.0;
    int reuse = 0;                    // IQN-ILS couplingReuse (time steps of columns kept)
    int nFaces() const { return (int)fc.size(); }
};

struct PartitionedPair {
    std::string name;
    double K = 1.0;                   // temperature jump coefficient
    PartitionedSide dir, neu;
};

struct PartitionedModel {
    std::vector<PartitionedPair> pairs;
    std::vector<std::vector<double>> sourceBase;   // per region: source without the partitioned faces
    int maxCoupleIter = 1;
    double maxJumpRes = -1.0, maxFluxRes = -1.0;    // < 0: no residual control
};

struct CoupledSystem {
    std::vector<LduRegion> regions;
    std::vector<CoupledInterface> interfaces;
    std::shared_ptr<const FvModel> fv;   // cases read from disk; null otherwise
    std::shared_ptr<const PartitionedModel> part;   // partitioned set-ups read from disk
    long long nCells() const
    {
        long long n = 0;
        for (auto &r : regions) n += r.nCells;
        return n;
    }
    long long nFaces() const
    {
        long long n = 0;
        for (auto &r : regions) n += r.nFaces;
        return n;
    }
    long long nPairs() const
    {
        long long n = 0;
        for (auto &i : interfaces) n += (long long)i.a.w.size() + (long long)i.b.w.size();
        return n;
    }
    // Throws std::runtime_error if addressing is inconsistent.
    void check() const;
};

// Build and assemble a case from its dictionary text (regions +
// monolithicCoupledPatches).  Extra per-region data (k, cell sizes) are
// returned for the interface flux check.
struct CaseExtras {
    std::vector<std::vector<double>> k;
    std::vector<std::vector<double>> cellSize;   // 3 per region
    std::vector<int> ifaceNormalAxis;            // per interface
};

CoupledSystem buildCase(const std::string &dictText, CaseExtras *extras = nullptr);

// Read a multi-region case from disk (ASCII polyMesh, 0/ fields, constant/
// and system/ dictionaries) and assemble its coupled system.  `options` is
// dictionary text: fieldDir (sub-directory of 0/<region>/ holding T and k),
// interfaceProperties (file under constant/), ddt (steadyState | Euler |
// backward, overriding fvSchemes) and deltaT (overriding controlDict).
struct FoamCaseInfo {
    std::vector<std::vector<double>> volumes, k;   // per region, per cell
    std::vector<double> maxNonOrthDeg, rDeltaT, rhoCp;   // per region
};

CoupledSystem readFoamCase(const std::string &caseDir, const std::string &options, FoamCaseInfo *info = nullptr);

// splitmix64 hash -> uniform [0,1)
double hashU01(uint64_t seed, uint64_t index);

}  // namespace mrfg
