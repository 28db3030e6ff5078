// Reading an OpenFOAM / foam-extend multi-region case from disk: ASCII
// polyMesh, the 0/ fields and the constant/ and system/ dictionaries, then
// the finite-volume assembly of the temperature equations on that general
// (polyhedral, orthogonal) mesh into the coupled block system.
//
// Follows the reference's case layout and region types:
//   constant/multiRegionProperties        regions and their region types
//   constant/regionInterfaceProperties    monolithicCoupledPatches
//       (tutorials/conjugateHeatTransfer/flowOverHeatedPlate/constant/*)
//   conductTemperature::setCoupledEqns    (conductTemperature.C:135-152)
//   transportTemperature::setCoupledEqns  (transportTemperature.C:129-150),
//       phi = linearInterpolate(U) & Sf   (transportTemperature.C:86-94)
//   monolithicThermalDiffusivity::calcThermalDiffusivity
//                          (monolithicThermalDiffusivityFvPatchScalarField.C:167-277)
//   directMap face matching (directMapInterfaceToInterfaceMapping.C:150-175)
// Geometry, interpolation weights and delta coefficients follow foam-extend's
// primitiveMesh / surfaceInterpolation definitions; the first time step of a
// `backward` ddt scheme is the Euler step.
//
// Limits (each raises an error): internal faces must be in upper-triangular
// order; non-orthogonality above kMaxNonOrthDeg (the explicit non-orthogonal
// correction of `corrected` snGrad is not assembled); the two patches of an
// interface must lie in one plane (any orientation) and have convex faces,
// whose overlaps in that plane give the GGI weights.
#include "coupled_system.hpp"
#include "dictionary.hpp"

#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace mrfg {

namespace {

constexpr double kMaxNonOrthDeg = 1.0;

using Vec = std::array<double, 3>;

inline Vec operator+(const Vec &a, const Vec &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline Vec operator-(const Vec &a, const Vec &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline Vec operator*(double s, const Vec &a) { return {s * a[0], s * a[1], s * a[2]}; }
inline double dot(const Vec &a, const Vec &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline Vec cross(const Vec &a, const Vec &b)
{
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double mag(const Vec &a) { return std::sqrt(dot(a, a)); }

struct FoamError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

std::string readFile(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw FoamError("cannot open '" + path + "'");
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}

bool fileExists(const std::string &path)
{
    std::ifstream f(path);
    return (bool)f;
}

// ---------------------------------------------------------------------------
// polyMesh list files: FoamFile header, then "N ( item ... )"
// ---------------------------------------------------------------------------
class ListReader {
public:
    ListReader(const std::string &path) : path_(path), text_(readFile(path))
    {
        p_ = text_.c_str();
        e_ = p_ + text_.size();
        skipHeader();
    }
    int count()
    {
        long n = integer();
        if (n < 0) fail("negative list size");
        expect('(');
        return (int)n;
    }
    long integer()
    {
        skip();
        char *end;
        errno = 0;
        long v = std::strtol(p_, &end, 10);
        if (end == p_ || errno) fail("integer expected");
        p_ = end;
        return v;
    }
    double real()
    {
        skip();
        char *end;
        double v = std::strtod(p_, &end);
        if (end == p_) fail("number expected");
        p_ = end;
        return v;
    }
    void expect(char c)
    {
        skip();
        if (p_ >= e_ || *p_ != c) fail(std::string("'") + c + "' expected");
        ++p_;
    }

private:
    [[noreturn]] void fail(const std::string &what)
    {
        throw FoamError(path_ + ": " + what + " at byte " + std::to_string(p_ - text_.c_str()));
    }
    void skip()
    {
        for (;;) {
            while (p_ < e_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
            if (p_ + 1 < e_ && p_[0] == '/' && p_[1] == '/') {
                while (p_ < e_ && *p_ != '\n') ++p_;
            } else if (p_ + 1 < e_ && p_[0] == '/' && p_[1] == '*') {
                const char *q = std::strstr(p_ + 2, "*/");
                p_ = q ? q + 2 : e_;
            } else {
                return;
            }
        }
    }
    void skipHeader()
    {
        skip();
        if (e_ - p_ >= 8 && std::strncmp(p_, "FoamFile", 8) == 0) {
            const char *q = std::strchr(p_, '}');
            if (!q) fail("unterminated FoamFile header");
            p_ = q + 1;
        }
    }
    std::string path_, text_;
    const char *p_, *e_;
};

// ---------------------------------------------------------------------------
// Mesh: points, faces, owner / neighbour, patches, and the derived geometry
// ---------------------------------------------------------------------------
struct Patch {
    std::string name, type;
    int start = 0, size = 0;
};

struct PolyMesh {
    std::string name;
    std::vector<Vec> points;
    std::vector<int32_t> faceOff, faceVerts;
    std::vector<int32_t> owner, neighbour;   // neighbour: internal faces only
    std::vector<Patch> patches;
    int nCells = 0, nInternal = 0;
    // geometry
    std::vector<Vec> Cf, Sf, C;
    std::vector<double> magSf, V;
    std::vector<double> w, deltaCoeffs;      // internal faces
    double maxNonOrthDeg = 0.0;

    int nFaces() const { return (int)owner.size(); }
    const Patch &patch(const std::string &n) const
    {
        for (auto &p : patches)
            if (p.name == n) return p;
        throw FoamError("region '" + name + "': no patch '" + n + "'");
    }
    // unit normal and owner-centre-to-face normal distance of boundary face f
    double nDist(int f) const { return dot((1.0 / magSf[f]) * Sf[f], Cf[f] - C[owner[f]]); }
};

void readPolyMesh(const std::string &dir, PolyMesh &m)
{
    {
        ListReader r(dir + "/points");
        int n = r.count();
        m.points.resize(n);
        for (int i = 0; i < n; ++i) {
            r.expect('(');
            for (int a = 0; a < 3; ++a) m.points[i][a] = r.real();
            r.expect(')');
        }
        r.expect(')');
    }
    {
        ListReader r(dir + "/faces");
        int n = r.count();
        m.faceOff.assign(1, 0);
        m.faceOff.reserve(n + 1);
        m.faceVerts.reserve((size_t)n * 4);
        for (int i = 0; i < n; ++i) {
            long k = r.integer();
            if (k < 3) throw FoamError(dir + "/faces: face " + std::to_string(i) + " has fewer than 3 points");
            r.expect('(');
            for (long j = 0; j < k; ++j) {
                long v = r.integer();
                if (v < 0 || v >= (long)m.points.size())
                    throw FoamError(dir + "/faces: point index out of range in face " + std::to_string(i));
                m.faceVerts.push_back((int32_t)v);
            }
            r.expect(')');
            m.faceOff.push_back((int32_t)m.faceVerts.size());
        }
        r.expect(')');
    }
    auto labels = [&](const std::string &file, std::vector<int32_t> &out) {
        ListReader r(dir + "/" + file);
        int n = r.count();
        out.resize(n);
        for (int i = 0; i < n; ++i) out[i] = (int32_t)r.integer();
        r.expect(')');
    };
    labels("owner", m.owner);
    labels("neighbour", m.neighbour);
    const int nF = (int)m.faceOff.size() - 1;
    if ((int)m.owner.size() != nF) throw FoamError(dir + ": owner and faces differ in size");
    m.nInternal = (int)m.neighbour.size();
    m.nCells = 0;
    for (int32_t c : m.owner) m.nCells = std::max(m.nCells, c + 1);
    for (int32_t c : m.neighbour) m.nCells = std::max(m.nCells, c + 1);
    for (int f = 0; f < m.nInternal; ++f) {
        if (m.owner[f] >= m.neighbour[f])
            throw FoamError(dir + ": internal face " + std::to_string(f) + " has owner >= neighbour");
        if (f > 0 && (m.owner[f - 1] > m.owner[f] ||
                      (m.owner[f - 1] == m.owner[f] && m.neighbour[f - 1] >= m.neighbour[f])))
            throw FoamError(dir + ": internal faces are not in upper-triangular order");
    }
    // boundary: "N ( name { type ..; nFaces ..; startFace ..; } ... )"
    std::string b = readFile(dir + "/boundary");
    size_t h = b.find("FoamFile");
    if (h != std::string::npos) {
        size_t e = b.find('}', h);
        if (e == std::string::npos) throw FoamError(dir + "/boundary: unterminated FoamFile header");
        b = b.substr(e + 1);
    }
    Dict bd = parseDict("patches " + b + ";");
    int next = m.nInternal;
    for (const Value &v : bd.lookup("patches").tokens) {
        if (!v.isList()) continue;
        for (const Value &it : v.items) {
            if (!it.isDict()) continue;
            Patch p;
            p.name = it.dict->name;
            p.type = it.dict->word("type");
            p.size = (int)it.dict->number("nFaces");
            p.start = (int)it.dict->number("startFace");
            if (p.start != next) throw FoamError(dir + "/boundary: patch '" + p.name + "' does not follow the previous one");
            next += p.size;
            m.patches.push_back(p);
        }
    }
    if (next != nF) throw FoamError(dir + "/boundary: patches do not cover the boundary faces");
}

// primitiveMesh face centres / areas (triangle fan about the point average)
// and cell centres / volumes (pyramids about the face-centre average).
void computeGeometry(PolyMesh &m)
{
    const int nF = m.nFaces();
    m.Cf.resize(nF);
    m.Sf.resize(nF);
    m.magSf.resize(nF);
    for (int f = 0; f < nF; ++f) {
        const int32_t *v = &m.faceVerts[m.faceOff[f]];
        const int n = m.faceOff[f + 1] - m.faceOff[f];
        if (n == 3) {
            const Vec &a = m.points[v[0]], &b = m.points[v[1]], &c = m.points[v[2]];
            m.Cf[f] = (1.0 / 3.0) * (a + b + c);
            m.Sf[f] = 0.5 * cross(b - a, c - a);
        } else {
            Vec fc{0, 0, 0};
            for (int i = 0; i < n; ++i) fc = fc + m.points[v[i]];
            fc = (1.0 / n) * fc;
            Vec sumN{0, 0, 0}, sumAc{0, 0, 0};
            double sumA = 0.0;
            for (int i = 0; i < n; ++i) {
                const Vec &p = m.points[v[i]], &q = m.points[v[(i + 1) % n]];
                Vec c = p + q + fc;
                Vec nt = cross(q - p, fc - p);
                double a = mag(nt);
                sumN = sumN + nt;
                sumA += a;
                sumAc = sumAc + a * c;
            }
            m.Cf[f] = sumA > 0.0 ? (1.0 / (3.0 * sumA)) * sumAc : fc;
            m.Sf[f] = 0.5 * sumN;
        }
        m.magSf[f] = mag(m.Sf[f]);
        if (!(m.magSf[f] > 0.0)) throw FoamError("region '" + m.name + "': face " + std::to_string(f) + " has zero area");
    }
    std::vector<Vec> cEst(m.nCells, Vec{0, 0, 0});
    std::vector<int> nCf(m.nCells, 0);
    for (int f = 0; f < nF; ++f) {
        cEst[m.owner[f]] = cEst[m.owner[f]] + m.Cf[f];
        ++nCf[m.owner[f]];
        if (f < m.nInternal) {
            cEst[m.neighbour[f]] = cEst[m.neighbour[f]] + m.Cf[f];
            ++nCf[m.neighbour[f]];
        }
    }
    for (int c = 0; c < m.nCells; ++c) {
        if (nCf[c] < 4) throw FoamError("region '" + m.name + "': cell " + std::to_string(c) + " has fewer than 4 faces");
        cEst[c] = (1.0 / nCf[c]) * cEst[c];
    }
    std::vector<Vec> ctr(m.nCells, Vec{0, 0, 0});
    std::vector<double> vol(m.nCells, 0.0);
    auto pyramid = [&](int c, int f, double sign) {
        double pv = sign * dot(m.Sf[f], m.Cf[f] - cEst[c]);
        Vec pc = 0.75 * m.Cf[f] + 0.25 * cEst[c];
        ctr[c] = ctr[c] + pv * pc;
        vol[c] += pv;
    };
    for (int f = 0; f < nF; ++f) {
        pyramid(m.owner[f], f, 1.0);
        if (f < m.nInternal) pyramid(m.neighbour[f], f, -1.0);
    }
    m.C.resize(m.nCells);
    m.V.resize(m.nCells);
    for (int c = 0; c < m.nCells; ++c) {
        if (!(vol[c] > 0.0)) throw FoamError("region '" + m.name + "': cell " + std::to_string(c) + " has non-positive volume");
        m.C[c] = (1.0 / vol[c]) * ctr[c];
        m.V[c] = vol[c] / 3.0;
    }
    // surfaceInterpolation: linear weights and (non-orthogonal) delta coefficients
    m.w.resize(m.nInternal);
    m.deltaCoeffs.resize(m.nInternal);
    double minCos = 1.0;
    for (int f = 0; f < m.nInternal; ++f) {
        const Vec &co = m.C[m.owner[f]], &cn = m.C[m.neighbour[f]];
        double sfdOwn = std::fabs(dot(m.Sf[f], m.Cf[f] - co));
        double sfdNei = std::fabs(dot(m.Sf[f], cn - m.Cf[f]));
        m.w[f] = sfdNei / (sfdOwn + sfdNei);
        Vec delta = cn - co;
        Vec unit = (1.0 / m.magSf[f]) * m.Sf[f];
        double nd = dot(unit, delta);
        m.deltaCoeffs[f] = 1.0 / std::fmax(nd, 0.05 * mag(delta));
        minCos = std::fmin(minCos, nd / mag(delta));
    }
    m.maxNonOrthDeg = std::acos(std::fmax(-1.0, std::fmin(1.0, minCos))) * 180.0 / M_PI;
    if (m.maxNonOrthDeg > kMaxNonOrthDeg)
        throw FoamError("region '" + m.name + "': mesh non-orthogonality " + std::to_string(m.maxNonOrthDeg) +
                        " deg exceeds the supported " + std::to_string(kMaxNonOrthDeg) +
                        " deg (no explicit non-orthogonal correction)");
}

// ---------------------------------------------------------------------------
// Fields: internalField and per-patch boundary conditions
// ---------------------------------------------------------------------------
struct PatchBC {
    std::string type;
    std::vector<double> value;   // per face (scalar) or 3 per face (vector)
    std::string accType = "fixed";   // partitioned coupling: acceleration model
    double relax = 1.0;
    int reuse = 0;                   // IQN-ILS couplingReuse
};

struct FieldFile {
    std::vector<double> internal;              // 1 or 3 per cell
    std::map<std::string, PatchBC> patches;
};

const Value *lastOf(const std::vector<Value> &t, Value::Kind k)
{
    for (size_t i = t.size(); i-- > 0;)
        if (t[i].kind == k) return &t[i];
    return nullptr;
}

// "uniform X" / "uniform (x y z)" / "nonuniform List<..> N (...)" / "$internalField"
std::vector<double> fieldValues(const Dict &top, const std::vector<Value> &tok, int n, int nComp,
                                const std::string &what)
{
    std::vector<double> out;
    if (tok.size() == 1 && tok[0].isWord() && tok[0].asWord().size() > 1 && tok[0].asWord()[0] == '$')
        return fieldValues(top, top.lookup(tok[0].asWord().substr(1)).tokens, n, nComp, what);
    if (tok.empty() || !tok[0].isWord()) throw FoamError(what + ": value expected");
    const std::string &kind = tok[0].asWord();
    auto comp = [&](const Value &v, std::vector<double> &dst) {
        if (nComp == 1) {
            dst.push_back(v.asNumber());
        } else {
            if (!v.isList() || (int)v.items.size() != nComp) throw FoamError(what + ": vector value expected");
            for (const Value &x : v.items) dst.push_back(x.asNumber());
        }
    };
    if (kind == "uniform") {
        if (tok.size() < 2) throw FoamError(what + ": uniform value missing");
        std::vector<double> one;
        comp(tok[1], one);
        out.reserve((size_t)n * nComp);
        for (int i = 0; i < n; ++i) out.insert(out.end(), one.begin(), one.end());
    } else if (kind == "nonuniform") {
        const Value *l = lastOf(tok, Value::List);
        if (!l) throw FoamError(what + ": nonuniform list missing");
        if ((int)l->items.size() != n) throw FoamError(what + ": nonuniform list has the wrong size");
        for (const Value &v : l->items) comp(v, out);
    } else {
        throw FoamError(what + ": unsupported value form '" + kind + "'");
    }
    return out;
}

FieldFile readField(const std::string &path, const PolyMesh &m, int nComp)
{
    Dict d = parseDict(readFile(path));
    FieldFile F;
    F.internal = fieldValues(d, d.lookup("internalField").tokens, m.nCells, nComp, path + " internalField");
    const Dict &bf = d.subDict("boundaryField");
    for (const Patch &p : m.patches) {
        PatchBC bc;
        const Dict *pd = nullptr;
        for (auto &e : bf.entries)
            if (e.dict && e.key == p.name) pd = e.dict.get();
        if (!pd) throw FoamError(path + ": no boundary condition for patch '" + p.name + "'");
        bc.type = pd->word("type");
        if (pd->found("value") && bc.type != "empty")
            bc.value = fieldValues(d, pd->lookup("value").tokens, p.size, nComp, path + " patch " + p.name);
        bc.accType = pd->wordOrDefault("accType", "fixed");
        bc.relax = pd->numberOrDefault("relax", 1.0);
        bc.reuse = (int)pd->numberOrDefault("couplingReuse", 0);
        F.patches[p.name] = bc;
    }
    return F;
}

// the numeric value of a transportProperties entry "name [dims] value;" or
// "name name [dims] value;"
double property(const Dict &d, const std::string &key, const std::string &file)
{
    const Value *v = lastOf(d.lookup(key).tokens, Value::Number);
    if (!v) throw FoamError(file + ": no value for '" + key + "'");
    return v->number;
}

// ---------------------------------------------------------------------------
// Regions and assembly
// ---------------------------------------------------------------------------
// PartDirichlet / PartNeumann: the two sides of a partitioned interface
// (regionCoupledTemperatureJump, a relaxed fixed value; regionCoupledHeatFlux,
// a fixed gradient).  Their source terms are linear in the boundary value and
// are assembled per unit value (the driver of the coupling supplies the value).
enum class BcT { FixedValue, ZeroGradient, Empty, Coupled, PartDirichlet, PartNeumann };

struct FoamRegion {
    PolyMesh mesh;
    bool transport = false;
    double rhoCp = 0.0;                 // rho*cp (transport) / rho*cv (conduct)
    std::vector<double> T, k;           // cell values
    std::vector<BcT> bcT;               // per patch
    std::vector<std::vector<double>> Tb, kb;   // per patch face values (fixedValue T; k)
    std::vector<double> phi;            // internal faces
    std::vector<std::vector<double>> phiB;     // per patch
    std::vector<double> U;              // 3 per cell (transport regions)
    std::vector<int8_t> bcU, kFixed;    // per patch: FvUBc, k given on the patch
    std::vector<std::string> accType;   // per patch (partitioned sides)
    std::vector<double> relax;
    std::vector<int> reuse;
    std::vector<std::vector<double>> Ub;       // per patch: fixed U, 3 per face
    // coupled patches: face conductivity and delta coefficients, and the
    // boundary coefficients handed to the interface
    std::map<int, std::vector<double>> coupledK, coupledDelta, coupledCoeff;
};

struct FvMatrixG {
    std::vector<double> diag, upper, lower, source;
    bool asymmetric = false;
    std::vector<std::vector<double>> internalCoeffs, boundaryCoeffs;   // per patch
    void init(const PolyMesh &m, bool asym)
    {
        asymmetric = asym;
        diag.assign(m.nCells, 0.0);
        upper.assign(m.nInternal, 0.0);
        if (asym) lower.assign(m.nInternal, 0.0);
        source.assign(m.nCells, 0.0);
        internalCoeffs.resize(m.patches.size());
        boundaryCoeffs.resize(m.patches.size());
        for (size_t p = 0; p < m.patches.size(); ++p) {
            internalCoeffs[p].assign(m.patches[p].size, 0.0);
            boundaryCoeffs[p].assign(m.patches[p].size, 0.0);
        }
    }
    void negSumDiag(const PolyMesh &m)
    {
        const std::vector<double> &lo = asymmetric ? lower : upper;
        for (int f = 0; f < m.nInternal; ++f) {
            diag[m.owner[f]] -= lo[f];
            diag[m.neighbour[f]] -= upper[f];
        }
    }
};

FvMatrixG laplacian(const FoamRegion &R)
{
    const PolyMesh &m = R.mesh;
    FvMatrixG M;
    M.init(m, false);
    for (int f = 0; f < m.nInternal; ++f) {
        double kf = m.w[f] * R.k[m.owner[f]] + (1.0 - m.w[f]) * R.k[m.neighbour[f]];
        M.upper[f] = m.deltaCoeffs[f] * (kf * m.magSf[f]);
    }
    M.negSumDiag(m);
    for (size_t p = 0; p < m.patches.size(); ++p) {
        const Patch &P = m.patches[p];
        for (int i = 0; i < P.size; ++i) {
            const int f = P.start + i;
            switch (R.bcT[p]) {
            case BcT::FixedValue: {
                double deltab = 1.0 / m.nDist(f);
                double gammaMagSf = R.kb[p][i] * m.magSf[f];
                M.internalCoeffs[p][i] = gammaMagSf * (-deltab);
                M.boundaryCoeffs[p][i] = -gammaMagSf * (deltab * R.Tb[p][i]);
                break;
            }
            case BcT::Coupled: {
                double gammaMagSf = R.coupledK.at((int)p)[i] * m.magSf[f];
                double delta = R.coupledDelta.at((int)p)[i];
                M.internalCoeffs[p][i] = gammaMagSf * (-delta);
                M.boundaryCoeffs[p][i] = -gammaMagSf * delta;
                break;
            }
            case BcT::PartDirichlet: {   // fixed value, per unit value
                double deltab = 1.0 / m.nDist(f);
                double gammaMagSf = R.kb[p][i] * m.magSf[f];
                M.internalCoeffs[p][i] = gammaMagSf * (-deltab);
                M.boundaryCoeffs[p][i] = -gammaMagSf * (deltab * 1.0);
                break;
            }
            case BcT::PartNeumann: {     // fixed gradient, per unit gradient
                double gammaMagSf = R.kb[p][i] * m.magSf[f];
                M.internalCoeffs[p][i] = gammaMagSf * 0.0;
                M.boundaryCoeffs[p][i] = -gammaMagSf * 1.0;
                break;
            }
            default:
                break;
            }
        }
    }
    return M;
}

FvMatrixG divUpwind(const FoamRegion &R)
{
    const PolyMesh &m = R.mesh;
    FvMatrixG M;
    M.init(m, true);
    for (int f = 0; f < m.nInternal; ++f) {
        double w = R.phi[f] >= 0.0 ? 1.0 : 0.0;
        M.lower[f] = -w * R.phi[f];
        M.upper[f] = M.lower[f] + R.phi[f];
    }
    M.negSumDiag(m);
    for (size_t p = 0; p < m.patches.size(); ++p) {
        for (int i = 0; i < m.patches[p].size; ++i) {
            double ph = R.phiB[p][i];
            double w = ph >= 0.0 ? 1.0 : 0.0;
            switch (R.bcT[p]) {
            case BcT::FixedValue:
                M.internalCoeffs[p][i] = ph * 0.0;
                M.boundaryCoeffs[p][i] = -ph * R.Tb[p][i];
                break;
            case BcT::Coupled:
                M.internalCoeffs[p][i] = ph * w;
                M.boundaryCoeffs[p][i] = -ph * (1.0 - w);
                break;
            case BcT::ZeroGradient:
                M.internalCoeffs[p][i] = ph * 1.0;
                M.boundaryCoeffs[p][i] = -ph * 0.0;
                break;
            case BcT::PartDirichlet:
                M.internalCoeffs[p][i] = ph * 0.0;
                M.boundaryCoeffs[p][i] = -ph * 1.0;
                break;
            case BcT::PartNeumann:
                M.internalCoeffs[p][i] = ph * 1.0;
                M.boundaryCoeffs[p][i] = -ph * (1.0 / (1.0 / m.nDist(m.patches[p].start + i)));
                break;
            case BcT::Empty:
                break;
            }
        }
    }
    return M;
}


// The plane of a coupled patch: origin, unit normal (of the total area
// vector) and two in-plane unit vectors.
struct Frame {
    Vec o, n, e1, e2;
};

using P2 = std::array<double, 2>;

// A face projected into the interface frame: counter-clockwise vertices, its
// bounding box and its area.
struct Poly {
    std::vector<P2> v;
    double lo[2], hi[2], area;
};

inline double cross2(const P2 &a, const P2 &b, const P2 &c)   // (b - a) x (c - a)
{
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
}

double polyArea(const std::vector<P2> &v)
{
    double s = 0.0;
    for (size_t i = 0; i < v.size(); ++i) {
        const P2 &p = v[i], &q = v[(i + 1) % v.size()];
        s += p[0] * q[1] - q[0] * p[1];
    }
    return 0.5 * s;
}

Frame patchFrame(const PolyMesh &m, const Patch &P)
{
    Vec sum{0, 0, 0}, c{0, 0, 0};
    double area = 0.0;
    for (int i = 0; i < P.size; ++i) {
        const int f = P.start + i;
        sum = sum + m.Sf[f];
        area += m.magSf[f];
        c = c + m.magSf[f] * m.Cf[f];
    }
    if (!(area > 0.0)) throw FoamError("coupled patch '" + P.name + "' of region '" + m.name + "' has no faces");
    Frame F;
    F.o = (1.0 / area) * c;
    F.n = (1.0 / mag(sum)) * sum;
    // every face parallel to the plane and on it
    if (mag(sum) < (1.0 - 1e-9) * area)
        throw FoamError("coupled patch '" + P.name + "' of region '" + m.name + "' is not planar");
    for (int i = 0; i < P.size; ++i) {
        const int f = P.start + i;
        if (std::fabs(dot(F.n, m.Cf[f] - F.o)) > 1e-6 * std::sqrt(m.magSf[f]))
            throw FoamError("coupled patch '" + P.name + "' of region '" + m.name + "' is not planar");
    }
    int a = 0;
    for (int q = 1; q < 3; ++q)
        if (std::fabs(F.n[q]) < std::fabs(F.n[a])) a = q;
    Vec u{0, 0, 0};
    u[a] = 1.0;
    Vec e1 = u - dot(u, F.n) * F.n;
    F.e1 = (1.0 / mag(e1)) * e1;
    F.e2 = cross(F.n, F.e1);
    return F;
}

// the patch's faces in frame F (counter-clockwise; faces must be convex)
std::vector<Poly> patchPolys(const PolyMesh &m, const Patch &P, const Frame &F)
{
    std::vector<Poly> out(P.size);
    for (int i = 0; i < P.size; ++i) {
        const int f = P.start + i;
        Poly &q = out[i];
        for (int j = m.faceOff[f]; j < m.faceOff[f + 1]; ++j) {
            const Vec d = m.points[m.faceVerts[j]] - F.o;
            q.v.push_back(P2{dot(d, F.e1), dot(d, F.e2)});
        }
        if (polyArea(q.v) < 0.0) std::reverse(q.v.begin(), q.v.end());
        q.area = polyArea(q.v);
        const size_t n = q.v.size();
        for (size_t k = 0; k < n; ++k)
            if (cross2(q.v[k], q.v[(k + 1) % n], q.v[(k + 2) % n]) < -1e-12 * q.area)
                throw FoamError("coupled patch '" + P.name + "' of region '" + m.name + "': face " + std::to_string(i) +
                                " is not convex");
        q.lo[0] = q.lo[1] = 1e300;
        q.hi[0] = q.hi[1] = -1e300;
        for (const P2 &p : q.v)
            for (int d = 0; d < 2; ++d) {
                q.lo[d] = std::fmin(q.lo[d], p[d]);
                q.hi[d] = std::fmax(q.hi[d], p[d]);
            }
    }
    return out;
}

// Area of the intersection of two convex counter-clockwise polygons
// (Sutherland-Hodgman: the first clipped by each edge of the second).
double overlapArea(const std::vector<P2> &a, const std::vector<P2> &b)
{
    std::vector<P2> cur = a, nxt;
    for (size_t i = 0; i < b.size() && !cur.empty(); ++i) {
        const P2 &c0 = b[i], &c1 = b[(i + 1) % b.size()];
        nxt.clear();
        for (size_t j = 0; j < cur.size(); ++j) {
            const P2 &s0 = cur[j], &s1 = cur[(j + 1) % cur.size()];
            const double d0 = cross2(c0, c1, s0), d1 = cross2(c0, c1, s1);
            if (d0 >= 0.0) nxt.push_back(s0);
            if ((d0 >= 0.0) != (d1 >= 0.0)) {
                const double t = d0 / (d0 - d1);
                nxt.push_back(P2{s0[0] + t * (s1[0] - s0[0]), s0[1] + t * (s1[1] - s0[1])});
            }
        }
        cur.swap(nxt);
    }
    return cur.size() < 3 ? 0.0 : std::fabs(polyArea(cur));
}

// GGI weights of side A's faces gathering side B's faces: overlap areas of
// the faces in the interface plane, rescaled to sum 1 per covered face;
// overlaps below 1e-10 of the smaller face are dropped.  A bucket grid over
// B's bounding boxes keeps the search local.
void ggiWeights(const std::vector<Poly> &A, const std::vector<Poly> &B, InterfaceSide &S)
{
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300}, size = 0.0;
    for (const Poly &r : B) {
        for (int d = 0; d < 2; ++d) {
            lo[d] = std::fmin(lo[d], r.lo[d]);
            hi[d] = std::fmax(hi[d], r.hi[d]);
        }
        size += std::sqrt(r.area);
    }
    size = B.empty() ? 1.0 : 2.0 * size / (double)B.size();
    int nb[2];
    for (int d = 0; d < 2; ++d) nb[d] = std::max(1, std::min(4096, (int)std::ceil((hi[d] - lo[d]) / size)));
    auto bucket = [&](double x, int d) {
        int i = (int)std::floor((x - lo[d]) / (hi[d] - lo[d] + 1e-300) * nb[d]);
        return std::max(0, std::min(nb[d] - 1, i));
    };
    std::vector<std::vector<int>> grid((size_t)nb[0] * nb[1]);
    for (int j = 0; j < (int)B.size(); ++j)
        for (int y = bucket(B[j].lo[1], 1); y <= bucket(B[j].hi[1], 1); ++y)
            for (int x = bucket(B[j].lo[0], 0); x <= bucket(B[j].hi[0], 0); ++x) grid[(size_t)y * nb[0] + x].push_back(j);
    S.off.assign(A.size() + 1, 0);
    S.idx.clear();
    S.w.clear();
    std::vector<int> cand;
    for (size_t i = 0; i < A.size(); ++i) {
        const Poly &a = A[i];
        S.off[i] = (int32_t)S.w.size();
        cand.clear();
        for (int y = bucket(a.lo[1], 1); y <= bucket(a.hi[1], 1); ++y)
            for (int x = bucket(a.lo[0], 0); x <= bucket(a.hi[0], 0); ++x)
                for (int j : grid[(size_t)y * nb[0] + x]) cand.push_back(j);
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        const size_t p0 = S.w.size();
        double sum = 0.0;
        for (int j : cand) {
            const Poly &b = B[j];
            if (b.lo[0] >= a.hi[0] || b.hi[0] <= a.lo[0] || b.lo[1] >= a.hi[1] || b.hi[1] <= a.lo[1]) continue;
            const double ar = overlapArea(a.v, b.v);
            if (ar <= 1e-10 * std::fmin(a.area, b.area)) continue;
            S.idx.push_back(j);
            S.w.push_back(ar);
            sum += ar;
        }
        if (sum > 1e-12 * a.area) {
            for (size_t q = p0; q < S.w.size(); ++q) S.w[q] = S.w[q] / sum;
        } else {
            S.idx.resize(p0);
            S.w.resize(p0);
        }
    }
    S.off[A.size()] = (int32_t)S.w.size();
}

// calcThermalDiffusivity with radiation off, on general geometry: the owner
// side's normal distance, the GGI-interpolated neighbour distance and k.
void calcThermalDiffusivity(FoamRegion &own, int pOwn, const FoamRegion &nei, int pNei, InterfaceSide &S)
{
    const PolyMesh &mo = own.mesh, &mn = nei.mesh;
    const Patch &Po = mo.patches[pOwn], &Pn = mn.patches[pNei];
    std::vector<double> &kface = own.coupledK[pOwn];
    std::vector<double> &delta = own.coupledDelta[pOwn];
    kface.resize(Po.size);
    delta.resize(Po.size);
    S.fc.resize(Po.size);
    S.magSf.resize(Po.size);
    S.kOwn.resize(Po.size);
    S.kNei.resize(Po.size);
    for (int i = 0; i < Po.size; ++i) {
        const int f = Po.start + i;
        const int cO = mo.owner[f];
        S.fc[i] = cO;
        S.magSf[i] = mo.magSf[f];
        double dOwn = mo.nDist(f), dNei = 0.0, fNei = 0.0;
        for (int p = S.off[i]; p < S.off[i + 1]; ++p) {
            const int g = Pn.start + S.idx[p];
            dNei += S.w[p] * mn.nDist(g);
            fNei += S.w[p] * nei.k[mn.owner[g]];
        }
        if (S.off[i + 1] == S.off[i])
            throw FoamError("coupled patch '" + Po.name + "' of region '" + mo.name + "': face " +
                            std::to_string(i) + " has no partner face");
        double fOwn = own.k[cO];
        double magDelta = dOwn + dNei;       // magLongDelta
        double wP = dNei / magDelta;         // patch weights
        delta[i] = 1.0 / magDelta;           // patch deltaCoeffs
        double kOwn = fOwn / (1.0 - wP) / magDelta;
        double kNei = fNei / wP / magDelta;
        S.kOwn[i] = kOwn;
        S.kNei[i] = kNei;
        double k = kOwn * kNei;
        k /= delta[i] * (kOwn + kNei);
        double kHarm = 1.0 / (wP / fOwn + (1.0 - wP) / fNei);
        kface[i] = std::fmax(std::fmin(k, 100.0 * kHarm), 0.01 * kHarm);
    }
}

struct Options {
    std::string fieldDir;                    // under 0/<region>/ ("" = directly)
    std::string interfaceFile;               // under constant/
    std::string ddt;                         // "" = from fvSchemes
    std::string div;                         // "" = div(phi,T) from fvSchemes; "upwind" = assemble upwind
    double deltaT = 0.0;                     // 0 = from controlDict
};

void readRegion(const std::string &dir, const std::string &name, const std::string &type, const Options &opt,
                FoamRegion &R)
{
    PolyMesh &m = R.mesh;
    m.name = name;
    readPolyMesh(dir + "/constant/" + name + "/polyMesh", m);
    computeGeometry(m);
    R.transport = type == "transportTemperature";

    std::string tp = dir + "/constant/" + name + "/transportProperties";
    Dict props = parseDict(readFile(tp));
    const double rho = property(props, "rho", tp);
    R.rhoCp = rho * property(props, R.transport ? "cp" : "cv", tp);

    std::string fdir = dir + "/0/" + name + "/";
    if (!opt.fieldDir.empty()) fdir += opt.fieldDir + "/";
    else if (!fileExists(fdir + "T") && fileExists(fdir + "orig/monolithic/T")) fdir += "orig/monolithic/";
    FieldFile T = readField(fdir + "T", m, 1);
    R.T = T.internal;
    const size_t nP = m.patches.size();
    R.bcT.resize(nP);
    R.Tb.resize(nP);
    R.kb.resize(nP);
    R.accType.assign(nP, "fixed");
    R.relax.assign(nP, 1.0);
    R.reuse.assign(nP, 0);
    for (size_t p = 0; p < nP; ++p) {
        const Patch &P = m.patches[p];
        const PatchBC &bc = T.patches.at(P.name);
        if (bc.type == "fixedValue") {
            R.bcT[p] = BcT::FixedValue;
            if ((int)bc.value.size() != P.size) throw FoamError(fdir + "T: fixedValue patch '" + P.name + "' has no value");
            R.Tb[p] = bc.value;
        } else if (bc.type == "zeroGradient") {
            R.bcT[p] = BcT::ZeroGradient;
        } else if (bc.type == "empty") {
            R.bcT[p] = BcT::Empty;
        } else if (bc.type == "monolithicTemperature") {
            R.bcT[p] = BcT::Coupled;
        } else if (bc.type == "regionCoupledTemperatureJump" || bc.type == "regionCoupledHeatFlux") {
            R.bcT[p] = bc.type == "regionCoupledHeatFlux" ? BcT::PartNeumann : BcT::PartDirichlet;
            if ((int)bc.value.size() != P.size) throw FoamError(fdir + "T: patch '" + P.name + "' has no value");
            if (bc.accType != "fixed" && bc.accType != "aitken" && bc.accType != "IQN-ILS")
                throw FoamError(fdir + "T: patch '" + P.name + "': acceleration '" + bc.accType +
                                "' is not supported (fixed, aitken, IQN-ILS)");
            R.Tb[p] = bc.value;
            R.accType[p] = bc.accType;
            R.relax[p] = bc.relax;
            R.reuse[p] = bc.reuse;
        } else {
            throw FoamError(fdir + "T: patch '" + P.name + "': unsupported type '" + bc.type + "'");
        }
        if ((R.bcT[p] == BcT::Empty) != (P.type == "empty"))
            throw FoamError(fdir + "T: patch '" + P.name + "' is empty in only one of mesh and field");
    }
    // conductivity: 0/<region>/k, else transportProperties k (uniform)
    R.kFixed.assign(nP, 0);
    if (fileExists(fdir + "k")) {
        FieldFile K = readField(fdir + "k", m, 1);
        R.k = K.internal;
        for (size_t p = 0; p < nP; ++p) {
            const Patch &P = m.patches[p];
            const PatchBC &bc = K.patches.at(P.name);
            R.kFixed[p] = bc.type == "fixedValue";
            R.kb[p].resize(P.size);
            for (int i = 0; i < P.size; ++i)
                R.kb[p][i] = bc.type == "fixedValue" ? bc.value.at(i) : R.k[m.owner[P.start + i]];
        }
    } else {
        R.k.assign(m.nCells, property(props, "k", tp));
        for (size_t p = 0; p < nP; ++p) {
            const Patch &P = m.patches[p];
            R.kb[p].resize(P.size);
            for (int i = 0; i < P.size; ++i) R.kb[p][i] = R.k[m.owner[P.start + i]];
        }
    }
    for (double x : R.k)
        if (!(x > 0.0)) throw FoamError("region '" + name + "': conductivity must be positive");

    // flux: phi = linearInterpolate(U) & Sf, boundary U_b & Sf
    R.phiB.resize(nP);
    R.bcU.assign(nP, kUZero);
    R.Ub.assign(nP, {});
    if (R.transport) {
        std::string up = dir + "/0/" + name + "/U";
        FieldFile U = readField(up, m, 3);
        R.U = U.internal;
        auto Uc = [&](int c) { return Vec{U.internal[3 * c], U.internal[3 * c + 1], U.internal[3 * c + 2]}; };
        R.phi.resize(m.nInternal);
        for (int f = 0; f < m.nInternal; ++f) {
            Vec uf = m.w[f] * Uc(m.owner[f]) + (1.0 - m.w[f]) * Uc(m.neighbour[f]);
            R.phi[f] = dot(uf, m.Sf[f]);
        }
        for (size_t p = 0; p < nP; ++p) {
            const Patch &P = m.patches[p];
            const PatchBC &bc = U.patches.at(P.name);
            R.phiB[p].assign(P.size, 0.0);
            R.bcU[p] = bc.type == "fixedValue" ? kUFixedValue : bc.type == "zeroGradient" ? kUZeroGradient
                     : bc.type == "slip" ? kUSlip : kUZero;
            R.Ub[p].assign(3 * (size_t)P.size, 0.0);
            for (int i = 0; i < P.size; ++i) {
                const int f = P.start + i;
                Vec ub{0, 0, 0};
                if (bc.type == "fixedValue") {
                    ub = Vec{bc.value.at(3 * i), bc.value.at(3 * i + 1), bc.value.at(3 * i + 2)};
                    for (int a = 0; a < 3; ++a) R.Ub[p][3 * i + a] = ub[a];
                } else if (bc.type == "zeroGradient") ub = Uc(m.owner[f]);
                else if (bc.type == "slip") {
                    Vec n = (1.0 / m.magSf[f]) * m.Sf[f], u = Uc(m.owner[f]);
                    ub = u - dot(u, n) * n;
                } else if (bc.type != "noSlip" && bc.type != "empty")
                    throw FoamError(up + ": patch '" + P.name + "': unsupported type '" + bc.type + "'");
                R.phiB[p][i] = dot(ub, m.Sf[f]);
            }
        }
    } else {
        for (size_t p = 0; p < nP; ++p) R.phiB[p].assign(m.patches[p].size, 0.0);
    }
}

// the region's TEqn as setCoupledEqns builds it, without the ddt term (added
// by the caller as diag += a, source += a*T_old with a = rho*c*V/dt)
FvMatrixG setCoupledEqns(const FoamRegion &R)
{
    const PolyMesh &m = R.mesh;
    FvMatrixG lap = laplacian(R);
    FvMatrixG E;
    E.init(m, R.transport);
    if (R.transport) {
        // rho*cp*(fvm::ddt(T) + fvm::div(phi,T)) == fvm::laplacian(k,T)
        FvMatrixG A = divUpwind(R);
        for (int c = 0; c < m.nCells; ++c) {
            E.diag[c] = R.rhoCp * A.diag[c] - lap.diag[c];
            E.source[c] = R.rhoCp * A.source[c] - lap.source[c];
        }
        for (int f = 0; f < m.nInternal; ++f) {
            E.upper[f] = R.rhoCp * A.upper[f] - lap.upper[f];
            E.lower[f] = R.rhoCp * A.lower[f] - lap.upper[f];
        }
        for (size_t p = 0; p < m.patches.size(); ++p)
            for (int i = 0; i < m.patches[p].size; ++i) {
                E.internalCoeffs[p][i] = R.rhoCp * A.internalCoeffs[p][i] - lap.internalCoeffs[p][i];
                E.boundaryCoeffs[p][i] = R.rhoCp * A.boundaryCoeffs[p][i] - lap.boundaryCoeffs[p][i];
            }
    } else {
        // fvm::ddt(rho*cv, T) == fvm::laplacian(k, T)
        for (int c = 0; c < m.nCells; ++c) {
            E.diag[c] = 0.0 - lap.diag[c];
            E.source[c] = 0.0 - lap.source[c];
        }
        for (int f = 0; f < m.nInternal; ++f) E.upper[f] = 0.0 - lap.upper[f];
        for (size_t p = 0; p < m.patches.size(); ++p)
            for (int i = 0; i < m.patches[p].size; ++i) {
                E.internalCoeffs[p][i] = 0.0 - lap.internalCoeffs[p][i];
                E.boundaryCoeffs[p][i] = 0.0 - lap.boundaryCoeffs[p][i];
            }
    }
    return E;
}

// The model the device re-assembles from (geometry, boundary conditions,
// fields), in the loop orders of laplacian / divUpwind / setCoupledEqns and
// calcThermalDiffusivity above.
std::shared_ptr<const FvModel> buildFvModel(const std::vector<FoamRegion> &R, const std::vector<double> &rDeltaT,
                                            const std::vector<std::array<int, 4>> &sides, const CoupledSystem &sys)
{
    auto M = std::make_shared<FvModel>();
    M->regions.resize(R.size());
    std::vector<std::vector<int>> patchBase(R.size());
    for (size_t r = 0; r < R.size(); ++r) {
        const FoamRegion &rc = R[r];
        const PolyMesh &m = rc.mesh;
        FvRegionModel &F = M->regions[r];
        F.transport = rc.transport;
        F.rhoCp = rc.rhoCp;
        F.rDeltaT = rDeltaT[r];
        F.w = m.w;
        F.deltaCoeffs = m.deltaCoeffs;
        F.magSf.assign(m.magSf.begin(), m.magSf.begin() + m.nInternal);
        F.Sf.resize(3 * (size_t)m.nInternal);
        for (int f = 0; f < m.nInternal; ++f)
            for (int a = 0; a < 3; ++a) F.Sf[3 * (size_t)f + a] = m.Sf[f][a];
        F.V = m.V;
        for (size_t p = 0; p < m.patches.size(); ++p) {
            const Patch &P = m.patches[p];
            patchBase[r].push_back((int)F.bOwner.size());
            for (int i = 0; i < P.size; ++i) {
                const int f = P.start + i;
                F.bOwner.push_back(m.owner[f]);
                F.bT.push_back((int8_t)rc.bcT[p]);
                F.bU.push_back(rc.transport ? rc.bcU[p] : (int8_t)kUZero);
                F.bKFixed.push_back(rc.kFixed[p]);
                F.bNDist.push_back(m.nDist(f));
                F.bMagSf.push_back(m.magSf[f]);
                for (int a = 0; a < 3; ++a) F.bSf.push_back(m.Sf[f][a]);
                F.bTb.push_back(rc.bcT[p] == BcT::FixedValue ? rc.Tb[p][i] : 0.0);
                for (int a = 0; a < 3; ++a)
                    F.bUb.push_back(rc.transport && rc.bcU[p] == kUFixedValue ? rc.Ub[p][3 * i + a] : 0.0);
                F.bKb.push_back(rc.kb[p][i]);
                F.bSide.push_back(-1);
            }
        }
        F.k = rc.k;
        F.U = rc.U;
        F.T = rc.T;
    }
    M->sides.resize(2 * sides.size());
    int sideFace = 0;
    for (size_t i = 0; i < sides.size(); ++i)
        for (int q = 0; q < 2; ++q) {
            const int rO = sides[i][2 * q], pO = sides[i][2 * q + 1];
            const int rN = sides[i][2 * (1 - q)], pN = sides[i][2 * (1 - q) + 1];
            const PolyMesh &mo = R[rO].mesh, &mn = R[rN].mesh;
            const InterfaceSide &S = q ? sys.interfaces[i].b : sys.interfaces[i].a;
            FvSideModel &D = M->sides[2 * i + q];
            const Patch &Po = mo.patches[pO], &Pn = mn.patches[pN];
            for (int j = 0; j < Po.size; ++j) {
                D.nDist.push_back(mo.nDist(Po.start + j));
                D.bFace.push_back(patchBase[rO][pO] + j);
                M->regions[rO].bSide[patchBase[rO][pO] + j] = sideFace + j;
            }
            for (size_t p = 0; p < S.idx.size(); ++p) {
                const int g = Pn.start + S.idx[p];
                D.pairNDist.push_back(mn.nDist(g));
                D.pairCell.push_back(mn.owner[g]);
            }
            sideFace += Po.size;
        }
    return M;
}

// "regions ( (fluid (icoFluid transportTemperature)) (solid (conductTemperature)) );"
std::vector<std::pair<std::string, std::string>> regionList(const std::string &dir)
{
    std::string f = dir + "/constant/multiRegionProperties";
    Dict d = parseDict(readFile(f));
    std::vector<std::pair<std::string, std::string>> out;
    for (const Value &r : d.list("regions")) {
        if (!r.isList() || r.items.size() < 2 || !r.items[0].isWord()) throw FoamError(f + ": bad regions entry");
        std::string type;
        const Value &types = r.items[1];
        auto consider = [&](const Value &v) {
            if (v.isWord() && (v.asWord() == "transportTemperature" || v.asWord() == "conductTemperature"))
                type = v.asWord();
        };
        if (types.isList())
            for (const Value &v : types.items) consider(v);
        else
            consider(types);
        if (type.empty())
            throw FoamError(f + ": region '" + r.items[0].asWord() +
                            "' has no temperature region type (transportTemperature / conductTemperature)");
        out.push_back({r.items[0].asWord(), type});
    }
    return out;
}

}  // namespace

// The div(phi,T) entry of an fvSchemes dictionary (else its default), as
// words joined by spaces; "" when the file names no convection scheme.
static std::string divScheme(const Dict &fvSchemes)
{
    if (!fvSchemes.isDict("divSchemes")) return "";
    const Dict &d = fvSchemes.subDict("divSchemes");
    auto words = [](const std::vector<Value> &toks, size_t from) {
        std::string out;
        for (size_t i = from; i < toks.size(); ++i) {
            if (!toks[i].isWord()) continue;
            if (!out.empty()) out += " ";
            out += toks[i].text;
        }
        return out;
    };
    std::string def, named;
    for (const Entry &e : d.entries) {
        if (e.key == "default") def = words(e.tokens, 0);
        // "div(phi,T)" reads as the word div followed by the list (phi,T)
        if (e.key == "div" && !e.tokens.empty() && e.tokens[0].isList() && e.tokens[0].items.size() == 1 &&
            e.tokens[0].items[0].text == "phi,T")
            named = words(e.tokens, 1);
    }
    return named.empty() ? def : named;
}

CoupledSystem readFoamCase(const std::string &dir, const std::string &optionsText, FoamCaseInfo *info)
{
    Options opt;
    {
        Dict o = parseDict(optionsText);
        opt.fieldDir = o.wordOrDefault("fieldDir", "");
        opt.interfaceFile = o.wordOrDefault("interfaceProperties", "");
        opt.ddt = o.wordOrDefault("ddt", "");
        opt.div = o.wordOrDefault("div", "");
        if (!opt.div.empty() && opt.div != "upwind") throw FoamError("option div must be upwind, not '" + opt.div + "'");
        opt.deltaT = o.numberOrDefault("deltaT", 0.0);
    }
    auto regions = regionList(dir);
    // time step: controlDict deltaT, ddt scheme of the first region's fvSchemes
    double deltaT = opt.deltaT;
    if (deltaT <= 0.0) deltaT = parseDict(readFile(dir + "/system/controlDict")).number("deltaT");
    std::vector<FoamRegion> R(regions.size());
    std::vector<double> rDeltaT(regions.size(), 0.0);
    std::map<std::string, int> byName;
    for (size_t r = 0; r < regions.size(); ++r) {
        readRegion(dir, regions[r].first, regions[r].second, opt, R[r]);
        byName[regions[r].first] = (int)r;
        std::string ddt = opt.ddt;
        if (ddt.empty()) {
            std::string fs = dir + "/system/" + regions[r].first + "/fvSchemes";
            if (!fileExists(fs)) fs = dir + "/system/fvSchemes";
            ddt = parseDict(readFile(fs)).subDict("ddtSchemes").word("default");
        }
        if (ddt == "steadyState") rDeltaT[r] = 0.0;
        else if (ddt == "Euler" || ddt == "backward") rDeltaT[r] = 1.0 / deltaT;   // backward: first step
        else throw FoamError("ddt scheme '" + ddt + "' is not supported (steadyState, Euler, backward)");
        // convection (transportTemperature.C:133-139, fvm::div(phi, T)): the
        // upwind part is assembled; a scheme with an explicit correction
        // (linearUpwind, linear, ...) is refused unless the caller asks for
        // upwind in its place
        if (R[r].transport && opt.div.empty()) {
            std::string fs = dir + "/system/" + regions[r].first + "/fvSchemes";
            if (!fileExists(fs)) fs = dir + "/system/fvSchemes";
            const std::string sch = divScheme(parseDict(readFile(fs)));
            if (!sch.empty() && sch != "Gauss upwind" && sch != "bounded Gauss upwind")
                throw FoamError("region '" + regions[r].first + "': div(phi,T) scheme '" + sch +
                                "' is not assembled (Gauss upwind only; the option div upwind assembles upwind "
                                "convection in its place)");
        }
    }

    // monolithic interfaces
    std::string ifile = dir + "/constant/" + (opt.interfaceFile.empty() ? "regionInterfaceProperties" : opt.interfaceFile);
    if (opt.interfaceFile.empty() && !fileExists(ifile)) ifile = dir + "/constant/regionInterfacePropertiesMonolithic";
    CoupledSystem sys;
    std::vector<std::array<int, 4>> sides;   // region A, patch A, region B, patch B
    if (fileExists(ifile)) {
        Dict d = parseDict(readFile(ifile));
        if (d.found("monolithicCoupledPatches"))
            for (const Value &it : d.list("monolithicCoupledPatches")) {
                if (!it.isDict()) continue;
                const auto &pp = it.dict->list("coupledPatchPair");
                if (pp.size() != 2 || !pp[0].isList() || !pp[1].isList())
                    throw FoamError(ifile + ": coupledPatchPair must be ((regionA patchA) (regionB patchB))");
                int reg[2], pat[2];
                for (int s = 0; s < 2; ++s) {
                    const std::string &rn = pp[s].items.at(0).asWord(), &pn = pp[s].items.at(1).asWord();
                    auto f = byName.find(rn);
                    if (f == byName.end()) throw FoamError(ifile + ": unknown region '" + rn + "'");
                    reg[s] = f->second;
                    pat[s] = -1;
                    const PolyMesh &m = R[reg[s]].mesh;
                    for (size_t p = 0; p < m.patches.size(); ++p)
                        if (m.patches[p].name == pn) pat[s] = (int)p;
                    if (pat[s] < 0) throw FoamError(ifile + ": region '" + rn + "' has no patch '" + pn + "'");
                    if (R[reg[s]].bcT[pat[s]] != BcT::Coupled)
                        throw FoamError(ifile + ": patch '" + pn + "' of '" + rn + "' is not monolithicTemperature");
                }
                // both sides in the frame of side A's plane (side B faces the
                // other way and must lie in the same plane)
                const Patch &PA = R[reg[0]].mesh.patches[pat[0]], &PB = R[reg[1]].mesh.patches[pat[1]];
                const Frame FA = patchFrame(R[reg[0]].mesh, PA), FB = patchFrame(R[reg[1]].mesh, PB);
                if (dot(FA.n, FB.n) > -(1.0 - 1e-9) ||
                    std::fabs(dot(FA.n, FB.o - FA.o)) > 1e-6 * std::sqrt(R[reg[0]].mesh.magSf[PA.start]))
                    throw FoamError(ifile + ": coupled patches of '" + it.dict->name + "' are not in one plane");
                auto rA = patchPolys(R[reg[0]].mesh, PA, FA);
                auto rB = patchPolys(R[reg[1]].mesh, PB, FA);
                CoupledInterface I;
                I.name = it.dict->name;
                I.a.region = reg[0];
                I.b.region = reg[1];
                ggiWeights(rA, rB, I.a);
                ggiWeights(rB, rA, I.b);
                calcThermalDiffusivity(R[reg[0]], pat[0], R[reg[1]], pat[1], I.a);
                calcThermalDiffusivity(R[reg[1]], pat[1], R[reg[0]], pat[0], I.b);
                sys.interfaces.push_back(std::move(I));
                sides.push_back({reg[0], pat[0], reg[1], pat[1]});
            }
    }
    // partitioned interfaces (heatTransferInterface pairs) and DNA controls
    std::shared_ptr<PartitionedModel> partModel;
    std::vector<std::array<int, 2>> partPatch;   // per pair: Dirichlet patch, Neumann patch
    {
        bool anyPart = false;
        for (auto &rc : R)
            for (BcT t : rc.bcT) anyPart = anyPart || t == BcT::PartDirichlet || t == BcT::PartNeumann;
        if (anyPart) {
            std::string pfile = dir + "/constant/" + (opt.interfaceFile.empty() ? "regionInterfaceProperties" : opt.interfaceFile);
            bool has = fileExists(pfile) && parseDict(readFile(pfile)).found("partitionedCoupledPatches") &&
                       !parseDict(readFile(pfile)).list("partitionedCoupledPatches").empty();
            if (!has && opt.interfaceFile.empty()) pfile = dir + "/constant/regionInterfacePropertiesPartitioned";
            Dict d = parseDict(readFile(pfile));
            partModel = std::make_shared<PartitionedModel>();
            partModel->sourceBase.resize(R.size());
            for (const Value &it : d.list("partitionedCoupledPatches")) {
                if (!it.isDict()) continue;
                const Dict &pd = *it.dict;
                const std::string itype = pd.wordOrDefault("interfaceType", "heatTransferInterface");
                if (itype != "heatTransferInterface")
                    throw FoamError(pfile + ": interface type '" + itype + "' is not supported (heatTransferInterface)");
                const auto &pp = pd.list("coupledPatchPair");
                if (pp.size() != 2 || !pp[0].isList() || !pp[1].isList())
                    throw FoamError(pfile + ": coupledPatchPair must be ((regionA patchA) (regionB patchB))");
                int reg[2], pat[2];
                for (int q = 0; q < 2; ++q) {
                    const std::string &rn = pp[q].items.at(0).asWord(), &pn = pp[q].items.at(1).asWord();
                    auto f = byName.find(rn);
                    if (f == byName.end()) throw FoamError(pfile + ": unknown region '" + rn + "'");
                    reg[q] = f->second;
                    pat[q] = -1;
                    const PolyMesh &m = R[reg[q]].mesh;
                    for (size_t p = 0; p < m.patches.size(); ++p)
                        if (m.patches[p].name == pn) pat[q] = (int)p;
                    if (pat[q] < 0) throw FoamError(pfile + ": region '" + rn + "' has no patch '" + pn + "'");
                }
                const BcT t0 = R[reg[0]].bcT[pat[0]], t1 = R[reg[1]].bcT[pat[1]];
                if (!((t0 == BcT::PartDirichlet && t1 == BcT::PartNeumann) ||
                      (t0 == BcT::PartNeumann && t1 == BcT::PartDirichlet)))
                    throw FoamError(pfile + ": interface '" + pd.name + "' needs regionCoupledTemperatureJump on one "
                                    "side and regionCoupledHeatFlux on the other");
                const int qd = t0 == BcT::PartDirichlet ? 0 : 1, qn = 1 - qd;
                PartitionedPair P;
                P.name = pd.name;
                P.K = pd.isDict("heatTransferInterfaceCoeffs")
                          ? property(pd.subDict("heatTransferInterfaceCoeffs"), "K", pfile) : 1.0;
                const Patch &PD = R[reg[qd]].mesh.patches[pat[qd]], &PN = R[reg[qn]].mesh.patches[pat[qn]];
                const Frame FD = patchFrame(R[reg[qd]].mesh, PD), FN = patchFrame(R[reg[qn]].mesh, PN);
                if (dot(FD.n, FN.n) > -(1.0 - 1e-9) ||
                    std::fabs(dot(FD.n, FN.o - FD.o)) > 1e-6 * std::sqrt(R[reg[qd]].mesh.magSf[PD.start]))
                    throw FoamError(pfile + ": coupled patches of '" + pd.name + "' are not in one plane");
                auto pD = patchPolys(R[reg[qd]].mesh, PD, FD);
                auto pN = patchPolys(R[reg[qn]].mesh, PN, FD);
                InterfaceSide gD, gN;
                ggiWeights(pD, pN, gD);
                ggiWeights(pN, pD, gN);
                auto fill = [&](PartitionedSide &S, int rr, int pa, const InterfaceSide &g) {
                    const FoamRegion &rc = R[rr];
                    const PolyMesh &m = rc.mesh;
                    const Patch &Pt = m.patches[pa];
                    S.region = rr;
                    for (int i = 0; i < Pt.size; ++i) {
                        const int f = Pt.start + i;
                        S.fc.push_back(m.owner[f]);
                        S.kb.push_back(rc.kb[pa][i]);
                        S.dc.push_back(1.0 / m.nDist(f));
                        S.value0.push_back(rc.Tb[pa][i]);
                        if (g.off[i + 1] == g.off[i])
                            throw FoamError(pfile + ": face " + std::to_string(i) + " of patch '" + Pt.name +
                                            "' has no partner face");
                    }
                    S.off = g.off;
                    S.idx = g.idx;
                    S.w = g.w;
                    S.accType = rc.accType[pa] == "aitken" ? 1 : rc.accType[pa] == "IQN-ILS" ? 2 : 0;
                    S.relax = rc.relax[pa];
                    S.reuse = rc.reuse[pa];
                };
                fill(P.dir, reg[qd], pat[qd], gD);
                fill(P.neu, reg[qn], pat[qn], gN);
                partModel->pairs.push_back(std::move(P));
                partPatch.push_back({pat[qd], pat[qn]});
            }
            for (size_t r = 0; r < R.size(); ++r)
                for (size_t p = 0; p < R[r].bcT.size(); ++p)
                    if ((R[r].bcT[p] == BcT::PartDirichlet || R[r].bcT[p] == BcT::PartNeumann) &&
                        R[r].mesh.patches[p].size > 0) {
                        bool found = false;
                        for (size_t i = 0; i < partPatch.size(); ++i)
                            for (int q = 0; q < 2; ++q)
                                found = found || ((q ? partModel->pairs[i].neu : partModel->pairs[i].dir).region ==
                                                      (int)r && partPatch[i][q] == (int)p);
                        if (!found)
                            throw FoamError("partitioned patch '" + R[r].mesh.patches[p].name + "' of region '" +
                                            R[r].mesh.name + "' has no partitionedCoupledPatches entry");
                    }
            // DNA controls of T (constant/multiRegionProperties, dnaControl::read)
            Dict mrp = parseDict(readFile(dir + "/constant/multiRegionProperties"));
            if (mrp.isDict("DNA") && mrp.subDict("DNA").isDict("T")) {
                const Dict &dt = mrp.subDict("DNA").subDict("T");
                partModel->maxCoupleIter = (int)dt.numberOrDefault("maxCoupleIter", 1);
                if (dt.isDict("residualControl")) {
                    const Dict &rcd = dt.subDict("residualControl");
                    if (!rcd.found("maxJumpRes") || !rcd.found("maxFluxRes"))
                        throw FoamError("DNA residualControl of T: maxJumpRes and maxFluxRes are required");
                    partModel->maxJumpRes = rcd.number("maxJumpRes");
                    partModel->maxFluxRes = rcd.number("maxFluxRes");
                }
            }
        }
    }
    for (size_t r = 0; r < R.size(); ++r)
        for (size_t p = 0; p < R[r].bcT.size(); ++p)
            if (R[r].bcT[p] == BcT::Coupled && !R[r].coupledK.count((int)p) && R[r].mesh.patches[p].size > 0)
                throw FoamError("coupled patch '" + R[r].mesh.patches[p].name + "' of region '" + R[r].mesh.name +
                                "' has no monolithicCoupledPatches entry");

    sys.regions.resize(R.size());
    for (size_t r = 0; r < R.size(); ++r) {
        const FoamRegion &rc = R[r];
        const PolyMesh &m = rc.mesh;
        FvMatrixG E = setCoupledEqns(rc);
        LduRegion &L = sys.regions[r];
        L.name = m.name;
        L.nCells = m.nCells;
        L.nFaces = m.nInternal;
        L.lAddr.assign(m.owner.begin(), m.owner.begin() + m.nInternal);
        L.uAddr = m.neighbour;
        L.symmetric = !E.asymmetric;
        L.diag = E.diag;
        L.upper = E.upper;
        L.lower = E.asymmetric ? E.lower : E.upper;
        L.source = E.source;
        L.psi = rc.T;
        // addBoundaryDiag for every patch, addBoundarySource(couples=false)
        // for the others; the coupled patches' boundaryCoeffs become the
        // interface coefficients
        for (size_t p = 0; p < m.patches.size(); ++p)
            for (int i = 0; i < m.patches[p].size; ++i)
                L.diag[m.owner[m.patches[p].start + i]] += E.internalCoeffs[p][i];
        for (size_t p = 0; p < m.patches.size(); ++p) {
            if (rc.bcT[p] == BcT::Coupled || rc.bcT[p] == BcT::PartDirichlet || rc.bcT[p] == BcT::PartNeumann)
                continue;
            for (int i = 0; i < m.patches[p].size; ++i)
                L.source[m.owner[m.patches[p].start + i]] += E.boundaryCoeffs[p][i];
        }
        if (partModel) {
            // partitioned sides: per unit value, then the initial values
            // (Dirichlet: the patch value; Neumann: zero gradient)
            partModel->sourceBase[r] = L.source;
            for (auto &pp : partModel->pairs)
                for (PartitionedSide *ps : {&pp.dir, &pp.neu}) {
                    if (ps->region != (int)r) continue;
                    const int pat = ps == &pp.dir ? partPatch[&pp - &partModel->pairs[0]][0]
                                                  : partPatch[&pp - &partModel->pairs[0]][1];
                    ps->coeff = E.boundaryCoeffs[pat];
                }
        }
        for (size_t i = 0; i < sides.size(); ++i)
            for (int s = 0; s < 2; ++s) {
                if (sides[i][2 * s] != (int)r) continue;
                InterfaceSide &S = s ? sys.interfaces[i].b : sys.interfaces[i].a;
                S.coeff = E.boundaryCoeffs[sides[i][2 * s + 1]];
            }
        // Euler ddt (fvm::ddt with coefficient rho*c): a = rho*c*V/dt on the
        // diagonal, a*T_old in the source; a `backward` run continues from here
        if (rDeltaT[r] > 0.0) {
            L.diagBase = L.diag;
            L.sourceBase = L.source;
            L.ddtCoeff.resize(m.nCells);
            for (int c = 0; c < m.nCells; ++c) {
                L.ddtCoeff[c] = rDeltaT[r] * rc.rhoCp * m.V[c];
                L.diag[c] = L.diagBase[c] + L.ddtCoeff[c];
                L.source[c] = L.sourceBase[c] + L.ddtCoeff[c] * rc.T[c];
            }
        }
    }
    if (partModel) {
        // the initial boundary values: Dirichlet sides their patch values,
        // Neumann sides a zero gradient
        for (auto &pp : partModel->pairs)
            for (const PartitionedSide *ps : {&pp.dir, &pp.neu}) {
                LduRegion &L = sys.regions[ps->region];
                for (int i = 0; i < ps->nFaces(); ++i)
                    L.source[ps->fc[i]] += ps->coeff[i] * (ps == &pp.dir ? ps->value0[i] : 0.0);
            }
        // a transient run keeps IQN-ILS columns within one time step only
        bool transient = false;
        for (auto &L : sys.regions) transient = transient || L.transient();
        for (auto &pp : partModel->pairs)
            for (const PartitionedSide *ps : {&pp.dir, &pp.neu})
                if (transient && ps->accType == 2 && ps->reuse > 0)
                    throw FoamError("IQN-ILS with couplingReuse > 0 is not supported in transient runs "
                                    "(couplingReuse 0 keeps the columns of the current time step)");
        sys.part = partModel;
        sys.fv = nullptr;   // device re-assembly covers monolithic set-ups
    } else {
        sys.fv = buildFvModel(R, rDeltaT, sides, sys);
    }
    if (info) {
        info->volumes.clear();
        info->maxNonOrthDeg.clear();
        info->rDeltaT = rDeltaT;
        info->rhoCp.clear();
        info->k.clear();
        for (auto &rc : R) {
            info->volumes.push_back(rc.mesh.V);
            info->maxNonOrthDeg.push_back(rc.mesh.maxNonOrthDeg);
            info->rhoCp.push_back(rc.rhoCp);
            info->k.push_back(rc.k);
        }
    }
    sys.check();
    return sys;
}

}  // namespace mrfg
