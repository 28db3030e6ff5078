// Batch: value refresh of committed systems defined from LDU arrays.
//
// In the reference's time loop every time step re-assembles each region's
// fvMatrix (regionType::setCoupledEqns) and calls coupledFvMatrix::solve
// (multiRegionSystem.C:115-153, driven by multiRegionFoam.C:94-117): the
// mesh, and with it the addressing, stays the same, only the values change.
// Here the structure -- layout, DILU tile schedule, codes, inboxes -- is
// built once at commit; a refresh stages the new values on the device and
// the placement pass (assembly.hip, k_fv_place) gathers them into the
// solver's arrays through the maps the layout builder recorded (where every
// ELL entry, DILU record slot and interface-row face came from).
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "batch.hpp"
#include "dilu.hpp"

namespace mrfg {

#define HIP_OK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
    } while (0)

RefreshMaps::~RefreshMaps()
{
    if (stage) {
        std::lock_guard<std::recursive_mutex> lk(deviceMutex());
        (void)hipHostFree(stage);
    }
}

void Batch::buildRefresh()
{
    const int nS = (int)sys_.size();
    std::unique_ptr<RefreshMaps> F(new RefreshMaps);
    F->cellBase.assign(nS, 0);
    F->faceBase.assign(nS, 0);
    F->sideBase.assign(nS, 0);
    F->dirty.assign(nS, 0);
    F->changed.assign(nS, 0);
    std::vector<int32_t> cellOfPos(nPos_, -1), ellSrc, fSrc, bSrc, ifSrc;
    std::vector<double> faceVal, sideCoeff, diag, source;
    int cellBase = 0, faceBase = 0, sideBase = 0;
    for (int q = 0; q < nS; ++q) {
        const CoupledSystem &S = *sys_[q];
        const SystemLayout &L = *lay_[q];
        F->cellBase[q] = cellBase;
        F->faceBase[q] = faceBase;
        F->sideBase[q] = sideBase;
        // the committed values, numbered as the layout's source maps are:
        // face f of region r at 2*(faceOffset[r] + f) (+0 lower, +1 upper),
        // interface-side faces in interface order, side a then side b
        for (const LduRegion &R : S.regions) {
            diag.insert(diag.end(), R.diag.begin(), R.diag.end());
            source.insert(source.end(), R.source.begin(), R.source.end());
            for (int f = 0; f < R.nFaces; ++f) {
                faceVal.push_back(R.lower[f]);
                faceVal.push_back(R.upper[f]);
            }
        }
        int nSideQ = 0;
        for (const CoupledInterface &I : S.interfaces)
            for (const InterfaceSide *IS : {&I.a, &I.b}) {
                sideCoeff.insert(sideCoeff.end(), IS->coeff.begin(), IS->coeff.end());
                nSideQ += IS->nFaces();
            }
        const int p0 = posBegin(q);
        for (int p = 0; p < L.nPos; ++p)
            if (L.cellOfPos[p] >= 0) cellOfPos[p0 + p] = cellBase + L.cellOfPos[p];
        for (int32_t v : L.ellSrc) ellSrc.push_back(v >= 0 ? v + 2 * faceBase : -1);
        for (int s = 0; s < L.nSlices; ++s)
            for (int kk = 0; kk < K_; ++kk)
                for (int l = 0; l < kSlice; ++l) {
                    int fs = -1, bs = -1;
                    if (kk < L.K) {
                        const size_t e = ((size_t)s * L.K + kk) * kSlice + l;
                        fs = L.fSrc[e] >= 0 ? L.fSrc[e] + 2 * faceBase : -1;
                        bs = L.bSrc[e] >= 0 ? L.bSrc[e] + 2 * faceBase : -1;
                    }
                    fSrc.push_back(fs);
                    bSrc.push_back(bs);
                }
        for (int32_t v : L.ifSrc) ifSrc.push_back(v + sideBase);
        cellBase += (int)S.nCells();
        faceBase += (int)S.nFaces();
        sideBase += nSideQ;
    }
    // every index the placement pass follows stays inside its array
    if (ellSrc.size() != m64_.ellVal.n || fSrc.size() != (size_t)nSlices_ * K_ * kSlice ||
        ifSrc.size() != m64_.ifCoeff.n || (int)diag.size() != cellBase || (int)faceVal.size() != 2 * faceBase ||
        (int)sideCoeff.size() != sideBase)
        throw std::runtime_error("value refresh maps do not match the batch layout");
    auto inRange = [](const std::vector<int32_t> &v, long long lo, long long hi, const char *what) {
        for (int32_t x : v)
            if (x < lo || x >= hi) throw std::runtime_error(std::string("value refresh: index out of range in ") + what);
    };
    inRange(cellOfPos, -1, cellBase, "position cell");
    inRange(ellSrc, -1, 2LL * faceBase, "ELL source");
    inRange(fSrc, -1, 2LL * faceBase, "forward record source");
    inRange(bSrc, -1, 2LL * faceBase, "backward record source");
    inRange(ifSrc, 0, sideBase, "interface source");
    F->nCells = cellBase;
    F->nFaces = faceBase;
    F->nSides = sideBase;
    hipStream_t st = stream_;
    F->cellOfPos.upload(cellOfPos, st);
    F->ellSrc.upload(ellSrc.empty() ? std::vector<int32_t>{-1} : ellSrc, st);
    F->fSrc.upload(fSrc, st);
    F->bSrc.upload(bSrc, st);
    F->ifSrc.upload(ifSrc.empty() ? std::vector<int32_t>{0} : ifSrc, st);
    F->faceVal.upload(faceVal.empty() ? std::vector<double>{0.0} : faceVal, st);
    F->sideCoeff.upload(sideCoeff.empty() ? std::vector<double>{0.0} : sideCoeff, st);
    F->diag.upload(diag, st);
    F->source.upload(source, st);
    {
        std::lock_guard<std::recursive_mutex> lk(deviceMutex());
        const size_t n = 2 * (size_t)faceBase + 2 * (size_t)cellBase + (size_t)sideBase + 1;
        HIP_OK(hipHostMalloc((void **)&F->stage, n * sizeof(double), hipHostMallocDefault));
    }
    HIP_OK(hipStreamSynchronize(st));
    rf_ = std::move(F);
}

namespace {

void checkRefreshable(bool transient, bool model, int sys)
{
    if (model)
        throw std::runtime_error("system " + std::to_string(sys) + " was read from disk: its values come from its "
                                 "fields (mrfg_set_field, mrfg_assemble)");
    if (transient)
        throw std::runtime_error("value refresh of a batch with ddt terms: hand over each step's full matrix on a "
                                 "steady commit, or step with mrfg_time_steps");
}

}  // namespace

void Batch::updateRegion(int sys, int region, const double *diag, const double *lower, const double *upper,
                         const double *source)
{
    HIP_OK(hipSetDevice(device_));
    if (sys < 0 || sys >= (int)sys_.size()) throw std::out_of_range("system index");
    checkRefreshable(transient_, hasModel(sys), sys);
    const CoupledSystem &S = *sys_[sys];
    if (region < 0 || region >= (int)S.regions.size()) throw std::out_of_range("region index");
    const LduRegion &R = S.regions[region];
    if (!diag || !source || (R.nFaces > 0 && !upper)) throw std::invalid_argument("null region array");
    if (lower && R.symmetric)
        throw std::invalid_argument("region " + std::to_string(region) + " was committed symmetric: pass lower = NULL");
    if (!lower && !R.symmetric)
        throw std::invalid_argument("region " + std::to_string(region) + " was committed asymmetric: lower is needed");
    if (!rf_) buildRefresh();
    const SystemLayout &L = *lay_[sys];
    const size_t c0 = (size_t)rf_->cellBase[sys] + L.regionOffset[region];
    const size_t f0 = (size_t)rf_->faceBase[sys] + L.faceOffset[region];
    // staged in pinned memory in the device arrays' layout, then one DMA each
    double *sf = rf_->stage + 2 * f0;
    double *sd = rf_->stage + 2 * (size_t)rf_->nFaces + c0;
    double *ss = rf_->stage + 2 * (size_t)rf_->nFaces + (size_t)rf_->nCells + c0;
    const double *lo = lower ? lower : upper;
    for (int f = 0; f < R.nFaces; ++f) {
        sf[2 * (size_t)f] = lo[f];
        sf[2 * (size_t)f + 1] = upper[f];
    }
    std::memcpy(sd, diag, R.nCells * sizeof(double));
    std::memcpy(ss, source, R.nCells * sizeof(double));
    HIP_OK(hipMemcpyAsync(rf_->diag.p + c0, sd, R.nCells * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIP_OK(hipMemcpyAsync(rf_->source.p + c0, ss, R.nCells * sizeof(double), hipMemcpyHostToDevice, stream_));
    if (R.nFaces > 0)
        HIP_OK(hipMemcpyAsync(rf_->faceVal.p + 2 * f0, sf, 2 * (size_t)R.nFaces * sizeof(double),
                              hipMemcpyHostToDevice, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    rf_->dirty[sys] = 1;
}

void Batch::updateInterface(int sys, int iface, const double *coeffA, const double *coeffB)
{
    HIP_OK(hipSetDevice(device_));
    if (sys < 0 || sys >= (int)sys_.size()) throw std::out_of_range("system index");
    checkRefreshable(transient_, hasModel(sys), sys);
    const CoupledSystem &S = *sys_[sys];
    if (iface < 0 || iface >= (int)S.interfaces.size()) throw std::out_of_range("interface index");
    if (!rf_) buildRefresh();
    size_t off = rf_->sideBase[sys];
    for (int i = 0; i < iface; ++i) off += S.interfaces[i].a.nFaces() + S.interfaces[i].b.nFaces();
    const CoupledInterface &I = S.interfaces[iface];
    double *st = rf_->stage + 2 * (size_t)rf_->nFaces + 2 * (size_t)rf_->nCells + off;
    if (coeffA && I.a.nFaces() > 0) {
        std::memcpy(st, coeffA, I.a.nFaces() * sizeof(double));
        HIP_OK(hipMemcpyAsync(rf_->sideCoeff.p + off, st, I.a.nFaces() * sizeof(double), hipMemcpyHostToDevice,
                              stream_));
    }
    if (coeffB && I.b.nFaces() > 0) {
        std::memcpy(st + I.a.nFaces(), coeffB, I.b.nFaces() * sizeof(double));
        HIP_OK(hipMemcpyAsync(rf_->sideCoeff.p + off + I.a.nFaces(), st + I.a.nFaces(),
                              I.b.nFaces() * sizeof(double), hipMemcpyHostToDevice, stream_));
    }
    HIP_OK(hipStreamSynchronize(stream_));
    rf_->dirty[sys] = 1;
}

void Batch::applyRefresh()
{
    if (!rf_) return;
    bool any = false;
    for (char d : rf_->dirty) any = any || d;
    if (!any) return;
    HIP_OK(hipSetDevice(device_));
    FvPlace q;
    q.K = K_;
    q.G = packChunks(K_ + 1, 2);
    q.cellOfPos = rf_->cellOfPos.p;
    q.ellOff = ellOff_.p;
    q.ellSrc = rf_->ellSrc.p;
    q.fSrc = rf_->fSrc.p;
    q.bSrc = rf_->bSrc.p;
    q.ifMask = ifMask_.p;
    q.ifBase = ifBase_.p;
    q.ifRowFace = ifRowFace_.p;
    q.ifSrc = rf_->ifSrc.p;
    q.faceVal = rf_->faceVal.p;
    q.sideCoeff = rf_->sideCoeff.p;
    q.diag = rf_->diag.p;
    q.source = rf_->source.p;
    q.ddtA = q.diagBase = q.srcBase = nullptr;
    q.ellVal = m64_.ellVal.p;
    q.diagOut = m64_.diag.p;
    q.b = b64_.p;
    q.sumA = sumA_.p;
    q.ifCoeff = m64_.ifCoeff.p;
    q.fPack = (double *)m64_.fPack.p;
    q.bPack = (double *)m64_.bPack.p;
    q.prodPack = (double *)prodPack_.p;
    q.ddtAOut = q.diagBaseOut = q.srcBaseOut = nullptr;
    // each staged system's positions (systems are contiguous in position space)
    const int nS = (int)sys_.size();
    for (int a = 0; a < nS;) {
        if (!rf_->dirty[a]) {
            ++a;
            continue;
        }
        int b = a;
        while (b < nS && rf_->dirty[b]) ++b;
        q.pos0 = posBegin(a);
        q.nPos = b < nS ? posBegin(b) : nPos_;
        HIP_OK(launchFvPlace(q, stream_));
        for (int k = a; k < b; ++k) {
            rf_->dirty[k] = 0;
            rf_->changed[k] = 1;
        }
        a = b;
    }
    HIP_OK(hipStreamSynchronize(stream_));
    factorised_ = false;
    haveFloat_ = false;
    hostStale_ = true;
}

}  // namespace mrfg
