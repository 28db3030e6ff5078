// Device-resident batch of coupled block systems and the solve drivers.
//
// A batch holds one or more independent coupled systems (each: regions +
// monolithic interfaces) concatenated in one position space.  Kernels run
// over all systems at once; Krylov scalars and convergence are per system,
// so a single system is a batch of one and many uneven systems share every
// launch.  Matrices, factors and fields stay resident between solves.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "assembly.hpp"
#include "coupled_system.hpp"
#include "device.hpp"
#include "dictionary.hpp"
#include "kernels.hpp"
#include "layout.hpp"

namespace mrfg {

// Solver controls of an fvSolution solver entry, e.g.
//   Tcoupled { solver BiCGStab; preconditioner { preconditioner Cholesky; }
//              tolerance 1e-10; relTol 0; minIter 0; maxIter 1000;
//              precision mixed; refinement { innerRelTol 1e-4; ... } }
struct SolverControls {
    enum Kind { BiCGStab = 0, PCG = 1 } solver = BiCGStab;
    enum Precision { FP64 = 0, Mixed = 1 } precision = FP64;
    double tolerance = 1e-6;
    double relTol = 0.0;
    int minIter = 0;
    int maxIter = 1000;
    double innerRelTol = 1e-4;
    int innerMaxIter = 1000;
    int maxOuter = 50;
    // mixed: the FP32 inner solve's preconditioner.  DILU: the FP64 path's
    // DILU in FP32.  TileDILU: every box tile's DILU alone, the couplings
    // across tile edges left out, so the sweeps have no cross-tile waits
    // (regions on general waves keep the full DILU).  The outer FP64
    // refinement sets the result's accuracy either way.
    enum InnerPre { InnerDILU = 0, InnerTileDILU = 1 } innerPre = InnerTileDILU;
    int checkInterval = 8;    // iterations between host convergence checks
    bool graphs = true;       // replay each check interval's iterations as one captured HIP graph
    bool autotune = true;     // time the SpMV launch configurations on the batch once, use the fastest
};

// The batch kernels' launch configuration chosen by the autotuner: blocks of
// 4, 8 or 16 slices (256, 512 or 1024 threads) and the SpMV's ELL chunk;
// the SpMV times (ms) of the choice and of the fixed default (4 slices,
// chunk 8).  Results do not depend on it (row sums in entry order, exact
// reductions).
struct Tuning {
    bool tuned = false;
    int slices = 4, ellChunk = 8;
    double bestMs = 0.0, defaultMs = 0.0;
};

SolverControls parseControls(const Dict &d);

struct SolvePerf {
    double initialResidual = 0, finalResidual = 0;
    int nIterations = 0, converged = 0, singular = 0, nOuter = 0;
    double seconds = 0;       // device time of the solve (setup + iterations)
    double bytes = 0;         // algorithmic bytes moved (canonical model)
};

// Held around every device allocation and free, stream creation and
// destruction, and each HIP-graph capture of the library: a free (an
// implicit device synchronisation) or a stream change in one thread while
// another thread captures would invalidate that capture.
std::recursive_mutex &deviceMutex();

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;      // elements in use
    size_t cap = 0;    // elements allocated
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        std::lock_guard<std::recursive_mutex> lk(deviceMutex());
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        cap = 0;
    }
    void alloc(size_t count);
    // copy h (pageable: the copy is synchronous for the host) on stream st
    void upload(const std::vector<T> &h, hipStream_t st);
    void zero(hipStream_t st);
};

template <typename T>
struct DevMatrix {
    DevBuf<T> ellVal, diag, ifCoeff, ifPairW;
    DevBuf<typename Pack<T>::V> fPack, bPack;   // DILU records
};

// The finite-volume models of a batch's systems read from disk, numbered
// batch-wide (systems without a model contribute nothing), with the maps
// that put assembled values into the solver's arrays.
struct FvBatch {
    std::vector<int> cellBase, faceBase, sideBase, regionBase;   // per system, -1 without a model
    std::vector<FvRegionParams> hRegion;
    int nCells = 0, nFaces = 0, nSideFaces = 0;
    long long nBFaces = 0, nPairs = 0, nPosModel = 0, nEll = 0, nSlots = 0, nIfEntries = 0;
    DevBuf<FvRegionParams> region;
    DevBuf<int32_t> cellRegion, faceRegion, own, nei, cfOff, cf, cbOff, bSide, sCell, sOff, sPairCell;
    DevBuf<int8_t> bT, bU, bKFixed;
    DevBuf<double> w, dc, magSf, Sf, V, k, U, T, bNDist, bMagSf, bSf, bTb, bUb, bKb, sNDist, sW, sPairNDist;
    DevBuf<double> lapU, aLo, aUp, faceVal, sK, sDelta, sKOwn, sKNei, ifCoeff, diag, source, ddtA, diagBase, srcBase;
    DevBuf<int32_t> cellOfPos, ellSrc, fSrc, bSrc, ifSrc;
    bool assembled = false;
};

// Value refresh of systems defined from LDU arrays: the values a drop-in
// caller hands over per solve (diag, lower / upper, source per region;
// interface coefficients per side), numbered batch-wide, and the maps the
// placement pass gathers them through (built on the first refresh).
struct RefreshMaps {
    std::vector<int> cellBase, faceBase, sideBase;   // per system
    int nCells = 0, nFaces = 0, nSides = 0;
    DevBuf<int32_t> cellOfPos, ellSrc, fSrc, bSrc, ifSrc;
    DevBuf<double> faceVal, sideCoeff, diag, source;
    std::vector<char> dirty;     // per system: staged values not yet placed
    std::vector<char> changed;   // per system: values differ from the committed ones
    // pinned host staging of the values in the device arrays' layout
    // (faceVal, diag, source, sideCoeff), so each upload is one DMA
    double *stage = nullptr;
    RefreshMaps() = default;
    RefreshMaps(const RefreshMaps &) = delete;
    RefreshMaps &operator=(const RefreshMaps &) = delete;
    ~RefreshMaps();
};

class Batch {
public:
    explicit Batch(int device);
    ~Batch();

    // Replace the batch content.  Builds layouts (or takes the given ones,
    // one per system, built by buildLayout without preferLevels), uploads
    // FP64 (and, on demand, FP32) matrices and the initial guesses.
    void setSystems(const std::vector<const CoupledSystem *> &systems,
                    std::vector<std::unique_ptr<SystemLayout>> *layouts = nullptr);
    int nSystems() const { return (int)sys_.size(); }

    // Called for each system as soon as it has stopped iterating for good
    // (at the convergence check that finds it, while the others continue),
    // with its performance record; copyPsiNow may read its field then.
    using RetireFn = std::function<void(int sys, const SolvePerf &perf)>;

    // Solve every system (or those with active[q] != 0; the others keep
    // their fields); perf gets one entry per system.
    void solve(const SolverControls &c, std::vector<SolvePerf> &perf, const std::vector<char> *active = nullptr,
               const RetireFn *retire = nullptr);
    // The field of one region of a system that no longer iterates, copied
    // without waiting for the batch's stream (the other systems' kernels
    // may be running; this system's part of the field is final).
    void copyPsiNow(int sys, int region, double *psi);

    // Time stepping of the transient systems (regions with ddt coefficients;
    // steady systems are re-solved from their current field): per step the
    // ddt terms are rebuilt on the device from the current and previous
    // fields, then every system is solved from its current field.  Euler,
    // or backward (constant time step; the first step of a run is Euler).
    // perf gets nSteps x nSystems entries.  FP64 only.
    void timeSteps(const SolverControls &c, bool backward, int nSteps, std::vector<SolvePerf> &perf);
    bool transient() const { return transient_; }
    // The ddt terms of the next time step, rebuilt on the device from the
    // current and previous fields (diagonal, source b, previous field); the
    // systems are then solved by the caller.
    void beginTimeStep(bool backward);

    // Reset every system's field to the committed initial guess (device
    // copy); a following time-step run starts anew.
    void resetPsi();
    // Per-region initial guess / solution of one system (original cell order).
    void setPsi(int sys, int region, const double *psi);
    void getPsi(int sys, int region, double *psi);

    // Kernel-level entry points (used by tests and the profiler harness).
    void spmv(const double *xHost, double *yHost, int sys);           // y = A x
    void diluFactor(int sys, double *rDHost);                        // rD per cell
    void precondition(const double *rHost, double *wHost, int sys);   // w = M^-1 r
    // FP32 records, rD from the FP64 factor; tileLocal: the mixed solver's
    // tile-local DILU (each box tile alone)
    void precondition32(const float *rHost, float *wHost, int sys, bool tileLocal = false);
    // per cell of system sys (concatenated regions): the batch-wide index of
    // its structured-box tile wave, or -1 for a cell on a general wave
    void tilesOf(int sys, int *tileOfCell) const;
    // Heat flow into each side of interface `iface` of system `sys` from the
    // current FP64 field, on the device: Q = sum coeff*(T_neighbour - T_own)
    // per side (a conserving interface gives qa + qb = 0).
    void interfaceFlux(int sys, int iface, double *qa, double *qb);
    // The same on a caller-given device scratch (nScratch doubles: the
    // reduction partials, then the two results); allocates nothing.  The
    // interface's face lists must be resident (setSystems uploads them).
    void interfaceFluxScratch(int sys, int iface, double *qa, double *qb, double *scratch, size_t nScratch);
    // monolithicTemperature's patch values of one side (0 = a, 1 = b) from
    // the current FP64 field, on the device: the neighbour field tn (the GGI
    // interpolate of the other side's face-cell values,
    // monolithicCouplingFvPatchField::patchNeighbourField) and the wall value
    // tw = (kOwn To + kNei tn) / (kOwn + kNei) (calcTemperature, radiation
    // off).  Needs the side's wall conductances.
    void interfacePatchValues(int sys, int iface, int side, double *tn, double *tw);
    // Replace the wall conductances of one interface side (after commit)
    void setInterfaceWall(int sys, int iface, int side, const double *kOwn, const double *kNei);

    // Time `reps` repetitions of one kernel on resident data; returns the
    // average device time per launch (ms), measured with HIP events on the
    // library's stream.  which: 0 SpMV, 1 DILU forward sweep, 2 DILU
    // backward sweep, 3 DILU factor.
    double timeKernel(int which, int reps);
    // Run one forward/backward pair with per-wave tracing; out receives, per
    // wave in dispatch order, [start, end, slow-path hits, steps] with times
    // in ns relative to the earliest start (4*nWaves values per sweep,
    // forward sweep first).  Returns the number of waves.
    int traceDilu(long long *out, int maxWaves);

    const SystemLayout &layout(int sys) const { return *lay_[sys]; }
    // Time k_spmv_plain in every launch configuration on this batch and keep
    // the fastest (run by solve once per setSystems when controls ask for it)
    void autotune();
    const Tuning &tuning() const { return tune_; }
    void forceTuning(int slices, int ellChunk)
    {
        setTuning(slices, ellChunk);
        tune_.tuned = true;
    }
    // device field and right-hand side (batch positions) and the position of
    // global cell `cell` of system `sys`
    const double *fieldDevice() const { return x64_.p; }
    double *rhsDevice() { return b64_.p; }
    size_t nPositions() const { return (size_t)nPos_; }
    int position(int sys, int cell) const { return posBegin(sys) + lay_.at(sys)->posOfCell.at(cell); }
    hipStream_t stream() const { return stream_; }
    // Canonical algorithmic bytes of one iteration / one SpMV / one DILU
    // apply for system `sys` at value size sv.
    double bytesIter(int sys, int sv) const;
    double bytesSpmv(int sys, int sv) const;
    double bytesDilu(int sys, int sv) const;

    // Value refresh on the committed structure (systems defined from LDU
    // arrays, steady): new values of one region (lower == nullptr: symmetric,
    // the upper values) or of one interface's coefficients (a null side is
    // left unchanged).  They are staged on the device and put in place --
    // SpMV values, DILU records and products, diagonal, source, row sums,
    // interface coefficients -- before the next solve or kernel call; no
    // layout is rebuilt.  applyRefresh() places what is staged.
    void updateRegion(int sys, int region, const double *diag, const double *lower, const double *upper,
                      const double *source);
    void updateInterface(int sys, int iface, const double *coeffA, const double *coeffB);
    void applyRefresh();

    // Force the level schedule for every region (general-mesh path); applies
    // at the next setSystems.
    void setPreferLevels(bool v) { preferLevels_ = v; }

    // Finite-volume models (systems read from disk).  setField replaces a
    // field of one region: "k" (per cell), "U" (3 per cell, transport
    // regions), "T" (per cell; also the system's initial guess for that
    // region, as when the case is read), "rhoCp" (one value).  assemble()
    // re-assembles every system that has a model from its current fields,
    // on the device, and puts the values in place; nothing is uploaded but
    // the fields set since.  The getters return the device-assembled region
    // matrix and interface coefficients in the original order.
    bool hasModel(int sys) const;
    void setField(int sys, int region, const std::string &name, const double *v, size_t n);
    void assemble();
    void getRegionMatrix(int sys, int region, double *diag, double *lower, double *upper, double *source);
    void getInterfaceCoeff(int sys, int iface, int side, double *coeff);
    // Average device time (ms) of one assembly pass over the batch (which 7:
    // faces, interface sides and cells; 8: placement into the solver's
    // arrays) and the bytes each pass reads and writes, every array once.
    double timeAssembly(int which, int reps);
    void assemblyBytes(double &assemble, double &place) const;

private:
    // copies ordered on the library stream, waited for (never on the legacy
    // stream, which must not wait on another slot's stream while that one
    // captures a graph)
    void copyToHost(void *dst, const void *src, size_t bytes) const;
    void copyToDevice(void *dst, const void *src, size_t bytes) const;
    // the FP32 copies; tileLocal: box tiles' records without the
    // coefficients across tile edges, rD of the tile-local factor
    void ensureFloat(bool tileLocal);
    void factorise();
    void initState(const SolverControls &c, const std::vector<char> *active = nullptr);
    void readState(std::vector<SysState> &out);
    void runFp64(const SolverControls &c);
    // early retirement of finished systems during a solve
    const RetireFn *retire_ = nullptr;
    std::vector<char> retired_;
    std::chrono::steady_clock::time_point solveStart_;
    SolvePerf perfFp64(int q, const SysState &s) const;
    std::vector<int> newlyFinished(const std::vector<SysState> &hs);
    // `chunk` BiCGStab / PCG iterations on the library stream: one launch of
    // a HIP graph captured from the iteration's kernel chain (kept per
    // kind and chunk length until the batch's buffers change), or the
    // kernels launched one by one
    template <typename T>
    void runChunk(bool pcg, int chunk, bool useGraph, const MatView<T> &A, const DiluView<T> &D, const Krylov<T> &K);
    struct IterGraph {
        int kind = -1, chunk = 0;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<IterGraph> graphs_;
    void dropGraphs();
    // Mixed FP32/FP64 iterative refinement; a system whose outer residual
    // falls by less than kStallFactor in an outer step continues in FP64.
    static constexpr double kStallFactor = 0.5;
    // host checks of an inner FP32 solve without progress before it stops
    static constexpr int kStallWindow = 3;
    void runMixed(const SolverControls &c, std::vector<int> &innerIters, std::vector<int> &fp64Iters);
    void checkDeviceError();
    void markAllRunning();
    template <typename T>
    MatView<T> view(const DevMatrix<T> &m) const;
    template <typename T>
    DiluView<T> diluView(DevMatrix<T> &m, void *hF, void *hB);
    template <typename T>
    void bindKrylov(Krylov<T> &K);
    int posBegin(int sys) const { return sliceBegin_[sys] * kSlice; }
    void setDdtDiag(double coefft);
    void buildFv();
    FvDev fvDevView() const;
    FvPlace fvPlaceView();
    // the host copies of the FP64 values after the device changed them
    void refreshHostCopies();
    // the FP32 DILU records (host) from the FP64 ones (tileLocal: without the
    // coefficients across box tile edges)
    void derivePack32(bool tileLocal);

    int device_;
    hipStream_t stream_ = nullptr;
    DevBuf<int8_t> gate_;   // per system: the mixed refinement's outer step applies
    hipStream_t side_ = nullptr;               // box DILU sweeps beside general ones
    hipEvent_t fork_ = nullptr, join_ = nullptr;
    bool preferLevels_ = false;
    std::vector<std::unique_ptr<SystemLayout>> lay_;
    std::vector<const CoupledSystem *> sys_;
    std::vector<int> sliceBegin_;
    int nSlices_ = 0, nPos_ = 0, nBlk_ = 0, nWaves_ = 0, K_ = 3, nHaloF_ = 0, nHaloB_ = 0;
    int nGenWaves_ = 0, nBoxWaves_ = 0, nBoxIn_ = 0;
    bool boxLone_ = false;   // every box tile alone in its region
    bool factorised_ = false, haveFloat_ = false;
    bool tileLocal32_ = false;   // the FP32 records are those of the tile-local DILU
    // segments per box tile in the tile-local DILU, each swept by a wave of
    // its own (a shorter chain of steps, slightly more inner iterations): as
    // many as keep the lone waves within kTileSegWaves, at most 3.  The waves
    // may outnumber the CUs several times over: duct3d (135 tiles) takes 3
    // segments (33.6 ms per mixed solve against 37.1 with 2 and 39.9 with 1;
    // 4 needed more iterations than they saved), exchanger3 (585 tiles) 3
    // (362.8 -> 328.6 ms), batch64 2 (656 -> 558 ms; a budget of 16384 waves,
    // giving it 3, was slower: 590 ms)
    static constexpr int kTileSegWaves = 8192;
    int tileSegments() const
    {
        return nBoxWaves_ > 0 ? std::max(1, std::min(3, kTileSegWaves / nBoxWaves_)) : 1;
    }

    // structure (shared by both precisions)
    DevBuf<int32_t> ellOff_, ellCol_, ifBase_, ifRowFace_, ifPairOff_, ifPairCol_;
    DevBuf<unsigned long long> ifMask_;
    DevBuf<int32_t> cellOfPos_;
    DevBuf<BlockEntry> blocks_;
    DevBuf<SysInfo> info_;
    // blocks of 8 and 16 slices (blocks_ / info_ hold those of 4), with the
    // systems' ranges of block partials in each
    struct BlockCfg {
        int nBlk = 0;
        DevBuf<BlockEntry> blocks;
        DevBuf<SysInfo> info;
    };
    std::array<BlockCfg, 2> bcfg_;
    Tuning tune_;
    void launchCfg(int slices, const BlockEntry *&blocks, int &nBlk, const SysInfo *&info, int &threads) const;
    void setTuning(int slices, int ellChunk);
    // DILU schedule
    DevBuf<int32_t> order_, waveSlice_, waveSteps_, waveSys_, waveOut_, boxOrder_, boxMeta_;
    std::vector<int32_t> hBoxMeta_, hBoxWaves_;   // host copies: box tile geometry, the box waves
    DevBuf<double> rdLocal_;                      // rD of the tile-local factor
    DevBuf<unsigned long long> rowMask_;   // per slice: lanes holding a cell
    DevBuf<int4> fCodes_, bCodes_;
    DevBuf<double2> prodPack_;
    int haloMaskF_ = 0, haloMaskB_ = 0;
    DevBuf<unsigned long long> haloF64_, haloB64_, haloC64_;
    DevBuf<unsigned int> haloF32_, haloB32_;
    DevBuf<int> err_;
    // host copies of the FP64 values, kept to derive the FP32 copies
    std::vector<double> hEllVal_, hDiag_, hIfCoeff_, hIfPairW_, hFVal_, hBVal_, hFPack_, hBPack_;
    std::vector<float> hFPack32_, hBPack32_;
    DevMatrix<double> m64_;
    DevMatrix<float> m32_;
    DevBuf<double> sumA_, b64_, x64_, x0_;
    DevBuf<double> v64_[12];
    DevBuf<float> v32_[14];
    DevBuf<SysState> st64_, st32_;
    // per system, per interface, per side: the face lists in batch positions
    struct IfaceDev {
        int nFaces = 0;
        bool wall = false;   // kOwn / kNei known
        DevBuf<int32_t> fcPos, off, pairPos;
        DevBuf<double> coeff, w, kOwn, kNei;
    };
    std::vector<std::vector<std::array<std::unique_ptr<IfaceDev>, 2>>> ifDev_;
    // the face lists of one interface side in batch positions, uploaded on
    // first use (every side is uploaded at setSystems)
    IfaceDev &ifaceSlot(int sys, int iface, int side);
    // copy coefficients (and wall conductances) the device changed since
    // the commit -- assembled from fields, or refreshed -- into the slot
    void ifaceSync(int sys, int iface, int side, IfaceDev &slot);
    DevBuf<double> partials_;
    DevBuf<double> fluxBuf_;   // interface-flux check: partials and results
    // time stepping (position space): ddt coefficient, diagonal and source
    // without the ddt term, previous field
    DevBuf<double> ddtA_, diagBase_, srcBase_, xPrev_;
    bool transient_ = false;
    int stepsDone_ = 0;
    double diagCoefft_ = 1.0;
    std::vector<SysInfo> hInfo_;
    std::unique_ptr<FvBatch> fv_;
    std::unique_ptr<RefreshMaps> rf_;
    void buildRefresh();
    bool hostStale_ = false;
};

}  // namespace mrfg
