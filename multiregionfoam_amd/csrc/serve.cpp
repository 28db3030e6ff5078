#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <stdexcept>
#include <thread>

namespace mrfg {

namespace {

double wallSeconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// the host cores this process may use: OMP_NUM_THREADS when set (the share
// a scheduler gives the process), else the hardware's
int coreShare()
{
    if (const char *e = std::getenv("OMP_NUM_THREADS")) {
        const int n = std::atoi(e);
        if (n > 0) return n;
    }
    const unsigned h = std::thread::hardware_concurrency();
    return h > 0 ? (int)h : 8;
}

}  // namespace

size_t Server::footprint(const CoupledSystem &s)
{
    // padded rows ~ 1.15 N; per row: 12 FP64 + 14 FP32 vectors, matrix
    // diagonal / source / sumA / psi, DILU records and codes of both
    // precisions, inboxes; per face: two ELL entries (value + column) in
    // both precisions; per GGI pair: column + weight
    const double N = (double)s.nCells() * 1.15 + 64.0 * 64.0, F = (double)s.nFaces(), P = (double)s.nPairs();
    const double perRow = 12 * 8 + 14 * 4 + 5 * 8 + 2 * 32 + 2 * 16 + 16 * 4 + 3 * 8 + 64;
    const double perFace = 2 * (8 + 4 + 4);
    const double perPair = 2 * (8 + 4 + 4);
    return (size_t)((N * perRow + F * perFace + P * perPair) * 1.25) + (16u << 20);
}

Server::Server(int device, size_t hbmBudget, int nWorkers, int maxBatch, const SolverControls &c, bool keepFields,
               int nPrep)
    : device_(device), budget_(hbmBudget), maxBatch_(std::max(1, maxBatch)), ctrl_(c), keepFields_(keepFields),
      t0_(wallSeconds())
{
    if (nWorkers < 1) throw std::invalid_argument("a server needs at least one worker");
    if (nPrep <= 0) nPrep = std::max(1, std::min(16, coreShare() - nWorkers));
    for (int w = 0; w < nPrep; ++w) prep_.emplace_back(&Server::prepare, this);
    for (int w = 0; w < nWorkers; ++w) workers_.emplace_back(&Server::worker, this);
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cvArrive_.notify_all();
    cvReady_.notify_all();
    for (auto &t : prep_) t.join();
    for (auto &t : workers_) t.join();
}

double Server::now() const { return wallSeconds() - t0_; }

int Server::submit(const std::string &caseText)
{
    std::lock_guard<std::mutex> lk(mu_);
    const int id = (int)reqs_.size();
    reqs_.emplace_back(new ServeRequest);
    reqs_.back()->text = caseText;
    reqs_.back()->tSubmit = now();
    arrivals_.push_back(id);
    ++pending_;
    cvArrive_.notify_one();
    return id;
}

void Server::waitAll()
{
    std::unique_lock<std::mutex> lk(mu_);
    cvDone_.wait(lk, [&] { return pending_ == 0; });
}

bool Server::result(int id, ServeResult &out) const
{
    std::lock_guard<std::mutex> lk(mu_);
    if (id < 0 || id >= (int)reqs_.size()) return false;
    const ServeRequest &R = *reqs_[id];
    out.status = R.status;
    out.tSubmit = R.tSubmit;
    out.tAdmit = R.tAdmit;
    out.tDone = R.tDone;
    out.tReady = R.tReady;
    out.tSolve = R.tSolve;
    out.batchSize = R.batchSize;
    out.perf = R.perf;
    out.error = R.error;
    return true;
}

bool Server::copyPsi(int id, int region, double *out) const
{
    std::lock_guard<std::mutex> lk(mu_);
    if (id < 0 || id >= (int)reqs_.size()) return false;
    const ServeRequest &R = *reqs_[id];
    if (R.status != 2 || region < 0 || region >= (int)R.psi.size()) return false;
    std::copy(R.psi[region].begin(), R.psi[region].end(), out);
    return true;
}

int Server::nRequests() const
{
    std::lock_guard<std::mutex> lk(mu_);
    return (int)reqs_.size();
}

void Server::fail(int id, const std::string &why)
{
    ServeRequest &R = *reqs_[id];
    R.status = -1;
    R.error = why;
    R.tDone = now();
    R.sys.reset();
    R.layout.reset();
    --pending_;
    cvDone_.notify_all();
}

// Assemble and lay out the oldest unprepared request (host work, in
// parallel with the other preparation threads and the device).
void Server::prepare()
{
    for (;;) {
        int id;
        std::string text;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cvArrive_.wait(lk, [&] { return stop_ || !arrivals_.empty(); });
            if (stop_) return;
            id = arrivals_.front();
            arrivals_.pop_front();
            text = reqs_[id]->text;
        }
        try {
            std::unique_ptr<CoupledSystem> sys(new CoupledSystem(buildCase(text)));
            const size_t bytes = footprint(*sys);
            std::unique_ptr<SystemLayout> lay;
            if (bytes <= budget_) lay.reset(new SystemLayout(buildLayout(*sys)));
            std::lock_guard<std::mutex> lk(mu_);
            if (bytes > budget_) {
                fail(id, "request needs more device memory than the serving budget");
                continue;
            }
            ServeRequest &R = *reqs_[id];
            R.sys = std::move(sys);
            R.layout = std::move(lay);
            R.bytes = bytes;
            R.tReady = now();
            ready_.insert(id);
            cvReady_.notify_all();
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> lk(mu_);
            fail(id, e.what());
        }
    }
}

// One device slot: admit the oldest prepared requests that fit the memory
// left, solve them together, retire each as soon as it has converged.
void Server::worker()
{
    std::unique_ptr<Batch> batch;
    std::string noDevice;
    try {
        batch.reset(new Batch(device_));
    } catch (const std::exception &e) {
        noDevice = e.what();
    }
    // the systems the batch points at, kept until it is refilled
    std::vector<std::unique_ptr<CoupledSystem>> held;
    for (;;) {
        std::vector<int> ids;
        std::vector<const CoupledSystem *> ptrs;
        std::vector<std::unique_ptr<SystemLayout>> layouts;
        size_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cvReady_.wait(lk, [&] {
                if (stop_) return true;
                if (ready_.empty()) return false;
                return inUse_ + reqs_[*ready_.begin()]->bytes <= budget_ || inFlight_ == 0;
            });
            if (stop_) return;
            if (!batch) {
                fail(*ready_.begin(), "no device for this worker: " + noDevice);
                ready_.erase(ready_.begin());
                continue;
            }
            // the oldest first; then, in arrival order, those that still fit
            for (auto it = ready_.begin(); it != ready_.end() && (int)ids.size() < maxBatch_;) {
                const size_t b = reqs_[*it]->bytes;
                if (!ids.empty() && inUse_ + b > budget_) break;
                inUse_ += b;
                bytes += b;
                ids.push_back(*it);
                it = ready_.erase(it);
            }
            ++inFlight_;
            const double t = now();
            held.clear();
            for (int id : ids) {
                ServeRequest &R = *reqs_[id];
                R.tAdmit = t;
                R.status = 1;
                R.batchSize = (int)ids.size();
                ptrs.push_back(R.sys.get());
                held.push_back(std::move(R.sys));
                layouts.push_back(std::move(R.layout));
            }
        }
        std::vector<char> retired(ids.size(), 0);
        auto retire = [&](int q, const SolvePerf &p) {
            std::vector<std::vector<double>> f;
            if (keepFields_)
                for (size_t r = 0; r < held[q]->regions.size(); ++r) {
                    f.emplace_back(held[q]->regions[r].nCells);
                    batch->copyPsiNow(q, (int)r, f.back().data());
                }
            std::lock_guard<std::mutex> lk(mu_);
            ServeRequest &R = *reqs_[ids[q]];
            R.perf = p;
            R.psi = std::move(f);
            R.status = 2;
            R.tDone = now();
            retired[q] = 1;
            --pending_;
            cvDone_.notify_all();
        };
        std::string err;
        try {
            batch->setSystems(ptrs, &layouts);
            {
                std::lock_guard<std::mutex> lk(mu_);
                const double t = now();
                for (int id : ids) reqs_[id]->tSolve = t;
            }
            std::vector<SolvePerf> perf;
            const Batch::RetireFn fn = retire;
            batch->solve(ctrl_, perf, nullptr, &fn);
            for (size_t q = 0; q < ids.size(); ++q)
                if (!retired[q]) retire((int)q, perf[q]);
        } catch (const std::exception &e) {
            err = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (!err.empty())
                for (size_t q = 0; q < ids.size(); ++q)
                    if (!retired[q]) fail(ids[q], err);
            inUse_ -= bytes;
            --inFlight_;
        }
        cvReady_.notify_all();
    }
}

}  // namespace mrfg
