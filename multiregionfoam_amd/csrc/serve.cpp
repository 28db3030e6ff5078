#include "serve.hpp"

#include <chrono>
#include <stdexcept>
#include <thread>

namespace mrfg {

namespace {

double wallSeconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
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

Server::Server(int device, size_t hbmBudget, int nWorkers, int maxBatch, const SolverControls &c, bool keepFields)
    : device_(device), budget_(hbmBudget), maxBatch_(std::max(1, maxBatch)), ctrl_(c), keepFields_(keepFields),
      t0_(wallSeconds())
{
    if (nWorkers < 1) throw std::invalid_argument("a server needs at least one worker");
    for (int w = 0; w < nWorkers; ++w) threads_.emplace_back(&Server::worker, this, w);
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : threads_) t.join();
}

double Server::now() const { return wallSeconds() - t0_; }

int Server::submit(const std::string &caseText)
{
    std::lock_guard<std::mutex> lk(mu_);
    int id = (int)reqs_.size();
    reqs_.emplace_back(new ServeRequest);
    reqs_.back()->text = caseText;
    reqs_.back()->tSubmit = now();
    queue_.push_back(id);
    ++pending_;
    cv_.notify_one();
    return id;
}

void Server::waitAll()
{
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return pending_ == 0; });
}

const ServeRequest &Server::request(int id) const
{
    std::lock_guard<std::mutex> lk(mu_);
    return *reqs_.at(id);
}

int Server::nRequests() const
{
    std::lock_guard<std::mutex> lk(mu_);
    return (int)reqs_.size();
}

void Server::worker(int)
{
    std::unique_ptr<Batch> batchPtr;
    try {
        batchPtr.reset(new Batch(device_));
    } catch (const std::exception &) {
        batchPtr.reset();
    }
    for (;;) {
        // 1. take the oldest request and assemble it (outside the lock; the
        // request objects themselves never move)
        int first;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
            if (stop_ && queue_.empty()) return;
            first = queue_.front();
            queue_.pop_front();
        }
        auto req = [&](int id) {
            std::lock_guard<std::mutex> lk(mu_);
            return reqs_[id].get();
        };
        std::vector<int> ids{first};
        std::vector<std::unique_ptr<CoupledSystem>> systems;
        auto assemble = [&](int id) -> bool {
            ServeRequest &R = *req(id);
            if (!batchPtr) {
                std::lock_guard<std::mutex> lk(mu_);
                R.status = -1;
                R.error = "no device for this worker";
                R.tDone = now();
                --pending_;
                done_.notify_all();
                return false;
            }
            try {
                systems.emplace_back(new CoupledSystem(buildCase(R.text)));
                R.bytes = footprint(*systems.back());
                return true;
            } catch (const std::exception &e) {
                std::lock_guard<std::mutex> lk(mu_);
                R.status = -1;
                R.error = e.what();
                R.tDone = now();
                --pending_;
                done_.notify_all();
                return false;
            }
        };
        if (!assemble(first)) continue;
        Batch &batch = *batchPtr;
        size_t bytes = req(first)->bytes;
        // 2. admission by HBM budget; while waiting, more arrivals may join
        {
            std::unique_lock<std::mutex> lk(mu_);
            if (bytes > budget_) {
                ServeRequest &R = *reqs_[first];   // (lock held)
                R.status = -1;
                R.error = "request needs more device memory than the serving budget";
                R.tDone = now();
                --pending_;
                done_.notify_all();
                continue;
            }
            cv_.wait(lk, [&] { return inUse_ + bytes <= budget_ || inFlight_ == 0; });
            inUse_ += bytes;
            ++inFlight_;
        }
        // 3. batch the requests that are already waiting, as far as they fit
        while ((int)ids.size() < maxBatch_) {
            int next;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (queue_.empty()) break;
                next = queue_.front();
                queue_.pop_front();
            }
            if (!assemble(next)) continue;
            size_t b = req(next)->bytes;
            std::lock_guard<std::mutex> lk(mu_);
            if (inUse_ + b > budget_) {   // does not fit now: back to the head of the queue
                queue_.push_front(next);
                systems.pop_back();
                cv_.notify_one();
                break;
            }
            inUse_ += b;
            bytes += b;
            ids.push_back(next);
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            const double tAdmit = now();
            for (int id : ids) {
                reqs_[id]->tAdmit = tAdmit;
                reqs_[id]->status = 1;
                reqs_[id]->batchSize = (int)ids.size();
            }
        }
        // 4. solve the batch on this worker's stream
        std::vector<SolvePerf> perf;
        std::string err;
        std::vector<std::vector<std::vector<double>>> fields(ids.size());
        try {
            std::vector<const CoupledSystem *> ptrs;
            for (auto &s : systems) ptrs.push_back(s.get());
            batch.setSystems(ptrs);
            batch.solve(ctrl_, perf);
            if (keepFields_)
                for (size_t q = 0; q < ids.size(); ++q)
                    for (size_t r = 0; r < systems[q]->regions.size(); ++r) {
                        fields[q].emplace_back(systems[q]->regions[r].nCells);
                        batch.getPsi((int)q, (int)r, fields[q].back().data());
                    }
        } catch (const std::exception &e) {
            err = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            const double t = now();
            for (size_t q = 0; q < ids.size(); ++q) {
                ServeRequest &R = *reqs_[ids[q]];
                R.tDone = t;
                if (err.empty()) {
                    R.status = 2;
                    R.perf = perf[q];
                    R.psi = std::move(fields[q]);
                } else {
                    R.status = -1;
                    R.error = err;
                }
                --pending_;
            }
            inUse_ -= bytes;
            --inFlight_;
        }
        cv_.notify_all();
        done_.notify_all();
    }
}

}  // namespace mrfg
