// Serving: a request loop over coupled solves (continuous batching).
//
// Requests (case dictionaries) are queued as they arrive.  Worker threads,
// each with its own device-resident batch and HIP stream, take the oldest
// request, assemble it on the host, and admit it when its device footprint
// fits the HBM budget left by the requests in flight; a worker that finds
// more requests waiting takes them into the same batch (up to maxBatch
// systems), so one set of launches serves all of them.  Batches of different
// workers run concurrently on the device.  A system's result does not depend
// on which batch or worker solved it (per-system reductions are exact and in
// a fixed order).
#pragma once

#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch.hpp"

namespace mrfg {

struct ServeRequest {
    std::string text;
    double tSubmit = 0, tAdmit = 0, tDone = 0;   // seconds since the server started
    int status = 0;                              // 0 queued, 1 running, 2 done, -1 failed
    int batchSize = 0;
    size_t bytes = 0;
    SolvePerf perf;
    std::string error;
    std::vector<std::vector<double>> psi;        // per region, when fields are kept
};

class Server {
public:
    Server(int device, size_t hbmBudget, int nWorkers, int maxBatch, const SolverControls &c, bool keepFields);
    ~Server();
    int submit(const std::string &caseText);
    void waitAll();
    const ServeRequest &request(int id) const;
    int nRequests() const;
    // device bytes of one coupled system's layout and vectors (both precisions)
    static size_t footprint(const CoupledSystem &s);

private:
    void worker(int w);
    double now() const;

    int device_;
    size_t budget_, inUse_ = 0;
    int maxBatch_;
    SolverControls ctrl_;
    bool keepFields_;
    std::vector<std::unique_ptr<ServeRequest>> reqs_;
    std::deque<int> queue_;
    int inFlight_ = 0, pending_ = 0;
    bool stop_ = false;
    mutable std::mutex mu_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> threads_;
    double t0_;
};

}  // namespace mrfg
