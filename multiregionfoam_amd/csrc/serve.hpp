// Serving: a request loop over coupled solves with continuous batching.
//
// Requests (case dictionaries) are queued as they arrive.  Two pools work
// on them:
//   * preparation threads assemble each request on the host and build its
//     device layout (DILU tile schedule, sliced ELL, interface rows), in
//     arrival order, off the device's critical path;
//   * device workers, each with its own resident batch and HIP stream
//     (a slot), take the oldest prepared requests whose HBM footprint fits
//     the budget left by the requests in flight -- up to maxBatch of them --
//     and solve them with one set of launches.  Slots run concurrently on
//     the device, so a request that arrives while others run is admitted as
//     soon as a slot and memory are free.
// A request is retired as soon as its own system has converged, at the
// batch's next convergence check: its field and performance record are
// published while the rest of its batch continues.  A system's result does
// not depend on which batch, slot or time it ran in (per-system reductions
// are exact and in a fixed order).
#pragma once

#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "batch.hpp"

namespace mrfg {

struct ServeRequest {
    std::string text;
    double tSubmit = 0, tAdmit = 0, tDone = 0;   // seconds since the server started
    double tReady = 0, tSolve = 0;               // prepared; device layout uploaded, solve starts
    int status = 0;                              // 0 waiting, 1 running, 2 done, -1 failed
    int batchSize = 0;
    size_t bytes = 0;
    SolvePerf perf;
    std::string error;
    std::vector<std::vector<double>> psi;        // per region, when fields are kept
    // the prepared request: assembled system and its device layout
    std::unique_ptr<CoupledSystem> sys;
    std::unique_ptr<SystemLayout> layout;
};

// A copy of a request's public state, taken under the server's lock.
struct ServeResult {
    int status = 0;
    double tSubmit = 0, tAdmit = 0, tDone = 0, tReady = 0, tSolve = 0;
    int batchSize = 0;
    SolvePerf perf;
    std::string error;
};

class Server {
public:
    // nPrep = 0: one preparation thread per host core share (OMP_NUM_THREADS,
    // else the hardware's), less the device workers, at least one
    Server(int device, size_t hbmBudget, int nWorkers, int maxBatch, const SolverControls &c, bool keepFields,
           int nPrep = 0);
    ~Server();
    int submit(const std::string &caseText);
    void waitAll();
    // false if id is unknown
    bool result(int id, ServeResult &out) const;
    // copy region `region`'s kept field of a finished request; false if none
    bool copyPsi(int id, int region, double *out) const;
    int nRequests() const;
    // device bytes of one coupled system's layout and vectors (both precisions)
    static size_t footprint(const CoupledSystem &s);

private:
    void prepare();
    void worker();
    double now() const;
    void fail(int id, const std::string &why);   // under mu_

    int device_;
    size_t budget_, inUse_ = 0;
    int maxBatch_;
    SolverControls ctrl_;
    bool keepFields_;
    std::vector<std::unique_ptr<ServeRequest>> reqs_;
    std::deque<int> arrivals_;    // submitted, not yet prepared
    std::set<int> ready_;         // prepared, waiting for a slot (oldest first)
    int inFlight_ = 0, pending_ = 0;
    bool stop_ = false;
    mutable std::mutex mu_;
    std::condition_variable cvArrive_;   // arrivals for the preparation threads
    std::condition_variable cvReady_;    // prepared requests or released memory, for the workers
    std::condition_variable cvDone_;     // finished requests, for waitAll
    std::vector<std::thread> prep_, workers_;
    double t0_;
};

}  // namespace mrfg
