#pragma once
// Request server: continuous batching of simulate / signal / Jacobian requests
// on one GPU.  No counterpart in the reference (its host is a Julia session
// that calls the library one operation at a time).
//
// Requests reference objects that are already resident in HBM (tissue maps,
// sequences, trajectories, coil maps, residuals).  A dispatcher thread:
//   * admits pending requests while their estimated device memory fits the
//     budget -- in arrival order, except that a request that does not fit
//     lets later ones that do fit pass until it has waited 0.5 s, after which
//     it holds the queue until it fits, except for small requests (at most
//     1/64 of the budget), which still pass; at most 128 requests at once;
//   * splits each admitted request's simulations (the echoes, and for J^H.y
//     the perturbed simulations of its finite differences) into work items
//     (voxel blocks, and slice-position groups for pSSFP) and runs them in
//     work-table launches: each launch takes items from every active
//     simulation in fair-share order (least GPU time used first), so one
//     launch serves several requests, small requests finish within a launch
//     or two, and large ones progress alongside instead of blocking the queue.
//     Requests that arrive join the next launch; up to three launches are in
//     flight on three streams, so the tail of one overlaps the others, and the
//     items of a launch go longest first;
//   * when a request's simulations are complete (device events, no host
//     wait), runs its signal or J^H.y contraction on one of S slot streams.
// Completion is detected with an event per request; results stay in HBM
// until released, and count against the memory budget until then.  A voxel's echoes are computed by the same arithmetic whatever
// else shares a launch, so a request's result is the same alone and in a mix.

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "compas/compas.h"

namespace compas {

enum struct RequestKind {
    SimulateSignal,  // simulate_magnetization + magnetization_to_signal
    JacobianHermitian,  // simulate (FISP), finite-difference derivatives, J^H.y
};

struct ServeRequest {
    RequestKind kind = RequestKind::SimulateSignal;
    TissueParameters parameters;
    std::shared_ptr<FISPSequence> fisp;  // exactly one of fisp / pssfp
    std::shared_ptr<pSSFPSequence> pssfp;
    std::shared_ptr<CartesianTrajectory> cartesian;  // exactly one of cartesian / spiral
    std::shared_ptr<SpiralTrajectory> spiral;
    Array<cfloat, 2> coils;
    Array<cfloat, 3> residual;  // J^H.y input y [nc][nr][ns]
    SimulateSignalMethod signal_method = SimulateSignalMethod::MatmulFast;
    JacobianComputeMethod jacobian_method = JacobianComputeMethod::GemmFast;
    bool return_echos = true;  // the result frame carries the echoes (wire submissions ask explicitly)
};

struct ServeResult {
    Array<cfloat, 2> echos;  // [nr][nv]
    Array<cfloat, 3> signal;  // SimulateSignal: [nc][nr][ns]
    Array<cfloat, 2> jhy;  // JacobianHermitian: [4 or 6][nv]
    std::vector<Array<cfloat, 2>> derivatives;  // JacobianHermitian: finite differences T1, T2 (, B1, B0)
    double t_arrival = 0, t_start = 0, t_done = 0;  // seconds since the server started
    std::string error;
    bool return_echos = true;  // copied from the request
};

class Server {
  public:
    Server(const CompasContext& base, int slots, size_t memory_budget);
    ~Server();  // waits for every admitted request
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    int64_t submit(ServeRequest request);
    // Blocks until the request is done (true) or the timeout passes (false).
    bool wait(int64_t ticket, double timeout_seconds);
    // Result of a finished request; the server keeps it until release().
    ServeResult result(int64_t ticket);
    void release(int64_t ticket);

    // Per-kernel HIP-event timing on every slot stream (merged when read).
    void set_kernel_timing(bool on);
    std::map<std::string, KernelStats> collect_timings();

    size_t estimate_bytes(const ServeRequest& r) const;
    double now() const;
    const CompasContext& base() const { return base_; }  // where callers upload request data
    uint64_t id() const { return id_; }  // unique per server in the process
    int slots() const { return int(slots_.size()); }

    // Work-table launches so far: launches, work items, the most requests
    // one launch served, and the launches that served two or more requests.
    struct BatchStats {
        int64_t launches = 0, items = 0, max_requests = 0, shared_launches = 0;
    };
    BatchStats batch_stats();

    struct SimJob;
    struct Event;

  private:
    struct Entry {
        ServeRequest request;
        ServeResult result;
        size_t bytes = 0;  // estimated device memory, counted against the budget from admission
        size_t held_bytes = 0;  // of which the results, counted until release
        int slot = -1;
        bool table_path = false;  // simulations in work-table launches (else all calls on a slot)
        std::vector<std::shared_ptr<SimJob>> jobs;  // [0]: the echoes; then one per finite difference
        size_t next_job = 0;  // jobs activated so far
        size_t jobs_done = 0;  // jobs whose completion step is enqueued
        std::vector<Array<cfloat, 2>> derivs;  // finite differences, in the order of the jobs
        hipEvent_t sim_done = nullptr;  // simulations complete (recorded on the simulation stream)
        hipEvent_t done_event = nullptr;
        hipEvent_t inputs_ready = nullptr;  // recorded on the base stream at submission
        bool done = false;
    };

    void dispatch_loop();
    void launch(Entry& e, int slot);  // every call of the request on a slot stream
    void post_process(Entry& e, int slot);  // the request's contraction after its simulations
    void admit(int64_t ticket, Entry& e);
    void activate_jobs(int64_t ticket, Entry& e);
    void build_and_launch_batch();
    void complete_jobs();
    void fail_simulations(const std::string& error);  // every request of admitted_ ends with `error`
    void finish_entry(Entry& e, const std::string& error);

    CompasContext base_;
    uint64_t id_;
    std::vector<CompasContext> slots_;
    std::vector<bool> slot_busy_;
    std::vector<CompasContext> sim_;  // streams of the work-table launches; sim_[0] holds the job memory
    size_t budget_, in_flight_bytes_ = 0;
    std::chrono::steady_clock::time_point t0_;

    // work-table launches: kSimStreams streams, at most one launch in flight
    // on each; tables staged in pinned host memory and copied to a ring of
    // device buffers, each reusable once its launch's event has completed
    static constexpr int kSimStreams = 3;
    static constexpr int kRing = 2 * kSimStreams;
    static constexpr size_t kRingBytes = size_t(1) << 20;
    void* host_ring_[kRing] = {};
    Array<uint8_t, 1> dev_ring_[kRing];
    std::shared_ptr<Event> ring_busy_[kRing];
    std::shared_ptr<Event> stream_busy_[kSimStreams];  // each stream's last launch
    hipEvent_t alloc_event_ = nullptr;  // recorded on sim_[0] after job memory is set up
    std::vector<std::shared_ptr<SimJob>> active_jobs_;
    std::vector<int64_t> admitted_;  // table-path requests whose simulations are not all complete
    std::deque<int64_t> post_queue_;  // simulations complete or enqueued, waiting for a slot
    BatchStats stats_;

    std::mutex mutex_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<int64_t> pending_;
    std::vector<int64_t> running_;
    std::map<int64_t, std::unique_ptr<Entry>> entries_;
    int64_t next_ticket_ = 0;
    bool stop_ = false;
    std::thread thread_;
};

}  // namespace compas
