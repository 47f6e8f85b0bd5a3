#pragma once
// Request server: continuous batching of simulate / signal / Jacobian requests
// on one GPU.  No counterpart in the reference (its host is a Julia session
// that calls the library one operation at a time).
//
// Requests reference objects that are already resident in HBM (tissue maps,
// sequences, trajectories, coil maps, residuals).  A dispatcher thread admits
// requests in arrival order while (a) a slot is free and (b) the estimated
// device memory of the in-flight requests stays within the budget.  Each slot
// is a private runtime (its own HIP stream and memory pool), so admitted
// requests run concurrently on the device and a request that arrives while
// others run starts as soon as a slot and memory free up.  Completion is
// detected with an event per request; results stay in HBM until taken.

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

  private:
    struct Entry {
        ServeRequest request;
        ServeResult result;
        size_t bytes = 0;
        int slot = -1;
        hipEvent_t done_event = nullptr;
        hipEvent_t inputs_ready = nullptr;  // recorded on the base stream at submission
        bool done = false;
    };

    void dispatch_loop();
    void launch(Entry& e, int slot);

    CompasContext base_;
    uint64_t id_;
    std::vector<CompasContext> slots_;
    std::vector<bool> slot_busy_;
    size_t budget_, in_flight_bytes_ = 0;
    std::chrono::steady_clock::time_point t0_;

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
