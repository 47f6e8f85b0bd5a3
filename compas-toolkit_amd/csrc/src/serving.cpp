// Request server (compas/serving.h): continuous batching on one GPU.
//
// Slots are private runtimes (own stream, own stream-ordered pool).  The
// dispatcher thread admits pending requests in arrival order when a slot is
// free and the in-flight memory estimate fits the budget (a request larger
// than the whole budget runs when nothing else does), enqueues the request's
// library calls on the slot's stream, records an event after them, and polls
// the events of running requests; a finished request frees its slot and its
// memory share at once, so the next arrival starts while the others run.

#include "compas/serving.h"

#include <algorithm>

#include "compas/internal.h"

namespace compas {

namespace {

const Trajectory& trajectory_of(const ServeRequest& r) {
    if (r.cartesian) return *r.cartesian;
    COMPAS_CHECK(r.spiral != nullptr);
    return *r.spiral;
}

Array<cfloat, 2> simulate(const CompasContext& ctx, const ServeRequest& r) {
    if (r.fisp) return simulate_magnetization(ctx, r.parameters, *r.fisp);
    COMPAS_CHECK(r.pssfp != nullptr);
    return simulate_magnetization(ctx, r.parameters, *r.pssfp);
}

Array<cfloat, 2> derivative(const CompasContext& ctx, int field, const Array<cfloat, 2>& echos, const ServeRequest& r) {
    if (r.fisp) return simulate_magnetization_derivative(ctx, field, echos, r.parameters, *r.fisp);
    return simulate_magnetization_derivative(ctx, field, echos, r.parameters, *r.pssfp);
}

// A view of a finished result on another runtime: same memory, kept alive by
// the view and read in that runtime's stream order.  When the view goes, the
// producing runtime's stream waits (an event, no host wait) for the work
// queued on the view's runtime before the memory returns to its pool.
template<typename T, size_t N>
Array<T, N> rehome(const Array<T, N>& a, const std::shared_ptr<Runtime>& rt) {
    if (!a.buffer()) return a;
    std::shared_ptr<Buffer> orig = a.buffer();
    auto release = [orig, rt] {
        rt->activate();
        hipEvent_t ev;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            rt->synchronize();  // no event: order by waiting instead
            return;
        }
        COMPAS_HIP_CHECK(hipEventRecord(ev, rt->stream()));
        COMPAS_HIP_CHECK(hipStreamWaitEvent(orig->runtime()->stream(), ev, 0));
        (void)hipEventDestroy(ev);
    };
    auto view = std::make_shared<Buffer>(rt, orig->data(), orig->nbytes(), release, false);
    return Array<T, N>(std::move(view), a.sizes());
}

}  // namespace

namespace {
std::atomic<uint64_t> g_next_server_id {1};
}

Server::Server(const CompasContext& base, int slots, size_t memory_budget) :
    base_(base), id_(g_next_server_id++), budget_(memory_budget), t0_(std::chrono::steady_clock::now()) {
    COMPAS_CHECK(slots >= 1 && slots <= 32);
    base.activate();
    for (int i = 0; i < slots; i++) {
        auto rt = std::make_shared<Runtime>(base.device());
        // no autotuning trials on the request path: request sizes vary per
        // request, so every request would time every tile variant; shapes
        // tuned elsewhere keep their tile, new ones run the default
        rt->set_autotune(false);
        slots_.emplace_back(rt);
    }
    slot_busy_.assign(size_t(slots), false);
    thread_ = std::thread([this] { dispatch_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lock(mutex_);
        stop_ = true;
    }
    cv_work_.notify_all();
    thread_.join();
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

size_t Server::estimate_bytes(const ServeRequest& r) const {
    const auto& t = trajectory_of(r);
    const size_t nv = size_t(r.parameters.nvoxels), nr = size_t(t.nreadouts), ns = size_t(t.samples_per_readout);
    const size_t nc = size_t(r.coils.size(0));
    const size_t echo = 8 * nr * nv;
    size_t bytes = echo + (size_t(16) << 20);
    if (r.pssfp) {  // partial echoes of the slice-position groups
        bytes += size_t(pssfp_scratch_bytes(slots_.front(), int(r.pssfp->z.size(0)), int(nv),
                                            int(r.pssfp->RF_train.size(0))));
    }
    if (r.kind == RequestKind::SimulateSignal) {
        bytes += 8 * nc * nr * ns * 17;  // signal + split-K partials
        bytes += r.cartesian ? 8 * nc * ns * nv : echo;  // operand W, or the spiral amplitudes
    } else {
        const size_t nder = r.cartesian ? 2 : 4;
        bytes += nder * echo + echo;  // derivatives + one perturbed simulation in flight
        bytes += 16 * nv * ns + 2 * echo;  // decay operand, partial sums / per-coil buffers
    }
    return bytes;
}

int64_t Server::submit(ServeRequest request) {
    trajectory_of(request);  // validates
    COMPAS_CHECK(request.fisp || request.pssfp);
    auto e = std::make_unique<Entry>();
    e->bytes = estimate_bytes(request);
    // the request's inputs may still be in flight on the base stream (uploads,
    // a residual from compute_residual, arrays shared with another framework
    // and ordered onto it): the slot waits for this point before its first kernel
    base_.activate();
    COMPAS_HIP_CHECK(hipEventCreateWithFlags(&e->inputs_ready, hipEventDisableTiming));
    COMPAS_HIP_CHECK(hipEventRecord(e->inputs_ready, base_.stream()));
    e->result.return_echos = request.return_echos;
    e->request = std::move(request);
    std::lock_guard<std::mutex> lock(mutex_);
    COMPAS_CHECK(!stop_);
    const int64_t ticket = next_ticket_++;
    e->result.t_arrival = now();
    entries_[ticket] = std::move(e);
    pending_.push_back(ticket);
    cv_work_.notify_one();
    return ticket;
}

void Server::launch(Entry& e, int slot) {
    const CompasContext& ctx = slots_[size_t(slot)];
    ctx.activate();
    COMPAS_HIP_CHECK(hipStreamWaitEvent(ctx.stream(), e.inputs_ready, 0));
    const ServeRequest& r = e.request;
    if (r.kind == RequestKind::SimulateSignal) {
        e.result.echos = simulate(ctx, r);
        e.result.signal = magnetization_to_signal(ctx, e.result.echos, r.parameters, trajectory_of(r), r.coils,
                                                  r.signal_method);
    } else {
        auto echos = simulate(ctx, r);
        auto d1 = derivative(ctx, TissueParameterField::T1, echos, r);
        auto d2 = derivative(ctx, TissueParameterField::T2, echos, r);
        if (r.cartesian) {
            e.result.jhy = compute_jacobian_hermitian(ctx, echos, d1, d2, r.parameters, *r.cartesian, r.coils,
                                                      r.residual, r.jacobian_method);
        } else {
            auto db1 = derivative(ctx, TissueParameterField::B1, echos, r);
            auto db0 = derivative(ctx, TissueParameterField::B0, echos, r);
            e.result.jhy = compute_jacobian_hermitian(ctx, echos, d1, d2, db1, db0, r.parameters, *r.spiral, r.coils,
                                                      r.residual, r.jacobian_method);
        }
        e.result.echos = echos;
    }
    COMPAS_HIP_CHECK(hipEventCreateWithFlags(&e.done_event, hipEventDisableTiming));
    COMPAS_HIP_CHECK(hipEventRecord(e.done_event, ctx.stream()));
}

void Server::dispatch_loop() {
    if (!slots_.empty()) slots_[0].activate();
    std::unique_lock<std::mutex> lock(mutex_);
    for (;;) {
        // ---- completions
        for (size_t i = 0; i < running_.size();) {
            Entry& e = *entries_[running_[i]];
            const hipError_t q = hipEventQuery(e.done_event);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            if (q != hipSuccess) e.result.error = hipGetErrorString(q);
            // the results are final: hand them to the base runtime, so reading
            // them does not queue behind the next request on this slot's stream
            e.result.echos = rehome(e.result.echos, base_.runtime());
            e.result.signal = rehome(e.result.signal, base_.runtime());
            e.result.jhy = rehome(e.result.jhy, base_.runtime());
            e.result.t_done = now();
            (void)hipEventDestroy(e.done_event);
            e.done_event = nullptr;
            (void)hipEventDestroy(e.inputs_ready);
            e.inputs_ready = nullptr;
            slot_busy_[size_t(e.slot)] = false;
            in_flight_bytes_ -= e.bytes;
            e.request = ServeRequest {};  // drop the references to the inputs
            e.done = true;
            running_.erase(running_.begin() + std::ptrdiff_t(i));
            cv_done_.notify_all();
        }
        // ---- admissions, in arrival order
        while (!pending_.empty()) {
            const auto free_slot = std::find(slot_busy_.begin(), slot_busy_.end(), false);
            if (free_slot == slot_busy_.end()) break;
            const int64_t ticket = pending_.front();
            Entry& e = *entries_[ticket];
            if (!running_.empty() && in_flight_bytes_ + e.bytes > budget_) break;
            pending_.pop_front();
            const int slot = int(free_slot - slot_busy_.begin());
            slot_busy_[size_t(slot)] = true;
            in_flight_bytes_ += e.bytes;
            e.slot = slot;
            e.result.t_start = now();
            lock.unlock();
            std::string error;
            try {
                launch(e, slot);
            } catch (const std::exception& ex) {
                error = ex.what();
            }
            lock.lock();
            if (error.empty()) {
                running_.push_back(ticket);
            } else {
                // the slot's stream may hold part of the request: let it drain
                (void)hipStreamSynchronize(slots_[size_t(slot)].stream());
                (void)hipEventDestroy(e.inputs_ready);
                e.inputs_ready = nullptr;
                e.result.error = error;
                e.result.t_done = now();
                slot_busy_[size_t(slot)] = false;
                in_flight_bytes_ -= e.bytes;
                e.request = ServeRequest {};
                e.done = true;
                cv_done_.notify_all();
            }
        }
        if (stop_ && pending_.empty() && running_.empty()) break;
        if (running_.empty()) {
            cv_work_.wait(lock, [&] { return stop_ || !pending_.empty(); });
        } else {
            cv_work_.wait_for(lock, std::chrono::microseconds(50));
        }
    }
}

void Server::set_kernel_timing(bool on) {
    for (auto& c : slots_) c.runtime()->set_timing(on);
}

std::map<std::string, KernelStats> Server::collect_timings() {
    std::map<std::string, KernelStats> out;
    for (auto& c : slots_) {
        for (auto& kv : c.runtime()->collect_timings()) {
            auto& o = out[kv.first];
            o.total_ms += kv.second.total_ms;
            o.count += kv.second.count;
        }
    }
    return out;
}

bool Server::wait(int64_t ticket, double timeout_seconds) {
    std::unique_lock<std::mutex> lock(mutex_);
    auto it = entries_.find(ticket);
    COMPAS_CHECK(it != entries_.end());
    Entry* e = it->second.get();
    return cv_done_.wait_for(lock, std::chrono::duration<double>(timeout_seconds), [&] { return e->done; });
}

ServeResult Server::result(int64_t ticket) {
    std::lock_guard<std::mutex> lock(mutex_);
    auto it = entries_.find(ticket);
    COMPAS_CHECK(it != entries_.end());
    COMPAS_CHECK(it->second->done);
    return it->second->result;
}

void Server::release(int64_t ticket) {
    std::lock_guard<std::mutex> lock(mutex_);
    auto it = entries_.find(ticket);
    COMPAS_CHECK(it != entries_.end());
    COMPAS_CHECK(it->second->done);
    entries_.erase(it);
}

}  // namespace compas
