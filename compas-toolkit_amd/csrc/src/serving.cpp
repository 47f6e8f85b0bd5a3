// Request server (compas/serving.h): continuous batching on one GPU.
//
// Three stages, all driven by one dispatcher thread that never waits on the
// host for the GPU (device events order the streams):
//   1. Admission by memory budget (size-aware, see serving.h).  An admitted
//      request's simulations become jobs: the echoes, and for J^H.y one
//      perturbed simulation per finite-difference field.  Job memory (echoes,
//      FISP RF tables, pSSFP partial planes, perturbed tissue copies) comes
//      from the pool of sim_[0], set up on its stream after the request's
//      inputs are ready.
//   2. Work-table launches.  Each launch takes up to batch_waves() work items
//      (one wave each) from the active jobs, least GPU time used first (each
//      job's virtual time grows by its item cost), groups them by kernel
//      instance and launches fisp_epg_table / pssfp_iso_table over a table
//      uploaded from pinned memory.  Launches alternate over two streams with
//      at most two in flight.  When a job's last item is launched, its
//      completion step is enqueued on sim_[0] (after the other stream's
//      launches that held its items): the pSSFP group sum, and for a
//      finite-difference job (sim(p + delta e_f) - echoes) / delta.
//   3. Post-processing.  Once every job of a request is complete, the
//      request's contraction (signal, or J^H.y) runs on a free slot stream
//      that waits for the simulation stream's event.
// Requests whose simulation has no work-table kernel (FISP with more than 32
// states) run every call on a slot stream, as one unit.

#include "compas/serving.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <set>

#include "compas/device/fisp_common.h"
#include "compas/device/pssfp_common.h"
#include "compas/internal.h"

namespace compas {

namespace {

constexpr float kDelta = 1e-4f;  // finite-difference step (derivative.h:21-47 default)
// Work items (waves) per launch.  pSSFP: 2048, two waves per SIMD of a
// 256-CU GPU (the FP64 recursion needs the second wave to hide its latency).
// FISP: 1024, one wave per SIMD -- a FISP wave issues VALU on most of its
// SIMD's cycles alone, and the contractions of other requests run beside it
// (measured on the simulate + signal steps of 50176 voxels x 1120 TRs: 2.26
// ms per step with 1024, 2.38 with 2048).  COMPAS_SERVER_BATCH_WAVES sets
// both (measurement).
int batch_waves(bool pssfp) {
    static const int forced = [] {
        const char* e = std::getenv("COMPAS_SERVER_BATCH_WAVES");
        const int v = e ? std::atoi(e) : 0;
        return v >= 64 && v <= 65536 ? v : 0;
    }();
    if (forced) return forced;
    return pssfp ? 2048 : 1024;
}

constexpr int kMaxBatchJobs = 256;  // simulations per launch
constexpr int kMaxActive = 128;  // admitted requests at once
constexpr double kMaxPassSeconds = 0.5;  // a request that does not fit may be passed this long after arrival
constexpr size_t kSmallShare = 64;  // requests of at most budget / kSmallShare pass a held queue

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
    if (r.fisp) return simulate_magnetization_derivative(ctx, field, echos, r.parameters, *r.fisp, kDelta);
    return simulate_magnetization_derivative(ctx, field, echos, r.parameters, *r.pssfp, kDelta);
}

// fields of the finite differences of a J^H.y request, in job order
int fd_field(size_t i) {
    static const int fields[4] = {TissueParameterField::T1, TissueParameterField::T2, TissueParameterField::B1,
                                  TissueParameterField::B0};
    return fields[i];
}

size_t fd_count(const ServeRequest& r) {
    return r.kind == RequestKind::JacobianHermitian ? (r.cartesian ? 2 : 4) : 0;
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

size_t align16(size_t n) {
    return (n + 15) & ~size_t(15);
}

}  // namespace

struct Server::Event {
    hipEvent_t ev = nullptr;
    Event() { COMPAS_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming)); }
    ~Event() {
        if (ev) (void)hipEventDestroy(ev);
    }
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    bool complete() const { return hipEventQuery(ev) != hipErrorNotReady; }
};

// One simulation of a request in the work tables.
struct Server::SimJob {
    int64_t ticket = -1;
    size_t index = 0;  // position in the request's jobs (0: the echoes)
    int field = -1;  // finite-difference field of this job, -1 for the echoes
    bool pssfp = false;
    int bucket = 0;  // kernel instance: FISP states per lane, pSSFP slice positions per lane
    int vblocks = 0, groups = 1;
    int64_t items = 0, next = 0;  // work items, and the first not yet launched
    double item_cost = 1.0, vtime = 0.0;  // fair share: GPU work launched so far
    bool completed = false;  // completion step enqueued
    TissueParameters params;  // the request's, or the perturbed copy
    kernels::FispArgs fisp {};
    kernels::PssfpTableJob pss {};
    Array<cfloat, 2> echos, partial;
    Array<float4, 1> table;
    std::shared_ptr<Event> last[Server::kSimStreams];  // last launch holding items of this job, per stream
};

namespace {
std::atomic<uint64_t> g_next_server_id {1};

// the job's simulation set up on `sim` (its stream already ordered after the
// request's inputs): output arrays, RF table or partial planes, the kernel
// arguments and the work items
std::shared_ptr<Server::SimJob> make_job(const CompasContext& sim, const ServeRequest& r, int64_t ticket, size_t k) {
    auto j = std::make_shared<Server::SimJob>();
    j->ticket = ticket;
    j->index = k;
    j->field = k == 0 ? -1 : fd_field(k - 1);
    j->params = k == 0 ? r.parameters : perturbed_parameters(sim, r.parameters, j->field, kDelta);
    const int nv = j->params.nvoxels;
    j->vblocks = int(dev::div_ceil(int64_t(nv), int64_t(64)));
    if (r.fisp) {
        const FISPSequence& seq = *r.fisp;
        const int nTR = int(seq.RF_train.size(0)), nslices = int(seq.sliceprofiles.size(0));
        COMPAS_CHECK(fisp_table_supported(seq.max_state));
        COMPAS_CHECK(seq.sliceprofiles.size(1) == nTR);  // fisp.cu:22
        COMPAS_CHECK(seq.undersampling_factor >= 1);
        COMPAS_CHECK(seq.repetitions >= 1);
        const int64_t nr = int64_t(nTR) * seq.undersampling_factor;
        j->bucket = seq.max_state;
        j->echos = sim.empty<cfloat, 2>({nr, nv});
        if (nv > 0 && nr > 0 && nslices == 0) sim.fill_zero(j->echos.data(), j->echos.nbytes());
        if (nv > 0 && nr > 0 && nslices > 0) {
            j->table = fisp_rf_table(sim, seq);
            kernels::FispArgs& a = j->fisp;
            a.tissue = dev::tissue_view(j->params);
            a.table = j->table.data();
            a.nTR = nTR;
            a.nslices = nslices;
            a.TR = seq.TR;
            a.TE = seq.TE;
            a.TW = seq.TW;
            a.TI = seq.TI;
            a.N = seq.max_state;
            a.usf = seq.undersampling_factor;
            a.repetitions = seq.repetitions;
            a.inversion = seq.inversion_prepulse;
            a.wait_spoiling = seq.wait_spoiling;
            a.echos = j->echos.data();
            j->items = j->vblocks;
            j->item_cost = double(nTR) * nslices * seq.repetitions * (26.0 * seq.max_state + 90.0);
        }
    } else {
        const pSSFPSequence& seq = *r.pssfp;
        const int nTR = int(seq.RF_train.size(0)), nz = int(seq.z.size(0)), nRF = int(seq.gamma_dt_RF.size(0));
        COMPAS_CHECK(nTR > 0);  // pssfp.h:176
        j->pssfp = true;
        j->echos = sim.empty<cfloat, 2>({nTR, nv});
        if (nv > 0 && nz == 0) sim.fill_zero(j->echos.data(), j->echos.nbytes());
        if (nv > 0 && nz > 0) {
            const int zpl = pssfp_table_zpl(nz);
            j->bucket = zpl;
            j->groups = (nz + zpl - 1) / zpl;
            const int64_t plane = int64_t(nTR) * nv;
            if (j->groups > 1) j->partial = sim.empty<cfloat, 2>({j->groups - 1, plane});
            kernels::PssfpArgs& a = j->pss.a;
            a.tissue = dev::tissue_view(j->params);
            a.RF = seq.RF_train.data();
            a.gamma_dt_RF = seq.gamma_dt_RF.data();
            a.z = seq.z.data();
            a.nz = nz;
            a.nTR = nTR;
            a.nRF = nRF;
            a.dt_ex = seq.dt.ex;
            a.dt_inv = seq.dt.inv;
            a.dt_pr = seq.dt.pr;
            a.g_ex = seq.gamma_dt_GRz.ex;
            a.g_pr = seq.gamma_dt_GRz.pr;
            a.echos = j->echos.data();
            j->pss.partial = j->partial.data();
            j->items = int64_t(j->vblocks) * j->groups;
            j->item_cost = double(nTR + 1) * std::max(nRF, 1) * (zpl * 50.0 + 12.0) + double(nTR) * zpl * 40.0;
        }
    }
    return j;
}
}  // namespace

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
    for (int i = 0; i < kSimStreams; i++) {
        auto rt = std::make_shared<Runtime>(base.device());
        rt->set_autotune(false);
        sim_.emplace_back(rt);
    }
    for (int r = 0; r < kRing; r++) {
        COMPAS_HIP_CHECK(hipHostMalloc(&host_ring_[r], kRingBytes, hipHostMallocDefault));
        dev_ring_[r] = sim_[0].empty<uint8_t, 1>({int64_t(kRingBytes)});
    }
    sim_[0].activate();
    COMPAS_HIP_CHECK(hipEventCreateWithFlags(&alloc_event_, hipEventDisableTiming));
    COMPAS_HIP_CHECK(hipEventRecord(alloc_event_, sim_[0].stream()));
    thread_ = std::thread([this] { dispatch_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lock(mutex_);
        stop_ = true;
    }
    cv_work_.notify_all();
    thread_.join();
    for (auto& s : sim_) (void)hipStreamSynchronize(s.stream());
    for (auto& r : ring_busy_) r.reset();
    for (int r = 0; r < kRing; r++) {
        if (host_ring_[r]) (void)hipHostFree(host_ring_[r]);
    }
    if (alloc_event_) (void)hipEventDestroy(alloc_event_);
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
    if (r.pssfp) {  // partial planes of one simulation's slice-position groups
        const int nz = int(r.pssfp->z.size(0));
        if (nz > 0) {
            const int zpl = pssfp_table_zpl(nz);
            bytes += size_t((nz + zpl - 1) / zpl - 1) * echo;
        }
    }
    if (r.kind == RequestKind::SimulateSignal) {
        bytes += 8 * nc * nr * ns * 17;  // signal + split-K partials
        bytes += r.cartesian ? 8 * nc * ns * nv : echo;  // operand W, or the spiral amplitudes
    } else {
        const size_t nder = fd_count(r);
        // derivatives (each computed in place over its perturbed simulation),
        // the decay operand and the partial sums of the contraction
        bytes += nder * echo + 16 * nv * ns + 2 * echo;
    }
    return bytes;
}

int64_t Server::submit(ServeRequest request) {
    trajectory_of(request);  // validates
    COMPAS_CHECK(request.fisp || request.pssfp);
    auto e = std::make_unique<Entry>();
    e->bytes = estimate_bytes(request);
    e->table_path = request.pssfp ? true : fisp_table_supported(request.fisp->max_state);
    // the request's inputs may still be in flight on the base stream (uploads,
    // a residual from compute_residual, arrays shared with another framework
    // and ordered onto it): its first kernels wait for this point
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

// Every call of a request on a slot stream (requests without a work-table
// kernel for their simulation).
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
            e.result.derivatives = {d1, d2};
        } else {
            auto db1 = derivative(ctx, TissueParameterField::B1, echos, r);
            auto db0 = derivative(ctx, TissueParameterField::B0, echos, r);
            e.result.jhy = compute_jacobian_hermitian(ctx, echos, d1, d2, db1, db0, r.parameters, *r.spiral, r.coils,
                                                      r.residual, r.jacobian_method);
            e.result.derivatives = {d1, d2, db1, db0};
        }
        e.result.echos = echos;
    }
    COMPAS_HIP_CHECK(hipEventCreateWithFlags(&e.done_event, hipEventDisableTiming));
    COMPAS_HIP_CHECK(hipEventRecord(e.done_event, ctx.stream()));
}

// The request's contraction on a slot stream, after its simulations.
void Server::post_process(Entry& e, int slot) {
    const CompasContext& ctx = slots_[size_t(slot)];
    ctx.activate();
    COMPAS_HIP_CHECK(hipStreamWaitEvent(ctx.stream(), e.inputs_ready, 0));
    COMPAS_HIP_CHECK(hipStreamWaitEvent(ctx.stream(), e.sim_done, 0));
    const ServeRequest& r = e.request;
    const Array<cfloat, 2>& echos = e.jobs[0]->echos;
    if (r.kind == RequestKind::SimulateSignal) {
        e.result.signal = magnetization_to_signal(ctx, echos, r.parameters, trajectory_of(r), r.coils, r.signal_method);
    } else if (r.cartesian) {
        e.result.jhy = compute_jacobian_hermitian(ctx, echos, e.derivs[0], e.derivs[1], r.parameters, *r.cartesian,
                                                  r.coils, r.residual, r.jacobian_method);
    } else {
        e.result.jhy = compute_jacobian_hermitian(ctx, echos, e.derivs[0], e.derivs[1], e.derivs[2], e.derivs[3],
                                                  r.parameters, *r.spiral, r.coils, r.residual, r.jacobian_method);
    }
    e.result.echos = echos;
    e.result.derivatives = e.derivs;
    COMPAS_HIP_CHECK(hipEventCreateWithFlags(&e.done_event, hipEventDisableTiming));
    COMPAS_HIP_CHECK(hipEventRecord(e.done_event, ctx.stream()));
}

// Sets up the request's next simulations (all FISP ones at once; pSSFP one at
// a time, since each holds partial planes of its slice-position groups).
void Server::activate_jobs(int64_t ticket, Entry& e) {
    const CompasContext& sim = sim_[0];
    bool any = false;
    while (e.next_job < e.jobs.size()) {
        if (e.request.pssfp) {
            bool busy = false;
            for (size_t k = 0; k < e.next_job; k++) busy |= e.jobs[k] && !e.jobs[k]->completed;
            if (busy) break;
        }
        const size_t k = e.next_job++;
        auto job = make_job(sim, e.request, ticket, k);
        e.jobs[k] = job;
        // a new simulation starts at the least virtual time of the active ones
        double vmin = 0.0;
        bool first = true;
        for (auto& a : active_jobs_) {
            if (a->next < a->items) {
                vmin = first ? a->vtime : std::min(vmin, a->vtime);
                first = false;
            }
        }
        job->vtime = vmin;
        if (job->items > 0) active_jobs_.push_back(job);
        any = true;
    }
    if (any) COMPAS_HIP_CHECK(hipEventRecord(alloc_event_, sim.stream()));
}

void Server::admit(int64_t ticket, Entry& e) {
    sim_[0].activate();
    COMPAS_HIP_CHECK(hipStreamWaitEvent(sim_[0].stream(), e.inputs_ready, 0));
    e.jobs.assign(1 + fd_count(e.request), nullptr);
    e.derivs.assign(fd_count(e.request), Array<cfloat, 2>());
    activate_jobs(ticket, e);
}

// Completion steps (on sim_[0]) of the jobs whose items are all launched; a
// finite-difference job completes after the request's echoes.
void Server::complete_jobs() {
    const CompasContext& sim = sim_[0];
    sim.activate();
    for (size_t i = 0; i < admitted_.size();) {
        Entry* e;
        {
            std::lock_guard<std::mutex> lock(mutex_);
            e = entries_.at(admitted_[i]).get();
        }
        bool progressed = false;
        for (size_t k = 0; k < e->next_job; k++) {
            SimJob& j = *e->jobs[k];
            if (j.completed || j.next < j.items) continue;
            if (k > 0 && !e->jobs[0]->completed) continue;
            for (int st = 1; st < kSimStreams; st++) {
                if (j.last[st]) COMPAS_HIP_CHECK(hipStreamWaitEvent(sim.stream(), j.last[st]->ev, 0));
            }
            if (j.pssfp && j.groups > 1) {
                const int64_t plane = j.echos.numel();
                pssfp_sum_groups(sim, j.echos.data(), j.partial.data(), j.groups - 1, plane, plane);
            }
            j.partial = Array<cfloat, 2>();  // back to sim_[0]'s pool, in its stream order
            if (j.field >= 0) {
                e->derivs[k - 1] = finite_difference(sim, e->jobs[0]->echos, j.echos, kDelta);
                j.echos = Array<cfloat, 2>();
            }
            j.completed = true;
            e->jobs_done++;
            progressed = true;
        }
        if (progressed) {
            active_jobs_.erase(std::remove_if(active_jobs_.begin(), active_jobs_.end(),
                                              [](const std::shared_ptr<SimJob>& j) { return j->completed; }),
                               active_jobs_.end());
        }
        if (e->jobs_done == e->jobs.size()) {
            COMPAS_HIP_CHECK(hipEventCreateWithFlags(&e->sim_done, hipEventDisableTiming));
            COMPAS_HIP_CHECK(hipEventRecord(e->sim_done, sim.stream()));
            post_queue_.push_back(admitted_[i]);
            admitted_.erase(admitted_.begin() + std::ptrdiff_t(i));
            continue;
        }
        if (progressed) activate_jobs(admitted_[i], *e);
        i++;
    }
}

// One work-table launch over the active jobs, on a stream with no launch in
// flight, if a ring slot is free.
void Server::build_and_launch_batch() {
    int ring = -1, s = -1;
    for (int r = 0; r < kRing; r++) {
        if (ring_busy_[r] && ring_busy_[r]->complete()) ring_busy_[r].reset();
        if (!ring_busy_[r] && ring < 0) ring = r;
    }
    for (int st = 0; st < kSimStreams; st++) {
        if (stream_busy_[st] && stream_busy_[st]->complete()) stream_busy_[st].reset();
        if (!stream_busy_[st] && s < 0) s = st;
    }
    if (ring < 0 || s < 0) return;

    // One kernel instance per launch (a stream runs its kernels one after the
    // other, so two instances in one launch would serialise): the instance
    // of the least-served job.  Then fair share within it: repeatedly the job
    // with the least virtual time gives one item.
    int best = -1;
    for (size_t i = 0; i < active_jobs_.size(); i++) {
        const SimJob& j = *active_jobs_[i];
        if (j.next < j.items && (best < 0 || j.vtime < active_jobs_[size_t(best)]->vtime)) best = int(i);
    }
    if (best < 0) return;
    const bool want_pssfp = active_jobs_[size_t(best)]->pssfp;
    const int want_bucket = active_jobs_[size_t(best)]->bucket;
    using Key = std::pair<double, size_t>;
    std::priority_queue<Key, std::vector<Key>, std::greater<Key>> heap;
    for (size_t i = 0; i < active_jobs_.size(); i++) {
        const SimJob& j = *active_jobs_[i];
        if (j.next < j.items && j.pssfp == want_pssfp && j.bucket == want_bucket) heap.push({j.vtime, i});
    }
    std::vector<int> local(active_jobs_.size(), -1);  // job -> index among this launch's jobs
    std::vector<size_t> jobs;  // active_jobs_ indices in the launch
    std::vector<std::pair<size_t, int64_t>> picks;  // (launch job, item)
    const int max_items = batch_waves(want_pssfp);
    while (!heap.empty() && int(picks.size()) < max_items) {
        const size_t i = heap.top().second;
        heap.pop();
        SimJob& j = *active_jobs_[i];
        if (local[i] < 0) {
            if (int(jobs.size()) >= kMaxBatchJobs) continue;
            local[i] = int(jobs.size());
            jobs.push_back(i);
        }
        picks.push_back({size_t(local[i]), j.next++});
        j.vtime += j.item_cost;
        if (j.next < j.items) heap.push({j.vtime, i});
    }

    // longest items first, so a launch ends with its short items and the
    // next launch's waves fill the SIMDs its long items leave
    std::stable_sort(picks.begin(), picks.end(), [&](const auto& x, const auto& y) {
        return active_jobs_[jobs[x.first]]->item_cost > active_jobs_[jobs[y.first]]->item_cost;
    });

    // table layout: per kernel instance (FISP states / pSSFP positions per
    // lane) its job descriptors, then its items
    struct Instance {
        bool pssfp;
        int bucket;
        std::vector<int> job_ids;  // launch jobs of this instance
        std::vector<int> slot_of;  // launch job -> descriptor index in this instance
        size_t jobs_off = 0, items_off = 0;
        int nitems = 0;
    };
    std::vector<Instance> inst;
    std::vector<int> inst_of(jobs.size(), -1);
    for (size_t lj = 0; lj < jobs.size(); lj++) {
        const SimJob& j = *active_jobs_[jobs[lj]];
        int found = -1;
        for (size_t q = 0; q < inst.size(); q++) {
            if (inst[q].pssfp == j.pssfp && inst[q].bucket == j.bucket) found = int(q);
        }
        if (found < 0) {
            inst.push_back({j.pssfp, j.bucket, {}, std::vector<int>(jobs.size(), -1), 0, 0, 0});
            found = int(inst.size()) - 1;
        }
        inst_of[lj] = found;
        inst[size_t(found)].slot_of[lj] = int(inst[size_t(found)].job_ids.size());
        inst[size_t(found)].job_ids.push_back(int(lj));
    }
    for (auto& p : picks) inst[size_t(inst_of[p.first])].nitems++;
    size_t off = 0;
    for (auto& q : inst) {
        q.jobs_off = off;
        off = align16(off + q.job_ids.size() * (q.pssfp ? sizeof(kernels::PssfpTableJob) : sizeof(kernels::FispArgs)));
        q.items_off = off;
        off = align16(off + size_t(q.nitems) * (q.pssfp ? sizeof(int4) : sizeof(int2)));
    }
    COMPAS_CHECK(off <= kRingBytes);
    char* host = static_cast<char*>(host_ring_[ring]);
    for (auto& q : inst) {
        for (size_t d = 0; d < q.job_ids.size(); d++) {
            const SimJob& j = *active_jobs_[jobs[size_t(q.job_ids[d])]];
            if (q.pssfp) {
                std::memcpy(host + q.jobs_off + d * sizeof(kernels::PssfpTableJob), &j.pss, sizeof(j.pss));
            } else {
                std::memcpy(host + q.jobs_off + d * sizeof(kernels::FispArgs), &j.fisp, sizeof(j.fisp));
            }
        }
    }
    std::vector<int> fill(inst.size(), 0);
    std::set<int64_t> tickets;
    for (auto& p : picks) {
        const size_t qi = size_t(inst_of[p.first]);
        Instance& q = inst[qi];
        const SimJob& j = *active_jobs_[jobs[p.first]];
        tickets.insert(j.ticket);
        const int d = q.slot_of[p.first];
        const int n = fill[qi]++;
        if (q.pssfp) {
            const int vb = int(p.second % j.vblocks), g = int(p.second / j.vblocks);
            reinterpret_cast<int4*>(host + q.items_off)[n] = make_int4(d, vb, g, 0);
        } else {
            reinterpret_cast<int2*>(host + q.items_off)[n] = make_int2(d, int(p.second));
        }
    }

    const CompasContext& ctx = sim_[size_t(s)];
    ctx.activate();
    if (s != 0) COMPAS_HIP_CHECK(hipStreamWaitEvent(ctx.stream(), alloc_event_, 0));
    uint8_t* dev = dev_ring_[ring].data();
    COMPAS_HIP_CHECK(hipMemcpyAsync(dev, host, off, hipMemcpyHostToDevice, ctx.stream()));
    for (auto& q : inst) {
        if (q.pssfp) {
            launch_pssfp_table(ctx, q.bucket, reinterpret_cast<const kernels::PssfpTableJob*>(dev + q.jobs_off),
                               reinterpret_cast<const int4*>(dev + q.items_off), q.nitems);
        } else {
            launch_fisp_table(ctx, q.bucket, reinterpret_cast<const kernels::FispArgs*>(dev + q.jobs_off),
                              reinterpret_cast<const int2*>(dev + q.items_off), q.nitems);
        }
    }
    auto ev = std::make_shared<Event>();
    COMPAS_HIP_CHECK(hipEventRecord(ev->ev, ctx.stream()));
    ring_busy_[ring] = ev;
    stream_busy_[s] = ev;
    for (size_t lj = 0; lj < jobs.size(); lj++) active_jobs_[jobs[lj]]->last[s] = ev;

    std::lock_guard<std::mutex> lock(mutex_);
    stats_.launches++;
    stats_.items += int64_t(picks.size());
    stats_.max_requests = std::max<int64_t>(stats_.max_requests, int64_t(tickets.size()));
    if (tickets.size() >= 2) stats_.shared_launches++;
}

// Under the lock: the request is finished (results in place or error set).
void Server::finish_entry(Entry& e, const std::string& error) {
    if (!error.empty()) e.result.error = error;
    // the results are final: hand them to the base runtime, so reading them
    // does not queue behind later work on the streams that produced them
    e.result.echos = rehome(e.result.echos, base_.runtime());
    e.result.signal = rehome(e.result.signal, base_.runtime());
    e.result.jhy = rehome(e.result.jhy, base_.runtime());
    for (auto& d : e.result.derivatives) d = rehome(d, base_.runtime());
    e.result.t_done = now();
    for (hipEvent_t* ev : {&e.done_event, &e.inputs_ready, &e.sim_done}) {
        if (*ev) (void)hipEventDestroy(*ev);
        *ev = nullptr;
    }
    if (e.slot >= 0) slot_busy_[size_t(e.slot)] = false;
    // the results stay in HBM until released: they keep their share of the
    // budget until then, the rest of the request's estimate is free now
    size_t held = 0;
    auto add = [&](const auto& a) { held += size_t(a.numel()) * sizeof(cfloat); };
    add(e.result.echos);
    add(e.result.signal);
    add(e.result.jhy);
    for (auto& d : e.result.derivatives) add(d);
    e.held_bytes = std::min(held, e.bytes);
    in_flight_bytes_ -= e.bytes - e.held_bytes;
    e.request = ServeRequest {};  // drop the references to the inputs
    e.jobs.clear();
    e.derivs.clear();
    e.done = true;
    cv_done_.notify_all();
}

void Server::fail_simulations(const std::string& error) {
    // the failed launch's neighbours may still run: let every simulation
    // stream drain before the jobs' memory and the table ring are reused
    for (auto& c : sim_) (void)hipStreamSynchronize(c.stream());
    (void)hipGetLastError();
    std::vector<int64_t> failed;
    failed.swap(admitted_);
    active_jobs_.clear();  // every active job belongs to a request of admitted_
    std::lock_guard<std::mutex> lock(mutex_);
    for (int64_t t : failed) finish_entry(*entries_.at(t), error);
}

void Server::dispatch_loop() {
    sim_[0].activate();
    std::unique_lock<std::mutex> lock(mutex_);
    for (;;) {
        // ---- completions of requests whose last kernels were enqueued on a slot
        for (size_t i = 0; i < running_.size();) {
            Entry& e = *entries_[running_[i]];
            const hipError_t q = hipEventQuery(e.done_event);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            finish_entry(e, q == hipSuccess ? std::string() : std::string(hipGetErrorString(q)));
            running_.erase(running_.begin() + std::ptrdiff_t(i));
        }

        // ---- admissions by memory budget: arrival order; later requests may
        // pass one that does not fit until it has waited kMaxPassSeconds, after
        // which it holds the queue until it fits -- except for small requests
        // (at most budget / kSmallShare), which still pass: they hold their
        // memory for tens of milliseconds, so they delay the held request
        // little, and holding them would queue them behind its wait
        std::vector<std::pair<int64_t, Entry*>> admitted_now;
        {
            size_t active = admitted_.size() + post_queue_.size() + running_.size();
            const double t = now();
            const size_t small = budget_ / kSmallShare;
            bool held = false;
            for (auto it = pending_.begin(); it != pending_.end() && active < size_t(kMaxActive);) {
                Entry& e = *entries_[*it];
                const bool nothing_running = active == 0 && admitted_now.empty();
                const bool may_pass = !held || e.bytes <= small;
                if (may_pass && (nothing_running || in_flight_bytes_ + e.bytes <= budget_)) {
                    in_flight_bytes_ += e.bytes;
                    e.result.t_start = t;
                    admitted_now.push_back({*it, &e});
                    it = pending_.erase(it);
                    active++;
                    continue;
                }
                if (!held && t - e.result.t_arrival > kMaxPassSeconds) held = true;
                ++it;
            }
        }
        lock.unlock();

        for (auto& [ticket, e] : admitted_now) {
            if (!e->table_path) {
                post_queue_.push_back(ticket);  // every call on a slot
                continue;
            }
            std::string error;
            try {
                admit(ticket, *e);
            } catch (const std::exception& ex) {
                error = ex.what();
            }
            if (error.empty()) {
                admitted_.push_back(ticket);
            } else {
                // the simulation stream may hold part of the request's set-up
                (void)hipStreamSynchronize(sim_[0].stream());
                for (auto& j : e->jobs) {
                    if (!j) continue;
                    active_jobs_.erase(std::remove(active_jobs_.begin(), active_jobs_.end(), j), active_jobs_.end());
                }
                lock.lock();
                finish_entry(*e, error);
                lock.unlock();
            }
        }

        // ---- work-table launches and the completion steps they allow; an
        // error there (a failed launch, device memory exhausted) fails the
        // requests whose simulations are not complete, and the server goes on
        try {
            for (int round = 0; round < 2; round++) {
                complete_jobs();
                build_and_launch_batch();
            }
            complete_jobs();
        } catch (const std::exception& ex) {
            fail_simulations(ex.what());
        }

        // ---- contractions (or whole requests) on free slots
        lock.lock();
        while (!post_queue_.empty()) {
            const auto free_slot = std::find(slot_busy_.begin(), slot_busy_.end(), false);
            if (free_slot == slot_busy_.end()) break;
            const int64_t ticket = post_queue_.front();
            post_queue_.pop_front();
            Entry& e = *entries_[ticket];
            const int slot = int(free_slot - slot_busy_.begin());
            slot_busy_[size_t(slot)] = true;
            e.slot = slot;
            lock.unlock();
            std::string error;
            try {
                if (e.table_path) {
                    post_process(e, slot);
                } else {
                    launch(e, slot);
                }
            } catch (const std::exception& ex) {
                error = ex.what();
            }
            lock.lock();
            if (error.empty()) {
                running_.push_back(ticket);
            } else {
                (void)hipStreamSynchronize(slots_[size_t(slot)].stream());
                finish_entry(e, error);
            }
        }

        const bool idle = running_.empty() && admitted_.empty() && post_queue_.empty();
        if (stop_ && pending_.empty() && idle) break;
        if (idle && pending_.empty()) {
            cv_work_.wait(lock, [&] { return stop_ || !pending_.empty(); });
        } else {
            cv_work_.wait_for(lock, std::chrono::microseconds(50));
        }
    }
}

Server::BatchStats Server::batch_stats() {
    std::lock_guard<std::mutex> lock(mutex_);
    return stats_;
}

void Server::set_kernel_timing(bool on) {
    for (auto& c : slots_) c.runtime()->set_timing(on);
    for (auto& c : sim_) c.runtime()->set_timing(on);
}

std::map<std::string, KernelStats> Server::collect_timings() {
    std::map<std::string, KernelStats> out;
    auto merge = [&](const CompasContext& c) {
        for (auto& kv : c.runtime()->collect_timings()) {
            auto& o = out[kv.first];
            o.total_ms += kv.second.total_ms;
            o.count += kv.second.count;
        }
    };
    for (auto& c : slots_) merge(c);
    for (auto& c : sim_) merge(c);
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
    in_flight_bytes_ -= it->second->held_bytes;
    entries_.erase(it);
}

}  // namespace compas
