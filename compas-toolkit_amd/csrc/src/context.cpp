#include "compas/core/context.h"

#include <set>
#include <sstream>

namespace compas {

void raise_error(const char* message, const char* file, int line, const char* func) {
    std::stringstream ss;
    ss << "COMPAS error occurred (" << file << ":" << line << ":" << func << "): " << message;
    throw CompasException(ss.str());
}

void check_launch(const char* name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::stringstream ss;
        ss << "kernel launch failed (" << name << "): " << hipGetErrorString(e);
        throw CompasException(ss.str());
    }
}

// Every live runtime, so that an allocation failing in one of them can
// reclaim the blocks cached by the others (several runtimes per device exist
// when a request server runs one per slot).
static std::mutex g_live_mutex;
static std::set<Runtime*>& live_runtimes() {
    static auto* s = new std::set<Runtime*>();  // never destroyed
    return *s;
}

Runtime::Runtime(int device) : device_(device) {
    COMPAS_HIP_CHECK(hipSetDevice(device));
    // default priority on purpose: streams of different priorities are not mixed
    COMPAS_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    std::lock_guard<std::mutex> lock(g_live_mutex);
    live_runtimes().insert(this);
}

Runtime::~Runtime() {
    // The per-device runtimes of the registry live until process exit.  A
    // private runtime (a server slot) ends when its last buffer is released:
    // all its blocks are cached by then, so give them back with its stream.
    {
        std::lock_guard<std::mutex> lock(g_live_mutex);
        live_runtimes().erase(this);
    }
    (void)hipSetDevice(device_);
    (void)hipStreamSynchronize(stream_);
    for (auto& kv : free_blocks_) (void)hipFree(kv.second);
    free_blocks_.clear();
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (auto e : event_pool_) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(stream_);
}

void Runtime::activate() const {
    COMPAS_HIP_CHECK(hipSetDevice(device_));
}

static size_t round_block(size_t n) {
    // 256-byte granularity below 1 MiB, 2 MiB above: limits fragmentation
    // while keeping recycled blocks close in size.
    if (n == 0) return 256;
    if (n < (size_t(1) << 20)) return (n + 255) & ~size_t(255);
    size_t g = size_t(2) << 20;
    return (n + g - 1) / g * g;
}

void* Runtime::allocate(size_t nbytes) {
    size_t n = round_block(nbytes);
    {
        std::lock_guard<std::mutex> lock(mutex_);
        auto it = free_blocks_.lower_bound(n);
        // accept a cached block up to 25% larger than requested
        if (it != free_blocks_.end() && it->first <= n + n / 4) {
            void* p = it->second;
            size_t sz = it->first;
            free_blocks_.erase(it);
            cached_ -= sz;
            in_use_ += sz;
            if (capture_id_ >= 0) graph_blocks_[p] = capture_id_;
            return p;
        }
    }
    activate();
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess && capturing()) {
        // no trimming while a capture is open (it synchronises the stream)
        (void)hipGetLastError();
        std::stringstream ss;
        ss << "device allocation of " << nbytes << " bytes failed during graph capture";
        throw CompasException(ss.str());
    }
    if (e != hipSuccess) {
        // out of memory: give the cache back and retry, then the caches of the
        // other runtimes on this device
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            {
                // holding the registry lock keeps the others alive while they trim
                std::lock_guard<std::mutex> lock(g_live_mutex);
                for (Runtime* r : live_runtimes())
                    if (r != this && r->device_ == device_) r->trim();
            }
            activate();
            e = hipMalloc(&p, n);
        }
    }
    if (e != hipSuccess) {
        // the caller may handle this (a request that does not fit ends with an
        // error): clear the thread's last error, or the next launch check
        // would report it as that launch's failure
        (void)hipGetLastError();
        std::stringstream ss;
        ss << "device allocation of " << nbytes << " bytes failed: " << hipGetErrorString(e);
        throw CompasException(ss.str());
    }
    std::lock_guard<std::mutex> lock(mutex_);
    in_use_ += n;
    return p;
}

void Runtime::release(void* ptr, size_t nbytes) {
    if (!ptr) return;
    size_t n = round_block(nbytes);
    std::lock_guard<std::mutex> lock(mutex_);
    auto g = graph_blocks_.find(ptr);
    if (g != graph_blocks_.end()) {  // a graph's replays still write it: hold it for the graph
        graph_held_[g->second].emplace_back(ptr, n);
        graph_blocks_.erase(g);
        return;
    }
    in_use_ -= n;
    cached_ += n;
    free_blocks_.emplace(n, ptr);
}

void Runtime::begin_capture() {
    activate();
    {
        std::lock_guard<std::mutex> lock(mutex_);
        if (capture_id_ >= 0) throw CompasException("a graph capture is already open on this context");
        capture_id_ = next_graph_id_++;
        timing_before_capture_ = timing_;
        timing_ = false;  // no timing events inside a graph
    }
    hipError_t e = hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lock(mutex_);
        capture_id_ = -1;
        timing_ = timing_before_capture_;
        COMPAS_HIP_CHECK(e);
    }
}

hipGraphExec_t Runtime::end_capture(int* id) {
    activate();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(stream_, &graph);
    int cid;
    {
        std::lock_guard<std::mutex> lock(mutex_);
        cid = capture_id_;
        capture_id_ = -1;
        timing_ = timing_before_capture_;
    }
    if (e != hipSuccess || graph == nullptr) {
        (void)hipGetLastError();
        release_graph(cid);
        std::stringstream ss;
        ss << "graph capture failed (a call synchronised or was not capturable): " << hipGetErrorString(e);
        throw CompasException(ss.str());
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        release_graph(cid);
        COMPAS_HIP_CHECK(e);
    }
    *id = cid;
    return exec;
}

void Runtime::abort_capture() {
    activate();
    hipGraph_t graph = nullptr;
    (void)hipStreamEndCapture(stream_, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    int cid;
    {
        std::lock_guard<std::mutex> lock(mutex_);
        cid = capture_id_;
        capture_id_ = -1;
        timing_ = timing_before_capture_;
    }
    if (cid >= 0) release_graph(cid);
}

void Runtime::release_graph(int id) {
    std::lock_guard<std::mutex> lock(mutex_);
    auto h = graph_held_.find(id);
    if (h != graph_held_.end()) {
        // stream order: later work on this stream runs after the graph's last launch
        for (auto& pn : h->second) {
            in_use_ -= pn.second;
            cached_ += pn.second;
            free_blocks_.emplace(pn.second, pn.first);
        }
        graph_held_.erase(h);
    }
    // blocks of the capture still held by arrays return to the pool as usual
    for (auto it = graph_blocks_.begin(); it != graph_blocks_.end();) {
        if (it->second == id) it = graph_blocks_.erase(it);
        else ++it;
    }
}

CapturedGraph::~CapturedGraph() {
    (void)hipSetDevice(rt_->device());
    (void)hipGraphExecDestroy(exec_);
    rt_->release_graph(id_);
}

void CapturedGraph::launch() const {
    rt_->activate();
    COMPAS_HIP_CHECK(hipGraphLaunch(exec_, rt_->stream()));
}

std::shared_ptr<CapturedGraph> end_capture(const CompasContext& ctx) {
    int id = -1;
    hipGraphExec_t exec = ctx.runtime()->end_capture(&id);
    return std::make_shared<CapturedGraph>(ctx.runtime(), exec, id);
}

void Runtime::trim() {
    std::multimap<size_t, void*> blocks;
    {
        std::lock_guard<std::mutex> lock(mutex_);
        blocks.swap(free_blocks_);
        cached_ = 0;
    }
    if (blocks.empty()) return;
    activate();
    // pending kernels may still read recycled memory of other arrays; freeing
    // cached blocks is only safe once the stream has drained
    (void)hipStreamSynchronize(stream_);
    for (auto& kv : blocks) (void)hipFree(kv.second);
}

void Runtime::synchronize() {
    activate();
    COMPAS_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Runtime::set_timing(bool on) {
    std::lock_guard<std::mutex> lock(mutex_);
    timing_ = on;
}

hipEvent_t Runtime::take_event() {
    if (!event_pool_.empty()) {
        hipEvent_t e = event_pool_.back();
        event_pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    COMPAS_HIP_CHECK(hipEventCreate(&e));
    return e;
}

int Runtime::begin_kernel(const char* name) {
    std::lock_guard<std::mutex> lock(mutex_);
    Pending p;
    p.name = name;
    p.start = take_event();
    p.stop = take_event();
    COMPAS_HIP_CHECK(hipEventRecord(p.start, stream_));
    pending_.push_back(p);
    return int(pending_.size() - 1);
}

void Runtime::end_kernel(int token) {
    std::lock_guard<std::mutex> lock(mutex_);
    COMPAS_HIP_CHECK(hipEventRecord(pending_[size_t(token)].stop, stream_));
}

std::map<std::string, KernelStats> Runtime::collect_timings() {
    synchronize();
    std::lock_guard<std::mutex> lock(mutex_);
    for (auto& p : pending_) {
        float ms = 0;
        COMPAS_HIP_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
        auto& s = stats_[p.name];
        s.total_ms += ms;
        s.count += 1;
        event_pool_.push_back(p.start);
        event_pool_.push_back(p.stop);
    }
    pending_.clear();
    auto out = stats_;
    stats_.clear();
    return out;
}

static std::mutex g_registry_mutex;
static std::map<int, std::shared_ptr<Runtime>>& registry() {
    static auto* r = new std::map<int, std::shared_ptr<Runtime>>();  // never destroyed
    return *r;
}

std::shared_ptr<Runtime> runtime_for_device(int device) {
    std::lock_guard<std::mutex> lock(g_registry_mutex);
    auto& reg = registry();
    auto it = reg.find(device);
    if (it != reg.end()) return it->second;
    auto rt = std::make_shared<Runtime>(device);
    reg[device] = rt;
    return rt;
}

Buffer::Buffer(std::shared_ptr<Runtime> rt, size_t nbytes) : rt_(std::move(rt)), nbytes_(nbytes) {
    ptr_ = rt_->allocate(nbytes);
}

Buffer::Buffer(
    std::shared_ptr<Runtime> rt, void* external, size_t nbytes, std::function<void()> release, bool sync_first) :
    rt_(std::move(rt)),
    ptr_(external),
    nbytes_(nbytes),
    external_(true),
    sync_first_(sync_first),
    release_(std::move(release)) {}

Buffer::~Buffer() {
    if (!external_) {
        rt_->release(ptr_, nbytes_);
        return;
    }
    if (release_) {
        try {
            // queued kernels may still read the memory: let them finish before
            // its owner may reuse it (arrays a captured graph reads must outlive
            // the graph, wrapped or not)
            if (sync_first_ && !rt_->capturing()) rt_->synchronize();
            release_();
        } catch (...) {  // a destructor must not throw; the owner's hook failed
        }
    }
}

void CompasContext::check_wrappable(const void* device_ptr, size_t nbytes) const {
    if (nbytes == 0) return;
    if (device_ptr == nullptr) COMPAS_ERROR("null device pointer");
    if (reinterpret_cast<uintptr_t>(device_ptr) % 256 != 0)
        COMPAS_ERROR("wrapped device memory must be aligned to 256 bytes");
    activate();
    hipPointerAttribute_t attr {};
    if (hipPointerGetAttributes(&attr, device_ptr) != hipSuccess) {
        (void)hipGetLastError();
        COMPAS_ERROR("pointer is not device memory known to HIP");
    }
    if (attr.type != hipMemoryTypeDevice) COMPAS_ERROR("pointer is not device memory");
    if (attr.device != device()) COMPAS_ERROR("device memory belongs to another GPU than the context's");
}

void CompasContext::wait_for_stream(hipStream_t other) const {
    activate();
    hipEvent_t ev;
    COMPAS_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    COMPAS_HIP_CHECK(hipEventRecord(ev, other));
    COMPAS_HIP_CHECK(hipStreamWaitEvent(stream(), ev, 0));
    COMPAS_HIP_CHECK(hipEventDestroy(ev));
}

void CompasContext::make_stream_wait(hipStream_t other) const {
    activate();
    hipEvent_t ev;
    COMPAS_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    COMPAS_HIP_CHECK(hipEventRecord(ev, stream()));
    COMPAS_HIP_CHECK(hipStreamWaitEvent(other, ev, 0));
    COMPAS_HIP_CHECK(hipEventDestroy(ev));
}

void copy_device_to_host(const std::shared_ptr<Buffer>& buf, void* host, size_t nbytes) {
    COMPAS_CHECK(nbytes <= buf->nbytes());
    auto& rt = buf->runtime();
    rt->activate();
    COMPAS_HIP_CHECK(hipMemcpyAsync(host, buf->data(), nbytes, hipMemcpyDeviceToHost, rt->stream()));
    COMPAS_HIP_CHECK(hipStreamSynchronize(rt->stream()));
}

void CompasContext::upload(const std::shared_ptr<Buffer>& buf, const void* host, size_t nbytes) const {
    COMPAS_CHECK(host != nullptr);
    activate();
    COMPAS_HIP_CHECK(hipMemcpyAsync(buf->data(), host, nbytes, hipMemcpyHostToDevice, stream()));
    // the caller may free or reuse the host memory as soon as this returns
    COMPAS_HIP_CHECK(hipStreamSynchronize(stream()));
}

CompasContext CompasContext::with_device(int index) const {
    int count = 0;
    COMPAS_HIP_CHECK(hipGetDeviceCount(&count));
    COMPAS_CHECK(count > 0);
    int dev = ((index % count) + count) % count;
    return CompasContext(runtime_for_device(dev));
}

CompasContext make_context(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        throw CompasException("no HIP device available (compas requires an AMD GPU)");
    }
    return CompasContext(runtime_for_device(((device % count) + count) % count));
}

}  // namespace compas
