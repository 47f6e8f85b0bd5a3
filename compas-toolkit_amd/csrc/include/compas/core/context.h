#pragma once
// Native single-GPU runtime behind the compas:: API.
//
// The reference delegates scheduling, memory and streams to the kmm runtime
// (include/compas/core/context.h:16-125).  Here one Runtime per device owns:
//   * one in-order HIP stream (default priority) on which every kernel of the
//     library is enqueued, so results are returned asynchronously and read
//     back with a blocking copy (as in the reference);
//   * a caching HBM pool: freed blocks are recycled by later allocations on
//     the same stream, which keeps allocation off the critical path and is
//     safe because stream order serialises the reuse;
//   * an optional kernel timer (HIP events around each launch) used by the
//     benchmark to measure per-kernel durations on the library's own stream;
//   * graph capture of the library calls enqueued on the stream
//     (hipStreamBeginCapture, relaxed mode): blocks allocated while a graph
//     is captured belong to it while it lives -- released arrays of the
//     capture are held back from the pool until the graph is destroyed, so a
//     replay never writes memory that other work has been given.

#include <array>
#include <cstdint>
#include <functional>
#include <map>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "compas/core/complex.h"

namespace compas {

using index_t = int;

struct CompasException: std::runtime_error {
    explicit CompasException(const std::string& msg) : std::runtime_error(msg) {}
};

[[noreturn]] void raise_error(const char* message, const char* file, int line, const char* func);

#define COMPAS_ERROR(message) ::compas::raise_error((message), __FILE__, __LINE__, __func__)
#define COMPAS_CHECK(...)                                    \
    do {                                                     \
        if (!(__VA_ARGS__)) {                                \
            COMPAS_ERROR("assertion failed: " #__VA_ARGS__); \
        }                                                    \
    } while (0)
#define COMPAS_HIP_CHECK(expr)                                                               \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess) {                                                              \
            (void)hipGetLastError(); /* a handled failure must not resurface at the next   \
                                        launch check of this thread */                       \
            ::compas::raise_error(hipGetErrorString(_e), __FILE__, __LINE__, #expr);         \
        }                                                                                    \
    } while (0)

struct KernelStats {
    double total_ms = 0;
    int64_t count = 0;
};

class Runtime {
  public:
    explicit Runtime(int device);
    ~Runtime();

    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }

    void* allocate(size_t nbytes);
    void release(void* ptr, size_t nbytes);
    void trim();  // return cached blocks to the driver
    size_t bytes_in_use() const { return in_use_; }
    size_t bytes_cached() const { return cached_; }

    void synchronize();
    void activate() const;  // make this runtime's device current on the calling thread

    // Kernel timing.  When enabled, `begin_kernel`/`end_kernel` record HIP
    // events around each launch on the runtime's stream; `collect` folds the
    // elapsed times into per-kernel statistics (after a synchronize).
    void set_timing(bool on);
    bool timing() const { return timing_; }
    int begin_kernel(const char* name);
    void end_kernel(int token);
    std::map<std::string, KernelStats> collect_timings();

    // Graph capture on this runtime's stream (one capture at a time).  The
    // calls between begin and end must not synchronise the host (reads of
    // arrays, uploads, first-time autotuning); run the sequence once eagerly
    // before capturing it.  Arrays the captured calls return are written by
    // every launch of the graph.
    void begin_capture();
    hipGraphExec_t end_capture(int* id);  // instantiated graph; `id` names its memory
    void abort_capture();
    void release_graph(int id);  // the graph is gone: its held blocks return to the pool
    bool capturing() const { return capture_id_ >= 0; }
    // whether the contraction autotuner may time trials on this runtime
    bool autotune() const { return autotune_; }
    void set_autotune(bool on) { autotune_ = on; }

  private:
    int device_;
    hipStream_t stream_ = nullptr;
    std::mutex mutex_;
    std::multimap<size_t, void*> free_blocks_;
    size_t in_use_ = 0, cached_ = 0;
    bool timing_ = false;
    bool autotune_ = true;
    bool timing_before_capture_ = false;
    int capture_id_ = -1, next_graph_id_ = 0;
    std::unordered_map<void*, int> graph_blocks_;  // block allocated during a capture -> its graph
    std::map<int, std::vector<std::pair<void*, size_t>>> graph_held_;  // released blocks kept for a graph
    struct Pending {
        std::string name;
        hipEvent_t start, stop;
    };
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> event_pool_;
    std::map<std::string, KernelStats> stats_;
    hipEvent_t take_event();
};

std::shared_ptr<Runtime> runtime_for_device(int device);

// Device buffer with shared ownership; released into the runtime's pool, or,
// for memory the library does not own (wrapped from another framework), handed
// back through its release hook when the last reference goes.
class Buffer {
  public:
    Buffer(std::shared_ptr<Runtime> rt, size_t nbytes);
    // sync_first: wait for the runtime's queued work before the hook (memory a
    // caller owns); false when the hook orders the release itself
    Buffer(std::shared_ptr<Runtime> rt, void* external, size_t nbytes, std::function<void()> release,
           bool sync_first = true);
    ~Buffer();
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;

    void* data() const { return ptr_; }
    size_t nbytes() const { return nbytes_; }
    bool external() const { return external_; }
    const std::shared_ptr<Runtime>& runtime() const { return rt_; }

  private:
    std::shared_ptr<Runtime> rt_;
    void* ptr_ = nullptr;
    size_t nbytes_ = 0;
    bool external_ = false;
    bool sync_first_ = true;
    std::function<void()> release_;
};

// Asynchronous device array (row-major, last index fastest).  Copies share the
// buffer, like the reference's Array (context.h:16-17).
template<typename T, size_t N = 1>
class Array {
  public:
    Array() = default;
    Array(std::shared_ptr<Buffer> buf, std::array<int64_t, N> sizes) : buf_(std::move(buf)), sizes_(sizes) {}

    T* data() const { return buf_ ? static_cast<T*>(buf_->data()) : nullptr; }
    int64_t size(size_t axis) const { return sizes_[axis]; }
    const std::array<int64_t, N>& sizes() const { return sizes_; }
    int64_t numel() const {
        int64_t n = 1;
        for (auto s : sizes_) n *= s;
        return n;
    }
    size_t nbytes() const { return size_t(numel()) * sizeof(T); }
    bool is_empty() const { return numel() == 0; }
    const std::shared_ptr<Buffer>& buffer() const { return buf_; }

    void copy_to(T* host) const;
    std::vector<T> copy_to_vector() const {
        std::vector<T> out(static_cast<size_t>(numel()));
        copy_to(out.data());
        return out;
    }
    void synchronize() const {
        if (buf_) buf_->runtime()->synchronize();
    }

  private:
    std::shared_ptr<Buffer> buf_;
    std::array<int64_t, N> sizes_ {};
};

void copy_device_to_host(const std::shared_ptr<Buffer>& buf, void* host, size_t nbytes);

template<typename T, size_t N>
void Array<T, N>::copy_to(T* host) const {
    if (numel() == 0) return;
    copy_device_to_host(buf_, host, nbytes());
}

// Host-side view used by make_* functions (pointer + sizes).
template<typename T, size_t N = 1>
struct View {
    const T* ptr = nullptr;
    std::array<int64_t, N> sizes {};
    View() = default;
    View(const T* p, std::array<int64_t, N> s) : ptr(p), sizes(s) {}
    const T* data() const { return ptr; }
    int64_t size(size_t i = 0) const { return sizes[i]; }
    int64_t numel() const {
        int64_t n = 1;
        for (auto s : sizes) n *= s;
        return n;
    }
    bool is_empty() const { return ptr == nullptr || numel() == 0; }
};
template<typename T>
using HostView = View<T, 1>;

class CompasContext {
  public:
    explicit CompasContext(std::shared_ptr<Runtime> rt) : rt_(std::move(rt)) {}

    CompasContext with_device(int index) const;

    template<typename T, size_t N>
    Array<T, N> allocate(View<T, N> content) const {
        return allocate_from<T, N>(content.data(), content.sizes);
    }
    template<typename T>
    Array<T, 1> allocate(const std::vector<T>& content) const {
        return allocate_from<T, 1>(content.data(), {int64_t(content.size())});
    }
    template<typename T, size_t N>
    Array<T, N> allocate_from(const T* host, std::array<int64_t, N> sizes) const {
        Array<T, N> a = empty<T, N>(sizes);
        if (a.numel() > 0) upload(a.buffer(), host, a.nbytes());
        return a;
    }
    template<typename T, size_t N>
    Array<T, N> empty(std::array<int64_t, N> sizes) const {
        int64_t n = 1;
        for (auto s : sizes) {
            COMPAS_CHECK(s >= 0);
            n *= s;
        }
        auto buf = std::make_shared<Buffer>(rt_, size_t(n) * sizeof(T));
        return Array<T, N>(std::move(buf), sizes);
    }
    template<typename T, size_t N>
    Array<T, N> zeros(std::array<int64_t, N> sizes) const {
        auto a = empty<T, N>(sizes);
        fill_zero(a.data(), a.nbytes());
        return a;
    }

    void synchronize() const { rt_->synchronize(); }
    hipStream_t stream() const { return rt_->stream(); }
    int device() const { return rt_->device(); }
    const std::shared_ptr<Runtime>& runtime() const { return rt_; }
    void activate() const { rt_->activate(); }

    void upload(const std::shared_ptr<Buffer>& buf, const void* host, size_t nbytes) const;
    void fill_zero(void* ptr, size_t nbytes) const;

    // Array over device memory owned elsewhere (no copy): `device_ptr` must be
    // device memory of this context's GPU, aligned to 256 bytes; `release` runs
    // once the last array over it is gone.
    template<typename T, size_t N>
    Array<T, N> wrap(void* device_ptr, std::array<int64_t, N> sizes, std::function<void()> release) const {
        int64_t n = 1;
        for (auto s : sizes) {
            COMPAS_CHECK(s >= 0);
            n *= s;
        }
        check_wrappable(device_ptr, size_t(n) * sizeof(T));
        auto buf = std::make_shared<Buffer>(rt_, device_ptr, size_t(n) * sizeof(T), std::move(release));
        return Array<T, N>(std::move(buf), sizes);
    }
    void check_wrappable(const void* device_ptr, size_t nbytes) const;

    // Stream ordering with work outside the library: this context's stream
    // waits for what is queued on `other` so far, or `other` waits for what
    // is queued on this context's stream so far (an event, no host wait).
    void wait_for_stream(hipStream_t other) const;
    void make_stream_wait(hipStream_t other) const;

  private:
    std::shared_ptr<Runtime> rt_;
};

CompasContext make_context(int device = 0);

// An instantiated graph of library calls captured on one runtime's stream;
// launch() replays it on that stream.
class CapturedGraph {
  public:
    CapturedGraph(std::shared_ptr<Runtime> rt, hipGraphExec_t exec, int id) : rt_(std::move(rt)), exec_(exec), id_(id) {}
    ~CapturedGraph();
    CapturedGraph(const CapturedGraph&) = delete;
    CapturedGraph& operator=(const CapturedGraph&) = delete;
    void launch() const;

  private:
    std::shared_ptr<Runtime> rt_;
    hipGraphExec_t exec_;
    int id_;
};

std::shared_ptr<CapturedGraph> end_capture(const CompasContext& ctx);

// RAII helper that times one launch when the runtime's timer is enabled.
struct KernelScope {
    Runtime* rt;
    int token;
    KernelScope(const CompasContext& ctx, const char* name) : rt(ctx.runtime().get()), token(-1) {
        if (rt->timing()) token = rt->begin_kernel(name);
    }
    ~KernelScope() {
        if (token >= 0) rt->end_kernel(token);
    }
};

void check_launch(const char* name);

}  // namespace compas
