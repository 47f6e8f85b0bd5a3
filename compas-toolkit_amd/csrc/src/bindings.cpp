// C-ABI of libcompas_amd.so (declared in include/compas_amd.h).
//
// Mirrors the reference's julia-bindings/src/bindings.cpp:19-411 entry point
// by entry point.  Handle objects are type-checked on every call (the
// reference casts blindly), and errors go through jl_errorf when the host is
// Julia, or through a per-thread message otherwise.

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <typeinfo>

#include "compas/compas.h"
#include "compas/internal.h"
#include "compas/serving.h"
#include "compas/wire.h"
#include "compas_amd.h"

// Provided by libjulia when loaded from CompasToolkit.jl; absent otherwise.
extern "C" void jl_errorf(const char* fmt, ...) __attribute__((weak, noreturn));

namespace {

using namespace compas;

thread_local std::string g_last_error;

struct ContextHandle {
    CompasContext ctx;
};

struct Handle {
    virtual ~Handle() = default;
    virtual const char* name() const = 0;
};

template<typename T>
struct ObjectHandle: Handle, T {
    template<typename... Args>
    explicit ObjectHandle(Args&&... args) : T(std::forward<Args>(args)...) {}
    const char* name() const override { return typeid(T).name(); }
};

enum DType { DT_FLOAT = 0, DT_COMPLEX = 1, DT_BF16 = 2 };

template<typename T>
constexpr int dtype_of() {
    return std::is_same<T, cfloat>::value ? DT_COMPLEX : std::is_same<T, bfloat16_t>::value ? DT_BF16 : DT_FLOAT;
}

const char* dtype_name(int dtype) {
    return dtype == DT_COMPLEX ? "complex" : dtype == DT_BF16 ? "bfloat16" : "float";
}

size_t dtype_size(int dtype) {
    return dtype == DT_COMPLEX ? sizeof(cfloat) : dtype == DT_BF16 ? sizeof(uint16_t) : sizeof(float);
}

struct ArrayHandle: Handle {
    std::shared_ptr<Buffer> buf;
    int dtype = DT_FLOAT;
    int rank = 0;
    int64_t sizes[3] = {0, 0, 0};

    const char* name() const override {
        return dtype == DT_FLOAT ? "Array<float>" : dtype == DT_BF16 ? "Array<bfloat16>" : "Array<cfloat>";
    }

    int64_t numel() const {
        int64_t n = 1;
        for (int i = 0; i < rank; i++) n *= sizes[i];
        return n;
    }

    template<typename T, size_t N>
    Array<T, N> as(const char* what) const {
        const int want = dtype_of<T>();
        if (dtype != want || rank != int(N)) {
            std::stringstream ss;
            ss << "invalid array for `" << what << "`: expecting " << dtype_name(want)
               << " rank " << N << ", gotten " << name() << " rank " << rank;
            throw CompasException(ss.str());
        }
        std::array<int64_t, N> s {};
        for (size_t i = 0; i < N; i++) s[i] = sizes[i];
        return Array<T, N>(buf, s);
    }

    template<typename T, size_t N>
    static ArrayHandle* wrap(const Array<T, N>& a) {
        auto* h = new ArrayHandle();
        h->buf = a.buffer();
        h->dtype = dtype_of<T>();
        h->rank = int(N);
        for (size_t i = 0; i < N; i++) h->sizes[i] = a.size(i);
        return h;
    }
};

template<typename T>
const T& unwrap(const void* p, const char* what) {
    if (p == nullptr) throw CompasException(std::string("null handle passed as `") + what + "`");
    const auto* h = static_cast<const Handle*>(p);
    if (const auto* t = dynamic_cast<const T*>(h)) return *t;
    std::stringstream ss;
    ss << "invalid object for `" << what << "` (expecting: " << typeid(T).name() << ", gotten: " << h->name() << ")";
    throw CompasException(ss.str());
}

const ArrayHandle& array(const void* p, const char* what) {
    return unwrap<ArrayHandle>(p, what);
}

// Slice positions of a pSSFP sequence.  CompasToolkit.jl:614 converts z to
// ComplexF32 before the call, and the reference binding reads that complex
// array as Array<float> (julia-bindings/src/bindings.cpp:186,194), i.e. it
// simulates the first nz floats [z0, 0, z1, 0, ...] of the complex buffer.
// Here a complex z stands for its real parts (its imaginary parts must be
// zero, as Julia's conversion of real positions leaves them); a float z (C++
// and other callers) is taken as is.
Array<float, 1> slice_positions(const ArrayHandle& z) {
    if (z.dtype == DT_FLOAT) return z.as<float, 1>("z");
    auto zc = z.as<cfloat, 1>("z");
    std::vector<cfloat> host = zc.copy_to_vector();
    std::vector<float> re(host.size());
    for (size_t i = 0; i < host.size(); i++) {
        if (host[i].imag() != 0.0f)
            throw CompasException("pSSFP slice positions `z` must be real (a complex z needs zero imaginary parts)");
        re[i] = host[i].real();
    }
    return CompasContext(zc.buffer()->runtime()).allocate(re);
}

const CompasContext& context(const void* p) {
    if (p == nullptr) throw CompasException("null context");
    return static_cast<const ContextHandle*>(p)->ctx;
}

void report(const char* msg) {
    if (jl_errorf) jl_errorf("COMPAS error occurred: %s\n", msg);
    g_last_error = msg;
}

template<typename R>
R default_value() {
    if constexpr (std::is_void<R>::value) {
        return;
    } else {
        return R {};
    }
}

template<typename F>
auto catch_exceptions(F fun) -> decltype(fun()) {
    using R = decltype(fun());
    try {
        g_last_error.clear();
        return fun();
    } catch (const std::exception& e) {
        report(e.what());
    } catch (...) {
        report("unknown exception");
    }
    return default_value<R>();
}

template<typename T>
CompasObjectHandle* new_object(T&& value) {
    auto* h = new ObjectHandle<std::decay_t<T>>(std::forward<T>(value));
    return reinterpret_cast<CompasObjectHandle*>(static_cast<Handle*>(h));
}

template<typename T, size_t N>
CompasArrayHandle* new_array(const Array<T, N>& a) {
    return reinterpret_cast<CompasArrayHandle*>(static_cast<Handle*>(ArrayHandle::wrap(a)));
}

const CompasContextHandle* new_context(const CompasContext& ctx) {
    return reinterpret_cast<const CompasContextHandle*>(new ContextHandle {ctx});
}

View<float> host_view(const float* p, int n) {
    if (p == nullptr) return {};
    return View<float>(p, {n});
}

template<typename T>
const ArrayHandle* make_array_impl(const CompasContext& ctx, const T* data, int rank, const int64_t* sizes) {
    if (rank < 1 || rank > 3) COMPAS_ERROR("cannot support rank > 3");  // bindings.cpp:63-64
    COMPAS_CHECK(sizes != nullptr);
    for (int i = 0; i < rank; i++) COMPAS_CHECK(sizes[i] >= 0);
    if (rank == 1) return ArrayHandle::wrap(ctx.allocate_from<T, 1>(data, {sizes[0]}));
    if (rank == 2) return ArrayHandle::wrap(ctx.allocate_from<T, 2>(data, {sizes[0], sizes[1]}));
    return ArrayHandle::wrap(ctx.allocate_from<T, 3>(data, {sizes[0], sizes[1], sizes[2]}));
}

void read_array_impl(const ArrayHandle& a, void* dest, int64_t length, size_t elem, int dtype) {
    if (a.dtype != dtype) throw CompasException("array element type does not match the read function");
    COMPAS_CHECK(length >= 0 && length <= a.numel());
    if (length == 0) return;
    COMPAS_CHECK(dest != nullptr);
    copy_device_to_host(a.buf, dest, size_t(length) * elem);
}

SimulateSignalMethod signal_method(int m) {
    if (m < 0 || m > 5) COMPAS_ERROR("invalid value for `SimulateSignalMethod`");
    return SimulateSignalMethod(m);
}

GemmComputeMethod gemm_method(int m) {
    if (m < 0 || m > 2) COMPAS_ERROR("invalid value for `GemmComputeMethod`");
    return GemmComputeMethod(m);
}

JacobianComputeMethod jacobian_method(int m) {
    if (m < 0 || m > 4) COMPAS_ERROR("invalid value for `JacobianComputeMethod`");
    return JacobianComputeMethod(m);
}

// "name count total_ms" lines into buf (NUL-terminated, truncated); returns bytes needed
int format_timings(const std::map<std::string, KernelStats>& stats, char* buf, int buflen) {
    std::stringstream ss;
    for (auto& kv : stats) {
        char line[256];
        std::snprintf(line, sizeof(line), "%s %lld %.6f\n", kv.first.c_str(), (long long)kv.second.count,
                      kv.second.total_ms);
        ss << line;
    }
    const std::string s = ss.str();
    if (buf && buflen > 0) {
        const size_t n = std::min(size_t(buflen - 1), s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return int(s.size() + 1);
}

}  // namespace

namespace {

template<bool ADJOINT>
CompasArrayHandle* spiral_product(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* dT1,
    const CompasArrayHandle* dT2, const CompasArrayHandle* dB1, const CompasArrayHandle* dB0,
    const CompasObjectHandle* parameters, const CompasObjectHandle* trajectory, const CompasArrayHandle* coils,
    const CompasArrayHandle* vector, JacobianComputeMethod method) {
    return catch_exceptions([&] {
        auto e = array(echos, "echos").as<cfloat, 2>("echos");
        auto d1 = array(dT1, "delta_echos_T1").as<cfloat, 2>("delta_echos_T1");
        auto d2 = array(dT2, "delta_echos_T2").as<cfloat, 2>("delta_echos_T2");
        auto db1 = array(dB1, "delta_echos_B1").as<cfloat, 2>("delta_echos_B1");
        auto db0 = array(dB0, "delta_echos_B0").as<cfloat, 2>("delta_echos_B0");
        const auto& p = unwrap<TissueParameters>(parameters, "parameters");
        const auto& t = unwrap<SpiralTrajectory>(trajectory, "trajectory");
        auto c = array(coils, "coils").as<cfloat, 2>("coils");
        if constexpr (ADJOINT) {
            return new_array(compute_jacobian_hermitian(context(ctx), e, d1, d2, db1, db0, p, t, c,
                                                        array(vector, "vector").as<cfloat, 3>("vector"), method));
        } else {
            return new_array(compute_jacobian(context(ctx), e, d1, d2, db1, db0, p, t, c,
                                              array(vector, "vector").as<cfloat, 2>("vector"), method));
        }
    });
}

}  // namespace

namespace {

template<typename Sequence>
int tangents_impl(
    const CompasContextHandle* ctx, const int32_t* fields, int32_t nfields, int32_t with_echos,
    const CompasObjectHandle* parameters, const CompasObjectHandle* sequence, CompasArrayHandle** out) {
    if (nfields < 0 || (nfields > 0 && fields == nullptr)) throw CompasException("invalid field list");
    if (out == nullptr) throw CompasException("null output list");
    std::vector<int> f(fields, fields + nfields);
    Array<cfloat, 2> echos;
    auto d = simulate_magnetization_tangents(context(ctx), f, unwrap<TissueParameters>(parameters, "parameters"),
                                             unwrap<Sequence>(sequence, "sequence"), with_echos ? &echos : nullptr);
    int n = 0;
    if (with_echos) out[n++] = new_array(echos);
    for (auto& a : d) out[n++] = new_array(a);
    return n;
}

}  // namespace

namespace {

// DLPack (legacy, unversioned) ABI: the exchange structs of dlpack.h v0.8,
// restated here so the C header carries no dlpack.h dependency.
struct DLDevice {
    int32_t device_type;
    int32_t device_id;
};
struct DLDataType {
    uint8_t code;
    uint8_t bits;
    uint16_t lanes;
};
struct DLTensor {
    void* data;
    DLDevice device;
    int32_t ndim;
    DLDataType dtype;
    int64_t* shape;
    int64_t* strides;
    uint64_t byte_offset;
};
struct DLManagedTensor {
    DLTensor dl_tensor;
    void* manager_ctx;
    void (*deleter)(DLManagedTensor*);
};
constexpr int32_t DL_ROCM = 10;
constexpr uint8_t DL_FLOAT = 2, DL_BFLOAT = 4, DL_COMPLEX = 5;

// An exported array: the DLPack view and a reference to the buffer
struct ExportedArray {
    DLManagedTensor managed;
    int64_t shape[3];
    std::shared_ptr<Buffer> buf;
};

void delete_exported(DLManagedTensor* m) {
    delete static_cast<ExportedArray*>(m->manager_ctx);
}

ArrayHandle* wrap_external(
    const CompasContext& ctx, void* ptr, int dtype, int rank, const int64_t* sizes, std::function<void()> release) {
    if (dtype != DT_FLOAT && dtype != DT_COMPLEX) throw CompasException("dtype must be 0 (float) or 1 (complex)");
    if (rank < 1 || rank > 3) COMPAS_ERROR("cannot support rank > 3");
    COMPAS_CHECK(sizes != nullptr);
    int64_t n = 1;
    for (int i = 0; i < rank; i++) {
        COMPAS_CHECK(sizes[i] >= 0);
        n *= sizes[i];
    }
    const size_t nbytes = size_t(n) * (dtype == DT_COMPLEX ? sizeof(cfloat) : sizeof(float));
    ctx.check_wrappable(ptr, nbytes);
    auto h = std::make_unique<ArrayHandle>();
    h->dtype = dtype;
    h->rank = rank;
    for (int i = 0; i < rank; i++) h->sizes[i] = sizes[i];
    h->buf = std::make_shared<Buffer>(ctx.runtime(), ptr, nbytes, std::move(release));  // last: owns from here
    return h.release();
}

}  // namespace

extern "C" {

const char* compas_version() {
    return "compas-amd 0.1.0 (gfx950)";
}

const CompasContextHandle* compas_make_context() {
    return catch_exceptions([&] { return new_context(make_context(0)); });
}

const CompasContextHandle* compas_copy_context_for_device(CompasContextHandle* ctx, int device) {
    return catch_exceptions([&] { return new_context(context(ctx).with_device(device)); });
}

void compas_destroy_context(const CompasContextHandle* ctx) {
    catch_exceptions([&] { delete reinterpret_cast<const ContextHandle*>(ctx); });
}

void compas_destroy_object(const CompasObjectHandle* obj) {
    catch_exceptions([&] { delete reinterpret_cast<const Handle*>(obj); });
}

void compas_synchronize(const CompasContextHandle* ctx) {
    catch_exceptions([&] { context(ctx).synchronize(); });
}

const CompasArrayHandle* compas_make_array_float(
    const CompasContextHandle* ctx, const float* data, int rank, int64_t* sizes) {
    return catch_exceptions([&] {
        return reinterpret_cast<const CompasArrayHandle*>(
            static_cast<const Handle*>(make_array_impl<float>(context(ctx), data, rank, sizes)));
    });
}

void compas_read_array_float(const CompasContextHandle* ctx, const CompasArrayHandle* a, float* dest, int64_t length) {
    catch_exceptions([&] {
        (void)context(ctx);
        read_array_impl(array(a, "array"), dest, length, sizeof(float), DT_FLOAT);
    });
}

const CompasArrayHandle* compas_make_array_complex(
    const CompasContextHandle* ctx, const compas_cfloat* data, int rank, int64_t* sizes) {
    return catch_exceptions([&] {
        return reinterpret_cast<const CompasArrayHandle*>(static_cast<const Handle*>(
            make_array_impl<cfloat>(context(ctx), reinterpret_cast<const cfloat*>(data), rank, sizes)));
    });
}

void compas_read_array_complex(
    const CompasContextHandle* ctx, const CompasArrayHandle* a, compas_cfloat* dest, int64_t length) {
    catch_exceptions([&] {
        (void)context(ctx);
        read_array_impl(array(a, "array"), dest, length, sizeof(cfloat), DT_COMPLEX);
    });
}

const CompasArrayHandle* compas_make_array_bf16(
    const CompasContextHandle* ctx, const uint16_t* data, int rank, int64_t* sizes) {
    return catch_exceptions([&] {
        return reinterpret_cast<const CompasArrayHandle*>(static_cast<const Handle*>(
            make_array_impl<bfloat16_t>(context(ctx), reinterpret_cast<const bfloat16_t*>(data), rank, sizes)));
    });
}

void compas_read_array_bf16(const CompasContextHandle* ctx, const CompasArrayHandle* a, uint16_t* dest, int64_t length) {
    catch_exceptions([&] {
        (void)context(ctx);
        read_array_impl(array(a, "array"), dest, length, sizeof(uint16_t), DT_BF16);
    });
}

void compas_write_array_complex(
    const CompasContextHandle* ctx, const CompasArrayHandle* a, const compas_cfloat* src, int64_t length) {
    catch_exceptions([&] {
        const CompasContext& c = context(ctx);
        const ArrayHandle& h = array(a, "array");
        if (h.dtype != DT_COMPLEX) throw CompasException("array element type does not match the write function");
        COMPAS_CHECK(length >= 0 && length <= h.numel());
        if (length == 0) return;
        COMPAS_CHECK(src != nullptr);
        c.upload(h.buf, src, size_t(length) * sizeof(cfloat));
    });
}

void compas_copy_array(const CompasContextHandle* ctx, const CompasArrayHandle* dst, const CompasArrayHandle* src) {
    catch_exceptions([&] {
        const CompasContext& c = context(ctx);
        const ArrayHandle& d = array(dst, "dst");
        const ArrayHandle& s = array(src, "src");
        if (d.dtype != s.dtype || d.numel() != s.numel()) throw CompasException("arrays differ in type or size");
        if (d.numel() == 0) return;
        c.activate();
        const size_t elem = dtype_size(d.dtype);
        COMPAS_HIP_CHECK(hipMemcpyAsync(d.buf->data(), s.buf->data(), size_t(d.numel()) * elem,
                                        hipMemcpyDeviceToDevice, c.stream()));
    });
}

void* compas_context_stream(const CompasContextHandle* ctx) {
    return catch_exceptions([&]() -> void* { return static_cast<void*>(context(ctx).stream()); });
}

void compas_context_wait_stream(const CompasContextHandle* ctx, void* stream) {
    catch_exceptions([&] { context(ctx).wait_for_stream(static_cast<hipStream_t>(stream)); });
}

void compas_stream_wait_context(const CompasContextHandle* ctx, void* stream) {
    catch_exceptions([&] { context(ctx).make_stream_wait(static_cast<hipStream_t>(stream)); });
}

const CompasArrayHandle* compas_wrap_device_array(
    const CompasContextHandle* ctx, void* device_ptr, int32_t dtype, int32_t rank, const int64_t* sizes,
    compas_release_fn release, void* user) {
    return catch_exceptions([&]() -> const CompasArrayHandle* {
        std::function<void()> hook;
        if (release) hook = [release, user] { release(user); };
        return reinterpret_cast<const CompasArrayHandle*>(
            static_cast<const Handle*>(wrap_external(context(ctx), device_ptr, dtype, rank, sizes, std::move(hook))));
    });
}

int32_t compas_array_device(const CompasArrayHandle* a) {
    return catch_exceptions([&]() -> int32_t { return array(a, "array").buf->runtime()->device(); });
}

void compas_dlpack_delete(void* dl_managed_tensor) {
    catch_exceptions([&] {
        auto* m = static_cast<DLManagedTensor*>(dl_managed_tensor);
        if (m && m->deleter) m->deleter(m);
    });
}

void* compas_array_device_pointer(const CompasArrayHandle* a) {
    return catch_exceptions([&]() -> void* { return array(a, "array").buf->data(); });
}

const CompasArrayHandle* compas_array_from_dlpack(const CompasContextHandle* ctx, void* managed_tensor) {
    return catch_exceptions([&]() -> const CompasArrayHandle* {
        if (managed_tensor == nullptr) throw CompasException("null DLPack tensor");
        auto* m = static_cast<DLManagedTensor*>(managed_tensor);
        const DLTensor& t = m->dl_tensor;
        const CompasContext& c = context(ctx);
        if (t.device.device_type != DL_ROCM) throw CompasException("DLPack tensor is not ROCm device memory");
        if (t.device.device_id != c.device()) throw CompasException("DLPack tensor is on another GPU than the context's");
        int dtype;
        if (t.dtype.code == DL_FLOAT && t.dtype.bits == 32 && t.dtype.lanes == 1) dtype = DT_FLOAT;
        else if (t.dtype.code == DL_COMPLEX && t.dtype.bits == 64 && t.dtype.lanes == 1) dtype = DT_COMPLEX;
        else throw CompasException("DLPack tensor must be float32 or complex64");
        if (t.ndim < 1 || t.ndim > 3) throw CompasException("DLPack tensor must have rank 1 to 3");
        if (t.strides) {  // row-major and compact (size-1 axes may carry any stride)
            int64_t expect = 1;
            for (int i = t.ndim - 1; i >= 0; i--) {
                if (t.shape[i] != 1 && t.strides[i] != expect) throw CompasException("DLPack tensor is not contiguous");
                expect *= t.shape[i];
            }
        }
        void* data = static_cast<char*>(t.data) + t.byte_offset;
        // the array owns the tensor once built (the hook runs only from its
        // buffer); when a check fails the caller still owns it
        auto* h = wrap_external(c, data, dtype, t.ndim, t.shape, [m] {
            if (m->deleter) m->deleter(m);
        });
        return reinterpret_cast<const CompasArrayHandle*>(static_cast<const Handle*>(h));
    });
}

void* compas_array_to_dlpack(const CompasArrayHandle* a) {
    return catch_exceptions([&]() -> void* {
        const ArrayHandle& h = array(a, "array");
        auto* e = new ExportedArray();
        e->buf = h.buf;
        for (int i = 0; i < 3; i++) e->shape[i] = h.sizes[i];
        DLTensor& t = e->managed.dl_tensor;
        t.data = h.buf->data();
        t.device = {DL_ROCM, h.buf->runtime()->device()};
        t.ndim = h.rank;
        t.dtype = h.dtype == DT_COMPLEX ? DLDataType {DL_COMPLEX, 64, 1}
                  : h.dtype == DT_BF16  ? DLDataType {DL_BFLOAT, 16, 1}
                                        : DLDataType {DL_FLOAT, 32, 1};
        t.shape = e->shape;
        t.strides = nullptr;  // row-major, compact
        t.byte_offset = 0;
        e->managed.manager_ctx = e;
        e->managed.deleter = &delete_exported;
        return &e->managed;
    });
}

void compas_graph_begin(const CompasContextHandle* ctx) {
    catch_exceptions([&] { context(ctx).runtime()->begin_capture(); });
}

CompasGraphHandle* compas_graph_end(const CompasContextHandle* ctx) {
    return catch_exceptions([&]() -> CompasGraphHandle* {
        auto g = end_capture(context(ctx));
        return reinterpret_cast<CompasGraphHandle*>(new std::shared_ptr<CapturedGraph>(std::move(g)));
    });
}

void compas_graph_abort(const CompasContextHandle* ctx) {
    catch_exceptions([&] { context(ctx).runtime()->abort_capture(); });
}

void compas_graph_launch(const CompasGraphHandle* graph) {
    catch_exceptions([&] {
        if (graph == nullptr) throw CompasException("null graph");
        (*reinterpret_cast<const std::shared_ptr<CapturedGraph>*>(graph))->launch();
    });
}

void compas_graph_destroy(const CompasGraphHandle* graph) {
    catch_exceptions([&] { delete reinterpret_cast<const std::shared_ptr<CapturedGraph>*>(graph); });
}

void compas_destroy_array(const CompasArrayHandle* a) {
    catch_exceptions([&] { delete reinterpret_cast<const Handle*>(a); });
}

const CompasObjectHandle* compas_make_tissue_parameters(
    const CompasContextHandle* ctx, int nvoxels, const float* T1, const float* T2, const float* B1,
    const float* B0, const float* rho_x, const float* rho_y, const float* x, const float* y, const float* z) {
    return catch_exceptions([&]() -> const CompasObjectHandle* {
        COMPAS_CHECK(nvoxels >= 0);
        // bindings.cpp:131 -- one chunk holding every voxel
        return new_object(make_tissue_parameters(
            context(ctx), nvoxels, nvoxels, host_view(T1, nvoxels), host_view(T2, nvoxels), host_view(B1, nvoxels),
            host_view(B0, nvoxels), host_view(rho_x, nvoxels), host_view(rho_y, nvoxels), host_view(x, nvoxels),
            host_view(y, nvoxels), host_view(z, nvoxels)));
    });
}

CompasObjectHandle* compas_make_fisp_sequence(
    const CompasArrayHandle* RF_train, const CompasArrayHandle* sliceprofiles, float TR, float TE, float TW,
    int max_state, float TI, int undersampling_factor, int repetitions) {
    return catch_exceptions([&] {
        auto rf = array(RF_train, "RF_train").as<cfloat, 1>("RF_train");
        auto prof = array(sliceprofiles, "sliceprofiles").as<cfloat, 2>("sliceprofiles");
        return new_object(FISPSequence(rf, prof, TR, TE, TW, max_state, TI, undersampling_factor, repetitions));
    });
}

CompasObjectHandle* compas_make_pssfp_sequence(
    const CompasArrayHandle* RF_train, float TR, const CompasArrayHandle* gamma_dt_RF, float dt_ex, float dt_inv,
    float dt_pr, float g_ex, float g_inv, float g_pr, const CompasArrayHandle* z) {
    return catch_exceptions([&] {
        auto rf = array(RF_train, "RF_train").as<cfloat, 1>("RF_train");
        auto g = array(gamma_dt_RF, "gamma_dt_RF").as<cfloat, 1>("gamma_dt_RF");
        auto zz = slice_positions(array(z, "z"));
        COMPAS_CHECK(rf.size(0) > 0);
        return new_object(pSSFPSequence(rf, TR, g, RepetitionData {dt_ex, dt_inv, dt_pr},
                                        RepetitionData {g_ex, g_inv, g_pr}, zz));
    });
}

CompasObjectHandle* compas_make_cartesian_trajectory(
    int nreadouts, int samples_per_readout, float delta_t, CompasArrayHandle* k_start, float delta_k) {
    return catch_exceptions([&] {
        auto ks = array(k_start, "k_start").as<cfloat, 1>("k_start");
        COMPAS_CHECK(ks.size(0) == nreadouts);
        return new_object(CartesianTrajectory(nreadouts, samples_per_readout, delta_t, ks, cfloat(delta_k, 0.0f)));
    });
}

CompasObjectHandle* compas_make_spiral_trajectory(
    int nreadouts, int samples_per_readout, float delta_t, CompasArrayHandle* k_start, CompasArrayHandle* delta_k) {
    return catch_exceptions([&] {
        auto ks = array(k_start, "k_start").as<cfloat, 1>("k_start");
        auto dk = array(delta_k, "delta_k").as<cfloat, 1>("delta_k");
        COMPAS_CHECK(ks.size(0) == nreadouts && dk.size(0) == nreadouts);
        return new_object(SpiralTrajectory(nreadouts, samples_per_readout, delta_t, ks, dk));
    });
}

CompasArrayHandle* compas_simulate_magnetization_fisp(
    const CompasContextHandle* ctx, const CompasObjectHandle* parameters, const CompasObjectHandle* sequence) {
    return catch_exceptions([&] {
        return new_array(simulate_magnetization(
            context(ctx), unwrap<TissueParameters>(parameters, "parameters"),
            unwrap<FISPSequence>(sequence, "sequence")));
    });
}

int compas_simulate_magnetization_fisp_noalloc(
    const CompasContextHandle* ctx, const CompasObjectHandle* parameters, const CompasObjectHandle* sequence,
    compas_cfloat* echos, int64_t capacity) {
    return catch_exceptions([&]() -> int {
        simulate_magnetization_into(
            context(ctx), unwrap<TissueParameters>(parameters, "parameters"),
            unwrap<FISPSequence>(sequence, "sequence"), reinterpret_cast<cfloat*>(echos), capacity);
        return 1;
    });
}

CompasArrayHandle* compas_simulate_magnetization_pssfp(
    const CompasContextHandle* ctx, const CompasObjectHandle* parameters, const CompasObjectHandle* sequence) {
    return catch_exceptions([&] {
        return new_array(simulate_magnetization(
            context(ctx), unwrap<TissueParameters>(parameters, "parameters"),
            unwrap<pSSFPSequence>(sequence, "sequence")));
    });
}

CompasArrayHandle* compas_magnetization_to_signal_cartesian(
    const CompasContextHandle* ctx, int ncoils, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    const CompasArrayHandle* coils, const CompasObjectHandle* trajectory, int32_t low_precision) {
    (void)ncoils;  // unused, as in the reference (the coil count comes from `coils`)
    return catch_exceptions([&] {
        auto method = low_precision == 0 ? SimulateSignalMethod::MatmulFast : SimulateSignalMethod::MatmulLow;
        return new_array(magnetization_to_signal(
            context(ctx), array(echos, "echos").as<cfloat, 2>("echos"),
            unwrap<TissueParameters>(parameters, "parameters"),
            unwrap<CartesianTrajectory>(trajectory, "trajectory"), array(coils, "coils").as<cfloat, 2>("coils"),
            method));
    });
}

CompasArrayHandle* compas_magnetization_to_signal_spiral(
    const CompasContextHandle* ctx, int ncoils, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    const CompasArrayHandle* coils, const CompasObjectHandle* trajectory) {
    (void)ncoils;
    return catch_exceptions([&] {
        return new_array(magnetization_to_signal(
            context(ctx), array(echos, "echos").as<cfloat, 2>("echos"),
            unwrap<TissueParameters>(parameters, "parameters"), unwrap<SpiralTrajectory>(trajectory, "trajectory"),
            array(coils, "coils").as<cfloat, 2>("coils")));
    });
}

CompasArrayHandle* compas_compute_jacobian(
    const CompasContextHandle* ctx, int ncoils, const CompasArrayHandle* echos, const CompasArrayHandle* dT1,
    const CompasArrayHandle* dT2, const CompasObjectHandle* parameters, const CompasArrayHandle* coils,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* vector, int32_t low_precision) {
    (void)ncoils;
    return catch_exceptions([&] {
        auto method = low_precision == 0 ? JacobianComputeMethod::GemmFast : JacobianComputeMethod::GemmLow;
        return new_array(compute_jacobian(
            context(ctx), array(echos, "echos").as<cfloat, 2>("echos"),
            array(dT1, "delta_echos_T1").as<cfloat, 2>("delta_echos_T1"),
            array(dT2, "delta_echos_T2").as<cfloat, 2>("delta_echos_T2"),
            unwrap<TissueParameters>(parameters, "parameters"),
            unwrap<CartesianTrajectory>(trajectory, "trajectory"), array(coils, "coils").as<cfloat, 2>("coils"),
            array(vector, "vector").as<cfloat, 2>("vector"), method));
    });
}

CompasArrayHandle* compas_compute_jacobian_hermitian(
    const CompasContextHandle* ctx, int ncoils, const CompasArrayHandle* echos, const CompasArrayHandle* dT1,
    const CompasArrayHandle* dT2, const CompasObjectHandle* parameters, const CompasObjectHandle* trajectory,
    const CompasArrayHandle* coils, const CompasArrayHandle* vector, int32_t low_precision) {
    (void)ncoils;
    return catch_exceptions([&] {
        auto method = low_precision == 0 ? JacobianComputeMethod::GemmFast : JacobianComputeMethod::GemmLow;
        return new_array(compute_jacobian_hermitian(
            context(ctx), array(echos, "echos").as<cfloat, 2>("echos"),
            array(dT1, "delta_echos_T1").as<cfloat, 2>("delta_echos_T1"),
            array(dT2, "delta_echos_T2").as<cfloat, 2>("delta_echos_T2"),
            unwrap<TissueParameters>(parameters, "parameters"),
            unwrap<CartesianTrajectory>(trajectory, "trajectory"), array(coils, "coils").as<cfloat, 2>("coils"),
            array(vector, "vector").as<cfloat, 3>("vector"), method));
    });
}

CompasArrayHandle* phase_encoding(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    const CompasObjectHandle* trajectory) {
    return catch_exceptions([&] {
        return new_array(compas::phase_encoding(
            context(ctx), array(echos, "echos").as<cfloat, 2>("echos"),
            unwrap<TissueParameters>(parameters, "parameters"),
            unwrap<CartesianTrajectory>(trajectory, "trajectory")));
    });
}

CompasArrayHandle* compas_compute_residual(
    const CompasContextHandle* ctx, const CompasArrayHandle* lhs, const CompasArrayHandle* rhs, float* sum) {
    return catch_exceptions([&] {
        return new_array(compute_residual(
            context(ctx), array(lhs, "lhs").as<cfloat, 3>("lhs"), array(rhs, "rhs").as<cfloat, 3>("rhs"), sum));
    });
}

CompasArrayHandle* compas_simulate_magnetization_derivative_pssfp(
    const CompasContextHandle* ctx, int field, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    float delta, const CompasObjectHandle* sequence) {
    return catch_exceptions([&] {
        return new_array(simulate_magnetization_derivative(
            context(ctx), field, array(echos, "echos").as<cfloat, 2>("echos"),
            unwrap<TissueParameters>(parameters, "parameters"), unwrap<pSSFPSequence>(sequence, "sequence"), delta));
    });
}

CompasArrayHandle* compas_simulate_magnetization_derivative_fisp(
    const CompasContextHandle* ctx, int field, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    float delta, const CompasObjectHandle* sequence) {
    return catch_exceptions([&] {
        return new_array(simulate_magnetization_derivative(
            context(ctx), field, array(echos, "echos").as<cfloat, 2>("echos"),
            unwrap<TissueParameters>(parameters, "parameters"), unwrap<FISPSequence>(sequence, "sequence"), delta));
    });
}

// ---- extensions ----

const char* compas_last_error() {
    return g_last_error.empty() ? nullptr : g_last_error.c_str();
}

void compas_clear_error() {
    g_last_error.clear();
}

CompasObjectHandle* compas_make_fisp_sequence_ex(
    const CompasArrayHandle* RF_train, const CompasArrayHandle* sliceprofiles, float TR, float TE, float TW,
    int max_state, float TI, int undersampling_factor, int repetitions, int inversion_prepulse, int wait_spoiling) {
    return catch_exceptions([&] {
        auto rf = array(RF_train, "RF_train").as<cfloat, 1>("RF_train");
        auto prof = array(sliceprofiles, "sliceprofiles").as<cfloat, 2>("sliceprofiles");
        return new_object(FISPSequence(rf, prof, TR, TE, TW, max_state, TI, undersampling_factor, repetitions,
                                       inversion_prepulse != 0, wait_spoiling != 0));
    });
}

int compas_compute_gemm(
    const CompasContextHandle* ctx, const CompasArrayHandle* result, const CompasArrayHandle* lhs,
    const CompasArrayHandle* rhs, float alpha, float beta, int kind) {
    return catch_exceptions([&] {
        const auto method = gemm_method(kind);
        auto out = array(result, "result").as<float, 2>("result");
        const ArrayHandle& l = array(lhs, "lhs");
        if (l.dtype == DT_BF16) {
            compute_gemm(context(ctx), out, l.as<bfloat16_t, 2>("lhs"),
                         array(rhs, "rhs").as<bfloat16_t, 2>("rhs"), alpha, beta, method);
        } else {
            compute_gemm(context(ctx), out, l.as<float, 2>("lhs"), array(rhs, "rhs").as<float, 2>("rhs"), alpha,
                         beta, method);
        }
        return 1;
    });
}

int compas_compute_complex_gemm(
    const CompasContextHandle* ctx, const CompasArrayHandle* result, const CompasArrayHandle* lhs,
    const CompasArrayHandle* rhs, float alpha, float beta, int kind) {
    return catch_exceptions([&] {
        const auto method = gemm_method(kind);
        auto out = array(result, "result").as<float, 3>("result");
        const ArrayHandle& l = array(lhs, "lhs");
        if (l.dtype == DT_BF16) {
            compute_complex_gemm(context(ctx), out, l.as<bfloat16_t, 3>("lhs"),
                                 array(rhs, "rhs").as<bfloat16_t, 3>("rhs"), alpha, beta, method);
        } else {
            compute_complex_gemm(context(ctx), out, l.as<float, 3>("lhs"), array(rhs, "rhs").as<float, 3>("rhs"),
                                 alpha, beta, method);
        }
        return 1;
    });
}

int compas_convert_complex_to_planar(
    const CompasContextHandle* ctx, const CompasArrayHandle* output, const CompasArrayHandle* input) {
    return catch_exceptions([&] {
        const ArrayHandle& o = array(output, "output");
        auto in = array(input, "input").as<cfloat, 2>("input");
        if (o.dtype == DT_BF16) {
            convert_complex_to_planar(context(ctx), o.as<bfloat16_t, 3>("output"), in);
        } else {
            convert_complex_to_planar(context(ctx), o.as<float, 3>("output"), in);
        }
        return 1;
    });
}

int compas_convert_planar_to_complex(
    const CompasContextHandle* ctx, const CompasArrayHandle* output, const CompasArrayHandle* input) {
    return catch_exceptions([&] {
        convert_planar_to_complex(context(ctx), array(output, "output").as<cfloat, 2>("output"),
                                  array(input, "input").as<float, 3>("input"));
        return 1;
    });
}

int compas_array_info(const CompasArrayHandle* a, int* dtype, int64_t* sizes3) {
    return catch_exceptions([&] {
        const auto& h = array(a, "array");
        if (dtype) *dtype = h.dtype;
        if (sizes3)
            for (int i = 0; i < 3; i++) sizes3[i] = i < h.rank ? h.sizes[i] : 0;
        return h.rank;
    });
}

CompasArrayHandle* compas_magnetization_to_signal_method(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    const CompasArrayHandle* coils, const CompasObjectHandle* trajectory, int method) {
    return catch_exceptions([&] {
        return new_array(magnetization_to_signal(
            context(ctx), array(echos, "echos").as<cfloat, 2>("echos"),
            unwrap<TissueParameters>(parameters, "parameters"), unwrap<Trajectory>(trajectory, "trajectory"),
            array(coils, "coils").as<cfloat, 2>("coils"), signal_method(method)));
    });
}

CompasArrayHandle* compas_compute_jacobian_method(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* dT1,
    const CompasArrayHandle* dT2, const CompasObjectHandle* parameters, const CompasArrayHandle* coils,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* vector, int method) {
    return catch_exceptions([&] {
        return new_array(compute_jacobian(
            context(ctx), array(echos, "echos").as<cfloat, 2>("echos"),
            array(dT1, "delta_echos_T1").as<cfloat, 2>("delta_echos_T1"),
            array(dT2, "delta_echos_T2").as<cfloat, 2>("delta_echos_T2"),
            unwrap<TissueParameters>(parameters, "parameters"),
            unwrap<CartesianTrajectory>(trajectory, "trajectory"), array(coils, "coils").as<cfloat, 2>("coils"),
            array(vector, "vector").as<cfloat, 2>("vector"), jacobian_method(method)));
    });
}

CompasArrayHandle* compas_compute_jacobian_hermitian_method(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* dT1,
    const CompasArrayHandle* dT2, const CompasObjectHandle* parameters, const CompasObjectHandle* trajectory,
    const CompasArrayHandle* coils, const CompasArrayHandle* vector, int method) {
    return catch_exceptions([&] {
        return new_array(compute_jacobian_hermitian(
            context(ctx), array(echos, "echos").as<cfloat, 2>("echos"),
            array(dT1, "delta_echos_T1").as<cfloat, 2>("delta_echos_T1"),
            array(dT2, "delta_echos_T2").as<cfloat, 2>("delta_echos_T2"),
            unwrap<TissueParameters>(parameters, "parameters"),
            unwrap<CartesianTrajectory>(trajectory, "trajectory"), array(coils, "coils").as<cfloat, 2>("coils"),
            array(vector, "vector").as<cfloat, 3>("vector"), jacobian_method(method)));
    });
}

CompasArrayHandle* compas_compute_jacobian_spiral(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* dT1,
    const CompasArrayHandle* dT2, const CompasArrayHandle* dB1, const CompasArrayHandle* dB0,
    const CompasObjectHandle* parameters, const CompasObjectHandle* trajectory, const CompasArrayHandle* coils,
    const CompasArrayHandle* vector) {
    return spiral_product<false>(ctx, echos, dT1, dT2, dB1, dB0, parameters, trajectory, coils, vector,
                                 JacobianComputeMethod::GemmFast);
}

CompasArrayHandle* compas_compute_jacobian_hermitian_spiral(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* dT1,
    const CompasArrayHandle* dT2, const CompasArrayHandle* dB1, const CompasArrayHandle* dB0,
    const CompasObjectHandle* parameters, const CompasObjectHandle* trajectory, const CompasArrayHandle* coils,
    const CompasArrayHandle* vector) {
    return spiral_product<true>(ctx, echos, dT1, dT2, dB1, dB0, parameters, trajectory, coils, vector,
                                JacobianComputeMethod::GemmFast);
}

CompasArrayHandle* compas_compute_jacobian_spiral_method(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* dT1,
    const CompasArrayHandle* dT2, const CompasArrayHandle* dB1, const CompasArrayHandle* dB0,
    const CompasObjectHandle* parameters, const CompasObjectHandle* trajectory, const CompasArrayHandle* coils,
    const CompasArrayHandle* vector, int32_t method) {
    return catch_exceptions([&] {
        return spiral_product<false>(ctx, echos, dT1, dT2, dB1, dB0, parameters, trajectory, coils, vector,
                                     jacobian_method(method));
    });
}

CompasArrayHandle* compas_compute_jacobian_hermitian_spiral_method(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* dT1,
    const CompasArrayHandle* dT2, const CompasArrayHandle* dB1, const CompasArrayHandle* dB0,
    const CompasObjectHandle* parameters, const CompasObjectHandle* trajectory, const CompasArrayHandle* coils,
    const CompasArrayHandle* vector, int32_t method) {
    return catch_exceptions([&] {
        return spiral_product<true>(ctx, echos, dT1, dT2, dB1, dB0, parameters, trajectory, coils, vector,
                                    jacobian_method(method));
    });
}

CompasArrayHandle* compas_simulate_magnetization_tangent_fisp(
    const CompasContextHandle* ctx, int field, const CompasObjectHandle* parameters,
    const CompasObjectHandle* sequence) {
    return catch_exceptions([&] {
        return new_array(simulate_magnetization_tangent(context(ctx), field,
                                                        unwrap<TissueParameters>(parameters, "parameters"),
                                                        unwrap<FISPSequence>(sequence, "sequence")));
    });
}

CompasArrayHandle* compas_simulate_magnetization_tangent_pssfp(
    const CompasContextHandle* ctx, int field, const CompasObjectHandle* parameters,
    const CompasObjectHandle* sequence) {
    return catch_exceptions([&] {
        return new_array(simulate_magnetization_tangent(context(ctx), field,
                                                        unwrap<TissueParameters>(parameters, "parameters"),
                                                        unwrap<pSSFPSequence>(sequence, "sequence")));
    });
}

int compas_simulate_magnetization_tangents_fisp(
    const CompasContextHandle* ctx, const int32_t* fields, int32_t nfields, int32_t with_echos,
    const CompasObjectHandle* parameters, const CompasObjectHandle* sequence, CompasArrayHandle** out) {
    return catch_exceptions(
        [&] { return tangents_impl<FISPSequence>(ctx, fields, nfields, with_echos, parameters, sequence, out); });
}

int compas_simulate_magnetization_tangents_pssfp(
    const CompasContextHandle* ctx, const int32_t* fields, int32_t nfields, int32_t with_echos,
    const CompasObjectHandle* parameters, const CompasObjectHandle* sequence, CompasArrayHandle** out) {
    return catch_exceptions(
        [&] { return tangents_impl<pSSFPSequence>(ctx, fields, nfields, with_echos, parameters, sequence, out); });
}

// ---- request server

namespace {

struct ServerHandle {
    std::unique_ptr<Server> server;
};

// The result frame of the last size query on this thread, so the usual
// query-then-copy pair of compas_server_result_wire reads the result back once.
thread_local uint64_t cached_server = 0;  // Server::id(), never 0
thread_local int64_t cached_ticket = -1;
thread_local std::vector<uint8_t> cached;

Server& server_of(CompasServerHandle* s) {
    if (s == nullptr) throw CompasException("null server");
    return *reinterpret_cast<ServerHandle*>(s)->server;
}

void request_objects(ServeRequest& r, const void* parameters, const void* sequence, const void* trajectory,
                     const void* coils) {
    r.parameters = unwrap<TissueParameters>(parameters, "parameters");
    if (sequence == nullptr) throw CompasException("null handle passed as `sequence`");
    const auto* sh = static_cast<const Handle*>(sequence);
    if (const auto* f = dynamic_cast<const FISPSequence*>(sh)) {
        r.fisp = std::make_shared<FISPSequence>(*f);
    } else if (const auto* q = dynamic_cast<const pSSFPSequence*>(sh)) {
        r.pssfp = std::make_shared<pSSFPSequence>(*q);
    } else {
        throw CompasException("invalid object for `sequence` (expecting a FISP or pSSFP sequence)");
    }
    if (trajectory == nullptr) throw CompasException("null handle passed as `trajectory`");
    const auto* th = static_cast<const Handle*>(trajectory);
    if (const auto* c = dynamic_cast<const CartesianTrajectory*>(th)) {
        r.cartesian = std::make_shared<CartesianTrajectory>(*c);
    } else if (const auto* sp = dynamic_cast<const SpiralTrajectory*>(th)) {
        r.spiral = std::make_shared<SpiralTrajectory>(*sp);
    } else {
        throw CompasException("invalid object for `trajectory` (expecting a Cartesian or spiral trajectory)");
    }
    r.coils = array(coils, "coils").as<cfloat, 2>("coils");
}

}  // namespace

CompasServerHandle* compas_server_create(const CompasContextHandle* ctx, int slots, int64_t memory_budget) {
    return catch_exceptions([&]() -> CompasServerHandle* {
        COMPAS_CHECK(memory_budget > 0);
        auto* h = new ServerHandle {std::make_unique<Server>(context(ctx), slots, size_t(memory_budget))};
        return reinterpret_cast<CompasServerHandle*>(h);
    });
}

void compas_server_destroy(CompasServerHandle* server) {
    catch_exceptions([&] {
        auto* h = reinterpret_cast<ServerHandle*>(server);
        if (h && h->server && cached_server == h->server->id()) {
            cached_server = 0;
            cached_ticket = -1;
            std::vector<uint8_t>().swap(cached);
        }
        delete h;
    });
}

int64_t compas_server_submit_simulate_signal(
    CompasServerHandle* server, const CompasObjectHandle* parameters, const CompasObjectHandle* sequence,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* coils, int method) {
    const int64_t t = catch_exceptions([&]() -> int64_t {
        ServeRequest r;
        r.kind = RequestKind::SimulateSignal;
        request_objects(r, parameters, sequence, trajectory, coils);
        r.signal_method = signal_method(method);
        return server_of(server).submit(std::move(r));
    });
    return g_last_error.empty() ? t : -1;
}

int64_t compas_server_submit_jacobian_hermitian(
    CompasServerHandle* server, const CompasObjectHandle* parameters, const CompasObjectHandle* sequence,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* coils, const CompasArrayHandle* residual,
    int method) {
    const int64_t t = catch_exceptions([&]() -> int64_t {
        ServeRequest r;
        r.kind = RequestKind::JacobianHermitian;
        request_objects(r, parameters, sequence, trajectory, coils);
        r.residual = array(residual, "residual").as<cfloat, 3>("residual");
        r.jacobian_method = jacobian_method(method);
        return server_of(server).submit(std::move(r));
    });
    return g_last_error.empty() ? t : -1;
}

int64_t compas_server_submit_wire(CompasServerHandle* server, const void* frame, int64_t nbytes) {
    const int64_t t = catch_exceptions([&]() -> int64_t {
        Server& s = server_of(server);
        if (nbytes < 0) throw CompasException("negative frame size");
        const wire::Frame f = wire::parse(frame, size_t(nbytes), wire::REQUEST_MAGIC);
        return s.submit(wire::upload_request(s.base(), f));
    });
    return g_last_error.empty() ? t : -1;
}

int64_t compas_server_result_wire(CompasServerHandle* server, int64_t ticket, void* buf, int64_t capacity) {
    const int64_t n = catch_exceptions([&]() -> int64_t {
        Server& s = server_of(server);
        if (cached_server != s.id() || cached_ticket != ticket) {
            if (!s.wait(ticket, 0.0)) throw CompasException("request has not finished");
            cached = wire::encode_result(ticket, s.result(ticket), s.base());
            cached_server = s.id();
            cached_ticket = ticket;
        }
        const int64_t n = int64_t(cached.size());
        if (buf != nullptr && capacity >= n) {
            std::memcpy(buf, cached.data(), cached.size());
            cached_server = 0;
            cached_ticket = -1;
            std::vector<uint8_t>().swap(cached);
        }
        return n;
    });
    return g_last_error.empty() ? n : -1;
}

int64_t compas_wire_describe(const void* frame, int64_t nbytes, char* out, int64_t capacity) {
    const int64_t n = catch_exceptions([&]() -> int64_t {
        if (nbytes < 0) throw CompasException("negative frame size");
        const std::string s = wire::describe_request(wire::parse(frame, size_t(nbytes), wire::REQUEST_MAGIC));
        if (out != nullptr && capacity > 0) {
            const size_t k = std::min(size_t(capacity - 1), s.size());
            std::memcpy(out, s.data(), k);
            out[k] = '\0';
        }
        return int64_t(s.size() + 1);
    });
    return g_last_error.empty() ? n : -1;
}

int compas_server_wait(CompasServerHandle* server, int64_t ticket, double timeout_seconds, double* times) {
    const int rc = catch_exceptions([&]() -> int {
        Server& s = server_of(server);
        if (!s.wait(ticket, timeout_seconds)) return 0;
        const ServeResult r = s.result(ticket);
        if (times) {
            times[0] = r.t_arrival;
            times[1] = r.t_start;
            times[2] = r.t_done;
        }
        if (!r.error.empty()) throw CompasException("request failed: " + r.error);
        return 1;
    });
    return g_last_error.empty() ? rc : -1;
}

CompasArrayHandle* compas_server_result(CompasServerHandle* server, int64_t ticket, int which) {
    return catch_exceptions([&]() -> CompasArrayHandle* {
        const ServeResult r = server_of(server).result(ticket);
        if (which == 0) return new_array(r.echos);
        if (which == 1) {
            if (r.signal.buffer()) return new_array(r.signal);
            return new_array(r.jhy);
        }
        if (which >= 2 && size_t(which - 2) < r.derivatives.size()) return new_array(r.derivatives[size_t(which - 2)]);
        throw CompasException(
            "`which` must be 0 (echos), 1 (signal or J^H.y) or 2.. (finite differences of a J^H.y request)");
    });
}

int compas_server_batch_stats(CompasServerHandle* server, int64_t* out4) {
    return catch_exceptions([&] {
        COMPAS_CHECK(out4 != nullptr);
        const Server::BatchStats b = server_of(server).batch_stats();
        out4[0] = b.launches;
        out4[1] = b.items;
        out4[2] = b.max_requests;
        out4[3] = b.shared_launches;
        return 1;
    });
}

void compas_server_release(CompasServerHandle* server, int64_t ticket) {
    catch_exceptions([&] { server_of(server).release(ticket); });
}

void compas_server_set_kernel_timing(CompasServerHandle* server, int enabled) {
    catch_exceptions([&] { server_of(server).set_kernel_timing(enabled != 0); });
}

int compas_server_read_kernel_timings(CompasServerHandle* server, char* buf, int buflen) {
    return catch_exceptions([&] { return format_timings(server_of(server).collect_timings(), buf, buflen); });
}

double compas_server_now(CompasServerHandle* server) {
    return catch_exceptions([&] { return server_of(server).now(); });
}

void compas_set_kernel_timing(const CompasContextHandle* ctx, int enabled) {
    catch_exceptions([&] { context(ctx).runtime()->set_timing(enabled != 0); });
}

int compas_read_kernel_timings(const CompasContextHandle* ctx, char* buf, int buflen) {
    return catch_exceptions([&] { return format_timings(context(ctx).runtime()->collect_timings(), buf, buflen); });
}

void compas_set_contraction_variant(int variant) {
    catch_exceptions([&] { set_contraction_variant(variant); });
}

void compas_set_contraction_precision(int precision) {
    catch_exceptions([&] { set_contraction_precision_override(precision); });
}

int compas_contraction_variant_count() {
    return contraction_variant_count();
}

int compas_contraction_tuning_report(char* buf, int buflen) {
    return catch_exceptions([&] {
        const std::string s = contraction_tuning_report();
        if (buf && buflen > 0) {
            const size_t n = std::min(size_t(buflen - 1), s.size());
            std::memcpy(buf, s.data(), n);
            buf[n] = '\0';
        }
        return int(s.size() + 1);
    });
}

float compas_bandwidth_probe(const CompasContextHandle* ctx, int64_t nbytes) {
    return catch_exceptions([&] { return bandwidth_probe(context(ctx), nbytes); });
}

void compas_trim_memory(const CompasContextHandle* ctx) {
    catch_exceptions([&] { context(ctx).runtime()->trim(); });
}

}  // extern "C"
