// Wire format of the request server (frame layout in compas/wire.h).

#include "compas/wire.h"

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <set>
#include <sstream>

namespace compas {
namespace wire {

namespace {

constexpr size_t HEADER_BYTES = 16;
constexpr size_t ENTRY_BYTES = 48;
constexpr uint32_t MAX_ENTRIES = 64;
constexpr int64_t MAX_DIM = int64_t(1) << 40;

size_t elem_size(uint32_t dtype) {
    switch (dtype) {
        case F32: return 4;
        case C64: return 8;
        case I32: return 4;
        case F64: return 8;
        case U8: return 1;
        case I64: return 8;
        default: return 0;
    }
}

const char* dtype_name(uint32_t dtype) {
    static const char* names[] = {"f32", "c64", "i32", "f64", "u8", "i64"};
    return dtype <= I64 ? names[dtype] : "?";
}

template<typename T>
T load(const uint8_t* p) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}

[[noreturn]] void bad(const std::string& what) {
    throw CompasException("malformed frame: " + what);
}

size_t pad8(size_t n) {
    return (n + 7) & ~size_t(7);
}

}  // namespace

Frame parse(const void* data, size_t nbytes, uint32_t magic) {
    if (data == nullptr && nbytes > 0) bad("null buffer");
    const auto* p = static_cast<const uint8_t*>(data);
    if (nbytes < HEADER_BYTES) bad("shorter than its header");
    Frame f;
    f.magic = load<uint32_t>(p);
    if (f.magic != magic) bad("wrong magic (not a request/result frame of this format)");
    const uint16_t version = load<uint16_t>(p + 4);
    if (version != VERSION) bad("unsupported version " + std::to_string(version));
    f.kind = load<uint16_t>(p + 6);
    const uint32_t count = load<uint32_t>(p + 8);
    if (count > MAX_ENTRIES) bad("too many entries");
    size_t off = HEADER_BYTES;
    std::set<std::string> seen;
    for (uint32_t i = 0; i < count; i++) {
        if (nbytes - off < ENTRY_BYTES) bad("entry header past the end");
        const uint8_t* h = p + off;
        const void* nul = std::memchr(h, 0, 16);
        if (nul == nullptr) bad("entry name not terminated");
        Entry e;
        e.name.assign(reinterpret_cast<const char*>(h), static_cast<const uint8_t*>(nul) - h);
        if (e.name.empty()) bad("empty entry name");
        if (!seen.insert(e.name).second) bad("entry `" + e.name + "` appears twice");
        e.dtype = load<uint32_t>(h + 16);
        const uint32_t rank = load<uint32_t>(h + 20);
        if (elem_size(e.dtype) == 0) bad("entry `" + e.name + "` has an unknown dtype");
        if (rank > 3) bad("entry `" + e.name + "` has rank above 3");
        e.rank = int(rank);
        uint64_t n = 1;
        for (int d = 0; d < 3; d++) {
            e.dims[d] = load<int64_t>(h + 24 + 8 * d);
            if (d < e.rank) {
                if (e.dims[d] < 0 || e.dims[d] > MAX_DIM) bad("entry `" + e.name + "` has an invalid size");
                n *= uint64_t(e.dims[d]);
                if (n > uint64_t(MAX_DIM)) bad("entry `" + e.name + "` is too large");
            } else {
                e.dims[d] = 1;
            }
        }
        off += ENTRY_BYTES;
        e.nbytes = size_t(n) * elem_size(e.dtype);
        if (pad8(e.nbytes) > nbytes - off) bad("entry `" + e.name + "` runs past the end");
        e.data = p + off;
        off += pad8(e.nbytes);
        f.entries.push_back(std::move(e));
    }
    if (off != nbytes) bad("trailing bytes after the last entry");
    return f;
}

const Entry* Frame::find(const char* name) const {
    for (const auto& e : entries)
        if (e.name == name) return &e;
    return nullptr;
}

const Entry& Frame::get(const char* name, uint32_t dtype, int rank) const {
    const Entry* e = find(name);
    if (e == nullptr) throw CompasException(std::string("request has no entry `") + name + "`");
    if (e->dtype != dtype || e->rank != rank) {
        std::stringstream ss;
        ss << "entry `" << name << "` must be " << dtype_name(dtype) << " of rank " << rank << ", is "
           << dtype_name(e->dtype) << " of rank " << e->rank;
        throw CompasException(ss.str());
    }
    return *e;
}

float Frame::f32(const char* name) const {
    return load<float>(get(name, F32, 0).data);
}

int32_t Frame::i32(const char* name) const {
    return load<int32_t>(get(name, I32, 0).data);
}

int32_t Frame::i32(const char* name, int32_t fallback) const {
    return find(name) ? i32(name) : fallback;
}

Writer::Writer(uint32_t magic, uint16_t kind) : buf_(HEADER_BYTES, 0) {
    std::memcpy(buf_.data(), &magic, 4);
    std::memcpy(buf_.data() + 4, &VERSION, 2);
    std::memcpy(buf_.data() + 6, &kind, 2);
}

void Writer::add(const char* name, uint32_t dtype, int rank, const int64_t* dims, const void* data) {
    const size_t len = std::strlen(name);
    COMPAS_CHECK(len >= 1 && len <= 15 && rank >= 0 && rank <= 3 && elem_size(dtype) > 0);
    uint8_t h[ENTRY_BYTES] = {};
    std::memcpy(h, name, len);
    const uint32_t r = uint32_t(rank);
    std::memcpy(h + 16, &dtype, 4);
    std::memcpy(h + 20, &r, 4);
    size_t n = 1;
    for (int d = 0; d < 3; d++) {
        const int64_t v = d < rank ? dims[d] : 1;
        if (d < rank) n *= size_t(v);
        std::memcpy(h + 24 + 8 * d, &v, 8);
    }
    const size_t nbytes = n * elem_size(dtype);
    buf_.insert(buf_.end(), h, h + ENTRY_BYTES);
    const size_t at = buf_.size();
    buf_.resize(at + pad8(nbytes), 0);
    if (nbytes) std::memcpy(buf_.data() + at, data, nbytes);
    count_++;
}

std::vector<uint8_t> Writer::finish() {
    std::memcpy(buf_.data() + 8, &count_, 4);
    return std::move(buf_);
}

namespace {

const char* const TISSUE_FIELDS[] = {"T1", "T2", "B1", "B0", "rho_x", "rho_y", "x", "y", "z"};

// The request's entries, checked for presence, type and consistent sizes.
struct HostRequest {
    int kind = 0, sequence = 0, trajectory = 0, signal_method = 0, jacobian_method = 0, return_echos = 0;
    int nv = 0;
    const Entry* tissue[9] = {};
    const Entry *RF = nullptr, *profiles = nullptr, *gamma_dt_RF = nullptr, *slice_z = nullptr;
    float TR = 0, TE = 0, TW = 0, TI = 0;
    int max_state = 0, usf = 1, reps = 1, inversion = 1, spoil = 1;
    float dt[3] = {}, gdt[3] = {};
    int nr = 0, ns = 0;
    float delta_t = 0;
    const Entry *k_start = nullptr, *delta_k = nullptr;
    cfloat delta_k0 {0, 0};
    const Entry *coils = nullptr, *residual = nullptr;
};

HostRequest read_request(const Frame& f) {
    if (f.magic != REQUEST_MAGIC) throw CompasException("not a request frame");
    HostRequest r;
    r.kind = f.kind;
    if (r.kind != 0 && r.kind != 1) throw CompasException("request kind must be 0 (simulate + signal) or 1 (J^H.y)");
    r.sequence = f.i32("sequence_type");
    r.trajectory = f.i32("trajectory_type");
    if (r.sequence != 0 && r.sequence != 1) throw CompasException("sequence_type must be 0 (FISP) or 1 (pSSFP)");
    if (r.trajectory != 0 && r.trajectory != 1)
        throw CompasException("trajectory_type must be 0 (Cartesian) or 1 (spiral)");
    r.signal_method = f.i32("signal_method", int(SimulateSignalMethod::MatmulFast));
    r.jacobian_method = f.i32("jacobian_method", int(JacobianComputeMethod::GemmFast));
    if (r.signal_method < 0 || r.signal_method > 5) throw CompasException("invalid signal_method");
    if (r.jacobian_method < 0 || r.jacobian_method > 4) throw CompasException("invalid jacobian_method");
    r.return_echos = f.i32("return_echos", 0);

    const Entry& T1 = f.get("T1", F32, 1);
    if (T1.dims[0] > int64_t(INT32_MAX)) throw CompasException("too many voxels");
    r.nv = int(T1.dims[0]);
    for (int i = 0; i < 9; i++) {
        const char* name = TISSUE_FIELDS[i];
        const bool optional = (i == 2 || i == 3 || i == 8);
        if (optional && f.find(name) == nullptr) continue;
        r.tissue[i] = &f.get(name, F32, 1);
        if (r.tissue[i]->dims[0] != r.nv) throw CompasException(std::string("tissue field ") + name + " has another length than T1");
    }

    r.RF = &f.get("RF", C64, 1);
    const int64_t nTR = r.RF->dims[0];
    r.TR = f.f32("TR");
    if (r.sequence == 0) {
        r.profiles = &f.get("sliceprofiles", C64, 2);
        if (r.profiles->dims[1] != nTR) throw CompasException("sliceprofiles must be [nslices][nTR]");
        r.TE = f.f32("TE");
        r.TW = f.f32("TW");
        r.TI = f.f32("TI");
        r.max_state = f.i32("max_state");
        r.usf = f.i32("undersampling", 1);
        r.reps = f.i32("repetitions", 1);
        r.inversion = f.i32("inversion", 1);
        r.spoil = f.i32("wait_spoiling", 1);
    } else {
        r.gamma_dt_RF = &f.get("gamma_dt_RF", C64, 1);
        const Entry& dt = f.get("dt", F32, 1);
        const Entry& g = f.get("gamma_dt_GRz", F32, 1);
        if (dt.dims[0] != 3 || g.dims[0] != 3) throw CompasException("dt and gamma_dt_GRz hold (ex, inv, pr)");
        std::memcpy(r.dt, dt.data, 12);
        std::memcpy(r.gdt, g.data, 12);
        r.slice_z = &f.get("slice_z", F32, 1);
    }

    r.k_start = &f.get("k_start", C64, 1);
    r.nr = int(r.k_start->dims[0]);
    r.ns = f.i32("samples");
    if (r.ns < 0) throw CompasException("samples must be non-negative");
    r.delta_t = f.f32("delta_t");
    const Entry* dk = f.find("delta_k");
    if (dk == nullptr) throw CompasException("request has no entry `delta_k`");
    if (r.trajectory == 0) {
        if (dk->rank != 0 || (dk->dtype != F32 && dk->dtype != C64))
            throw CompasException("Cartesian delta_k must be an f32 or c64 scalar");
        if (dk->dtype == F32) r.delta_k0 = {load<float>(dk->data), 0.0f};
        else r.delta_k0 = load<cfloat>(dk->data);
    } else {
        r.delta_k = &f.get("delta_k", C64, 1);
        if (r.delta_k->dims[0] != r.nr) throw CompasException("spiral delta_k must hold one step per readout");
    }

    r.coils = &f.get("coils", C64, 2);
    if (r.coils->dims[1] != r.nv) throw CompasException("coils must be [ncoils][nvoxels]");
    if (r.kind == 1) {
        r.residual = &f.get("residual", C64, 3);
        if (r.residual->dims[0] != r.coils->dims[0] || r.residual->dims[1] != r.nr || r.residual->dims[2] != r.ns)
            throw CompasException("residual must be [ncoils][nreadouts][samples]");
    }
    return r;
}

template<typename T, size_t N>
Array<T, N> upload(const CompasContext& ctx, const Entry& e) {
    std::array<int64_t, N> sizes {};
    for (size_t i = 0; i < N; i++) sizes[i] = e.dims[i];
    // entry data is only 8-byte aligned within the frame: the copy reads bytes
    return ctx.allocate_from<T, N>(reinterpret_cast<const T*>(e.data), sizes);
}

// The array read back on `stream`: the request has finished (its completion
// event has fired), so the copy need not queue behind the later work of the
// slot stream that produced it.
template<size_t N>
void put_array(Writer& w, const char* name, const Array<cfloat, N>& a, hipStream_t stream, int device) {
    if (!a.buffer()) return;  // not produced by this kind of request
    std::vector<cfloat> host(static_cast<size_t>(a.numel()));
    if (!host.empty()) {
        COMPAS_HIP_CHECK(hipSetDevice(device));
        COMPAS_HIP_CHECK(hipMemcpyAsync(host.data(), a.data(), a.nbytes(), hipMemcpyDeviceToHost, stream));
        COMPAS_HIP_CHECK(hipStreamSynchronize(stream));
    }
    int64_t dims[3] = {1, 1, 1};
    for (size_t i = 0; i < N; i++) dims[i] = a.size(i);
    w.add(name, C64, int(N), dims, host.data());
}

}  // namespace

ServeRequest upload_request(const CompasContext& ctx, const Frame& f) {
    const HostRequest h = read_request(f);
    ServeRequest r;
    r.kind = h.kind == 0 ? RequestKind::SimulateSignal : RequestKind::JacobianHermitian;
    r.signal_method = SimulateSignalMethod(h.signal_method);
    r.jacobian_method = JacobianComputeMethod(h.jacobian_method);
    r.return_echos = h.return_echos != 0;

    // tissue fields are copied out of the frame (alignment), then uploaded as one map
    std::vector<std::vector<float>> fields(9);
    View<float> views[9];
    for (int i = 0; i < 9; i++) {
        if (!h.tissue[i]) continue;
        fields[i].resize(size_t(h.nv));
        if (h.nv) std::memcpy(fields[i].data(), h.tissue[i]->data, size_t(h.nv) * sizeof(float));
        views[i] = View<float>(fields[i].data(), {h.nv});
    }
    r.parameters = make_tissue_parameters(ctx, h.nv, h.nv, views[0], views[1], views[2], views[3], views[4], views[5],
                                          views[6], views[7], views[8]);

    auto RF = upload<cfloat, 1>(ctx, *h.RF);
    if (h.sequence == 0) {
        r.fisp = std::make_shared<FISPSequence>(RF, upload<cfloat, 2>(ctx, *h.profiles), h.TR, h.TE, h.TW,
                                                h.max_state, h.TI, h.usf, h.reps, h.inversion != 0, h.spoil != 0);
    } else {
        r.pssfp = std::make_shared<pSSFPSequence>(RF, h.TR, upload<cfloat, 1>(ctx, *h.gamma_dt_RF),
                                                  RepetitionData {h.dt[0], h.dt[1], h.dt[2]},
                                                  RepetitionData {h.gdt[0], h.gdt[1], h.gdt[2]},
                                                  upload<float, 1>(ctx, *h.slice_z));
    }
    auto k_start = upload<cfloat, 1>(ctx, *h.k_start);
    if (h.trajectory == 0) {
        r.cartesian = std::make_shared<CartesianTrajectory>(h.nr, h.ns, h.delta_t, k_start, h.delta_k0);
    } else {
        r.spiral = std::make_shared<SpiralTrajectory>(h.nr, h.ns, h.delta_t, k_start, upload<cfloat, 1>(ctx, *h.delta_k));
    }
    r.coils = upload<cfloat, 2>(ctx, *h.coils);
    if (h.residual) r.residual = upload<cfloat, 3>(ctx, *h.residual);
    return r;
}

std::string describe_request(const Frame& f) {
    const HostRequest h = read_request(f);
    std::stringstream ss;
    ss << "{\"kind\": \"" << (h.kind == 0 ? "simulate_signal" : "jacobian_hermitian") << "\", \"sequence\": \""
       << (h.sequence == 0 ? "fisp" : "pssfp") << "\", \"trajectory\": \"" << (h.trajectory == 0 ? "cartesian" : "spiral")
       << "\", \"nvoxels\": " << h.nv << ", \"nTR\": " << h.RF->dims[0] << ", \"nreadouts\": " << h.nr
       << ", \"samples\": " << h.ns << ", \"ncoils\": " << h.coils->dims[0] << ", \"signal_method\": " << h.signal_method
       << ", \"jacobian_method\": " << h.jacobian_method << ", \"return_echos\": " << h.return_echos
       << ", \"fields\": [";
    bool first = true;
    for (int i = 0; i < 9; i++)
        if (h.tissue[i]) {
            ss << (first ? "" : ", ") << "\"" << TISSUE_FIELDS[i] << "\"";
            first = false;
        }
    ss << "]";
    if (h.sequence == 0) {
        ss << ", \"nslices\": " << h.profiles->dims[0] << ", \"max_state\": " << h.max_state;
    } else {
        ss << ", \"nRF\": " << h.gamma_dt_RF->dims[0] << ", \"nz\": " << h.slice_z->dims[0];
    }
    // checksums of the payloads as read (sum of the values in double)
    auto sum = [](const Entry& e) {
        double s = 0;
        const size_t n = e.nbytes / 4;
        for (size_t i = 0; i < n; i++) s += double(load<float>(e.data + 4 * i));
        return s;
    };
    // (JSON has no NaN or infinity: non-finite sums print as null)
    auto num = [](double v) {
        if (!std::isfinite(v)) return std::string("null");
        std::ostringstream o;
        o.precision(17);
        o << v;
        return o.str();
    };
    ss << ", \"sum_T1\": " << num(sum(*h.tissue[0])) << ", \"sum_RF\": " << num(sum(*h.RF))
       << ", \"sum_coils\": " << num(sum(*h.coils));
    if (h.residual) ss << ", \"sum_residual\": " << num(sum(*h.residual));
    ss << "}";
    return ss.str();
}

std::vector<uint8_t> encode_result(int64_t ticket, const ServeResult& result, const CompasContext& readback) {
    const bool failed = !result.error.empty();
    Writer w(RESULT_MAGIC, failed ? 1 : 0);
    w.scalar_i64("ticket", ticket);
    w.scalar_f64("t_arrival", result.t_arrival);
    w.scalar_f64("t_start", result.t_start);
    w.scalar_f64("t_done", result.t_done);
    if (failed) {
        const int64_t n = int64_t(result.error.size());
        w.add("error", U8, 1, &n, result.error.data());
        return w.finish();
    }
    const hipStream_t st = readback.stream();
    if (result.return_echos) put_array(w, "echos", result.echos, st, readback.device());
    put_array(w, "signal", result.signal, st, readback.device());
    put_array(w, "jhy", result.jhy, st, readback.device());
    return w.finish();
}

}  // namespace wire
}  // namespace compas
