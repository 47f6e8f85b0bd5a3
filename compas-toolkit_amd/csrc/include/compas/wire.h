#pragma once
// Wire format of the request server: one self-describing binary frame per
// request and per result, so requests can cross a process or network
// boundary (a socket, a queue, a file) instead of being built through the
// library's object API.  No counterpart in the reference, which has no wire
// or on-disk format: its arrays cross only through compas_make_array_* copies
// (julia-bindings/src/bindings.cpp:50-107).
//
// Frame (little-endian, every part 8-byte aligned):
//   header  u32 magic ("CMPQ" request, "CMPR" result), u16 version (1),
//           u16 kind, u32 entry count, u32 zero
//   entry   char name[16] (NUL-terminated), u32 dtype, u32 rank (0..3),
//           i64 dims[3] (unused dims 1), then prod(dims[:rank]) elements,
//           zero-padded to a multiple of 8 bytes
//   dtypes  0 f32, 1 c64 (re, im f32), 2 i32, 3 f64, 4 u8, 5 i64
//
// Request kinds: 0 simulate + signal, 1 J^H.y.  Entries:
//   sequence_type i32 (0 FISP, 1 pSSFP), trajectory_type i32 (0 Cartesian,
//   1 spiral), signal_method i32 and jacobian_method i32 (optional, the
//   enumerators of SimulateSignalMethod / JacobianComputeMethod),
//   return_echos i32 (optional, 0: the result carries no echoes);
//   tissue  T1 T2 rho_x rho_y x y f32[nv]; B1 B0 z f32[nv] optional (absent:
//           the field is absent, as with a NULL pointer in the C-ABI);
//   FISP    RF c64[nTR], sliceprofiles c64[nslices][nTR], TR TE TW TI f32,
//           max_state i32; undersampling repetitions inversion
//           wait_spoiling i32 optional (1 1 1 1);
//   pSSFP   RF c64[nTR], TR f32, gamma_dt_RF c64[nRF], dt f32[3] and
//           gamma_dt_GRz f32[3] (ex, inv, pr), slice_z f32[nz];
//   trajectory  samples i32, delta_t f32, k_start c64[nr], delta_k f32 or
//           c64 (Cartesian) / c64[nr] (spiral);
//   coils c64[nc][nv]; J^H.y: residual c64[nc][nr][ns].
// Result kinds: 0 done, 1 failed.  Entries: ticket i64, t_arrival t_start
// t_done f64 (seconds since the server started); done: signal
// c64[nc][nr][ns] (simulate + signal) or jhy c64[rows][nv] (J^H.y), and
// echos c64[nr][nv] when the request set return_echos;
// failed: error u8[len] (UTF-8).
//
// Parsing checks every size against the frame before reading it; a
// malformed frame is an error, never a read outside the buffer.

#include <cstdint>
#include <string>
#include <vector>

#include "compas/serving.h"

namespace compas {
namespace wire {

constexpr uint32_t REQUEST_MAGIC = 0x51504d43u;  // "CMPQ"
constexpr uint32_t RESULT_MAGIC = 0x52504d43u;  // "CMPR"
constexpr uint16_t VERSION = 1;
enum DType : uint32_t { F32 = 0, C64 = 1, I32 = 2, F64 = 3, U8 = 4, I64 = 5 };

struct Entry {
    std::string name;
    uint32_t dtype = 0;
    int rank = 0;
    int64_t dims[3] = {1, 1, 1};
    const uint8_t* data = nullptr;
    size_t nbytes = 0;
    int64_t count() const {
        int64_t n = 1;
        for (int i = 0; i < rank; i++) n *= dims[i];
        return n;
    }
};

// A parsed frame: entries point into the caller's buffer.
struct Frame {
    uint32_t magic = 0;
    uint16_t kind = 0;
    std::vector<Entry> entries;

    const Entry* find(const char* name) const;
    // the entry, checked for type and rank (throws when missing or different)
    const Entry& get(const char* name, uint32_t dtype, int rank) const;
    float f32(const char* name) const;
    int32_t i32(const char* name, int32_t fallback) const;  // optional scalar
    int32_t i32(const char* name) const;
};

Frame parse(const void* data, size_t nbytes, uint32_t magic);

class Writer {
  public:
    Writer(uint32_t magic, uint16_t kind);
    void add(const char* name, uint32_t dtype, int rank, const int64_t* dims, const void* data);
    void scalar_f64(const char* name, double v) { add(name, F64, 0, nullptr, &v); }
    void scalar_i64(const char* name, int64_t v) { add(name, I64, 0, nullptr, &v); }
    std::vector<uint8_t> finish();

  private:
    std::vector<uint8_t> buf_;
    uint32_t count_ = 0;
};

// The request of a frame, with its arrays uploaded on `ctx`.
ServeRequest upload_request(const CompasContext& ctx, const Frame& request);

// One-line JSON summary of a request frame (kind, types, sizes); checks the
// request's entries as upload_request does, without a GPU.
std::string describe_request(const Frame& request);

// Result frame of a finished request; its arrays are read back on the
// stream of `readback` (not behind later work of the producing slot).
std::vector<uint8_t> encode_result(int64_t ticket, const ServeResult& result, const CompasContext& readback);

}  // namespace wire
}  // namespace compas
