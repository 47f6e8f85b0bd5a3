/*
 * C-ABI of the MI355X compas simulation hot path (libcompas_amd.so).
 *
 * The first 26 entry points are the exports of the reference's Julia
 * bindings, with the same names, argument order and meaning
 * (julia-bindings/src/bindings.cpp, line of each cited below), so the
 * @ccall wrappers of CompasToolkit.jl bind to this library unchanged.
 * Handles are opaque pointers owned by the caller and released with the
 * matching destroy function.  Array layout is row-major (C order, last index
 * fastest), complex values are interleaved float pairs.
 *
 * Errors: every entry point catches exceptions.  When the process provides
 * `jl_errorf` (a Julia host), it is called with the message and does not
 * return, exactly as in the reference (julia-bindings/src/common.h:50-63).
 * Otherwise the message is stored per thread, the call returns NULL (or
 * nothing), and compas_last_error() returns the message.
 */
#ifndef COMPAS_AMD_H
#define COMPAS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct CompasContextHandle CompasContextHandle; /* compas::CompasContext */
typedef struct CompasObjectHandle CompasObjectHandle;   /* parameters, sequences, trajectories */
typedef struct CompasArrayHandle CompasArrayHandle;     /* device array */
typedef struct CompasServerHandle CompasServerHandle;   /* request server (extension) */
typedef struct CompasGraphHandle CompasGraphHandle;     /* captured graph (extension) */

typedef struct {
    float re;
    float im;
} compas_cfloat;

/* ---- context and arrays ---------------------------------------------- */

/* bindings.cpp:19 */
const char* compas_version(void);
/* bindings.cpp:23 -- context on device 0 */
const CompasContextHandle* compas_make_context(void);
/* bindings.cpp:30 -- copy bound to `device` (modulo the device count) */
const CompasContextHandle* compas_copy_context_for_device(CompasContextHandle* ctx, int device);
/* bindings.cpp:38 */
void compas_destroy_context(const CompasContextHandle* ctx);
/* bindings.cpp:42 -- parameters, sequences and trajectories */
void compas_destroy_object(const CompasObjectHandle* obj);
/* bindings.cpp:46 -- waits for all work queued on the context */
void compas_synchronize(const CompasContextHandle* ctx);
/* bindings.cpp:50 -- copies `data` (rank <= 3, sizes in C order) to the device */
const CompasArrayHandle* compas_make_array_float(
    const CompasContextHandle* ctx, const float* data, int rank, int64_t* sizes);
/* bindings.cpp:69 -- blocking copy of `length` elements to `dest` */
void compas_read_array_float(
    const CompasContextHandle* ctx, const CompasArrayHandle* array, float* dest, int64_t length);
/* bindings.cpp:80 */
const CompasArrayHandle* compas_make_array_complex(
    const CompasContextHandle* ctx, const compas_cfloat* data, int rank, int64_t* sizes);
/* bindings.cpp:98 */
void compas_read_array_complex(
    const CompasContextHandle* ctx, const CompasArrayHandle* array, compas_cfloat* dest, int64_t length);
/* bindings.cpp:109 */
void compas_destroy_array(const CompasArrayHandle* array);

/* ---- data model --------------------------------------------------------- */

/* bindings.cpp:117 -- every field present; nvoxels floats each */
const CompasObjectHandle* compas_make_tissue_parameters(
    const CompasContextHandle* ctx, int nvoxels, const float* T1, const float* T2, const float* B1,
    const float* B0, const float* rho_x, const float* rho_y, const float* x, const float* y, const float* z);
/* bindings.cpp:151 -- RF_train [nTR] complex, sliceprofiles [nslices][nTR] complex */
CompasObjectHandle* compas_make_fisp_sequence(
    const CompasArrayHandle* RF_train, const CompasArrayHandle* sliceprofiles, float TR, float TE, float TW,
    int max_state, float TI, int undersampling_factor, int repetitions);
/* bindings.cpp:176 -- z is a float array [nz] */
CompasObjectHandle* compas_make_pssfp_sequence(
    const CompasArrayHandle* RF_train, float TR, const CompasArrayHandle* gamma_dt_RF, float dt_ex, float dt_inv,
    float dt_pr, float gamma_dt_GRz_ex, float gamma_dt_GRz_inv, float gamma_dt_GRz_pr, const CompasArrayHandle* z);
/* bindings.cpp:198 -- real delta_k, as passed by CompasToolkit.jl */
CompasObjectHandle* compas_make_cartesian_trajectory(
    int nreadouts, int samples_per_readout, float delta_t, CompasArrayHandle* k_start, float delta_k);
/* bindings.cpp:214 */
CompasObjectHandle* compas_make_spiral_trajectory(
    int nreadouts, int samples_per_readout, float delta_t, CompasArrayHandle* k_start, CompasArrayHandle* delta_k);

/* ---- compute ------------------------------------------------------------ */

/* bindings.cpp:230 -- returns echos [nreadouts][nvoxels] complex */
CompasArrayHandle* compas_simulate_magnetization_fisp(
    const CompasContextHandle* ctx, const CompasObjectHandle* parameters, const CompasObjectHandle* sequence);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): compas_simulate_magnetization_fisp into the caller's
 * device buffer `echos` ([nTR * undersampling][nvoxels] complex64, room for
 * `capacity` values, on the context's GPU), with no allocation.  The per-TR
 * RF table goes to a static device array of the library holding at most
 * 65536 (slice profile, TR) pairs per GPU; larger sequences are rejected
 * with an error.  Every call on one GPU shares that array: calls ordered on
 * one stream are safe, calls that overlap on one GPU race.  Returns 1, or 0
 * on error (message through jl_errorf or compas_last_error). */
int compas_simulate_magnetization_fisp_noalloc(
    const CompasContextHandle* ctx, const CompasObjectHandle* parameters, const CompasObjectHandle* sequence,
    compas_cfloat* echos, int64_t capacity);
/* bindings.cpp:243 */
CompasArrayHandle* compas_simulate_magnetization_pssfp(
    const CompasContextHandle* ctx, const CompasObjectHandle* parameters, const CompasObjectHandle* sequence);
/* bindings.cpp:256 -- signal [ncoils][nreadouts][nsamples]; low_precision 0 = MatmulFast, else MatmulLow */
CompasArrayHandle* compas_magnetization_to_signal_cartesian(
    const CompasContextHandle* ctx, int ncoils, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    const CompasArrayHandle* coils, const CompasObjectHandle* trajectory, int32_t low_precision);
/* bindings.cpp:278 */
CompasArrayHandle* compas_magnetization_to_signal_spiral(
    const CompasContextHandle* ctx, int ncoils, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    const CompasArrayHandle* coils, const CompasObjectHandle* trajectory);
/* bindings.cpp:296 -- vector [4][nvoxels]; low_precision 0 = GemmFast, else GemmLow */
CompasArrayHandle* compas_compute_jacobian(
    const CompasContextHandle* ctx, int ncoils, const CompasArrayHandle* echos, const CompasArrayHandle* delta_echos_T1,
    const CompasArrayHandle* delta_echos_T2, const CompasObjectHandle* parameters, const CompasArrayHandle* coils,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* vector, int32_t low_precision);
/* bindings.cpp:324 -- note: trajectory before coils; vector [ncoils][nreadouts][nsamples] */
CompasArrayHandle* compas_compute_jacobian_hermitian(
    const CompasContextHandle* ctx, int ncoils, const CompasArrayHandle* echos, const CompasArrayHandle* delta_echos_T1,
    const CompasArrayHandle* delta_echos_T2, const CompasObjectHandle* parameters, const CompasObjectHandle* trajectory,
    const CompasArrayHandle* coils, const CompasArrayHandle* vector, int32_t low_precision);
/* bindings.cpp:352 -- exported without the compas_ prefix, as in the reference */
CompasArrayHandle* phase_encoding(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    const CompasObjectHandle* trajectory);
/* bindings.cpp:367 -- returns lhs - rhs, writes sum |lhs - rhs|^2 to *sum */
CompasArrayHandle* compas_compute_residual(
    const CompasContextHandle* ctx, const CompasArrayHandle* lhs, const CompasArrayHandle* rhs, float* sum);
/* bindings.cpp:376 */
CompasArrayHandle* compas_simulate_magnetization_derivative_pssfp(
    const CompasContextHandle* ctx, int field, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    float delta, const CompasObjectHandle* sequence);
/* bindings.cpp:395 */
CompasArrayHandle* compas_simulate_magnetization_derivative_fisp(
    const CompasContextHandle* ctx, int field, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    float delta, const CompasObjectHandle* sequence);

/* ---- extensions (no counterpart in the reference) ------------------------ */

/* Error channel for hosts without jl_errorf (Python ctypes, C, Go, ...). */
const char* compas_last_error(void);
void compas_clear_error(void);

/* compas_make_fisp_sequence with the two FISPSequence flags of the C++ API
 * (include/compas/sequences/fisp.h:37-41) that the reference's C-ABI fixes
 * to true: the inversion prepulse and spoiling before the next repetition. */
CompasObjectHandle* compas_make_fisp_sequence_ex(
    const CompasArrayHandle* RF_train, const CompasArrayHandle* sliceprofiles, float TR, float TE, float TW,
    int max_state, float TI, int undersampling_factor, int repetitions, int inversion_prepulse, int wait_spoiling);

/* Shape of an array handle: dtype 0 = float32, 1 = complex64, 2 = bfloat16.
 * Returns rank. */
int compas_array_info(const CompasArrayHandle* array, int* dtype, int64_t* sizes3);

/* GEMM utility (include/compas/utils/gemm.h:7-71 and src/utils/gemm.cu:8-254
 * in the reference, a C++ API there with no C-ABI): result = alpha * lhs *
 * transpose(rhs) + beta * result, written in place into `result`.
 * compas_compute_gemm: result float [m][n], lhs [m][k], rhs [n][k], both
 * float or both bfloat16.  compas_compute_complex_gemm: planar complex
 * result float [2][m][n], lhs [2][m][k], rhs [2][n][k] (real plane first).
 * kind = GemmComputeMethod {Pedantic, Regular, Fast} as 0..2: for float
 * operands FP32 FMA on VALU, BF16 hi + lo on MFMA (FP32-level), BF16 on MFMA
 * (one product); bfloat16 operands run exact products with FP32
 * accumulation for every method.  beta = 0 does not read `result`.
 * compas_convert_complex_to_planar: complex [n][m] into float or bfloat16
 * [2][n][m]; compas_convert_planar_to_complex: float [2][n][m] into complex
 * [n][m] (gemm_kernels.cuh:9-32).  bfloat16 arrays hold raw bf16 bits
 * (compas_make_array_bf16 / compas_read_array_bf16).  Each returns 1, or 0 on
 * error. */
const CompasArrayHandle* compas_make_array_bf16(
    const CompasContextHandle* ctx, const uint16_t* data, int rank, int64_t* sizes);
void compas_read_array_bf16(
    const CompasContextHandle* ctx, const CompasArrayHandle* array, uint16_t* dest, int64_t length);
int compas_compute_gemm(
    const CompasContextHandle* ctx, const CompasArrayHandle* result, const CompasArrayHandle* lhs,
    const CompasArrayHandle* rhs, float alpha, float beta, int kind);
int compas_compute_complex_gemm(
    const CompasContextHandle* ctx, const CompasArrayHandle* result, const CompasArrayHandle* lhs,
    const CompasArrayHandle* rhs, float alpha, float beta, int kind);
int compas_convert_complex_to_planar(
    const CompasContextHandle* ctx, const CompasArrayHandle* output, const CompasArrayHandle* input);
int compas_convert_planar_to_complex(
    const CompasContextHandle* ctx, const CompasArrayHandle* output, const CompasArrayHandle* input);

/* Method-selecting variants of the signal and Jacobian entry points, exposing
 * the C++ enums SimulateSignalMethod {Naive, Direct, MatmulPedantic, Matmul,
 * MatmulFast, MatmulLow} and JacobianComputeMethod {Naive, Direct, Gemm,
 * GemmFast, GemmLow} as integers 0..5 / 0..4. */
CompasArrayHandle* compas_magnetization_to_signal_method(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasObjectHandle* parameters,
    const CompasArrayHandle* coils, const CompasObjectHandle* trajectory, int method);
CompasArrayHandle* compas_compute_jacobian_method(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* delta_echos_T1,
    const CompasArrayHandle* delta_echos_T2, const CompasObjectHandle* parameters, const CompasArrayHandle* coils,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* vector, int method);
CompasArrayHandle* compas_compute_jacobian_hermitian_method(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* delta_echos_T1,
    const CompasArrayHandle* delta_echos_T2, const CompasObjectHandle* parameters, const CompasObjectHandle* trajectory,
    const CompasArrayHandle* coils, const CompasArrayHandle* vector, int method);

/* Jacobian products on a spiral trajectory with six parameters (T1, T2, B1,
 * B0, rho_x, rho_y): the reference's Jacobian is Cartesian with four, so these
 * have no counterpart there; they extend its conventions (T rows scaled by T,
 * rho rows multiply m and i m) to the derivatives of its spiral signal.
 * echos and the four derivative arrays are [nreadouts][nvoxels]; J.v takes
 * vector [6][nvoxels] and returns [ncoils][nreadouts][nsamples]; J^H.y takes
 * [ncoils][nreadouts][nsamples] and returns [6][nvoxels].  These two use
 * JacobianComputeMethod GemmFast (mixed precision, as the reference's
 * bindings choose for the Cartesian products when low_precision is 0); the
 * _method variants take the enumerator (0 Naive .. 4 GemmLow).  Naive,
 * Direct and Gemm sum on FP32 VALU; GemmFast and GemmLow run J^H.y on MFMA
 * (BF16 hi + lo, or hi only). */
CompasArrayHandle* compas_compute_jacobian_spiral(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* delta_echos_T1,
    const CompasArrayHandle* delta_echos_T2, const CompasArrayHandle* delta_echos_B1,
    const CompasArrayHandle* delta_echos_B0, const CompasObjectHandle* parameters,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* coils, const CompasArrayHandle* vector);
CompasArrayHandle* compas_compute_jacobian_hermitian_spiral(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* delta_echos_T1,
    const CompasArrayHandle* delta_echos_T2, const CompasArrayHandle* delta_echos_B1,
    const CompasArrayHandle* delta_echos_B0, const CompasObjectHandle* parameters,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* coils, const CompasArrayHandle* vector);
CompasArrayHandle* compas_compute_jacobian_spiral_method(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* delta_echos_T1,
    const CompasArrayHandle* delta_echos_T2, const CompasArrayHandle* delta_echos_B1,
    const CompasArrayHandle* delta_echos_B0, const CompasObjectHandle* parameters,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* coils, const CompasArrayHandle* vector,
    int32_t method);
CompasArrayHandle* compas_compute_jacobian_hermitian_spiral_method(
    const CompasContextHandle* ctx, const CompasArrayHandle* echos, const CompasArrayHandle* delta_echos_T1,
    const CompasArrayHandle* delta_echos_T2, const CompasArrayHandle* delta_echos_B1,
    const CompasArrayHandle* delta_echos_B0, const CompasObjectHandle* parameters,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* coils, const CompasArrayHandle* vector,
    int32_t method);

/* Forward-mode derivative of the FISP echoes [nreadouts][nvoxels] with
 * respect to one tissue parameter (field: 0 T1, 1 T2, 2 B1, 3 B0): the exact
 * derivative of the discrete recursion, computed beside it in one pass.  No
 * counterpart in the reference, which differentiates by FP32 finite
 * differences (compas_simulate_magnetization_derivative_fisp). */
CompasArrayHandle* compas_simulate_magnetization_tangent_fisp(
    const CompasContextHandle* ctx, int field, const CompasObjectHandle* parameters,
    const CompasObjectHandle* sequence);

/* The same for the pSSFP isochromat simulation (FP64 state, FP32 result);
 * replaces compas_simulate_magnetization_derivative_pssfp's finite
 * differences (reference julia-bindings/src/bindings.cpp:376). */
CompasArrayHandle* compas_simulate_magnetization_tangent_pssfp(
    const CompasContextHandle* ctx, int field, const CompasObjectHandle* parameters,
    const CompasObjectHandle* sequence);

/* Derivatives with respect to nfields distinct fields (0 T1, 1 T2, 2 B1,
 * 3 B0) in one call, and with with_echos != 0 the echoes as well: writes
 * the echoes (if asked) and then one [nreadouts][nvoxels] array per field, in
 * the order given, to out[] (room for nfields + 1 handles) and returns the
 * number written (0 on error).  pSSFP runs T1 + T2 and all four fields, with
 * or without the echoes, as one pass sharing the isochromat recursion -- where
 * the reference runs one simulation per field plus the echoes
 * (compas_simulate_magnetization_derivative_pssfp, julia-bindings/src/
 * bindings.cpp:376, called once per field).  FISP runs one pass per field. */
int compas_simulate_magnetization_tangents_fisp(
    const CompasContextHandle* ctx, const int32_t* fields, int32_t nfields, int32_t with_echos,
    const CompasObjectHandle* parameters, const CompasObjectHandle* sequence, CompasArrayHandle** out);
int compas_simulate_magnetization_tangents_pssfp(
    const CompasContextHandle* ctx, const int32_t* fields, int32_t nfields, int32_t with_echos,
    const CompasObjectHandle* parameters, const CompasObjectHandle* sequence, CompasArrayHandle** out);

/* Graph capture (no counterpart in the reference, whose kmm runtime
 * schedules every call afresh): the library calls made on `ctx` between
 * compas_graph_begin and compas_graph_end are recorded, not run, into one HIP
 * graph that compas_graph_launch replays on the context's stream.  Arrays
 * those calls return are written by every launch; memory the captured calls
 * allocate stays reserved for the graph until compas_graph_destroy.  The
 * recorded calls must not synchronise the host (no reads or uploads of
 * arrays, no compute_residual): run the sequence once before capturing it so
 * that first-call autotuning is done.  compas_graph_abort ends a capture
 * that failed.  compas_write_array_complex and compas_copy_array
 * (stream-ordered device copy) refresh a graph's input arrays between
 * launches. */
void compas_graph_begin(const CompasContextHandle* ctx);
CompasGraphHandle* compas_graph_end(const CompasContextHandle* ctx);
void compas_graph_abort(const CompasContextHandle* ctx);
void compas_graph_launch(const CompasGraphHandle* graph);
void compas_graph_destroy(const CompasGraphHandle* graph);
void compas_write_array_complex(
    const CompasContextHandle* ctx, const CompasArrayHandle* array, const compas_cfloat* src, int64_t length);
void compas_copy_array(const CompasContextHandle* ctx, const CompasArrayHandle* dst, const CompasArrayHandle* src);

/* Zero-copy interop (no counterpart in the reference, where arrays cross
 * only by copies: compas_make_array_* and compas_read_array_*, julia-
 * bindings/src/bindings.cpp:50-107).
 *
 * compas_wrap_device_array: an array over device memory owned by the caller
 * (a Julia ROCArray, a torch tensor, ...), no copy.  dtype 0 float, 1
 * complex; the memory must be device memory of the context's GPU, aligned to
 * 256 bytes, row-major.  `release(user)` runs once the last array over it is
 * destroyed, after the context's stream has finished the work queued on it
 * (NULL: nothing is called).
 * compas_array_device_pointer: the device address of an array, for the
 * caller to wrap (valid while the array handle lives).
 * compas_array_from_dlpack / compas_array_to_dlpack: the same through DLPack
 * (legacy DLManagedTensor, device type kDLROCM, float32 or complex64,
 * compact row-major).  from_dlpack takes ownership of the tensor on success
 * and calls its deleter when the last array over it is gone; on failure
 * (NULL returned) the caller keeps it.  to_dlpack returns a new
 * DLManagedTensor whose deleter drops its reference to the array's memory.
 *
 * Stream order: the library runs on the context's stream
 * (compas_context_stream).  Before the library reads memory written on
 * another stream, call compas_context_wait_stream(ctx, that stream); before
 * another stream reads a library result, call compas_stream_wait_context(ctx,
 * that stream).  Both record an event and enqueue a wait; the host does not
 * block. */
typedef void (*compas_release_fn)(void* user);
void* compas_context_stream(const CompasContextHandle* ctx);
void compas_context_wait_stream(const CompasContextHandle* ctx, void* stream);
void compas_stream_wait_context(const CompasContextHandle* ctx, void* stream);
const CompasArrayHandle* compas_wrap_device_array(
    const CompasContextHandle* ctx, void* device_ptr, int32_t dtype, int32_t rank, const int64_t* sizes,
    compas_release_fn release, void* user);
void* compas_array_device_pointer(const CompasArrayHandle* array);
int32_t compas_array_device(const CompasArrayHandle* array); /* GPU ordinal of the array's memory */
void compas_dlpack_delete(void* dl_managed_tensor);         /* calls its deleter (unconsumed exports) */
const CompasArrayHandle* compas_array_from_dlpack(const CompasContextHandle* ctx, void* dl_managed_tensor);
void* compas_array_to_dlpack(const CompasArrayHandle* array);

/* Request server (no counterpart in the reference, whose host is a Julia
 * session calling one operation at a time): continuous batching on one GPU.
 * Pending requests are admitted while the estimated device memory of the
 * admitted ones fits `memory_budget` bytes (arrival order; a request that
 * does not fit lets later ones pass for 0.5 s, then holds the queue for all
 * but small requests of at most 1/64 of the budget; 128 admitted at most).  The
 * simulations of all admitted requests run as work items (voxel blocks,
 * slice-position groups) of shared work-table launches in fair-share order,
 * so one launch serves several requests; a request's contraction (signal or
 * J^H.y) then runs on one of `slots` private streams.  Requests reference
 * resident objects (parameters, sequence, trajectory, coils, residual
 * handles).  Tickets are >= 0; -1 means the submission failed
 * (compas_last_error).
 *   simulate_signal: simulate_magnetization + magnetization_to_signal
 *     (FISP or pSSFP, Cartesian or spiral), method = SimulateSignalMethod;
 *   jacobian_hermitian: simulation, finite-difference derivatives and J^H.y
 *     (Cartesian: 4 parameters; spiral: 6), method = JacobianComputeMethod.
 * compas_server_wait returns 1 when done (times = arrival, start, done in
 * seconds on the server clock; start = admission), 0 on timeout, -1 if the
 * request failed.
 * compas_server_result: which = 0 echos, 1 signal or J^H.y, 2.. the finite
 * differences of a J^H.y request (T1, T2, then B1, B0 for spiral) (new
 * handle); results stay in HBM until compas_server_release.
 * compas_server_batch_stats: out4 = work-table launches, work items, the
 * most requests one launch served, launches that served two or more.
 * Destroying the server waits for the running and pending requests. */
CompasServerHandle* compas_server_create(const CompasContextHandle* ctx, int slots, int64_t memory_budget);
void compas_server_destroy(CompasServerHandle* server);
int64_t compas_server_submit_simulate_signal(
    CompasServerHandle* server, const CompasObjectHandle* parameters, const CompasObjectHandle* sequence,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* coils, int method);
int64_t compas_server_submit_jacobian_hermitian(
    CompasServerHandle* server, const CompasObjectHandle* parameters, const CompasObjectHandle* sequence,
    const CompasObjectHandle* trajectory, const CompasArrayHandle* coils, const CompasArrayHandle* residual,
    int method);
int compas_server_wait(CompasServerHandle* server, int64_t ticket, double timeout_seconds, double* times);
CompasArrayHandle* compas_server_result(CompasServerHandle* server, int64_t ticket, int which);
void compas_server_release(CompasServerHandle* server, int64_t ticket);
int compas_server_batch_stats(CompasServerHandle* server, int64_t* out4);
double compas_server_now(CompasServerHandle* server);
/* Per-kernel HIP-event timing on every slot stream of the server; the read
 * has the format of compas_read_kernel_timings (merged over the slots). */
void compas_server_set_kernel_timing(CompasServerHandle* server, int enabled);
int compas_server_read_kernel_timings(CompasServerHandle* server, char* buf, int buflen);

/* Wire format of the server (no counterpart in the reference, which has no
 * wire or on-disk format): a request arrives as one self-describing binary
 * frame of named, typed arrays (layout in csrc/include/compas/wire.h) and
 * its result leaves as one.  compas_server_submit_wire parses and checks the
 * frame (every size against the buffer), uploads its arrays and submits the
 * request: a ticket, or -1 (compas_last_error) for a malformed or
 * inconsistent frame.  compas_server_result_wire encodes a finished
 * request's result (or its error) into buf when `capacity` suffices and
 * returns the frame's size in bytes (-1 if the request is unknown or not
 * finished).  compas_wire_describe checks a request frame without a GPU and
 * writes a one-line JSON summary (kind, sizes, payload sums) into out;
 * returns the bytes needed, -1 on a malformed frame. */
int64_t compas_server_submit_wire(CompasServerHandle* server, const void* frame, int64_t nbytes);
int64_t compas_server_result_wire(CompasServerHandle* server, int64_t ticket, void* buf, int64_t capacity);
int64_t compas_wire_describe(const void* frame, int64_t nbytes, char* out, int64_t capacity);

/* Per-kernel timing with HIP events on the context's stream.  While enabled,
 * every library launch is bracketed by events; compas_read_kernel_timings
 * synchronizes, writes "name count total_ms\n" lines into buf (NUL-terminated,
 * truncated to buflen) and resets the counters.  Returns bytes needed. */
void compas_set_kernel_timing(const CompasContextHandle* ctx, int enabled);
int compas_read_kernel_timings(const CompasContextHandle* ctx, char* buf, int buflen);

/* MFMA contraction tile variants.  By default each new contraction shape is
 * timed once over all variants and the fastest is cached (set the environment
 * variable COMPAS_AUTOTUNE=0 to use a fixed default).  A variant >= 0 forces
 * that variant for every call (tests use this to check each one); -1 restores
 * autotuning.  The report lists "M=.. N=.. K=.. prec=.. -> variant" lines. */
void compas_set_contraction_variant(int variant);
int compas_contraction_variant_count(void);
int compas_contraction_tuning_report(char* buf, int buflen);

/* Overrides the contraction precision that every signal / Jacobian method
 * selects, process-wide, for side-by-side accuracy and speed studies:
 * 0 FP32 VALU, 1 BF16 hi+lo (3 MFMA products), 2 FP16 hi+lo with per-row
 * power-of-two scaling, 3 BF16 x1, 4 FP16 x1 (scaled); -1 clears it. */
void compas_set_contraction_precision(int precision);

/* Streaming copy of `nbytes` between two device buffers (kernel
 * "stream_copy"); returns milliseconds.  Measures achievable HBM bandwidth
 * and serves as the known-size reference for hardware traffic counters. */
float compas_bandwidth_probe(const CompasContextHandle* ctx, int64_t nbytes);

/* Release cached device memory held by the context's pool. */
void compas_trim_memory(const CompasContextHandle* ctx);

#ifdef __cplusplus
}
#endif

#endif /* COMPAS_AMD_H */
