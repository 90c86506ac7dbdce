// capi_fixed_grid.cpp -- extern "C" entry points of the fixed-grid path:
// status plumbing, stage launchers, the workspace, device-frame views, the
// one-shot API and the device plumbing helpers.
//
// Behavioural contract follows the reference's host API:
//   stage launchers   include/beamformer/cuda_stage_api.hpp:14-42
//   workspace         src/cuda_beamformer.cu:545-928
//   device frame      src/cuda_beamformer.cu:930-1079, cuda_frame.hpp:62-230
//   one-shot          src/cuda_beamformer.cu:1081-1173
#include <hip/hip_runtime.h>

#include <array>
#include <chrono>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>

#include "bf_common.hpp"
#include "bf_kernels.hpp"

namespace bfhip {

namespace {
thread_local std::string g_last_error;
}

bf_status set_error(bf_status code, const char* msg) {
    g_last_error = msg ? msg : "";
    return code;
}

namespace {

using Clock = std::chrono::steady_clock;

template <typename T>
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t count) { allocate(count); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void allocate(size_t count) {
        release();
        if (count == 0) count = 1;
        hip_check(hipMalloc(reinterpret_cast<void**>(&ptr_), count * sizeof(T)), "hipMalloc");
        count_ = count;
    }
    void release() {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        count_ = 0;
    }
    T* get() const { return ptr_; }
    bool empty() const { return ptr_ == nullptr; }

  private:
    T* ptr_ = nullptr;
    size_t count_ = 0;
};

class Stream {
  public:
    Stream() { hip_check(hipStreamCreate(&s_), "hipStreamCreate"); }
    ~Stream() {
        if (s_) (void)hipStreamDestroy(s_);
    }
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;
    hipStream_t get() const { return s_; }

  private:
    hipStream_t s_ = nullptr;
};

class Event {
  public:
    Event() { hip_check(hipEventCreate(&e_), "hipEventCreate"); }
    ~Event() {
        if (e_) (void)hipEventDestroy(e_);
    }
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    hipEvent_t get() const { return e_; }

  private:
    hipEvent_t e_ = nullptr;
};

double elapsed(const Event& a, const Event& b) {
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, a.get(), b.get()), "hipEventElapsedTime");
    return ms;
}

void require_ptr(const void* p, const char* what) {
    if (p == nullptr) invalid(std::string(what) + " must not be null");
}

void require_aligned(const void* p, size_t align, const char* what) {
    if ((reinterpret_cast<uintptr_t>(p) % align) != 0)
        invalid(std::string(what) + " must be " + std::to_string(align) + "-byte aligned");
}

bool is_reference_kernel(int k) { return k == BF_KERNEL_DIRECT || k == BF_KERNEL_TILED; }

void validate_kernel(int k) {
    if (k < BF_KERNEL_DIRECT || k > BF_KERNEL_AUTO) invalid("unknown beamformer kernel selector");
}

size_t weight_storage_count(const bf_dims& d, int kernel) {
    return kernel == BF_KERNEL_TILED ? tiled_weight_count(d) : d.n_beams * d.n_freq * d.n_ant;
}

// Reference src/cuda_beamformer.cu:418-433 for the reference's selectors;
// the additions accept any positive window length.
void validate_integration(int kernel, size_t S) {
    if (S == 0) invalid("integration_spectra must be positive");
    if (is_reference_kernel(kernel)) {
        if (S != 10 && S != 320)
            invalid("the Direct and Tiled selectors integrate over 10 or 320 spectra only");
        if (kernel == BF_KERNEL_DIRECT && S != 320)
            invalid("the Direct selector integrates over 320 spectra only");
    }
    if ((kernel == BF_KERNEL_SPLIT_F16 || kernel == BF_KERNEL_F16) && !(S == 10 || S % 4 == 0))
        invalid("the MFMA kernels support integration over 10 spectra or a multiple of 4");
}

// S == 0 selects the per-spectrum output.
int resolve(int kernel, const bf_dims& d, size_t S) {
    if (kernel == BF_KERNEL_AUTO)
        return mfma_supported(d.n_ant, d.n_beams, S) ? BF_KERNEL_SPLIT_F16 : BF_KERNEL_FP32;
    return kernel;
}

// `mask` (optional): [T][F] frame-validity bytes; only the MFMA kernels
// apply it.
void run_beamformer(int kernel, hipStream_t stream, const uint8_t* x, const float* w,
                    float* out, const bf_dims& d, size_t S, const uint8_t* mask = nullptr) {
    require_ptr(x, "packed_voltage");
    require_ptr(w, "weights");
    require_ptr(out, "output");
    require_aligned(x, 16, "packed_voltage");
    require_aligned(w, 8, "weights");
    require_aligned(out, 4, "output");
    const int resolved = resolve(kernel, d, S);
    if (mask != nullptr && resolved != BF_KERNEL_SPLIT_F16 && resolved != BF_KERNEL_F16)
        invalid("loss masks are applied by the MFMA kernels (SPLIT_F16, F16, or AUTO where "
                "it selects one)");
    switch (resolved) {
    case BF_KERNEL_DIRECT:
    case BF_KERNEL_FP32:
        launch_fp32_fused(stream, x, w, WeightLayout::Canonical, out, d.n_time, d.n_freq,
                          d.n_ant, d.n_beams, S);
        return;
    case BF_KERNEL_TILED:
        launch_fp32_fused(stream, x, w, WeightLayout::Tiled, out, d.n_time, d.n_freq, d.n_ant,
                          d.n_beams, S);
        return;
    case BF_KERNEL_SPLIT_F16:
        launch_mfma_fused(stream, true, x, w, out, d.n_time, d.n_freq, d.n_ant, d.n_beams, S,
                          mask);
        return;
    case BF_KERNEL_F16:
        launch_mfma_fused(stream, false, x, w, out, d.n_time, d.n_freq, d.n_ant, d.n_beams, S,
                          mask);
        return;
    default:
        invalid("unknown beamformer kernel selector");
    }
}

bf_dims integrated_dims_of(const bf_dims& d, size_t S) {
    return bf_dims{integrated_time_count(d.n_time, S), d.n_freq, d.n_ant, d.n_beams};
}

// ---------------------------------------------------------- frame views
size_t dtype_size(int t) {
    switch (t) {
    case BF_DTYPE_PACKED_INT4X2:
    case BF_DTYPE_SIGNED_INT8: return 1;
    case BF_DTYPE_COMPLEX_FLOAT32: return 8;
    case BF_DTYPE_FLOAT32: return 4;
    default: invalid("unknown frame element type");
    }
}

bool same_shard(const bf_shard& a, const bf_shard& b) {
    return a.shard_id == b.shard_id && a.shard_count == b.shard_count &&
           a.local_frequency_count == b.local_frequency_count &&
           a.absolute_frequency_start == b.absolute_frequency_start &&
           a.timestamp_start == b.timestamp_start && a.timestamp_step == b.timestamp_step &&
           a.loss_mask_id == b.loss_mask_id &&
           (a.loss_mask_independent != 0) == (b.loss_mask_independent != 0);
}

// Reference include/beamformer/cuda_frame.hpp:62-149.
size_t element_count(const bf_frame_view& v) {
    if (v.rank == 0 || v.rank > 4) invalid("a frame view has rank 1 to 4");
    size_t n = 1;
    for (size_t i = 0; i < v.rank; ++i) {
        if (v.shape[i] == 0 || n > std::numeric_limits<size_t>::max() / v.shape[i])
            invalid("a frame view's shape has a zero extent or its element count overflows");
        n *= v.shape[i];
    }
    for (size_t i = v.rank; i < 4; ++i)
        if (v.shape[i] != 0) invalid("a frame view's shape entries past its rank must be 0");
    return n;
}

size_t minimum_storage(const bf_frame_view& v) {
    element_count(v);
    size_t last = dtype_size(v.type);
    for (size_t i = 0; i < v.rank; ++i) {
        if (v.strides_bytes[i] == 0) invalid("a frame view's strides must be nonzero");
        const size_t extent = v.shape[i] - 1;
        if (extent > (std::numeric_limits<size_t>::max() - last) / v.strides_bytes[i])
            invalid("a frame view's byte extent overflows");
        last += extent * v.strides_bytes[i];
    }
    for (size_t i = v.rank; i < 4; ++i)
        if (v.strides_bytes[i] != 0) invalid("a frame view's strides past its rank must be 0");
    return last;
}

bool is_contiguous(const bf_frame_view& v) {
    element_count(v);
    size_t stride = dtype_size(v.type);
    for (size_t i = v.rank; i-- > 0;) {
        if (v.strides_bytes[i] != stride) return false;
        stride *= v.shape[i];
    }
    return true;
}

void validate_frame(const bf_frame_view* v, const char* label) {
    if (v == nullptr) invalid(std::string(label) + " frame must not be null");
    if (v->device_data == nullptr) invalid("a frame view's device pointer is null");
    if (v->bytes < minimum_storage(*v))
        invalid("a frame view's byte count is below what its shape and strides address");
    validate_shard(v->shard);
}

void validate_frame_shape(const bf_frame_view& v, int type, size_t rank,
                          const std::array<size_t, 4>& shape, const char* label) {
    if (!is_contiguous(v)) invalid("the frame view must be densely packed (row-major strides)");
    bool same = v.type == type && v.rank == rank;
    for (int i = 0; i < 4; ++i) same = same && v.shape[i] == shape[i];
    if (!same) invalid(std::string(label) + " shape or dtype does not match");
}

// ------------------------------------------------ non-allocating launcher
// The f32 integrated intermediate of bf_launch_packed_quantized_integrated_
// beamformer_noalloc lives in this file-scope device array: every GPU the
// library runs on holds its own copy, at its own address, until the process
// ends.  Host code never passes `scratch` itself to a kernel (in host code
// the name is a stand-in, not a device address); it uses the address
// hipGetSymbolAddress returns, once per device, with that device current.
__device__ __align__(256) float scratch[BF_NOALLOC_SCRATCH_BYTES / sizeof(float)];

constexpr int kMaxScratchDevices = 256;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxScratchDevices];

float* scratch_for_current_device() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxScratchDevices)
        fail(BF_ERR_RUNTIME, "device ordinal " + std::to_string(dev) +
                                 " is beyond the devices the scratch table covers");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                  "hipGetSymbolAddress(scratch)");
        if (p == nullptr) fail(BF_ERR_RUNTIME, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace bfhip

using namespace bfhip;

// ===================================================================== C ABI
extern "C" {

const char* bf_last_error_message(void) { return g_last_error.c_str(); }

const char* bf_version(void) { return "beamform_project_amd 0.1 (gfx950)"; }

bf_status bf_validate_dimensions(const bf_dims* dims) {
    return guarded([&] { validate_dimensions(deref_dims(dims)); });
}

bf_status bf_integrated_time_count(size_t n_time, size_t spectra, size_t* out) {
    return guarded([&] {
        require_ptr(out, "out");
        *out = integrated_time_count(n_time, spectra);
    });
}

bf_status bf_quantization_parameter_count(const bf_dims* d, size_t* out) {
    return guarded([&] {
        require_ptr(out, "out");
        validate_dimensions(deref_dims(d));
        *out = quantization_parameter_count(*d);
    });
}

bf_status bf_resolve_kernel(int kernel, const bf_dims* dims, size_t S, int* resolved) {
    return guarded([&] {
        validate_kernel(kernel);
        require_ptr(resolved, "resolved");
        *resolved = resolve(kernel, deref_dims(dims), S);
    });
}

bf_status bf_launch_packed_beamformer(int kernel, void* stream, const uint8_t* packed,
                                      const float* weights, float* intensity,
                                      const bf_dims* dims) {
    return guarded([&] {
        validate_kernel(kernel);
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        run_beamformer(kernel, static_cast<hipStream_t>(stream), packed, weights, intensity, d,
                       0);
    });
}

bf_status bf_launch_packed_integrated_beamformer(int kernel, void* stream,
                                                 const uint8_t* packed, const float* weights,
                                                 float* integrated, const bf_dims* dims,
                                                 size_t S) {
    return guarded([&] {
        validate_kernel(kernel);
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        validate_integration(kernel, S);
        run_beamformer(kernel, static_cast<hipStream_t>(stream), packed, weights, integrated, d,
                       S);
    });
}

bf_status bf_launch_packed_beamformer_masked(int kernel, void* stream, const uint8_t* packed,
                                             const uint8_t* loss_mask, const float* weights,
                                             float* intensity, const bf_dims* dims) {
    return guarded([&] {
        validate_kernel(kernel);
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        require_ptr(loss_mask, "loss_mask");
        run_beamformer(kernel, static_cast<hipStream_t>(stream), packed, weights, intensity, d,
                       0, loss_mask);
    });
}

bf_status bf_launch_packed_integrated_beamformer_masked(int kernel, void* stream,
                                                        const uint8_t* packed,
                                                        const uint8_t* loss_mask,
                                                        const float* weights, float* integrated,
                                                        const bf_dims* dims, size_t S) {
    return guarded([&] {
        validate_kernel(kernel);
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        validate_integration(kernel, S);
        require_ptr(loss_mask, "loss_mask");
        run_beamformer(kernel, static_cast<hipStream_t>(stream), packed, weights, integrated, d,
                       S, loss_mask);
    });
}

bf_status bf_autotune_packed_beamformer(int kernel, void* stream, const uint8_t* packed,
                                        const float* weights, float* output,
                                        const bf_dims* dims, size_t S, int repeats,
                                        bf_tune_entry* entries, size_t max_entries,
                                        size_t* n_entries) {
    return guarded([&] {
        validate_kernel(kernel);
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        if (S != 0) validate_integration(kernel, S);
        const int k = resolve(kernel, d, S);
        if (k != BF_KERNEL_SPLIT_F16 && k != BF_KERNEL_F16)
            invalid("the autotuner covers the SplitF16 and F16 MFMA kernels");
        require_ptr(packed, "packed_voltage");
        require_ptr(weights, "weights");
        require_ptr(output, "output");
        require_aligned(packed, 16, "packed_voltage");
        require_aligned(weights, 8, "weights");
        require_aligned(output, 4, "output");
        const auto res = autotune_mfma_fused(static_cast<hipStream_t>(stream),
                                             k == BF_KERNEL_SPLIT_F16, packed, weights, output,
                                             d.n_time, d.n_freq, d.n_ant, d.n_beams, S,
                                             repeats);
        if (n_entries) *n_entries = res.size();
        for (size_t i = 0; entries && i < res.size() && i < max_entries; ++i)
            entries[i] = bf_tune_entry{res[i].cfg.nt * 8, res[i].cfg.waves, res[i].cfg.db,
                                       res[i].cfg.resident, (double)res[i].ms};
    });
}

bf_status bf_autotune_clear(void) {
    return guarded([&] { clear_mfma_tuning(); });
}

bf_status bf_launch_quantize_integrated_intensity(void* stream, const float* integrated,
                                                  int8_t* codes, float* params,
                                                  const bf_dims* integrated_dims) {
    return guarded([&] {
        const bf_dims& d = deref_dims(integrated_dims);
        validate_dimensions(d);
        require_ptr(integrated, "integrated_intensity");
        require_ptr(codes, "quantized_intensity");
        require_ptr(params, "parameters");
        launch_quantize(static_cast<hipStream_t>(stream), integrated, codes, params, d.n_time,
                        d.n_freq, d.n_beams);
    });
}

bf_status bf_launch_packed_quantized_integrated_beamformer_noalloc(
    int kernel, void* stream, const uint8_t* packed, const float* weights, int8_t* codes,
    float* params, const bf_dims* dims, size_t S) {
    return guarded([&] {
        validate_kernel(kernel);
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        validate_integration(kernel, S);
        require_ptr(codes, "quantized_intensity");
        require_ptr(params, "parameters");
        const bf_dims id = integrated_dims_of(d, S);
        const size_t need = id.n_time * id.n_freq * id.n_beams * sizeof(float);
        if (need > BF_NOALLOC_SCRATCH_BYTES)
            invalid("the integrated intermediate of " + std::to_string(need) +
                    " bytes exceeds the " + std::to_string(BF_NOALLOC_SCRATCH_BYTES) +
                    "-byte scratch of the non-allocating launcher; split the frame or use "
                    "bf_launch_packed_integrated_beamformer with a caller buffer");
        float* mid = scratch_for_current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        // Every window of the intermediate is written by the beamformer
        // before the quantiser reads it, so nothing needs clearing.
        run_beamformer(kernel, s, packed, weights, mid, d, S);
        launch_quantize(s, mid, codes, params, id.n_time, id.n_freq, id.n_beams);
    });
}

// ------------------------------------------------------------- workspace
}  // extern "C"

struct bf_workspace {
    bf_dims capacity{};
    int kernel = BF_KERNEL_DIRECT;
    size_t spectra = 0;  // 0: no temporal integration
    int output = BF_OUTPUT_FLOAT32;
    DeviceBuffer<uint8_t> voltage;
    DeviceBuffer<float> weights;
    DeviceBuffer<float> intensity;  // allocated on first per-spectrum use
    DeviceBuffer<float> integrated;
    DeviceBuffer<int8_t> codes;
    DeviceBuffer<float> params;
    Stream stream;
    Event start, transfer_end, kernel_end, quantization_end, result_end;
    double setup_ms = 0.0;
    size_t loaded_voltage = 0;
    size_t loaded_weights = 0;

    // Reference src/cuda_beamformer.cu:573-579.
    void validate_request(const bf_dims& d) const {
        validate_dimensions(d);
        if (d.n_freq != capacity.n_freq || d.n_ant != capacity.n_ant ||
            d.n_time > capacity.n_time || d.n_beams > capacity.n_beams)
            invalid("the request exceeds the workspace capacity (n_time and n_beams may shrink; n_freq and n_ant are fixed)");
    }
    // Reference src/cuda_beamformer.cu:581-605.
    void require_loaded(const bf_dims& d) const {
        if (loaded_voltage < d.n_time * d.n_freq * d.n_ant)
            fail(BF_ERR_LOGIC, "packed voltage has not been uploaded for these dimensions");
        require_weights(d);
    }
    void require_weights(const bf_dims& d) const {
        if (loaded_weights != weight_storage_count(d, kernel))
            fail(BF_ERR_LOGIC, "weights have not been uploaded for these dimensions");
    }
    void require_integration() const {
        if (spectra == 0 || integrated.empty())
            fail(BF_ERR_LOGIC, "this workspace was created without temporal integration");
    }
    void require_quantization() const {
        require_integration();
        if (output != BF_OUTPUT_INT8 || codes.empty() || params.empty())
            fail(BF_ERR_LOGIC, "this workspace was created without int8 output");
    }
    float* intensity_buffer() {
        if (intensity.empty())
            intensity.allocate(capacity.n_time * capacity.n_freq * capacity.n_beams);
        return intensity.get();
    }
};

extern "C" {

bf_status bf_workspace_create(const bf_dims* capacity, int kernel, size_t spectra, int output,
                              bf_workspace** out) {
    return guarded([&] {
        require_ptr(out, "out");
        *out = nullptr;
        validate_kernel(kernel);
        const bf_dims& cap = deref_dims(capacity);
        validate_dimensions(cap);
        if (spectra != 0) validate_integration(kernel, spectra);
        if (output != BF_OUTPUT_FLOAT32 && output != BF_OUTPUT_INT8)
            invalid("unknown workspace output selector");
        if (output == BF_OUTPUT_INT8 && spectra == 0)
            invalid("int8 quantization requires a temporal-integration configuration");
        const auto t0 = Clock::now();
        auto ws = std::make_unique<bf_workspace>();
        ws->capacity = cap;
        ws->kernel = kernel;
        ws->spectra = spectra;
        ws->output = output;
        ws->voltage.allocate(cap.n_time * cap.n_freq * cap.n_ant);
        ws->weights.allocate(2 * weight_storage_count(cap, kernel));
        if (spectra != 0) {
            const bf_dims idims = integrated_dims_of(cap, spectra);
            const size_t n = idims.n_time * idims.n_freq * idims.n_beams;
            ws->integrated.allocate(n);
            if (output == BF_OUTPUT_INT8) {
                ws->codes.allocate(n);
                ws->params.allocate(2 * quantization_parameter_count(idims));
            }
        } else {
            ws->intensity_buffer();
        }
        ws->setup_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        *out = ws.release();
    });
}

bf_status bf_workspace_destroy(bf_workspace* ws) {
    return guarded([&] { delete ws; });
}

bf_status bf_workspace_setup_ms(const bf_workspace* ws, double* out) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        require_ptr(out, "out");
        *out = ws->setup_ms;
    });
}

bf_status bf_workspace_kernel(const bf_workspace* ws, int* out) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        require_ptr(out, "out");
        *out = ws->kernel;
    });
}

bf_status bf_workspace_has_temporal_integration(const bf_workspace* ws, int* out) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        require_ptr(out, "out");
        *out = ws->spectra != 0;
    });
}

bf_status bf_workspace_has_int8_quantization(const bf_workspace* ws, int* out) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        require_ptr(out, "out");
        *out = ws->output == BF_OUTPUT_INT8;
    });
}

bf_status bf_workspace_packed_voltage_capacity_bytes(const bf_workspace* ws, size_t* out) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        require_ptr(out, "out");
        *out = ws->capacity.n_time * ws->capacity.n_freq * ws->capacity.n_ant;
    });
}

bf_status bf_workspace_upload_packed_voltage(bf_workspace* ws, const uint8_t* host,
                                             size_t host_bytes, const bf_dims* dims,
                                             double* ms) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        const bf_dims& d = deref_dims(dims);
        ws->validate_request(d);
        const size_t count = d.n_time * d.n_freq * d.n_ant;
        if (host == nullptr || host_bytes < count)
            invalid("packed voltage is smaller than dimensions");
        hipStream_t s = ws->stream.get();
        hip_check(hipEventRecord(ws->start.get(), s), "hipEventRecord");
        hip_check(hipMemcpyAsync(ws->voltage.get(), host, count, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync packed voltage host to device");
        hip_check(hipEventRecord(ws->transfer_end.get(), s), "hipEventRecord");
        hip_check(hipEventSynchronize(ws->transfer_end.get()), "hipEventSynchronize");
        ws->loaded_voltage = count;
        if (ms) *ms = elapsed(ws->start, ws->transfer_end);
    });
}

bf_status bf_workspace_upload_weights(bf_workspace* ws, const float* host, size_t count,
                                      const bf_dims* dims, double* ms) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        const bf_dims& d = deref_dims(dims);
        ws->validate_request(d);
        const size_t need = weight_storage_count(d, ws->kernel);
        if (host == nullptr || count != need)
            invalid("weight count does not match selected kernel layout");
        hipStream_t s = ws->stream.get();
        hip_check(hipEventRecord(ws->start.get(), s), "hipEventRecord");
        hip_check(hipMemcpyAsync(ws->weights.get(), host, need * 2 * sizeof(float),
                                 hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync weights host to device");
        hip_check(hipEventRecord(ws->transfer_end.get(), s), "hipEventRecord");
        hip_check(hipEventSynchronize(ws->transfer_end.get()), "hipEventSynchronize");
        ws->loaded_weights = need;
        if (ms) *ms = elapsed(ws->start, ws->transfer_end);
    });
}

bf_status bf_workspace_run_kernel(bf_workspace* ws, const bf_dims* dims, double* ms) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        const bf_dims& d = deref_dims(dims);
        ws->validate_request(d);
        ws->require_loaded(d);
        hipStream_t s = ws->stream.get();
        hip_check(hipEventRecord(ws->start.get(), s), "hipEventRecord");
        run_beamformer(ws->kernel, s, ws->voltage.get(), ws->weights.get(),
                       ws->intensity_buffer(), d, 0);
        hip_check(hipEventRecord(ws->kernel_end.get(), s), "hipEventRecord");
        hip_check(hipEventSynchronize(ws->kernel_end.get()), "hipEventSynchronize");
        if (ms) *ms = elapsed(ws->start, ws->kernel_end);
    });
}

bf_status bf_workspace_download_intensity(bf_workspace* ws, float* host, size_t host_count,
                                          const bf_dims* dims, double* ms) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        const bf_dims& d = deref_dims(dims);
        ws->validate_request(d);
        const size_t count = d.n_time * d.n_freq * d.n_beams;
        if (host == nullptr || host_count < count)
            invalid("intensity output is smaller than dimensions");
        hipStream_t s = ws->stream.get();
        hip_check(hipEventRecord(ws->start.get(), s), "hipEventRecord");
        hip_check(hipMemcpyAsync(host, ws->intensity_buffer(), count * sizeof(float),
                                 hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync intensity device to host");
        hip_check(hipEventRecord(ws->result_end.get(), s), "hipEventRecord");
        hip_check(hipEventSynchronize(ws->result_end.get()), "hipEventSynchronize");
        if (ms) *ms = elapsed(ws->start, ws->result_end);
    });
}

bf_status bf_workspace_run_pipeline(bf_workspace* ws, const uint8_t* packed,
                                    size_t packed_bytes, float* intensity,
                                    size_t intensity_count, const bf_dims* dims,
                                    bf_timings* timings) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        const bf_dims& d = deref_dims(dims);
        ws->validate_request(d);
        const size_t n_in = d.n_time * d.n_freq * d.n_ant;
        const size_t n_out = d.n_time * d.n_freq * d.n_beams;
        if (packed == nullptr || packed_bytes < n_in)
            invalid("packed voltage is smaller than dimensions");
        if (intensity == nullptr || intensity_count < n_out)
            invalid("intensity output is smaller than dimensions");
        ws->require_weights(d);
        hipStream_t s = ws->stream.get();
        float* dev_out = ws->intensity_buffer();
        hip_check(hipEventRecord(ws->start.get(), s), "hipEventRecord");
        hip_check(hipMemcpyAsync(ws->voltage.get(), packed, n_in, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync pipeline packed voltage host to device");
        hip_check(hipEventRecord(ws->transfer_end.get(), s), "hipEventRecord");
        run_beamformer(ws->kernel, s, ws->voltage.get(), ws->weights.get(), dev_out, d, 0);
        hip_check(hipEventRecord(ws->kernel_end.get(), s), "hipEventRecord");
        hip_check(hipMemcpyAsync(intensity, dev_out, n_out * sizeof(float),
                                 hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync pipeline intensity device to host");
        hip_check(hipEventRecord(ws->result_end.get(), s), "hipEventRecord");
        hip_check(hipEventSynchronize(ws->result_end.get()), "hipEventSynchronize");
        ws->loaded_voltage = n_in;
        if (timings) {
            *timings = bf_timings{};
            timings->host_to_device_ms = elapsed(ws->start, ws->transfer_end);
            timings->kernel_ms = elapsed(ws->transfer_end, ws->kernel_end);
            timings->device_to_host_ms = elapsed(ws->kernel_end, ws->result_end);
        }
    });
}

bf_status bf_workspace_run_integrated_pipeline(bf_workspace* ws, const uint8_t* packed,
                                               size_t packed_bytes, float* integrated,
                                               size_t integrated_count, const bf_dims* dims,
                                               bf_timings* timings) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        const bf_dims& d = deref_dims(dims);
        ws->validate_request(d);
        ws->require_integration();
        const size_t n_in = d.n_time * d.n_freq * d.n_ant;
        const bf_dims id = integrated_dims_of(d, ws->spectra);
        const size_t n_out = id.n_time * id.n_freq * id.n_beams;
        if (packed == nullptr || packed_bytes < n_in)
            invalid("packed voltage is smaller than dimensions");
        if (integrated == nullptr || integrated_count < n_out)
            invalid("integrated intensity output is smaller than dimensions");
        ws->require_weights(d);
        hipStream_t s = ws->stream.get();
        hip_check(hipEventRecord(ws->start.get(), s), "hipEventRecord");
        hip_check(hipMemcpyAsync(ws->voltage.get(), packed, n_in, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync integrated pipeline packed voltage host to device");
        hip_check(hipEventRecord(ws->transfer_end.get(), s), "hipEventRecord");
        run_beamformer(ws->kernel, s, ws->voltage.get(), ws->weights.get(),
                       ws->integrated.get(), d, ws->spectra);
        hip_check(hipEventRecord(ws->kernel_end.get(), s), "hipEventRecord");
        hip_check(hipMemcpyAsync(integrated, ws->integrated.get(), n_out * sizeof(float),
                                 hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync integrated pipeline intensity device to host");
        hip_check(hipEventRecord(ws->result_end.get(), s), "hipEventRecord");
        hip_check(hipEventSynchronize(ws->result_end.get()), "hipEventSynchronize");
        ws->loaded_voltage = n_in;
        if (timings) {
            *timings = bf_timings{};
            timings->host_to_device_ms = elapsed(ws->start, ws->transfer_end);
            timings->kernel_ms = elapsed(ws->transfer_end, ws->kernel_end);
            timings->device_to_host_ms = elapsed(ws->kernel_end, ws->result_end);
        }
    });
}

bf_status bf_workspace_run_quantized_integrated_pipeline(
    bf_workspace* ws, const uint8_t* packed, size_t packed_bytes, int8_t* codes,
    size_t code_count, float* parameters, size_t parameter_count, const bf_dims* dims,
    bf_timings* timings) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        const bf_dims& d = deref_dims(dims);
        ws->validate_request(d);
        ws->require_quantization();
        const size_t n_in = d.n_time * d.n_freq * d.n_ant;
        const bf_dims id = integrated_dims_of(d, ws->spectra);
        const size_t n_out = id.n_time * id.n_freq * id.n_beams;
        const size_t n_par = quantization_parameter_count(id);
        if (packed == nullptr || packed_bytes < n_in)
            invalid("packed voltage is smaller than dimensions");
        if (codes == nullptr || parameters == nullptr || code_count < n_out ||
            parameter_count < n_par)
            invalid("quantized output is smaller than dimensions");
        ws->require_weights(d);
        hipStream_t s = ws->stream.get();
        hip_check(hipEventRecord(ws->start.get(), s), "hipEventRecord");
        hip_check(hipMemcpyAsync(ws->voltage.get(), packed, n_in, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync quantized pipeline packed voltage host to device");
        hip_check(hipEventRecord(ws->transfer_end.get(), s), "hipEventRecord");
        run_beamformer(ws->kernel, s, ws->voltage.get(), ws->weights.get(),
                       ws->integrated.get(), d, ws->spectra);
        hip_check(hipEventRecord(ws->kernel_end.get(), s), "hipEventRecord");
        launch_quantize(s, ws->integrated.get(), ws->codes.get(), ws->params.get(), id.n_time,
                        id.n_freq, id.n_beams);
        hip_check(hipEventRecord(ws->quantization_end.get(), s), "hipEventRecord");
        hip_check(hipMemcpyAsync(codes, ws->codes.get(), n_out, hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync quantized pipeline codes device to host");
        hip_check(hipMemcpyAsync(parameters, ws->params.get(), n_par * 2 * sizeof(float),
                                 hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync quantized pipeline parameters device to host");
        hip_check(hipEventRecord(ws->result_end.get(), s), "hipEventRecord");
        hip_check(hipEventSynchronize(ws->result_end.get()), "hipEventSynchronize");
        ws->loaded_voltage = n_in;
        if (timings) {
            *timings = bf_timings{};
            timings->host_to_device_ms = elapsed(ws->start, ws->transfer_end);
            timings->kernel_ms = elapsed(ws->transfer_end, ws->kernel_end);
            timings->quantization_ms = elapsed(ws->kernel_end, ws->quantization_end);
            timings->device_to_host_ms = elapsed(ws->quantization_end, ws->result_end);
        }
    });
}

bf_status bf_workspace_run_device_frame(bf_workspace* ws, const bf_frame_view* voltage,
                                        const bf_frame_view* weights, bf_frame_view* output,
                                        bf_frame_view* qparams, bf_timings* timings) {
    return guarded([&] {
        require_ptr(ws, "workspace");
        validate_frame(voltage, "voltage");
        validate_frame(weights, "weights");
        validate_frame(output, "output");
        if (voltage->frame_id != output->frame_id)
            invalid("output frame_id does not match input");
        if (!same_shard(voltage->shard, output->shard))
            invalid("output shard metadata does not match input");
        if (voltage->type != BF_DTYPE_PACKED_INT4X2 || voltage->rank != 3)
            invalid("offline device frame voltage must be packed int4x2 with rank 3");
        const bf_dims d{voltage->shape[0], voltage->shape[1], voltage->shape[2],
                        output->shape[2]};
        ws->validate_request(d);
        if (voltage->shape[1] != ws->capacity.n_freq || output->rank != 3 ||
            output->shape[1] != d.n_freq || output->shape[2] != d.n_beams)
            invalid("offline device frame dimensions do not match");
        const bool tiled = ws->kernel == BF_KERNEL_TILED;
        const std::array<size_t, 4> wshape =
            tiled ? std::array<size_t, 4>{d.n_freq, ceil_div(d.n_beams, 32), d.n_ant, 32}
                  : std::array<size_t, 4>{d.n_beams, d.n_freq, d.n_ant, 0};
        validate_frame_shape(*weights, BF_DTYPE_COMPLEX_FLOAT32, tiled ? 4 : 3, wshape,
                             "offline device frame weights");
        if (!same_shard(voltage->shard, weights->shard))
            invalid("offline device frame weights shard does not match input");
        bf_dims od = d;
        if (ws->spectra != 0) od.n_time = integrated_time_count(d.n_time, ws->spectra);
        const int out_type =
            ws->output == BF_OUTPUT_INT8 ? BF_DTYPE_SIGNED_INT8 : BF_DTYPE_FLOAT32;
        validate_frame_shape(*output, out_type, 3, {od.n_time, od.n_freq, od.n_beams, 0},
                             "offline device frame output");
        if (ws->output == BF_OUTPUT_INT8) {
            if (ws->spectra == 0 || qparams == nullptr)
                invalid("offline int8 output requires temporal integration parameters");
            validate_frame(qparams, "quantization parameters");
            if (voltage->frame_id != qparams->frame_id)
                invalid("quantization parameters frame_id does not match input");
            if (!same_shard(voltage->shard, qparams->shard))
                invalid("quantization parameters shard metadata does not match input");
            validate_frame_shape(*qparams, BF_DTYPE_FLOAT32, 4,
                                 {od.n_time, ceil_div(od.n_freq, 16), ceil_div(od.n_beams, 16),
                                  2},
                                 "offline quantization parameters");
        } else if (qparams != nullptr) {
            invalid("float32 offline output must not provide quantization parameters");
        }
        hipStream_t s = ws->stream.get();
        hip_check(hipEventRecord(ws->start.get(), s), "hipEventRecord");
        hip_check(hipEventRecord(ws->transfer_end.get(), s), "hipEventRecord");
        const uint8_t* x = static_cast<const uint8_t*>(voltage->device_data);
        const float* w = static_cast<const float*>(weights->device_data);
        float* produced = nullptr;
        if (ws->spectra != 0) {
            run_beamformer(ws->kernel, s, x, w, ws->integrated.get(), d, ws->spectra);
            produced = ws->integrated.get();
        } else {
            produced = ws->intensity_buffer();
            run_beamformer(ws->kernel, s, x, w, produced, d, 0);
        }
        hip_check(hipEventRecord(ws->kernel_end.get(), s), "hipEventRecord");
        const Event* producer_end = &ws->kernel_end;
        size_t out_bytes = od.n_time * od.n_freq * od.n_beams * sizeof(float);
        const void* src = produced;
        if (ws->output == BF_OUTPUT_INT8) {
            launch_quantize(s, ws->integrated.get(), ws->codes.get(), ws->params.get(),
                            od.n_time, od.n_freq, od.n_beams);
            hip_check(hipEventRecord(ws->quantization_end.get(), s), "hipEventRecord");
            producer_end = &ws->quantization_end;
            out_bytes = od.n_time * od.n_freq * od.n_beams;
            src = ws->codes.get();
        }
        hip_check(hipMemcpyAsync(output->device_data, src, out_bytes, hipMemcpyDeviceToDevice, s),
                  "hipMemcpyAsync offline frame output device to device");
        if (ws->output == BF_OUTPUT_INT8) {
            hip_check(hipMemcpyAsync(qparams->device_data, ws->params.get(),
                                     quantization_parameter_count(od) * 2 * sizeof(float),
                                     hipMemcpyDeviceToDevice, s),
                      "hipMemcpyAsync offline frame parameters device to device");
        }
        hip_check(hipEventRecord(ws->result_end.get(), s), "hipEventRecord");
        hip_check(hipEventSynchronize(ws->result_end.get()), "hipEventSynchronize");
        if (timings) {
            *timings = bf_timings{};
            timings->kernel_ms = elapsed(ws->transfer_end, ws->kernel_end);
            if (ws->output == BF_OUTPUT_INT8)
                timings->quantization_ms = elapsed(ws->kernel_end, ws->quantization_end);
            timings->device_to_device_ms = elapsed(*producer_end, ws->result_end);
        }
    });
}

// ------------------------------------------------- offline frame runner
}  // extern "C"

// One offline frame from host buffers: temporary device copies of the
// caller's voltages and weights, the workspace's device-frame path, and the
// result copied back (src/cuda_offline_runner.cu:130-228).
struct bf_offline_frame_runner {
    bf_workspace* ws = nullptr;
    bf_dims capacity{};
    int kernel = BF_KERNEL_DIRECT;
    size_t spectra = 0;
    int output = BF_OUTPUT_FLOAT32;
    ~bf_offline_frame_runner() {
        if (ws) (void)bf_workspace_destroy(ws);
    }
};

namespace {

bf_frame_view contiguous_view(void* data, int type, size_t rank,
                              const std::array<size_t, 4>& shape, uint64_t frame_id,
                              const bf_shard& shard) {
    bf_frame_view v{};
    v.device_data = data;
    v.type = type;
    v.rank = rank;
    size_t stride = dtype_size(type);
    for (size_t i = rank; i-- > 0;) {
        v.shape[i] = shape[i];
        v.strides_bytes[i] = stride;
        stride *= shape[i];
    }
    v.bytes = stride;
    v.frame_id = frame_id;
    v.shard = shard;
    return v;
}

void rethrow_status(bf_status st) {
    if (st != BF_OK) fail(st, std::string(bf_last_error_message()));
}

}  // namespace

extern "C" {

bf_status bf_offline_frame_runner_create(const bf_dims* capacity, int kernel,
                                         size_t integration_spectra, int output,
                                         bf_offline_frame_runner** out) {
    return guarded([&] {
        require_ptr(out, "out");
        *out = nullptr;
        auto r = std::make_unique<bf_offline_frame_runner>();
        rethrow_status(bf_workspace_create(capacity, kernel, integration_spectra, output, &r->ws));
        r->capacity = *capacity;
        r->kernel = kernel;
        r->spectra = integration_spectra;
        r->output = output;
        *out = r.release();
    });
}

bf_status bf_offline_frame_runner_destroy(bf_offline_frame_runner* r) {
    return guarded([&] { delete r; });
}

bf_status bf_offline_frame_runner_config(const bf_offline_frame_runner* r, int* kernel,
                                         int* output, int* has_temporal_integration) {
    return guarded([&] {
        require_ptr(r, "offline frame runner");
        if (kernel) *kernel = r->kernel;
        if (output) *output = r->output;
        if (has_temporal_integration) *has_temporal_integration = r->spectra != 0;
    });
}

bf_status bf_offline_frame_runner_run(bf_offline_frame_runner* r, const uint8_t* packed,
                                      size_t packed_bytes, const float* weights,
                                      size_t weight_count, const bf_dims* dims,
                                      uint64_t frame_id, const bf_shard* shard,
                                      float* float_output, size_t float_count,
                                      int8_t* codes, size_t code_count, float* parameters,
                                      size_t parameter_count, bf_dims* output_dims,
                                      bf_timings* timings) {
    return guarded([&] {
        require_ptr(r, "offline frame runner");
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        require_ptr(shard, "shard");
        rethrow_status(bf_validate_shard_descriptor(shard));
        const bf_dims& cap = r->capacity;
        if (d.n_freq != cap.n_freq || d.n_ant != cap.n_ant || d.n_time > cap.n_time ||
            d.n_beams > cap.n_beams)
            invalid("the frame's dimensions exceed the runner's capacity");
        if (packed == nullptr || packed_bytes != d.n_time * d.n_freq * d.n_ant)
            invalid("packed voltage byte count does not match the frame's dimensions");
        const bool tiled = r->kernel == BF_KERNEL_TILED;
        const size_t wcount = tiled ? d.n_freq * ceil_div(d.n_beams, 32) * d.n_ant * 32
                                    : d.n_beams * d.n_freq * d.n_ant;
        if (weights == nullptr || weight_count != wcount)
            invalid("weight count does not match the kernel's weight layout");
        bf_dims od = d;
        if (r->spectra != 0) od.n_time = integrated_time_count(d.n_time, r->spectra);
        const size_t n_out = od.n_time * od.n_freq * od.n_beams;
        const bool int8 = r->output == BF_OUTPUT_INT8;
        const size_t n_par = int8 ? quantization_parameter_count(od) : 0;
        if (int8) {
            if (codes == nullptr || code_count < n_out)
                invalid("code buffer is smaller than the frame's integrated output");
            if (parameters == nullptr || parameter_count < n_par)
                invalid("parameter buffer is smaller than the frame's chunk count");
        } else if (float_output == nullptr || float_count < n_out) {
            invalid("float output buffer is smaller than the frame's output");
        }
        bf_timings t{};
        rethrow_status(bf_workspace_setup_ms(r->ws, &t.setup_ms));
        DeviceBuffer<uint8_t> dx(packed_bytes);
        DeviceBuffer<float> dw(2 * wcount);
        const auto h0 = Clock::now();
        hip_check(hipMemcpy(dx.get(), packed, packed_bytes, hipMemcpyHostToDevice),
                  "hipMemcpy offline voltage host to device");
        hip_check(hipMemcpy(dw.get(), weights, 2 * wcount * sizeof(float),
                            hipMemcpyHostToDevice),
                  "hipMemcpy offline weights host to device");
        t.host_to_device_ms =
            std::chrono::duration<double, std::milli>(Clock::now() - h0).count();
        const bf_frame_view vin =
            contiguous_view(dx.get(), BF_DTYPE_PACKED_INT4X2, 3,
                            {d.n_time, d.n_freq, d.n_ant, 0}, frame_id, *shard);
        const bf_frame_view vw =
            tiled ? contiguous_view(dw.get(), BF_DTYPE_COMPLEX_FLOAT32, 4,
                                    {d.n_freq, ceil_div(d.n_beams, 32), d.n_ant, 32}, frame_id,
                                    *shard)
                  : contiguous_view(dw.get(), BF_DTYPE_COMPLEX_FLOAT32, 3,
                                    {d.n_beams, d.n_freq, d.n_ant, 0}, frame_id, *shard);
        DeviceBuffer<float> dout_f;
        DeviceBuffer<int8_t> dout_q;
        DeviceBuffer<float> dpar;
        bf_frame_view vout{}, vpar{};
        if (int8) {
            dout_q.allocate(n_out);
            dpar.allocate(2 * n_par);
            vout = contiguous_view(dout_q.get(), BF_DTYPE_SIGNED_INT8, 3,
                                   {od.n_time, od.n_freq, od.n_beams, 0}, frame_id, *shard);
            vpar = contiguous_view(dpar.get(), BF_DTYPE_FLOAT32, 4,
                                   {od.n_time, ceil_div(od.n_freq, 16), ceil_div(od.n_beams, 16),
                                    2},
                                   frame_id, *shard);
        } else {
            dout_f.allocate(n_out);
            vout = contiguous_view(dout_f.get(), BF_DTYPE_FLOAT32, 3,
                                   {od.n_time, od.n_freq, od.n_beams, 0}, frame_id, *shard);
        }
        bf_timings dt{};
        rethrow_status(bf_workspace_run_device_frame(r->ws, &vin, &vw, &vout,
                                                     int8 ? &vpar : nullptr, &dt));
        t.kernel_ms = dt.kernel_ms;
        t.temporal_integration_ms = dt.temporal_integration_ms;
        t.quantization_ms = dt.quantization_ms;
        t.device_to_device_ms = dt.device_to_device_ms;
        const auto d0 = Clock::now();
        if (int8) {
            hip_check(hipMemcpy(codes, dout_q.get(), n_out, hipMemcpyDeviceToHost),
                      "hipMemcpy offline int8 output device to host");
            hip_check(hipMemcpy(parameters, dpar.get(), 2 * n_par * sizeof(float),
                                hipMemcpyDeviceToHost),
                      "hipMemcpy offline quantization parameters device to host");
        } else {
            hip_check(hipMemcpy(float_output, dout_f.get(), n_out * sizeof(float),
                                hipMemcpyDeviceToHost),
                      "hipMemcpy offline float output device to host");
        }
        t.device_to_host_ms =
            std::chrono::duration<double, std::milli>(Clock::now() - d0).count();
        if (output_dims) *output_dims = od;
        if (timings) *timings = t;
    });
}

// ------------------------------------------------------------ one-shot
bf_status bf_beamform_packed_intensity(const uint8_t* packed, size_t packed_bytes,
                                       const float* weights, size_t weight_count,
                                       const bf_dims* dims, int kernel, float* out,
                                       size_t out_count, bf_timings* timings) {
    return guarded([&] {
        validate_kernel(kernel);
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        if (packed == nullptr || packed_bytes != d.n_time * d.n_freq * d.n_ant)
            invalid("packed voltage size does not match dimensions");
        if (weights == nullptr || weight_count != weight_storage_count(d, kernel))
            invalid("weight count does not match selected kernel layout");
        if (out == nullptr || out_count < d.n_time * d.n_freq * d.n_beams)
            invalid("intensity output is smaller than dimensions");
        bf_workspace* ws = nullptr;
        if (bf_workspace_create(&d, kernel, 0, BF_OUTPUT_FLOAT32, &ws) != BF_OK)
            fail(BF_ERR_RUNTIME, g_last_error);
        std::unique_ptr<bf_workspace> owner(ws);
        bf_timings t{};
        t.setup_ms = ws->setup_ms;
        double a = 0, b = 0, k = 0, c = 0;
        auto step = [&](bf_status st) {
            if (st != BF_OK) fail(st, g_last_error);
        };
        step(bf_workspace_upload_packed_voltage(ws, packed, packed_bytes, &d, &a));
        step(bf_workspace_upload_weights(ws, weights, weight_count, &d, &b));
        step(bf_workspace_run_kernel(ws, &d, &k));
        step(bf_workspace_download_intensity(ws, out, out_count, &d, &c));
        t.host_to_device_ms = a + b;
        t.kernel_ms = k;
        t.device_to_host_ms = c;
        if (timings) *timings = t;
    });
}

bf_status bf_beamform_packed_integrated_intensity(const uint8_t* packed, size_t packed_bytes,
                                                  const float* weights, size_t weight_count,
                                                  const bf_dims* dims, size_t S, int kernel,
                                                  float* out, size_t out_count,
                                                  bf_timings* timings) {
    return guarded([&] {
        validate_kernel(kernel);
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        validate_integration(kernel, S);
        if (packed == nullptr || packed_bytes != d.n_time * d.n_freq * d.n_ant)
            invalid("packed voltage size does not match dimensions");
        if (weights == nullptr || weight_count != weight_storage_count(d, kernel))
            invalid("weight count does not match selected kernel layout");
        bf_workspace* ws = nullptr;
        if (bf_workspace_create(&d, kernel, S, BF_OUTPUT_FLOAT32, &ws) != BF_OK)
            fail(BF_ERR_RUNTIME, g_last_error);
        std::unique_ptr<bf_workspace> owner(ws);
        bf_timings t{};
        t.setup_ms = ws->setup_ms;
        double b = 0;
        auto step = [&](bf_status st) {
            if (st != BF_OK) fail(st, g_last_error);
        };
        step(bf_workspace_upload_weights(ws, weights, weight_count, &d, &b));
        bf_timings p{};
        step(bf_workspace_run_integrated_pipeline(ws, packed, packed_bytes, out, out_count, &d,
                                                  &p));
        t.host_to_device_ms = b + p.host_to_device_ms;
        t.kernel_ms = p.kernel_ms;
        t.device_to_host_ms = p.device_to_host_ms;
        if (timings) *timings = t;
    });
}

bf_status bf_beamform_packed_quantized_integrated_intensity(
    const uint8_t* packed, size_t packed_bytes, const float* weights, size_t weight_count,
    const bf_dims* dims, size_t S, int kernel, int8_t* codes, size_t code_count,
    float* parameters, size_t parameter_count, bf_timings* timings) {
    return guarded([&] {
        validate_kernel(kernel);
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        validate_integration(kernel, S);
        if (packed == nullptr || packed_bytes != d.n_time * d.n_freq * d.n_ant)
            invalid("packed voltage size does not match dimensions");
        if (weights == nullptr || weight_count != weight_storage_count(d, kernel))
            invalid("weight count does not match selected kernel layout");
        bf_workspace* ws = nullptr;
        if (bf_workspace_create(&d, kernel, S, BF_OUTPUT_INT8, &ws) != BF_OK)
            fail(BF_ERR_RUNTIME, g_last_error);
        std::unique_ptr<bf_workspace> owner(ws);
        bf_timings t{};
        t.setup_ms = ws->setup_ms;
        double b = 0;
        auto step = [&](bf_status st) {
            if (st != BF_OK) fail(st, g_last_error);
        };
        step(bf_workspace_upload_weights(ws, weights, weight_count, &d, &b));
        bf_timings p{};
        step(bf_workspace_run_quantized_integrated_pipeline(
            ws, packed, packed_bytes, codes, code_count, parameters, parameter_count, &d, &p));
        t.host_to_device_ms = b + p.host_to_device_ms;
        t.kernel_ms = p.kernel_ms;
        t.quantization_ms = p.quantization_ms;
        t.device_to_host_ms = p.device_to_host_ms;
        if (timings) *timings = t;
    });
}

// ------------------------------------------------------------ plumbing
bf_status bf_device_count(int* out) {
    return guarded([&] {
        require_ptr(out, "out");
        int n = 0;
        const hipError_t e = hipGetDeviceCount(&n);
        *out = e == hipSuccess ? n : 0;
    });
}

bf_status bf_device_info(char* name, size_t name_len, size_t* mem, int* cus, int* clock_khz) {
    return guarded([&] {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t p{};
        hip_check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
        if (name && name_len) {
            std::strncpy(name, p.gcnArchName, name_len - 1);
            name[name_len - 1] = '\0';
        }
        if (mem) *mem = p.totalGlobalMem;
        if (cus) *cus = p.multiProcessorCount;
        if (clock_khz) *clock_khz = p.clockRate;
    });
}

bf_status bf_device_malloc(size_t bytes, void** out) {
    return guarded([&] {
        require_ptr(out, "out");
        *out = nullptr;
        hip_check(hipMalloc(out, bytes ? bytes : 1), "hipMalloc");
    });
}

bf_status bf_device_free(void* p) {
    return guarded([&] { hip_check(hipFree(p), "hipFree"); });
}

bf_status bf_memcpy_htod(void* dst, const void* src, size_t n) {
    return guarded([&] { hip_check(hipMemcpy(dst, src, n, hipMemcpyHostToDevice), "hipMemcpy H2D"); });
}

bf_status bf_memcpy_dtoh(void* dst, const void* src, size_t n) {
    return guarded([&] { hip_check(hipMemcpy(dst, src, n, hipMemcpyDeviceToHost), "hipMemcpy D2H"); });
}

bf_status bf_memcpy_dtod(void* dst, const void* src, size_t n) {
    return guarded([&] { hip_check(hipMemcpy(dst, src, n, hipMemcpyDeviceToDevice), "hipMemcpy D2D"); });
}

bf_status bf_memset(void* dst, int value, size_t n) {
    return guarded([&] { hip_check(hipMemset(dst, value, n), "hipMemset"); });
}

bf_status bf_stream_create(void** out) {
    return guarded([&] {
        require_ptr(out, "out");
        hipStream_t s = nullptr;
        hip_check(hipStreamCreate(&s), "hipStreamCreate");
        *out = s;
    });
}

bf_status bf_stream_destroy(void* s) {
    return guarded([&] { hip_check(hipStreamDestroy(static_cast<hipStream_t>(s)), "hipStreamDestroy"); });
}

bf_status bf_stream_synchronize(void* s) {
    return guarded([&] {
        hip_check(hipStreamSynchronize(static_cast<hipStream_t>(s)), "hipStreamSynchronize");
    });
}

bf_status bf_device_synchronize(void) {
    return guarded([&] { hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); });
}

bf_status bf_event_create(void** out) {
    return guarded([&] {
        require_ptr(out, "out");
        hipEvent_t e = nullptr;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        *out = e;
    });
}

bf_status bf_event_destroy(void* e) {
    return guarded([&] { hip_check(hipEventDestroy(static_cast<hipEvent_t>(e)), "hipEventDestroy"); });
}

bf_status bf_event_record(void* e, void* s) {
    return guarded([&] {
        hip_check(hipEventRecord(static_cast<hipEvent_t>(e), static_cast<hipStream_t>(s)),
                  "hipEventRecord");
    });
}

bf_status bf_event_elapsed_ms(void* a, void* b, float* ms) {
    return guarded([&] {
        require_ptr(ms, "ms");
        hip_check(hipEventSynchronize(static_cast<hipEvent_t>(b)), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)),
                  "hipEventElapsedTime");
    });
}

// ------------------------------------------------------------ generators
bf_status bf_generate_uniform_int4(void* stream, uint8_t* out, size_t count, uint64_t seed) {
    return guarded([&] {
        require_ptr(out, "out");
        launch_uniform_int4(static_cast<hipStream_t>(stream), out, count, seed);
    });
}

bf_status bf_generate_point_source_noise(void* stream, uint8_t* out, size_t T, size_t F,
                                         size_t A, const float* pos, const float* freqs,
                                         const float* dir3, float amplitude, float sigma,
                                         uint64_t seed) {
    return guarded([&] {
        require_ptr(out, "out");
        require_ptr(pos, "positions");
        require_ptr(freqs, "frequencies");
        require_ptr(dir3, "direction");
        if (T == 0 || F == 0 || A == 0) invalid("generator dimensions must be positive");
        launch_point_source_noise(static_cast<hipStream_t>(stream), out, T, F, A, pos, freqs,
                                  dir3[0], dir3[1], dir3[2], amplitude, sigma, seed);
    });
}

}  // extern "C"
