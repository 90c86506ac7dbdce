/*
 * bf_capi.h -- the C-ABI of libbeamform_hip.so, the MI355X (gfx950) build of
 * the CHARTS voltage beamformer / dynamic tracker hot path.
 *
 * Plain C: pointers, sizes and an opaque handle type per stateful object.
 * Device pointers are HIP device allocations; `stream` is a hipStream_t
 * passed as void* (NULL = the default stream).  Every entry point returns a
 * bf_status; on failure a thread-local message is available from
 * bf_last_error_message().  The status codes map one-to-one onto the
 * exception types the reference's C++ API throws:
 *
 *   BF_ERR_INVALID_ARGUMENT  std::invalid_argument (contract violations)
 *   BF_ERR_OUT_OF_RANGE      std::out_of_range
 *   BF_ERR_OVERFLOW          std::overflow_error (grid sizes)
 *   BF_ERR_LOGIC             std::logic_error (workspace misuse)
 *   BF_ERR_RUNTIME           std::runtime_error (HIP runtime failures)
 *
 * Validation always happens before anything is launched.  The C++ wrappers
 * in include/beamformer/ rethrow these as the same exception types.
 *
 * Layouts (reference include/beamformer/indexing.hpp:13-66):
 *   packed voltage [t][f][a] u8 (low nibble real, high nibble imag, int4)
 *   weights        [b][f][a] complex f32 (re, im) -- "canonical"
 *   tiled weights  [f][ceil(B/32)][a][32] complex f32, zero padded
 *   intensity      [t][f][b] f32;  integrated [w][f][b] f32
 *   int8 codes     [w][f][b];  params [wt][ft][bt] {offset, scale} f32
 */
#ifndef BF_CAPI_H
#define BF_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */
typedef enum bf_status {
    BF_OK = 0,
    BF_ERR_INVALID_ARGUMENT = 1,
    BF_ERR_OUT_OF_RANGE = 2,
    BF_ERR_OVERFLOW = 3,
    BF_ERR_LOGIC = 4,
    BF_ERR_RUNTIME = 5
} bf_status;

/* Message of the last failing call on this thread ("" if none). */
const char* bf_last_error_message(void);
/* Library version string. */
const char* bf_version(void);

/* ------------------------------------------------------- dimensions */
/* Reference: include/beamformer/config.hpp:74-79 (Dimensions). */
typedef struct bf_dims {
    size_t n_time;
    size_t n_freq;
    size_t n_ant;
    size_t n_beams;
} bf_dims;

/* Reference: include/beamformer/config.hpp:81-94 (validate_dimensions):
 * n_time > 0, n_freq == 336, n_ant in {32,64,128,256}, 1 <= n_beams <= 128. */
bf_status bf_validate_dimensions(const bf_dims* dims);

/* Reference: src/temporal_integration.cpp:16-26 (integrated_time_count). */
bf_status bf_integrated_time_count(size_t n_time, size_t integration_spectra,
                                   size_t* out);
/* Reference: src/quantization.cpp:34-47 (quantization_parameter_count). */
bf_status bf_quantization_parameter_count(const bf_dims* integrated_dims, size_t* out);

/* ----------------------------------------------------- kernel choice */
/* Reference: include/beamformer/cuda_stage_api.hpp:14-17
 * (CudaBeamformerKernel {Direct, Tiled}).  DIRECT reads canonical weights,
 * TILED the tiled layout; both run the plain FP32 fused kernel and keep the
 * reference's integration rules (Direct: 320 only; Tiled: 10 or 320).
 * The remaining selectors are additions of this build; all take canonical
 * weights and accept any integration_spectra >= 1:
 *   FP32       plain FP32 fused kernel (VALU FMA)
 *   SPLIT_F16  FP16 MFMA, weights split hi+lo, FP32 accumulate (~FP32 accuracy)
 *   F16        FP16 MFMA, single FP16 weights (~1e-3 relative)
 *   AUTO       SPLIT_F16 where supported, else FP32 */
typedef enum bf_kernel {
    BF_KERNEL_DIRECT = 0,
    BF_KERNEL_TILED = 1,
    BF_KERNEL_FP32 = 2,
    BF_KERNEL_SPLIT_F16 = 3,
    BF_KERNEL_F16 = 4,
    BF_KERNEL_AUTO = 5
} bf_kernel;

/* ---------------------------------------------------- stage launchers */
/* Reference: include/beamformer/cuda_stage_api.hpp:19-24
 * (launch_packed_beamformer).  Caller-owned device buffers; asynchronous on
 * `stream`; allocates nothing and does not synchronize.  `weights` points at
 * interleaved complex f32 in the layout the kernel selector names. */
bf_status bf_launch_packed_beamformer(int kernel, void* stream,
                                      const uint8_t* packed_voltage,
                                      const float* weights, float* intensity,
                                      const bf_dims* dims);

/* Reference: include/beamformer/cuda_stage_api.hpp:26-34
 * (launch_packed_integrated_beamformer): fused beamform + |.|^2 + temporal
 * integration writing [W][F][B] directly. */
bf_status bf_launch_packed_integrated_beamformer(int kernel, void* stream,
                                                 const uint8_t* packed_voltage,
                                                 const float* weights,
                                                 float* integrated_intensity,
                                                 const bf_dims* dims,
                                                 size_t integration_spectra);

/* Reference: include/beamformer/cuda_stage_api.hpp:36-42
 * (launch_quantize_integrated_intensity): 1x16x16 chunk int8 codes; codes are
 * bit-identical to the reference CPU quantiser applied to the same floats.
 * `parameters` holds {offset, scale} pairs. */
bf_status bf_launch_quantize_integrated_intensity(void* stream,
                                                  const float* integrated_intensity,
                                                  int8_t* quantized_intensity,
                                                  float* parameters,
                                                  const bf_dims* integrated_dims);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): beamform + integrate + int8-quantise device-resident
 * voltages straight into caller-owned codes [W][F][B] and {offset, scale}
 * parameters, with the f32 integrated intermediate in a file-scope device
 * array of BF_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends).
 * Async on `stream`; calls no hipMalloc/hipFree.  Results equal
 * bf_launch_packed_integrated_beamformer followed by
 * bf_launch_quantize_integrated_intensity (and the workspace's int8 device
 * frame) bit for bit.  Rejects, before launch, problems whose intermediate
 * W*F*B*4 bytes exceeds the scratch (BF_ERR_INVALID_ARGUMENT).  Every call
 * on one GPU shares that GPU's scratch: calls that overlap on one GPU (on
 * different streams, or from different threads) race. */
#define BF_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
bf_status bf_launch_packed_quantized_integrated_beamformer_noalloc(
    int kernel, void* stream, const uint8_t* packed_voltage, const float* weights,
    int8_t* quantized_intensity, float* parameters, const bf_dims* dims,
    size_t integration_spectra);

/* Which kernel AUTO resolves to for these dimensions (a bf_kernel value). */
/* The stage launchers with a loss mask: one validity byte per [t][f] frame
 * (device memory, 0 = lost; include/beamformer/formats.hpp:22-32).  A lost
 * frame's spectrum contributes nothing: its per-spectrum power is 0 and
 * integrated windows sum the valid spectra only.  The payload is read, never
 * rewritten.  The reference defines the mask but never applies it; no
 * reference launcher corresponds.  MFMA kernels only (SPLIT_F16, F16, or
 * AUTO where it selects one); other selectors return INVALID_ARGUMENT. */
bf_status bf_launch_packed_beamformer_masked(int kernel, void* stream, const uint8_t* packed,
                                             const uint8_t* loss_mask, const float* weights,
                                             float* intensity, const bf_dims* dims);
bf_status bf_launch_packed_integrated_beamformer_masked(int kernel, void* stream,
                                                        const uint8_t* packed,
                                                        const uint8_t* loss_mask,
                                                        const float* weights, float* integrated,
                                                        const bf_dims* dims,
                                                        size_t integration_spectra);

bf_status bf_resolve_kernel(int kernel, const bf_dims* dims, size_t integration_spectra,
                            int* resolved);

/* Autotuner of the MFMA kernels (addition of this build; the reference
 * picks its CUDA block shapes statically).  Times every instantiated tile
 * shape -- beams per wave, waves per workgroup, single or double LDS stage,
 * chunk sizing -- on the caller's device buffers (the layouts of
 * bf_launch_packed_integrated_beamformer; integration_spectra 0 = the
 * per-spectrum launcher), `repeats` timed launches each, and makes the
 * fastest the shape of every later launch of the same problem (T, F, A, B,
 * S, kernel) in this process.  Synchronous; allocates nothing on the
 * device; leaves a valid result in `output` (every shape computes the same
 * sums).  entries may be NULL; n_entries receives the candidate count. */
typedef struct bf_tune_entry {
    int beams_per_wave;
    int waves_per_workgroup;
    int double_buffered;
    int resident_workgroups;
    double median_ms;
} bf_tune_entry;

bf_status bf_autotune_packed_beamformer(int kernel, void* stream, const uint8_t* packed_voltage,
                                        const float* weights, float* output,
                                        const bf_dims* dims, size_t integration_spectra,
                                        int repeats, bf_tune_entry* entries,
                                        size_t max_entries, size_t* n_entries);
/* Forgets every tuned shape (later launches use the defaults). */
bf_status bf_autotune_clear(void);

/* ------------------------------------------------------- timings */
/* Reference: include/beamformer/cuda_beamformer.hpp:22-31. */
typedef struct bf_timings {
    double setup_ms;
    double host_to_device_ms;
    double kernel_ms;
    double temporal_integration_ms;
    double quantization_ms;
    double device_to_device_ms;
    double device_to_host_ms;
} bf_timings;

/* ------------------------------------------------------- workspace */
/* Reference: include/beamformer/cuda_beamformer.hpp:44-92
 * (CudaBeamformerWorkspace).  Owns device buffers sized at creation, one
 * stream and its events.  Methods are synchronous and report ms.  Request
 * dims may shrink n_time and n_beams but not n_freq or n_ant.  Not
 * thread-safe: one workspace per host thread. */
typedef struct bf_workspace bf_workspace;

typedef enum bf_output {
    BF_OUTPUT_FLOAT32 = 0,
    BF_OUTPUT_INT8 = 1
} bf_output;

/* integration_spectra == 0 means "no temporal integration". */
bf_status bf_workspace_create(const bf_dims* capacity, int kernel,
                              size_t integration_spectra, int output,
                              bf_workspace** out);
bf_status bf_workspace_destroy(bf_workspace* ws);
bf_status bf_workspace_setup_ms(const bf_workspace* ws, double* out);
bf_status bf_workspace_kernel(const bf_workspace* ws, int* out);
bf_status bf_workspace_has_temporal_integration(const bf_workspace* ws, int* out);
bf_status bf_workspace_has_int8_quantization(const bf_workspace* ws, int* out);
bf_status bf_workspace_packed_voltage_capacity_bytes(const bf_workspace* ws, size_t* out);
bf_status bf_workspace_upload_packed_voltage(bf_workspace* ws, const uint8_t* host,
                                             size_t host_bytes, const bf_dims* dims,
                                             double* ms);
/* weight_count counts complex values in the kernel's layout. */
bf_status bf_workspace_upload_weights(bf_workspace* ws, const float* host,
                                      size_t weight_count, const bf_dims* dims,
                                      double* ms);
bf_status bf_workspace_run_kernel(bf_workspace* ws, const bf_dims* dims, double* ms);
bf_status bf_workspace_download_intensity(bf_workspace* ws, float* host,
                                          size_t host_count, const bf_dims* dims,
                                          double* ms);
bf_status bf_workspace_run_pipeline(bf_workspace* ws, const uint8_t* packed,
                                    size_t packed_bytes, float* intensity,
                                    size_t intensity_count, const bf_dims* dims,
                                    bf_timings* timings);
bf_status bf_workspace_run_integrated_pipeline(bf_workspace* ws, const uint8_t* packed,
                                               size_t packed_bytes, float* integrated,
                                               size_t integrated_count,
                                               const bf_dims* dims, bf_timings* timings);
bf_status bf_workspace_run_quantized_integrated_pipeline(
    bf_workspace* ws, const uint8_t* packed, size_t packed_bytes, int8_t* codes,
    size_t code_count, float* parameters, size_t parameter_count, const bf_dims* dims,
    bf_timings* timings);

/* ------------------------------------------------ device frame views */
/* Reference: include/beamformer/cuda_frame.hpp:16-60 (CudaDataType,
 * CudaBufferView, CudaFrameMetadata, CudaFrameView) and
 * include/beamformer/config.hpp:29-41 (ShardDescriptor). */
typedef enum bf_data_type {
    BF_DTYPE_PACKED_INT4X2 = 0,
    BF_DTYPE_COMPLEX_FLOAT32 = 1,
    BF_DTYPE_FLOAT32 = 2,
    BF_DTYPE_SIGNED_INT8 = 3
} bf_data_type;

typedef struct bf_shard {
    size_t shard_id;
    size_t shard_count;
    size_t local_frequency_count;
    size_t absolute_frequency_start;
    uint64_t timestamp_start;
    uint64_t timestamp_step;
    uint64_t loss_mask_id;
    int loss_mask_independent;
} bf_shard;

typedef struct bf_frame_view {
    void* device_data;
    size_t bytes;
    int type; /* bf_data_type */
    size_t rank;
    size_t shape[4];
    size_t strides_bytes[4];
    uint64_t frame_id;
    bf_shard shard;
} bf_frame_view;

/* Reference: src/cuda_beamformer.cu:930-1079 (run_device_frame): input and
 * weights are caller device buffers, the final output is copied device to
 * device into `output`; `quantization_parameters` is required for int8
 * output and must be NULL otherwise. */
bf_status bf_workspace_run_device_frame(bf_workspace* ws, const bf_frame_view* voltage,
                                        const bf_frame_view* weights,
                                        bf_frame_view* output,
                                        bf_frame_view* quantization_parameters,
                                        bf_timings* timings);

/* Reference: CudaOfflineFrameRunner (include/beamformer/cuda_offline_runner.hpp:
 * 17-52, src/cuda_offline_runner.cu:130-228): one offline frame from host
 * buffers.  Each run makes temporary device copies of the caller's voltages
 * and weights, runs bf_workspace_run_device_frame on its own workspace and
 * copies the result back: float output [T or W][F][B] (FLOAT32 runners) or
 * int8 codes and {offset, scale} parameters (INT8 runners, which need
 * integration).  Frame dims may shrink n_time and n_beams below the
 * capacity, not change n_freq or n_ant; `weight_count` counts complex
 * values in the kernel's layout (tiled for TILED).  `output_dims` (may be
 * NULL) receives the output's dims.  Synchronous; not thread-safe. */
typedef struct bf_offline_frame_runner bf_offline_frame_runner;
bf_status bf_offline_frame_runner_create(const bf_dims* capacity, int kernel,
                                         size_t integration_spectra, int output,
                                         bf_offline_frame_runner** out);
bf_status bf_offline_frame_runner_destroy(bf_offline_frame_runner* runner);
bf_status bf_offline_frame_runner_config(const bf_offline_frame_runner* runner, int* kernel,
                                         int* output, int* has_temporal_integration);
bf_status bf_offline_frame_runner_run(bf_offline_frame_runner* runner, const uint8_t* packed,
                                      size_t packed_bytes, const float* weights,
                                      size_t weight_count, const bf_dims* dims,
                                      uint64_t frame_id, const bf_shard* shard,
                                      float* float_output, size_t float_count,
                                      int8_t* codes, size_t code_count, float* parameters,
                                      size_t parameter_count, bf_dims* output_dims,
                                      bf_timings* timings);

/* --------------------------------------------------- one-shot API */
/* Reference: include/beamformer/cuda_beamformer.hpp:94-123 and
 * src/cuda_beamformer.cu:1081-1173: allocate-and-run on host buffers.
 * `timings` may be NULL. */
bf_status bf_beamform_packed_intensity(const uint8_t* packed, size_t packed_bytes,
                                       const float* weights, size_t weight_count,
                                       const bf_dims* dims, int kernel, float* out,
                                       size_t out_count, bf_timings* timings);
bf_status bf_beamform_packed_integrated_intensity(const uint8_t* packed,
                                                  size_t packed_bytes,
                                                  const float* weights,
                                                  size_t weight_count,
                                                  const bf_dims* dims,
                                                  size_t integration_spectra, int kernel,
                                                  float* out, size_t out_count,
                                                  bf_timings* timings);
bf_status bf_beamform_packed_quantized_integrated_intensity(
    const uint8_t* packed, size_t packed_bytes, const float* weights, size_t weight_count,
    const bf_dims* dims, size_t integration_spectra, int kernel, int8_t* codes,
    size_t code_count, float* parameters, size_t parameter_count, bf_timings* timings);

/* ------------------------------------------------------- dynamic tracker */
/* Reference: include/beamformer/beam_tracker.hpp:33-57 (TrackerConfig).
 * The trajectory advances the direction cosines linearly per spectrum:
 * l = l0 + t * rate_l, m = m0 + t * rate_m, n = sqrt(1 - l^2 - m^2), and a
 * window uses the direction of its first spectrum. */
typedef struct bf_tracker_config {
    float direction_start[3];
    float direction_rate_per_sample[2];
    size_t integration_spectra;
} bf_tracker_config;

/* Reference: include/beamformer/cuda_beam_tracker_v5.hpp:45-49
 * (V5ExecutionConfig).  `kernel` picks FP32, SPLIT_F16, F16 or AUTO
 * (= SPLIT_F16); `enable_graph` captures the batched stream's launch in a
 * hipGraph. */
typedef struct bf_tracker_exec {
    int kernel;
    int enable_graph;
} bf_tracker_exec;

/* Reference: cuda_beam_tracker_v5_device_resident
 * (include/beamformer/cuda_beam_tracker_v5.hpp:80-85): single-beam tracker
 * on device buffers, per-spectrum [T][F][1] output, the reference's
 * hard-coded geometry (default_positions at 0.6 m, channels at 300 MHz +
 * k * 300 kHz).  Validation (n_beams == 1, trajectory on the unit disk for
 * every window) happens before launch; synchronises only if stream is NULL.
 * The 16-byte trajectory row travels in a stream-ordered allocation on
 * `stream` (hipMallocAsync / hipFreeAsync); nothing else is allocated.
 * `exec` may be NULL (AUTO: the FP32 stream kernel for this one-beam,
 * per-spectrum case). */
bf_status bf_beam_tracker_device_resident(const uint8_t* d_packed, float* d_intensity,
                                          const bf_dims* dims,
                                          const bf_tracker_config* tracker,
                                          const bf_tracker_exec* exec, void* stream);

/* Reference: cuda_beam_tracker_v5_into (cuda_beam_tracker_v5.hpp:74-78):
 * host buffers in and out, synchronous. */
bf_status bf_beam_tracker_into(const uint8_t* packed, size_t packed_bytes,
                               const bf_dims* dims, const bf_tracker_config* tracker,
                               float* intensity, size_t intensity_count,
                               const bf_tracker_exec* exec);

/* Reference: BatchedTrackerStreamV5 (cuda_beam_tracker_v5.hpp:90-125).
 * `single_window_dims.n_time` is the window length; a batch holds
 * batch_size windows.  process_batch honours first_window_index: the batch's
 * windows are the absolute windows first_window_index .. +batch_size-1 of
 * the trajectory (the reference ignores the index). */
typedef struct bf_tracker_stream bf_tracker_stream;
bf_status bf_tracker_stream_create(const bf_dims* single_window_dims,
                                   const bf_tracker_config* tracker, size_t batch_size,
                                   const bf_tracker_exec* exec, bf_tracker_stream** out);
bf_status bf_tracker_stream_destroy(bf_tracker_stream* s);
bf_status bf_tracker_stream_process_batch(bf_tracker_stream* s, size_t first_window_index,
                                          const uint8_t* host_packed, float* host_intensity);
bf_status bf_tracker_stream_process_batch_kernel_only(bf_tracker_stream* s,
                                                      size_t first_window_index);
bf_status bf_tracker_stream_last_kernel_time_ms(const bf_tracker_stream* s, float* out);
bf_status bf_tracker_stream_device_packed_buffer(bf_tracker_stream* s, uint8_t** out);
bf_status bf_tracker_stream_device_intensity_buffer(bf_tracker_stream* s, float** out);
bf_status bf_tracker_stream_device_stream(bf_tracker_stream* s, void** out);
bf_status bf_tracker_stream_sizes(const bf_tracker_stream* s, size_t* batch_size,
                                  size_t* window_bytes, size_t* batch_voltage_bytes,
                                  size_t* batch_output_floats);

/* Multi-beam tracker (addition of this build; the reference tracker is
 * single-beam).  Weights are regenerated per window on the device from the
 * direction of every beam: either a device table [n_windows][n_beams][3] of
 * unit vectors, or per-beam linear trajectories [n_beams][4] =
 * (l0, m0, rate_l, rate_m) evaluated at absolute spectrum
 * (window_offset + w) * integration_spectra.  Positions [n_ant][3] and
 * frequencies [n_freq] are device tables; NULL selects the reference's
 * default geometry (0.6 m grid, 300 MHz + k * 300 kHz).  Output is
 * integrated [W][F][B] (integrate != 0) or per spectrum [T][F][B].
 * n_freq may be any positive count (e.g. a full 672-channel band);
 * n_ant in {32, 64, 128, 256}; 1 <= n_beams <= 128. */
typedef struct bf_multibeam_args {
    size_t n_time;
    size_t n_freq;
    size_t n_ant;
    size_t n_beams;
    size_t integration_spectra;
    const float* positions;     /* device [n_ant][3] or NULL */
    const float* frequencies;   /* device [n_freq] or NULL */
    const float* directions;    /* device [W][n_beams][3] or NULL */
    const float* trajectories;  /* device [n_beams][4] or NULL */
    size_t window_offset;
    int integrate;
    int kernel;                 /* FP32, SPLIT_F16, F16 or AUTO */
} bf_multibeam_args;

bf_status bf_multibeam_tracker_launch(void* stream, const uint8_t* d_packed, float* d_out,
                                      const bf_multibeam_args* args);

/* ---------------------------------------------------- adaptive DOA */
/* Adaptive direction finding on a uniform linear array (addition of this
 * build; it restates the primitives of the reference's CPU DOA tracker,
 * src/cpu_opt_beam_tracker.cpp, on the GPU).  Per update of `snapshots`
 * spectra and per channel: exact integer covariance R = sum x x^H (int8
 * MFMA, int32 accumulate), forward spatial smoothing over sub-arrays of
 * `subarray` elements, R_block = R_smoothed / (K L), exponential fold
 * R = forgetting * R + (1 - forgetting) * R_block (the first update after
 * create/reset copies), Bartlett a^H R a, and Capon 1 / (a^H R_L^-1 a) on
 * the diagonally loaded R_L = R + diagonal_load * tr(R) / M * I (FP64),
 * which equals the reference's recursion over R_block loaded by
 * diagonal_load * tr(R_block) / M before its fold (Capon only,
 * src/cpu_opt_beam_tracker.cpp:1879-1892); scans over n_theta angles uniformly spaced in
 * [theta_min, theta_max] with a_m = exp(-j 2 pi f / c * m * spacing * sin theta).
 * Broadband spectra sum the channels; peaks are local maxima refined with a
 * 3-point parabola.  n_ant in {32, 64}; 1 <= subarray <= n_ant;
 * snapshots <= 65536. */
typedef struct bf_doa_config {
    size_t n_freq;
    size_t n_ant;
    size_t snapshots;
    size_t subarray;
    size_t n_theta;
    size_t max_peaks;
    double forgetting;
    double diagonal_load;
    double theta_min_rad;
    double theta_max_rad;
    double element_spacing_m;
    const float* frequencies; /* host [n_freq], copied at create */
} bf_doa_config;

typedef struct bf_doa bf_doa;
bf_status bf_doa_create(const bf_doa_config* config, bf_doa** out);
bf_status bf_doa_destroy(bf_doa* doa);
/* The next update starts the recursion afresh (copies R_block). */
bf_status bf_doa_reset(bf_doa* doa);
/* Enqueues one update on `stream` over device voltages [n_time][F][A],
 * snapshots t0 .. t0 + snapshots - 1. */
bf_status bf_doa_update(bf_doa* doa, void* stream, const uint8_t* d_packed, size_t n_time,
                        size_t t0);
/* Synchronous reads of the last update.  Any pointer may be NULL.
 * broadband: [2][n_theta] (Capon, Bartlett); peaks: [2][max_peaks][2]
 * (theta_rad, power); n_peaks: [2]. */
bf_status bf_doa_read_results(bf_doa* doa, double* broadband, double* peaks, int* n_peaks);
/* Per-channel spectra [F][n_theta] each, Cholesky success flags [F]. */
bf_status bf_doa_read_channel_spectra(bf_doa* doa, double* capon, double* bartlett,
                                      int* chol_ok);
/* Exact smoothed covariance of the last update, [F][M][M][2] int32. */
bf_status bf_doa_read_covariance(bf_doa* doa, int32_t* rs);
/* Device time of the last update's kernels (waits for it).  Reference
 * analogue: CudaBeamformerTimings (cuda_beamformer.hpp:22-31). */
typedef struct bf_doa_timings {
    double covariance_ms;
    double fold_scan_ms;
    double estimate_ms;
} bf_doa_timings;
bf_status bf_doa_last_timings(bf_doa* doa, bf_doa_timings* timings);
/* Checkpoint / resume of the recursion: the unloaded R [F][M][M][2] FP64
 * and whether an update has been folded. */
bf_status bf_doa_get_state(bf_doa* doa, double* state, int* initialised);
bf_status bf_doa_set_state(bf_doa* doa, const double* state, int initialised);

/* ------------------------------------ geometry, beam grids, weights */
/* Producers of the beamformer's weight inputs (src/geometry.cpp:33-451,
 * src/weights.cpp:11-122), same arithmetic and messages.  Positions and
 * directions are [n][3] float, (u, v) centres [n][2]; weights are
 * interleaved complex f32. */
/* regular_array (geometry.cpp:33-51): x = column * spacing, y = row * spacing. */
bf_status bf_regular_array(size_t rows, size_t columns, float spacing_m, float* out);
/* default_positions (geometry.cpp:53-67): 4x8, 8x8, 8x16, 16x16 grids. */
bf_status bf_default_positions(size_t n_ant, float spacing_m, float* out);
/* constant_frequencies / channelized_frequencies (geometry.cpp:69-95). */
bf_status bf_constant_frequencies(size_t n_freq, float frequency_hz, float* out);
bf_status bf_channelized_frequencies(size_t n_freq, float start_hz, float channel_width_hz,
                                     float* out);
/* direction_from_lm (geometry.cpp:97-106). */
bf_status bf_direction_from_lm(float l, float m, float* out);
/* default_beam_grid, the line of n_beams beams at l_step (geometry.cpp:108-125). */
bf_status bf_default_beam_grid(size_t n_beams, float l_step, float m, float* out);
/* centered_integer_range (geometry.cpp:127-137). */
bf_status bf_centered_integer_range(size_t count, long long* out);
/* select_fft_beam_centers_rectangular (geometry.cpp:139-215): centers
 * [n_beams][2]; shape = {n_u, n_v, n_bank_u, n_bank_v}. */
bf_status bf_select_fft_beam_centers_rectangular(float du_fft_u, float dv_fft_v,
                                                 size_t n_beams, size_t m_side, size_t n_side,
                                                 float* centers, size_t* shape);
/* fft_beam_grid, the FFT-bin grid (geometry.cpp:217-249). */
bf_status bf_fft_beam_grid(size_t n_ant, size_t n_beams, float spacing_m,
                           float design_frequency_hz, float* out);
/* estimate_nbeams_hex_formula (geometry.cpp:251-273). */
bf_status bf_estimate_nbeams_hex_formula(float spacing_m, size_t m_side, size_t n_side,
                                         float bw_e_deg, float bw_h_deg, float wavelength_m,
                                         size_t* out);
/* generate_hex_targets_cropped_fov, the hexagonal grid (geometry.cpp:275-330):
 * up to n_beams (u, v) targets in *n_out; out may be NULL when n_beams = 0. */
bf_status bf_generate_hex_targets_cropped_fov(size_t n_beams, float u_max, float v_max,
                                              float* out, size_t* n_out);
/* rectangular_beam_grid (geometry.cpp:332-366): n_ant beams. */
bf_status bf_rectangular_beam_grid(size_t n_ant, float spacing_m, float design_frequency_hz,
                                   float* out);
/* load_positions / load_frequencies / load_directions (geometry.cpp:368-451). */
bf_status bf_load_positions(const char* path, size_t expected_count, float* out);
bf_status bf_load_frequencies(const char* path, size_t expected_count, float* out);
bf_status bf_load_directions(const char* path, size_t expected_count, float* out);
/* generate_weights [B][F][A] / generate_tiled_weights [F][ceil(B/32)][A][32]
 * on the host (weights.cpp:11-122). */
bf_status bf_generate_weights(const bf_dims* dims, const float* positions,
                              const float* frequencies, const float* directions, float* out,
                              size_t n_complex);
bf_status bf_generate_tiled_weights(const bf_dims* dims, const float* positions,
                                    const float* frequencies, const float* directions,
                                    float* out, size_t n_complex);
/* The same weights generated on the device into d_weights (canonical, or
 * tiled when `tiled`), asynchronous on `stream`; inputs are device arrays
 * and are not validated here (the host forms validate). */
bf_status bf_generate_weights_device(void* stream, const bf_dims* dims,
                                     const float* d_positions, const float* d_frequencies,
                                     const float* d_directions, float* d_weights, int tiled);

/* ---------------------------------------- adaptive planar DOA tracker */
/* GPU counterpart of the reference's CpuOptBeamTracker
 * (include/beamformer/cpu_opt_beam_tracker.hpp:39-196,
 * src/cpu_opt_beam_tracker.cpp:1084-2494): per integration window, the
 * per-channel covariance recursion, a G x G coarse (l, m) scan around the
 * previous estimate summed over channels, refinement_levels 3 x 3
 * halvings around the running best, quadratic peak interpolation, and the
 * window's intensity [T][F][1] along the estimate.  With
 * coarse_grid_resolution <= 1 the scan is off and each window follows the
 * seeded trajectory (the open-loop tracker).  The covariance state
 * persists across runs (reset clears it); the first window of every run is
 * searched around the trajectory's window-0 direction.  A = 32 or 64;
 * coarse_grid_resolution <= 64. */
enum bf_estimator { BF_ESTIMATOR_BARTLETT = 0, BF_ESTIMATOR_CAPON = 1 };

/* Reference: CpuOptTrackerConfig (cpu_opt_beam_tracker.hpp:50-88); `kernel`
 * (FP32, SPLIT_F16, F16 or AUTO = SPLIT_F16) computes the emitted
 * intensities. */
typedef struct bf_opt_tracker_config {
    int estimator;
    size_t coarse_grid_resolution;
    size_t refinement_levels;
    float search_fov_l;
    float search_fov_m;
    size_t spatial_smoothing_subarray_size;
    float forgetting_factor;
    float diagonal_load;
    int enable_quadratic_peak_interp;
    size_t integration_spectra;
    int kernel;
} bf_opt_tracker_config;

typedef struct bf_opt_tracker bf_opt_tracker;
/* The reference's default-constructed CpuOptTrackerConfig. */
bf_status bf_opt_tracker_default_config(bf_opt_tracker_config* config);
/* Reference: CpuOptBeamTracker::CpuOptBeamTracker (.cpp:1084-1320).
 * positions [n_ant][3] m and frequencies [n_freq] Hz are host arrays,
 * copied; NULL selects default_positions / channelized_frequencies. */
bf_status bf_opt_tracker_create(const float* positions, const float* frequencies,
                                const bf_dims* dims, const bf_opt_tracker_config* config,
                                bf_opt_tracker** out);
bf_status bf_opt_tracker_destroy(bf_opt_tracker* tracker);
/* Reference: seed_trajectory (.cpp:1335-1357). */
bf_status bf_opt_tracker_seed_trajectory(bf_opt_tracker* tracker, const float* direction_start,
                                         const float* direction_rate_per_sample);
/* Clears the covariance recursion (the next window starts it afresh). */
bf_status bf_opt_tracker_reset(bf_opt_tracker* tracker);
/* Reference: run_into (.cpp:1619-2457) on device buffers, packed
 * [T][F][A] -> intensity [T][F][1]; asynchronous on `stream`. */
bf_status bf_opt_tracker_run_device(bf_opt_tracker* tracker, void* stream,
                                    const uint8_t* d_packed, float* d_intensity);
/* run_into on host buffers (size checks of .cpp:1626-1633). */
bf_status bf_opt_tracker_run(bf_opt_tracker* tracker, const uint8_t* packed,
                             size_t packed_count, float* intensity, size_t intensity_count);
/* Reference: window_direction (.cpp:1328-1333) for every window of the last
 * run, [n][3]; dirs may be NULL to query the count. */
bf_status bf_opt_tracker_window_directions(bf_opt_tracker* tracker, float* dirs,
                                           size_t capacity, size_t* n_windows);
/* Reference: cpu_opt_beam_tracker_packed_intensity_into, stateless
 * (.cpp:2471-2486): default geometry, a fresh recursion, the trajectory's
 * integration_spectra. */
bf_status bf_opt_tracker_packed_intensity(const uint8_t* packed, size_t packed_count,
                                          const bf_dims* dims,
                                          const bf_tracker_config* trajectory,
                                          const bf_opt_tracker_config* config, float* intensity,
                                          size_t intensity_count);

/* ------------------------------------------- streaming tracker server */
/* Continuous-batching server for tracking requests over one voltage stream
 * (addition of this build; a superset of the reference's
 * BatchedTrackerStreamV5, include/beamformer/cuda_beam_tracker_v5.hpp:90-125,
 * which runs fixed batches of one request synchronously).
 *
 * Voltages [n][F][A] are ingested into a device ring of ring_spectra
 * spectra (spectrum t at row t mod ring_spectra).  A request tracks
 * n_beams beams along linear trajectories [n_beams][4] =
 * (l0, m0, rate_l, rate_m), evaluated at each window's first spectrum
 * counted from the request's start, over n_spectra spectra in windows of
 * `window` spectra (the last may be partial), and produces integrated
 * power [window][F][n_beams] f32 -- the multi-beam tracker's output for the
 * same span -- into an output ring of output_windows windows in the
 * server's HBM arena.  Requests wait in first-in first-out order until a
 * request slot and arena memory are free (admission by HBM budget).  Each
 * step processes, for every admitted request, the windows whose last
 * spectrum was ingested since the previous step, in one captured launch
 * (hipGraph when enable_graph).  Per-window latency runs from the ingest
 * copy that landed a window's last spectrum to the end of the step that
 * produced it (device events).  One host thread drives a server; calls
 * that overlap on one server race.
 *
 * With loss_masks set, the server keeps one validity byte per frame
 * (include/beamformer/formats.hpp:22-32: [t][F], 0 = lost) in a second
 * ring [ring_spectra][F]; a lost frame contributes nothing to any window
 * (bit for bit the result of zeroing its voltages).  Spectra may also be
 * written in place by the caller's producer (bf_server_acquire /
 * bf_server_commit) instead of copied in by bf_server_ingest. */
typedef struct bf_server bf_server;

typedef struct bf_server_config {
    size_t n_freq;            /* channels F */
    size_t n_ant;             /* 32 or 64 */
    size_t ring_spectra;      /* power of two */
    size_t max_window;        /* longest window accepted, <= ring_spectra / 2 */
    size_t max_step_spectra;  /* most spectra one step releases */
    size_t hbm_budget_bytes;  /* arena for request outputs */
    size_t max_requests;      /* request slots, 1..256 */
    const float* positions;   /* host [n_ant][3] m, copied; NULL: reference grid */
    const float* frequencies; /* host [n_freq] Hz, copied; NULL: reference plan */
    int kernel;               /* SPLIT_F16, F16 or AUTO (= SPLIT_F16) */
    int enable_graph;
    int loss_masks;           /* 1: keep a validity ring and honour loss masks */
    const bf_shard* shard;    /* the stream's shard (copied); with frequencies NULL its
                                 channels of the reference plan; NULL: shard 0 */
    size_t max_active_beams;  /* compute budget: a request is admitted only while the
                                 beams of the active requests plus its own stay within
                                 it (FIFO, like the HBM budget); 0: no cap */
} bf_server_config;

/* Ring rows of consecutive spectra handed to a producer: voltages
 * [n_spectra][F][A] and, with loss masks, validity bytes [n_spectra][F]
 * (NULL without).  Device pointers. */
typedef struct bf_ring_span {
    long long first_spectrum;
    size_t n_spectra;
    uint8_t* voltages;
    uint8_t* valid;
} bf_ring_span;

typedef struct bf_server_request {
    size_t n_beams;              /* 1..64 */
    size_t window;               /* 1..max_window */
    size_t n_spectra;            /* duration */
    long long start;             /* first spectrum; -1: the head at admission */
    size_t output_windows;       /* output ring depth; 0: every window */
    const float* trajectories;   /* host [n_beams][4], copied */
} bf_server_request;

enum bf_server_request_state {
    BF_REQUEST_PENDING = 0,
    BF_REQUEST_ACTIVE = 1,
    BF_REQUEST_COMPLETE = 2,
    BF_REQUEST_REJECTED = 3 /* its start left the ring before admission */
};

typedef struct bf_request_status {
    int state;
    long long start;          /* first spectrum once admitted */
    size_t windows_total;
    size_t windows_issued;    /* windows handed to a step */
    size_t windows_done;      /* windows whose outputs are complete */
    size_t output_windows;
    size_t output_bytes;
} bf_request_status;

typedef struct bf_server_stats {
    size_t windows_timed;
    double latency_p50_ms;
    double latency_p90_ms;
    double latency_p99_ms;
    double latency_max_ms;
    double latency_mean_ms;
    size_t steps;
    size_t windows_processed;
    double beam_samples;      /* sum over processed windows of spectra * F * B */
    double gpu_busy_ms;       /* sum of step durations */
    size_t requests_submitted;
    size_t requests_admitted;
    size_t requests_completed;
    size_t requests_rejected;
    size_t requests_pending;
    size_t requests_in_slots;
    size_t arena_bytes_in_use;
    size_t arena_peak_bytes;
    long long head;           /* spectra ingested */
    long long processed_head; /* spectra released to steps */
    double mean_admission_wait_spectra;
} bf_server_stats;

bf_status bf_server_create(const bf_server_config* config, bf_server** out);
bf_status bf_server_destroy(bf_server* server);
/* Queues a request; it is admitted by a later bf_server_step. */
bf_status bf_server_submit(bf_server* server, const bf_server_request* request,
                           uint64_t* id);
/* Appends n_spectra spectra [n][F][A] (device memory when src_on_device,
 * else host) to the ring: asynchronous for a device source, complete on
 * return for a host source.  Fails with BF_ERR_LOGIC when it would
 * overwrite spectra of windows no step has processed yet. */
bf_status bf_server_ingest(bf_server* server, const uint8_t* src, size_t n_spectra,
                           int src_on_device);
/* As bf_server_ingest, with the spectra's validity bytes valid [n][F]
 * (same memory kind as src); needs a server with loss_masks.  Spectra
 * ingested without a mask are valid in every channel. */
bf_status bf_server_ingest_masked(bf_server* server, const uint8_t* src, const uint8_t* valid,
                                  size_t n_spectra, int src_on_device);
/* Zero-copy producer interface: hands out the ring rows of the next
 * n_spectra spectra (one span, two when they wrap; spans[2]) for the
 * caller to write on a stream of its own.  Blocks the host until no step
 * still reads those rows; BF_ERR_LOGIC when no step launched so far
 * would release them.  With loss masks the producer writes the validity
 * bytes of every spectrum too.  Until bf_server_commit, no other ingest
 * or acquire is accepted. */
bf_status bf_server_acquire(bf_server* server, size_t n_spectra, bf_ring_span* spans,
                            size_t* n_spans);
/* Publishes the first n_spectra acquired spectra (0 abandons the span):
 * steps wait for the work queued on producer_stream (a hipStream_t; NULL
 * is the null stream) up to this call. */
bf_status bf_server_commit(bf_server* server, size_t n_spectra, void* producer_stream);
/* Admits what fits, then enqueues one step over the spectra ingested since
 * the previous step (at most max_step_spectra); returns the windows it
 * scheduled.  A request admitted with an explicit start behind the
 * processed head has all its windows that the ring already holds scheduled
 * by this step, even when no new spectra were ingested.  Asynchronous. */
bf_status bf_server_step(bf_server* server, size_t* windows_scheduled);
bf_status bf_server_request_status(bf_server* server, uint64_t id,
                                   bf_request_status* status);
/* Copies windows [first, first + n) of a request, [n][F][n_beams] f32;
 * waits for the step that produced them. */
bf_status bf_server_read_output(bf_server* server, uint64_t id, size_t first_window,
                                size_t n_windows, float* host_out);
/* Frees a request's slot and outputs (cancels it if still running). */
bf_status bf_server_release(bf_server* server, uint64_t id);
bf_status bf_server_synchronize(bf_server* server);
bf_status bf_server_get_stats(bf_server* server, bf_server_stats* stats);
bf_status bf_server_reset_stats(bf_server* server);

/* Native load generator: Poisson request arrivals at arrival_rate_hz over
 * run_seconds of stream time, each request with n_beams uniform in
 * [min_beams, max_beams], window window_a or window_b with equal odds and
 * duration uniform in [min_duration_s, max_duration_s]; trajectories start
 * uniform in |l|, |m| <= 0.5 and drift at most 15 degrees per hour.
 * Voltages are a device pool of pool_spectra spectra (one point source
 * plus noise) cycled into the ring chunk_spectra at a time, at the
 * stream's real-time rate when paced (else as fast as the steps run);
 * every chunk is followed by the steps it releases.  Finished requests are
 * released. */
typedef struct bf_server_sim_config {
    bf_server_config server;
    uint64_t seed;
    double cadence_s;
    double run_seconds;
    size_t chunk_spectra;
    double arrival_rate_hz;
    size_t min_beams;
    size_t max_beams;
    double min_duration_s;
    double max_duration_s;
    size_t window_a;
    size_t window_b;
    size_t output_windows_cap;
    int paced;
    size_t pool_spectra;
    float source_amplitude;
    float noise_sigma;
} bf_server_sim_config;

typedef struct bf_server_sim_report {
    bf_server_stats stats;
    double wall_seconds;
    double stream_seconds;
    double beam_samples_per_s;     /* processed beam-samples / wall time */
    double mean_requests_in_slots;
    size_t requests_generated;
    double max_pacing_lag_ms;
    int graph_used;
} bf_server_sim_report;

bf_status bf_server_simulate(const bf_server_sim_config* config, bf_server_sim_report* report);

/* --------------------------------------------------- device plumbing */
/* Minimal HIP plumbing so hosts without a HIP toolchain (ctypes, JNI, cgo)
 * can own device buffers and streams.  Reference: the CUDA runtime calls
 * the reference makes inside src/cuda_beamformer.cu:33-106. */
bf_status bf_device_count(int* out);
bf_status bf_device_info(char* name, size_t name_len, size_t* global_memory_bytes,
                         int* compute_units, int* clock_khz);
bf_status bf_device_malloc(size_t bytes, void** out);
bf_status bf_device_free(void* ptr);
bf_status bf_memcpy_htod(void* dst, const void* src, size_t bytes);
bf_status bf_memcpy_dtoh(void* dst, const void* src, size_t bytes);
bf_status bf_memcpy_dtod(void* dst, const void* src, size_t bytes);
bf_status bf_memset(void* dst, int value, size_t bytes);
bf_status bf_stream_create(void** out);
bf_status bf_stream_destroy(void* stream);
bf_status bf_stream_synchronize(void* stream);
bf_status bf_device_synchronize(void);
/* Timing events on `stream`. */
bf_status bf_event_create(void** out);
bf_status bf_event_destroy(void* event);
bf_status bf_event_record(void* event, void* stream);
bf_status bf_event_elapsed_ms(void* start, void* end, float* ms);

/* ------------------------------------------------ synthetic inputs */
/* Seeded device generators for benchmark-size inputs (no reference
 * counterpart; the reference generates on the host in
 * src/synthetic_data.cpp).  Uniform: both nibbles uniform in [-8, 7] from a
 * splitmix64 hash of (index, seed). */
bf_status bf_generate_uniform_int4(void* stream, uint8_t* out, size_t count,
                                   uint64_t seed);
/* Point source plus complex Gaussian noise, rounded and clipped to int4:
 * x = amplitude * exp(-j k p.d) + N(0, sigma^2) per component. */
bf_status bf_generate_point_source_noise(void* stream, uint8_t* out, size_t n_time,
                                         size_t n_freq, size_t n_ant,
                                         const float* positions_device,
                                         const float* frequencies_device,
                                         const float* direction3, float amplitude,
                                         float sigma, uint64_t seed);

/* --------------------------------------------------- on-disk formats
 * Host-only readers and writers of the reference's files, byte for byte:
 * raw little-endian arrays whose size must match the dimensions exactly,
 * and key=value metadata text.  Reference: include/beamformer/io.hpp,
 * src/io.cpp:60-342, include/beamformer/formats.hpp:13-77.  IO failures
 * (cannot open, wrong file size, missing or unsupported metadata field)
 * return BF_ERR_RUNTIME as the reference throws std::runtime_error; size
 * and contract violations return BF_ERR_INVALID_ARGUMENT. */

/* Reference include/beamformer/quantization.hpp:75-81
 * (QuantizedOutputMetadata); parameters_file is a NUL-terminated name. */
typedef struct bf_quantized_output_metadata {
    bf_dims output_dims;
    size_t input_n_time;
    size_t integration_spectra;
    bf_shard shard;
    char parameters_file[1024];
} bf_quantized_output_metadata;

/* default_shard_descriptors()[shard_id], config.hpp:65-72. */
bf_status bf_default_shard_descriptor(size_t shard_id, bf_shard* out);
/* validate_shard_descriptor, config.hpp:43-63. */
bf_status bf_validate_shard_descriptor(const bf_shard* shard);

/* read_packed_voltage / write_packed_voltage, io.cpp:60-72: [t][f][a] bytes. */
bf_status bf_read_packed_voltage(const char* path, const bf_dims* dims, uint8_t* out,
                                 size_t out_bytes);
bf_status bf_write_packed_voltage(const char* path, const uint8_t* voltage, size_t bytes,
                                  const bf_dims* dims);
/* read_weights / write_weights, io.cpp:74-86: [b][f][a] complex f32. */
bf_status bf_read_weights(const char* path, const bf_dims* dims, float* out,
                          size_t out_complex);
bf_status bf_write_weights(const char* path, const float* weights, size_t n_complex,
                           const bf_dims* dims);
/* read_tiled_weights / write_tiled_weights, io.cpp:88-101. */
bf_status bf_read_tiled_weights(const char* path, const bf_dims* dims, float* out,
                                size_t out_complex);
bf_status bf_write_tiled_weights(const char* path, const float* weights, size_t n_complex,
                                 const bf_dims* dims);
/* write_intensities, io.cpp:103-111: [t][f][b] f32. */
bf_status bf_write_intensities(const char* path, const float* intensity, size_t count,
                               const bf_dims* dims);
/* read/write_quantized_intensities, io.cpp:113-129: [w][f][b] int8. */
bf_status bf_read_quantized_intensities(const char* path, const bf_dims* dims, int8_t* out,
                                        size_t out_count);
bf_status bf_write_quantized_intensities(const char* path, const int8_t* codes, size_t count,
                                         const bf_dims* dims);
/* read/write_quantization_parameters, io.cpp:131-146: {offset, scale} f32. */
bf_status bf_read_quantization_parameters(const char* path, const bf_dims* integrated_dims,
                                          float* out, size_t out_params);
bf_status bf_write_quantization_parameters(const char* path, const float* params,
                                           size_t n_params, const bf_dims* integrated_dims);
/* beamformer-quantized-int8-v1 metadata, io.cpp:149-242. */
bf_status bf_write_quantized_output_metadata(const char* path,
                                             const bf_quantized_output_metadata* metadata);
bf_status bf_read_quantized_output_metadata(const char* path,
                                            bf_quantized_output_metadata* out);
/* write_packed_shard / read_packed_shard, io.cpp:244-342: payload bytes,
 * one validity byte (0 or 1) per [t][f] frame, beamformer-packed-shard-v1
 * metadata. */
bf_status bf_write_packed_shard(const char* payload_path, const char* metadata_path,
                                const char* loss_mask_path, const bf_shard* shard,
                                const uint8_t* payload, size_t payload_bytes,
                                const uint8_t* loss_mask, size_t loss_mask_bytes,
                                const bf_dims* dims);
bf_status bf_read_packed_shard(const char* payload_path, const char* metadata_path,
                               const char* loss_mask_path, const bf_dims* dims,
                               bf_shard* shard_out, uint8_t* payload,
                               size_t payload_capacity, uint8_t* loss_mask,
                               size_t loss_mask_capacity);
/* validate_packed_shard / validate_packed_shards (synthetic_data.cpp:139-171):
 * a shard's descriptor, payload size and 0/1 mask; a two-shard set adds slot
 * identity, distinct frequency origins and independent masks (shards[2],
 * payloads[2], loss_masks[2]). */
bf_status bf_validate_packed_shard(const bf_shard* shard, const uint8_t* payload,
                                   size_t payload_bytes, const uint8_t* loss_mask,
                                   size_t loss_mask_bytes, const bf_dims* dims);
bf_status bf_validate_packed_shards(const bf_shard* shards, const uint8_t* const* payloads,
                                    size_t payload_bytes, const uint8_t* const* loss_masks,
                                    size_t loss_mask_bytes, const bf_dims* dims);
/* make_loss_mask (synthetic_data.cpp:121-137): [T][F] bytes, 0 with
 * probability loss_probability, from the same seeded generator. */
bf_status bf_make_loss_mask(const bf_dims* dims, uint32_t seed, float loss_probability,
                            uint8_t* out, size_t n);
/* The shard's channel frequencies: channels absolute_frequency_start ..
 * + 335 of the full-band plan start_hz + k * channel_width_hz. */
bf_status bf_shard_frequencies(const bf_shard* shard, float start_hz, float channel_width_hz,
                               float* out);

#ifdef __cplusplus
}
#endif

#endif /* BF_CAPI_H */
