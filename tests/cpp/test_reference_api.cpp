// test_reference_api.cpp -- the reference-named C++ API
// (include/beamformer/reference_api.hpp) on a GPU: the workspace, the
// one-shot functions, the offline frame runner and the device-frame path
// agree bit for bit on the same inputs; the V2, FWS, V3, V4 and V5 tracker
// entry points and the batched V5 stream give identical results; contract
// violations raise the reference's exception types.  Prints OK on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "beamformer/reference_api.hpp"

namespace bf = beamformer;

static int failures = 0;
#define EXPECT(cond)                                                                 \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
            ++failures;                                                              \
        }                                                                            \
    } while (0)

template <typename Exc, typename Fn>
static bool throws(Fn fn) {
    try {
        fn();
    } catch (const Exc&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

static std::uint32_t lcg(std::uint32_t& s) {
    s = s * 1664525u + 1013904223u;
    return s >> 8;
}

int main() {
    const bf::Dimensions dims{640, 336, 64, 20};
    std::uint32_t seed = 7;
    bf::PackedVoltage x(dims.n_time * dims.n_freq * dims.n_ant);
    for (auto& b : x) b = static_cast<std::uint8_t>(lcg(seed) & 0xFF);
    bf::Weights w(dims.n_beams * dims.n_freq * dims.n_ant);
    for (std::size_t i = 0; i < w.size(); ++i) {
        const double ph = 0.001 * static_cast<double>(lcg(seed) % 6283);
        w[i] = bf::ComplexFloat{static_cast<float>(std::cos(ph)), static_cast<float>(std::sin(ph))};
    }
    const bf::TemporalIntegrationConfig s10 = bf::integration_after_upchan;
    const auto kSplit = bf::CudaBeamformerKernel::SplitF16;

    // workspace pipelines vs one-shot
    bf::CudaBeamformerWorkspace ws(dims, kSplit, s10, bf::CudaBeamformerOutput::QuantizedInt8);
    EXPECT(ws.has_temporal_integration() && ws.has_int8_quantization());
    EXPECT(ws.packed_voltage_capacity_bytes() == x.size());
    ws.upload_weights(w, dims);
    const std::size_t W = bf::integrated_time_count(dims.n_time, 10);
    bf::IntegratedIntensities integ(W * dims.n_freq * dims.n_beams);
    ws.run_integrated_pipeline(x, integ, dims);
    bf::CudaBeamformerTimings t1;
    const auto one = bf::cuda_beamform_packed_integrated_intensity(x, w, dims, s10, &t1, kSplit);
    EXPECT(one == integ);
    EXPECT(t1.kernel_ms > 0.0);
    bf::QuantizedIntegratedOutput q;
    const bf::Dimensions od{W, dims.n_freq, dims.n_ant, dims.n_beams};
    q.codes.resize(W * dims.n_freq * dims.n_beams);
    q.parameters.resize(bf::quantization_parameter_count(od));
    ws.run_quantized_integrated_pipeline(x, q, dims);
    const auto q1 = bf::cuda_beamform_packed_quantized_integrated_intensity(x, w, dims, s10,
                                                                            nullptr, kSplit);
    EXPECT(q1.codes == q.codes);
    EXPECT(q1.parameters.size() == q.parameters.size() &&
           std::memcmp(q1.parameters.data(), q.parameters.data(),
                       q.parameters.size() * sizeof(bf::Int8QuantizationParameters)) == 0);

    // offline frame runner: float and int8, a smaller frame than capacity
    const bf::Dimensions small{320, 336, 64, 12};
    bf::PackedVoltage xs(x.begin(), x.begin() + small.n_time * small.n_freq * small.n_ant);
    bf::Weights ws_(w.begin(), w.begin() + small.n_beams * small.n_freq * small.n_ant);
    bf::ShardDescriptor shard;
    bf::CudaOfflineFrameRunner rf(dims, kSplit, s10);
    const auto r1 = rf.run(xs, ws_, small, 42, shard);
    EXPECT(r1.output_dims.n_time == 32 && r1.output_dims.n_beams == 12);
    EXPECT(r1.float32_output == bf::cuda_beamform_packed_integrated_intensity(
                                    xs, ws_, small, s10, nullptr, kSplit));
    EXPECT(r1.timings.host_to_device_ms >= 0.0 && r1.timings.kernel_ms > 0.0);
    bf::CudaOfflineFrameRunner rq(dims, kSplit, s10, bf::CudaBeamformerOutput::QuantizedInt8);
    const auto r2 = rq.run(xs, ws_, small, 43, shard);
    const auto q2 = bf::cuda_beamform_packed_quantized_integrated_intensity(xs, ws_, small, s10,
                                                                            nullptr, kSplit);
    EXPECT(r2.quantized_output.codes == q2.codes);
    EXPECT(std::memcmp(r2.quantized_output.parameters.data(), q2.parameters.data(),
                       q2.parameters.size() * 8) == 0);
    EXPECT(throws<std::invalid_argument>([&] {
        rf.run(xs, ws_, bf::Dimensions{1280, 336, 64, 12}, 1, shard);  // beyond capacity
    }));
    bf::ShardDescriptor bad = shard;
    bad.timestamp_step = 0;
    EXPECT(throws<std::invalid_argument>([&] { rf.run(xs, ws_, small, 1, bad); }));
    EXPECT(throws<std::invalid_argument>([&] {
        bf::CudaOfflineFrameRunner(dims, bf::CudaBeamformerKernel::Direct, std::nullopt,
                                   bf::CudaBeamformerOutput::QuantizedInt8);
    }));

    // trackers: every generation's entry point runs the one tracker
    const bf::Dimensions td{1280, 336, 64, 1};
    bf::PackedVoltage tx(td.n_time * td.n_freq * td.n_ant);
    for (auto& b : tx) b = static_cast<std::uint8_t>(lcg(seed) & 0xFF);
    bf::TrackerConfig tc;
    tc.trajectory.direction_start = bf::direction_from_lm(0.01F, -0.02F);
    tc.trajectory.direction_rate_per_sample = {1e-6F, 2e-6F};
    const auto v5 = bf::cuda_beam_tracker_v5(tx, td, tc);
    EXPECT(v5.size() == td.n_time * td.n_freq);
    EXPECT(bf::cuda_tracker_v2_packed_intensity(tx, td, tc, bf::CudaTrackerKernelV2::Fused) == v5);
    EXPECT(bf::cuda_beam_tracker_fused_warp_shuffle(tx, td, tc) == v5);
    EXPECT(bf::cuda_beam_tracker_v3(tx, td, tc) == v5);
    bf::V4ExecutionConfig v4c;
    v4c.mode = bf::V4KernelMode::BlockReduction;
    EXPECT(bf::cuda_beam_tracker_v4(tx, td, tc, v4c) == v5);
    bf::Intensities st;
    bf::cuda_beam_tracker_v4_stream(tx, td, tc, st, 3);
    EXPECT(st == v5);
    // batched V5 stream: 2 windows per batch, batches at windows 0 and 2
    bf::BatchedTrackerStreamV5 bs(bf::Dimensions{320, 336, 64, 1}, tc, 2);
    EXPECT(bs.batch_size() == 2 && bs.window_bytes() == 320 * 336 * 64);
    std::vector<float> got(bs.batch_output_floats());
    for (std::size_t first : {0, 2}) {
        bs.process_batch(first, tx.data() + first * bs.window_bytes(), got.data());
        EXPECT(std::memcmp(got.data(), v5.data() + first * 320 * 336, got.size() * 4) == 0);
    }
    EXPECT(bs.last_kernel_time_ms() > 0.0F);
    bf::TrackerConfig multi = tc;
    EXPECT(throws<std::invalid_argument>(
        [&] { bf::cuda_beam_tracker_v5(tx, bf::Dimensions{1280, 336, 64, 2}, multi); }));

    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
