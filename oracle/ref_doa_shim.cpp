// ref_doa_shim.cpp -- the reference's FP32 DOA primitives, compiled from the
// reference's own source file (src/cpu_opt_beam_tracker.cpp, included
// below from the reference tree by oracle/Makefile's include path) into
// oracle/_ref/libbfref_doa.so.  TEST AND BASELINE INFRASTRUCTURE ONLY: it
// times the reference CPU path of the adaptive-DOA update on the host.
//
// One update composes the primitives in the order of
// CpuOptBeamTracker::run_into (src/cpu_opt_beam_tracker.cpp:1853-1910,
// 2063-2160), parallel over channels as there (OpenMP, static schedule):
//   decode the K snapshots of the channel (decode_snapshot_into, :523-536),
//   spatial_smoothed_covariance_into (:559-658; with P_sub < n_ant it also
//     applies the reference's forward-backward step),
//   fold R = alpha R + (1 - alpha) R_block (:1880-1892) for the Bartlett
//     state, and for the Capon state after apply_diagonal_load_into
//     (:244-253), trace,
//   cholesky_factorize (:398-421) of the Capon state,
//   then per angle theta of a 1-D scan (direction (sin, 0, cos)):
//     steer_one (:96-106), bartlett_power_hermitian_trace (:287-306) and
//     capon_power_from_factor (:423-437),
// and sums the per-channel spectra over channels in channel order.
#include "cpu_opt_beam_tracker.cpp"

#include <cstdint>
#include <vector>

#include <omp.h>

namespace {

using beamformer::Cfloat;

}  // namespace

extern "C" int ref_doa_update_f32(const std::uint8_t* x, std::size_t T, std::size_t F,
                                  std::size_t A, std::size_t t0, std::size_t K, std::size_t M,
                                  float alpha, float load, int first, const float* freqs,
                                  const float* xpos, std::size_t n_theta, const double* theta,
                                  float* state_bartlett, float* state_capon,
                                  float* p_bartlett, float* p_capon) {
    try {
        namespace bf = beamformer;
        if (t0 + K > T || M == 0 || M > A) return -1;
        const bool smoothing = M < A;
        const std::size_t L = smoothing ? A - M + 1 : 1;
        std::vector<bf::Vec3> positions(A);
        for (std::size_t a = 0; a < A; ++a) positions[a] = {xpos[a], 0.0F, 0.0F};
        std::vector<float> per_f_b(F * n_theta), per_f_c(F * n_theta);
        auto* Rb_all = reinterpret_cast<Cfloat*>(state_bartlett);
        auto* Rc_all = reinterpret_cast<Cfloat*>(state_capon);
#pragma omp parallel
        {
            std::vector<Cfloat> snap(K * A), block(M * M), loaded(M * M), fac(M * M),
                sub(L * M * M), a(M), v(M);
#pragma omp for schedule(static)
            for (std::ptrdiff_t fi = 0; fi < static_cast<std::ptrdiff_t>(F); ++fi) {
                const std::size_t f = static_cast<std::size_t>(fi);
                for (std::size_t k = 0; k < K; ++k)
                    for (std::size_t ai = 0; ai < A; ++ai) {
                        const auto s = bf::unpack_complex_int4(x[((t0 + k) * F + f) * A + ai]);
                        snap[k * A + ai] =
                            Cfloat{static_cast<float>(s.real), static_cast<float>(s.imag)};
                    }
                bf::spatial_smoothed_covariance_into(snap.data(), K, A, A, smoothing ? M : 0, M,
                                                     block.data(), sub.data(), L * M * M);
                loaded = block;
                bf::apply_diagonal_load_into(loaded.data(), M, load);
                Cfloat* Rb = Rb_all + f * M * M;
                Cfloat* Rc = Rc_all + f * M * M;
                const float a_comp = 1.0F - alpha;
                for (std::size_t i = 0; i < M * M; ++i) {
                    Rb[i] = first ? block[i] : alpha * Rb[i] + a_comp * block[i];
                    Rc[i] = first ? loaded[i] : alpha * Rc[i] + a_comp * loaded[i];
                }
                float tr = 0.0F;
                for (std::size_t i = 0; i < M; ++i) tr += Rb[i * M + i].real();
                std::copy(Rc, Rc + M * M, fac.begin());
                const bool ok = bf::cholesky_factorize(fac.data(), M);
                const double k_f =
                    bf::two_pi * static_cast<double>(freqs[f]) / bf::speed_of_light_m_per_s;
                for (std::size_t th = 0; th < n_theta; ++th) {
                    const bf::Vec3 dir{static_cast<float>(std::sin(theta[th])), 0.0F,
                                       static_cast<float>(std::cos(theta[th]))};
                    for (std::size_t m = 0; m < M; ++m) a[m] = bf::steer_one(positions, k_f, dir, m);
                    per_f_b[f * n_theta + th] =
                        bf::bartlett_power_hermitian_trace(Rb, a.data(), M, tr);
                    per_f_c[f * n_theta + th] =
                        ok ? bf::capon_power_from_factor(fac.data(), a.data(), M, v.data())
                           : 0.0F;
                }
            }
        }
        for (std::size_t th = 0; th < n_theta; ++th) {
            float sb = 0.0F, sc = 0.0F;
            for (std::size_t f = 0; f < F; ++f) {
                sb += per_f_b[f * n_theta + th];
                sc += per_f_c[f * n_theta + th];
            }
            p_bartlett[th] = sb;
            p_capon[th] = sc;
        }
        return 0;
    } catch (...) {
        return -3;
    }
}
