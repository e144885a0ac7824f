// Host-side checks of the C++ cusat_dsp API (no GPU needed): the acceptance
// rule (lib/qa_rx_fec_chain.cc:9-17), the decision-directed SNR estimators
// (modulation/dd_snr.cc semantics), argument errors raised before any
// device work, and the MODCOD table (acm_config.h).
#include <cmath>
#include <random>
#include <stdexcept>
#include <string>

#include "cusat_dsp/acm_config.h"
#include "cusat_dsp/ldpc_decode.h"
#include "cusat_dsp/rx_fec_chain.h"
#include "cusat_fec.h"
#include "qa_util.h"

using namespace cusat_dsp;

int main()
{
    return qa_run({
        {"acceptance_requires_ldpc_and_bch",
         [] {
             QA_CHECK(rx_fec_chain::frame_passes_fec_checks(25, true));
             QA_CHECK(rx_fec_chain::frame_passes_fec_checks(0, true));
             QA_CHECK(!rx_fec_chain::frame_passes_fec_checks(-1, true));
             QA_CHECK(!rx_fec_chain::frame_passes_fec_checks(3, false));
             QA_CHECK(!rx_fec_chain::frame_passes_fec_checks(-1, false));
         }},
        {"dd_snr_estimates",
         [] {
             const float a = std::sqrt(0.5f);
             std::vector<complexf> clean(1000), noisy(200000);
             std::mt19937 rng(3);
             std::uniform_int_distribution<int> bit(0, 1);
             for (auto& s : clean) s = complexf(bit(rng) ? a : -a, bit(rng) ? a : -a);
             QA_CHECK_EQ(qpsk_snr_dd(clean.data(), (int)clean.size()), 1e6f);  // error-free: 60 dB
             // 15 dB: decisions are almost never wrong, so the estimate is the true SNR
             const float sigma = std::sqrt(std::pow(10.0f, -1.5f) / 2.0f);
             std::normal_distribution<float> nd(0.0f, sigma);
             for (auto& s : noisy) s = complexf((bit(rng) ? a : -a) + nd(rng), (bit(rng) ? a : -a) + nd(rng));
             const float est = 10.0f * std::log10(qpsk_snr_dd(noisy.data(), (int)noisy.size()));
             QA_CHECK(std::fabs(est - 15.0f) < 0.1f);
             // the general estimator on the QPSK points agrees with the QPSK one at stride 1
             const complexf pts[4] = {{a, a}, {a, -a}, {-a, a}, {-a, -a}};
             const float g = constellation_snr_dd(pts, 4, noisy.data(), (int)noisy.size(), 1);
             QA_CHECK(std::fabs(10.0f * std::log10(g) - est) < 1e-3f);
             std::vector<complexf> bad(4, complexf(NAN, 0.0f));
             QA_CHECK_EQ(qpsk_snr_dd(bad.data(), 4), 0.0f);
         }},
        {"invalid_arguments_throw",
         [] {
             bool thrown = false;
             try {
                 rx_fec_chain c(FECFRAME_SHORT, RATE_9_10, MOD_QPSK);  // no short 9/10 code
             } catch (const std::invalid_argument&) {
                 thrown = true;
             }
             QA_CHECK(thrown);
             thrown = false;
             try {
                 ldpc_decode d(FECFRAME_MEDIUM, RATE_1_2);
             } catch (const std::invalid_argument&) {
                 thrown = true;
             }
             QA_CHECK(thrown);
             thrown = false;
             try {
                 rx_fec_chain c(FECFRAME_NORMAL, RATE_3_4, MOD_8PSK, 25, 64, 0, 4);
             } catch (const std::invalid_argument&) {
                 thrown = true;
             }
             QA_CHECK(thrown);
         }},
        {"modcod_table",
         [] {
             const auto& t = get_supported_modcods();
             QA_CHECK_EQ(t.size(), (size_t)66);
             const modcod_def_t* m = find_modcod(14);
             QA_CHECK(m && m->constellation == MOD_8PSK && m->code_rate == RATE_3_4);
             QA_CHECK(m && std::string(m->name) == "8psk3/4" && m->min_snr_db == 8.11f);
             QA_CHECK(find_modcod(0) == nullptr);
             QA_CHECK(find_modcod(67) == nullptr);
             const modcod_def_t* l = find_modcod_by_name("16apsk2/3l");
             QA_CHECK(l && l->modcod_idx == 62 && l->code_rate == RATE_20_30);
             QA_CHECK(find_modcod_by_name("qpsk7/8") == nullptr);
             // short frames: DVB-S2 MODCODs whose code exists there
             QA_CHECK(modcod_supports_framesize(*find_modcod(14), FECFRAME_SHORT));
             QA_CHECK(!modcod_supports_framesize(*find_modcod(11), FECFRAME_SHORT));  // no short 9/10
             QA_CHECK(!modcod_supports_framesize(*find_modcod(31), FECFRAME_SHORT));  // S2X index
             QA_CHECK(modcod_supports_framesize(*find_modcod(31), FECFRAME_NORMAL));
             // every index unique, and every MODCOD a code this library decodes
             for (size_t a = 0; a < t.size(); a++) {
                 for (size_t b = a + 1; b < t.size(); b++) QA_CHECK(t[a].modcod_idx != t[b].modcod_idx);
                 cusat_fec_info inf;
                 QA_CHECK_EQ(cusat_fec_query(FECFRAME_NORMAL, t[a].code_rate, t[a].constellation, 0, &inf), CUSAT_OK);
             }
         }},
    });
}
