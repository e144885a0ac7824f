/*
 * MODCOD definitions of the cusat_dsp C++ API: the reference's MODCOD index
 * (the key its PL layer signals and rx_fec_pool routes on), constellation,
 * code rate, name, spectral efficiency and measured Es/N0 threshold.
 * The index -> (constellation, rate) pairs and the threshold values are the
 * reference's (lib/cusat_dsp/acm_config.h:50-117): a receiver built on this
 * library must route the same PLS-signalled index to the same code.
 */
#ifndef CUSAT_AMD_DSP_ACM_CONFIG_H
#define CUSAT_AMD_DSP_ACM_CONFIG_H

#include <cstdint>
#include <cstring>
#include <vector>

#include "modcod_types.h"

// A translation unit that already includes the reference's acm_config.h
// (its ACM controller does) keeps that header's inline definitions, which
// this library's out-of-line ones equal.
#ifndef INCLUDED_CUSAT_DSP_ACM_CONFIG_H
namespace cusat_dsp {

// Headroom constants of the reference's ACM controller (acm_config.h:24-30).
inline constexpr float ACM_DEFAULT_GUARD_DB = 2.5f;
inline constexpr float ACM_DEFAULT_BOOTSTRAP_MARGIN_DB = 2.0f;

struct modcod_def_t {
    uint8_t modcod_idx;
    modulation_t constellation;
    code_rate_t code_rate;
    const char* name;
    float spectral_eff;  // bits/symbol
    float min_snr_db;    // Es/N0 of post-BCH FER <= 1e-5 (normal frames, 25 iterations)
};

/** Every supported MODCOD, ascending spectral efficiency. */
const std::vector<modcod_def_t>& get_supported_modcods();

/** The definition of a MODCOD index, or nullptr. */
const modcod_def_t* find_modcod(uint8_t modcod_idx);

/** The definition of a MODCOD name ("qpsk3/5", "16apsk2/3l", ...), or nullptr. */
const modcod_def_t* find_modcod_by_name(const char* name);

/** True when FEC parameters exist for (frame size, rate)
 *  (fec/code_params.cc has_fec_params). */
bool has_fec_params(fecframe_size_t framesize, code_rate_t rate);

/** The MODCOD may be sent in this frame size: its code exists there, and
 *  short / medium frames carry DVB-S2 MODCODs (index 1..28) only
 *  (acm_config.h:140-160). */
inline bool modcod_supports_framesize(const modcod_def_t& mc, fecframe_size_t fs)
{
    return has_fec_params(fs, mc.code_rate) && (fs == FECFRAME_NORMAL || mc.modcod_idx < 29);
}

}  // namespace cusat_dsp
#endif

#endif
