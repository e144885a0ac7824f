/*
 * PLFRAME layout record of the cusat_dsp C++ API, as the receiver's PL
 * layer hands it to rx_fec_pool::submit: the data members of the
 * reference's pl_layout (lib/cusat_dsp/framing/plframe.h:38-53), same names,
 * types and order.  PL framing itself (PLS code tables, pilot and slot
 * arithmetic) is not part of this library; rx_fec_pool reads `modcod`.
 */
#ifndef CUSAT_AMD_DSP_PLFRAME_H
#define CUSAT_AMD_DSP_PLFRAME_H

#include <cstdint>

#include "modcod_types.h"

// A translation unit that already includes the reference's
// framing/plframe.h (for its PLS functions) gets that pl_layout, which has
// the same members.
#ifndef INCLUDED_CUSAT_DSP_FRAMING_PLFRAME_H
namespace cusat_dsp {

struct pl_layout {
    uint8_t pls_code = 0;                     // 8-bit PLS code of the PLHEADER
    uint8_t modcod = 0;                       // MODCOD index (acm_config.h)
    fecframe_size_t fec_size = FECFRAME_NORMAL;
    bool pilots = false;                      // pilot blocks present
    bool dummy = false;                       // dummy PLFRAME (no payload)
    bool reserved = false;                    // reserved code with a fixed length
    uint8_t bits_per_symbol = 0;              // 0 when no supported constellation
    uint16_t slots = 0;                       // S

    int data_symbols() const { return 90 * slots; }  // XFECFRAME length
};

}  // namespace cusat_dsp
#endif

#endif
