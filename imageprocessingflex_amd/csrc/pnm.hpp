// Launch interface of the Netpbm / PFM raster unpacking (pnm.hip).
#pragma once

#include "common.hpp"

namespace ipfg {

// Layouts of a binary raster as the file stores it.
enum PnmRaster {
    PNM_BITS = 0,      // P4: rows of (w + 7) / 8 bytes, most significant bit first, 1 = black
    PNM_BYTES = 1,     // P5 / P6 with maxval < 256
    PNM_WORDS_BE = 2,  // P5 / P6 with maxval >= 256: big-endian 16-bit samples
    PFM_LE = 3,        // PFM, negative scale: little-endian floats, bottom row first
    PFM_BE = 4         // PFM, positive scale: big-endian floats, bottom row first
};
// Unpacks `raw` (the file's raster bytes, on the device) into dst (uint8 or
// uint16 for PNM, float32 for PFM), rescaling PNM samples from maxval to the
// type's range as the reference does.
hipError_t launch_pnm_unpack(const uint8_t* raw, int layout, int maxval, const DevImage& dst, hipStream_t s);

}  // namespace ipfg
