// The opaque cujpegxl_encoder handle shared by the C ABI and the harness.
#pragma once

#include "cujpegxl/cujpegxl.h"
#include "encoder.h"

struct cujpegxl_encoder {
    cujpegxl_config config;
    cjx::Encoder engine;
};
