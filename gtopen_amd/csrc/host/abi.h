// Definitions behind the C ABI's opaque handles (include/gtopen_amd.h).
#pragma once

#include <memory>

#include "../../../include/gtopen_amd.h"
#include "host.h"

struct gto_spot {
    std::unique_ptr<gto::Spot> spot;
    gto_spot_view view;
};
