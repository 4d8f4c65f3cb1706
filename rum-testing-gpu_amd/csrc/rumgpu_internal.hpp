// Internal declarations shared by the librumgpu translation units.
#pragma once
#include "common.hpp"
#include "../../include/rumgpu.h"

// Takes the device address of the fixed scratch array behind rumgpu_lb64_noalloc for `device`
// (which must be current); called once per device by rumgpu_ctx_create.
int rg_noalloc_setup(int device);
