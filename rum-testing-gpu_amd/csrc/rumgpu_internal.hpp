// Internal declarations shared by the librumgpu translation units.
#pragma once
#include "common.hpp"
#include "../../include/rumgpu.h"
