// Internal declarations shared by the engine's translation units.
#pragma once

#include "physim_gpu.h"

// Records a formatted error message for pg_last_error() and returns `code`.
int pg_set_error(int code, const char* fmt, ...);
