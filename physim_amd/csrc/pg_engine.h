// Internal declarations shared by the engine's translation units.
#pragma once

#include "physim_gpu.h"

// Records a formatted error message for pg_last_error() and returns `code`.
int pg_set_error(int code, const char* fmt, ...);

// The plugin's engine (created on first use, PHYSIM_GPU_DEVICE selects the
// device) and the device transform of one of its gravity elements: false
// when `element` is not a live GPU gravity element of this library.
pg_engine* pg_plugin_engine(void);
bool pg_plugin_element_params(const void* element, pg_params* out);
