// Device blocks a context keeps for the library's scheduler between runs:
// the trace rounds' buffers of each engine system.  Allocating and freeing
// them per batch costs milliseconds (hipFree waits for the device).  The
// context frees what is parked when its genotypes change or it is destroyed.
// Defined with the context (saige_capi.hip; the CPU engine stand-in under
// tests/native defines its own).
#pragma once

#include <cstdint>

#include "../../include/saige_hip.h"

namespace saige {
// a parked device block of exactly `bytes`, taken off the context, or nullptr
void* ctx_unpark(saige_ctx* c, int64_t bytes);
// hands a device block from saige_alloc_device to the context
void ctx_park(saige_ctx* c, void* p, int64_t bytes);
}  // namespace saige
