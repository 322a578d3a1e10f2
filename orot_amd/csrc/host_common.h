// host_common.h -- host-side helpers shared by the C-ABI layer and the engine.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/orot/deflate.h"

namespace orot_gpu {

// Message of the last failure on the calling thread (orot_gpu_last_error).
void set_error(const char* where, hipError_t e);
void set_error_msg(const std::string& m);
const char* last_error_message();

// The allocator of deflate_set_allocator (default: posix_memalign / free),
// used for the host memory of the decompress side.
deflate_allocator host_allocator();
void set_allocator(const deflate_allocator& a);

}  // namespace orot_gpu
