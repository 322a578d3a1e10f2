// host_common.h -- host-side helpers shared by the C-ABI layer and the engine.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace orot_gpu {

// Message of the last failure on the calling thread (orot_gpu_last_error).
void set_error(const char* where, hipError_t e);
void set_error_msg(const std::string& m);
const char* last_error_message();

}  // namespace orot_gpu
