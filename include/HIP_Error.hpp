// HIP_Error.hpp -- the exception the HIP backend throws for HIP runtime
// errors, as the reference's CUDA backend throws from CPRA_CUDA_TRY
// (include/CUDA_Error_Check.hpp:69-82).  Shared by HIP_Impl.hpp and
// HIP_Allocator.hpp.
#pragma once
#include <stdexcept>
#include <string>

namespace CPRA {

struct hip_error : public std::runtime_error {
    int code;
    hip_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

}  // namespace CPRA
