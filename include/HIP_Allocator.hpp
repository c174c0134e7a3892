// HIP_Allocator.hpp -- HipAllocator, the HIP backend's AllocatorInterface
// (the reference's CudaAllocator / CudaManagedAllocator,
// include/Allocator.hpp:18-52), and NewHIPAllocator for the reference's
// NewAllocator switch (include/Allocator.hpp:73-93).
//
// Standalone, HIP_Impl.hpp includes this header.  Inside the reference tree
// (CPRA_HIP_IN_REFERENCE_TREE) the reference's include/Allocator.hpp includes
// it at its end, after AllocatorInterface (that header has no include guard,
// so this one must not include it); HIP_Impl.hpp, included before, declares
// NewHIPAllocator for the NewAllocator case
//   case IMPL_TYPE::HIP: return NewHIPAllocator(gpu_managed);
// whose default, gpu_managed = true, gives managed memory that host code
// (fread, the RealToComplex / ComplexToReal loops) reads and writes directly,
// as with the CUDA backend.
#pragma once
#include <cstdint>
#include <memory>
#ifndef CPRA_HIP_IN_REFERENCE_TREE
#include "cpra_impl_interface.hpp"
#endif

#include "HIP_Error.hpp"
#include "HIP_Impl.hpp"
#include "cpra_hip.h"

namespace CPRA {

class HipAllocator final : public AllocatorInterface {
public:
    explicit HipAllocator(bool managed = false) : managed_(managed) {}
    void* allocate(uint64_t alloc_bytes, int alignment = 64) const override
    {
        (void)alignment;  // HIP allocations are at least 256-byte aligned
        void* p = nullptr;
        const int rc = cpra_hip_malloc(nullptr, &p, alloc_bytes, managed_ ? 1 : 0);
        if (rc < 0) throw hip_error(rc, cpra_hip_last_error(nullptr));
        return p;
    }
    void deallocate(void* p) const override
    {
        const int rc = cpra_hip_free(nullptr, p);
        if (rc < 0) throw hip_error(rc, cpra_hip_last_error(nullptr));
    }
    bool managed() const { return managed_; }

private:
    bool managed_;
};

// Device memory by default (declared in HIP_Impl.hpp): on MI355X device
// memory with explicit Memcpy is the fast path.  gpu_managed = true gives
// hipMallocManaged memory, which the host reads and writes directly (the
// reference's CUDA default).
inline std::unique_ptr<AllocatorInterface> NewHIPAllocator(bool gpu_managed)
{
    return std::make_unique<HipAllocator>(gpu_managed);
}

}  // namespace CPRA
