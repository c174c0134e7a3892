// HIP_Impl.hpp -- HipImpl<T>, the MI355X backend of CPRA::CpraImpl<T>.
//
// Sibling of the reference's MklImpl<T> (include/MKL_Impl.hpp) and
// CudaImpl<T> (include/CUDA_Impl.cuh).  Every method forwards to the C ABI of
// libcpra_hip.so (include/cpra_hip.h); this header needs no HIP headers and
// compiles with any C++17 host compiler.  Link with -lcpra_hip.
//
// Inside the reference tree define CPRA_HIP_IN_REFERENCE_TREE so the
// reference's include/CPRA_Impl.hpp provides CpraImpl<T> and IMPL_TYPE (with
// HIP added to the enum).  NewHIPAllocator is declared here and defined in
// HIP_Allocator.hpp, which needs the reference's AllocatorInterface: the
// reference's include/Allocator.hpp (no include guard, so this header does
// not include it) includes HIP_Allocator.hpp at its end.
//
// Semantics follow the MKL backend by default (the parity target):
// SpaceConstraint zeroes where (double)S < 0.9 and keeps negative real parts;
// DataConstraint leaves A == 0 pixels untouched.  set_space_mode() and
// set_data_skip_zero() select the CUDA backend's variants.
//
// Errors: argument and geometry errors make a method return false (message in
// last_error()); HIP runtime errors throw CPRA::hip_error, like the CUDA
// backend's CPRA_CUDA_TRY (include/CUDA_Error_Check.hpp:69-82).
#pragma once
#include <complex>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#ifdef CPRA_HIP_IN_REFERENCE_TREE
#include "CPRA_Impl.hpp"
#else
#include "cpra_impl_interface.hpp"
#endif

#include "HIP_Error.hpp"
#include "cpra_hip.h"
#include "cpra_hip_engine.h"

namespace CPRA {

class AllocatorInterface;
// Defined in HIP_Allocator.hpp: device memory by default, hipMallocManaged
// memory with gpu_managed = true (the reference's NewAllocator default).
inline std::unique_ptr<AllocatorInterface> NewHIPAllocator(bool gpu_managed = false);

template <typename T> class HipImpl final : public CpraImpl<T> {
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>,
                  "HipImpl<T> supports float and double");

public:
    HipImpl(uint64_t m, uint64_t n, uint64_t l, uint64_t batch_size, uint64_t seed = 0)
        : seed_(seed)
    {
        const int rc = cpra_hip_create(&h_, std::is_same_v<T, float> ? CPRA_HIP_F32 : CPRA_HIP_F64,
                                       m, n, l, batch_size);
        if (rc != CPRA_HIP_OK) throw hip_error(rc, cpra_hip_last_error(nullptr));
    }
    ~HipImpl() override
    {
        if (h_) cpra_hip_destroy(h_);
    }
    HipImpl(const HipImpl&) = delete;
    HipImpl& operator=(const HipImpl&) = delete;

    bool Initialize(T* flat_data_ptr, uint64_t num) override
    {
        return ok(cpra_hip_initialize(h_, flat_data_ptr, num, seed_++));
    }
    bool Forward2D(std::complex<T>* x) override { return ok(cpra_hip_forward2d(h_, x)); }
    bool Backward2D(std::complex<T>* x) override { return ok(cpra_hip_backward2d(h_, x)); }
    bool FFTShift2D(std::complex<T>* x, size_t M, size_t N, size_t Batch) override
    {
        return ok(cpra_hip_fftshift2d(h_, x, M, N, Batch));
    }
    bool Forward3D(std::complex<T>* x) override { return ok(cpra_hip_forward3d(h_, x)); }
    bool Backward3D(std::complex<T>* x) override { return ok(cpra_hip_backward3d(h_, x)); }
    bool FFTShift3D(std::complex<T>* x, size_t M, size_t N, size_t L, size_t Batch) override
    {
        return ok(cpra_hip_fftshift3d(h_, x, M, N, L, Batch));
    }
    bool SpaceConstraint(std::complex<T>* x, T* S, uint64_t num, uint64_t batch_size) override
    {
        return ok(cpra_hip_space_constraint(h_, x, S, num, batch_size, space_mode_));
    }
    bool DataConstraint(std::complex<T>* x, T* A, uint64_t num, uint64_t batch_size) override
    {
        return ok(cpra_hip_data_constraint_real(h_, x, A, num, batch_size, skip_zero_));
    }
    bool DataConstraint(std::complex<T>* x, std::complex<T>* Cc, uint64_t num,
                        uint64_t batch_size) override
    {
        return ok(cpra_hip_data_constraint_complex(h_, x, Cc, num, batch_size));
    }
    bool MergeAddData(std::complex<T>* src, std::complex<T>* dst, T alpha, T beta,
                      uint64_t num) override
    {
        return ok(cpra_hip_merge_add(h_, src, dst, (double)alpha, (double)beta, num));
    }
    bool Normalization(std::complex<T>* x, T norm, uint64_t num) override
    {
        return ok(cpra_hip_normalize(h_, x, (double)norm, num));
    }
    bool Real2DTo3DInterpolation(T* slices, T* vol, T* angles, uint64_t m, uint64_t n, uint64_t p,
                                 uint64_t l) override
    {
        return ok(cpra_hip_real2d_to_3d(h_, slices, vol, angles, m, n, p, l));
    }
    bool Complex2DTo3DInterpolation(std::complex<T>* slices, std::complex<T>* vol, T* angles,
                                    uint64_t m, uint64_t n, uint64_t p, uint64_t l) override
    {
        return ok(cpra_hip_complex2d_to_3d(h_, slices, vol, angles, m, n, p, l));
    }
    bool ConvergeError(std::complex<T>* old_, std::complex<T>* new_, T* host_err, uint64_t m,
                       uint64_t n, uint64_t l = 1, uint64_t batch_size = 1) override
    {
        return ok(cpra_hip_converge_error(h_, old_, new_, host_err, m, n, l, batch_size));
    }
    bool Memcpy(void* dst, void* src, uint64_t bytes) override
    {
        return ok(cpra_hip_memcpy(h_, dst, src, bytes));
    }
    bool Sync() override { return ok(cpra_hip_sync(h_)); }

    // --- backend-specific -------------------------------------------------
    void set_space_mode(int mode) { space_mode_ = mode; }
    void set_data_skip_zero(bool skip) { skip_zero_ = skip ? 1 : 0; }
    const char* last_error() const { return cpra_hip_last_error(h_); }
    cpra_hip_handle handle() const { return h_; }

private:
    bool ok(int rc)
    {
        if (rc == CPRA_HIP_OK) return true;
        if (rc == CPRA_HIP_FALSE || rc == CPRA_HIP_ERR_ARG || rc == CPRA_HIP_ERR_UNSUPPORTED)
            return false;
        throw hip_error(rc, cpra_hip_last_error(h_));
    }

    cpra_hip_handle h_ = nullptr;
    uint64_t seed_ = 0;
    int space_mode_ = CPRA_HIP_SPACE_MKL;
    int skip_zero_ = 1;
};

template <typename T>
std::unique_ptr<CpraImpl<T>> NewHIPImpl(uint64_t m, uint64_t n, uint64_t l, uint64_t batch_size)
{
    return std::make_unique<HipImpl<T>>(m, n, l, batch_size);
}

}  // namespace CPRA

#ifndef CPRA_HIP_IN_REFERENCE_TREE
#include "HIP_Allocator.hpp"
#endif
