// cpra_impl_interface.hpp -- the backend interface HipImpl<T> implements.
//
// Inside the reference tree HIP_Impl.hpp includes the reference's own
// include/CPRA_Impl.hpp (define CPRA_HIP_IN_REFERENCE_TREE).  For a standalone
// build this header declares an interface with the same virtual signatures as
// CPRA::CpraImpl<T> (include/CPRA_Impl.hpp:18-63 of the reference) and the
// IMPL_TYPE enum extended with HIP, so HipImpl<T> compiles and is tested
// without the reference sources.  The signatures have to be identical for the
// class to be a drop-in; the bodies are this project's own.
#pragma once
#include <complex>
#include <cstddef>
#include <cstdint>

namespace CPRA {

enum IMPL_TYPE { CUDA, MKL, HIP };

template <typename T> class CpraImpl {
public:
    CpraImpl() {}
    virtual bool Initialize(T* flat_data_ptr, uint64_t num) = 0;
    virtual bool Forward2D(std::complex<T>* flat_input) = 0;
    virtual bool Backward2D(std::complex<T>* flat_input) = 0;
    virtual bool FFTShift2D(std::complex<T>* flat_input, size_t M, size_t N, size_t Batch) = 0;
    virtual bool Forward3D(std::complex<T>* flat_input) = 0;
    virtual bool Backward3D(std::complex<T>* flat_input) = 0;
    virtual bool FFTShift3D(std::complex<T>* flat_input, size_t M, size_t N, size_t L,
                            size_t Batch) = 0;
    virtual bool SpaceConstraint(std::complex<T>* flat_src_data, T* flat_constr_data,
                                 uint64_t num, uint64_t batch_size) = 0;
    virtual bool DataConstraint(std::complex<T>* flat_src_data, T* flat_constr_data,
                                uint64_t num, uint64_t batch_size) = 0;
    virtual bool DataConstraint(std::complex<T>* flat_src_data,
                                std::complex<T>* flat_constr_data, uint64_t num,
                                uint64_t batch_size) = 0;
    virtual bool MergeAddData(std::complex<T>* flat_src, std::complex<T>* flat_dst, T alpha,
                              T beta, uint64_t num) = 0;
    virtual bool Normalization(std::complex<T>* flat_src, T norm, uint64_t num) = 0;
    virtual bool Real2DTo3DInterpolation(T* flat_2d_src, T* flat_3D_dst, T* angles, uint64_t m,
                                         uint64_t n, uint64_t p, uint64_t l) = 0;
    virtual bool Complex2DTo3DInterpolation(std::complex<T>* flat_2d_src,
                                            std::complex<T>* flat_3D_dst, T* angles,
                                            uint64_t m, uint64_t n, uint64_t p,
                                            uint64_t l) = 0;
    virtual bool ConvergeError(std::complex<T>* flat_old, std::complex<T>* flat_new,
                               T* flat_error, uint64_t m, uint64_t n, uint64_t l = 1,
                               uint64_t batch_size = 1) = 0;
    virtual bool Memcpy(void* dst, void* src, uint64_t bytes) = 0;
    virtual bool Sync() = 0;
    virtual ~CpraImpl() {}
};

class AllocatorInterface {
public:
    AllocatorInterface() {}
    virtual ~AllocatorInterface() {}
    virtual void* allocate(uint64_t alloc_bytes, int alignment = 64) const = 0;
    virtual void deallocate(void* dealloc_ptr) const = 0;
};

}  // namespace CPRA
