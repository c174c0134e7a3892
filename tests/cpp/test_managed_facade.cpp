// The reference's facade usage over HipImpl<float> with the reference's
// default allocator kind: managed memory (include/CPRA.hpp:19 calls
// NewAllocator(type_) with gpu_managed = true, include/Allocator.hpp:73-82).
//
// As the reference's programs do (src/CPRA_MKL.cpp:10-232,
// include/detail/CPRA.inl:8-48), the host fills the buffers with fread and
// the RealToComplex loop (include/CPRA.hpp:31-39), the reconstruction runs
// through the CpraImpl<float> interface (Step A's difference map,
// src/CPRA_MKL.cpp:35-55, then ConvergeError into a managed error buffer and
// the final projection, :126-140), and the host reads the results back with
// the ComplexToReal loop (include/CPRA.hpp:40-47) and fwrite, with no
// Memcpy between host and device.
//
// usage: test_managed_facade <dir> M N B preit beta
//   reads  <dir>/A.bin (float M*N), <dir>/S.bin (float M*N),
//          <dir>/psi_re.bin, <dir>/psi_im.bin (float M*N*B each)
//   writes <dir>/psi.bin (complex float M*N*B, after the final projection),
//          <dir>/real.bin (float M*N*B, its real part via ComplexToReal),
//          <dir>/err.bin (float B, ConvergeError of the last DM iteration)
// Exit status: 0 done, 1 error, 2 no GPU.
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>

#include "HIP_Impl.hpp"

using namespace CPRA;
using cf = std::complex<float>;

static bool read_file(const std::string& name, float* dst, size_t count)
{
    FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) return false;
    const size_t got = std::fread(dst, sizeof(float), count, f);
    std::fclose(f);
    return got == count;
}

static bool write_file(const std::string& name, const void* src, size_t bytes)
{
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) return false;
    const size_t put = std::fwrite(src, 1, bytes, f);
    std::fclose(f);
    return put == bytes;
}

int main(int argc, char** argv)
{
    if (argc != 7) {
        std::printf("usage: %s dir M N B preit beta\n", argv[0]);
        return 1;
    }
    const std::string dir = argv[1];
    const uint64_t M = std::strtoull(argv[2], nullptr, 10), N = std::strtoull(argv[3], nullptr, 10);
    const uint64_t B = std::strtoull(argv[4], nullptr, 10);
    const int preit = std::atoi(argv[5]);
    const float beta = std::strtof(argv[6], nullptr);
    if (cpra_hip_device_count() < 1) {
        std::printf("no GPU\n");
        return 2;
    }
    const uint64_t px = M * N, num = px * B;
    try {
        // the facade: one impl, one allocator of the reference's default kind
        auto impl = NewHIPImpl<float>(M, N, 1, B);
        auto alloc = NewHIPAllocator(true);
        auto* A = (float*)alloc->allocate(sizeof(float) * px);
        auto* S = (float*)alloc->allocate(sizeof(float) * px);
        auto* re = (float*)alloc->allocate(sizeof(float) * num);
        auto* im = (float*)alloc->allocate(sizeof(float) * num);
        auto* psi = (cf*)alloc->allocate(sizeof(cf) * num);
        auto* t1 = (cf*)alloc->allocate(sizeof(cf) * num);
        auto* t2 = (cf*)alloc->allocate(sizeof(cf) * num);
        auto* old_record = (cf*)alloc->allocate(sizeof(cf) * num);
        auto* err = (float*)alloc->allocate(sizeof(float) * B);
        auto* real = (float*)alloc->allocate(sizeof(float) * num);
        // ReadMatrixFromFile: fread straight into the allocated buffers
        if (!read_file(dir + "/A.bin", A, px) || !read_file(dir + "/S.bin", S, px) ||
            !read_file(dir + "/psi_re.bin", re, num) || !read_file(dir + "/psi_im.bin", im, num)) {
            std::printf("reading the inputs failed\n");
            return 1;
        }
        // RealToComplex-style host loop over the managed buffers
        for (uint64_t i = 0; i < num; ++i) psi[i] = cf(re[i], im[i]);
        // Step A (src/CPRA_MKL.cpp:35-55) with the convergence error of the
        // last iteration (:85-91, 112-121)
        const uint64_t n = num;
        impl->Memcpy(t1, psi, sizeof(cf) * n);
        impl->Memcpy(t2, psi, sizeof(cf) * n);
        for (int i = 0; i < preit; ++i) {
            if (i == preit - 1) impl->Memcpy(old_record, psi, sizeof(cf) * n);
            impl->Forward2D(t1);
            impl->DataConstraint(t1, A, n, B);
            impl->Backward2D(t1);
            impl->MergeAddData(psi, t1, -1.0f / beta, 1.0f + 1.0f / beta, n);
            impl->SpaceConstraint(t1, S, n, B);
            impl->SpaceConstraint(t2, S, n, B);
            impl->MergeAddData(psi, t2, 1.0f / beta, 1.0f - 1.0f / beta, n);
            impl->Forward2D(t2);
            impl->DataConstraint(t2, A, n, B);
            impl->Backward2D(t2);
            impl->MergeAddData(t1, psi, beta, 1.0f, n);
            impl->MergeAddData(t2, psi, -1.0f * beta, 1.0f, n);
            if (i == preit - 1) {
                impl->Memcpy(t1, psi, sizeof(cf) * n);
                impl->Forward2D(old_record);
                impl->Forward2D(t1);
                impl->ConvergeError(old_record, t1, err, M, N, 1, B);
            }
        }
        // final projection (:126-140)
        impl->Memcpy(t1, psi, sizeof(cf) * n);
        impl->Forward2D(t1);
        impl->DataConstraint(t1, A, n, B);
        impl->Backward2D(t1);
        impl->MergeAddData(psi, t1, -1.0f / beta, 1.0f + 1.0f / beta, n);
        impl->SpaceConstraint(t1, S, n, B);
        impl->Memcpy(psi, t1, sizeof(cf) * n);
        impl->Sync();
        // ComplexToReal-style host loop, then WriteMatrixToFile (fwrite)
        for (uint64_t i = 0; i < num; ++i) real[i] = psi[i].real();
        if (!write_file(dir + "/psi.bin", psi, sizeof(cf) * num) ||
            !write_file(dir + "/real.bin", real, sizeof(float) * num) ||
            !write_file(dir + "/err.bin", err, sizeof(float) * B)) {
            std::printf("writing the outputs failed\n");
            return 1;
        }
        for (void* p : {(void*)A, (void*)S, (void*)re, (void*)im, (void*)psi, (void*)t1, (void*)t2,
                        (void*)old_record, (void*)err, (void*)real})
            alloc->deallocate(p);
    } catch (const hip_error& e) {
        std::printf("hip_error %d: %s\n", e.code, e.what());
        return 1;
    }
    std::printf("done\n");
    return 0;
}
