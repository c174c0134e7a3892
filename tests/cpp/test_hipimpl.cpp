// The reference's unit tests (test/unitest_mkl.cpp, test/unitest_cuda.cu)
// restated against HipImpl<T> through the CpraImpl<T> interface, plus a
// short run of the reference DM composition (src/CPRA_MKL.cpp:94-110).
// Host-only C++17; links libcpra_hip.so.  Exit status 0 = all passed.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "HIP_Impl.hpp"

static int g_fail = 0;
#define EXPECT(cond)                                                       \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
            ++g_fail;                                                      \
        }                                                                  \
    } while (0)

using namespace CPRA;

static void test_initialize()
{
    auto obj = NewHIPImpl<float>(1, 1, 1, 1);
    auto alloc = NewHIPAllocator();
    float* d = (float*)alloc->allocate(sizeof(float) * 1000);
    EXPECT(obj->Initialize(d, 1000));
    std::vector<float> h(1000);
    EXPECT(obj->Memcpy(h.data(), d, sizeof(float) * 1000));
    EXPECT(obj->Sync());
    for (float v : h) EXPECT(v >= 0 && v <= 1);
    alloc->deallocate(d);
    std::printf("Test_HIP_Initialize done\n");
}

static void test_norm_merge()
{
    auto obj = NewHIPImpl<float>(1, 1, 1, 1);
    auto alloc = NewHIPAllocator();
    auto* d = (std::complex<float>*)alloc->allocate(sizeof(std::complex<float>) * 1000);
    std::vector<std::complex<float>> h(1000, {10, 10});
    obj->Memcpy(d, h.data(), sizeof(std::complex<float>) * 1000);
    EXPECT(obj->MergeAddData(d, d, 0.5f, 0.5f, 1000));
    obj->Memcpy(h.data(), d, sizeof(std::complex<float>) * 1000);
    obj->Sync();
    for (auto& v : h) EXPECT(v.real() == 10 && v.imag() == 10);
    EXPECT(obj->Normalization(d, 10.f, 1000));
    obj->Memcpy(h.data(), d, sizeof(std::complex<float>) * 1000);
    obj->Sync();
    for (auto& v : h) EXPECT(v.real() == 1 && v.imag() == 1);
    EXPECT(!obj->MergeAddData(d, d, 0.f, 1.f, 1000));  // alpha == 0 -> false
    EXPECT(!obj->MergeAddData(d, d, 1.f, 1.f, 0));     // num == 0 -> false
    alloc->deallocate(d);
    std::printf("Test_Norm_Merge done\n");
}

static void test_2d_dft()
{
    auto obj = NewHIPImpl<double>(10, 10, 1, 1);
    auto alloc = NewHIPAllocator();
    auto* d = (std::complex<double>*)alloc->allocate(sizeof(std::complex<double>) * 100);
    std::vector<std::complex<double>> h(100, {1, 0});
    obj->Memcpy(d, h.data(), sizeof(std::complex<double>) * 100);
    EXPECT(obj->Forward2D(d));
    obj->Memcpy(h.data(), d, sizeof(std::complex<double>) * 100);
    obj->Sync();
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j) {
            const double want = (i == 5 && j == 5) ? 100.0 : 0.0;
            EXPECT(std::fabs(h[i * 10 + j].real() - want) < 1e-13);
            EXPECT(std::fabs(h[i * 10 + j].imag()) < 1e-13);
        }
    EXPECT(obj->Backward2D(d));
    obj->Memcpy(h.data(), d, sizeof(std::complex<double>) * 100);
    obj->Sync();
    for (auto& v : h) {
        EXPECT(std::fabs(v.real() - 1.0) < 1e-14);
        EXPECT(std::fabs(v.imag()) < 1e-14);
    }
    alloc->deallocate(d);
    std::printf("Test_2D_DFT done\n");
}

// Reference DM composition on 64x64, 2 random starts sharing one pattern:
// the exact solution is a fixed point (src/CPRA_MKL.cpp:94-110).
static void test_dm_fixed_point()
{
    const int M = 64, N = 64, B = 2;
    const float beta = 0.9f;
    auto obj = NewHIPImpl<float>(M, N, 1, B);
    auto alloc = NewHIPAllocator();
    const size_t px = (size_t)M * N;
    std::vector<std::complex<float>> xo(px * B);
    std::vector<float> S(px), A(px);
    for (int r = 0; r < M; ++r)
        for (int c = 0; c < N; ++c) {
            const bool in = (r - M / 2) * (r - M / 2) + (c - N / 2) * (c - N / 2) <= 100;
            S[r * N + c] = in ? 1.f : 0.f;
            const float v = in ? 0.25f + 0.5f * (float)((r * 7 + c * 13) % 17) / 17.f : 0.f;
            xo[r * N + c] = {v, 0};
            xo[px + r * N + c] = {v, 0};
        }
    auto* dA = (std::complex<float>*)alloc->allocate(sizeof(std::complex<float>) * px);
    obj->Memcpy(dA, xo.data(), sizeof(std::complex<float>) * px);
    auto amp = std::make_unique<HipImpl<float>>(M, N, 1, 1);
    amp->Forward2D(dA);
    std::vector<std::complex<float>> F(px);
    amp->Memcpy(F.data(), dA, sizeof(std::complex<float>) * px);
    amp->Sync();
    for (size_t i = 0; i < px; ++i) A[i] = std::abs(F[i]);
    auto* psi = (std::complex<float>*)alloc->allocate(sizeof(std::complex<float>) * px * B);
    auto* t1 = (std::complex<float>*)alloc->allocate(sizeof(std::complex<float>) * px * B);
    auto* t2 = (std::complex<float>*)alloc->allocate(sizeof(std::complex<float>) * px * B);
    float* dS = (float*)alloc->allocate(sizeof(float) * px);
    float* dAm = (float*)alloc->allocate(sizeof(float) * px);
    obj->Memcpy(psi, xo.data(), sizeof(std::complex<float>) * px * B);
    obj->Memcpy(dS, S.data(), sizeof(float) * px);
    obj->Memcpy(dAm, A.data(), sizeof(float) * px);
    obj->Memcpy(t1, psi, sizeof(std::complex<float>) * px * B);
    obj->Memcpy(t2, psi, sizeof(std::complex<float>) * px * B);
    const uint64_t num = px * B;
    for (int it = 0; it < 5; ++it) {
        obj->Forward2D(t1);
        obj->DataConstraint(t1, dAm, num, B);
        obj->Backward2D(t1);
        obj->MergeAddData(psi, t1, -1.0f / beta, 1.0f + 1.0f / beta, num);
        obj->SpaceConstraint(t1, dS, num, B);
        obj->SpaceConstraint(t2, dS, num, B);
        obj->MergeAddData(psi, t2, 1.0f / beta, 1.0f - 1.0f / beta, num);
        obj->Forward2D(t2);
        obj->DataConstraint(t2, dAm, num, B);
        obj->Backward2D(t2);
        obj->MergeAddData(t1, psi, beta, 1.0f, num);
        obj->MergeAddData(t2, psi, -1.0f * beta, 1.0f, num);
    }
    std::vector<std::complex<float>> out(px * B);
    obj->Memcpy(out.data(), psi, sizeof(std::complex<float>) * px * B);
    obj->Sync();
    double err = 0, ref = 0;
    for (size_t i = 0; i < px * B; ++i) {
        err += std::norm(std::complex<double>(out[i]) - std::complex<double>(xo[i]));
        ref += std::norm(std::complex<double>(xo[i]));
    }
    std::printf("DM fixed point rel err %.3e\n", std::sqrt(err / ref));
    EXPECT(std::sqrt(err / ref) < 1e-5);
    std::vector<float> e(B);
    EXPECT(obj->ConvergeError(t1, psi, e.data(), M, N, 1, B));
    for (auto& x : {psi, t1, t2}) alloc->deallocate(x);
    alloc->deallocate(dS);
    alloc->deallocate(dAm);
    alloc->deallocate(dA);
    std::printf("Test_DM_FixedPoint done\n");
}

static void test_errors()
{
    auto obj = NewHIPImpl<float>(9, 8, 1, 1);  // odd: shift-folded FFT unsupported
    std::complex<float>* d = nullptr;
    EXPECT(!obj->Forward2D(d));                   // NULL -> false, no throw
    EXPECT(!obj->SpaceConstraint(nullptr, nullptr, 0, 1));
    std::printf("Test_Errors done\n");
}

int main()
{
    try {
        test_initialize();
        test_norm_merge();
        test_2d_dft();
        test_dm_fixed_point();
        test_errors();
    } catch (const hip_error& e) {
        std::printf("hip_error %d: %s\n", e.code, e.what());
        return 2;
    }
    std::printf("%s (%d failures)\n", g_fail ? "FAILED" : "PASSED", g_fail);
    return g_fail ? 1 : 0;
}
