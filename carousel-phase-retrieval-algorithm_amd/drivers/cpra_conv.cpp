// cpra_conv.cpp -- the reference's conventional 3D reconstruction driver
// (src/CONV_MKL.cpp:10-174, identical to src/CONV_CUDA.cu) on the HIP backend.
//
// Same command line (include/Command_Parser.hpp:21-36, of which it reads M,
// N, L, iter, Batch, Total and Beta), same inputs (raw float32 files: the
// data constraint and the space constraint, [L][M][N] each, shared by every
// batch member, include/detail/CPRA.inl:8-29) and outputs per batch member
// k = 1, 2, ...: <output_reconstruction>.<k> with the real part of the
// reconstructed volume [L][M][N] float32 and <output_error>.<k> with its
// final convergence error (src/CONV_MKL.cpp:140-159).
//
// One batch of n volumes follows src/CONV_MKL.cpp:11-101: t1 = t2 = psi
// (:33-34); `iter` difference-map iterations with Forward3D/Backward3D in the
// modulus projection and t1, t2 carried across iterations (:47-63); the
// Fourier-domain change of the last iteration, ||F psi_before - F psi_after||
// / ||F psi_after|| over the whole volume (:40-45, 65-74); the final
// projection (:78-88).  The iterations run through the fused volume engine
// cpra_hip_dm3d_run (four HBM passes per iteration: z lines with the modulus
// projection, y lines, x rows with the update, y lines) for l, m, n in
// {64, 128, 192, 256, 384, 512}; other sizes run the reference's composition
// of HipImpl primitives.
//
// Deviations, each deliberate:
//  * every batch member gets its own output pair; the reference writes the
//    last member of each batch only, with an error file of four floats read
//    from a one-float buffer (src/CONV_MKL.cpp:145-158);
//  * with iter = 0 the error is 0 (the reference leaves it uninitialised);
//  * --init FILE reads the initial guess psi [Batch][L][M][N] complex64
//    (member i of a batch uses member i of the file), --seed S seeds the
//    uniform initialisation otherwise (the reference draws from an unseeded
//    generator).
#include <algorithm>
#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "HIP_Impl.hpp"
#include "cpra_hip_engine.h"
#include "driver_common.hpp"

using namespace CPRA;
using namespace cpra_driver;
using cf = std::complex<float>;

namespace {

bool fast_size(int v)
{
    return v == 64 || v == 128 || v == 192 || v == 256 || v == 384 || v == 512;
}

// One batch of n volumes (src/CONV_MKL.cpp:11-101).  A, S: device [L][M][N];
// mask: device u8 [n][L][M][N] (the support of S for every member).  Returns
// the real parts [n][L][M][N] and the errors [n].
void reconstruct(const Options& o, int n, const float* A, float* S, const uint8_t* mask,
                 const std::vector<cf>* init, std::vector<float>& real_out,
                 std::vector<float>& err_out)
{
    const uint64_t M = o.M, N = o.N, L = o.L, vox = M * N * L, per = vox * n;
    const float B = o.Beta;
    auto impl = NewHIPImpl<float>(M, N, L, n);
    auto* hip = static_cast<HipImpl<float>*>(impl.get());
    cpra_hip_handle h = hip->handle();
    auto alloc = NewHIPAllocator();
    cf* psi = (cf*)alloc->allocate(sizeof(cf) * per);
    cf* t1 = (cf*)alloc->allocate(sizeof(cf) * per);
    cf* t2 = (cf*)alloc->allocate(sizeof(cf) * per);
    cf* old = (cf*)alloc->allocate(sizeof(cf) * per);
    auto chk = [&](bool ok, const char* what) {
        if (!ok) throw std::runtime_error(std::string(what) + ": " + cpra_hip_last_error(h));
    };
    if (init) {
        impl->Memcpy(psi, (void*)init->data(), sizeof(cf) * per);
        chk(impl->Sync(), "Sync");
    } else {
        check(cpra_hip_initialize(h, reinterpret_cast<float*>(psi), 2 * per, o.seed), h, "Initialize");
    }
    cpra_hip_dm_params dp;
    cpra_hip_dm_default_params(&dp);
    dp.beta = B;
    dp.textbook = 0;
    dp.space_mode = CPRA_HIP_DM_SPACE_MKL;
    dp.dc_skip_zero = 1;
    dp.use_graph = 1;
    const bool fused = fast_size(o.M) && fast_size(o.N) && fast_size(o.L);
    if (!fused)
        std::cout << "Volume " << L << " x " << M << " x " << N
                  << " runs the primitive composition (the fused engine takes 64..512)." << std::endl;
    auto dm = [&](int iters) {
        if (iters <= 0) return;
        if (fused) {
            dp.iters = iters;
            check(cpra_hip_dm3d_run(h, psi, t1, t2, A, 1, mask, &dp), h, "difference map");
            return;
        }
        float* Ap = const_cast<float*>(A);
        for (int i = 0; i < iters; ++i) {  // src/CONV_MKL.cpp:47-63
            chk(impl->Forward3D(t1), "Forward3D");
            chk(impl->DataConstraint(t1, Ap, per, n), "DataConstraint");
            chk(impl->Backward3D(t1), "Backward3D");
            chk(impl->MergeAddData(psi, t1, -1.f / B, 1.f + 1.f / B, per), "MergeAddData");
            chk(impl->SpaceConstraint(t1, S, per, n), "SpaceConstraint");
            chk(impl->SpaceConstraint(t2, S, per, n), "SpaceConstraint");
            chk(impl->MergeAddData(psi, t2, 1.f / B, 1.f - 1.f / B, per), "MergeAddData");
            chk(impl->Forward3D(t2), "Forward3D");
            chk(impl->DataConstraint(t2, Ap, per, n), "DataConstraint");
            chk(impl->Backward3D(t2), "Backward3D");
            chk(impl->MergeAddData(t1, psi, B, 1.f, per), "MergeAddData");
            chk(impl->MergeAddData(t2, psi, -B, 1.f, per), "MergeAddData");
        }
    };

    std::vector<float> err((size_t)n, 0.f);
    const auto start = std::chrono::system_clock::now();
    chk(impl->Memcpy(t1, psi, sizeof(cf) * per), "Memcpy");
    chk(impl->Memcpy(t2, psi, sizeof(cf) * per), "Memcpy");
    if (o.Iter >= 1) {
        dm(o.Iter - 1);
        chk(impl->Memcpy(old, psi, sizeof(cf) * per), "Memcpy");
        dm(1);
        chk(impl->Memcpy(t1, psi, sizeof(cf) * per), "Memcpy");
        chk(impl->Forward3D(old), "Forward3D");
        chk(impl->Forward3D(t1), "Forward3D");
        chk(impl->ConvergeError(old, t1, err.data(), M, N, L, n), "ConvergeError");
    }
    // final projection (:78-88)
    chk(impl->Memcpy(t1, psi, sizeof(cf) * per), "Memcpy");
    chk(impl->Forward3D(t1), "Forward3D");
    chk(impl->DataConstraint(t1, const_cast<float*>(A), per, n), "DataConstraint");
    chk(impl->Backward3D(t1), "Backward3D");
    chk(impl->MergeAddData(psi, t1, -1.f / B, 1.f + 1.f / B, per), "MergeAddData");
    chk(impl->SpaceConstraint(t1, S, per, n), "SpaceConstraint");
    chk(impl->Memcpy(psi, t1, sizeof(cf) * per), "Memcpy");
    chk(impl->Sync(), "Sync");
    const auto end = std::chrono::system_clock::now();
    const std::chrono::duration<float> el = end - start;
    const std::time_t end_time = std::chrono::system_clock::to_time_t(end);
    std::cout << "Finished computation at " << std::ctime(&end_time) << "elapsed time: " << el.count() << "s\n";

    std::vector<cf> host(per);
    chk(impl->Memcpy(host.data(), psi, sizeof(cf) * per), "Memcpy");
    chk(impl->Sync(), "Sync");
    real_out.resize(per);
    for (size_t i = 0; i < host.size(); ++i) real_out[i] = host[i].real();
    err_out = err;
    alloc->deallocate(psi);
    alloc->deallocate(t1);
    alloc->deallocate(t2);
    alloc->deallocate(old);
}

}  // namespace

int main(int argc, char** argv)
{
    program_name() = "cpra_conv";
    const Options o = parse(argc, argv, 1);
    std::cout << "Conventional 3D reconstruction, HIP implementation (MI355X backend)." << std::endl;
    std::cout << "Object dimension is " << o.M << ", " << o.N << ", " << o.L << "." << std::endl;
    std::cout << "Batch size is " << o.Batch << "." << std::endl;
    std::cout << "Shrinkwrap hyper-param beta is " << o.Beta << "." << std::endl;
    std::cout << "Data constraint file is " << o.data_constraint << "." << std::endl;
    std::cout << "Space constraint file is " << o.space_constraint << "." << std::endl;
    std::cout << "Output reconstructions file is " << o.output_reconstruction << "." << std::endl;
    std::cout << "Output errors file is " << o.output_error << "." << std::endl;
    try {
        const uint64_t M = o.M, N = o.N, L = o.L, vox = M * N * L;
        const std::vector<float> A = read_floats(o.data_constraint, vox);
        const std::vector<float> S = read_floats(o.space_constraint, vox);
        std::vector<cf> init;
        if (!o.init.empty()) {
            std::vector<float> raw = read_floats(o.init, 2 * vox * (uint64_t)o.Batch);
            init.resize(vox * (uint64_t)o.Batch);
            std::memcpy(init.data(), raw.data(), raw.size() * sizeof(float));
        }
        auto impl0 = NewHIPImpl<float>(M, N, L, 1);
        cpra_hip_handle h0 = static_cast<HipImpl<float>*>(impl0.get())->handle();
        auto alloc = NewHIPAllocator();
        float* dA = (float*)alloc->allocate(sizeof(float) * vox);
        float* dS = (float*)alloc->allocate(sizeof(float) * vox);
        impl0->Memcpy(dA, (void*)A.data(), sizeof(float) * vox);
        impl0->Memcpy(dS, (void*)S.data(), sizeof(float) * vox);
        // support mask of S, !(S < 0.9) (MKL_Impl.inl:80), for every member
        const uint64_t Bmax = (uint64_t)o.Batch;
        uint8_t* mask = (uint8_t*)alloc->allocate(vox * Bmax);
        check(cpra_hip_support_from_constraint(h0, dS, mask, vox), h0, "support mask");
        for (uint64_t i = 1; i < Bmax; ++i) impl0->Memcpy(mask + i * vox, mask, vox);
        impl0->Sync();

        const int total = std::max(o.Batch, o.Total);
        int cnt = total, file_cnt = 1;
        while (cnt > 0) {
            const int n = std::min(cnt, o.Batch);
            std::cout << "Reconstrucint batch size " << n << ". Total batch left " << cnt << "." << std::endl;
            std::vector<cf> init_n;
            if (!init.empty()) init_n.assign(init.begin(), init.begin() + (size_t)vox * n);
            std::vector<float> rec, err;
            reconstruct(o, n, dA, dS, mask, init.empty() ? nullptr : &init_n, rec, err);
            for (int i = 0; i < n; ++i) {
                write_floats(o.output_reconstruction + "." + std::to_string(file_cnt), &rec[(size_t)i * vox], vox);
                write_floats(o.output_error + "." + std::to_string(file_cnt), &err[i], 1);
                ++file_cnt;
            }
            cnt -= n;
        }
        alloc->deallocate(mask);
        alloc->deallocate(dA);
        alloc->deallocate(dS);
    } catch (const std::exception& e) {
        std::cerr << "cpra_conv: " << e.what() << std::endl;
        return 1;
    }
    return EXIT_SUCCESS;
}
