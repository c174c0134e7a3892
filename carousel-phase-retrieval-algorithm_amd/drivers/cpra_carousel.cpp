// cpra_carousel.cpp -- the reference's carousel reconstruction driver
// (src/CPRA_MKL.cpp:10-232, identical to src/CPRA_CUDA.cu) on the HIP backend.
//
// Same command line (the options of include/Command_Parser.hpp:21-36, parsed
// here without Boost), same inputs (raw float32 files, data constraint and
// space constraint [P][M][N], include/detail/CPRA.inl:8-29) and same outputs
// (per batch member k = 1, 2, ...: <output_reconstruction>.<k> with the real
// part of the reconstructed projections [P][M][N] float32, and
// <output_error>.<k> with the final convergence error of each projection,
// src/CPRA_MKL.cpp:193-216).
//
// The reconstruction follows src/CPRA_MKL.cpp:10-154:
//   Step A  (:35-55)   `preit` difference-map iterations on projection 0;
//   Step B  (:58-124)  `epi` sequential sweeps over the projections: warm
//                      start from the neighbours (:63-76), t1 = t2 = psi_p,
//                      `iter` DM iterations, and in the last sweep the
//                      Fourier-domain change of the last iteration (:85-91,
//                      112-121);
//   final projection  (:126-140).
// The DM iterations (:94-110, t1 and t2 carried across iterations) run
// through the fused engine cpra_hip_dm_run (reference variant, MKL
// SpaceConstraint semantics on the mask of the float constraint, MKL
// DataConstraint with the A == 0 skip): two HBM passes per iteration instead
// of eleven primitive calls.  Pattern sizes the engine does not take (it
// takes 64, 128, 192, 256, 384, 512) run the reference's composition of
// HipImpl primitives.
//
// Deviations, each deliberate:
//  * <output_error>.<k> holds P floats; the reference writes sizeof(float)*P
//    elements of a P-element buffer (src/CPRA_MKL.cpp:212-214).
//  * P must be at least 2 (the reference's warm start reads psi_{-1} for
//    P == 1).
//  * --init FILE reads the initial guess psi [P][Batch][M][N] complex64 (batch
//    member i of a smaller last batch uses member i of each projection block),
//    --seed S seeds the uniform initialisation otherwise (the reference draws
//    from an unseeded generator).
#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
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

struct DeviceData {
    float* A = nullptr;       // [P][M][N] data constraint
    float* S = nullptr;       // [P][M][N] space constraint (float, for the final projection)
    uint8_t* mask = nullptr;  // [P][Batch][M][N] support mask of S, repeated over the batch
};

// One batch of n reconstructions (src/CPRA_MKL.cpp:10-154).  Returns the real
// parts [P][n][M][N] and the errors [P][n].
void reconstruct(const Options& o, int n, const DeviceData& dd, const std::vector<cf>* init,
                 std::vector<float>& real_out, std::vector<float>& err_out)
{
    const uint64_t M = o.M, N = o.N, P = o.P, px = M * N, per = px * n;
    const float B = o.Beta;
    auto impl = NewHIPImpl<float>(M, N, o.L, n);
    auto* hip = static_cast<HipImpl<float>*>(impl.get());
    cpra_hip_handle h = hip->handle();
    auto alloc = NewHIPAllocator();
    cf* psi = (cf*)alloc->allocate(sizeof(cf) * per * P);
    cf* t1 = (cf*)alloc->allocate(sizeof(cf) * per);
    cf* t2 = (cf*)alloc->allocate(sizeof(cf) * per);
    cf* old = (cf*)alloc->allocate(sizeof(cf) * per);
    if (init) {
        // psi_p member i <- init[p][i]
        std::vector<cf> host(per * P);
        for (uint64_t p = 0; p < P; ++p)
            std::memcpy(&host[p * per], &(*init)[p * px * o.Batch], sizeof(cf) * per);
        impl->Memcpy(psi, host.data(), sizeof(cf) * per * P);
        impl->Sync();  // `host` is released at the end of this block
    } else {
        check(cpra_hip_initialize(h, reinterpret_cast<float*>(psi), 2 * per * P, o.seed), h, "Initialize");
    }
    std::vector<float> err((size_t)P * n, 0.f);
    cpra_hip_dm_params dp;
    cpra_hip_dm_default_params(&dp);
    dp.beta = B;
    dp.textbook = 0;
    dp.space_mode = CPRA_HIP_DM_SPACE_MKL;
    dp.dc_skip_zero = 1;
    dp.use_graph = 1;
    auto chk = [&](bool ok, const char* what) {
        if (!ok) throw std::runtime_error(std::string(what) + ": " + cpra_hip_last_error(h));
    };
    // sizes outside the fused engine's {64..512}^2 run the reference's own
    // composition of primitives (src/CPRA_MKL.cpp:94-110), whose FFT passes
    // take any size
    bool fused = true;
    auto dm_compose = [&](uint64_t p, int iters) {
        cf* cur = psi + per * p;
        float* Ap = dd.A + px * p;
        float* Sp = dd.S + px * p;
        for (int i = 0; i < iters; ++i) {
            chk(impl->Forward2D(t1), "Forward2D");
            chk(impl->DataConstraint(t1, Ap, per, n), "DataConstraint");
            chk(impl->Backward2D(t1), "Backward2D");
            chk(impl->MergeAddData(cur, t1, -1.f / B, 1.f + 1.f / B, per), "MergeAddData");
            chk(impl->SpaceConstraint(t1, Sp, per, n), "SpaceConstraint");
            chk(impl->SpaceConstraint(t2, Sp, per, n), "SpaceConstraint");
            chk(impl->MergeAddData(cur, t2, 1.f / B, 1.f - 1.f / B, per), "MergeAddData");
            chk(impl->Forward2D(t2), "Forward2D");
            chk(impl->DataConstraint(t2, Ap, per, n), "DataConstraint");
            chk(impl->Backward2D(t2), "Backward2D");
            chk(impl->MergeAddData(t1, cur, B, 1.f, per), "MergeAddData");
            chk(impl->MergeAddData(t2, cur, -B, 1.f, per), "MergeAddData");
        }
    };
    auto dm = [&](uint64_t p, int iters) {
        if (iters <= 0) return;
        if (fused) {
            dp.iters = iters;
            const int rc = cpra_hip_dm_run(h, psi + per * p, t1, t2, dd.A + px * p, 1, dd.mask + per * p, &dp);
            if (rc == CPRA_HIP_ERR_UNSUPPORTED) {
                fused = false;
                std::cout << "Pattern size " << M << " x " << N
                          << " runs the primitive composition (the fused engine takes 64..512)." << std::endl;
            } else {
                check(rc, h, "difference map");
                return;
            }
        }
        dm_compose(p, iters);
    };

    const auto start = std::chrono::system_clock::now();
    // Step A (:35-55): projection 0, t1 = t2 = psi_0
    chk(impl->Memcpy(t1, psi, sizeof(cf) * per), "Memcpy");
    chk(impl->Memcpy(t2, psi, sizeof(cf) * per), "Memcpy");
    dm(0, o.PreIter);
    // Step B (:58-124)
    for (int e = 0; e < o.Epi; ++e) {
        for (uint64_t p = 0; p < P; ++p) {
            cf* cur = psi + per * p;
            if (p == P - 1) {
                chk(impl->MergeAddData(psi + per * (p - 1), cur, 0.5f, 0.5f, per), "MergeAddData");
            } else if (p == 0) {
                chk(impl->MergeAddData(psi + per * (p + 1), cur, 0.5f, 0.5f, per), "MergeAddData");
            } else {
                chk(impl->MergeAddData(psi + per * (p - 1), cur, 0.5f, 0.f, per), "MergeAddData");
                chk(impl->MergeAddData(psi + per * (p + 1), cur, 0.5f, 1.f, per), "MergeAddData");
            }
            chk(impl->Memcpy(t1, cur, sizeof(cf) * per), "Memcpy");
            chk(impl->Memcpy(t2, cur, sizeof(cf) * per), "Memcpy");
            if (e == o.Epi - 1 && o.Iter >= 1) {
                dm(p, o.Iter - 1);
                chk(impl->Memcpy(old, cur, sizeof(cf) * per), "Memcpy");
                dm(p, 1);
                // the Fourier-domain change of the last iteration (:112-121);
                // t1 is free again (the reference reuses it the same way)
                chk(impl->Memcpy(t1, cur, sizeof(cf) * per), "Memcpy");
                chk(impl->Forward2D(old), "Forward2D");
                chk(impl->Forward2D(t1), "Forward2D");
                chk(impl->ConvergeError(old, t1, err.data() + (size_t)n * p, M, N, 1, n), "ConvergeError");
            } else {
                dm(p, o.Iter);
            }
        }
    }
    // final projection (:126-140)
    for (uint64_t p = 0; p < P; ++p) {
        cf* cur = psi + per * p;
        chk(impl->Memcpy(t1, cur, sizeof(cf) * per), "Memcpy");
        chk(impl->Forward2D(t1), "Forward2D");
        chk(impl->DataConstraint(t1, dd.A + px * p, per, n), "DataConstraint");
        chk(impl->Backward2D(t1), "Backward2D");
        chk(impl->MergeAddData(cur, t1, -1.f / B, 1.f + 1.f / B, per), "MergeAddData");
        chk(impl->SpaceConstraint(t1, dd.S + px * p, per, n), "SpaceConstraint");
        chk(impl->Memcpy(cur, t1, sizeof(cf) * per), "Memcpy");
    }
    chk(impl->Sync(), "Sync");
    const auto end = std::chrono::system_clock::now();
    const std::chrono::duration<float> el = end - start;
    const std::time_t end_time = std::chrono::system_clock::to_time_t(end);
    std::cout << "Finished computation at " << std::ctime(&end_time) << "elapsed time: " << el.count() << "s\n";

    std::vector<cf> host(per * P);
    chk(impl->Memcpy(host.data(), psi, sizeof(cf) * per * P), "Memcpy");
    chk(impl->Sync(), "Sync");
    real_out.resize(per * P);
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
    program_name() = "cpra_carousel";
    const Options o = parse(argc, argv, 2);
    std::cout << "CPRA HIP implementation (MI355X backend)." << std::endl;
    std::cout << "Object dimension is " << o.M << ", " << o.N << ", " << o.L << "." << std::endl;
    std::cout << "Number of projected object is " << o.P << "." << std::endl;
    std::cout << "Batch size is " << o.Batch << "." << std::endl;
    std::cout << "Shrinkwrap hyper-param beta is " << o.Beta << "." << std::endl;
    std::cout << "Data constraint file is " << o.data_constraint << "." << std::endl;
    std::cout << "Space constraint file is " << o.space_constraint << "." << std::endl;
    std::cout << "Output reconstructions file is " << o.output_reconstruction << "." << std::endl;
    std::cout << "Output errors file is " << o.output_error << "." << std::endl;
    try {
        const uint64_t M = o.M, N = o.N, P = o.P, px = M * N;
        const std::vector<float> A = read_floats(o.data_constraint, px * P);
        const std::vector<float> S = read_floats(o.space_constraint, px * P);
        std::vector<cf> init;
        if (!o.init.empty()) {
            std::vector<float> raw = read_floats(o.init, 2 * px * P * (uint64_t)o.Batch);
            init.resize(px * P * (uint64_t)o.Batch);
            std::memcpy(init.data(), raw.data(), raw.size() * sizeof(float));
        }
        // device copies of the constraints and the support mask
        auto impl0 = NewHIPImpl<float>(M, N, 1, 1);
        cpra_hip_handle h0 = static_cast<HipImpl<float>*>(impl0.get())->handle();
        auto alloc = NewHIPAllocator();
        DeviceData dd;
        dd.A = (float*)alloc->allocate(sizeof(float) * px * P);
        dd.S = (float*)alloc->allocate(sizeof(float) * px * P);
        const uint64_t Bmax = (uint64_t)o.Batch;
        dd.mask = (uint8_t*)alloc->allocate(px * P * Bmax);
        impl0->Memcpy(dd.A, (void*)A.data(), sizeof(float) * px * P);
        impl0->Memcpy(dd.S, (void*)S.data(), sizeof(float) * px * P);
        // mask[p][i] = !(S_p < 0.9) for every batch member i
        uint8_t* m1 = (uint8_t*)alloc->allocate(px * P);
        check(cpra_hip_support_from_constraint(h0, dd.S, m1, px * P), h0, "support mask");
        for (uint64_t p = 0; p < P; ++p)
            for (uint64_t i = 0; i < Bmax; ++i)
                impl0->Memcpy(dd.mask + (p * Bmax + i) * px, m1 + p * px, px);
        impl0->Sync();

        const int total = std::max(o.Batch, o.Total);
        int cnt = total, file_cnt = 1;
        std::vector<float> per_rec(px * P), per_err(P);
        while (cnt > 0) {
            const int n = std::min(cnt, o.Batch);
            std::cout << "Reconstrucint batch size " << n << ". Total batch left " << cnt << "." << std::endl;
            // the mask rows of a batch of n: [p][i] for i < n of the Bmax-wide layout
            DeviceData db = dd;
            uint8_t* mb = nullptr;
            if ((uint64_t)n != Bmax) {
                mb = (uint8_t*)alloc->allocate(px * P * n);
                for (uint64_t p = 0; p < P; ++p)
                    impl0->Memcpy(mb + p * px * n, dd.mask + p * px * Bmax, px * n);
                impl0->Sync();
                db.mask = mb;
            }
            std::vector<float> rec, err;
            reconstruct(o, n, db, init.empty() ? nullptr : &init, rec, err);
            if (mb) alloc->deallocate(mb);
            for (int i = 0; i < n; ++i) {
                for (uint64_t p = 0; p < P; ++p) {
                    per_err[p] = err[p * n + i];
                    std::memcpy(&per_rec[p * px], &rec[(p * n + i) * px], sizeof(float) * px);
                }
                write_floats(o.output_reconstruction + "." + std::to_string(file_cnt), per_rec.data(), px * P);
                write_floats(o.output_error + "." + std::to_string(file_cnt), per_err.data(), P);
                ++file_cnt;
            }
            cnt -= n;
        }
        alloc->deallocate(m1);
        alloc->deallocate(dd.A);
        alloc->deallocate(dd.S);
        alloc->deallocate(dd.mask);
    } catch (const std::exception& e) {
        std::cerr << "cpra_carousel: " << e.what() << std::endl;
        return 1;
    }
    return EXIT_SUCCESS;
}
