// amploss.hip -- amplitude loss L = sum_k (|F x|_k - A_k)^2 (F = Forward2D /
// Forward3D, centred, unnormalised), its hand-written adjoint and fused
// gradient descent.
//
// grad = dL/d(re x) + i dL/d(im x) = 2 F^H(F x - A (F x)/|F x|), with the
// residual taken as 0 where |F x| = 0 (no A == 0 skip: the loss there is A^2
// whatever the phase).  F^H of the centred transform is its unnormalised
// centred inverse, so the adjoint is a residual kernel between a forward and an
// inverse transform built from the same FFT passes.  The gradient is
// recomputed from x at every call; no transform of x is kept between calls.
//
// Fused 2D gradient descent (FP32, fast sizes): with the iterate kept in the
// row-spectral domain Y = DFT_r(x), the step x <- x - mu*grad is linear, so
// Y_new = IDFT_c(X/M - 2*mu*N*R) with X = DFT_c(Y) and R the residual of X:
// one column pass per step (read Y, read A, write Y: 20 B/px) that also sums
// the loss of the incoming iterate.
#include <cmath>
#include <vector>

#include "common.hpp"
#include "fft_core.hpp"
#include "plans.hpp"

namespace cpra {

// XCD-aware tile order: consecutive workgroups go round-robin over the 8 XCDs,
// so XCD x runs tiles [x*G/8, (x+1)*G/8) in order (G a multiple of 8).
__device__ __forceinline__ unsigned gd_tile(int xcd_remap)
{
    const unsigned b = blockIdx.x;
    return xcd_remap ? (b & 7u) * (gridDim.x >> 3) + (b >> 3) : b;
}

// Residual and loss partials over a batch of spectra X [B][per]; A has
// amp_batch planes of `per` values (pattern b uses A[b % amp_batch]).  Each
// block of the y grid handles one pattern; partials[b*gridDim.x + blk].
template <typename T>
__global__ __launch_bounds__(256) void k_amp_residual(c2<T>* __restrict__ X, const T* __restrict__ A,
                                                      int64_t per, int64_t amp_batch, int write_res,
                                                      double* __restrict__ partials)
{
    const int64_t b = blockIdx.y;
    c2<T>* Xb = X + b * per;
    const T* Ab = A + (b % amp_batch) * per;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < per;
         i += (int64_t)gridDim.x * blockDim.x) {
        const c2<T> x = Xb[i];
        const T a = Ab[i];
        const T m = sqrt(x.x * x.x + x.y * x.y);
        const T d = m - a;
        acc += (double)d * (double)d;
        if (write_res) {
            const T s = m > (T)0 ? (T)1 - a / m : (T)0;
            Xb[i] = mk<c2<T>>(x.x * s, x.y * s);
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[b * gridDim.x + blockIdx.x] = red[0];
}

__global__ void k_sum_partials(const double* __restrict__ part, int nblk, int64_t B,
                               double* __restrict__ out)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double s = 0.0;
    for (int k = 0; k < nblk; ++k) s += part[b * nblk + k];
    out[b] = s;
}

// Fused GD pass along the outermost transformed axis (see the file header).
// A "column" here is a line of length M with element stride W: 2D columns
// (W = N) or 3D z-lines (W = M*N, the state then being DFT over (y, x)).
// Block = C consecutive lines of one pattern.
//  GD_STEP:      Y <- IDFT(X/M - step*R)
//  GD_STEP_LOSS: as GD_STEP, plus loss partials of the incoming iterate
//  GD_LOSS:      loss partials only (Y untouched)
enum GdMode { GD_STEP = 0, GD_STEP_LOSS = 1, GD_LOSS = 2 };

// TILED (2D, square patterns): Y and A in the tile-major layout (each
// block's M x C tile contiguous, row stride C; Tiled<M> in engine_kernels.hpp).
template <int M, int MODE, bool TILED = false>
__global__ __launch_bounds__(SizeCfg<M>::TPB) void k_gd_cols(c2<float>* Y, const float* __restrict__ A,
                                                            int64_t amp_batch, int64_t W, float invM,
                                                            float step, float* __restrict__ loss_part,
                                                            int xcd_remap)
{
    using V = c2<float>;
    using PL = typename SizeCfg<M>::P;
    constexpr int E = PL::E, TT = PL::T, C = SizeCfg<M>::C, TPB = SizeCfg<M>::TPB;
    __shared__ V lds[M * C];
    __shared__ float red[TPB];
    const int tid = threadIdx.x;
    const int c = tid % C, t = tid / C;
    const unsigned tile = gd_tile(xcd_remap);
    const unsigned tiles_per_pattern = (unsigned)(W / C);
    const unsigned b = tile / tiles_per_pattern;
    const int64_t col0 = (int64_t)(tile - b * tiles_per_pattern) * C;
    const size_t pat = (size_t)M * (size_t)W;
    const size_t blk0 = TILED ? (size_t)tile * M * C : b * pat + col0;
    const size_t ablk0 = (size_t)(b % (unsigned)amp_batch) * pat +
                         (TILED ? (size_t)(tile - b * tiles_per_pattern) * M * C : (size_t)col0);
    c2<float>* Yp = Y + blk0;
    const float* Ap = A + ablk0;
    // element m of this lane sits at line position t + rot(m)*T: a uniform
    // base per m plus one per-lane offset (row stride W, or C when tiled)
    const unsigned RS = TILED ? (unsigned)C : (unsigned)W;
    const unsigned toff = (unsigned)t * RS + (unsigned)c;  // < T*W <= 2^24
    auto rowbase = [&](int m) { return (size_t)(PL::rot(m) * TT) * (size_t)RS; };
    V v[E];
    float av[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = (Yp + rowbase(m))[toff];
    });
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        av[m] = (Ap + rowbase(m))[toff];
    });
    LdsCol<V, C, false> acc{lds + c};
    fft_line<PL, -1>(v, t, g_tw<M>(), acc);
    float loss = 0.f;
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const V x = v[m];
        const float a = av[m];
        const float mag = sqrtf(x.x * x.x + x.y * x.y);
        if constexpr (MODE != GD_STEP) {
            const float d = mag - a;
            loss += d * d;
        }
        if constexpr (MODE != GD_LOSS) {
            // residual R = X - A X/|X| = r X (0 where X = 0)
            const float r = mag > 0.f ? 1.f - a / mag : 0.f;
            v[m] = cscale(x, invM - step * r);
        }
    });
    if constexpr (MODE != GD_LOSS) {
        fft_line<PL, +1>(v, t, g_tw<M>(), acc);
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            (Yp + rowbase(m))[toff] = v[m];
        });
    }
    if constexpr (MODE != GD_STEP) {
        red[tid] = loss;
        __syncthreads();
        for (int s = TPB / 2; s > 0; s >>= 1) {
            if (tid < s) red[tid] += red[tid + s];
            __syncthreads();
        }
        if (tid == 0) loss_part[tile] = red[0];
    }
}

__global__ void k_loss_fin(const float* __restrict__ part, int tiles_per_pattern, int64_t B,
                           double* __restrict__ out)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double s = 0.0;
    for (int k = 0; k < tiles_per_pattern; ++k) s += part[b * tiles_per_pattern + k];
    out[b] = s;
}

// forward transform of ws (dims) with the shift folded, optional residual
// and inverse (unnormalised) scaled by `inv_scale`.
struct Dims {
    int rank;  // 2 or 3
    int L, M, N;
    int64_t B;
};

static void forward_passes(cpra_hip_ctx* h, void* w, const Dims& d, bool dbl)
{
    if (d.rank == 2) {
        fft_pass_contig(h, w, d.N, d.B * d.M, -1, true, 1.0, dbl);
        fft_pass_strided(h, w, d.M, d.B * d.N, d.N, d.N, (int64_t)d.M * d.N, -1, true, 1.0, dbl);
    } else {
        const int64_t MN = (int64_t)d.M * d.N;
        fft_pass_contig(h, w, d.N, d.B * d.L * d.M, -1, true, 1.0, dbl);
        fft_pass_strided(h, w, d.M, d.B * d.L * d.N, d.N, d.N, MN, -1, true, 1.0, dbl);
        fft_pass_strided(h, w, d.L, d.B * MN, MN, MN, (int64_t)d.L * MN, -1, true, 1.0, dbl);
    }
}

static void inverse_passes(cpra_hip_ctx* h, void* w, const Dims& d, bool dbl, double scale)
{
    if (d.rank == 2) {
        fft_pass_contig(h, w, d.N, d.B * d.M, +1, true, 1.0, dbl);
        fft_pass_strided(h, w, d.M, d.B * d.N, d.N, d.N, (int64_t)d.M * d.N, +1, true, scale, dbl);
    } else {
        const int64_t MN = (int64_t)d.M * d.N;
        fft_pass_contig(h, w, d.N, d.B * d.L * d.M, +1, true, 1.0, dbl);
        fft_pass_strided(h, w, d.M, d.B * d.L * d.N, d.N, d.N, MN, +1, true, 1.0, dbl);
        fft_pass_strided(h, w, d.L, d.B * MN, MN, MN, (int64_t)d.L * MN, +1, true, scale, dbl);
    }
}

// loss (and gradient if grad != NULL) of x; loss_host receives B doubles.
static void amp_loss_grad(cpra_hip_ctx* h, const void* x, const void* amp, uint64_t amp_batch,
                          void* grad, double* loss_host, const Dims& d)
{
    const bool dbl = h->dtype == CPRA_HIP_F64;
    const size_t es = dbl ? sizeof(c2<double>) : sizeof(c2<float>);
    const int64_t per = (int64_t)d.L * d.M * d.N;
    void* w = grad;
    if (!w) {
        h->ws_z.ensure(es * (size_t)per * d.B);
        w = h->ws_z.ptr;
    }
    hipStream_t st = h->stream;
    CPRA_CHECK_HIP(hipMemcpyAsync(w, x, es * (size_t)per * d.B, hipMemcpyDeviceToDevice, st));
    forward_passes(h, w, d, dbl);
    const unsigned nblk = grid_1d(per, 256, 256);
    CPRA_ARG(d.B <= 65535, "amplitude loss: batch > 65535");
    h->ws_part.ensure(sizeof(double) * (nblk * (size_t)d.B + (size_t)d.B));
    double* part = static_cast<double*>(h->ws_part.ptr);
    double* tot = part + nblk * (size_t)d.B;
    const int wr = grad ? 1 : 0;
    if (dbl)
        hipLaunchKernelGGL(k_amp_residual<double>, dim3(nblk, (unsigned)d.B), dim3(256), 0, st,
                           static_cast<c2<double>*>(w), static_cast<const double*>(amp), per, (int64_t)amp_batch, wr, part);
    else
        hipLaunchKernelGGL(k_amp_residual<float>, dim3(nblk, (unsigned)d.B), dim3(256), 0, st,
                           static_cast<c2<float>*>(w), static_cast<const float*>(amp), per, (int64_t)amp_batch, wr, part);
    check_launch("k_amp_residual");
    hipLaunchKernelGGL(k_sum_partials, dim3(grid_1d(d.B, 64)), dim3(64), 0, st, part, (int)nblk, d.B, tot);
    check_launch("k_sum_partials");
    if (grad) inverse_passes(h, w, d, dbl, 2.0);  // 2 F^H R
    if (loss_host) {
        CPRA_CHECK_HIP(hipMemcpyAsync(loss_host, tot, sizeof(double) * d.B, hipMemcpyDeviceToHost, st));
        CPRA_CHECK_HIP(hipStreamSynchronize(st));
    }
}

template <typename F> static int aguard(cpra_hip_ctx* h, F&& f)
{
    try {
        return f();
    } catch (const Status& s) {
        if (h) h->last_error = s.msg;
        return s.code;
    } catch (const std::exception& e) {
        if (h) h->last_error = e.what();
        return CPRA_HIP_ERR_HIP;
    }
}

static Dims dims2(cpra_hip_ctx* h) { return Dims{2, 1, (int)h->m, (int)h->n, (int64_t)h->batch}; }
static Dims dims3(cpra_hip_ctx* h) { return Dims{3, (int)h->l, (int)h->m, (int)h->n, (int64_t)h->batch}; }

static void check_even(const Dims& d, const char* who)
{
    if ((d.M & 1) || (d.N & 1) || (d.rank == 3 && (d.L & 1)))
        fail(CPRA_HIP_ERR_UNSUPPORTED, "%s: sizes must be even (centred transform)", who);
}

// Fused gradient descent, 2D (lines along M, W = N) or 3D (lines along L,
// W = M*N).  Entry: the inner axes are transformed in place (Y); exit: back.
template <int LEN, bool TL>
static void gd_launch_t(hipStream_t st, int mode, unsigned grid, c2<float>* Y, const float* A,
                        int64_t amp_batch, int64_t W, float invM, float step, float* part, int remap)
{
    if (mode == GD_STEP)
        hipLaunchKernelGGL((k_gd_cols<LEN, GD_STEP, TL>), dim3(grid), dim3(SizeCfg<LEN>::TPB), 0, st, Y, A, amp_batch, W, invM, step, part, remap);
    else if (mode == GD_STEP_LOSS)
        hipLaunchKernelGGL((k_gd_cols<LEN, GD_STEP_LOSS, TL>), dim3(grid), dim3(SizeCfg<LEN>::TPB), 0, st, Y, A, amp_batch, W, invM, step, part, remap);
    else
        hipLaunchKernelGGL((k_gd_cols<LEN, GD_LOSS, TL>), dim3(grid), dim3(SizeCfg<LEN>::TPB), 0, st, Y, A, amp_batch, W, invM, step, part, remap);
}

template <int LEN>
static void gd_launch(hipStream_t st, int mode, unsigned grid, c2<float>* Y, const float* A,
                      int64_t amp_batch, int64_t W, float invM, float step, float* part, int remap,
                      bool tiled)
{
    if (tiled) gd_launch_t<LEN, true>(st, mode, grid, Y, A, amp_batch, W, invM, step, part, remap);
    else gd_launch_t<LEN, false>(st, mode, grid, Y, A, amp_batch, W, invM, step, part, remap);
}

static int line_tile(int len)
{
    switch (len) {
    case 64: return SizeCfg<64>::C;
    case 128: return SizeCfg<128>::C;
    case 192: return SizeCfg<192>::C;
    case 256: return SizeCfg<256>::C;
    case 384: return SizeCfg<384>::C;
    case 512: return SizeCfg<512>::C;
    }
    return 0;
}

static void gd_run(cpra_hip_ctx* h, void* x, const float* amp, uint64_t amp_batch, int iters, float mu,
                   double* loss_host, const Dims& d, const char* who)
{
    CPRA_ARG(amp_batch >= 1 && (uint64_t)d.B % amp_batch == 0, "%s: amp_batch must divide batch", who);
    CPRA_ARG(iters >= 0, "%s: iters < 0", who);
    CPRA_ARG(std::isfinite(mu), "%s: step size not finite", who);
    if (h->dtype != CPRA_HIP_F32) fail(CPRA_HIP_ERR_UNSUPPORTED, "%s: the fused descent runs in FP32", who);
    check_even(d, who);
    // the fused pass runs along the outermost transformed axis
    const int len = d.rank == 2 ? d.M : d.L;
    const int64_t W = d.rank == 2 ? (int64_t)d.N : (int64_t)d.M * d.N;
    const int C = line_tile(len);
    if (!C) fail(CPRA_HIP_ERR_UNSUPPORTED, "%s: outer length %d not in {64,128,192,256,384,512}", who, len);
    if (W % C) fail(CPRA_HIP_ERR_UNSUPPORTED, "%s: inner size %lld not a multiple of %d", who, (long long)W, C);
    CPRA_ARG((uint64_t)W * 32u < (1ull << 32), "%s: inner plane too large", who);
    const int64_t B = d.B;
    const int64_t tiles = B * (W / C);
    CPRA_ARG(tiles < (1ll << 31), "%s: batch too large", who);
    hipStream_t st = h->stream;
    c2<float>* Y = static_cast<c2<float>*>(x);
    const float* Ause = amp;
    // 2D square patterns: the state lives tile-major in a workspace (the
    // column pass then reads and writes contiguous tiles), with a tile-major
    // copy of the amplitudes; the tuning key tiled=0 keeps it in x
    const bool tiled = d.rank == 2 && d.M == d.N && engine_tiled_state(h);
    // enter: Y = centred DFT over the inner axes
    if (tiled) {
        const size_t px = (size_t)d.M * d.N;
        h->ws_state.ensure(sizeof(c2<float>) * px * (size_t)B);
        h->ws_amp_t.ensure(sizeof(float) * px * amp_batch);
        Y = static_cast<c2<float>*>(h->ws_state.ptr);
        Ause = static_cast<const float*>(h->ws_amp_t.ptr);
        engine_amp_retile(st, d.N, (int64_t)amp_batch, amp, static_cast<float*>(h->ws_amp_t.ptr));
        engine_retile(st, d.N, B, -1, x, Y);
    } else if (d.rank == 2) {
        fft_pass_contig(h, x, d.N, B * d.M, -1, true, 1.0, false);
    } else {
        fft_pass_contig(h, x, d.N, B * d.L * d.M, -1, true, 1.0, false);
        fft_pass_strided(h, x, d.M, B * d.L * d.N, d.N, d.N, (int64_t)d.M * d.N, -1, true, 1.0, false);
    }
    const size_t part_bytes = ((sizeof(float) * (size_t)tiles + 255) / 256) * 256;
    h->ws_part.ensure(part_bytes + sizeof(double) * (size_t)B);
    float* part = static_cast<float*>(h->ws_part.ptr);
    double* tot = reinterpret_cast<double*>(static_cast<char*>(h->ws_part.ptr) + part_bytes);
    const float invM = 1.0f / (float)len;
    const float step = (float)(2.0 * (double)mu * (double)W);  // 2 mu * (size of the inner axes)
    const unsigned grid = (unsigned)tiles;
    // XCD-aware tile order helps the strided row-major tiles; contiguous
    // tiles stream best in dispatch order
    const int remap = (!tiled && grid % 8u == 0u) ? 1 : 0;
    KernelTimer kt;
    auto launch = [&](int mode, float stp) {
        kt.begin(h, st, KT_GD);
        switch (len) {
        case 64: gd_launch<64>(st, mode, grid, Y, Ause, (int64_t)amp_batch, W, invM, stp, part, remap, tiled); break;
        case 128: gd_launch<128>(st, mode, grid, Y, Ause, (int64_t)amp_batch, W, invM, stp, part, remap, tiled); break;
        case 192: gd_launch<192>(st, mode, grid, Y, Ause, (int64_t)amp_batch, W, invM, stp, part, remap, tiled); break;
        case 256: gd_launch<256>(st, mode, grid, Y, Ause, (int64_t)amp_batch, W, invM, stp, part, remap, tiled); break;
        case 384: gd_launch<384>(st, mode, grid, Y, Ause, (int64_t)amp_batch, W, invM, stp, part, remap, tiled); break;
        case 512: gd_launch<512>(st, mode, grid, Y, Ause, (int64_t)amp_batch, W, invM, stp, part, remap, tiled); break;
        }
        check_launch("k_gd_cols");
        kt.end(h, st);
    };
    for (int i = 0; i < iters; ++i) launch(GD_STEP, step);
    if (loss_host) {
        launch(GD_LOSS, 0.f);  // loss of the final iterate
        hipLaunchKernelGGL(k_loss_fin, dim3(grid_1d(B, 64)), dim3(64), 0, st, part, (int)(W / C), B, tot);
        check_launch("k_loss_fin");
    }
    // leave: inverse over the inner axes, scaled by 1/(their size)
    if (tiled) {
        engine_retile(st, d.N, B, +1, Y, x);
    } else if (d.rank == 2) {
        fft_pass_contig(h, x, d.N, B * d.M, +1, true, 1.0 / (double)d.N, false);
    } else {
        fft_pass_contig(h, x, d.N, B * d.L * d.M, +1, true, 1.0, false);
        fft_pass_strided(h, x, d.M, B * d.L * d.N, d.N, d.N, (int64_t)d.M * d.N, +1, true,
                         1.0 / ((double)d.M * d.N), false);
    }
    if (loss_host) {
        CPRA_CHECK_HIP(hipMemcpyAsync(loss_host, tot, sizeof(double) * B, hipMemcpyDeviceToHost, st));
        CPRA_CHECK_HIP(hipStreamSynchronize(st));
    }
    kt.collect(h);
}

}  // namespace cpra

using namespace cpra;

extern "C" {

int cpra_hip_amp_loss(cpra_hip_handle h, const void* x, const void* amp, uint64_t amp_batch,
                      double* loss_host)
{
    return aguard(h, [&]() -> int {
        CPRA_ARG(h && x && amp && loss_host, "cpra_hip_amp_loss: NULL argument");
        CPRA_ARG(amp_batch >= 1 && h->batch % amp_batch == 0, "cpra_hip_amp_loss: amp_batch must divide batch");
        const Dims d = dims2(h);
        check_even(d, "cpra_hip_amp_loss");
        amp_loss_grad(h, x, amp, amp_batch, nullptr, loss_host, d);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_amp_grad(cpra_hip_handle h, const void* x, const void* amp, uint64_t amp_batch,
                      void* grad, double* loss_host)
{
    return aguard(h, [&]() -> int {
        CPRA_ARG(h && x && amp && grad, "cpra_hip_amp_grad: NULL argument");
        CPRA_ARG(grad != x, "cpra_hip_amp_grad: grad must not alias x");
        CPRA_ARG(amp_batch >= 1 && h->batch % amp_batch == 0, "cpra_hip_amp_grad: amp_batch must divide batch");
        const Dims d = dims2(h);
        check_even(d, "cpra_hip_amp_grad");
        amp_loss_grad(h, x, amp, amp_batch, grad, loss_host, d);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_amp_loss_3d(cpra_hip_handle h, const void* x, const void* amp, uint64_t amp_batch,
                         double* loss_host)
{
    return aguard(h, [&]() -> int {
        CPRA_ARG(h && x && amp && loss_host, "cpra_hip_amp_loss_3d: NULL argument");
        CPRA_ARG(amp_batch >= 1 && h->batch % amp_batch == 0, "cpra_hip_amp_loss_3d: amp_batch must divide batch");
        const Dims d = dims3(h);
        check_even(d, "cpra_hip_amp_loss_3d");
        amp_loss_grad(h, x, amp, amp_batch, nullptr, loss_host, d);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_amp_grad_3d(cpra_hip_handle h, const void* x, const void* amp, uint64_t amp_batch,
                         void* grad, double* loss_host)
{
    return aguard(h, [&]() -> int {
        CPRA_ARG(h && x && amp && grad, "cpra_hip_amp_grad_3d: NULL argument");
        CPRA_ARG(grad != x, "cpra_hip_amp_grad_3d: grad must not alias x");
        CPRA_ARG(amp_batch >= 1 && h->batch % amp_batch == 0, "cpra_hip_amp_grad_3d: amp_batch must divide batch");
        const Dims d = dims3(h);
        check_even(d, "cpra_hip_amp_grad_3d");
        amp_loss_grad(h, x, amp, amp_batch, grad, loss_host, d);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_amp_gd_run(cpra_hip_handle h, void* x, const float* amp, uint64_t amp_batch,
                        int iters, float mu, double* loss_host)
{
    return aguard(h, [&]() -> int {
        CPRA_ARG(h && x && amp, "cpra_hip_amp_gd_run: NULL argument");
        gd_run(h, x, amp, amp_batch, iters, mu, loss_host, dims2(h), "cpra_hip_amp_gd_run");
        return CPRA_HIP_OK;
    });
}

int cpra_hip_amp_gd_run_3d(cpra_hip_handle h, void* x, const float* amp, uint64_t amp_batch,
                           int iters, float mu, double* loss_host)
{
    return aguard(h, [&]() -> int {
        CPRA_ARG(h && x && amp, "cpra_hip_amp_gd_run_3d: NULL argument");
        gd_run(h, x, amp, amp_batch, iters, mu, loss_host, dims3(h), "cpra_hip_amp_gd_run_3d");
        return CPRA_HIP_OK;
    });
}

}  // extern "C"
