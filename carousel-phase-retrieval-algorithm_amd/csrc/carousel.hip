// carousel.hip -- 3D carousel reconstruction: batched 2D HIO on P projections
// made consistent through one 3D object every outer iteration.
//
// Projection p (angle th_p about the z axis) is the sum of the object along the
// beam; by the Fourier-slice theorem its centred 2D DFT is the central slice of
// the object's centred 3D DFT that merge3d.hip samples (rows u along z, columns
// v along (cos th, sin th)).  One outer iteration:
//   1. `inner` HIO iterations on every projection (measured amplitudes, 2D
//      support, real-positive inside);
//   2. F_p = Forward2D(x_p);
//   3. V = R^T(F) / R^T(1)                       (gather-trilinear merge; the
//      footprint list is built once per run, each merge is a plain gather);
//   4. v = Backward3D(V); v <- (Re v)^+ * S3 (Im v <- 0); V = Forward3D(v);
//   5. F_p = R V                                  (re-projection);
//   6. x_p = Backward2D(F_p).
// All projections advance together (the Jacobi form of the reference's
// sequential sweep, src/CPRA_MKL.cpp:58-124).
//
// MI355X mapping: the projections stay in the row-spectral domain
// Y = DFT_rows(x) for the whole run, so steps 2 and 6 are one column pass each
// (the row transforms cancel against the HIO engine's).  Step 4 is five passes,
// not seven: the z-line kernel runs IDFT_z, the constraint and DFT_z on each
// line in registers, between the inverse and forward (y, x) passes.  The
// outer body (12 launches for inner = 1; 10 for n <= 128, where one kernel per
// direction transforms whole z-planes in LDS) is captured in a hipGraph, up to
// ten outer iterations per graph, which the handle keeps across calls and
// replays; its working set (the projections, the transposed slices and the
// volume: 56 MB at 180 x 128^2 and 128^3) stays in the Infinity Cache.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "fft_core.hpp"
#include "plans.hpp"

namespace cpra {

// ZL_ROUNDTRIP_NO_FORWARD is a volume_constraint mode only (the z-line
// kernel runs ZL_ROUNDTRIP)
enum ZlMode { ZL_ROUNDTRIP = 0, ZL_INVERSE = 1, ZL_ROUNDTRIP_NO_FORWARD = 2 };

// The 3D support [L][W] packed for the z-line kernel, once per run: for the
// kernel's block b of C consecutive lines and plane z, C bits (bit c: line
// b*C + c holds support at z) in C/8 bytes at (b*L + z)*C/8.  The kernel's
// byte-per-voxel reads of the mask took whole 128-byte lines for C bytes of
// each (fetch 1.34 times the kernel's algorithmic bytes at 128^3); a block's
// packed support is L*C/8 contiguous bytes.
template <int C>
__global__ __launch_bounds__(256) void k_pack_zline_support(const uint8_t* __restrict__ S, int64_t W, int L,
                                                            uint8_t* __restrict__ bits)
{
    const int64_t total = W / 8 * L;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = i % (C / 8), bz = i / (C / 8);
        const int64_t z = bz % L, b = bz / L;
        const uint8_t* src = S + z * W + b * C + j * 8;
        unsigned v = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) v |= (src[k] != 0 ? 1u : 0u) << k;
        bits[i] = (uint8_t)v;
    }
}

// z-lines of the volume [L][W] (W = M*N points per plane), after the inverse
// x and y passes: IDFT_z (scale), the real-positive support projection, and
// (ZL_ROUNDTRIP) DFT_z back.  Block = C consecutive lines; Sbits: the
// support packed by k_pack_zline_support<C>.
template <int L, int MODE>
__global__ __launch_bounds__(SizeCfg<L>::TPB) void k_zline_constraint(c2<float>* V,
                                                                     const uint8_t* __restrict__ Sbits,
                                                                     int64_t W, float scale)
{
    using Vt = c2<float>;
    using PL = typename SizeCfg<L>::P;
    constexpr int E = PL::E, TT = PL::T, C = SizeCfg<L>::C;
    static_assert(C % 8 == 0, "whole bytes of packed support per plane");
    __shared__ Vt lds[L * C];
    const int tid = threadIdx.x;
    const int c = tid % C, t = tid / C;
    const int64_t col0 = (int64_t)blockIdx.x * C;
    Vt* Vp = V + col0;
    const uint8_t* Sb = Sbits + (size_t)blockIdx.x * (L * C / 8) + (unsigned)t * (C / 8) + (c >> 3);
    const unsigned toff = (unsigned)t * (unsigned)W + (unsigned)c;  // < T*W <= 2^24
    // the volume is in natural order along z; rot() places the centring shift
    auto zbase = [&](int m) { return (size_t)(PL::rot(m) * TT) * (size_t)W; };
    Vt v[E];
    uint8_t sm[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = (Vp + zbase(m))[toff];
        sm[m] = (uint8_t)((Sb[PL::rot(m) * TT * (C / 8)] >> (c & 7)) & 1);
    });
    LdsCol<Vt, C, false> acc{lds + c};
    fft_line<PL, +1>(v, t, g_tw<L>(), acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const float re = v[m].x * scale;
        v[m] = mk<Vt>(sm[m] && re > 0.f ? re : 0.f, 0.f);
    });
    if constexpr (MODE == ZL_ROUNDTRIP) fft_line<PL, -1>(v, t, g_tw<L>(), acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        (Vp + zbase(m))[toff] = v[m];
    });
}

static int zline_tile(int n)
{
    switch (n) {
    case 64: return SizeCfg<64>::C;
    case 128: return SizeCfg<128>::C;
    case 192: return SizeCfg<192>::C;
    case 256: return SizeCfg<256>::C;
    case 384: return SizeCfg<384>::C;
    case 512: return SizeCfg<512>::C;
    }
    return 0;
}

// Centred 2D transform along x and y of every z-plane of the volume [N][N][N],
// one workgroup per plane, for N <= 128, where an N x N complex64 plane fits in
// LDS (136 KiB at 128 with the padding below).  This replaces two volume
// passes with one cache round trip.
// LDS layouts, with row span SP = N + N/16 and the LdsRow padding of one value
// per 16 (pad(i) = i + i/16):
//   NAT(y, x) = y*SP + pad(x)  (x-lines contiguous)
//   TRN(y, x) = x*SP + pad(y)  (y-lines contiguous)
// The plane is staged through LDS with whole-row (512-byte) global accesses;
// loading each line straight into registers (64-byte segments) measured
// slower.  An axis step reads each line from the layout where it is
// contiguous.  The line's T threads sit in one wave, so the FFT exchanges in
// place within the line's own span with wave-level ordering only.  After a
// workgroup barrier the step writes the other layout, or its own span when no
// transpose follows.  With 8 lines x 8 threads per wave at N = 128, both the
// in-place reads and the transposed writes touch 64 distinct 8-byte slots
// modulo 512 bytes.
// XFIRST runs x then y, the order of the inverse passes; otherwise y then x,
// the order of the forward passes.  Each line goes through the same plan,
// twiddles and centring fold as in the one-axis kernels of fft.hip.
// angles per group of the fused plane transform and re-projection (the
// group's (cos, sin) table sits in LDS beside the plane)
constexpr int kPlaneExtractMaxAngles = 256;

template <int N> struct PlaneGeom {
    using PL = typename SizeCfg<N>::P;
    static constexpr int E = PL::E, TT = PL::T, TPB = N * TT, SP = N + N / 16, K = N * N / TPB;
    static_assert(64 % TT == 0 && TPB % 64 == 0 && N % 8 == 0, "plane layout");
    static __device__ __forceinline__ int pad(int i) { return i + (i >> 4); }
};

// The plane Pp (N x N) transformed along both axes in lds; the result is left
// in NAT layout.
template <int N, int DIR, bool XFIRST>
__device__ __forceinline__ void plane2d_body(const c2<float>* Pp, c2<float>* lds)
{
    using Vt = c2<float>;
    using G = PlaneGeom<N>;
    using PL = typename G::PL;
    constexpr int E = G::E, TT = G::TT, TPB = G::TPB, SP = G::SP, K = G::K;
    const int tid = threadIdx.x;
    auto pad = [](int i) { return G::pad(i); };

    Vt g[K];
    if constexpr (XFIRST) {
        // coalesced rows into NAT
        sfor<K>([&](auto kc) { constexpr int k = decltype(kc)::value; g[k] = Pp[tid + k * TPB]; });
        sfor<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int i = tid + k * TPB;
            lds[(i / N) * SP + pad(i % N)] = g[k];
        });
    } else {
        // 8 x 8 blocks (8 consecutive x per row, 8 rows per wave) into TRN
        const int wv = tid >> 6, xl = tid & 7, yl = (tid >> 3) & 7;
        auto blk_y = [&](int k) { return ((wv + k * (TPB / 64)) / (N / 8)) * 8 + yl; };
        auto blk_x = [&](int k) { return ((wv + k * (TPB / 64)) % (N / 8)) * 8 + xl; };
        sfor<K>([&](auto kc) { constexpr int k = decltype(kc)::value; g[k] = Pp[blk_y(k) * N + blk_x(k)]; });
        sfor<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            lds[blk_x(k) * SP + pad(blk_y(k))] = g[k];
        });
    }
    __syncthreads();

    const int t = tid % TT, line = tid / TT;
    Vt* lb = lds + line * SP;
    auto axis = [&](bool transpose_out) {
        Vt v[E];
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            v[m] = lb[pad(t + PL::rot(m) * TT)];
        });
        LdsRow<Vt, N, true> acc{lb};
        fft_line<PL, DIR>(v, t, g_tw<N>(), acc);
        __syncthreads();  // every exchange done before the transposed writes
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const int pos = t + PL::rot(m) * TT;
            if (transpose_out) lds[pos * SP + pad(line)] = v[m];
            else lb[pad(pos)] = v[m];
        });
        __syncthreads();
    };
    // XFIRST: NAT -> (x) -> TRN -> (y) -> NAT; else TRN -> (y) -> NAT -> (x) -> NAT
    axis(true);
    axis(XFIRST);
}

template <int N>
__device__ __forceinline__ void plane2d_store(c2<float>* Pp, const c2<float>* lds)
{
    using G = PlaneGeom<N>;
    sfor<G::K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int i = threadIdx.x + k * G::TPB;
        Pp[i] = lds[(i / N) * G::SP + G::pad(i % N)];
    });
}

template <int N, int DIR, bool XFIRST>
__global__ __launch_bounds__(N * SizeCfg<N>::P::T) void k_plane2d(c2<float>* V)
{
    __shared__ c2<float> lds[N * PlaneGeom<N>::SP];
    c2<float>* Pp = V + (size_t)blockIdx.x * (N * N);
    plane2d_body<N, DIR, XFIRST>(Pp, lds);
    plane2d_store<N>(Pp, lds);
}

// The forward (y, x) plane transform fused with the re-projection of the
// plane: block (u, angle group g of `groups`) transforms plane z = u in LDS
// and samples its angles' slice rows u from there (the bilinear weights and
// FP64 geometry of k_extract_plane, so the same bits), which the separate
// extraction would read back from V.  Each group transforms the plane itself:
// the groups put the CUs that one block per plane leaves idle to work on the
// sampling.  V is only read: the next outer iteration's merge overwrites it,
// and the last outer iteration of a run, whose V is the run's result, runs
// the forward plane kernel and the separate extraction instead.  (A write of
// the plane by one group would race with another group of the same plane
// that has not loaded it yet.)
template <int N>
__global__ __launch_bounds__(N * SizeCfg<N>::P::T) void k_plane2d_extract(
    const c2<float>* __restrict__ V, const double* __restrict__ trig, int P, int groups,
    c2<float>* __restrict__ out)
{
    using Vt = c2<float>;
    using G = PlaneGeom<N>;
    __shared__ Vt lds[N * G::SP];
    __shared__ double tl[2 * kPlaneExtractMaxAngles];
    const int u = (int)blockIdx.x / groups, g = (int)blockIdx.x % groups;
    const int p0 = (int)((int64_t)g * P / groups), p1 = (int)((int64_t)(g + 1) * P / groups);
    for (int i = threadIdx.x; i < 2 * (p1 - p0); i += blockDim.x) tl[i] = trig[2 * p0 + i];
    plane2d_body<N, -1, false>(V + (size_t)u * (N * N), lds);  // ends with a barrier
    const double c = (double)(N / 2);
    for (int i = threadIdx.x; i < (p1 - p0) * N; i += blockDim.x) {
        const int pl = i / N, v = i % N;
        const int p = p0 + pl;
        const double pxd = c + ((double)v - c) * tl[2 * pl], pyd = c + ((double)v - c) * tl[2 * pl + 1];
        const double fxd = floor(pxd), fyd = floor(pyd);
        const float fx = (float)fxd, fy = (float)fyd;
        const float ax = (float)(pxd - fxd), ay = (float)(pyd - fyd);
        Vt acc = mk<Vt>(0.f, 0.f);
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dxx = 0; dxx < 2; ++dxx) {
                const int xi = (int)fx + dxx, yi = (int)fy + dy;
                const float w = (dxx ? ax : 1.f - ax) * (dy ? ay : 1.f - ay);
                if (xi >= 0 && xi < N && yi >= 0 && yi < N && w != 0.f) {
                    const Vt a = lds[yi * G::SP + G::pad(xi)];
                    acc = mk<Vt>(acc.x + w * a.x, acc.y + w * a.y);
                }
            }
        out[((int64_t)p * N + u) * N + v] = acc;
    }
}

// true when the fused kernel covers (n, P): n <= 128 planes in LDS and at
// most kPlaneExtractMaxAngles angles per group
static bool plane2d_extract_launch(hipStream_t st, int n, int P, const c2<float>* V, const double* trig,
                                   c2<float>* out)
{
    int groups = 1;
    while (n * groups < 256 && groups * 2 <= P) groups *= 2;
    while ((P + groups - 1) / groups > kPlaneExtractMaxAngles) groups *= 2;
    const dim3 grid((unsigned)(n * groups));
    switch (n) {
    case 64:
        hipLaunchKernelGGL((k_plane2d_extract<64>), grid, dim3(64 * SizeCfg<64>::P::T), 0, st, V, trig, P,
                           groups, out);
        break;
    case 128:
        hipLaunchKernelGGL((k_plane2d_extract<128>), grid, dim3(128 * SizeCfg<128>::P::T), 0, st, V, trig,
                           P, groups, out);
        break;
    default: return false;
    }
    check_launch("k_plane2d_extract");
    return true;
}

// true when the plane kernel covers size n (the plane fits in LDS)
static bool plane2d_launch(hipStream_t st, int n, int dir, c2<float>* V)
{
    const dim3 grid((unsigned)n);
    switch (n) {
    case 64:
        if (dir > 0) hipLaunchKernelGGL((k_plane2d<64, +1, true>), grid, dim3(64 * SizeCfg<64>::P::T), 0, st, V);
        else hipLaunchKernelGGL((k_plane2d<64, -1, false>), grid, dim3(64 * SizeCfg<64>::P::T), 0, st, V);
        break;
    case 128:
        if (dir > 0) hipLaunchKernelGGL((k_plane2d<128, +1, true>), grid, dim3(128 * SizeCfg<128>::P::T), 0, st, V);
        else hipLaunchKernelGGL((k_plane2d<128, -1, false>), grid, dim3(128 * SizeCfg<128>::P::T), 0, st, V);
        break;
    default: return false;
    }
    check_launch("k_plane2d");
    return true;
}

template <int L>
static void zline_n(hipStream_t st, int mode, c2<float>* V, const uint8_t* S, int64_t W, float scale)
{
    const unsigned grid = (unsigned)(W / SizeCfg<L>::C);
    if (mode == ZL_ROUNDTRIP)
        hipLaunchKernelGGL((k_zline_constraint<L, ZL_ROUNDTRIP>), dim3(grid), dim3(SizeCfg<L>::TPB), 0, st, V, S, W, scale);
    else
        hipLaunchKernelGGL((k_zline_constraint<L, ZL_INVERSE>), dim3(grid), dim3(SizeCfg<L>::TPB), 0, st, V, S, W, scale);
}

template <int L>
static void pack_zline_support_n(hipStream_t st, const uint8_t* S, int64_t W, uint8_t* bits)
{
    const int64_t total = W / 8 * L;
    hipLaunchKernelGGL((k_pack_zline_support<SizeCfg<L>::C>), dim3((unsigned)grid_1d(total, 256, 4096)), dim3(256),
                       0, st, S, W, L, bits);
}

// The 3D support S [n][n][n] packed into h->ws_zsup for zline_launch
// (n^3 / 8 bytes); returns the packed bits.
static const uint8_t* pack_zline_support(cpra_hip_ctx* h, hipStream_t st, int n, const uint8_t* S)
{
    const int64_t W = (int64_t)n * n;
    h->ws_zsup.ensure((size_t)(W / 8 * n));
    uint8_t* bits = static_cast<uint8_t*>(h->ws_zsup.ptr);
    switch (n) {
    case 64: pack_zline_support_n<64>(st, S, W, bits); break;
    case 128: pack_zline_support_n<128>(st, S, W, bits); break;
    case 192: pack_zline_support_n<192>(st, S, W, bits); break;
    case 256: pack_zline_support_n<256>(st, S, W, bits); break;
    case 384: pack_zline_support_n<384>(st, S, W, bits); break;
    case 512: pack_zline_support_n<512>(st, S, W, bits); break;
    default: fail(CPRA_HIP_ERR_UNSUPPORTED, "carousel: n=%d has no z-line kernel", n);
    }
    check_launch("k_pack_zline_support");
    return bits;
}

// S: the support packed by pack_zline_support
static void zline_launch(hipStream_t st, int n, int mode, c2<float>* V, const uint8_t* S)
{
    const int64_t W = (int64_t)n * n;
    const float scale = (float)(1.0 / ((double)n * n * n));
    switch (n) {
    case 64: zline_n<64>(st, mode, V, S, W, scale); break;
    case 128: zline_n<128>(st, mode, V, S, W, scale); break;
    case 192: zline_n<192>(st, mode, V, S, W, scale); break;
    case 256: zline_n<256>(st, mode, V, S, W, scale); break;
    case 384: zline_n<384>(st, mode, V, S, W, scale); break;
    case 512: zline_n<512>(st, mode, V, S, W, scale); break;
    }
    check_launch("k_zline_constraint");
}

// First HIO column pass of an outer iteration after the extraction, which
// leaves the projections' 2D spectra F in the state buffer.  The unfused
// sequence is a column IDFT pass (Y = IDFT_c(F)/n) and the column kernel's
// DFT_c(Y), which gives F back up to rounding.  This kernel instead runs the
// modulus step on F itself and writes both Y = IDFT_c(F)/n (in place) and
// Z = IDFT_c(modulus(F)) for the row kernel: one launch fewer, and 28 B/px
// (read F, A; write Y, Z) against 16 + 20.  Modulus step as in the engine's
// column kernel: F * A/|F| scaled by 1/n^2, F kept (scaled) where |F| = 0 or,
// with skip_zero, where A = 0.  Dense row-major state, block = C columns of
// one projection.
template <int N>
__global__ __launch_bounds__(SizeCfg<N>::TPB) void k_car_cols_from_spectrum(
    c2<float>* YF, c2<float>* __restrict__ Z, const float* __restrict__ A, int64_t amp_batch,
    float scale, float inv_n, int skip_zero)
{
    using V = c2<float>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, C = SizeCfg<N>::C;
    constexpr unsigned TPP = N / C;  // column tiles per projection
    __shared__ V lds[N * C];
    const int tid = threadIdx.x;
    const int c = tid % C, t = tid / C;
    const unsigned b = blockIdx.x / TPP;
    const int col0 = (int)(blockIdx.x - b * TPP) * C;
    const size_t pat = (size_t)N * N;
    V* Yp = YF + b * pat + col0;
    V* Zp = Z + b * pat + col0;
    const float* Ap = A + (size_t)(b % (unsigned)amp_batch) * pat + col0;
    // element m of this lane: row t + rot(m)*T (the centring shift folded in)
    const unsigned toff = (unsigned)t * N + (unsigned)c;
    auto rowbase = [&](int m) { return (size_t)(PL::rot(m) * TT) * (size_t)N; };
    V v[E], u[E];
    float av[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = (Yp + rowbase(m))[toff];
    });
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        av[m] = (Ap + rowbase(m))[toff];
    });
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const float a = av[m];
        const V x = v[m];
        const float m2 = x.x * x.x + x.y * x.y;
        float sc = scale;
        if (!(skip_zero && a == 0.f) && m2 > 0.f) sc = scale * a * rsqrtf(m2);
        u[m] = cscale(x, sc);
    });
    LdsCol<V, C, false> acc{lds + c};
    fft_line<PL, +1>(v, t, g_tw<N>(), acc);
    fft_line<PL, +1>(u, t, g_tw<N>(), acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        (Yp + rowbase(m))[toff] = cscale(v[m], inv_n);
        (Zp + rowbase(m))[toff] = u[m];
    });
}

template <int N>
static void cols_from_spectrum_n(hipStream_t st, int P, c2<float>* Y, c2<float>* Z, const float* A,
                                 int64_t amp_batch, int skip_zero)
{
    const float scale = (float)(1.0 / ((double)N * N));
    const float inv_n = (float)(1.0 / (double)N);
    hipLaunchKernelGGL((k_car_cols_from_spectrum<N>), dim3((unsigned)(P * (N / SizeCfg<N>::C))),
                       dim3(SizeCfg<N>::TPB), 0, st, Y, Z, A, amp_batch, scale, inv_n, skip_zero);
}

static void cols_from_spectrum_launch(hipStream_t st, int n, int P, c2<float>* Y, c2<float>* Z,
                                      const float* A, int64_t amp_batch, int skip_zero)
{
    switch (n) {
    case 64: cols_from_spectrum_n<64>(st, P, Y, Z, A, amp_batch, skip_zero); break;
    case 128: cols_from_spectrum_n<128>(st, P, Y, Z, A, amp_batch, skip_zero); break;
    case 192: cols_from_spectrum_n<192>(st, P, Y, Z, A, amp_batch, skip_zero); break;
    case 256: cols_from_spectrum_n<256>(st, P, Y, Z, A, amp_batch, skip_zero); break;
    case 384: cols_from_spectrum_n<384>(st, P, Y, Z, A, amp_batch, skip_zero); break;
    case 512: cols_from_spectrum_n<512>(st, P, Y, Z, A, amp_batch, skip_zero); break;
    default: fail(CPRA_HIP_ERR_UNSUPPORTED, "carousel: n=%d has no column kernel", n);
    }
    check_launch("k_car_cols_from_spectrum");
}

struct Carousel {
    cpra_hip_ctx* h;
    int n, P;
    c2<float>* Y;  // projections, row-spectral state [P][n][n]
    const float* amp;
    int64_t amp_batch;
    const uint8_t* sup2d;
    const float* angles;
    const uint8_t* sup3d;  // packed by pack_zline_support
    c2<float>* vol;        // [n][n][n]
    cpra_hip_hio_params hp;
    int inner;

    // inverse (x, y) passes, fused z-line constraint, forward (y, x) passes
    // (n <= 128: one whole-plane kernel per direction for the x and y passes)
    // mode ZL_ROUNDTRIP_NO_FORWARD: the round trip without the forward (y, x)
    // transform, which the caller runs fused with the re-projection
    void volume_constraint(int mode)
    {
        const int64_t W = (int64_t)n * n;
        const bool plane = engine_plane_fft(h);
        const bool forward = mode == ZL_ROUNDTRIP;
        if (mode == ZL_ROUNDTRIP_NO_FORWARD) mode = ZL_ROUNDTRIP;
        if (!(plane && plane2d_launch(h->stream, n, +1, vol))) {
            fft_pass_contig(h, vol, n, (int64_t)n * n, +1, true, 1.0, false);            // x
            fft_pass_strided(h, vol, n, (int64_t)n * n, n, n, W, +1, true, 1.0, false);  // y
        }
        zline_launch(h->stream, n, mode, vol, sup3d);                                // z
        if (forward && !(plane && plane2d_launch(h->stream, n, -1, vol))) {
            fft_pass_strided(h, vol, n, (int64_t)n * n, n, n, W, -1, true, 1.0, false);
            fft_pass_contig(h, vol, n, (int64_t)n * n, -1, true, 1.0, false);
        }
    }

    // F -> Y = DFT_rows(Backward2D(F)): the column IDFT scaled 1/n
    void to_row_spectral(KernelTimer* kt)
    {
        const int64_t W = (int64_t)n * n;
        if (kt) kt->begin(h, h->stream, KT_PFFT);
        fft_pass_strided(h, Y, n, (int64_t)P * n, n, n, W, +1, true, 1.0 / (double)n, false);
        if (kt) kt->end(h, h->stream);
    }

    // One HIO iteration whose input is F (left in Y by the extraction): with
    // dense rows in one launch per pass the column pass starts from F itself;
    // otherwise the column IDFT pass restores the row-spectral state first.
    void hio_from_spectrum(KernelTimer* kt)
    {
        if (engine_hio_dense_whole(h)) {
            if (kt) kt->begin(h, h->stream, KT_COL);
            cols_from_spectrum_launch(h->stream, n, P, Y, static_cast<c2<float>*>(h->ws_z.ptr), amp, amp_batch,
                                      hp.dc_skip_zero);
            if (kt) kt->end(h, h->stream);
            engine_hio_rows(h, h->stream, Y, sup2d, &hp, kt);
            return;
        }
        to_row_spectral(kt);
        engine_hio_iteration(h, h->stream, Y, amp, amp_batch, sup2d, &hp, kt);
    }

    // One outer iteration on h->stream.  On exit Y holds F, the 2D spectra of
    // the re-projected slices; on entry it holds F too when from_spectrum
    // (every outer iteration but the first), else the row-spectral state.
    // last: the run's last outer iteration, which leaves the constrained 3D
    // spectrum in vol (the others may leave vol stale: the next merge
    // overwrites it).
    void outer(KernelTimer* kt, bool from_spectrum, bool last)
    {
        const int64_t W = (int64_t)n * n;
        if (from_spectrum && inner == 0) to_row_spectral(kt);
        for (int i = 0; i < inner; ++i) {
            if (i == 0 && from_spectrum) hio_from_spectrum(kt);
            else engine_hio_iteration(h, h->stream, Y, amp, amp_batch, sup2d, &hp, kt);
        }
        // Y -> F = Forward2D(x): the column DFT of the row-spectral state,
        // written straight into the merge's layout [P][v][u] where a kernel
        // for n exists and tuning key 6 allows it (the extraction below
        // overwrites Y in full, so Y need not hold F)
        if (kt) kt->begin(h, h->stream, KT_PFFT);
        void* sT = merge_lines_ws(h, n, P);
        const bool lines = engine_merge_lines(h) && fft_pass_strided_lines_fwd(h, Y, sT, n, (int64_t)P * n, n, n, W);
        if (!lines) fft_pass_strided(h, Y, n, (int64_t)P * n, n, n, W, -1, true, 1.0, false);
        if (kt) kt->end(h, h->stream);
        if (kt) kt->begin(h, h->stream, KT_MERGE);
        if (lines) merge_apply_lines_c64(h, sT, vol, n, P);
        else merge_apply_c64(h, Y, vol, n, P);
        if (kt) kt->end(h, h->stream);
        // the forward plane transform and the re-projection in one kernel
        // where the planes fit in LDS and the gather list's trig table of
        // these angles is there; not in the last iteration, which must leave
        // the transformed volume in vol
        const bool fused = !last && engine_plane_fft(h) && engine_plane_extract(h) && n <= 128 &&
                           h->mplan_n == n && h->mplan_P == P && h->ws_trig.ptr;
        if (kt) kt->begin(h, h->stream, KT_VOL);
        volume_constraint(fused ? ZL_ROUNDTRIP_NO_FORWARD : ZL_ROUNDTRIP);
        if (kt) kt->end(h, h->stream);
        if (kt) kt->begin(h, h->stream, KT_EXTRACT);
        if (!(fused && plane2d_extract_launch(h->stream, n, P, vol, static_cast<const double*>(h->ws_trig.ptr),
                                              Y)))
            extract_c64(h, vol, Y, angles, n, P);
        if (kt) kt->end(h, h->stream);
    }
};

template <typename F> static int cguard(cpra_hip_ctx* h, F&& f)
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

}  // namespace cpra

using namespace cpra;

extern "C" {

void cpra_hip_carousel_default_params(cpra_hip_carousel_params* p)
{
    if (!p) return;
    p->outer_iters = 50;
    p->inner_iters = 1;
    p->beta = 0.9f;
    p->dc_skip_zero = 1;
    p->use_graph = 1;
}

int cpra_hip_carousel3d_run(cpra_hip_handle h, void* x, const float* amp, uint64_t amp_batch,
                            const uint8_t* support2d, const float* angles, const uint8_t* support3d,
                            void* vol, const cpra_hip_carousel_params* p, float* rfactor_host)
{
    return cguard(h, [&]() -> int {
        CPRA_ARG(h && x && amp && support2d && angles && support3d && vol && p,
                 "cpra_hip_carousel3d_run: NULL argument");
        CPRA_ARG(amp_batch >= 1 && h->batch % amp_batch == 0,
                 "cpra_hip_carousel3d_run: amp_batch must divide the projection count");
        CPRA_ARG(p->outer_iters >= 0 && p->inner_iters >= 0, "cpra_hip_carousel3d_run: negative iteration count");
        CPRA_ARG(x != vol, "cpra_hip_carousel3d_run: vol must not alias x");
        if (h->dtype != CPRA_HIP_F32) fail(CPRA_HIP_ERR_UNSUPPORTED, "cpra_hip_carousel3d_run: FP32 handles only");
        if (h->m != h->n)
            fail(CPRA_HIP_ERR_UNSUPPORTED, "cpra_hip_carousel3d_run: projections must be square (m == n)");
        const int n = (int)h->n;
        if (!zline_tile(n))
            fail(CPRA_HIP_ERR_UNSUPPORTED, "cpra_hip_carousel3d_run: n=%d not in {64,128,192,256,384,512}", n);
        Carousel cr;
        cr.h = h;
        cr.n = n;
        cr.P = (int)h->batch;
        cr.Y = static_cast<c2<float>*>(x);
        cr.amp = amp;
        cr.amp_batch = (int64_t)amp_batch;
        cr.sup2d = support2d;
        cr.angles = angles;
        cr.sup3d = nullptr;  // packed below
        cr.vol = static_cast<c2<float>*>(vol);
        cpra_hip_hio_default_params(&cr.hp);
        cr.hp.beta = p->beta;
        cr.hp.real_positive = 1;
        cr.hp.dc_skip_zero = p->dc_skip_zero;
        cr.hp.error_reduction = 0;
        cr.hp.sw_every = 0;
        cr.inner = p->inner_iters;
        hipStream_t st = h->stream;
        engine_hio_prepare(h, st, support2d);
        if (p->outer_iters > 0) merge_plan_c64(h, angles, n, cr.P);  // the footprint, once per run
        cr.sup3d = pack_zline_support(h, st, n, support3d);
        // enter the row-spectral state of the projections
        fft_pass_contig(h, x, n, (int64_t)cr.P * n, -1, true, 1.0, false);
        KernelTimer kt;
        KernelTimer* ktp = h->timing ? &kt : nullptr;
        const int outer_iters = p->outer_iters;
        int done = 0;
        if (outer_iters > 0) {
            cr.outer(ktp, false, outer_iters == 1);  // eager: sizes every workspace before any capture
            done = 1;
            // two more eager iterations queue GPU work that covers the host's
            // capture and graph update below
            for (; done < 3 && outer_iters - done >= 12; ++done) cr.outer(ktp, true, false);
        }
        const int rest = outer_iters - done;
        // a body of kBody outer iterations: each graph launch costs about
        // 4.5 us of idle GPU at its boundary, a third of an outer iteration's
        // kernel time at 180 x 128^2, so one launch covers several.  The body
        // has the same length in every call, so the graph kept by the handle
        // is re-pointed, not rebuilt; shorter runs and the remainder, which
        // holds at least the last iteration, run eagerly
        constexpr int kBody = 10;
        if (p->use_graph && !h->timing && !h->capturing && !debug_launches() && rest > kBody) {
            hipStream_t gst = graph_stream(h);
            hipEvent_t ev0 = nullptr, ev1 = nullptr;
            if (gst != st) {
                CPRA_CHECK_HIP(hipEventCreateWithFlags(&ev0, hipEventDisableTiming));
                CPRA_CHECK_HIP(hipEventCreateWithFlags(&ev1, hipEventDisableTiming));
                CPRA_CHECK_HIP(hipEventRecord(ev0, st));
                CPRA_CHECK_HIP(hipStreamWaitEvent(gst, ev0, 0));
            }
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            const int body = kBody;
            // what the body's launches depend on: when a call repeats the
            // kept graph's key, the graph replays as it is, with no capture
            // and no update (their host time, about half a millisecond for
            // a body of 100 launches at 180 x 128^2, would otherwise leave
            // the GPU idle)
            std::vector<uint64_t> key;
            engine_state_key(h, key);
            for (const void* ptr : {(const void*)cr.Y, (const void*)amp, (const void*)support2d,
                                    (const void*)angles, (const void*)support3d, (const void*)vol, (const void*)gst})
                key.push_back((uint64_t)(uintptr_t)ptr);
            uint32_t beta_bits;
            std::memcpy(&beta_bits, &cr.hp.beta, sizeof(beta_bits));
            for (uint64_t v : {(uint64_t)n, (uint64_t)cr.P, (uint64_t)amp_batch, (uint64_t)beta_bits,
                               (uint64_t)cr.hp.dc_skip_zero, (uint64_t)cr.inner, (uint64_t)body})
                key.push_back(v);
            KeptGraph& kg = h->kept[KG_CAROUSEL];
            if (kg.exec && kg.key == key) {
                exec = kg.exec;
                // no replay may still be in flight when the next call changes
                // the body's inputs: queue behind the last one
                if (kg.done) CPRA_CHECK_HIP(hipStreamWaitEvent(gst, kg.done, 0));
            } else {
                h->stream = gst;
                try {
                    CPRA_CHECK_HIP(hipStreamBeginCapture(gst, hipStreamCaptureModeThreadLocal));
                    for (int i = 0; i < body; ++i) cr.outer(nullptr, true, false);
                    CPRA_CHECK_HIP(hipStreamEndCapture(gst, &graph));
                } catch (...) {
                    h->stream = st;
                    throw;
                }
                h->stream = st;
                kg.key.clear();
                try {
                    exec = kept_graph_exec(h, kg, graph);
                } catch (...) {
                    (void)hipGraphDestroy(graph);
                    throw;
                }
                (void)hipGraphDestroy(graph);
                graph = nullptr;
                kg.key = std::move(key);
            }
            const int replays = (rest - 1) / body;  // the last iteration stays outside
            for (int i = 0; i < replays; ++i) CPRA_CHECK_HIP(hipGraphLaunch(exec, gst));
            kept_graph_launched(kg, gst);
            h->stream = gst;
            try {
                for (int i = replays * body; i < rest; ++i) cr.outer(nullptr, true, i == rest - 1);
            } catch (...) {
                h->stream = st;
                throw;
            }
            h->stream = st;
            if (ev0) {
                CPRA_CHECK_HIP(hipEventRecord(ev1, gst));
                CPRA_CHECK_HIP(hipStreamWaitEvent(st, ev1, 0));
                CPRA_CHECK_HIP(hipEventSynchronize(ev1));
                (void)hipEventDestroy(ev0);
                (void)hipEventDestroy(ev1);
            }
        } else {
            for (int i = 0; i < rest; ++i) cr.outer(ktp, true, i == rest - 1);
        }
        if (p->outer_iters > 0) cr.to_row_spectral(ktp);  // F -> the row-spectral state
        // leave: x = IDFT_rows(Y) / n; vol = the constrained real-space object
        fft_pass_contig(h, x, n, (int64_t)cr.P * n, +1, true, 1.0 / (double)n, false);
        if (p->outer_iters > 0) cr.volume_constraint(ZL_INVERSE);
        kt.collect(h);
        if (rfactor_host) {
            const int rc = cpra_hip_rfactor(h, x, amp, amp_batch, rfactor_host);
            if (rc < 0) return rc;
        }
        return CPRA_HIP_OK;
    });
}

}  // extern "C"
