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
// MI355X mapping.  Slice row u lies in the plane z = u, and the merge and the
// re-projection combine whole lines along the rotation axis with weights that
// do not depend on the position along it.  Both therefore commute with the
// centred transforms along that axis: with D the DFT along u or z and D^-1
// its normalised inverse,
//   D^-1 R^T(F) / R^T(1) = R^T(D^-1 F) / R^T(1),     R(D V) = D R(V).
// The projections stay in the row-spectral domain Y = DFT_rows(x) = D^-1 F
// for the whole run (u real, v spectral).  So the merge gathers Y itself into
// a volume that is spectral in (x, y) and real in z, the 3D constraint's z
// transforms cancel against the projections' column transforms, and the
// re-projection of the constrained volume is Y's next state directly:
//   vol = R^T(Y) / R^T(1);   v = (Re IDFT_xy(vol) / n^2)^+ * S3;
//   Y' = R(DFT_xy(v))
// (IDFT_xy, DFT_xy unnormalised; the row-spectral state carries no scale).
// Steps 2-6 are a transpose of Y into the merge's layout, the merge, and for
// n <= 128 one kernel per z-plane: IDFT_xy, the constraint and DFT_xy in LDS,
// then the plane's rows of every slice sampled from there.  Above 128 they
// are an x pass, one y-line kernel (IDFT_y, the constraint and DFT_y in
// registers) and an x pass back, and the re-projection is a kernel of its
// own.  The outer body is captured in a
// hipGraph, ten outer iterations per graph, which the handle keeps across
// calls and replays; its working set (the projections, the transposed slices
// and the volume: 56 MB at 180 x 128^2 and 128^3) stays in the Infinity Cache.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "fft_core.hpp"
#include "plans.hpp"

namespace cpra {

// The 3D support S [n][n][n] (bytes, nonzero = inside) packed one bit per
// voxel along x, once per run: word (z*n + y)*(n/32) + x/32 holds bit x % 32.
__global__ __launch_bounds__(256) void k_pack_plane_support(const uint8_t* __restrict__ S, int64_t words,
                                                           unsigned* __restrict__ bits)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t* src = S + i * 32;
        unsigned w = 0;
#pragma unroll
        for (int k = 0; k < 32; ++k) w |= (src[k] != 0 ? 1u : 0u) << k;
        bits[i] = w;
    }
}

// The y-lines of the x-transformed volume [N][N][N], for the sizes whose
// planes do not fit in LDS (and tuning plane=0): IDFT_y, the constraint
// (Re(w) * scale)^+ inside the support and 0 elsewhere (Im 0), and with FWD
// DFT_y, per line in registers, in place.  Block (z, tile): C consecutive x
// of plane z in the strided-line mapping of the column kernels (T threads per
// line, element m at row t + rot(m) * T).  sbits: k_pack_plane_support's
// words.  One pass for what the inverse y pass, a constraint pass and the
// forward y pass did.
template <int N, bool FWD>
__global__ __launch_bounds__(SizeCfg<N>::TPB) void k_ystep_constraint(c2<float>* V,
                                                                     const unsigned* __restrict__ sbits,
                                                                     float scale)
{
    using Vt = c2<float>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, C = SizeCfg<N>::C;
    static_assert(N % C == 0, "whole column tiles");
    __shared__ Vt lds[N * C];
    const int c = (int)threadIdx.x % C, t = (int)threadIdx.x / C;
    const int z = (int)blockIdx.x / (N / C), x = ((int)blockIdx.x % (N / C)) * C + c;
    Vt* Vp = V + (size_t)z * N * N + x;
    const unsigned* sw = sbits + (size_t)z * N * (N / 32) + (x >> 5);
    const unsigned bit = 1u << (x & 31);
    auto row = [&](int m) { return t + PL::rot(m) * TT; };
    Vt v[E];
    unsigned w[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = Vp[(size_t)row(m) * N];
        w[m] = sw[row(m) * (N / 32)];
    });
    LdsCol<Vt, C, false> acc{lds + c};
    fft_line<PL, +1>(v, t, g_tw<N>(), acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const float re = v[m].x * scale;
        v[m] = mk<Vt>((w[m] & bit) && re > 0.f ? re : 0.f, 0.f);
    });
    if constexpr (FWD) fft_line<PL, -1>(v, t, g_tw<N>(), acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        Vp[(size_t)row(m) * N] = v[m];
    });
}

template <int N>
static void ystep_n(hipStream_t st, c2<float>* V, const unsigned* sbits, bool fwd)
{
    const dim3 grid((unsigned)(N * (N / SizeCfg<N>::C)));
    const float scale = (float)(1.0 / ((double)N * N));
    if (fwd)
        hipLaunchKernelGGL((k_ystep_constraint<N, true>), grid, dim3(SizeCfg<N>::TPB), 0, st, V, sbits, scale);
    else
        hipLaunchKernelGGL((k_ystep_constraint<N, false>), grid, dim3(SizeCfg<N>::TPB), 0, st, V, sbits, scale);
}

static void ystep_launch(hipStream_t st, int n, c2<float>* V, const unsigned* sbits, bool fwd)
{
    switch (n) {
    case 64: ystep_n<64>(st, V, sbits, fwd); break;
    case 128: ystep_n<128>(st, V, sbits, fwd); break;
    case 192: ystep_n<192>(st, V, sbits, fwd); break;
    case 256: ystep_n<256>(st, V, sbits, fwd); break;
    case 384: ystep_n<384>(st, V, sbits, fwd); break;
    case 512: ystep_n<512>(st, V, sbits, fwd); break;
    default: fail(CPRA_HIP_ERR_UNSUPPORTED, "carousel: no y-line kernel for n=%d", n);
    }
    check_launch("k_ystep_constraint");
}

// Centred 2D transforms along x and y of the z-planes of the volume
// [N][N][N], one workgroup per plane, for N <= 128, where an N x N complex64
// plane fits in LDS (136 KiB at 128 with the padding below).
// LDS layouts, with row span SP = N + N/16 and the LdsRow padding of one value
// per 16 (pad(i) = i + i/16):
//   NAT(y, x) = y*SP + pad(x)  (x-lines contiguous)
//   TRN(y, x) = x*SP + pad(y)  (y-lines contiguous)
// The plane is staged through LDS with whole-row (512-byte) global accesses;
// loading each line straight into registers (64-byte segments) measured
// slower.  An axis step reads each line from the layout where it is
// contiguous.  The line's T threads sit in one wave, so the FFT exchanges in
// place within the line's own span with wave-level ordering only.  After a
// workgroup barrier the step writes the other layout.  With 8 lines x 8
// threads per wave at N = 128, both the in-place reads and the transposed
// writes touch 64 distinct 8-byte slots modulo 512 bytes.  Each line goes
// through the same plan, twiddles and centring fold as in the one-axis
// kernels of fft.hip.
// angles per group of the plane kernel's re-projection (the group's
// (cos, sin) table sits in LDS beside the plane)
constexpr int kPlaneExtractMaxAngles = 256;

template <int N> struct PlaneGeom {
    using PL = typename SizeCfg<N>::P;
    static constexpr int E = PL::E, TT = PL::T, TPB = N * TT, SP = N + N / 16, K = N * N / TPB;
    static_assert(64 % TT == 0 && TPB % 64 == 0 && N % 32 == 0 && TPB % N == 0 && K == E, "plane layout");
    static __device__ __forceinline__ int pad(int i) { return i + (i >> 4); }
};

// The plane Pp (N x N, row-major [y][x]) into lds in NAT layout.
template <int N>
__device__ __forceinline__ void plane2d_load(const c2<float>* Pp, c2<float>* lds)
{
    using G = PlaneGeom<N>;
    c2<float> g[G::K];
    sfor<G::K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        g[k] = Pp[threadIdx.x + k * G::TPB];
    });
    sfor<G::K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int i = threadIdx.x + k * G::TPB;
        lds[(i / N) * G::SP + G::pad(i % N)] = g[k];
    });
    __syncthreads();
}

// One axis step on the plane in lds: every line of the current layout (the
// rows of lds: x-lines in NAT, y-lines in TRN) transformed with sign DIR,
// unnormalised, in registers with its exchanges in the line's own span; then
// written transposed (TRANSPOSE: NAT <-> TRN) or back in place.
template <int N, int DIR, bool TRANSPOSE>
__device__ __forceinline__ void plane2d_step(c2<float>* lds)
{
    using Vt = c2<float>;
    using G = PlaneGeom<N>;
    using PL = typename G::PL;
    constexpr int E = G::E, TT = G::TT, SP = G::SP;
    auto pad = [](int i) { return G::pad(i); };
    const int t = (int)threadIdx.x % TT, line = (int)threadIdx.x / TT;
    Vt* lb = lds + line * SP;
    Vt v[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = lb[pad(t + PL::rot(m) * TT)];
    });
    LdsRow<Vt, N, true> acc{lb};
    fft_line<PL, DIR>(v, t, g_tw<N>(), acc);
    if constexpr (TRANSPOSE) __syncthreads();  // every exchange done before the transposed writes
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const int pos = t + PL::rot(m) * TT;
        if constexpr (TRANSPOSE) lds[pos * SP + pad(line)] = v[m];
        else lb[pad(pos)] = v[m];  // the line's own span: its exchanges are this wave's
    });
    __syncthreads();
}

// This thread's support words of plane z for plane2d_ystep: element m of
// y-line x = tid / T sits at y = tid % T + rot(m) * T, and its bit is bit
// x % 32 of word (y, x / 32) of the plane's words (N/32 per row,
// k_pack_plane_support).  Loaded at the kernel's start, so their latency
// overlaps the plane's load.
template <int N>
__device__ __forceinline__ void plane_support_words(const unsigned* __restrict__ sw,
                                                    unsigned (&w)[PlaneGeom<N>::E])
{
    using G = PlaneGeom<N>;
    using PL = typename G::PL;
    const int t = (int)threadIdx.x % G::TT, x = (int)threadIdx.x / G::TT;
    sfor<G::E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        w[m] = sw[(t + PL::rot(m) * G::TT) * (N / 32) + (x >> 5)];
    });
}

// The y-lines (TRN rows) of the x-transformed plane: IDFT_y, the constraint
// (Re(w) * scale)^+ inside the support and 0 elsewhere (Im 0), and with FWD
// DFT_y, all in registers, one line per T threads; written transposed (NAT).
// Taking the constraint between the two y transforms saves the plane a round
// trip through LDS and a pass of its own.  w: plane_support_words.
template <int N, bool FWD>
__device__ __forceinline__ void plane2d_ystep(c2<float>* lds, const unsigned (&w)[PlaneGeom<N>::E], float scale)
{
    using Vt = c2<float>;
    using G = PlaneGeom<N>;
    using PL = typename G::PL;
    constexpr int E = G::E, TT = G::TT, SP = G::SP;
    auto pad = [](int i) { return G::pad(i); };
    const int t = (int)threadIdx.x % TT, line = (int)threadIdx.x / TT;  // line = x
    Vt* lb = lds + line * SP;
    Vt v[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = lb[pad(t + PL::rot(m) * TT)];
    });
    LdsRow<Vt, N, true> acc{lb};
    fft_line<PL, +1>(v, t, g_tw<N>(), acc);
    const unsigned bit = 1u << (line & 31);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const float re = v[m].x * scale;
        v[m] = mk<Vt>((w[m] & bit) && re > 0.f ? re : 0.f, 0.f);
    });
    if constexpr (FWD) fft_line<PL, -1>(v, t, g_tw<N>(), acc);
    __syncthreads();  // every exchange done before the transposed writes
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        lds[(t + PL::rot(m) * TT) * SP + pad(line)] = v[m];
    });
    __syncthreads();
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

// The 3D constraint of plane z and its share of the re-projection: block
// (z, angle group g of `groups`) loads plane z of the merged volume (spectral
// in x and y, real in z), runs IDFT_x, then IDFT_y, the constraint (scale
// 1/n^2) and DFT_y on each y-line in registers, then DFT_x, all in LDS, and
// samples row u = z of its angles' slices from there with
// the bilinear weights and FP64 geometry of k_extract_plane.  The groups put
// the CUs that one block per plane leaves idle to work on the sampling; each
// group transforms the plane itself.  Block b runs on XCD b mod 8, and the
// groups of a plane are blocks 8 apart, so they share that XCD's L2 and the
// plane is fetched from beyond it about once (its load phase measured 4.6
// against 7.7 us at 128^2 with the groups on neighbouring XCDs).  V is only
// read.  The grid is N * groups blocks (N a multiple of 8).
template <int N>
__global__ __launch_bounds__(N * SizeCfg<N>::P::T) void k_plane_constraint_extract(
    const c2<float>* __restrict__ V, const unsigned* __restrict__ sbits, float scale,
    const double* __restrict__ trig, int P, int groups, c2<float>* __restrict__ out)
{
    using Vt = c2<float>;
    using G = PlaneGeom<N>;
    __shared__ Vt lds[N * G::SP];
    __shared__ double tl[2 * kPlaneExtractMaxAngles];
    static_assert(N % 8 == 0, "planes dealt to the 8 XCDs");
    const int xcd = (int)blockIdx.x & 7, i8 = (int)blockIdx.x >> 3;
    const int u = (i8 / groups) * 8 + xcd, g = i8 % groups;
    const int p0 = (int)((int64_t)g * P / groups), p1 = (int)((int64_t)(g + 1) * P / groups);
    for (int i = threadIdx.x; i < 2 * (p1 - p0); i += blockDim.x) tl[i] = trig[2 * p0 + i];
    unsigned sw[G::E];
    plane_support_words<N>(sbits + (size_t)u * N * (N / 32), sw);
    plane2d_load<N>(V + (size_t)u * (N * N), lds);
    plane2d_step<N, +1, true>(lds);  // IDFT_x: NAT -> TRN
    plane2d_ystep<N, true>(lds, sw, scale);  // IDFT_y, constraint, DFT_y: TRN -> NAT
    plane2d_step<N, -1, false>(lds);  // DFT_x in place; ends with a barrier
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

// The constrained object of the merged volume in place, one block per plane:
// the load, IDFT_xy and constraint of k_plane_constraint_extract (the same
// bits), then the plane is stored.  The run's last outer iteration leaves
// this for the caller.
template <int N>
__global__ __launch_bounds__(N * SizeCfg<N>::P::T) void k_plane_constraint(c2<float>* V,
                                                                          const unsigned* __restrict__ sbits,
                                                                          float scale)
{
    __shared__ c2<float> lds[N * PlaneGeom<N>::SP];
    c2<float>* Pp = V + (size_t)blockIdx.x * (N * N);
    unsigned sw[PlaneGeom<N>::E];
    plane_support_words<N>(sbits + (size_t)blockIdx.x * N * (N / 32), sw);
    plane2d_load<N>(Pp, lds);  // every load of the plane done before the stores below
    plane2d_step<N, +1, true>(lds);
    plane2d_ystep<N, false>(lds, sw, scale);
    plane2d_store<N>(Pp, lds);
}

// whether the plane kernels cover size n (an n x n plane in LDS)
static bool plane_fits(int n) { return n == 64 || n == 128; }

// angle groups per plane: enough blocks to fill 256 CUs where P allows, at
// most kPlaneExtractMaxAngles angles per group
static int plane_groups(int n, int P)
{
    int groups = 1;
    while (n * groups < 256 && groups * 2 <= P) groups *= 2;
    while ((P + groups - 1) / groups > kPlaneExtractMaxAngles) groups *= 2;
    return groups;
}

// groups: angle groups per plane, 0 for plane_groups(n, P)
static void plane_constraint_extract_launch(hipStream_t st, int n, int P, int groups, const c2<float>* V,
                                            const unsigned* sbits, const double* trig, c2<float>* out)
{
    if (groups <= 0) groups = plane_groups(n, P);
    groups = std::min(groups, P);
    while ((P + groups - 1) / groups > kPlaneExtractMaxAngles) groups *= 2;
    const dim3 grid((unsigned)(n * groups));
    const float scale = (float)(1.0 / ((double)n * n));
    switch (n) {
    case 64:
        hipLaunchKernelGGL((k_plane_constraint_extract<64>), grid, dim3(64 * SizeCfg<64>::P::T), 0, st, V, sbits,
                           scale, trig, P, groups, out);
        break;
    case 128:
        hipLaunchKernelGGL((k_plane_constraint_extract<128>), grid, dim3(128 * SizeCfg<128>::P::T), 0, st, V,
                           sbits, scale, trig, P, groups, out);
        break;
    default: fail(CPRA_HIP_ERR_UNSUPPORTED, "carousel: no plane kernel for n=%d", n);
    }
    check_launch("k_plane_constraint_extract");
}

static void plane_constraint_launch(hipStream_t st, int n, c2<float>* V, const unsigned* sbits)
{
    const float scale = (float)(1.0 / ((double)n * n));
    switch (n) {
    case 64:
        hipLaunchKernelGGL((k_plane_constraint<64>), dim3(64), dim3(64 * SizeCfg<64>::P::T), 0, st, V, sbits, scale);
        break;
    case 128:
        hipLaunchKernelGGL((k_plane_constraint<128>), dim3(128), dim3(128 * SizeCfg<128>::P::T), 0, st, V, sbits,
                           scale);
        break;
    default: fail(CPRA_HIP_ERR_UNSUPPORTED, "carousel: no plane kernel for n=%d", n);
    }
    check_launch("k_plane_constraint");
}

// The 3D support S [n][n][n] packed by k_pack_plane_support into
// h->ws_zsup (n^3 / 8 bytes); returns the packed words.
static const unsigned* pack_plane_support(cpra_hip_ctx* h, hipStream_t st, int n, const uint8_t* S)
{
    const int64_t words = (int64_t)n * n * (n / 32);
    h->ws_zsup.ensure((size_t)words * sizeof(unsigned));
    unsigned* bits = static_cast<unsigned*>(h->ws_zsup.ptr);
    hipLaunchKernelGGL(k_pack_plane_support, dim3((unsigned)grid_1d(words, 256, 4096)), dim3(256), 0, st, S,
                       words, bits);
    check_launch("k_pack_plane_support");
    return bits;
}

struct Carousel {
    cpra_hip_ctx* h;
    int n, P;
    c2<float>* Y;  // projections, row-spectral state [P][n][n]
    const float* amp;
    int64_t amp_batch;
    const uint8_t* sup2d;
    const float* angles;
    const unsigned* sbits;   // the 3D support packed by pack_plane_support
    c2<float>* vol;          // [n][n][n]: spectral in (x, y), real in z
    cpra_hip_hio_params hp;
    int inner;

    // whole-plane kernels (tuning key plane, n <= 128) with the gather
    // list's trig table of these angles at hand
    bool planes() const
    {
        return engine_plane_fft(h) && plane_fits(n) && h->mplan_n == n && h->mplan_P == P && h->ws_trig.ptr;
    }

    // vol <- v = (Re IDFT_xy(vol) / n^2)^+ * S3, then with fwd DFT_y(v): the
    // x pass and one y-line kernel
    void constraint_passes(bool fwd)
    {
        fft_pass_contig(h, vol, n, (int64_t)n * n, +1, true, 1.0, false);  // x
        ystep_launch(h->stream, n, vol, sbits, fwd);
    }

    // One outer iteration on h->stream: Y (the row-spectral state) in and
    // out.  last: the run's last outer iteration, which leaves the
    // constrained object v in vol (the others leave the merged volume there,
    // or, with the passes, DFT_xy(v)).
    void outer(KernelTimer* kt, bool last)
    {
        const int64_t W = (int64_t)n * n;
        // the last HIO row pass stores Y straight into the merge's layout
        // [P][v][u] where it can (Y itself is not read again before the
        // re-projection overwrites it); else Y is transposed into it
        void* sT = merge_lines_ws(h, n, P);
        bool lines = false;
        for (int i = 0; i < inner; ++i)
            lines = engine_hio_iteration(h, h->stream, Y, amp, amp_batch, sup2d, &hp, kt,
                                         i == inner - 1 && engine_merge_lines(h) ? sT : nullptr);
        if (!lines) {
            if (kt) kt->begin(h, h->stream, KT_PTRANS);
            merge_transpose_c64(h, Y, sT, n, P);
            if (kt) kt->end(h, h->stream);
        }
        if (kt) kt->begin(h, h->stream, KT_MERGE);
        merge_apply_lines_c64(h, sT, vol, n, P);
        if (kt) kt->end(h, h->stream);
        if (planes()) {
            if (kt) kt->begin(h, h->stream, KT_VOL);
            plane_constraint_extract_launch(h->stream, n, P, engine_plane_groups(h), vol, sbits,
                                            static_cast<const double*>(h->ws_trig.ptr), Y);
            if (kt) kt->end(h, h->stream);
            if (last) plane_constraint_launch(h->stream, n, vol, sbits);
            return;
        }
        if (kt) kt->begin(h, h->stream, KT_VOL);
        constraint_passes(true);
        fft_pass_contig(h, vol, n, W, -1, true, 1.0, false);  // x
        if (kt) kt->end(h, h->stream);
        if (kt) kt->begin(h, h->stream, KT_EXTRACT);
        extract_c64(h, vol, Y, angles, n, P);
        if (kt) kt->end(h, h->stream);
        if (last) constraint_passes(false);  // DFT_xy(v) back to v
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
        if (!(n == 64 || n == 128 || n == 192 || n == 256 || n == 384 || n == 512))
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
        cr.sbits = nullptr;  // packed below
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
        cr.sbits = pack_plane_support(h, st, n, support3d);
        // enter the row-spectral state of the projections
        fft_pass_contig(h, x, n, (int64_t)cr.P * n, -1, true, 1.0, false);
        KernelTimer kt;
        KernelTimer* ktp = h->timing ? &kt : nullptr;
        const int outer_iters = p->outer_iters;
        int done = 0;
        if (outer_iters > 0) {
            cr.outer(ktp, outer_iters == 1);  // eager: sizes every workspace before any capture
            done = 1;
            // two more eager iterations queue GPU work that covers the host's
            // capture and graph update below
            for (; done < 3 && outer_iters - done >= 12; ++done) cr.outer(ktp, false);
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
                    for (int i = 0; i < body; ++i) cr.outer(nullptr, false);
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
                for (int i = replays * body; i < rest; ++i) cr.outer(nullptr, i == rest - 1);
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
            for (int i = 0; i < rest; ++i) cr.outer(ktp, i == rest - 1);
        }
        // leave: x = IDFT_rows(Y) / n; vol holds the constrained real-space
        // object of the last outer iteration
        fft_pass_contig(h, x, n, (int64_t)cr.P * n, +1, true, 1.0 / (double)n, false);
        kt.collect(h);
        if (rfactor_host) {
            const int rc = cpra_hip_rfactor(h, x, amp, amp_batch, rfactor_host);
            if (rc < 0) return rc;
        }
        return CPRA_HIP_OK;
    });
}

}  // extern "C"
