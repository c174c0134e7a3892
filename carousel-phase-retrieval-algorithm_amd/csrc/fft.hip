// fft.hip -- batched line FFT passes and the centred 2D/3D transforms of the
// C ABI (cpra_hip_forward2d/backward2d/forward3d/backward3d).
//
// A centred transform X = S(DFT(S(x))) (include/detail/MKL_Impl.inl:128-178)
// is separable: one pass per axis, each with the shift folded into its index
// map (fft_core.hpp), so a 2D transform is two HBM passes (rows, columns) and
// a 3D transform three (x, y, z), with no separate shift pass.
#include "common.hpp"
#include "plans.hpp"

namespace cpra {

// ---------------------------------------------------------------------------
// fast kernels (compile-time plans)
// ---------------------------------------------------------------------------
template <int N, int DIR, typename T, bool SHIFT>
__global__ __launch_bounds__(SizeCfg<N>::TPB) void k_fft_contig(c2<T>* __restrict__ data,
                                                                 int64_t nlines, T scale,
                                                                 int do_scale)
{
    using V = c2<T>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = SizeCfg<N>::LPB;
    constexpr int SPAN = LdsRow<V, N>::span();
    __shared__ V lds[LPB * SPAN];
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    const int64_t line = (int64_t)blockIdx.x * LPB + ll;
    const bool valid = line < nlines;
    V* base = data + (valid ? line : 0) * N;
    V v[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        constexpr int pm = SHIFT ? PL::rot(m) : m;
        v[m] = valid ? base[t + pm * TT] : mk<V>(0, 0);
    });
    LdsRow<V, N> acc{lds + ll * SPAN};
    fft_line<PL, DIR>(v, t, g_tw<N, T>(), acc);
    if (valid) {
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            constexpr int pm = SHIFT ? PL::rot(m) : m;
            base[t + pm * TT] = do_scale ? cscale(v[m], scale) : v[m];
        });
    }
}

// W = 2: tiles of twice the lines with twice the threads, so each row of a
// tile is a 2*C*sizeof(c2<T>)-byte segment (128 bytes for 512-point float
// lines instead of 64).
template <int N, int DIR, typename T, bool SHIFT, int W = 1>
__global__ __launch_bounds__(W * SizeCfg<N>::TPB) void k_fft_strided(
    c2<T>* __restrict__ data, int64_t nlines, int64_t es, int64_t G, int64_t GS, T scale,
    int do_scale)
{
    using V = c2<T>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, C = W * SizeCfg<N>::C;
    __shared__ V lds[N * C];
    const int tid = threadIdx.x;
    const int c = tid % C, t = tid / C;
    // the tile never straddles a line group (nlines % C == 0, G % C == 0), so
    // the block base is uniform; element m adds the uniform (pm*T*es) to one
    // per-lane offset (c + t*es)
    const int64_t line0 = (int64_t)blockIdx.x * C;
    V* blk = data + (line0 / G) * GS + (line0 % G);
    const size_t loff = (size_t)c + (size_t)t * (size_t)es;
    V v[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        constexpr int pm = SHIFT ? PL::rot(m) : m;
        v[m] = (blk + (size_t)(pm * PL::T) * (size_t)es)[loff];
    });
    LdsCol<V, C> acc{lds + c};
    fft_line<PL, DIR>(v, t, g_tw<N, T>(), acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        constexpr int pm = SHIFT ? PL::rot(m) : m;
        (blk + (size_t)(pm * PL::T) * (size_t)es)[loff] = do_scale ? cscale(v[m], scale) : v[m];
    });
}

// ---------------------------------------------------------------------------
// generic kernel: any n whose prime factors are <= 64, one line per block,
// whole line in LDS, runtime radix sequence (fallback for non-fast sizes)
// ---------------------------------------------------------------------------
struct GenericPlan {
    int n;
    int nrad;
    int rad[40];
};

template <int R, int DIR, typename V>
__device__ __forceinline__ void generic_bfly_fixed(V* b)
{
    Dft<R, DIR>::run(b);
}

template <int DIR, typename T>
__global__ __launch_bounds__(256) void k_fft_generic(c2<T>* __restrict__ data, int64_t es,
                                                     int64_t G, int64_t GS,
                                                     const c2<T>* __restrict__ tw, GenericPlan plan,
                                                     int shift, T scale, int do_scale)
{
    using V = c2<T>;
    extern __shared__ unsigned char smem_raw[];
    V* a = reinterpret_cast<V*>(smem_raw);
    V* b = a + plan.n;
    const int n = plan.n;
    const int64_t line = blockIdx.x;
    V* base = data + (line / G) * GS + (line % G);
    const int half = n / 2;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int p = shift ? (j + half) % n : j;
        a[j] = base[(int64_t)p * es];
    }
    __syncthreads();
    int ns = 1;
    V* in = a;
    V* out = b;
    for (int s = 0; s < plan.nrad; ++s) {
        const int R = plan.rad[s];
        const int nr = n / R;
        const int stride = n / (ns * R);
        for (int j = threadIdx.x; j < nr; j += blockDim.x) {
            const int k = j % ns;
            V v[64];
            for (int r = 0; r < R; ++r) {
                V x = in[j + r * nr];
                V w = tw[(int64_t)r * k * stride];
                if (DIR > 0) w.y = -w.y;
                v[r] = cmul(x, w);
            }
            switch (R) {
            case 2: generic_bfly_fixed<2, DIR>(v); break;
            case 3: generic_bfly_fixed<3, DIR>(v); break;
            case 4: generic_bfly_fixed<4, DIR>(v); break;
            case 5: generic_bfly_fixed<5, DIR>(v); break;
            default: {
                V o[64];
                const int rs = n / R;  // w_R^e = tw[e * n/R]
                for (int kk = 0; kk < R; ++kk) {
                    V acc = mk<V>(0, 0);
                    for (int q = 0; q < R; ++q) {
                        V w = tw[(int64_t)((q * kk) % R) * rs];
                        if (DIR > 0) w.y = -w.y;
                        acc = cadd(acc, cmul(v[q], w));
                    }
                    o[kk] = acc;
                }
                for (int kk = 0; kk < R; ++kk) v[kk] = o[kk];
            }
            }
            const int idxd = (j / ns) * ns * R + k;
            for (int r = 0; r < R; ++r) out[idxd + r * ns] = v[r];
        }
        __syncthreads();
        V* tmp = in;
        in = out;
        out = tmp;
        ns *= R;
    }
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int p = shift ? (j + half) % n : j;
        V x = in[j];
        base[(int64_t)p * es] = do_scale ? cscale(x, scale) : x;
    }
}

static GenericPlan make_generic_plan(int n)
{
    GenericPlan p{};
    p.n = n;
    p.nrad = 0;
    int m = n;
    while (m % 4 == 0) { p.rad[p.nrad++] = 4; m /= 4; }
    while (m % 2 == 0) { p.rad[p.nrad++] = 2; m /= 2; }
    while (m % 3 == 0) { p.rad[p.nrad++] = 3; m /= 3; }
    while (m % 5 == 0) { p.rad[p.nrad++] = 5; m /= 5; }
    for (int f = 7; m > 1 && f <= 64; f += 2)
        while (m % f == 0) { p.rad[p.nrad++] = f; m /= f; }
    if (m > 1) fail(CPRA_HIP_ERR_UNSUPPORTED, "FFT length %d has a prime factor above 64", n);
    return p;
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
bool fft_fast_size(int n)
{
    switch (n) {
    case 64: case 128: case 192: case 256: case 384: case 512: return true;
    default: return false;
    }
}

template <int N, typename T>
static void launch_contig(cpra_hip_ctx* h, void* data, int64_t lines, int dir, bool shift,
                          double scale)
{
    using V = c2<T>;
    const unsigned grid = (unsigned)((lines + SizeCfg<N>::LPB - 1) / SizeCfg<N>::LPB);
    const int ds = scale != 1.0;
    V* d = static_cast<V*>(data);
#define CPRA_LC(DIR, SH)                                                                      \
    hipLaunchKernelGGL((k_fft_contig<N, DIR, T, SH>), dim3(grid), dim3(SizeCfg<N>::TPB), 0,  \
                       h->stream, d, lines, (T)scale, ds)
    if (dir < 0) {
        if (shift) CPRA_LC(-1, true); else CPRA_LC(-1, false);
    } else {
        if (shift) CPRA_LC(+1, true); else CPRA_LC(+1, false);
    }
#undef CPRA_LC
    check_launch("k_fft_contig");
}

template <int N, typename T>
static void launch_strided(cpra_hip_ctx* h, void* data, int64_t lines, int64_t es, int64_t G,
                           int64_t GS, int dir, bool shift, double scale)
{
    using V = c2<T>;
    const int ds = scale != 1.0;
    V* d = static_cast<V*>(data);
    // 512-point float lines: double-width tiles (see k_fft_strided)
    constexpr int C2 = 2 * SizeCfg<N>::C;
    const bool wide = N == 512 && sizeof(T) == 4 && G % C2 == 0 && lines % C2 == 0;
#define CPRA_LS(DIR, SH, WW)                                                                        \
    hipLaunchKernelGGL((k_fft_strided<N, DIR, T, SH, WW>), dim3((unsigned)(lines / (WW * SizeCfg<N>::C))), \
                       dim3(WW * SizeCfg<N>::TPB), 0, h->stream, d, lines, es, G, GS, (T)scale, ds)
    if constexpr (N == 512 && sizeof(T) == 4) {
        if (wide) {
            if (dir < 0) {
                if (shift) CPRA_LS(-1, true, 2); else CPRA_LS(-1, false, 2);
            } else {
                if (shift) CPRA_LS(+1, true, 2); else CPRA_LS(+1, false, 2);
            }
            check_launch("k_fft_strided");
            return;
        }
    }
    if (dir < 0) {
        if (shift) CPRA_LS(-1, true, 1); else CPRA_LS(-1, false, 1);
    } else {
        if (shift) CPRA_LS(+1, true, 1); else CPRA_LS(+1, false, 1);
    }
#undef CPRA_LS
    check_launch("k_fft_strided");
}

template <typename T>
static void launch_generic(cpra_hip_ctx* h, void* data, int n, int64_t lines, int64_t es,
                          int64_t G, int64_t GS, int dir, bool shift, double scale)
{
    using V = c2<T>;
    const size_t lds = 2 * (size_t)n * sizeof(V);
    if (lds > 64 * 1024)
        fail(CPRA_HIP_ERR_UNSUPPORTED, "FFT length %d too long for the generic LDS kernel", n);
    GenericPlan plan = make_generic_plan(n);
    const V* tw = static_cast<const V*>(twiddle_table(h, n, sizeof(T) == 8));
    V* d = static_cast<V*>(data);
    const int ds = scale != 1.0;
    const unsigned block = (unsigned)(n / 2 < 64 ? 64 : (n / 2 > 256 ? 256 : ((n / 2 + 63) / 64) * 64));
    if (lines > 0x7fffffff) fail(CPRA_HIP_ERR_UNSUPPORTED, "too many FFT lines");
    if (dir < 0)
        hipLaunchKernelGGL((k_fft_generic<-1, T>), dim3((unsigned)lines), dim3(block), lds,
                           h->stream, d, es, G, GS, tw, plan, shift ? 1 : 0, (T)scale, ds);
    else
        hipLaunchKernelGGL((k_fft_generic<+1, T>), dim3((unsigned)lines), dim3(block), lds,
                           h->stream, d, es, G, GS, tw, plan, shift ? 1 : 0, (T)scale, ds);
    check_launch("k_fft_generic");
}

template <typename T>
static void contig_t(cpra_hip_ctx* h, void* data, int n, int64_t lines, int dir, bool shift,
                     double scale)
{
    switch (n) {
    case 64: launch_contig<64, T>(h, data, lines, dir, shift, scale); return;
    case 128: launch_contig<128, T>(h, data, lines, dir, shift, scale); return;
    case 192: launch_contig<192, T>(h, data, lines, dir, shift, scale); return;
    case 256: launch_contig<256, T>(h, data, lines, dir, shift, scale); return;
    case 384: launch_contig<384, T>(h, data, lines, dir, shift, scale); return;
    case 512: launch_contig<512, T>(h, data, lines, dir, shift, scale); return;
    default: launch_generic<T>(h, data, n, lines, 1, 1, n, dir, shift, scale);
    }
}

template <typename T>
static void strided_t(cpra_hip_ctx* h, void* data, int n, int64_t lines, int64_t es, int64_t G,
                      int64_t GS, int dir, bool shift, double scale)
{
    int C = 0;
    switch (n) {
    case 64: C = SizeCfg<64>::C; break;
    case 128: C = SizeCfg<128>::C; break;
    case 192: C = SizeCfg<192>::C; break;
    case 256: C = SizeCfg<256>::C; break;
    case 384: C = SizeCfg<384>::C; break;
    case 512: C = SizeCfg<512>::C; break;
    default: break;
    }
    if (C == 0 || G % C != 0 || lines % C != 0) {
        launch_generic<T>(h, data, n, lines, es, G, GS, dir, shift, scale);
        return;
    }
    switch (n) {
    case 64: launch_strided<64, T>(h, data, lines, es, G, GS, dir, shift, scale); return;
    case 128: launch_strided<128, T>(h, data, lines, es, G, GS, dir, shift, scale); return;
    case 192: launch_strided<192, T>(h, data, lines, es, G, GS, dir, shift, scale); return;
    case 256: launch_strided<256, T>(h, data, lines, es, G, GS, dir, shift, scale); return;
    case 384: launch_strided<384, T>(h, data, lines, es, G, GS, dir, shift, scale); return;
    case 512: launch_strided<512, T>(h, data, lines, es, G, GS, dir, shift, scale); return;
    }
}

void fft_pass_contig(cpra_hip_ctx* h, void* data, int n, int64_t lines, int dir, bool shift,
                     double scale, bool dbl)
{
    if (lines <= 0) return;
    if (dbl) contig_t<double>(h, data, n, lines, dir, shift, scale);
    else contig_t<float>(h, data, n, lines, dir, shift, scale);
}

void fft_pass_strided(cpra_hip_ctx* h, void* data, int n, int64_t lines, int64_t es, int64_t G,
                      int64_t GS, int dir, bool shift, double scale, bool dbl)
{
    if (lines <= 0) return;
    if (dbl) strided_t<double>(h, data, n, lines, es, G, GS, dir, shift, scale);
    else strided_t<float>(h, data, n, lines, es, G, GS, dir, shift, scale);
}

// Backward scale of the reference's DFTI descriptor (include/MKL_Impl.hpp:40),
// (T)1.0 / (m*n) evaluated in T.
static double backward_scale(int dtype, uint64_t count)
{
    if (dtype == CPRA_HIP_F32) return (double)((float)1.0 / (float)count);
    return 1.0 / (double)count;
}

void centred_2d(cpra_hip_ctx* h, void* x, int dir)
{
    const int M = (int)h->m, N = (int)h->n;
    const int64_t B = (int64_t)h->batch;
    const bool dbl = h->dtype == CPRA_HIP_F64;
    const double sc = dir > 0 ? backward_scale(h->dtype, h->m * h->n) : 1.0;
    fft_pass_contig(h, x, N, B * M, dir, true, 1.0, dbl);
    fft_pass_strided(h, x, M, B * N, N, N, (int64_t)M * N, dir, true, sc, dbl);
}

void centred_3d(cpra_hip_ctx* h, void* x, int dir)
{
    const int L = (int)h->l, M = (int)h->m, N = (int)h->n;
    const int64_t B = (int64_t)h->batch;
    const int64_t MN = (int64_t)M * N;
    const bool dbl = h->dtype == CPRA_HIP_F64;
    const double sc = dir > 0 ? backward_scale(h->dtype, h->l * h->m * h->n) : 1.0;
    fft_pass_contig(h, x, N, B * L * M, dir, true, 1.0, dbl);
    fft_pass_strided(h, x, M, B * L * N, N, N, MN, dir, true, 1.0, dbl);
    fft_pass_strided(h, x, L, B * MN, MN, MN, (int64_t)L * MN, dir, true, sc, dbl);
}

}  // namespace cpra
