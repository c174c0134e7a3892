// engine_kernels.hpp -- device kernels of the fused 2D engine (engine2d.hip).
//
// Kept in a header so the tuning variants (column-thread mapping, XCD-aware
// tile order, exact-grid specialisation) are instantiated from one source by
// the engine and its autotuner.
#pragma once
#include "fft_core.hpp"
#include "plans.hpp"

namespace cpra {

// Float <-> order-preserving unsigned key for atomicMax on floats.
__device__ __forceinline__ unsigned f2key(float f)
{
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(unsigned k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with its own L2.  With remapping, XCD x runs the contiguous tile range
// [x*G/8, (x+1)*G/8) in order, so neighbouring tiles that share cache lines
// (the 4-byte amplitude rows of adjacent column tiles) meet in one L2.
__device__ __forceinline__ unsigned tile_index(int xcd_remap)
{
    const unsigned b = blockIdx.x;
    if (!xcd_remap) return b;
    return (b & 7u) * (gridDim.x >> 3) + (b >> 3);
}

enum ColMode { COL_MOD = 0, COL_MOD_RF = 1, COL_BLUR = 2 };
enum ColMap { MAP_BLOCK = 0, MAP_WAVE = 1 };

// Thread <-> (column, t) mapping of the column kernels.  A block owns C
// consecutive columns of one pattern and has C*T = 256 threads.
//  MAP_BLOCK: lane-consecutive columns across the whole block (a 16-column
//             row segment is one 128-byte line per load); the column FFT's
//             LDS exchange spans the block's waves -> workgroup barriers.
//  MAP_WAVE:  each wave owns CW = 64/T whole columns, so the exchange stays
//             inside the wave (no workgroup barrier) at the price of 32-byte
//             row segments per wave load.
template <int M, int MAP> struct ColGeom {
    using PL = typename SizeCfg<M>::P;
    static constexpr int T = PL::T;
    static constexpr int C = SizeCfg<M>::C;
    static constexpr int CW = 64 / T;
    static_assert(MAP == MAP_BLOCK || C == 4 * CW, "MAP_WAVE needs 4 waves of CW columns");
    __device__ static int col(int tid)
    {
        if constexpr (MAP == MAP_BLOCK) return tid % C;
        else return (tid >> 6) * CW + (tid & 63) % CW;
    }
    __device__ static int t(int tid)
    {
        if constexpr (MAP == MAP_BLOCK) return tid / C;
        else return (tid & 63) / CW;
    }
};

// Column kernel: centred column DFT of the row-spectral state Y, then
//  COL_MOD / COL_MOD_RF: modulus projection with A (scale 1/(MN)); RF also
//    accumulates the R-factor sums of the incoming iterate per block;
//  COL_BLUR: multiply by the centred Gaussian transfer function (scale 1/(MN));
// then the inverse column DFT into Z.  Y and Z may alias (each block reads
// its columns fully before writing them).
// SPARSE (HIO iterations): rows of a pattern that hold no support pixel get
// the linear HIO update x <- x - beta x' for every pixel, which in the
// row-spectral domain is Y <- Y - beta*N*Z.  Those rows are finished here
// (Y rewritten in place from a re-read of the incoming Y; error reduction
// writes 0) and only rows with support pixels are written to Z for the row
// kernel (rowflag[b*M + row] != 0).
// The kernel body works on one tile (C columns of one pattern) given by the
// caller, with the caller's LDS (M*C elements; `red` 2*TPB floats for
// COL_MOD_RF), so both the single-size kernel below and the mixed-size
// kernels (one launch over several pattern sizes) run it.
template <int M, int MODE, int MAP, bool SPARSE>
__device__ __forceinline__ void eng_cols_tile(
    const unsigned tile, c2<float>* lds, float* red, const c2<float>* Y, c2<float>* Z,
    const float* __restrict__ A, int64_t amp_batch, int N, float scale, int skip_zero,
    float* __restrict__ rf_part, float gcoef_r, float gcoef_c, const unsigned* __restrict__ rowbits,
    float beta_n, int error_reduction, int pattern0)
{
    using V = c2<float>;
    using G = ColGeom<M, MAP>;
    using PL = typename G::PL;
    constexpr int E = PL::E, TT = PL::T, C = G::C;
    const int tid = threadIdx.x;
    const int c = G::col(tid), t = G::t(tid);
    // block-uniform pattern and first column; per-lane 32-bit offsets
    const unsigned tiles_per_pattern = (unsigned)N / C;
    const unsigned b = tile / tiles_per_pattern;
    const int col0 = (int)(tile - b * tiles_per_pattern) * C;
    const int col = col0 + c;
    const size_t pat = (size_t)M * (size_t)N;
    const c2<float>* Yp = Y + b * pat + col0;
    c2<float>* Zp = Z + b * pat + col0;
    // element m of this lane sits at row t + rot(m)*T: a wave-uniform row base
    // (rot(m)*T*N) plus one per-lane offset (t*N + c) shared by all m, so each
    // access is an SGPR base + one VGPR offset
    const unsigned toff = (unsigned)t * (unsigned)N + (unsigned)c;
    auto rowbase = [&](int m) { return (size_t)(PL::rot(m) * TT) * (size_t)N; };
    V v[E];
    float av[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = (Yp + rowbase(m))[toff];
    });
    if constexpr (MODE != COL_BLUR) {
        // the amplitudes are needed after the forward FFT; load them now so
        // their latency overlaps the transform
        // pattern0: index of this launch's first pattern in the whole batch
        // (chunked launches), for the amplitude broadcast b % amp_batch
        const float* Ap = A + (size_t)(((unsigned)pattern0 + b) % (unsigned)amp_batch) * pat + col0;
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            av[m] = (Ap + rowbase(m))[toff];
        });
    }
    float rnum = 0.f, rden = 0.f;
    auto body = [&](auto& acc) {
        fft_line<PL, -1>(v, t, g_tw<M>(), acc);
        if constexpr (MODE == COL_BLUR) {
            const float dc = (float)(col - N / 2);
            const float gc = gcoef_c * dc * dc;
            sfor<E>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                const int row = t + PL::rot(m) * TT;
                const float dr = (float)(row - M / 2);
                const float g = expf(gcoef_r * dr * dr + gc) * scale;
                v[m] = cscale(v[m], g);
            });
        } else {
            sfor<E>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                const float a = av[m];
                const V x = v[m];
                const float m2 = x.x * x.x + x.y * x.y;
                if constexpr (MODE == COL_MOD_RF) {
                    if (a > 0.f) {
                        rnum += fabsf(sqrtf(m2) - a);
                        rden += a;
                    }
                }
                // x * A/|x| (the reference's polar(A, arg x), MKL_Impl.inl:102-103)
                float s = scale;
                if (!(skip_zero && a == 0.f) && m2 > 0.f) s = scale * a * rsqrtf(m2);
                v[m] = cscale(x, s);
            });
        }
        fft_line<PL, +1>(v, t, g_tw<M>(), acc);
    };
    if constexpr (MAP == MAP_BLOCK) {
        LdsCol<V, C, false> acc{lds + c};
        body(acc);
    } else {
        const int wave = tid >> 6, cw = (tid & 63) % G::CW;
        LdsCol<V, G::CW, true> acc{lds + wave * M * G::CW + cw};
        body(acc);
    }
    if constexpr (SPARSE) {
        // row r of this pattern holds support pixels iff bit r of rowbits;
        // rows t + rot(m)*T of one register lie in one 32-bit word for every
        // plan (T <= 32 and T | 32), so the word index is wave-uniform
        static_assert(TT <= 32 && 32 % TT == 0, "row bits need T | 32");
        const unsigned* rb = rowbits + (size_t)b * (M / 32);
        c2<float>* Yw = const_cast<c2<float>*>(Yp);
        static_assert(E % 4 == 0, "E must be a multiple of 4");
        sfor<E / 4>([&](auto gc) {
            constexpr int g = decltype(gc)::value;
            bool sup[4];
            V y[4];
            sfor<4>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr int m = 4 * g + k;
                constexpr int r0 = PL::rot(m) * TT;
                const int row = t + r0;
                sup[k] = (rb[r0 >> 5] >> (row & 31)) & 1u;
                if (!sup[k] && !error_reduction) y[k] = (Yp + rowbase(m))[toff];
            });
            sfor<4>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr int m = 4 * g + k;
                if (sup[k]) (Zp + rowbase(m))[toff] = v[m];
                else if (error_reduction) (Yw + rowbase(m))[toff] = mk<V>(0.f, 0.f);
                else (Yw + rowbase(m))[toff] = mk<V>(y[k].x - beta_n * v[m].x, y[k].y - beta_n * v[m].y);
            });
            __builtin_amdgcn_sched_barrier(0);
        });
    } else {
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            (Zp + rowbase(m))[toff] = v[m];
        });
    }
    if constexpr (MODE == COL_MOD_RF) {
        constexpr int TPB = SizeCfg<M>::TPB;
        red[tid] = rnum;
        red[TPB + tid] = rden;
        __syncthreads();
        for (int s = TPB / 2; s > 0; s >>= 1) {
            if (tid < s) {
                red[tid] += red[tid + s];
                red[TPB + tid] += red[TPB + tid + s];
            }
            __syncthreads();
        }
        if (tid == 0) {
            rf_part[2 * tile] = red[0];
            rf_part[2 * tile + 1] = red[TPB];
        }
    }
}

template <int M, int MODE, int MAP, bool SPARSE = false>
__global__ __launch_bounds__(SizeCfg<M>::TPB) void k_eng_cols(
    const c2<float>* Y, c2<float>* Z, const float* __restrict__ A, int64_t amp_batch, int N,
    float scale, int skip_zero, float* __restrict__ rf_part, float gcoef_r, float gcoef_c,
    int xcd_remap, const unsigned* __restrict__ rowbits = nullptr, float beta_n = 0.f,
    int error_reduction = 0, int pattern0 = 0)
{
    __shared__ c2<float> lds[M * SizeCfg<M>::C];
    if constexpr (MODE == COL_MOD_RF) {
        __shared__ float red[2 * SizeCfg<M>::TPB];
        eng_cols_tile<M, MODE, MAP, SPARSE>(tile_index(xcd_remap), lds, red, Y, Z, A, amp_batch, N, scale,
                                            skip_zero, rf_part, gcoef_r, gcoef_c, rowbits, beta_n,
                                            error_reduction, pattern0);
    } else {
        eng_cols_tile<M, MODE, MAP, SPARSE>(tile_index(xcd_remap), lds, nullptr, Y, Z, A, amp_batch, N,
                                            scale, skip_zero, rf_part, gcoef_r, gcoef_c, rowbits, beta_n,
                                            error_reduction, pattern0);
    }
}

enum RowMode { ROW_HIO = 0, ROW_ER = 1 };

// Row kernel of the HIO iteration (Fienup 1982; orc_hio in the oracle):
// x' = IDFT_r(Z), x = IDFT_r(Y)/N, HIO update with the support (and
// positivity), Y <- DFT_r(x_new).  A line's T threads are in one wave, so the
// LDS exchanges use wave-scope ordering.  EXACT: the grid covers the lines
// exactly (no bounds checks).
// rowbits != NULL (sparse rows): lines without a support pixel were finished
// by the column kernel; they are skipped, and a wave whose lines are all
// finished returns at once (the row exchange never leaves the wave).
template <int N, int MODE, bool EXACT, bool SPARSE>
__global__ __launch_bounds__(SizeCfg<N>::TPB) void k_eng_rows_hio(
    const c2<float>* __restrict__ Z, c2<float>* __restrict__ Y, const uint8_t* __restrict__ mask,
    int64_t nlines, float beta, float invN, int real_positive,
    const unsigned* __restrict__ rowbits)
{
    using V = c2<float>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = SizeCfg<N>::LPB;
    constexpr int SPAN = LdsRow<V, N>::span();
    __shared__ V lds[LPB * SPAN];
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    const int64_t line = (int64_t)blockIdx.x * LPB + ll;
    bool valid = EXACT || line < nlines;
    if constexpr (SPARSE) {
        valid = valid && ((rowbits[line >> 5] >> (line & 31)) & 1u);
        if (__ballot(valid) == 0ull) return;
    }
    // block-uniform base, per-lane 32-bit offsets
    const size_t bbase = (size_t)blockIdx.x * LPB * N;
    const int lo = ll * N + t;
    const c2<float>* Zb = Z + bbase;
    c2<float>* Yb = Y + bbase;
    const uint8_t* Sb = mask + bbase;
    LdsRow<V, N, true> acc{lds + ll * SPAN};
    const V* tw = g_tw<N>();
    V xp[E], xo[E];
    uint8_t sm[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const int o = lo + PL::rot(m) * TT;
        xp[m] = valid ? Zb[o] : mk<V>(0, 0);
        xo[m] = valid ? Yb[o] : mk<V>(0, 0);
        sm[m] = valid ? Sb[o] : (uint8_t)0;
    });
    fft_line<PL, +1>(xp, t, tw, acc);
    fft_line<PL, +1>(xo, t, tw, acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const bool s = sm[m] != 0;
        const V p = xp[m];
        const bool ok = s && (!real_positive || p.x >= 0.f);
        V r;
        if constexpr (MODE == ROW_ER) {
            r = ok ? (real_positive ? mk<V>(p.x, 0.f) : p) : mk<V>(0.f, 0.f);
        } else {
            const V o = mk<V>(xo[m].x * invN - beta * p.x, xo[m].y * invN - beta * p.y);
            r = ok ? (real_positive ? mk<V>(p.x, 0.f) : p) : o;
        }
        xo[m] = r;
    });
    fft_line<PL, -1>(xo, t, tw, acc);
    if (valid) {
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            Yb[lo + PL::rot(m) * TT] = xo[m];
        });
    }
}

// Shrinkwrap, row stage 1: W = DFT_r(|x|), x = IDFT_r(Y)/N (FROM_STATE) or x
// read directly.
template <int N, bool FROM_STATE>
__global__ __launch_bounds__(SizeCfg<N>::TPB) void k_sw_mag(const c2<float>* __restrict__ Y,
                                                           c2<float>* __restrict__ W,
                                                           int64_t nlines, float invN)
{
    using V = c2<float>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = SizeCfg<N>::LPB;
    constexpr int SPAN = LdsRow<V, N>::span();
    __shared__ V lds[LPB * SPAN];
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    const int64_t line = (int64_t)blockIdx.x * LPB + ll;
    const bool valid = line < nlines;
    const int64_t base = (valid ? line : 0) * N;
    LdsRow<V, N, true> acc{lds + ll * SPAN};
    const V* tw = g_tw<N>();
    V v[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = valid ? Y[base + t + PL::rot(m) * TT] : mk<V>(0, 0);
    });
    if constexpr (FROM_STATE) fft_line<PL, +1>(v, t, tw, acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const float sc = FROM_STATE ? invN : 1.f;
        const float a = sqrtf(v[m].x * v[m].x + v[m].y * v[m].y) * sc;
        v[m] = mk<V>(a, 0.f);
    });
    fft_line<PL, -1>(v, t, tw, acc);
    if (valid) {
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            W[base + t + PL::rot(m) * TT] = v[m];
        });
    }
}

// Shrinkwrap, row stage 3: blurred image b = Re IDFT_r(W) and per-pattern
// maximum (reduced in the wave, then per block, then one atomic per block
// when the block's lines belong to one pattern).
template <int N>
__global__ __launch_bounds__(SizeCfg<N>::TPB) void k_sw_blurmax(const c2<float>* __restrict__ W,
                                                               float* __restrict__ bl,
                                                               unsigned* __restrict__ keys,
                                                               int64_t nlines, int M)
{
    using V = c2<float>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = SizeCfg<N>::LPB;
    constexpr int SPAN = LdsRow<V, N>::span();
    __shared__ V lds[LPB * SPAN];
    __shared__ unsigned bmax[LPB];
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    const int64_t line = (int64_t)blockIdx.x * LPB + ll;
    const bool valid = line < nlines;
    const int64_t base = (valid ? line : 0) * N;
    LdsRow<V, N, true> acc{lds + ll * SPAN};
    const V* tw = g_tw<N>();
    V v[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = valid ? W[base + t + PL::rot(m) * TT] : mk<V>(0, 0);
    });
    fft_line<PL, +1>(v, t, tw, acc);
    float mx = -INFINITY;
    if (valid) {
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            bl[base + t + PL::rot(m) * TT] = v[m].x;
            mx = fmaxf(mx, v[m].x);
        });
    }
    // max over the line's T lanes (consecutive lanes of one wave)
    for (int off = TT / 2; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    if (t == 0) bmax[ll] = valid ? f2key(mx) : 0u;
    __syncthreads();
    if (tid == 0) {
        const int64_t first = (int64_t)blockIdx.x * LPB;
        const int64_t last = first + LPB - 1 < nlines ? first + LPB - 1 : nlines - 1;
        if (first / M == last / M) {
            unsigned k = 0u;
            for (int i = 0; i < LPB; ++i) k = bmax[i] > k ? bmax[i] : k;
            atomicMax(&keys[first / M], k);
        } else {
            for (int i = 0; i < LPB && first + i < nlines; ++i)
                atomicMax(&keys[(first + i) / M], bmax[i]);
        }
    }
}

}  // namespace cpra

namespace cpra {

// ---------------------------------------------------------------------------
// difference map (reference variant, src/CPRA_MKL.cpp:94-110, and Elser's
// textbook map).  State: psi in pattern space, Y1 = DFT_r(t1) and
// Y2 = DFT_r(psi/beta + (1 - 1/beta) P_S t2) in the row-spectral domain (the
// textbook map uses t1 = t2 = psi).  The column kernel turns [Y1; Y2] (2B
// patterns, one launch) into [Z1; Z2] with Z = IDFT_c(P_M(DFT_c Y))/(MN), so
// IDFT_r(Z1) = P_M t1 and IDFT_r(Z2) = P_M(psi/beta + (1-1/beta) P_S t2).
// ---------------------------------------------------------------------------
enum SpaceMode { SPACE_MKL = 0, SPACE_POS = 1, SPACE_CPLX = 2 };

// P_S with a u8 support: SPACE_MKL is the reference's SpaceConstraint on the
// equivalent mask (im <- 0; re <- 0 outside, MKL_Impl.inl:79-81), SPACE_POS
// also clamps re < 0 to 0, SPACE_CPLX keeps the complex value inside.
template <int SPACE> __device__ __forceinline__ c2<float> ps_apply(c2<float> z, bool in)
{
    if constexpr (SPACE == SPACE_CPLX) return in ? z : mk<c2<float>>(0.f, 0.f);
    else if constexpr (SPACE == SPACE_POS) return mk<c2<float>>(in && z.x > 0.f ? z.x : 0.f, 0.f);
    else return mk<c2<float>>(in ? z.x : 0.f, 0.f);
}

// Entry: Y1 = DFT_r(t1 | psi), Y2 = DFT_r(psi/beta + (1-1/beta) P_S(t2 | psi)).
// Y1 and Y2 are the two halves of one 2B-pattern array (Y2 = Y1 + B*M*N).
template <int N, bool TEXTBOOK, int SPACE>
__global__ __launch_bounds__(SizeCfg<N>::TPB) void k_dm_entry(
    const c2<float>* __restrict__ psi, const c2<float>* __restrict__ t1,
    const c2<float>* __restrict__ t2, const uint8_t* __restrict__ mask,
    c2<float>* __restrict__ Y1, c2<float>* __restrict__ Y2, int64_t nlines, float beta)
{
    using V = c2<float>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = SizeCfg<N>::LPB;
    constexpr int SPAN = LdsRow<V, N>::span();
    __shared__ V lds[LPB * SPAN];
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    const int64_t line = (int64_t)blockIdx.x * LPB + ll;
    const bool valid = line < nlines;
    const size_t base = (size_t)(valid ? line : 0) * N + t;
    LdsRow<V, N, true> acc{lds + ll * SPAN};
    const V* tw = g_tw<N>();
    const float ib = 1.f / beta, cb = 1.f - 1.f / beta;
    V a[E], u[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const size_t o = base + PL::rot(m) * TT;
        const V p = valid ? psi[o] : mk<V>(0, 0);
        const V s1 = TEXTBOOK ? p : (valid ? t1[o] : mk<V>(0, 0));
        const V s2 = TEXTBOOK ? p : (valid ? t2[o] : mk<V>(0, 0));
        const V q = ps_apply<SPACE>(s2, valid && mask[o] != 0);
        a[m] = s1;
        u[m] = mk<V>(fmaf(cb, q.x, p.x * ib), fmaf(cb, q.y, p.y * ib));
    });
    fft_line<PL, -1>(a, t, tw, acc);
    fft_line<PL, -1>(u, t, tw, acc);
    if (valid) {
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const size_t o = base + PL::rot(m) * TT;
            Y1[o] = a[m];
            Y2[o] = u[m];
        });
    }
}

// One DM iteration's row pass (after the column kernel on [Y1; Y2]):
//   a   = IDFT_r(Z1) = P_M t1
//   t1' = P_S(-psi/beta + (1 + 1/beta) a)                 CPRA_MKL.cpp:95-99
//   t2' = IDFT_r(Z2) = P_M(psi/beta + (1 - 1/beta) P_S t2)      :102-106
//   psi' = (psi + beta t1') - beta t2'                          :109-110
// then Y1 = DFT_r(t1' | psi'), Y2 = DFT_r(psi'/beta + (1-1/beta) P_S(t2' | psi'))
// for the next iteration; LAST writes psi', t1', t2' in pattern space instead.
// The body handles the LPB lines of block `vb` with the caller's LDS
// (LPB * LdsRow<V, N>::span() elements); k_dm_rows and the mixed-size kernel
// run it.
template <int N, bool TEXTBOOK, int SPACE, bool LAST>
__device__ __forceinline__ void dm_rows_block(
    const unsigned vb, c2<float>* lds, const c2<float>* __restrict__ Z1, const c2<float>* __restrict__ Z2,
    c2<float>* __restrict__ psi, const uint8_t* __restrict__ mask, c2<float>* __restrict__ Y1,
    c2<float>* __restrict__ Y2, c2<float>* __restrict__ t1out, c2<float>* __restrict__ t2out,
    int64_t nlines, float beta)
{
    using V = c2<float>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = SizeCfg<N>::LPB;
    constexpr int SPAN = LdsRow<V, N>::span();
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    const int64_t line = (int64_t)vb * LPB + ll;
    const bool valid = line < nlines;
    const size_t base = (size_t)(valid ? line : 0) * N + t;
    LdsRow<V, N, true> acc{lds + ll * SPAN};
    const V* tw = g_tw<N>();
    const float ib = 1.f / beta, pb = 1.f + 1.f / beta, cb = 1.f - 1.f / beta;
    V a[E], b[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const size_t o = base + PL::rot(m) * TT;
        a[m] = valid ? Z1[o] : mk<V>(0, 0);
        b[m] = valid ? Z2[o] : mk<V>(0, 0);
    });
    fft_line<PL, +1>(a, t, tw, acc);
    fft_line<PL, +1>(b, t, tw, acc);
    // a <- t1' (or the next Y1 source), b <- t2' (or next Y2 source)
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const size_t o = base + PL::rot(m) * TT;
        const bool in = valid && mask[o] != 0;
        const V p = valid ? psi[o] : mk<V>(0, 0);
        // explicit fmaf: the same operations whatever the kernel inlining this
        // body leaves the compiler free to contract (the launch modes of the
        // mixed-size engine agree bit for bit)
        const V w = mk<V>(fmaf(pb, a[m].x, -ib * p.x), fmaf(pb, a[m].y, -ib * p.y));
        const V t1n = ps_apply<SPACE>(w, in);
        const V t2n = b[m];
        V pn = mk<V>(fmaf(beta, t1n.x, p.x), fmaf(beta, t1n.y, p.y));
        pn = mk<V>(fmaf(-beta, t2n.x, pn.x), fmaf(-beta, t2n.y, pn.y));
        if (valid) psi[o] = pn;
        if constexpr (LAST) {
            a[m] = t1n;
            b[m] = t2n;
        } else {
            const V s1 = TEXTBOOK ? pn : t1n;
            const V q = ps_apply<SPACE>(TEXTBOOK ? pn : t2n, in);
            a[m] = s1;
            b[m] = mk<V>(fmaf(cb, q.x, pn.x * ib), fmaf(cb, q.y, pn.y * ib));
        }
    });
    if constexpr (LAST) {
        if (valid) {
            sfor<E>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                const size_t o = base + PL::rot(m) * TT;
                if (t1out) t1out[o] = a[m];
                if (t2out) t2out[o] = b[m];
            });
        }
    } else {
        fft_line<PL, -1>(a, t, tw, acc);
        fft_line<PL, -1>(b, t, tw, acc);
        if (valid) {
            sfor<E>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                const size_t o = base + PL::rot(m) * TT;
                Y1[o] = a[m];
                Y2[o] = b[m];
            });
        }
    }
}

template <int N, bool TEXTBOOK, int SPACE, bool LAST>
__global__ __launch_bounds__(SizeCfg<N>::TPB) void k_dm_rows(
    const c2<float>* __restrict__ Z1, const c2<float>* __restrict__ Z2,
    c2<float>* __restrict__ psi, const uint8_t* __restrict__ mask, c2<float>* __restrict__ Y1,
    c2<float>* __restrict__ Y2, c2<float>* __restrict__ t1out, c2<float>* __restrict__ t2out,
    int64_t nlines, float beta)
{
    __shared__ c2<float> lds[SizeCfg<N>::LPB * LdsRow<c2<float>, N>::span()];
    dm_rows_block<N, TEXTBOOK, SPACE, LAST>(blockIdx.x, lds, Z1, Z2, psi, mask, Y1, Y2, t1out, t2out,
                                            nlines, beta);
}

}  // namespace cpra
