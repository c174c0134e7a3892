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

// Tile-major layout of an N x N pattern (the HIO engine's state when tiled):
// the columns are cut into tiles of C = SizeCfg<N>::C (the column kernel's
// tile width) and element (r, c) sits at ((c / C) * N + r) * C + c % C.  A
// column-kernel block then reads and writes one contiguous N*C block instead
// of N row segments of C*8 bytes N*8 bytes apart, and a row-kernel wave's
// load of one register over its 64/T rows is one contiguous run.
// TT: threads per row line of the kernel that addresses the layout (the
// layout itself depends on C only).
template <int N, int TT = SizeCfg<N>::P::T> struct Tiled {
    static constexpr int C = SizeCfg<N>::C;
    static constexpr int T = TT;
    static_assert(C % T == 0 || T % C == 0, "tile width and row threads must nest");
    // offset in a pattern of element c = K*T + t of row r, split into a
    // compile-time part per register K (a wave-uniform base: the access
    // becomes SGPR base + one 32-bit VGPR offset, as the strides between
    // registers exceed the load instructions' immediate offsets) and the
    // lane part shared by all registers
    template <int K> static constexpr size_t kpart()
    {
        if constexpr (C % T == 0) return (size_t)((K / (C / T)) * N * C + (K % (C / T)) * T);
        else return (size_t)K * (T / C) * N * C;
    }
    // lane part for row r = r0 + rl: r0 (block-uniform) goes into the base
    // as r0 * C; rl and t come from threadIdx, so the compiler knows the
    // offset's range and can use it as a 32-bit VGPR offset
    __device__ static unsigned lane(int t, int rl)
    {
        if constexpr (C % T == 0) return (unsigned)(rl * C + t);
        else return (unsigned)((t / C) * N * C + rl * C + t % C);
    }
};

// complex64 access through a buffer resource: SGPR base, SGPR byte offset
// (soff, a compile-time constant per register) and one VGPR byte offset per
// lane.  Used where the strides between a lane's registers exceed the
// immediate offsets of global instructions, which would otherwise cost a
// 64-bit VGPR address per register.  Raw buffer (stride 0): num_records is a
// byte count; the kernels only form offsets inside the pattern.
struct BufC64 {
    __amdgpu_buffer_rsrc_t r;
    __device__ BufC64(const void* base, unsigned bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                              0x00020000))
    {
    }
    __device__ c2<float> ld(unsigned voff, unsigned soff) const
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        return __builtin_bit_cast(c2<float>, v);
    }
    __device__ void st(c2<float> x, unsigned voff, unsigned soff) const
    {
        using B = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0u, 0u, 0));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B, x), r, voff, soff, 0);
    }
};

enum ColMode { COL_MOD = 0, COL_MOD_RF = 1 };
enum ColMap { MAP_BLOCK = 0, MAP_WAVE = 1, MAP_WIDE = 2 };

// Thread <-> (column, t) mapping of the column kernels.  A block owns C
// consecutive columns of one pattern and has TPB = C*T threads.
//  MAP_BLOCK: lane-consecutive columns across the whole block (a 16-column
//             row segment is one 128-byte line per load); the column FFT's
//             LDS exchange spans the block's waves -> workgroup barriers.
//  MAP_WAVE:  each wave owns CW = 64/T whole columns, so the exchange stays
//             inside the wave (no workgroup barrier) at the price of 32-byte
//             row segments per wave load.
//  MAP_WIDE:  MAP_BLOCK over twice the columns with twice the threads (512):
//             256-byte row segments per load.  Column-strided HBM reads run
//             closer to the contiguous rate the wider the segment; 512-byte
//             segments (1024 threads) lose more to the lower block count.
template <int M, int MAP> struct ColGeom {
    using PL = typename SizeCfg<M>::P;
    static constexpr int T = PL::T;
    static constexpr int C = (MAP == MAP_WIDE ? 2 : 1) * SizeCfg<M>::C;
    static constexpr int TPB = (MAP == MAP_WIDE ? 2 : 1) * SizeCfg<M>::TPB;
    static constexpr int CW = 64 / T;
    static_assert(MAP != MAP_WAVE || C == 4 * CW, "MAP_WAVE needs 4 waves of CW columns");
    static_assert(C * T == TPB, "column tile");
    __device__ static int col(int tid)
    {
        if constexpr (MAP != MAP_WAVE) return tid % C;
        else return (tid >> 6) * CW + (tid & 63) % CW;
    }
    __device__ static int t(int tid)
    {
        if constexpr (MAP != MAP_WAVE) return tid / C;
        else return (tid & 63) / CW;
    }
};

// Column kernel: centred column DFT of the row-spectral state Y, then
//  COL_MOD / COL_MOD_RF: modulus projection with A (scale 1/(MN)); RF also
//    accumulates the R-factor sums of the incoming iterate per block;
// then the inverse column DFT into Z.  Y and Z may alias (each block reads
// its columns fully before writing them).
// SPARSE (HIO iterations): rows of a pattern that hold no support pixel get
// the linear HIO update x <- x - beta x' for every pixel, which in the
// row-spectral domain is Y <- Y - beta*N*Z.  Those rows are finished here
// (Y rewritten in place from the incoming Y kept in registers; error reduction
// writes 0) and only rows with support pixels are written to Z for the row
// kernel (rowflag[b*M + row] != 0).
// The kernel body works on one tile (C columns of one pattern) given by the
// caller, with the caller's LDS (M*C elements; `red` 2*TPB floats for
// COL_MOD_RF), so both the single-size kernel below and the mixed-size
// kernels (one launch over several pattern sizes) run it.
// TILED: Y, Z and A are in the tile-major layout (Tiled<M>, square patterns):
// the block's tile is one contiguous M x C block with row stride C.
template <int M, int MODE, int MAP, bool SPARSE, bool TILED = false>
__device__ __forceinline__ void eng_cols_tile(
    const unsigned tile, c2<float>* lds, float* red, const c2<float>* Y, c2<float>* Z,
    const float* __restrict__ A, int64_t amp_batch, int N, float scale, int skip_zero,
    float* __restrict__ rf_part, const unsigned* __restrict__ rowbits,
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
    const size_t pat = (size_t)M * (size_t)N;
    static_assert(!TILED || MAP != MAP_WIDE, "tiled state uses the layout's tile width");
    // row stride and the block's first element: row-major (stride N, column
    // col0 of pattern b) or tile-major (stride C, tile `tile` of the launch)
    const unsigned RS = TILED ? (unsigned)C : (unsigned)N;
    const size_t blk0 = TILED ? (size_t)tile * M * C : b * pat + col0;
    const c2<float>* Yp = Y + blk0;
    c2<float>* Zp = Z + blk0;
    // element m of this lane sits at row t + rot(m)*T: a wave-uniform row base
    // (rot(m)*T*RS) plus one per-lane offset (t*RS + c) shared by all m, so
    // each access is an SGPR base + one VGPR offset
    const unsigned toff = (unsigned)t * RS + (unsigned)c;
    auto rowbase = [&](int m) { return (size_t)(PL::rot(m) * TT) * (size_t)RS; };
    V v[E];
    float av[E];
    // sparse rows: the incoming Y stays in registers for the rows this kernel
    // finishes (32 more VGPRs, three waves per SIMD instead of four, but no
    // second read of Y: 128-133 against 146-150 us per pass at 35-50 %
    // support rows, 512 x 256^2)
    V y0[SPARSE ? E : 1];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = (Yp + rowbase(m))[toff];
        if constexpr (SPARSE) y0[m] = v[m];
    });
    // the amplitudes are needed after the forward FFT; load them now so their
    // latency overlaps the transform.  pattern0: index of this launch's first
    // pattern in the whole batch (chunked launches), for the amplitude
    // broadcast b % amp_batch.  Tile-major amplitudes are loaded
    // non-temporal: read once per iteration, they then leave the Infinity
    // Cache to Y and Z, which the row kernel reads next (512 x 256^2 in two
    // sub-batches: column pass 57.9 -> 56.6 us, row pass 48.8 -> 46.0,
    // 4,394 -> 4,554 it/s).  Row-major tiles read half of each 128-byte line
    // and the neighbouring tile the other half, which a non-temporal load
    // would fetch twice (86 against 57 us per launch there), so they keep
    // ordinary loads.
    const size_t abase = (size_t)(((unsigned)pattern0 + b) % (unsigned)amp_batch) * pat;
    const float* Ap = A + abase + (TILED ? (size_t)(tile - b * tiles_per_pattern) * M * C : (size_t)col0);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        if constexpr (TILED) av[m] = __builtin_nontemporal_load((Ap + rowbase(m)) + toff);
        else av[m] = (Ap + rowbase(m))[toff];
    });
    float rnum = 0.f, rden = 0.f;
    auto body = [&](auto& acc) {
        fft_line<PL, -1>(v, t, g_tw<M>(), acc);
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
        fft_line<PL, +1>(v, t, g_tw<M>(), acc);
    };
    if constexpr (MAP != MAP_WAVE) {
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
        // every row of the pattern listed (the adaptive mode's dense choice):
        // the dense kernel's stores, without the per-row tests
        unsigned every = 0xffffffffu;
        for (int w = 0; w < (M + 31) / 32; ++w) every &= rb[w];
        if (M % 32 == 0 && every == 0xffffffffu) {
            sfor<E>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                (Zp + rowbase(m))[toff] = v[m];
            });
        } else
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
                if (!sup[k] && !error_reduction) y[k] = y0[m];
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
        constexpr int TPB = G::TPB;
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

template <int M, int MODE, int MAP, bool SPARSE = false, bool TILED = false>
__global__ __launch_bounds__((ColGeom<M, MAP>::TPB)) void k_eng_cols(
    const c2<float>* Y, c2<float>* Z, const float* __restrict__ A, int64_t amp_batch, int N,
    float scale, int skip_zero, float* __restrict__ rf_part,
    int xcd_remap, const unsigned* __restrict__ rowbits = nullptr, float beta_n = 0.f,
    int error_reduction = 0, int pattern0 = 0)
{
    __shared__ c2<float> lds[M * ColGeom<M, MAP>::C];
    if constexpr (MODE == COL_MOD_RF) {
        __shared__ float red[2 * ColGeom<M, MAP>::TPB];
        eng_cols_tile<M, MODE, MAP, SPARSE, TILED>(tile_index(xcd_remap), lds, red, Y, Z, A, amp_batch, N, scale,
                                            skip_zero, rf_part, rowbits, beta_n,
                                            error_reduction, pattern0);
    } else {
        eng_cols_tile<M, MODE, MAP, SPARSE, TILED>(tile_index(xcd_remap), lds, nullptr, Y, Z, A, amp_batch, N,
                                            scale, skip_zero, rf_part, rowbits, beta_n,
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
// TILED: Z and Y in the tile-major layout of N x N patterns (Tiled<N>), the
// mask row-major; a block's LPB lines then lie in one pattern (LPB | N).
// PV: the radix-plan variant (HioRowPlan).
// The support bit of element m of this thread's line (position t + rot(m) T,
// T the plan's threads per line) from the line's N/32 packed words sw
// (k_pack_support: bit j of word w is pixel 32 w + j).  T divides 32 or 32
// divides T, so the word index is a compile-time constant (plus t / 32 when
// T > 32) and the bit a lane offset.
template <class PL, int m>
__device__ __forceinline__ bool support_bit(const unsigned* sw, int t)
{
    constexpr int TT = PL::T;
    constexpr int base = PL::rot(m) * TT;
    static_assert(32 % TT == 0 || TT % 32 == 0, "threads per line and 32 must nest");
    if constexpr (TT <= 32) {
        return (sw[base >> 5] >> (t + (base & 31))) & 1u;
    } else if constexpr (TT == 64) {
        return (((t >> 5) ? sw[(base >> 5) + 1] : sw[base >> 5]) >> (t & 31)) & 1u;
    } else {
        static_assert(TT <= 64, "at most 64 threads per line");
        return false;
    }
}

// The N/32 packed support words of a line into registers (16-byte loads when
// N is a multiple of 128, else 8-byte; every lane of the line loads the same
// words).
template <int N>
__device__ __forceinline__ void load_support_words(unsigned (&sw)[N / 32], const unsigned* __restrict__ p,
                                                   bool ok)
{
    if constexpr (N % 128 == 0) {
        sfor<N / 128>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const uint4 w = ok ? reinterpret_cast<const uint4*>(p)[q] : make_uint4(0, 0, 0, 0);
            sw[4 * q] = w.x;
            sw[4 * q + 1] = w.y;
            sw[4 * q + 2] = w.z;
            sw[4 * q + 3] = w.w;
        });
    } else {
        static_assert(N % 64 == 0, "whole 8-byte pairs");
        sfor<N / 64>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const uint2 w = ok ? reinterpret_cast<const uint2*>(p)[q] : make_uint2(0, 0);
            sw[2 * q] = w.x;
            sw[2 * q + 1] = w.y;
        });
    }
}

// Whether the HIO row kernel's block can stage its rows as [N][LPB + 1] in
// its LDS for the line-major store (LT).
template <int N, int PV> constexpr bool kRowsLinesFit =
    HioRowPlan<N, PV>::LPB * LdsRow<c2<float>, N>::span() >= N * (HioRowPlan<N, PV>::LPB + 1);

// sbits: the support packed by k_pack_support, N/32 words per line (line-
// major, whatever the state's layout).  LT: dense row-major patterns whose
// updated rows are stored line-major into lt (lt[p][v][u] = Y[p][u][v], the
// carousel merge's layout) instead of Y.
template <int N, int MODE, bool EXACT, bool SPARSE, bool TILED = false, int PV = 0, bool LT = false>
__global__ __launch_bounds__((HioRowPlan<N, PV>::TPB)) void k_eng_rows_hio(
    const c2<float>* __restrict__ Z, c2<float>* __restrict__ Y, const unsigned* __restrict__ sbits,
    int64_t nlines, float beta, float invN, int real_positive,
    const unsigned* __restrict__ rowbits, const unsigned* __restrict__ natural_order = nullptr,
    c2<float>* __restrict__ lt = nullptr)
{
    using V = c2<float>;
    using PL = typename HioRowPlan<N, PV>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = HioRowPlan<N, PV>::LPB;
    using TL = Tiled<N, TT>;
    constexpr int SPAN = LdsRow<V, N>::span();
    __shared__ V lds[LPB * SPAN];
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    // Sparse rows of tile-major patterns.  The dispatcher deals consecutive
    // blocks to the eight XCDs in turn, so in the natural order row block j
    // of every pattern lands on the same XCD, and the support rows (the
    // middle rows of every pattern) on a few XCDs only.  With whole groups of
    // eight patterns per launch, block i goes to pattern 8 g + i % 8 of group
    // g: each XCD walks through the row blocks of its own patterns in order,
    // which balances the XCDs whatever the support and keeps each XCD's reads
    // within one pattern.  Otherwise consecutive blocks take the same row
    // block of consecutive patterns, unless natural_order[0] != 0 (rows
    // judged dense, every row bit set), where the natural order measured
    // faster (152 against 188 us per pass at 512 x 256^2).  The grid is
    // whole patterns.
    unsigned vb = blockIdx.x;
    if constexpr (SPARSE && TILED) {
        constexpr unsigned BPP = N / LPB;
        const unsigned npat = gridDim.x / BPP;
        if (npat % 8 == 0) {
            const unsigned g = blockIdx.x / (8 * BPP);
            vb = (g * 8 + (blockIdx.x & 7u)) * BPP + (blockIdx.x >> 3) % BPP;
        } else if (natural_order == nullptr || natural_order[0] == 0u) {
            vb = (blockIdx.x % npat) * BPP + blockIdx.x / npat;
        }
    }
    const int64_t line = (int64_t)vb * LPB + ll;
    // inb: the line exists (loads); valid: it is also written (sparse rows
    // skip the lines the column kernel finished).  The loads test inb only,
    // which is a compile-time true for EXACT grids: per-lane conditions on
    // them cost the scheduler its batching of the loads (a third slower).
    const bool inb = EXACT || line < nlines;
    bool valid = inb;
    if constexpr (SPARSE) {
        valid = valid && ((rowbits[line >> 5] >> (line & 31)) & 1u);
        if (__ballot(valid) == 0ull) return;
    }
    static_assert(!TILED || (EXACT && N % LPB == 0), "tiled rows: whole blocks inside a pattern");
    // block-uniform base, per-lane 32-bit offsets
    const size_t bbase = (size_t)vb * LPB * N;
    const int lo = ll * N + t;
    // tiled state: pattern and first row of the block; element m at the
    // register base kb(m) plus the lane offset sl
    const size_t pbase = TILED ? ((size_t)vb * LPB / N) * N * N
                                     + (((size_t)vb * LPB) % N) * TL::C
                               : bbase;
    const unsigned sl = TILED ? TL::lane(t, ll) : (unsigned)lo;
    auto kb = [&](auto mc) -> size_t {
        constexpr int m = decltype(mc)::value;
        if constexpr (TILED) return TL::template kpart<PL::rot(m)>();
        else return (size_t)(PL::rot(m) * TT);
    };
    const c2<float>* Zb = Z + pbase;
    c2<float>* Yb = Y + pbase;
    unsigned sw[N / 32];
    load_support_words<N>(sw, sbits + (size_t)(inb ? line : 0) * (N / 32), inb);
    LdsRow<V, N, true> acc{lds + ll * SPAN};
    const V* tw = g_tw<N>();
    V xp[E], xo[E];
    const BufC64 zr(Zb, (unsigned)(N * N * sizeof(V))), yr(Yb, (unsigned)(N * N * sizeof(V)));
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        if constexpr (TILED) {
            constexpr unsigned kbytes = (unsigned)(TL::template kpart<PL::rot(m)>() * sizeof(V));
            xp[m] = inb ? zr.ld(sl * (unsigned)sizeof(V), kbytes) : mk<V>(0, 0);
            xo[m] = inb ? yr.ld(sl * (unsigned)sizeof(V), kbytes) : mk<V>(0, 0);
        } else {
            xp[m] = inb ? (Zb + kb(mc))[sl] : mk<V>(0, 0);
            xo[m] = inb ? (Yb + kb(mc))[sl] : mk<V>(0, 0);
        }
    });
    fft_line<PL, +1>(xp, t, tw, acc);
    fft_line<PL, +1>(xo, t, tw, acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const bool s = support_bit<PL, m>(sw, t);
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
    if constexpr (LT) {
        // the block's rows r0 .. r0 + LPB - 1 of pattern pb staged through LDS
        // as [v][LPB + 1], then each v's LPB values stored as one contiguous
        // run of lt[pb][v][r0 ..]
        constexpr int TPB = HioRowPlan<N, PV>::TPB, LS = LPB + 1;
        static_assert(!SPARSE && !TILED && EXACT && N % LPB == 0 && kRowsLinesFit<N, PV> && (N * LPB) % TPB == 0,
                      "line-major store");
        __syncthreads();  // every line's exchanges done before the LDS is reused
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            lds[(t + PL::rot(m) * TT) * LS + ll] = xo[m];
        });
        __syncthreads();
        const size_t row0 = (size_t)vb * LPB;
        c2<float>* ob = lt + (row0 / N) * N * N + row0 % N;
        sfor<N * LPB / TPB>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int i = tid + k * TPB;
            const int v = i / LPB, ul = i % LPB;
            ob[(size_t)v * N + ul] = lds[v * LS + ul];
        });
    } else if (valid) {
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            if constexpr (TILED) {
                constexpr unsigned kbytes = (unsigned)(TL::template kpart<PL::rot(m)>() * sizeof(V));
                yr.st(xo[m], sl * (unsigned)sizeof(V), kbytes);
            } else {
                (Yb + kb(mc))[sl] = xo[m];
            }
        });
    }
}

// Row FFT pass between the row-major and the tile-major layouts of N x N
// patterns (not in place): out = S_r DFT_r^DIR S_r in, times scale.  Entry
// (x -> tiled state, DIR = -1) and exit (tiled state -> x, DIR = +1, 1/N)
// of the tiled HIO engine.
template <int N, int DIR, bool IN_TILED, bool OUT_TILED>
__global__ __launch_bounds__(SizeCfg<N>::TPB) void k_rows_retile(const c2<float>* __restrict__ in,
                                                                 c2<float>* __restrict__ out,
                                                                 float scale)
{
    using V = c2<float>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = SizeCfg<N>::LPB;
    constexpr int SPAN = LdsRow<V, N>::span();
    static_assert(N % LPB == 0, "whole blocks inside a pattern");
    __shared__ V lds[LPB * SPAN];
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    const size_t line0 = (size_t)blockIdx.x * LPB;
    const size_t pbase = (line0 / N) * N * N;
    const size_t r0 = line0 % N;  // the block's first row
    const int lo = ll * N + t;
    LdsRow<V, N, true> acc{lds + ll * SPAN};
    V v[E];
    const c2<float>* ib = in + pbase + r0 * (IN_TILED ? Tiled<N>::C : N);
    c2<float>* ob = out + pbase + r0 * (OUT_TILED ? Tiled<N>::C : N);
    const unsigned il = IN_TILED ? Tiled<N>::lane(t, ll) : (unsigned)lo;
    const unsigned ol = OUT_TILED ? Tiled<N>::lane(t, ll) : (unsigned)lo;
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        constexpr size_t kb = IN_TILED ? Tiled<N>::template kpart<PL::rot(m)>() : (size_t)(PL::rot(m) * TT);
        v[m] = (ib + kb)[il];
    });
    fft_line<PL, DIR>(v, t, g_tw<N>(), acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        constexpr size_t kb = OUT_TILED ? Tiled<N>::template kpart<PL::rot(m)>() : (size_t)(PL::rot(m) * TT);
        (ob + kb)[ol] = cscale(v[m], scale);
    });
}

// Amplitudes [B][N][N] row-major -> tile-major (Tiled<N>).
template <int N>
__global__ __launch_bounds__(256) void k_amp_retile(const float* __restrict__ in,
                                                    float* __restrict__ out, int64_t total)
{
    constexpr int C = Tiled<N>::C;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;  // tiled index
    if (i >= total) return;
    const int64_t b = i / ((int64_t)N * N);
    const int64_t w = i - b * N * N;
    const int64_t tile = w / ((int64_t)N * C);
    const int64_t rr = (w / C) % N;
    const int64_t cc = w % C;
    out[i] = in[b * N * N + rr * N + tile * C + cc];
}

// The same copy, four columns per thread in row order: block = kRetileRows
// rows of one pattern, read as contiguous float4 rows and written as float4
// runs of a tile row (every tile width C is a multiple of 4).  Needs 16-byte
// aligned amplitudes; the index arithmetic is 32-bit within a pattern, with
// the divisors compile-time constants.
constexpr int kRetileRows = 16;
template <int N>
__global__ __launch_bounds__(256) void k_amp_retile4(const float* __restrict__ in,
                                                     float* __restrict__ out)
{
    constexpr int C = Tiled<N>::C, Q = N / 4, PER = kRetileRows * Q;
    static_assert(C % 4 == 0 && N % kRetileRows == 0, "float4 tile rows");
    const int64_t b = blockIdx.x / (N / kRetileRows);
    const int r0 = (int)(blockIdx.x % (N / kRetileRows)) * kRetileRows;
    const float4* src = reinterpret_cast<const float4*>(in + b * N * N) + r0 * Q;
    float* dst = out + b * N * N;
    for (int q = threadIdx.x; q < PER; q += 256) {
        const int r = r0 + q / Q, c = (q % Q) * 4;
        *reinterpret_cast<float4*>(dst + (c / C) * N * C + r * C + c % C) = src[q];
    }
}

// Shrinkwrap blur, separable form.  The centred Gaussian transfer function
// factors, G(kr, kc) = g_r(kr) g_c(kc) with g(k) = exp(-2 pi^2 sigma^2
// ((k - n/2)/n)^2), so b = Backward2D(G . Forward2D(|x|)) is a blur of every
// row by g_c followed by a blur of every column by g_r.  The row stage keeps
// only the real part (the row blur of a real row is real), so the image
// between the stages and the result are real: 12 + 8 bytes per pixel instead
// of the 16 + 16 + 12 of the 2D spectrum.
//
// Row stage: r1 = Re IDFT_r(g_c . DFT_r(|x|)) / N, with x = IDFT_r(Y)/N
// (FROM_STATE, Y tile-major when TILED) or x read directly; r1 row-major.
// A block takes 2 LPB lines: each thread group transforms line A (first half)
// and line B (second half) to |x| on its own, then blurs them together as
// |x_A| + i |x_B| (g_c is real and even about the centre, so the real and
// imaginary parts come out as the two blurred rows): four transforms per two
// rows instead of six.
template <int N, bool FROM_STATE, bool TILED = false>
__global__ __launch_bounds__(SizeCfg<N>::TPB) void k_sw_rowblur(const c2<float>* __restrict__ Y,
                                                               float* __restrict__ r1, int64_t nlines,
                                                               float invN, float gcoef_c)
{
    using V = c2<float>;
    using PL = typename SizeCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = SizeCfg<N>::LPB;
    constexpr int SPAN = LdsRow<V, N>::span();
    __shared__ V lds[LPB * SPAN];
    // the row factor g_c(k) / N per position, one expf per position and block
    // instead of one per element
    __shared__ float gtab[N];
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    LdsRow<V, N, true> acc{lds + ll * SPAN};
    const V* tw = g_tw<N>();
    const size_t lineA0 = (size_t)blockIdx.x * (2 * LPB), lineB0 = lineA0 + LPB;
    const bool hasB = (int64_t)lineB0 < nlines;  // block-uniform
    auto load = [&](size_t line0, V (&v)[E]) -> bool {
        const int64_t line = (int64_t)line0 + ll;
        const bool valid = line < nlines;
        const int64_t base = (valid ? line : 0) * N;
        const c2<float>* Yp = Y + (line0 / N) * N * N + (line0 % N) * Tiled<N>::C;
        const unsigned yl = TILED ? Tiled<N>::lane(t, ll) : 0u;
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            if constexpr (TILED)
                v[m] = valid ? (Yp + Tiled<N>::template kpart<PL::rot(m)>())[yl] : mk<V>(0, 0);
            else
                v[m] = valid ? Y[base + t + PL::rot(m) * TT] : mk<V>(0, 0);
        });
        return valid;
    };
    auto magnitude = [&](V (&v)[E]) {
        if constexpr (FROM_STATE) fft_line<PL, +1>(v, t, tw, acc);
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const float sc = FROM_STATE ? invN : 1.f;
            v[m] = mk<V>(sqrtf(v[m].x * v[m].x + v[m].y * v[m].y) * sc, 0.f);
        });
    };
    // both halves' loads leave before the factor table and the transforms
    // (41.2 against 43.4 us per half-batch update at 256^2, with the stores
    // below)
    V v[E], w[E];
    const bool okA = load(lineA0, v);
    bool okB = false;
    if (hasB) okB = load(lineB0, w);
    else sfor<E>([&](auto mc) { w[decltype(mc)::value] = mk<V>(0, 0); });
    for (int i = tid; i < N; i += blockDim.x) {
        const float d = (float)(i - N / 2);
        gtab[i] = expf(gcoef_c * d * d) * invN;
    }
    __syncthreads();
    magnitude(v);
    if (hasB) magnitude(w);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = mk<V>(v[m].x, w[m].x);
    });
    fft_line<PL, -1>(v, t, tw, acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = cscale(v[m], gtab[t + PL::rot(m) * TT]);
    });
    fft_line<PL, +1>(v, t, tw, acc);
    const int64_t baseA = (int64_t)(lineA0 + ll) * N, baseB = (int64_t)(lineB0 + ll) * N;
    {
        // the two rows through the line's LDS span, stored as 16-byte runs
        static_assert(N % (4 * TT) == 0 && (SPAN * 8) % 16 == 0, "float4 rows");
        float* fl = reinterpret_cast<float*>(lds + ll * SPAN);
        wave_lds_sync();
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            fl[t + PL::rot(m) * TT] = v[m].x;
            fl[N + t + PL::rot(m) * TT] = v[m].y;
        });
        wave_lds_sync();
        const float4* f4 = reinterpret_cast<const float4*>(fl);
        sfor<N / (4 * TT)>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (okA) reinterpret_cast<float4*>(r1 + baseA)[t + j * TT] = f4[t + j * TT];
            if (okB) reinterpret_cast<float4*>(r1 + baseB)[t + j * TT] = f4[N / 4 + t + j * TT];
        });
    }
}

// Column stage: b = Re IDFT_c(g_r . DFT_c(r1)) / M on tiles of one pattern
// (r1 and b row-major [B][M][N]), and the per-pattern maximum of b as an
// order-preserving key (one atomicMax per block).  PAIR: each line of the
// transform carries two adjacent real columns, a + i b; g_r is real and even
// about the centre, so the blur of a + i b is blur(a) + i blur(b): half the
// transforms, and 8-byte loads (2C columns per block, N % 2C == 0).
template <int M, bool PAIR>
__global__ __launch_bounds__(SizeCfg<M>::TPB) void k_sw_colblur(const float* __restrict__ r1,
                                                               float* __restrict__ bl,
                                                               unsigned* __restrict__ keys, int N,
                                                               float invM, float gcoef_r)
{
    using V = c2<float>;
    using PL = typename SizeCfg<M>::P;
    constexpr int E = PL::E, TT = PL::T, C = SizeCfg<M>::C, TPB = SizeCfg<M>::TPB;
    constexpr int CW = PAIR ? 2 * C : C;  // columns per block
    __shared__ V lds[M * C];
    __shared__ float wmax[TPB / 64];
    // the column factor g_r(k) / M per position (one expf per position)
    __shared__ float gtab[M];
    const int tid = threadIdx.x;
    const int c = tid % C, t = tid / C;
    const unsigned tiles_per_pattern = (unsigned)N / CW;
    const unsigned b = blockIdx.x / tiles_per_pattern;
    const int col0 = (int)(blockIdx.x - b * tiles_per_pattern) * CW;
    const size_t off = (size_t)b * M * N + col0 + (PAIR ? 2 * c : c);
    auto at = [&](int m) { return off + (size_t)(t + PL::rot(m) * TT) * (size_t)N; };
    // the loads leave before the factor table is formed; the row-blurred
    // image is dead after them, so they are non-temporal (29.9 -> 29.2 us per
    // launch at 256 x 256^2; the same hint on the threshold's loads of the
    // blurred image, which the cache still holds, cost 13.8 -> 21.3 us)
    V v[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        if constexpr (PAIR) {
            const float2 p = __builtin_bit_cast(
                float2, __builtin_nontemporal_load(reinterpret_cast<const unsigned long long*>(r1 + at(m))));
            v[m] = mk<V>(p.x, p.y);
        } else {
            v[m] = mk<V>(r1[at(m)], 0.f);
        }
    });
    for (int i = tid; i < M; i += TPB) {
        const float d = (float)(i - M / 2);
        gtab[i] = expf(gcoef_r * d * d) * invM;
    }
    __syncthreads();
    LdsCol<V, C> acc{lds + c};
    fft_line<PL, -1>(v, t, g_tw<M>(), acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = cscale(v[m], gtab[t + PL::rot(m) * TT]);
    });
    fft_line<PL, +1>(v, t, g_tw<M>(), acc);
    float mx = -INFINITY;
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        mx = fmaxf(mx, v[m].x);
        if constexpr (PAIR) {
            // the two columns as one 8-byte store
            *reinterpret_cast<float2*>(bl + at(m)) = make_float2(v[m].x, v[m].y);
            mx = fmaxf(mx, v[m].y);
        } else {
            bl[at(m)] = v[m].x;
        }
    });
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((tid & 63) == 0) wmax[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) {
        float bm = wmax[0];
        for (int w = 1; w < TPB / 64; ++w) bm = fmaxf(bm, wmax[w]);
        atomicMax(&keys[b], f2key(bm));
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
__global__ __launch_bounds__(DmRowCfg<N>::TPB) void k_dm_entry(
    const c2<float>* __restrict__ psi, const c2<float>* __restrict__ t1,
    const c2<float>* __restrict__ t2, const uint8_t* __restrict__ mask,
    c2<float>* __restrict__ Y1, c2<float>* __restrict__ Y2, int64_t nlines, float beta)
{
    using V = c2<float>;
    using PL = typename DmRowCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = DmRowCfg<N>::LPB;
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
// The body handles the LPB lines of block `vb` (DmRowCfg<N>) with the
// caller's LDS (LPB * LdsRow<V, N>::span() elements); k_dm_rows and the
// mixed-size kernel run it.
template <int N, bool TEXTBOOK, int SPACE, bool LAST>
__device__ __forceinline__ void dm_rows_block(
    const unsigned vb, c2<float>* lds, const c2<float>* __restrict__ Z1, const c2<float>* __restrict__ Z2,
    c2<float>* __restrict__ psi, const uint8_t* __restrict__ mask, c2<float>* __restrict__ Y1,
    c2<float>* __restrict__ Y2, c2<float>* __restrict__ t1out, c2<float>* __restrict__ t2out,
    int64_t nlines, float beta)
{
    using V = c2<float>;
    using PL = typename DmRowCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = DmRowCfg<N>::LPB;
    constexpr int SPAN = LdsRow<V, N>::span();
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    const int64_t line = (int64_t)vb * LPB + ll;
    const bool valid = line < nlines;
    const size_t base = (size_t)(valid ? line : 0) * N + t;
    LdsRow<V, N, true> acc{lds + ll * SPAN};
    const V* tw = g_tw<N>();
    const float ib = 1.f / beta, pb = 1.f + 1.f / beta, cb = 1.f - 1.f / beta;
    if constexpr (TEXTBOOK || N > 256) {
        // Textbook map: Y1 and Y2 both come from psi', so Z1 and Z2 are
        // transformed together and both forward transforms follow the update.
        // Lines longer than 256 take this order too: for them the compiler
        // allocates fewer registers to it than to the two-phase order below.
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
    } else {
        // Reference variant: two row arrays live at a time.  t1' is finished and
        // transformed into Y1 before Z2 is loaded; psi + beta t1' is carried to
        // the second half.  Explicit fmaf throughout: the same operations
        // whatever kernel inlines this body.
        static_assert(E <= 32, "support bits in one word");
        V a[E], acc1[E];
        unsigned inbits = 0u;  // bit m: element m lies in the support
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            a[m] = valid ? Z1[base + PL::rot(m) * TT] : mk<V>(0, 0);
        });
        fft_line<PL, +1>(a, t, tw, acc);
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const size_t o = base + PL::rot(m) * TT;
            const bool inm = valid && mask[o] != 0;
            inbits |= (inm ? 1u : 0u) << m;
            const V p = valid ? psi[o] : mk<V>(0, 0);
            const V w = mk<V>(fmaf(pb, a[m].x, -ib * p.x), fmaf(pb, a[m].y, -ib * p.y));
            const V t1n = ps_apply<SPACE>(w, inm);
            acc1[m] = mk<V>(fmaf(beta, t1n.x, p.x), fmaf(beta, t1n.y, p.y));  // psi + beta t1'
            a[m] = t1n;
        });
        if constexpr (LAST) {
            if (valid && t1out) {
                sfor<E>([&](auto mc) {
                    constexpr int m = decltype(mc)::value;
                    t1out[base + PL::rot(m) * TT] = a[m];
                });
            }
        } else {
            // next Y1 = DFT_r(t1')
            fft_line<PL, -1>(a, t, tw, acc);
            if (valid) {
                sfor<E>([&](auto mc) {
                    constexpr int m = decltype(mc)::value;
                    Y1[base + PL::rot(m) * TT] = a[m];
                });
            }
        }
        // second half: t2' = IDFT_r(Z2), psi' = (psi + beta t1') - beta t2'.
        // The barrier keeps the scheduler from hoisting the Z2 loads above the
        // first half (which would make three row arrays live at once).
        __builtin_amdgcn_sched_barrier(0);
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            a[m] = valid ? Z2[base + PL::rot(m) * TT] : mk<V>(0, 0);
        });
        fft_line<PL, +1>(a, t, tw, acc);
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const size_t o = base + PL::rot(m) * TT;
            const V t2n = a[m];
            const V pn = mk<V>(fmaf(-beta, t2n.x, acc1[m].x), fmaf(-beta, t2n.y, acc1[m].y));
            if (valid) psi[o] = pn;
            if constexpr (LAST) {
                if (valid && t2out) t2out[o] = t2n;
            } else {
                const V q = ps_apply<SPACE>(t2n, ((inbits >> m) & 1u) != 0u);
                a[m] = mk<V>(fmaf(cb, q.x, pn.x * ib), fmaf(cb, q.y, pn.y * ib));
            }
        });
        if constexpr (!LAST) {
            // next Y2 = DFT_r(psi'/beta + (1 - 1/beta) P_S t2')
            fft_line<PL, -1>(a, t, tw, acc);
            if (valid) {
                sfor<E>([&](auto mc) {
                    constexpr int m = decltype(mc)::value;
                    Y2[base + PL::rot(m) * TT] = a[m];
                });
            }
        }
    }
}

template <int N, bool TEXTBOOK, int SPACE, bool LAST>
__global__ __launch_bounds__(DmRowCfg<N>::TPB) void k_dm_rows(
    const c2<float>* __restrict__ Z1, const c2<float>* __restrict__ Z2,
    c2<float>* __restrict__ psi, const uint8_t* __restrict__ mask, c2<float>* __restrict__ Y1,
    c2<float>* __restrict__ Y2, c2<float>* __restrict__ t1out, c2<float>* __restrict__ t2out,
    int64_t nlines, float beta)
{
    __shared__ c2<float> lds[DmRowCfg<N>::LPB * LdsRow<c2<float>, N>::span()];
    dm_rows_block<N, TEXTBOOK, SPACE, LAST>(blockIdx.x, lds, Z1, Z2, psi, mask, Y1, Y2, t1out, t2out,
                                            nlines, beta);
}

}  // namespace cpra

namespace cpra {

// ---------------------------------------------------------------------------
// Elser's textbook difference map on 2D patterns in two passes of 36 and 33
// bytes per pixel (against 40 + 49 for the body above, which also serves
// the reference variant and the 3D map).
//
// With Y1 = DFT_r(psi) and Y2 = DFT_r(f_S psi), f_S psi = psi/beta +
// (1 - 1/beta) P_S psi, the column pass turns Y into Z with IDFT_r(Z) = P_M(.)
// in pattern space (Z1 for psi, Z2 for f_S psi).  Everything else the update
// needs from psi enters linearly, and IDFT_r(Y1) = N psi, so
//   IDFT_r(V) = (1 + 1/beta) P_M psi - psi/beta,   V = (1 + 1/beta) Z1 - Y1 / (beta N)
//   IDFT_r(W) = psi - beta P_M(f_S psi),           W = Y1 / N - beta Z2
//   psi'      = IDFT_r(W) + beta P_S(IDFT_r(V)).
// A column block holds Y1, Z1 and Z2 of the same columns, so it writes V and
// W instead of Z1 and Z2 (one read of the amplitudes for both), and the row
// pass reads V, W and the support only: psi is neither read nor written
// until the last iteration.  The values equal those of dm_rows_block's
// textbook map up to FP32 rounding (the linear combinations are formed in
// the row-spectral domain).
// ---------------------------------------------------------------------------

// Column pass: tile `tile` of the B-pattern batch (DmtColCfg<M>: C columns,
// 512 threads); Y2 = Y1 + half, W = V + half.  lds: M * C elements.
template <int M>
__device__ __forceinline__ void dmt_cols_tile(
    const unsigned tile, c2<float>* lds, const c2<float>* __restrict__ Y1, c2<float>* __restrict__ Vout,
    const float* __restrict__ A, int64_t amp_batch, int N, float scale, int skip_zero, float beta,
    size_t half)
{
    using V = c2<float>;
    using PL = typename DmtColCfg<M>::P;
    constexpr int E = PL::E, TT = PL::T, C = DmtColCfg<M>::C;
    static_assert(C * TT == DmtColCfg<M>::TPB, "column tile");
    const int tid = threadIdx.x;
    const int c = tid % C, t = tid / C;
    const unsigned tiles_per_pattern = (unsigned)N / C;
    const unsigned b = tile / tiles_per_pattern;
    const int col0 = (int)(tile - b * tiles_per_pattern) * C;
    const size_t pat = (size_t)M * (size_t)N;
    const size_t blk0 = b * pat + col0;
    // element m of this lane: row t + rot(m)*T, a wave-uniform row base plus
    // one per-lane offset shared by all m
    const unsigned toff = (unsigned)t * (unsigned)N + (unsigned)c;
    auto rowbase = [&](int m) { return (size_t)(PL::rot(m) * TT) * (size_t)N; };
    const float* Ap = A + (size_t)(b % (unsigned)amp_batch) * pat + col0;
    const float invN = 1.f / (float)N;
    const float ib = 1.f / beta, pb = 1.f + 1.f / beta;
    V v[E], y[E];
    float av[E];
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        y[m] = (Y1 + blk0 + rowbase(m))[toff];
        v[m] = y[m];
    });
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        av[m] = (Ap + rowbase(m))[toff];
    });
    LdsCol<V, C, false> acc{lds + c};
    // DFT_c, modulus projection with A (scale 1/(MN)), IDFT_c
    auto project = [&]() {
        fft_line<PL, -1>(v, t, g_tw<M>(), acc);
        sfor<E>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const float a = av[m];
            const V x = v[m];
            const float m2 = x.x * x.x + x.y * x.y;
            float s = scale;
            if (!(skip_zero && a == 0.f) && m2 > 0.f) s = scale * a * rsqrtf(m2);
            v[m] = cscale(x, s);
        });
        fft_line<PL, +1>(v, t, g_tw<M>(), acc);
    };
    project();  // v = Z1
    const float yv = -ib * invN;
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        (Vout + blk0 + rowbase(m))[toff] = mk<V>(fmaf(pb, v[m].x, yv * y[m].x), fmaf(pb, v[m].y, yv * y[m].y));
    });
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = (Y1 + half + blk0 + rowbase(m))[toff];
    });
    project();  // v = Z2
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        (Vout + half + blk0 + rowbase(m))[toff] =
            mk<V>(fmaf(-beta, v[m].x, invN * y[m].x), fmaf(-beta, v[m].y, invN * y[m].y));
    });
}

template <int M>
__global__ __launch_bounds__(DmtColCfg<M>::TPB) void k_dmt_cols(
    const c2<float>* __restrict__ Y, c2<float>* __restrict__ VW, const float* __restrict__ A,
    int64_t amp_batch, int N, float scale, int skip_zero, float beta, int64_t half)
{
    extern __shared__ unsigned char smem[];
    dmt_cols_tile<M>(blockIdx.x, reinterpret_cast<c2<float>*>(smem), Y, VW, A, amp_batch, N, scale,
                     skip_zero, beta, (size_t)half);
}

// Row pass: the LPB lines of block `vb` (DmtRowCfg<N>); W = V + half, Y2 = Y1 + half.
//   t1'  = P_S(IDFT_r(V))
//   psi' = IDFT_r(W) + beta t1'
// then Y1 = DFT_r(psi'), Y2 = DFT_r(psi'/beta + (1 - 1/beta) P_S psi') for
// the next iteration, or (LAST) psi' in pattern space.
template <int N, int SPACE, bool LAST>
__device__ __forceinline__ void dmt_rows_block(
    const unsigned vb, c2<float>* lds, const c2<float>* __restrict__ VW, c2<float>* __restrict__ psi,
    const uint8_t* __restrict__ mask, c2<float>* __restrict__ Y, int64_t nlines, float beta, size_t half)
{
    using V = c2<float>;
    using PL = typename DmtRowCfg<N>::P;
    constexpr int E = PL::E, TT = PL::T, LPB = DmtRowCfg<N>::LPB;
    constexpr int SPAN = LdsRow<V, N>::span();
    static_assert(E <= 32, "support bits in one word");
    const int tid = threadIdx.x;
    const int t = tid % TT, ll = tid / TT;
    const int64_t line = (int64_t)vb * LPB + ll;
    const bool valid = line < nlines;
    const size_t base = (size_t)(valid ? line : 0) * N + t;
    LdsRow<V, N, true> acc{lds + ll * SPAN};
    const V* tw = g_tw<N>();
    const float ib = 1.f / beta, cb = 1.f - 1.f / beta;
    V a[E], b[E];
    unsigned inbits = 0u;  // bit m: element m lies in the support
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        a[m] = valid ? VW[base + PL::rot(m) * TT] : mk<V>(0, 0);
    });
    fft_line<PL, +1>(a, t, tw, acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const bool inm = valid && mask[base + PL::rot(m) * TT] != 0;
        inbits |= (inm ? 1u : 0u) << m;
        a[m] = ps_apply<SPACE>(a[m], inm);  // t1'
    });
    // keep the scheduler from hoisting the W loads above the first transform
    __builtin_amdgcn_sched_barrier(0);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        b[m] = valid ? VW[half + base + PL::rot(m) * TT] : mk<V>(0, 0);
    });
    fft_line<PL, +1>(b, t, tw, acc);
    sfor<E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const V pn = mk<V>(fmaf(beta, a[m].x, b[m].x), fmaf(beta, a[m].y, b[m].y));  // psi'
        if constexpr (LAST) {
            if (valid) psi[base + PL::rot(m) * TT] = pn;
        } else {
            const V q = ps_apply<SPACE>(pn, ((inbits >> m) & 1u) != 0u);
            a[m] = pn;
            b[m] = mk<V>(fmaf(cb, q.x, pn.x * ib), fmaf(cb, q.y, pn.y * ib));
        }
    });
    if constexpr (!LAST) {
        fft_line<PL, -1>(a, t, tw, acc);
        fft_line<PL, -1>(b, t, tw, acc);
        if (valid) {
            sfor<E>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                Y[base + PL::rot(m) * TT] = a[m];
                Y[half + base + PL::rot(m) * TT] = b[m];
            });
        }
    }
}

template <int N, int SPACE, bool LAST>
__global__ __launch_bounds__(DmtRowCfg<N>::TPB) void k_dmt_rows(
    const c2<float>* __restrict__ VW, c2<float>* __restrict__ psi, const uint8_t* __restrict__ mask,
    c2<float>* __restrict__ Y, int64_t nlines, float beta, int64_t half)
{
    __shared__ c2<float> lds[DmtRowCfg<N>::LPB * LdsRow<c2<float>, N>::span()];
    dmt_rows_block<N, SPACE, LAST>(blockIdx.x, lds, VW, psi, mask, Y, nlines, beta, (size_t)half);
}

}  // namespace cpra
