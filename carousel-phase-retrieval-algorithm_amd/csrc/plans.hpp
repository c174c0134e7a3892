// plans.hpp -- compile-time FFT plans and tile shapes for the fast sizes.
//
// Each size has: the radix plan (fft_core.hpp), LPB = lines per block of the
// contiguous (row) kernels and C = lines per tile of the strided (column /
// pencil) kernels, both chosen so a block has TPB = 256 threads (4 wave64s).
// Radix 16 x 16 for 256 means one LDS exchange per transform; 192 and 384 use
// radix-3 factors inside the 24-point register DFTs.
#pragma once
#include "fft_core.hpp"

namespace cpra {

// Twiddle tables tw[k] = exp(-2 pi i k / N), evaluated at compile time in
// double (ct_cs) and rounded once to T; one device-global table per (N, T).
template <int N, typename T> struct TwTab {
    c2<T> v[N];
};
template <int N, typename T> constexpr TwTab<N, T> make_twtab()
{
    TwTab<N, T> t{};
    for (int k = 0; k < N; ++k) {
        const CS cs = ct_cs(k, N);
        t.v[k].x = (T)cs.c;
        t.v[k].y = (T)(-cs.s);
    }
    return t;
}
template <int N, typename T> __device__ const TwTab<N, T> kTw = make_twtab<N, T>();
template <int N, typename T = float> __device__ __forceinline__ const c2<T>* g_tw()
{
    return kTw<N, T>.v;
}

template <int N> struct SizeCfg;

template <> struct SizeCfg<64> {
    using P = Plan<64, 16, 16, 4>;
    static constexpr int TPB = 256, LPB = TPB / P::T, C = TPB / P::T;
};
template <> struct SizeCfg<128> {
    using P = Plan<128, 16, 16, 8>;
    static constexpr int TPB = 256, LPB = TPB / P::T, C = TPB / P::T;
};
template <> struct SizeCfg<192> {
    using P = Plan<192, 24, 24, 8>;
    static constexpr int TPB = 256, LPB = TPB / P::T, C = TPB / P::T;
};
template <> struct SizeCfg<256> {
    using P = Plan<256, 16, 16, 16>;
    static constexpr int TPB = 256, LPB = TPB / P::T, C = TPB / P::T;
};
template <> struct SizeCfg<384> {
    using P = Plan<384, 24, 6, 8, 8>;
    static constexpr int TPB = 256, LPB = TPB / P::T, C = TPB / P::T;
};
template <> struct SizeCfg<512> {
    using P = Plan<512, 16, 16, 16, 2>;
    static constexpr int TPB = 256, LPB = TPB / P::T, C = TPB / P::T;
};

// Plans of the difference-map row kernels (k_dm_entry and dm_rows_block).
// They hold two to four row arrays at a time, so where the plan above keeps
// 24 (192, 384) or 16 (256, 512) values per thread they take 12 or 8 instead:
// fewer registers per array, one more LDS exchange per transform.  The other
// row and column kernels keep the plans above (the column kernels' tile width
// C = 256 / T would narrow with a larger T).
template <int N> struct DmRowCfg {
    using P = typename SizeCfg<N>::P;
    static constexpr int TPB = 256, LPB = TPB / P::T;
};
template <> struct DmRowCfg<192> {
    using P = Plan<192, 12, 12, 4, 4>;
    static constexpr int TPB = 256, LPB = TPB / P::T;
};
template <> struct DmRowCfg<256> {
    using P = Plan<256, 8, 8, 8, 4>;
    static constexpr int TPB = 256, LPB = TPB / P::T;
};
template <> struct DmRowCfg<384> {
    using P = Plan<384, 12, 6, 4, 4, 4>;
    static constexpr int TPB = 256, LPB = TPB / P::T;
};
template <> struct DmRowCfg<512> {
    using P = Plan<512, 8, 8, 8, 8>;
    static constexpr int TPB = 256, LPB = TPB / P::T;
};

// Plan of the HIO row kernel (k_eng_rows_hio).  The 256-point rows keep the
// one-exchange 16 x 16 plan: 8 values per thread (Plan<256, 8, 8, 8, 4>)
// lowers the kernel from 125 to 96 VGPRs, but its two extra LDS exchanges
// cost what the fifth wave per SIMD gains (149.4 against 148.7 µs per
// 512-pattern pass).
template <int N> struct HioRowCfg {
    using P = typename SizeCfg<N>::P;
    static constexpr int TPB = 256, LPB = TPB / P::T;
};

// Radix-plan variants of the HIO row kernel that the autotuner times against
// each other (tuning key row_plan): variant 0 is HioRowCfg; variant 1, where
// a size has one, keeps 8 values per thread instead of 16 (fewer VGPRs, one
// more wave per SIMD, two more LDS exchanges per transform), the DM row
// kernels' plans.  Which one wins depends on the batch and the support, so
// it is measured rather than fixed.
template <int N, int PV> struct HioRowPlan : HioRowCfg<N> {};
template <int N> struct HasAltRowPlan {
    static constexpr bool value = false;
};
template <> struct HioRowPlan<256, 1> {
    using P = Plan<256, 8, 8, 8, 4>;
    static constexpr int TPB = 256, LPB = TPB / P::T;
};
template <> struct HasAltRowPlan<256> {
    static constexpr bool value = true;
};
template <> struct HioRowPlan<512, 1> {
    using P = Plan<512, 8, 8, 8, 8>;
    static constexpr int TPB = 256, LPB = TPB / P::T;
};
template <> struct HasAltRowPlan<512> {
    static constexpr bool value = true;
};

// Column tiles of the two-pass textbook difference map (dmt_cols_tile): one
// 512-thread block takes C = 512 / T columns of a pattern for both halves of
// [Y1; Y2], so it holds the incoming Y1, the transform in flight and the
// amplitudes at once.  The 192- and 384-point columns take the 12-value
// plans of the DM row kernels (the 24-value plans leave one or two waves per
// SIMD with that much state), the 64-point columns 8 values (16 would make
// the tile wider than the pattern).  LDS: M * C complex64, 32 or 48 KiB.
template <int N> struct DmtColCfg {
    using P = typename SizeCfg<N>::P;
    static constexpr int TPB = 512, C = TPB / P::T;
};
template <> struct DmtColCfg<64> {
    using P = Plan<64, 8, 8, 8>;
    static constexpr int TPB = 512, C = TPB / P::T;
};
template <> struct DmtColCfg<128> {
    using P = Plan<128, 8, 8, 8, 2>;
    static constexpr int TPB = 512, C = TPB / P::T;
};
template <> struct DmtColCfg<256> {
    using P = Plan<256, 8, 8, 8, 4>;
    static constexpr int TPB = 512, C = TPB / P::T;
};
template <> struct DmtColCfg<512> {
    using P = Plan<512, 8, 8, 8, 8>;
    static constexpr int TPB = 512, C = TPB / P::T;
};
template <> struct DmtColCfg<192> {
    using P = Plan<192, 12, 12, 4, 4>;
    static constexpr int TPB = 512, C = TPB / P::T;
};
template <> struct DmtColCfg<384> {
    using P = Plan<384, 12, 4, 4, 4, 6>;
    static constexpr int TPB = 512, C = TPB / P::T;
};

// Row plans of the two-pass textbook difference map (dmt_rows_block): the DM
// row plans, except at 384, where the radix order 4, 4, 4, 6 needs fewer
// registers than 6, 4, 4, 4.
template <int N> struct DmtRowCfg : DmRowCfg<N> {};
template <> struct DmtRowCfg<384> {
    using P = Plan<384, 12, 4, 4, 4, 6>;
    static constexpr int TPB = 256, LPB = TPB / P::T;
};

}  // namespace cpra
