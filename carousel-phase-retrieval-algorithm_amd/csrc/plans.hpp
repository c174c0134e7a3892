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

}  // namespace cpra
