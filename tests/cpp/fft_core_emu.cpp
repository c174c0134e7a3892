// Host emulation of the register/LDS Stockham FFT in csrc/fft_core.hpp.
// The T threads of a line are run one after another between the stage
// boundaries where the GPU code synchronises, so the plan arithmetic
// (radix sequence, twiddle indices, Stockham output positions, centring
// rotation) is checked on the CPU against a direct long-double DFT.
// Prints one line per plan: "<N> <E> <dir> <dtype> <max_rel_err>".
#include "../../carousel-phase-retrieval-algorithm_amd/csrc/fft_core.hpp"

#include <cmath>
#include <cstdio>
#include <vector>

using namespace cpra;

template <typename V> struct HostAcc {
    V* base;
    void st(int i, V x) { base[i] = x; }
    V ld(int i) const { return base[i]; }
};

template <class PL, int DIR, typename T> double run_plan(unsigned seed)
{
    using V = c2<T>;
    constexpr int N = PL::N, E = PL::E, TT = PL::T;
    std::vector<V> tw(N);
    for (int i = 0; i < N; ++i) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * i / N;
        tw[i] = mk<V>((T)std::cos(a), (T)std::sin(a));
    }
    std::vector<V> x(N);
    unsigned s = seed;
    for (int i = 0; i < N; ++i) {
        s = s * 1664525u + 1013904223u;
        const T re = (T)((s >> 8) / 16777216.0 - 0.5);
        s = s * 1664525u + 1013904223u;
        const T im = (T)((s >> 8) / 16777216.0 - 0.5);
        x[i] = mk<V>(re, im);
    }
    std::vector<std::vector<V>> v(TT, std::vector<V>(E));
    for (int t = 0; t < TT; ++t)
        for (int m = 0; m < E; ++m) v[t][m] = x[t + m * TT];
    std::vector<V> lds(N);
    HostAcc<V> acc{lds.data()};
    sfor<PL::NSTAGE>([&](auto sc) {
        constexpr int S = decltype(sc)::value;
        for (int t = 0; t < TT; ++t) {
            V(&r)[E] = *reinterpret_cast<V(*)[E]>(v[t].data());
            stage_compute<PL, S, DIR>(r, t, tw.data());
        }
        if constexpr (S + 1 < PL::NSTAGE) {
            for (int t = 0; t < TT; ++t) {
                V(&r)[E] = *reinterpret_cast<V(*)[E]>(v[t].data());
                stage_scatter<PL, S>(r, t, acc);
            }
            for (int t = 0; t < TT; ++t) {
                V(&r)[E] = *reinterpret_cast<V(*)[E]>(v[t].data());
                gather_natural<PL>(r, t, acc);
            }
        }
    });
    double maxerr = 0.0, maxref = 0.0;
    for (int k = 0; k < N; ++k) {
        long double sr = 0, si = 0;
        for (int n = 0; n < N; ++n) {
            const long double a =
                DIR * 2.0L * 3.14159265358979323846264338327950288L * (long double)((long)n * k % N) / N;
            sr += x[n].x * std::cos(a) - x[n].y * std::sin(a);
            si += x[n].x * std::sin(a) + x[n].y * std::cos(a);
        }
        const int t = k % TT, m = k / TT;
        const double dr = (double)(v[t][m].x - sr), di = (double)(v[t][m].y - si);
        maxerr = std::fmax(maxerr, std::sqrt(dr * dr + di * di));
        maxref = std::fmax(maxref, (double)std::sqrt(sr * sr + si * si));
    }
    return maxerr / maxref;
}

template <class PL> void both(const char* name)
{
    std::printf("%d %d -1 f32 %.3e %s\n", PL::N, PL::E, run_plan<PL, -1, float>(1u), name);
    std::printf("%d %d +1 f32 %.3e %s\n", PL::N, PL::E, run_plan<PL, +1, float>(2u), name);
    std::printf("%d %d -1 f64 %.3e %s\n", PL::N, PL::E, run_plan<PL, -1, double>(3u), name);
    std::printf("%d %d +1 f64 %.3e %s\n", PL::N, PL::E, run_plan<PL, +1, double>(4u), name);
}

int main()
{
    both<Plan<8, 8, 8>>("Plan<8, 8, 8>");
    both<Plan<10, 10, 10>>("Plan<10, 10, 10>");
    both<Plan<12, 12, 12>>("Plan<12, 12, 12>");
    both<Plan<16, 16, 16>>("Plan<16, 16, 16>");
    both<Plan<30, 30, 30>>("Plan<30, 30, 30>");
    both<Plan<32, 16, 16, 2>>("Plan<32, 16, 16, 2>");
    both<Plan<64, 16, 16, 4>>("Plan<64, 16, 16, 4>");
    both<Plan<64, 8, 8, 8>>("Plan<64, 8, 8, 8>");
    both<Plan<128, 16, 16, 8>>("Plan<128, 16, 16, 8>");
    both<Plan<128, 8, 8, 8, 2>>("Plan<128, 8, 8, 8, 2>");
    both<Plan<192, 24, 24, 8>>("Plan<192, 24, 24, 8>");
    both<Plan<192, 12, 12, 4, 4>>("Plan<192, 12, 12, 4, 4>");
    both<Plan<256, 16, 16, 16>>("Plan<256, 16, 16, 16>");
    both<Plan<256, 8, 8, 8, 4>>("Plan<256, 8, 8, 8, 4>");
    both<Plan<384, 24, 6, 8, 8>>("Plan<384, 24, 6, 8, 8>");
    both<Plan<384, 12, 6, 4, 4, 4>>("Plan<384, 12, 6, 4, 4, 4>");
    both<Plan<384, 12, 4, 4, 4, 6>>("Plan<384, 12, 4, 4, 4, 6>");
    both<Plan<512, 8, 8, 8, 8>>("Plan<512, 8, 8, 8, 8>");
    both<Plan<512, 16, 16, 16, 2>>("Plan<512, 16, 16, 16, 2>");
    both<Plan<512, 32, 32, 16>>("Plan<512, 32, 32, 16>");
    return 0;
}
