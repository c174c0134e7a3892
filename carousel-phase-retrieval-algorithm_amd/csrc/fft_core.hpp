// fft_core.hpp -- register/LDS Stockham FFT building blocks for gfx950.
//
// A line of N complex values is owned by T = N/E threads; thread t keeps the
// E values whose element index is t + m*T (m = 0..E-1) in registers v[m].
// A plan is a list of radices R_0..R_{s-1} (product N, each dividing E).
// Stage i performs E/R_i radix-R_i butterflies per thread in registers
// (Stockham autosort, Govindaraju et al., SC'08), and the stages exchange
// data through LDS.  Input and output are both in the natural register
// layout above, so loads and stores of a line are coalesced across the
// T threads and consecutive transforms on the same line chain in registers.
//
// The centring shift S (x[n] -> x[(n + N/2) mod N], include/detail/
// MKL_Impl.inl:180-209 applied per axis) is folded into the index map: the
// physical position of logical element t + m*T is t + ((m + E/2) mod E)*T,
// i.e. a compile-time renaming of the registers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#ifndef CPRA_HD
#define CPRA_HD __host__ __device__ __forceinline__
#endif

namespace cpra {

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
// Interleaved complex value with the layout of std::complex<T>; the
// alignment lets loads and stores become single 8/16-byte accesses.
template <typename T> struct alignas(2 * sizeof(T)) c2 {
    T x, y;
};
template <typename T> using cx = c2<T>;
template <typename V> struct Scalar;
template <typename T> struct Scalar<c2<T>> { using type = T; };

template <typename V> CPRA_HD V mk(typename Scalar<V>::type a, typename Scalar<V>::type b)
{
    V r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename V> CPRA_HD V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <typename V> CPRA_HD V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }
template <typename V> CPRA_HD V cmul(V a, V b)
{
    return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V, typename S> CPRA_HD V cscale(V a, S s)
{
    return mk<V>(a.x * s, a.y * s);
}
// multiply by DIR*i  (DIR = -1: -i, DIR = +1: +i)
template <int DIR, typename V> CPRA_HD V mul_i(V a)
{
    if constexpr (DIR > 0) return mk<V>(-a.y, a.x);
    else return mk<V>(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// compile-time loops and trigonometry
// ---------------------------------------------------------------------------
template <int... Is, typename F>
CPRA_HD void sfor_impl(std::integer_sequence<int, Is...>, F&& f)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F> CPRA_HD void sfor(F&& f)
{
    sfor_impl(std::make_integer_sequence<int, N>{}, f);
}

constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ct_sin_series(double x)
{
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 16; ++i) {
        term *= -x2 / (double)((2 * i) * (2 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double ct_cos_series(double x)
{
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 16; ++i) {
        term *= -x2 / (double)((2 * i - 1) * (2 * i));
        sum += term;
    }
    return sum;
}
struct CS {
    double c, s;
};
// cos and sin of 2*pi*p/q, reduced to the first octant before the series so
// that quadrant points are exact.
constexpr CS ct_cs(long p, long q)
{
    p %= q;
    if (p < 0) p += q;
    const long p4 = 4 * p;
    const long quad = p4 / q;
    const long rem = p4 - quad * q;  // angle = (pi/2) * (quad + rem/q)
    double c = 1.0, s = 0.0;
    if (2 * rem <= q) {
        const double a = (kPi / 2) * (double)rem / (double)q;
        c = ct_cos_series(a);
        s = ct_sin_series(a);
    } else {
        const double b = (kPi / 2) * (double)(q - rem) / (double)q;
        c = ct_sin_series(b);
        s = ct_cos_series(b);
    }
    switch (quad & 3) {
    case 0: return CS{c, s};
    case 1: return CS{-s, c};
    case 2: return CS{-c, -s};
    default: return CS{s, -c};
    }
}

// a * exp(DIR * 2*pi*i * E / R) with a compile-time exponent
template <int E, int R, int DIR, typename V> CPRA_HD V twc(V a)
{
    constexpr int e = ((E % R) + R) % R;
    using T = typename Scalar<V>::type;
    if constexpr (e == 0) {
        return a;
    } else if constexpr (4 * e == R) {
        return mul_i<DIR>(a);
    } else if constexpr (2 * e == R) {
        return mk<V>(-a.x, -a.y);
    } else if constexpr (4 * e == 3 * R) {
        return mul_i<-DIR>(a);
    } else if constexpr (8 * e == R || 8 * e == 3 * R || 8 * e == 5 * R || 8 * e == 7 * R) {
        // (+-1 +- i)/sqrt(2): two adds and two multiplies
        constexpr CS w = ct_cs(e, R);
        constexpr T h = (T)0.70710678118654752440084436210484903;
        constexpr T sc = (T)(w.c > 0 ? 1 : -1), ss = (T)((w.s > 0 ? 1 : -1) * DIR);
        // (a.x + i a.y)(sc + i ss)/sqrt2
        return mk<V>((sc * a.x - ss * a.y) * h, (ss * a.x + sc * a.y) * h);
    } else {
        constexpr CS w = ct_cs(e, R);
        constexpr T c = (T)w.c;
        constexpr T s = (T)(DIR * w.s);
        return mk<V>(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// ---------------------------------------------------------------------------
// small in-register DFTs (natural order in and out), sign DIR
// ---------------------------------------------------------------------------
constexpr int dft_split(int R)
{
    return (R % 4 == 0 && R > 4) ? 4 : (R % 2 == 0 ? 2 : (R % 3 == 0 ? 3 : (R % 5 == 0 ? 5 : R)));
}

template <int R, int DIR> struct Dft {
    static_assert(dft_split(R) != R, "prime radix without a direct kernel");
    template <typename V> CPRA_HD static void run(V* v)
    {
        constexpr int P = dft_split(R);
        constexpr int Q = R / P;
        V a[R];
        sfor<Q>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            V t[P];
            sfor<P>([&](auto n1c) {
                constexpr int n1 = decltype(n1c)::value;
                t[n1] = v[Q * n1 + n2];
            });
            Dft<P, DIR>::run(t);
            sfor<P>([&](auto k1c) {
                constexpr int k1 = decltype(k1c)::value;
                a[n2 * P + k1] = twc<n2 * k1, R, DIR>(t[k1]);
            });
        });
        sfor<P>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            V t[Q];
            sfor<Q>([&](auto n2c) {
                constexpr int n2 = decltype(n2c)::value;
                t[n2] = a[n2 * P + k1];
            });
            Dft<Q, DIR>::run(t);
            sfor<Q>([&](auto k2c) {
                constexpr int k2 = decltype(k2c)::value;
                v[k1 + P * k2] = t[k2];
            });
        });
    }
};

template <int DIR> struct Dft<1, DIR> {
    template <typename V> CPRA_HD static void run(V*) {}
};

template <int DIR> struct Dft<2, DIR> {
    template <typename V> CPRA_HD static void run(V* v)
    {
        const V a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int DIR> struct Dft<3, DIR> {
    template <typename V> CPRA_HD static void run(V* v)
    {
        using T = typename Scalar<V>::type;
        constexpr T h = (T)0.5;
        constexpr T s = (T)0.86602540378443864676372317075293618;  // sin(2pi/3)
        const V t1 = cadd(v[1], v[2]);
        const V t2 = mk<V>(v[0].x - h * t1.x, v[0].y - h * t1.y);
        const V d = csub(v[1], v[2]);
        const V t3 = mul_i<DIR>(mk<V>(s * d.x, s * d.y));
        v[0] = cadd(v[0], t1);
        v[1] = cadd(t2, t3);
        v[2] = csub(t2, t3);
    }
};

template <int DIR> struct Dft<4, DIR> {
    template <typename V> CPRA_HD static void run(V* v)
    {
        const V a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        const V c = cadd(v[1], v[3]), d = mul_i<DIR>(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, d);
        v[3] = csub(b, d);
    }
};

template <int DIR> struct Dft<5, DIR> {
    template <typename V> CPRA_HD static void run(V* v)
    {
        using T = typename Scalar<V>::type;
        constexpr T c1 = (T)0.30901699437494742410229341718281906;   // cos(2pi/5)
        constexpr T c2 = (T)-0.80901699437494742410229341718281906;  // cos(4pi/5)
        constexpr T s1 = (T)0.95105651629515357211643933337938214;   // sin(2pi/5)
        constexpr T s2 = (T)0.58778525229247312916870595463907277;   // sin(4pi/5)
        const V t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        const V t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        const V a1 = mk<V>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        const V a2 = mk<V>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        const V b1 = mul_i<DIR>(mk<V>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
        const V b2 = mul_i<DIR>(mk<V>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(a1, b1);
        v[4] = csub(a1, b1);
        v[2] = cadd(a2, b2);
        v[3] = csub(a2, b2);
    }
};

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
template <int N_, int E_, int... Rs> struct Plan {
    static constexpr int N = N_;
    static constexpr int E = E_;
    static constexpr int T = N_ / E_;
    static constexpr int NSTAGE = sizeof...(Rs);
    static constexpr int radix_arr[NSTAGE] = {Rs...};
    static constexpr int radix(int s) { return radix_arr[s]; }
    static constexpr int ns(int s)
    {
        int p = 1;
        for (int i = 0; i < s; ++i) p *= radix_arr[i];
        return p;
    }
    static constexpr bool valid()
    {
        int p = 1;
        for (int i = 0; i < NSTAGE; ++i) {
            if (E_ % radix_arr[i]) return false;
            p *= radix_arr[i];
        }
        return p == N_ && N_ % E_ == 0 && (E_ % 2 == 0);
    }
    static_assert(valid(), "invalid FFT plan");
    // register slot holding physical element t + m*T after the centring fold
    static constexpr int rot(int m) { return (m + E_ / 2) % E_; }
};

// Stage compute: twiddles + E/R butterflies of radix R in registers.
// tw: N-entry table tw[i] = exp(-2 pi i * i / N) in the plan's precision.
template <class PL, int S, int DIR, typename V>
CPRA_HD void stage_compute(V (&v)[PL::E], int t, const V* __restrict__ tw)
{
    constexpr int R = PL::radix(S);
    constexpr int NS = PL::ns(S);
    constexpr int Q = PL::E / R;
    constexpr int T = PL::T;
    constexpr int STRIDE = PL::N / (NS * R);
    sfor<Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        V b[R];
        sfor<R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            b[r] = v[q + r * Q];
        });
        if constexpr (NS > 1) {
            const int j = t + q * T;
            const int k = j % NS;
            sfor<R - 1>([&](auto rc) {
                constexpr int r = decltype(rc)::value + 1;
                V w = tw[r * k * STRIDE];
                if constexpr (DIR > 0) w.y = -w.y;
                b[r] = cmul(b[r], w);
            });
        }
        Dft<R, DIR>::run(b);
        sfor<R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            v[q + r * Q] = b[r];
        });
    });
}

// Stockham output positions of stage S for this thread's butterflies.
template <class PL, int S, typename V, class Acc>
CPRA_HD void stage_scatter(const V (&v)[PL::E], int t, Acc& acc)
{
    constexpr int R = PL::radix(S);
    constexpr int NS = PL::ns(S);
    constexpr int Q = PL::E / R;
    constexpr int T = PL::T;
    sfor<Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int j = t + q * T;
        const int idxd = (j / NS) * NS * R + (j % NS);
        sfor<R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            acc.st(idxd + r * NS, v[q + r * Q]);
        });
    });
}

template <class PL, typename V, class Acc>
CPRA_HD void gather_natural(V (&v)[PL::E], int t, Acc& acc)
{
    sfor<PL::E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        v[m] = acc.ld(t + m * PL::T);
    });
}

// Full transform of the line held in v (natural register layout in and out),
// unnormalised, sign DIR.  All threads of the block must call it (it syncs).
template <class PL, int DIR, typename V, class Acc>
CPRA_HD void fft_line(V (&v)[PL::E], int t, const V* __restrict__ tw, Acc& acc)
{
    sfor<PL::NSTAGE>([&](auto sc) {
        constexpr int S = decltype(sc)::value;
        stage_compute<PL, S, DIR>(v, t, tw);
        if constexpr (S + 1 < PL::NSTAGE) {
            stage_scatter<PL, S>(v, t, acc);
            acc.sync();
            gather_natural<PL>(v, t, acc);
            acc.sync();
        }
    });
}

// Orders this wave's LDS writes before its later LDS reads (and the reads
// before later writes).  Enough when every lane that exchanges data through
// an LDS region belongs to one wave: a wave's LDS instructions complete in
// issue order, so only compiler reordering has to be prevented.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS accessor for contiguous-line kernels: one region of N + N/16 values per
// line; the +1-per-16 padding keeps the radix-16 scatter conflict-free.
// WAVE: the T threads of a line sit in one wave (64 % T == 0), so the
// exchange needs no workgroup barrier.
template <typename V, int N, bool WAVE = false> struct LdsRow {
    V* base;
    static constexpr int span() { return N + N / 16; }
    CPRA_HD void st(int i, V x) { base[i + (i >> 4)] = x; }
    CPRA_HD V ld(int i) const { return base[i + (i >> 4)]; }
    __device__ __forceinline__ void sync()
    {
        if constexpr (WAVE) wave_lds_sync();
        else __syncthreads();
    }
};

// LDS accessor for strided-line kernels: element-major [N][C] so that the C
// lanes that share an element index touch consecutive words.
template <typename V, int C, bool WAVE = false> struct LdsCol {
    V* base;  // base + line
    CPRA_HD void st(int i, V x) { base[i * C] = x; }
    CPRA_HD V ld(int i) const { return base[i * C]; }
    __device__ __forceinline__ void sync()
    {
        if constexpr (WAVE) wave_lds_sync();
        else __syncthreads();
    }
};

}  // namespace cpra
