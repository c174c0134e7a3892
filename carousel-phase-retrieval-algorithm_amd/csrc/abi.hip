// abi.hip -- extern "C" entry points of libcpra_hip.so (lifetime, memory,
// streams, the pointwise CpraImpl primitives, shifts and ConvergeError) and
// the helpers shared with the engine translation units.
//
// Every entry point validates its arguments before any HIP call and maps
// exceptions to the status codes of include/cpra_hip.h.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace cpra {

static thread_local std::string g_thread_error;

template <typename F> static int guarded(cpra_hip_ctx* h, F&& f)
{
    try {
        return f();
    } catch (const Status& s) {
        if (h) h->last_error = s.msg;
        g_thread_error = s.msg;
        return s.code;
    } catch (const std::exception& e) {
        if (h) h->last_error = e.what();
        g_thread_error = e.what();
        return CPRA_HIP_ERR_HIP;
    }
}

size_t dtype_size(int dtype) { return dtype == CPRA_HIP_F64 ? 8 : 4; }

bool debug_launches()
{
    static const bool on = [] {
        const char* v = std::getenv("CPRA_HIP_DEBUG");
        return v && v[0] && v[0] != '0';
    }();
    return on;
}

const void* twiddle_table(cpra_hip_ctx* h, int n, bool dbl)
{
    const uint64_t key = ((uint64_t)n << 1) | (dbl ? 1u : 0u);
    auto it = h->twiddles.find(key);
    if (it != h->twiddles.end()) return it->second;
    const double two_pi = 6.283185307179586476925286766559;
    std::vector<double> tabd(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        // cos/sin of -2 pi k/n in double, via the first-octant argument
        const double a = two_pi * (double)k / (double)n;
        tabd[2 * k] = std::cos(a);
        tabd[2 * k + 1] = -std::sin(a);
    }
    void* dev = nullptr;
    if (dbl) {
        CPRA_CHECK_HIP(hipMalloc(&dev, sizeof(double) * 2 * n));
        CPRA_CHECK_HIP(hipMemcpy(dev, tabd.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tabf(2 * (size_t)n);
        for (size_t i = 0; i < tabf.size(); ++i) tabf[i] = (float)tabd[i];
        CPRA_CHECK_HIP(hipMalloc(&dev, sizeof(float) * 2 * n));
        CPRA_CHECK_HIP(hipMemcpy(dev, tabf.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    }
    h->twiddles[key] = dev;
    return dev;
}

void centred_2d(cpra_hip_ctx* h, void* x, int dir);  // fft.hip
void centred_3d(cpra_hip_ctx* h, void* x, int dir);  // fft.hip

// ---------------------------------------------------------------------------
// pointwise kernels
//
// The reference broadcasts a constraint array of per = num/batch values with
// index i % per.  The kernels walk j in [0, per) on the x grid and the
// repeats r in [0, reps) on the y grid (i = r*per + j), so each constraint
// value is loaded once per thread and the modulo disappears.
// ---------------------------------------------------------------------------
struct Bcast {
    uint64_t per;   // num / batch
    uint64_t reps;  // ceil(num / per)
    uint64_t num;
};

static Bcast make_bcast(uint64_t num, uint64_t batch)
{
    Bcast b;
    b.num = num;
    b.per = num / batch;
    b.reps = (num + b.per - 1) / b.per;
    return b;
}

#define CPRA_BCAST_LOOP(BODY)                                                            \
    const uint64_t rchunk = (bc.reps + gridDim.y - 1) / gridDim.y;                        \
    const uint64_t r0 = (uint64_t)blockIdx.y * rchunk;                                    \
    const uint64_t r1 = r0 + rchunk < bc.reps ? r0 + rchunk : bc.reps;                    \
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < bc.per;        \
         j += (uint64_t)gridDim.x * blockDim.x) {                                         \
        for (uint64_t r = r0; r < r1; ++r) {                                              \
            const uint64_t i = r * bc.per + j;                                            \
            if (i >= bc.num) break;                                                       \
            BODY                                                                          \
        }                                                                                 \
    }

template <typename T>
__global__ void k_space_mkl(c2<T>* __restrict__ x, const T* __restrict__ S, Bcast bc)
{
    // MKL_Impl.inl:77-82: the compare happens in double for both T
    const bool dummy = false;
    (void)dummy;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < bc.per;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const bool zero = (double)S[j] < 0.9;
        const uint64_t rchunk = (bc.reps + gridDim.y - 1) / gridDim.y;
        const uint64_t r0 = (uint64_t)blockIdx.y * rchunk;
        const uint64_t r1 = r0 + rchunk < bc.reps ? r0 + rchunk : bc.reps;
        for (uint64_t r = r0; r < r1; ++r) {
            const uint64_t i = r * bc.per + j;
            if (i >= bc.num) break;
            c2<T> v = x[i];
            v.y = (T)0;
            if (zero) v.x = (T)0;
            x[i] = v;
        }
    }
}

template <typename T>
__global__ void k_space_pos(c2<T>* __restrict__ x, const T* __restrict__ S, Bcast bc)
{
    // CUDA_Impl_Kernel.cuh:20-27
    CPRA_BCAST_LOOP({
        c2<T> v = x[i];
        v.y = (T)0;
        v.x = v.x < (T)0 ? (T)0 : v.x * S[j];
        x[i] = v;
    })
}

template <typename T>
__global__ void k_data_real(c2<T>* __restrict__ x, const T* __restrict__ A, Bcast bc, int skip_zero)
{
    // MKL_Impl.inl:93-104 (skip_zero) / CUDA_Impl_Kernel.cuh:33-37
    CPRA_BCAST_LOOP({
        const T a = A[j];
        if (!(skip_zero && a == (T)0)) {
            c2<T> v = x[i];
            const T m2 = v.x * v.x + v.y * v.y;
            if (m2 > (T)0) {
                const T s = a / sqrt(m2);
                v.x *= s;
                v.y *= s;
                x[i] = v;
            }
        }
    })
}

template <typename T>
__global__ void k_data_cplx(c2<T>* __restrict__ x, const c2<T>* __restrict__ Cc, Bcast bc)
{
    // MKL_Impl.inl:116-124
    CPRA_BCAST_LOOP({
        const c2<T> c = Cc[j];
        if (c.x * c.x + c.y * c.y != (T)0) x[i] = c;
    })
}

template <typename T>
__global__ void k_axpby(const c2<T>* src, c2<T>* dst, T alpha, T beta, uint64_t num)
{
    // MKL_Impl.inl:41-48 (caxpby with real alpha, beta); src may alias dst
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < num;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const c2<T> s = src[i];
        c2<T> d = dst[i];
        d.x = alpha * s.x + beta * d.x;
        d.y = alpha * s.y + beta * d.y;
        dst[i] = d;
    }
}

template <typename T> __global__ void k_scal(c2<T>* x, T a, uint64_t num)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < num;
         i += (uint64_t)gridDim.x * blockDim.x) {
        c2<T> v = x[i];
        v.x *= a;
        v.y *= a;
        x[i] = v;
    }
}

// Quadrant swap (MKL_Impl.inl:186-207): thread (col, row < M/2, b) swaps
// (row, col) with (row + M/2, (col + N/2) mod N).  One 16-byte element is a
// plain copy, so the permutation is bit-exact for any T.
template <typename E>
__global__ void k_shift2d(E* __restrict__ x, uint64_t M, uint64_t N, uint64_t B)
{
    const uint64_t halfM = M / 2, halfN = N / 2;
    const uint64_t total = halfM * N * B;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t col = i % N;
        const uint64_t rest = i / N;
        const uint64_t row = rest % halfM;
        const uint64_t b = rest / halfM;
        const uint64_t off = b * M * N;
        const uint64_t ccol = col < halfN ? col + halfN : col - halfN;
        E* p = x + off + row * N + col;
        E* q = x + off + (row + halfM) * N + ccol;
        const E a = *p;
        *p = *q;
        *q = a;
    }
}

// Octant swap (MKL_Impl.inl:218-249).
template <typename E>
__global__ void k_shift3d(E* __restrict__ x, uint64_t M, uint64_t N, uint64_t L, uint64_t B)
{
    const uint64_t halfL = L / 2, halfM = M / 2, halfN = N / 2;
    const uint64_t total = halfL * M * N * B;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t col = i % N;
        uint64_t rest = i / N;
        const uint64_t row = rest % M;
        rest /= M;
        const uint64_t dep = rest % halfL;
        const uint64_t b = rest / halfL;
        const uint64_t off = b * M * N * L;
        const uint64_t crow = row < halfM ? row + halfM : row - halfM;
        const uint64_t ccol = col < halfN ? col + halfN : col - halfN;
        E* p = x + off + (dep * M + row) * N + col;
        E* q = x + off + ((dep + halfL) * M + crow) * N + ccol;
        const E a = *p;
        *p = *q;
        *q = a;
    }
}

// ConvergeError part 1: old <- new - old and per-block partial sums of
// |old|^2 and |new|^2 (in double) for batch member blockIdx.y.
template <typename T>
__global__ __launch_bounds__(256) void k_conv_err(c2<T>* __restrict__ old_,
                                                  const c2<T>* __restrict__ nw, uint64_t per,
                                                  double* __restrict__ part)
{
    const uint64_t b = blockIdx.y;
    c2<T>* o = old_ + b * per;
    const c2<T>* w = nw + b * per;
    double so = 0.0, sn = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < per;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const c2<T> a = w[i];
        c2<T> d = o[i];
        d.x = a.x - d.x;
        d.y = a.y - d.y;
        o[i] = d;
        so += (double)d.x * d.x + (double)d.y * d.y;
        sn += (double)a.x * a.x + (double)a.y * a.y;
    }
    __shared__ double red[2][256];
    red[0][threadIdx.x] = so;
    red[1][threadIdx.x] = sn;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * (b * gridDim.x + blockIdx.x)] = red[0][0];
        part[2 * (b * gridDim.x + blockIdx.x) + 1] = red[1][0];
    }
}

// part 2: one thread per batch member sums its partials in a fixed order.
template <typename T>
__global__ void k_conv_err_fin(const double* __restrict__ part, unsigned nblk, uint64_t B,
                               T* __restrict__ err)
{
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double so = 0.0, sn = 0.0;
    for (unsigned k = 0; k < nblk; ++k) {
        so += part[2 * (b * nblk + k)];
        sn += part[2 * (b * nblk + k) + 1];
    }
    err[b] = (T)(sqrt(so) / sqrt(sn));
}

// Counter-based uniform [0, 1): splitmix64 of (seed, index), top 24 (float)
// or 53 (double) bits.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

template <typename T> __global__ void k_uniform(T* __restrict__ x, uint64_t num, uint64_t seed)
{
    const uint64_t key = splitmix64(seed ^ 0x6a09e667f3bcc909ull);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < num;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = splitmix64(key + i);
        if (sizeof(T) == 4)
            x[i] = (T)((float)(r >> 40) * (1.0f / 16777216.0f));
        else
            x[i] = (T)((double)(r >> 11) * (1.0 / 9007199254740992.0));
    }
}

static dim3 bcast_grid(const Bcast& bc)
{
    const unsigned gx = grid_1d(bc.per, 256, 65536);
    uint64_t want_y = (bc.num / 256 + gx - 1) / gx;  // enough blocks to fill the chip
    if (want_y < 1) want_y = 1;
    if (want_y > bc.reps) want_y = bc.reps;
    if (want_y > 65535) want_y = 65535;
    return dim3(gx, (unsigned)want_y);
}

}  // namespace cpra

using namespace cpra;

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
extern "C" {

const char* cpra_hip_version(void) { return "cpra-hip 0.1 (gfx950)"; }

const char* cpra_hip_last_error(cpra_hip_handle h)
{
    if (h) return h->last_error.c_str();
    return g_thread_error.c_str();
}

int cpra_hip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int cpra_hip_create(cpra_hip_handle* out, int dtype, uint64_t m, uint64_t n, uint64_t l,
                    uint64_t batch)
{
    return guarded(nullptr, [&]() -> int {
        CPRA_ARG(out != nullptr, "cpra_hip_create: out is NULL");
        *out = nullptr;
        CPRA_ARG(dtype == CPRA_HIP_F32 || dtype == CPRA_HIP_F64,
                 "cpra_hip_create: dtype must be CPRA_HIP_F32 or CPRA_HIP_F64");
        CPRA_ARG(m >= 1 && n >= 1 && l >= 1 && batch >= 1,
                 "cpra_hip_create: m, n, l and batch must be >= 1");
        CPRA_ARG(m <= (1u << 20) && n <= (1u << 20) && l <= (1u << 20),
                 "cpra_hip_create: dimension too large");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            fail(CPRA_HIP_ERR_NODEV, "cpra_hip_create: no GPU visible");
        auto* h = new cpra_hip_ctx();
        h->dtype = dtype;
        h->m = m;
        h->n = n;
        h->l = l;
        h->batch = batch;
        hipError_t e = hipGetDevice(&h->device);
        if (e == hipSuccess) e = hipStreamCreate(&h->own_stream);
        if (e != hipSuccess) {
            delete h;
            fail(CPRA_HIP_ERR_HIP, "cpra_hip_create: %s", hipGetErrorString(e));
        }
        h->stream = h->own_stream;
        *out = h;
        return CPRA_HIP_OK;
    });
}

int cpra_hip_destroy(cpra_hip_handle h)
{
    if (!h) return CPRA_HIP_ERR_ARG;
    (void)hipStreamSynchronize(h->stream);
    cpra_engine_release(h);
    if (h->capturing) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(h->stream, &g);
        if (g) (void)hipGraphDestroy(g);
    }
    for (auto& kv : h->user_graphs) {
        (void)hipGraphExecDestroy(kv.second.second);
        (void)hipGraphDestroy(kv.second.first);
    }
    for (auto& kv : h->twiddles) (void)hipFree(kv.second);
    h->ws_z.release();
    h->ws_state.release();
    h->ws_amp_t.release();
    h->ws_sup.release();
    h->ws_sbits.release();
    h->ws_blur.release();
    h->ws_keys.release();
    h->ws_part.release();
    h->ws_part2.release();
    h->ws_misc.release();
    h->ws_backup.release();
    h->ws_rows.release();
    h->ws_dm.release();
    h->ws_trig.release();
    h->ws_mcount.release();
    h->ws_mlist.release();
    h->ws_mheavy.release();
    h->ws_zsup.release();
    if (h->aux_stream) (void)hipStreamDestroy(h->aux_stream);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    delete h;
    return CPRA_HIP_OK;
}

int cpra_hip_set_stream(cpra_hip_handle h, void* stream)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "NULL handle");
        h->stream = static_cast<hipStream_t>(stream);
        return CPRA_HIP_OK;
    });
}

void* cpra_hip_get_stream(cpra_hip_handle h) { return h ? (void*)h->stream : nullptr; }

int cpra_hip_malloc(cpra_hip_handle h, void** ptr, uint64_t bytes, int managed)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(ptr, "cpra_hip_malloc: NULL argument");
        *ptr = nullptr;
        if (bytes == 0) return CPRA_HIP_FALSE;
        if (managed) CPRA_CHECK_HIP(hipMallocManaged(ptr, bytes));
        else CPRA_CHECK_HIP(hipMalloc(ptr, bytes));
        return CPRA_HIP_OK;
    });
}

int cpra_hip_free(cpra_hip_handle h, void* ptr)
{
    return guarded(h, [&]() -> int {
        if (!ptr) return CPRA_HIP_FALSE;
        CPRA_CHECK_HIP(hipFree(ptr));
        return CPRA_HIP_OK;
    });
}

int cpra_hip_memcpy(cpra_hip_handle h, void* dst, const void* src, uint64_t bytes)
{
    return guarded(h, [&]() -> int {
        if (bytes == 0) return CPRA_HIP_OK;
        CPRA_ARG(dst && src, "cpra_hip_memcpy: NULL pointer");
        if (h) CPRA_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, h->stream));
        else CPRA_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
        return CPRA_HIP_OK;
    });
}

int cpra_hip_sync(cpra_hip_handle h)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_sync: NULL handle");
        CPRA_ARG(!h->capturing, "cpra_hip_sync: the handle's stream is being captured");
        CPRA_CHECK_HIP(hipStreamSynchronize(h->stream));
        return CPRA_HIP_OK;
    });
}

int cpra_hip_graph_begin(cpra_hip_handle h)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_graph_begin: NULL handle");
        CPRA_ARG(!h->capturing, "cpra_hip_graph_begin: a capture is already open");
        CPRA_ARG(!h->timing, "cpra_hip_graph_begin: kernel timing is on");
        // relaxed: a workspace that still grows during the capture may call
        // hipMalloc (the allocation happens at capture time)
        CPRA_CHECK_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
        h->capturing = true;
        return CPRA_HIP_OK;
    });
}

int cpra_hip_graph_end(cpra_hip_handle h, uint64_t* graph_id)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h && graph_id, "cpra_hip_graph_end: NULL argument");
        CPRA_ARG(h->capturing, "cpra_hip_graph_end: no capture is open");
        h->capturing = false;
        hipGraph_t g = nullptr;
        CPRA_CHECK_HIP(hipStreamEndCapture(h->stream, &g));
        hipGraphExec_t e = nullptr;
        const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        if (err != hipSuccess) {
            (void)hipGraphDestroy(g);
            fail(CPRA_HIP_ERR_HIP, "cpra_hip_graph_end: instantiation failed: %s", hipGetErrorString(err));
        }
        const uint64_t id = h->next_graph++;
        h->user_graphs[id] = {g, e};
        *graph_id = id;
        return CPRA_HIP_OK;
    });
}

int cpra_hip_graph_launch(cpra_hip_handle h, uint64_t graph_id)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_graph_launch: NULL handle");
        CPRA_ARG(!h->capturing, "cpra_hip_graph_launch: a capture is open");
        auto it = h->user_graphs.find(graph_id);
        CPRA_ARG(it != h->user_graphs.end(), "cpra_hip_graph_launch: unknown graph %llu",
                 (unsigned long long)graph_id);
        CPRA_CHECK_HIP(hipGraphLaunch(it->second.second, h->stream));
        return CPRA_HIP_OK;
    });
}

int cpra_hip_graph_destroy(cpra_hip_handle h, uint64_t graph_id)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_graph_destroy: NULL handle");
        auto it = h->user_graphs.find(graph_id);
        CPRA_ARG(it != h->user_graphs.end(), "cpra_hip_graph_destroy: unknown graph %llu",
                 (unsigned long long)graph_id);
        CPRA_CHECK_HIP(hipGraphExecDestroy(it->second.second));
        CPRA_CHECK_HIP(hipGraphDestroy(it->second.first));
        h->user_graphs.erase(it);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_initialize(cpra_hip_handle h, void* data, uint64_t num, uint64_t seed)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_initialize: NULL handle");
        if (num == 0) return CPRA_HIP_FALSE;  // MKL_Impl.inl:11
        CPRA_ARG(data, "cpra_hip_initialize: NULL data");
        const unsigned g = grid_1d(num, 256, 8192);
        if (h->dtype == CPRA_HIP_F32)
            hipLaunchKernelGGL(k_uniform<float>, dim3(g), dim3(256), 0, h->stream, (float*)data, num, seed);
        else
            hipLaunchKernelGGL(k_uniform<double>, dim3(g), dim3(256), 0, h->stream, (double*)data, num, seed);
        check_launch("k_uniform");
        return CPRA_HIP_OK;
    });
}

static void check_even_2d(uint64_t M, uint64_t N, const char* who)
{
    CPRA_ARG(M >= 2 && N >= 2, "%s: sizes must be >= 2", who);
    if ((M & 1) || (N & 1))
        fail(CPRA_HIP_ERR_UNSUPPORTED, "%s: odd sizes are not supported (the shift is a quadrant swap)", who);
}

int cpra_hip_forward2d(cpra_hip_handle h, void* x)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h && x, "cpra_hip_forward2d: NULL argument");
        check_even_2d(h->m, h->n, "cpra_hip_forward2d");
        centred_2d(h, x, -1);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_backward2d(cpra_hip_handle h, void* x)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h && x, "cpra_hip_backward2d: NULL argument");
        check_even_2d(h->m, h->n, "cpra_hip_backward2d");
        centred_2d(h, x, +1);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_forward3d(cpra_hip_handle h, void* x)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h && x, "cpra_hip_forward3d: NULL argument");
        check_even_2d(h->m, h->n, "cpra_hip_forward3d");
        check_even_2d(h->l, h->n, "cpra_hip_forward3d");
        centred_3d(h, x, -1);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_backward3d(cpra_hip_handle h, void* x)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h && x, "cpra_hip_backward3d: NULL argument");
        check_even_2d(h->m, h->n, "cpra_hip_backward3d");
        check_even_2d(h->l, h->n, "cpra_hip_backward3d");
        centred_3d(h, x, +1);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_fftshift2d(cpra_hip_handle h, void* x, uint64_t M, uint64_t N, uint64_t batch)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h && x, "cpra_hip_fftshift2d: NULL argument");
        if (batch == 0) return CPRA_HIP_FALSE;
        check_even_2d(M, N, "cpra_hip_fftshift2d");
        const uint64_t total = M / 2 * N * batch;
        const unsigned g = grid_1d(total, 256, 65536);
        if (h->dtype == CPRA_HIP_F32)
            hipLaunchKernelGGL(k_shift2d<uint64_t>, dim3(g), dim3(256), 0, h->stream, (uint64_t*)x, M, N, batch);
        else
            hipLaunchKernelGGL(k_shift2d<c2<double>>, dim3(g), dim3(256), 0, h->stream, (c2<double>*)x, M, N, batch);
        check_launch("k_shift2d");
        return CPRA_HIP_OK;
    });
}

int cpra_hip_fftshift3d(cpra_hip_handle h, void* x, uint64_t M, uint64_t N, uint64_t L,
                        uint64_t batch)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h && x, "cpra_hip_fftshift3d: NULL argument");
        if (batch == 0) return CPRA_HIP_FALSE;
        check_even_2d(M, N, "cpra_hip_fftshift3d");
        check_even_2d(L, N, "cpra_hip_fftshift3d");
        const uint64_t total = L / 2 * M * N * batch;
        const unsigned g = grid_1d(total, 256, 65536);
        if (h->dtype == CPRA_HIP_F32)
            hipLaunchKernelGGL(k_shift3d<uint64_t>, dim3(g), dim3(256), 0, h->stream, (uint64_t*)x, M, N, L, batch);
        else
            hipLaunchKernelGGL(k_shift3d<c2<double>>, dim3(g), dim3(256), 0, h->stream, (c2<double>*)x, M, N, L, batch);
        check_launch("k_shift3d");
        return CPRA_HIP_OK;
    });
}

int cpra_hip_space_constraint(cpra_hip_handle h, void* x, const void* S, uint64_t num,
                              uint64_t batch, int mode)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_space_constraint: NULL handle");
        if (num == 0 || batch == 0) return CPRA_HIP_FALSE;  // MKL_Impl.inl:73
        CPRA_ARG(x && S, "cpra_hip_space_constraint: NULL pointer");
        CPRA_ARG(batch <= num, "cpra_hip_space_constraint: batch > num");
        CPRA_ARG(mode == CPRA_HIP_SPACE_MKL || mode == CPRA_HIP_SPACE_POSITIVE,
                 "cpra_hip_space_constraint: unknown mode %d", mode);
        const Bcast bc = make_bcast(num, batch);
        const dim3 g = bcast_grid(bc);
        if (h->dtype == CPRA_HIP_F32) {
            if (mode == CPRA_HIP_SPACE_MKL)
                hipLaunchKernelGGL(k_space_mkl<float>, g, dim3(256), 0, h->stream, (c2<float>*)x, (const float*)S, bc);
            else
                hipLaunchKernelGGL(k_space_pos<float>, g, dim3(256), 0, h->stream, (c2<float>*)x, (const float*)S, bc);
        } else {
            if (mode == CPRA_HIP_SPACE_MKL)
                hipLaunchKernelGGL(k_space_mkl<double>, g, dim3(256), 0, h->stream, (c2<double>*)x, (const double*)S, bc);
            else
                hipLaunchKernelGGL(k_space_pos<double>, g, dim3(256), 0, h->stream, (c2<double>*)x, (const double*)S, bc);
        }
        check_launch("k_space");
        return CPRA_HIP_OK;
    });
}

int cpra_hip_data_constraint_real(cpra_hip_handle h, void* x, const void* A, uint64_t num,
                                  uint64_t batch, int skip_zero)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_data_constraint_real: NULL handle");
        if (num == 0 || batch == 0) return CPRA_HIP_FALSE;  // MKL_Impl.inl:89
        CPRA_ARG(x && A, "cpra_hip_data_constraint_real: NULL pointer");
        CPRA_ARG(batch <= num, "cpra_hip_data_constraint_real: batch > num");
        const Bcast bc = make_bcast(num, batch);
        const dim3 g = bcast_grid(bc);
        if (h->dtype == CPRA_HIP_F32)
            hipLaunchKernelGGL(k_data_real<float>, g, dim3(256), 0, h->stream, (c2<float>*)x, (const float*)A, bc, skip_zero);
        else
            hipLaunchKernelGGL(k_data_real<double>, g, dim3(256), 0, h->stream, (c2<double>*)x, (const double*)A, bc, skip_zero);
        check_launch("k_data_real");
        return CPRA_HIP_OK;
    });
}

int cpra_hip_data_constraint_complex(cpra_hip_handle h, void* x, const void* Cc, uint64_t num,
                                     uint64_t batch)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_data_constraint_complex: NULL handle");
        if (num == 0 || batch == 0) return CPRA_HIP_FALSE;  // MKL_Impl.inl:112
        CPRA_ARG(x && Cc, "cpra_hip_data_constraint_complex: NULL pointer");
        CPRA_ARG(batch <= num, "cpra_hip_data_constraint_complex: batch > num");
        const Bcast bc = make_bcast(num, batch);
        const dim3 g = bcast_grid(bc);
        if (h->dtype == CPRA_HIP_F32)
            hipLaunchKernelGGL(k_data_cplx<float>, g, dim3(256), 0, h->stream, (c2<float>*)x, (const c2<float>*)Cc, bc);
        else
            hipLaunchKernelGGL(k_data_cplx<double>, g, dim3(256), 0, h->stream, (c2<double>*)x, (const c2<double>*)Cc, bc);
        check_launch("k_data_cplx");
        return CPRA_HIP_OK;
    });
}

int cpra_hip_merge_add(cpra_hip_handle h, const void* src, void* dst, double alpha, double beta,
                       uint64_t num)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_merge_add: NULL handle");
        if (alpha == 0.0 || num == 0) return CPRA_HIP_FALSE;  // MKL_Impl.inl:34
        CPRA_ARG(src && dst, "cpra_hip_merge_add: NULL pointer");
        const unsigned g = grid_1d(num, 256, 65536);
        if (h->dtype == CPRA_HIP_F32)
            hipLaunchKernelGGL(k_axpby<float>, dim3(g), dim3(256), 0, h->stream, (const c2<float>*)src, (c2<float>*)dst, (float)alpha, (float)beta, num);
        else
            hipLaunchKernelGGL(k_axpby<double>, dim3(g), dim3(256), 0, h->stream, (const c2<double>*)src, (c2<double>*)dst, alpha, beta, num);
        check_launch("k_axpby");
        return CPRA_HIP_OK;
    });
}

int cpra_hip_normalize(cpra_hip_handle h, void* x, double norm, uint64_t num)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_normalize: NULL handle");
        if (norm == 0.0 || num == 0) return CPRA_HIP_FALSE;  // MKL_Impl.inl:55
        CPRA_ARG(x, "cpra_hip_normalize: NULL pointer");
        const unsigned g = grid_1d(num, 256, 65536);
        if (h->dtype == CPRA_HIP_F32) {
            const float a = (float)(1.0 / (double)(float)norm);  // MKL_Impl.inl:57 a.real(1.0/norm)
            hipLaunchKernelGGL(k_scal<float>, dim3(g), dim3(256), 0, h->stream, (c2<float>*)x, a, num);
        } else {
            hipLaunchKernelGGL(k_scal<double>, dim3(g), dim3(256), 0, h->stream, (c2<double>*)x, 1.0 / norm, num);
        }
        check_launch("k_scal");
        return CPRA_HIP_OK;
    });
}

int cpra_hip_converge_error(cpra_hip_handle h, void* old_, const void* new_, void* host_err,
                            uint64_t m, uint64_t n, uint64_t l, uint64_t batch)
{
    return guarded(h, [&]() -> int {
        CPRA_ARG(h, "cpra_hip_converge_error: NULL handle");
        CPRA_ARG(!h->capturing, "cpra_hip_converge_error: its host result needs a synchronisation, "
                                "which a stream capture cannot hold");
        if (batch == 0 || m == 0 || n == 0 || l == 0) return CPRA_HIP_FALSE;
        CPRA_ARG(old_ && new_ && host_err, "cpra_hip_converge_error: NULL pointer");
        CPRA_ARG(batch <= 65535, "cpra_hip_converge_error: batch > 65535");
        const uint64_t per = m * n * l;
        const unsigned nblk = grid_1d(per, 256, 64);
        const size_t es = dtype_size(h->dtype);
        h->ws_misc.ensure(sizeof(double) * 2 * nblk * batch + es * batch + 64);
        double* part = static_cast<double*>(h->ws_misc.ptr);
        void* dev_err = static_cast<char*>(h->ws_misc.ptr) + sizeof(double) * 2 * nblk * batch;
        if (h->dtype == CPRA_HIP_F32) {
            hipLaunchKernelGGL(k_conv_err<float>, dim3(nblk, (unsigned)batch), dim3(256), 0, h->stream, (c2<float>*)old_, (const c2<float>*)new_, per, part);
            hipLaunchKernelGGL(k_conv_err_fin<float>, dim3(grid_1d(batch, 64)), dim3(64), 0, h->stream, part, nblk, batch, (float*)dev_err);
        } else {
            hipLaunchKernelGGL(k_conv_err<double>, dim3(nblk, (unsigned)batch), dim3(256), 0, h->stream, (c2<double>*)old_, (const c2<double>*)new_, per, part);
            hipLaunchKernelGGL(k_conv_err_fin<double>, dim3(grid_1d(batch, 64)), dim3(64), 0, h->stream, part, nblk, batch, (double*)dev_err);
        }
        check_launch("k_conv_err");
        CPRA_CHECK_HIP(hipMemcpyAsync(host_err, dev_err, es * batch, hipMemcpyDefault, h->stream));
        CPRA_CHECK_HIP(hipStreamSynchronize(h->stream));
        return CPRA_HIP_OK;
    });
}

}  // extern "C"
