// merge3d.hip -- carousel 2D -> 3D merge (gather, trilinear) and re-projection.
//
// Geometry (cube of side n, centre c = n/2): slice p at angle th_p is an
// [n][n] image with rows u along the rotation axis z and columns v; its sample
// (u, v) sits at the volume point x = c + (v-c) cos th, y = c + (v-c) sin th,
// z = u.  R (re-projection, cpra_hip_slice_extract) samples the volume
// [z][y][x] trilinearly at those points (zero outside); since z is on the grid
// the weights are bilinear in (x, y).  The merge is V = R^T(s) / R^T(1) where
// R^T(1) > 0, else 0.  It replaces the reference's racy nearest-voxel scatter
// (include/detail/CUDA_Impl_Kernel.cuh:68-119; a stub in the MKL backend,
// include/MKL_Impl.hpp:151-152) with a gather: every voxel column (x, y) finds
// the (p, v, w) samples whose footprint covers it -- the same for every z --
// and then sums w * s[p][z][v] over that list.
//
// Kernel shape: one wave64 per voxel column at a time; the lanes first walk
// 64 angles each (a sample s = v - c can touch (x, y) only if
// |s - s0| <= sqrt 2 with s0 the projection of (x - c, y - c) on the slice
// direction, so 4 candidates per angle), compact the non-zero weights into an
// LDS list with a ballot prefix, then walk z (lanes over z, slices transposed
// to [p][v][z] so the reads coalesce).  A block owns XT consecutive columns at
// one y and stages its [n][XT] output tile in LDS so the volume rows are
// written contiguously.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace cpra {

// s[p][z][v] -> sT[p][v][z], 32 x 32 LDS tiles
template <typename V>
__global__ __launch_bounds__(256) void k_transpose_slices(const V* __restrict__ s, V* __restrict__ sT,
                                                         int n)
{
    __shared__ V tile[32][33];
    const int p = blockIdx.z;
    const int v0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
    const V* sp = s + (int64_t)p * n * n;
    V* tp = sT + (int64_t)p * n * n;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int z = z0 + k, v = v0 + tx;
        if (z < n && v < n) tile[k][tx] = sp[(int64_t)z * n + v];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int v = v0 + k, z = z0 + tx;
        if (z < n && v < n) tp[(int64_t)v * n + z] = tile[tx][k];
    }
}

template <typename T> struct ElemOps;
template <> struct ElemOps<float> {
    using V = float;
    static __device__ __forceinline__ V zero() { return 0.f; }
    static __device__ __forceinline__ V fma(float w, V a, V acc) { return acc + w * a; }
    static __device__ __forceinline__ V scale(V a, float s) { return a * s; }
};
template <> struct ElemOps<double> {
    using V = double;
    static __device__ __forceinline__ V zero() { return 0.0; }
    static __device__ __forceinline__ V fma(double w, V a, V acc) { return acc + w * a; }
    static __device__ __forceinline__ V scale(V a, double s) { return a * s; }
};
template <typename T> struct ElemOps<c2<T>> {
    using V = c2<T>;
    static __device__ __forceinline__ V zero() { return mk<V>(0, 0); }
    static __device__ __forceinline__ V fma(T w, V a, V acc) { return mk<V>(acc.x + w * a.x, acc.y + w * a.y); }
    static __device__ __forceinline__ V scale(V a, T s) { return mk<V>(a.x * s, a.y * s); }
};

constexpr int kMergeMaxZPerLane = 8;  // n <= 512

// (cos th_p, sin th_p) in FP64 for the slice geometry.  The footprint
// offsets s*cos - dx and the weights derived from them are formed in FP64 and
// rounded to T only at the end: in FP32, |s| cos th carries an absolute error
// of ~4e-6 at s = 64, which is a large relative error on the small weights at
// the edge of a slice's footprint, and R^T(s)/R^T(1) divides by their sum.
template <typename T>
__global__ void k_trig_table(const T* __restrict__ angles, int P, double* __restrict__ trig)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double th = (double)angles[p];
    trig[2 * p] = cos(th);
    trig[2 * p + 1] = sin(th);
}

// E: element type (T or c2<T>), T: real type of the weights/angles.
template <typename E, typename T>
__global__ __launch_bounds__(256) void k_merge_gather(const E* __restrict__ sT, const double* __restrict__ trig,
                                                     int P, int n, int XT, E* __restrict__ vol)
{
    using Ops = ElemOps<E>;
    extern __shared__ unsigned char smem[];
    // per wave: list of up to 256 (p*n + v, w) entries; then the output tile
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int* lidx = reinterpret_cast<int*>(smem) + wave * 256;
    T* lw = reinterpret_cast<T*>(smem + 4 * 256 * sizeof(int)) + wave * 256;
    E* tile = reinterpret_cast<E*>(smem + 4 * 256 * (sizeof(int) + sizeof(T)));  // [n][XT]
    T* tw = reinterpret_cast<T*>(tile + (int64_t)n * XT);                       // [XT]
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * XT;
    const int c = n / 2;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int xl = wave; xl < XT; xl += 4) {
        const int x = x0 + xl;
        const double dx = (double)(x - c), dy = (double)(y - c);
        E acc[kMergeMaxZPerLane];
#pragma unroll
        for (int k = 0; k < kMergeMaxZPerLane; ++k) acc[k] = Ops::zero();
        T wsum = 0;
        for (int p0 = 0; p0 < P; p0 += 64) {
            const int p = p0 + lane;
            double cs = 0, sn = 0;
            bool live = p < P;
            if (live) {
                cs = trig[2 * p];
                sn = trig[2 * p + 1];
                const double dperp = -dx * sn + dy * cs;
                live = fabs(dperp) < 1.4142135623730951;
            }
            const double s0 = dx * cs + dy * sn;
            const int sb = (int)floor(s0) - 1;
            int cnt_base = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int s = sb + k;
                const int v = s + n / 2;
                T w = 0;
                if (live && v >= 0 && v < n) {
                    const double ex = (double)s * cs - dx, ey = (double)s * sn - dy;
                    const double wx = 1.0 - fabs(ex), wy = 1.0 - fabs(ey);
                    if (wx > 0.0 && wy > 0.0) w = (T)(wx * wy);
                }
                const bool has = w > (T)0;
                const uint64_t m = __ballot(has);
                if (has) {
                    const int pos = cnt_base + __popcll(m & lt);
                    lidx[pos] = p * n + v;
                    lw[pos] = w;
                }
                cnt_base += __popcll(m);
            }
            // the list is written and read by this wave only
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int e = 0; e < cnt_base; ++e) {
                const int pv = lidx[e];
                const T w = lw[e];
                wsum += w;
                const E* col = sT + (int64_t)pv * n;
#pragma unroll
                for (int k = 0; k < kMergeMaxZPerLane; ++k) {
                    const int z = lane + 64 * k;
                    if (z < n) acc[k] = Ops::fma(w, col[z], acc[k]);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        const T inv = wsum > (T)0 ? (T)1 / wsum : (T)0;
#pragma unroll
        for (int k = 0; k < kMergeMaxZPerLane; ++k) {
            const int z = lane + 64 * k;
            if (z < n) tile[(int64_t)z * XT + xl] = Ops::scale(acc[k], inv);
        }
        if (lane == 0) tw[xl] = wsum;
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < (int64_t)n * XT; i += blockDim.x) {
        const int z = (int)(i / XT), xl = (int)(i % XT);
        vol[((int64_t)z * n + y) * n + x0 + xl] = tile[i];
    }
}

// R: slices[p][u][v] = bilinear sample of the plane z = u at the slice point.
template <typename E, typename T>
__global__ __launch_bounds__(256) void k_slice_extract(const E* __restrict__ vol, const double* __restrict__ trig,
                                                      int P, int n, E* __restrict__ out)
{
    using Ops = ElemOps<E>;
    const int64_t total = (int64_t)P * n * n;
    const double c = (double)(n / 2);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int v = (int)(i % n);
        const int u = (int)((i / n) % n);
        const int p = (int)(i / ((int64_t)n * n));
        // the sample point in FP64 (see k_trig_table), the weights in T
        const double pxd = c + ((double)v - c) * trig[2 * p], pyd = c + ((double)v - c) * trig[2 * p + 1];
        const double fxd = floor(pxd), fyd = floor(pyd);
        const T fx = (T)fxd, fy = (T)fyd;
        const T ax = (T)(pxd - fxd), ay = (T)(pyd - fyd);
        E acc = Ops::zero();
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dxx = 0; dxx < 2; ++dxx) {
                const int xi = (int)fx + dxx, yi = (int)fy + dy;
                const T w = (dxx ? ax : (T)1 - ax) * (dy ? ay : (T)1 - ay);
                if (xi >= 0 && xi < n && yi >= 0 && yi < n && w != (T)0)
                    acc = Ops::fma(w, vol[((int64_t)u * n + yi) * n + xi], acc);
            }
        out[i] = acc;
    }
}

// LDS of the plane-staged extraction: the plane (at most 128 KiB) after a
// 4 KiB table of its angles' (cos, sin) (at most 256 angles per group); the
// CU has 160 KiB
constexpr size_t kExtractPlaneLds = 128 * 1024;
// the largest carousel size (the merge's LDS tile is n x kMergeXT elements)
constexpr int kMergeMaxN = 64 * kMergeMaxZPerLane;
constexpr size_t kExtractTrigLds = 4096;

// R with the plane in LDS (n*n elements fit): block (u, angle group) loads
// plane z = u once, coalesced, and computes the samples of its angles from
// LDS, writing each slice row along v.  Same arithmetic as k_slice_extract.
template <typename E, typename T>
__global__ __launch_bounds__(1024) void k_extract_plane(const E* __restrict__ vol, const double* __restrict__ trig,
                                                       int P, int n, int groups, E* __restrict__ out)
{
    using Ops = ElemOps<E>;
    extern __shared__ unsigned char smem[];
    double* tl = reinterpret_cast<double*>(smem);              // [2 * (p1 - p0)]
    E* plane = reinterpret_cast<E*>(smem + kExtractTrigLds);  // [n][n]
    const int ld = n;
    const int u = blockIdx.x / groups, g = blockIdx.x % groups;
    const int p0 = (int)((int64_t)g * P / groups), p1 = (int)((int64_t)(g + 1) * P / groups);
    const E* src = vol + (int64_t)u * n * n;
    // eight loads in flight per thread before their LDS stores
    for (int base = 0; base < n * n; base += 8 * (int)blockDim.x) {
        E r[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int i = base + j * (int)blockDim.x + (int)threadIdx.x;
            if (i < n * n) r[j] = src[i];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int i = base + j * (int)blockDim.x + (int)threadIdx.x;
            if (i < n * n) plane[i] = r[j];
        }
    }
    for (int i = threadIdx.x; i < 2 * (p1 - p0); i += blockDim.x) tl[i] = trig[2 * p0 + i];
    __syncthreads();
    const double c = (double)(n / 2);
    for (int i = threadIdx.x; i < (p1 - p0) * n; i += blockDim.x) {
        const int pl = i / n, v = i % n;
        const int p = p0 + pl;
        const double pxd = c + ((double)v - c) * tl[2 * pl], pyd = c + ((double)v - c) * tl[2 * pl + 1];
        const double fxd = floor(pxd), fyd = floor(pyd);
        const T fx = (T)fxd, fy = (T)fyd;
        const T ax = (T)(pxd - fxd), ay = (T)(pyd - fyd);
        E acc = Ops::zero();
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dxx = 0; dxx < 2; ++dxx) {
                const int xi = (int)fx + dxx, yi = (int)fy + dy;
                const T w = (dxx ? ax : (T)1 - ax) * (dy ? ay : (T)1 - ay);
                if (xi >= 0 && xi < n && yi >= 0 && yi < n && w != (T)0)
                    acc = Ops::fma(w, plane[yi * ld + xi], acc);
            }
        out[((int64_t)p * n + u) * n + v] = acc;
    }
}


template <typename E, typename T>
static void merge_launch(cpra_hip_ctx* h, const void* slices, void* vol, const void* angles, int n,
                         int P)
{
    const size_t per = (size_t)P * n * n;
    const size_t trig_bytes = ((sizeof(double) * 2 * (size_t)P + 255) / 256) * 256;
    h->ws_misc.ensure(trig_bytes + per * sizeof(E));
    double* trig = static_cast<double*>(h->ws_misc.ptr);
    E* sT = reinterpret_cast<E*>(static_cast<char*>(h->ws_misc.ptr) + trig_bytes);
    hipLaunchKernelGGL((k_trig_table<T>), dim3((P + 63) / 64), dim3(64), 0, h->stream,
                       static_cast<const T*>(angles), P, trig);
    check_launch("k_trig_table");
    hipLaunchKernelGGL((k_transpose_slices<E>), dim3((n + 31) / 32, (n + 31) / 32, P), dim3(256), 0,
                       h->stream, static_cast<const E*>(slices), sT, n);
    check_launch("k_transpose_slices");
    int XT = n < 64 ? n : 64;
    auto lds_of = [&](int xt) {
        return 4 * 256 * (sizeof(int) + sizeof(T)) + (size_t)n * xt * sizeof(E) + xt * sizeof(T);
    };
    while (XT > 1 && lds_of(XT) > 96 * 1024) XT /= 2;
    if (n % XT) fail(CPRA_HIP_ERR_UNSUPPORTED, "merge: n=%d not a multiple of the tile %d", n, XT);
    const size_t lds = lds_of(XT);
    hipLaunchKernelGGL((k_merge_gather<E, T>), dim3(n / XT, n), dim3(256), lds, h->stream, sT,
                       trig, P, n, XT, static_cast<E*>(vol));
    check_launch("k_merge_gather");
}

template <typename E, typename T>
static void extract_launch(cpra_hip_ctx* h, const void* vol, void* slices, const void* angles, int n,
                           int P, bool reuse_trig = false)
{
    const int64_t total = (int64_t)P * n * n;
    double* trig = static_cast<double*>(h->ws_trig.ptr);
    if (!reuse_trig) {
        h->mplan_n = 0;  // the table is rebuilt for these angles: drop the gather list
        h->ws_trig.ensure(sizeof(double) * 2 * (size_t)P);
        trig = static_cast<double*>(h->ws_trig.ptr);
        hipLaunchKernelGGL((k_trig_table<T>), dim3((P + 63) / 64), dim3(64), 0, h->stream,
                           static_cast<const T*>(angles), P, trig);
        check_launch("k_trig_table");
    }
    const size_t plane_bytes = sizeof(E) * (size_t)n * n;
    if (reuse_trig && plane_bytes <= kExtractPlaneLds) {
        // n planes x angle groups, about one block per CU: every group reads
        // its plane again, and the kernel is bound by that traffic (at 128^3 x
        // 180 angles: 2 groups 19 us, 4 groups 26 us); at most 256 angles per
        // group (the LDS table)
        int groups = 1;
        while ((n * groups < 256 && groups * 2 <= P) || (P + groups - 1) / groups > 256) groups *= 2;
        hipLaunchKernelGGL((k_extract_plane<E, T>), dim3(n * groups), dim3(1024),
                           kExtractTrigLds + plane_bytes, h->stream,
                           static_cast<const E*>(vol), trig, P, n, groups, static_cast<E*>(slices));
        check_launch("k_extract_plane");
        return;
    }
    hipLaunchKernelGGL((k_slice_extract<E, T>), dim3(grid_1d(total, 256, 1 << 16)), dim3(256), 0,
                       h->stream, static_cast<const E*>(vol), trig, P, n, static_cast<E*>(slices));
    check_launch("k_slice_extract");
}

static void check_merge_geometry(uint64_t m, uint64_t n, uint64_t p, uint64_t l, const char* who)
{
    CPRA_ARG(p >= 1, "%s: p must be >= 1", who);
    if (!(m == n && n == l))
        fail(CPRA_HIP_ERR_UNSUPPORTED, "%s: the carousel geometry needs a cube (m == n == l)", who);
    if (n < 2 || n > 64 * kMergeMaxZPerLane || (n & 1))
        fail(CPRA_HIP_ERR_UNSUPPORTED, "%s: n must be even and <= %d", who, 64 * kMergeMaxZPerLane);
}

// ---- merge with a precomputed gather list ----------------------------------
// The footprint of the slices on every voxel column depends on the angles
// only, so a run that merges many times (the carousel) lists it once: per
// column (x, y) the (p*n + v, w) entries in exactly the order k_merge_gather
// visits them, and 1/R^T(1).  The per-merge kernel is then a gather over
// that list (same summation order, same result bits).

// walks the candidates of column (x, y) in k_merge_gather's order; `emit`
// receives (list position, p*n + v, w) for each non-zero weight.  wl: this
// wave's 256-entry LDS scratch; wsum (uniform) is summed over it in list
// order per pass of 64 angles, as k_merge_gather does.
template <typename T, typename Emit>
__device__ __forceinline__ int footprint_walk(const double* __restrict__ trig, int P, int n, int x, int y,
                                              int lane, T* wl, T& wsum, Emit&& emit)
{
    const int c = n / 2;
    const double dx = (double)(x - c), dy = (double)(y - c);
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    int total = 0;
    for (int p0 = 0; p0 < P; p0 += 64) {
        const int p = p0 + lane;
        double cs = 0, sn = 0;
        bool live = p < P;
        if (live) {
            cs = trig[2 * p];
            sn = trig[2 * p + 1];
            live = fabs(-dx * sn + dy * cs) < 1.4142135623730951;
        }
        const int sb = (int)floor(dx * cs + dy * sn) - 1;
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int s = sb + k;
            const int v = s + n / 2;
            T w = 0;
            if (live && v >= 0 && v < n) {
                const double wx = 1.0 - fabs((double)s * cs - dx), wy = 1.0 - fabs((double)s * sn - dy);
                if (wx > 0.0 && wy > 0.0) w = (T)(wx * wy);
            }
            const bool has = w > (T)0;
            const uint64_t m = __ballot(has);
            if (has) {
                const int pos = cnt + __popcll(m & lt);
                wl[pos] = w;
                emit(total + pos, p * n + v, w);
            }
            cnt += __popcll(m);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int e = 0; e < cnt; ++e) wsum += wl[e];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        total += cnt;
    }
    return total;
}

template <typename T>
__global__ __launch_bounds__(256) void k_merge_count(const double* __restrict__ trig, int P, int n,
                                                     int* __restrict__ counts)
{
    __shared__ T wl[4][256];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = blockIdx.x * 4 + wave;
    if (col >= n * n) return;  // wave-uniform: no block barrier below
    T wsum = 0;
    const int cnt = footprint_walk<T>(trig, P, n, col % n, col / n, lane, wl[wave], wsum,
                                      [](int, int, T) {});
    if (lane == 0) counts[col] = cnt;
}

template <typename T>
__global__ __launch_bounds__(256) void k_merge_fill(const double* __restrict__ trig, int P, int n,
                                                    const int* __restrict__ offsets, int* __restrict__ lidx,
                                                    T* __restrict__ lw, T* __restrict__ inv)
{
    __shared__ T wl[4][256];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = blockIdx.x * 4 + wave;
    if (col >= n * n) return;
    const int base = offsets[col];
    T wsum = 0;
    footprint_walk<T>(trig, P, n, col % n, col / n, lane, wl[wave], wsum, [&](int pos, int pv, T w) {
        lidx[base + pos] = pv;
        lw[base + pos] = w;
    });
    if (lane == 0) inv[col] = wsum > (T)0 ? (T)1 / wsum : (T)0;
}

// Columns near the rotation axis are covered by every angle (at 180 x 128^2:
// up to 536 list entries against a mean of 5.5).  Columns with more than
// kMergeHeavy entries get a block of their own; the rest run one wave per
// column.
constexpr int kMergeHeavy = 32;
constexpr int kMergeXT = 16;     // light columns per block (one wave each)
constexpr int kHeavyWaves = 16;  // waves per heavy column
static_assert(kMergeXT == kHeavyWaves, "both block kinds use 1024 threads and n * 16 elements of LDS");

__device__ __forceinline__ float lane_bcast(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// acc[k] (lanes over z = lane + 64 k, KZ values per lane) += the cnt list
// entries e0, e0 + stride, ..., in list order.  Lane l loads entry l of each
// run of 64 (one vector load of indices and one of weights) and readlane
// hands them out, so the column loads of a round of R entries leave back to
// back: a list of c entries costs one index load and c / R column rounds,
// where loading the indices round by round put a dependent scalar load in
// front of every round.  The last round loads and adds only its live
// entries (a wave-uniform condition).  R: entries per round (R * KZ column
// values in flight per lane).
template <int KZ, int R, typename E, typename T>
__device__ __forceinline__ void merge_sum(const E* __restrict__ sT, const int* __restrict__ lidx,
                                          const T* __restrict__ lw, int e0, int cnt, int stride, int n,
                                          int lane, E (&acc)[KZ])
{
    using Ops = ElemOps<E>;
    for (int b0 = 0; b0 < cnt; b0 += 64) {
        const int m = min(64, cnt - b0);
        int li = 0;
        T lwv = (T)0;
        if (lane < m) {
            li = lidx[e0 + (b0 + lane) * stride];
            lwv = lw[e0 + (b0 + lane) * stride];
        }
        for (int j = 0; j < m; j += R) {
            const int live = m - j;
            const E* cp[R];
            T w[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int jr = min(j + r, m - 1);
                cp[r] = sT + (int64_t)__builtin_amdgcn_readlane(li, jr) * n;
                w[r] = lane_bcast(lwv, jr);
            }
#pragma unroll
            for (int k = 0; k < KZ; ++k) {
                const int z = lane + 64 * k;
                if (z < n) {
                    E v[R];
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (r < live) v[r] = cp[r][z];
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (r < live) acc[k] = Ops::fma(w[r], v[r], acc[k]);
                }
            }
        }
    }
}

// Entries per round of merge_sum.  Measured at 180 x 128^2 (KZ = 2), light
// and heavy columns alike: 8 per round 19.6-20.8 us per merge, 4 per round
// 20.4-20.7, 16 per round 27.5.
template <int KZ> constexpr int kMergeRound = KZ <= 2 ? 8 : 4;

// Light block: 16 consecutive x at one y, one wave per column, lanes over z;
// the columns are staged in LDS and stored z-major.  Heavy columns are left
// to their own blocks.
template <int KZ, typename E, typename T>
__device__ __forceinline__ void merge_light_block(unsigned t_id, const E* __restrict__ sT,
                                                  const int* __restrict__ offsets,
                                                  const int* __restrict__ lidx, const T* __restrict__ lw,
                                                  const T* __restrict__ inv, int n, E* __restrict__ vol,
                                                  E* tile, int* skip)
{
    using Ops = ElemOps<E>;
    constexpr int XT = kMergeXT;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tiles_x = n / XT;
    const int y = (int)t_id / tiles_x, x0 = ((int)t_id % tiles_x) * XT;
    {
        const int xl = wave;
        const int col = y * n + x0 + xl;
        const int e0 = offsets[col];
        const int e1 = offsets[col + 1];
        const bool heavy = e1 - e0 > kMergeHeavy;
        if (lane == 0) skip[xl] = heavy ? 1 : 0;
        E acc[KZ];
#pragma unroll
        for (int k = 0; k < KZ; ++k) acc[k] = Ops::zero();
        if (!heavy) merge_sum<KZ, kMergeRound<KZ>>(sT, lidx, lw, e0, e1 - e0, 1, n, lane, acc);
        const T s = inv[col];
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const int z = lane + 64 * k;
            if (z < n) tile[z * XT + xl] = Ops::scale(acc[k], s);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n * XT; i += blockDim.x) {
        const int z = i / XT, xl = i % XT;
        if (!skip[xl]) vol[((int64_t)z * n + y) * n + x0 + xl] = tile[i];
    }
}

// Heavy block: 16 waves on one column; wave w sums the entries e0 + w,
// e0 + w + 16, ..., then the 16 partial columns are added in LDS in a fixed
// order.
template <int KZ, typename E, typename T>
__device__ __forceinline__ void merge_heavy_block(int col, const E* __restrict__ sT,
                                                  const int* __restrict__ offsets,
                                                  const int* __restrict__ lidx, const T* __restrict__ lw,
                                                  const T* __restrict__ inv, int n, E* __restrict__ vol,
                                                  E* part)
{
    using Ops = ElemOps<E>;
    constexpr int NW = kHeavyWaves;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int e0 = offsets[col], e1 = offsets[col + 1];
    E acc[KZ];
#pragma unroll
    for (int k = 0; k < KZ; ++k) acc[k] = Ops::zero();
    const int cnt = e1 - e0 > wave ? (e1 - e0 - wave + NW - 1) / NW : 0;
    merge_sum<KZ, kMergeRound<KZ>>(sT, lidx, lw, e0 + wave, cnt, NW, n, lane, acc);
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
        const int z = lane + 64 * k;
        if (z < n) part[wave * n + z] = acc[k];
    }
    __syncthreads();
    const T s = inv[col];
    const int y = col / n, x = col % n;
    for (int z = threadIdx.x; z < n; z += blockDim.x) {
        E v = part[z];
#pragma unroll
        for (int w = 1; w < NW; ++w) v = Ops::fma((T)1, part[w * n + z], v);
        vol[((int64_t)z * n + y) * n + x] = Ops::scale(v, s);
    }
}

// One launch for the whole merge: the first heavy_blocks blocks (nheavy
// rounded up to a multiple of 8) take the heavy columns, longest lists
// first, so they overlap the light blocks; the light blocks follow, n/16 * n
// of them (a multiple of 8).  XCD-aware light order: block b runs on XCD
// b mod 8, and XCD x takes the rows of y in bands of `band` rows, the bands
// dealt to the XCDs in turn (rows [(8k + x) band, (8k + x + 1) band)), so
// voxel columns that read the same slice columns mostly share that XCD's L2.
// Every slice passes through the rotation axis, so the list lengths fall off
// as 1/r away from it: with one contiguous band of n/8 rows per XCD (band =
// n/8), the XCDs holding the middle rows had 1.5 times the mean work at
// 180 x 128^2; with 4-row bands every XCD is within 5 % of the mean.  Every
// voxel is written by exactly one block.  KZ: z values per lane, at least
// n / 64.
template <typename E, typename T, int KZ>
__global__ __launch_bounds__(1024) void k_merge_apply(const E* __restrict__ sT, const int* __restrict__ offsets,
                                                     const int* __restrict__ lidx, const T* __restrict__ lw,
                                                     const T* __restrict__ inv, const int* __restrict__ heavy,
                                                     int nheavy, int heavy_blocks, int n, int band,
                                                     E* __restrict__ vol)
{
    extern __shared__ unsigned char smem[];
    __shared__ int skip[kMergeXT];
    E* buf = reinterpret_cast<E*>(smem);  // [n][16] light tile or [16][n] heavy partials
    const int b = (int)blockIdx.x;
    if (b < heavy_blocks) {
        if (b < nheavy) merge_heavy_block<KZ, E, T>(heavy[b], sT, offsets, lidx, lw, inv, n, vol, buf);
        return;  // block-uniform
    }
    const unsigned lb = (unsigned)(b - heavy_blocks);
    const unsigned x = lb & 7u, i = lb >> 3;  // XCD, and its i-th tile
    const unsigned tx = (unsigned)n / kMergeXT;   // tiles per row
    const unsigned bt = (unsigned)band * tx;      // tiles per band
    const unsigned row = ((i / bt) * 8u + x) * (unsigned)band + (i % bt) / tx;
    merge_light_block<KZ, E, T>(row * tx + i % tx, sT, offsets, lidx, lw, inv, n, vol, buf, skip);
}

// z values per lane of k_merge_apply for size n (a power of two >= n / 64)
static int merge_kz(int n)
{
    const int need = (n + 63) / 64;
    return need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
}

template <int KZ>
static void merge_apply_launch(hipStream_t st, const c2<float>* sT, const int* offsets, const int* lidx,
                               const float* lw, const float* inv, const int* heavy, int nheavy, int heavy_blocks,
                               int n, void* vol)
{
    // 4-row bands where n / 8 is a multiple of 4, else one band per XCD
    const int band = (n / 8) % 4 == 0 ? 4 : n / 8;
    hipLaunchKernelGGL((k_merge_apply<c2<float>, float, KZ>), dim3(heavy_blocks + (n / 16) * n), dim3(1024),
                       sizeof(c2<float>) * (size_t)n * kMergeXT, st, sT, offsets, lidx, lw, inv, heavy, nheavy,
                       heavy_blocks, n, band, static_cast<c2<float>*>(vol));
}

// Builds the gather list of (angles, n, P) in the handle (synchronises once).
void merge_plan_c64(cpra_hip_ctx* h, const float* angles, int n, int P)
{
    if (h->capturing)
        fail(CPRA_HIP_ERR_ARG, "carousel merge: building the gather list needs a synchronisation, "
                               "which a stream capture cannot hold");
    check_merge_geometry(n, n, P, n, "carousel merge");
    if (n % 16) fail(CPRA_HIP_ERR_UNSUPPORTED, "carousel merge: n=%d not a multiple of 16", n);
    const int cols = n * n;
    h->ws_trig.ensure(sizeof(double) * 2 * (size_t)P);
    double* trig = static_cast<double*>(h->ws_trig.ptr);
    hipStream_t st = h->stream;
    hipLaunchKernelGGL((k_trig_table<float>), dim3((P + 63) / 64), dim3(64), 0, st, angles, P, trig);
    check_launch("k_trig_table");
    h->ws_mcount.ensure(sizeof(int) * ((size_t)cols + 1));
    int* counts = static_cast<int*>(h->ws_mcount.ptr);
    hipLaunchKernelGGL((k_merge_count<float>), dim3((cols + 3) / 4), dim3(256), 0, st, trig, P, n, counts);
    check_launch("k_merge_count");
    std::vector<int> off((size_t)cols + 1);
    CPRA_CHECK_HIP(hipMemcpyAsync(off.data(), counts, sizeof(int) * cols, hipMemcpyDeviceToHost, st));
    CPRA_CHECK_HIP(hipStreamSynchronize(st));
    int64_t run = 0;
    for (int i = 0; i < cols; ++i) {
        const int c = off[i];
        off[i] = (int)run;
        run += c;
    }
    off[cols] = (int)run;
    CPRA_ARG(run < (1ll << 31), "carousel merge: gather list too long");
    std::vector<int> heavy;
    for (int i = 0; i < cols; ++i)
        if (off[i + 1] - off[i] > kMergeHeavy) heavy.push_back(i);
    // longest lists first: the blocks that finish last start first
    std::stable_sort(heavy.begin(), heavy.end(),
                     [&](int a, int b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
    h->ws_mheavy.ensure(sizeof(int) * (heavy.empty() ? 1 : heavy.size()));
    if (!heavy.empty())
        CPRA_CHECK_HIP(hipMemcpyAsync(h->ws_mheavy.ptr, heavy.data(), sizeof(int) * heavy.size(),
                                      hipMemcpyHostToDevice, st));
    CPRA_CHECK_HIP(hipMemcpyAsync(counts, off.data(), sizeof(int) * ((size_t)cols + 1), hipMemcpyHostToDevice, st));
    const size_t ent = (size_t)(run > 0 ? run : 1);
    const size_t idx_bytes = ((sizeof(int) * ent + 255) / 256) * 256;
    const size_t w_bytes = ((sizeof(float) * ent + 255) / 256) * 256;
    h->ws_mlist.ensure(idx_bytes + w_bytes + sizeof(float) * (size_t)cols);
    char* base = static_cast<char*>(h->ws_mlist.ptr);
    int* lidx = reinterpret_cast<int*>(base);
    float* lw = reinterpret_cast<float*>(base + idx_bytes);
    float* inv = reinterpret_cast<float*>(base + idx_bytes + w_bytes);
    hipLaunchKernelGGL((k_merge_fill<float>), dim3((cols + 3) / 4), dim3(256), 0, st, trig, P, n, counts,
                       lidx, lw, inv);
    check_launch("k_merge_fill");
    CPRA_CHECK_HIP(hipStreamSynchronize(st));  // `off` and `heavy` are released on return
    // The merge's tile (n * 16 elements: 64 KiB at n = 512) beside its static
    // flags, and the extraction's trig table and plane (132 KiB at n = 128),
    // can exceed the default dynamic-LDS limit.  The attribute is a property
    // of the kernel for the whole process, so it is set to the largest size
    // the code launches, not to this handle's n: a handle of another size
    // must not lower it under one that is live.
    for (const void* f : {(const void*)k_merge_apply<c2<float>, float, 1>, (const void*)k_merge_apply<c2<float>, float, 2>,
                          (const void*)k_merge_apply<c2<float>, float, 4>, (const void*)k_merge_apply<c2<float>, float, 8>})
        CPRA_CHECK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(sizeof(c2<float>) * (size_t)kMergeMaxN * kMergeXT)));
    CPRA_CHECK_HIP(hipFuncSetAttribute((const void*)k_extract_plane<c2<float>, float>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(kExtractTrigLds + kExtractPlaneLds)));
    h->mplan_n = n;
    h->mplan_P = P;
    h->mplan_entries = run;
    h->mplan_heavy = (int)heavy.size();
}

// Workspace of the slices in the merge's layout [P][v][u] (line-major
// columns), valid until the next merge on the handle.
void* merge_lines_ws(cpra_hip_ctx* h, int n, int P)
{
    h->ws_misc.ensure((size_t)P * n * n * sizeof(c2<float>));
    return h->ws_misc.ptr;
}

// R^T(s)/R^T(1) through the list of merge_plan_c64, from slices already in the
// merge's layout sT[p][v][u].
void merge_apply_lines_c64(cpra_hip_ctx* h, const void* sT_, void* vol, int n, int P)
{
    CPRA_ARG(h->mplan_n == n && h->mplan_P == P, "carousel merge: no gather list for this geometry");
    using E = c2<float>;
    const E* sT = static_cast<const E*>(sT_);
    const size_t ent = (size_t)(h->mplan_entries > 0 ? h->mplan_entries : 1);
    const size_t idx_bytes = ((sizeof(int) * ent + 255) / 256) * 256;
    const size_t w_bytes = ((sizeof(float) * ent + 255) / 256) * 256;
    char* base = static_cast<char*>(h->ws_mlist.ptr);
    // n / 16 * n light tiles: a multiple of 8 for every n that is a multiple of 16
    const int nheavy = h->mplan_heavy;
    const int heavy_blocks = (nheavy + 7) / 8 * 8;
    const int* offsets = static_cast<const int*>(h->ws_mcount.ptr);
    const int* lidx = reinterpret_cast<const int*>(base);
    const float* lw = reinterpret_cast<const float*>(base + idx_bytes);
    const float* inv = reinterpret_cast<const float*>(base + idx_bytes + w_bytes);
    const int* heavy = static_cast<const int*>(h->ws_mheavy.ptr);
    switch (merge_kz(n)) {
    case 1: merge_apply_launch<1>(h->stream, sT, offsets, lidx, lw, inv, heavy, nheavy, heavy_blocks, n, vol); break;
    case 2: merge_apply_launch<2>(h->stream, sT, offsets, lidx, lw, inv, heavy, nheavy, heavy_blocks, n, vol); break;
    case 4: merge_apply_launch<4>(h->stream, sT, offsets, lidx, lw, inv, heavy, nheavy, heavy_blocks, n, vol); break;
    default: merge_apply_launch<8>(h->stream, sT, offsets, lidx, lw, inv, heavy, nheavy, heavy_blocks, n, vol); break;
    }
    check_launch("k_merge_apply");
}

// complex64 slices [P][u][v] into the merge's layout sT [P][v][u]
void merge_transpose_c64(cpra_hip_ctx* h, const void* slices, void* sT, int n, int P)
{
    using E = c2<float>;
    hipLaunchKernelGGL((k_transpose_slices<E>), dim3((n + 31) / 32, (n + 31) / 32, P), dim3(256), 0,
                       h->stream, static_cast<const E*>(slices), static_cast<E*>(sT), n);
    check_launch("k_transpose_slices");
}


// complex64 merge / extraction on the handle's stream (the carousel driver)
void merge_c64(cpra_hip_ctx* h, const void* slices, void* vol, const float* angles, int n, int P)
{
    check_merge_geometry(n, n, P, n, "carousel merge");
    merge_launch<c2<float>, float>(h, slices, vol, angles, n, P);
}

void extract_c64(cpra_hip_ctx* h, const void* vol, void* slices, const float* angles, int n, int P)
{
    check_merge_geometry(n, n, P, n, "carousel extraction");
    // the gather list's trig table of the same angles, when it is there
    const bool reuse = h->mplan_n == n && h->mplan_P == P && h->ws_trig.ptr;
    extract_launch<c2<float>, float>(h, vol, slices, angles, n, P, reuse);
}

template <typename F> static int mguard(cpra_hip_ctx* h, F&& f)
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

int cpra_hip_real2d_to_3d(cpra_hip_handle h, const void* slices, void* vol, const void* angles,
                          uint64_t m, uint64_t n, uint64_t p, uint64_t l)
{
    return mguard(h, [&]() -> int {
        CPRA_ARG(h && slices && vol && angles, "cpra_hip_real2d_to_3d: NULL argument");
        check_merge_geometry(m, n, p, l, "cpra_hip_real2d_to_3d");
        if (h->dtype == CPRA_HIP_F32) merge_launch<float, float>(h, slices, vol, angles, (int)n, (int)p);
        else merge_launch<double, double>(h, slices, vol, angles, (int)n, (int)p);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_complex2d_to_3d(cpra_hip_handle h, const void* slices, void* vol, const void* angles,
                             uint64_t m, uint64_t n, uint64_t p, uint64_t l)
{
    return mguard(h, [&]() -> int {
        CPRA_ARG(h && slices && vol && angles, "cpra_hip_complex2d_to_3d: NULL argument");
        check_merge_geometry(m, n, p, l, "cpra_hip_complex2d_to_3d");
        if (h->dtype == CPRA_HIP_F32) merge_launch<c2<float>, float>(h, slices, vol, angles, (int)n, (int)p);
        else merge_launch<c2<double>, double>(h, slices, vol, angles, (int)n, (int)p);
        return CPRA_HIP_OK;
    });
}

int cpra_hip_slice_extract(cpra_hip_handle h, const void* vol, void* slices, const void* angles,
                           uint64_t n, uint64_t p, int complex_data)
{
    return mguard(h, [&]() -> int {
        CPRA_ARG(h && slices && vol && angles, "cpra_hip_slice_extract: NULL argument");
        check_merge_geometry(n, n, p, n, "cpra_hip_slice_extract");
        if (h->dtype == CPRA_HIP_F32) {
            if (complex_data) extract_launch<c2<float>, float>(h, vol, slices, angles, (int)n, (int)p);
            else extract_launch<float, float>(h, vol, slices, angles, (int)n, (int)p);
        } else {
            if (complex_data) extract_launch<c2<double>, double>(h, vol, slices, angles, (int)n, (int)p);
            else extract_launch<double, double>(h, vol, slices, angles, (int)n, (int)p);
        }
        return CPRA_HIP_OK;
    });
}

}  // extern "C"
