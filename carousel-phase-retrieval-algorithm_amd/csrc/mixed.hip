// mixed.hip -- one difference-map iteration over patterns of several sizes.
//
// A batch holds size classes of square patterns (n in {64..512}).  The fused
// launch mode runs each iteration as two kernel launches for the whole batch:
// one column pass and one row pass whose blocks are partitioned by class;
// block b finds its class from the class's first block (uniform per block)
// and runs the class's size-specialised body (eng_cols_tile /
// dm_rows_block, the same bodies as the single-size engine) on dynamic LDS
// sized for the largest class.  The per-class mode launches each class on its
// own; the per-pattern mode launches every pattern on its own (the naive
// baseline).  All three modes compute the same values.
//
// The map is Elser's difference map (textbook, beta), the iteration body of
// cpra_hip_dm_run with textbook = 1 in its two-pass form (dmt_cols_tile,
// dmt_rows_block): [Y1; Y2] per class in the row-spectral domain, column
// pass [Y1; Y2] -> [V; W] (one block per column tile of a pattern for both
// halves), row pass -> Y1, Y2 (psi after the last iteration).
#include <vector>

#include "common.hpp"
#include "engine_kernels.hpp"

namespace cpra {

constexpr int kMaxClasses = CPRA_HIP_MAX_CLASSES;

struct MixedClass {
    c2<float>* Y;  // [Y1; Y2]: 2 * count patterns
    c2<float>* Z;  // [V; W]
    c2<float>* psi;
    const float* A;
    const uint8_t* mask;
    int n, count;
    unsigned cb0, rb0;  // first block of the class in the column / row grid
    float scale;        // 1 / n^2
};

struct MixedArgs {
    MixedClass c[kMaxClasses];
    int ncls;
    int skip_zero;
    float beta;
};

__device__ __forceinline__ int mixed_class(const MixedArgs& a, unsigned b, bool rows)
{
    int k = 0;
    for (int i = 1; i < a.ncls; ++i)
        if ((rows ? a.c[i].rb0 : a.c[i].cb0) <= b) k = i;
    return k;
}

// 512 threads per block (DmtColCfg), 128 VGPRs: two blocks, four waves per
// SIMD.
__global__ __launch_bounds__(512) void k_mixed_cols(const MixedArgs a)
{
    extern __shared__ unsigned char smem[];
    c2<float>* lds = reinterpret_cast<c2<float>*>(smem);
    const int k = mixed_class(a, blockIdx.x, false);
    const MixedClass& c = a.c[k];
    const unsigned tile = blockIdx.x - c.cb0;
#define CPRA_MC(NN)                                                                                    \
    dmt_cols_tile<NN>(tile, lds, c.Y, c.Z, c.A, c.count, NN, c.scale, a.skip_zero, a.beta, \
                                 (size_t)c.count * NN * NN)
    switch (c.n) {
    case 64: CPRA_MC(64); break;
    case 128: CPRA_MC(128); break;
    case 192: CPRA_MC(192); break;
    case 256: CPRA_MC(256); break;
    case 384: CPRA_MC(384); break;
    case 512: CPRA_MC(512); break;
    }
#undef CPRA_MC
}

// Three waves per SIMD (146-157 VGPRs, the 384-point body's): held to four
// waves (128 VGPRs) it spills about 64 bytes per lane, and the batch of
// 128^2 to 384^2 runs 848 instead of 925 it/s.
template <int SPACE, bool LAST>
__global__ __launch_bounds__(256) void k_mixed_dm_rows(const MixedArgs a)
{
    extern __shared__ unsigned char smem[];
    c2<float>* lds = reinterpret_cast<c2<float>*>(smem);
    const int k = mixed_class(a, blockIdx.x, true);
    const MixedClass& c = a.c[k];
    const unsigned vb = blockIdx.x - c.rb0;
    const size_t half = (size_t)c.count * c.n * c.n;
#define CPRA_MR(NN) \
    dmt_rows_block<NN, SPACE, LAST>(vb, lds, c.Z, c.psi, c.mask, c.Y, (int64_t)c.count * NN, a.beta, half)
    switch (c.n) {
    case 64: CPRA_MR(64); break;
    case 128: CPRA_MR(128); break;
    case 192: CPRA_MR(192); break;
    case 256: CPRA_MR(256); break;
    case 384: CPRA_MR(384); break;
    case 512: CPRA_MR(512); break;
    }
#undef CPRA_MR
}

// per-size geometry of the bodies
struct SizeGeo {
    int C, LPB;
    size_t col_lds, row_lds;
};

template <int N> static SizeGeo geo_n()
{
    SizeGeo g;
    g.C = DmtColCfg<N>::C;
    g.LPB = DmtRowCfg<N>::LPB;
    g.col_lds = sizeof(c2<float>) * (size_t)N * DmtColCfg<N>::C;
    g.row_lds = sizeof(c2<float>) * (size_t)DmtRowCfg<N>::LPB * LdsRow<c2<float>, N>::span();
    return g;
}

static SizeGeo size_geo(int n)
{
    switch (n) {
    case 64: return geo_n<64>();
    case 128: return geo_n<128>();
    case 192: return geo_n<192>();
    case 256: return geo_n<256>();
    case 384: return geo_n<384>();
    case 512: return geo_n<512>();
    }
    return SizeGeo{0, 0, 0, 0};
}

template <int N, int SPACE>
static void entry_n(hipStream_t st, const MixedClass& c, float beta)
{
    const int64_t lines = (int64_t)c.count * N;
    const size_t half = (size_t)c.count * N * N;
    constexpr int LPB = DmRowCfg<N>::LPB;
    hipLaunchKernelGGL((k_dm_entry<N, true, SPACE>), dim3((unsigned)((lines + LPB - 1) / LPB)),
                       dim3(DmRowCfg<N>::TPB), 0, st, c.psi, nullptr, nullptr, c.mask, c.Y, c.Y + half, lines, beta);
}

template <int N, int SPACE, bool LAST>
static void rows_n(hipStream_t st, const MixedClass& c, int j0, int cnt, float beta)
{
    // patterns [j0, j0 + cnt) of the class
    const size_t pp = (size_t)N * N, half = (size_t)c.count * pp;
    const int64_t lines = (int64_t)cnt * N;
    constexpr int LPB = DmtRowCfg<N>::LPB;
    hipLaunchKernelGGL((k_dmt_rows<N, SPACE, LAST>), dim3((unsigned)((lines + LPB - 1) / LPB)),
                       dim3(DmtRowCfg<N>::TPB), 0, st, c.Z + j0 * pp, c.psi + j0 * pp, c.mask + j0 * pp,
                       c.Y + j0 * pp, lines, beta, (int64_t)half);
}

// column pass over patterns [j0, j0 + cnt) of the class (both halves)
template <int N>
static void cols_n(hipStream_t st, const MixedClass& c, int j0, int cnt, int skip, float beta)
{
    const size_t pp = (size_t)N * N;
    const unsigned grid = (unsigned)((int64_t)cnt * N / DmtColCfg<N>::C);
    const size_t lds = sizeof(c2<float>) * (size_t)N * DmtColCfg<N>::C;
    hipLaunchKernelGGL((k_dmt_cols<N>), dim3(grid), dim3(DmtColCfg<N>::TPB), lds, st, c.Y + j0 * pp,
                       c.Z + j0 * pp, c.A + j0 * pp, (int64_t)cnt, N, c.scale, skip, beta,
                       (int64_t)c.count * pp);
}

#define CPRA_NSWITCH(NV, CALL)                         \
    switch (NV) {                                      \
    case 64: { constexpr int NN = 64; CALL; } break;   \
    case 128: { constexpr int NN = 128; CALL; } break; \
    case 192: { constexpr int NN = 192; CALL; } break; \
    case 256: { constexpr int NN = 256; CALL; } break; \
    case 384: { constexpr int NN = 384; CALL; } break; \
    case 512: { constexpr int NN = 512; CALL; } break; \
    }

struct MixedRun {
    cpra_hip_ctx* h;
    MixedArgs a;
    int space;
    int mode;
    unsigned col_grid = 0, row_grid = 0;
    size_t col_lds = 0, row_lds = 0;

    void entry(hipStream_t st)
    {
        for (int k = 0; k < a.ncls; ++k) {
            const MixedClass& c = a.c[k];
            if (space == SPACE_MKL) CPRA_NSWITCH(c.n, (entry_n<NN, SPACE_MKL>(st, c, a.beta)))
            else if (space == SPACE_POS) CPRA_NSWITCH(c.n, (entry_n<NN, SPACE_POS>(st, c, a.beta)))
            else CPRA_NSWITCH(c.n, (entry_n<NN, SPACE_CPLX>(st, c, a.beta)))
        }
        check_launch("k_dm_entry (mixed)");
    }

    template <bool LAST> void rows_fused(hipStream_t st)
    {
        if (space == SPACE_MKL)
            hipLaunchKernelGGL((k_mixed_dm_rows<SPACE_MKL, LAST>), dim3(row_grid), dim3(256), row_lds, st, a);
        else if (space == SPACE_POS)
            hipLaunchKernelGGL((k_mixed_dm_rows<SPACE_POS, LAST>), dim3(row_grid), dim3(256), row_lds, st, a);
        else
            hipLaunchKernelGGL((k_mixed_dm_rows<SPACE_CPLX, LAST>), dim3(row_grid), dim3(256), row_lds, st, a);
    }

    template <bool LAST> void rows_split(hipStream_t st, const MixedClass& c, int j0, int cnt)
    {
        if (space == SPACE_MKL) CPRA_NSWITCH(c.n, (rows_n<NN, SPACE_MKL, LAST>(st, c, j0, cnt, a.beta)))
        else if (space == SPACE_POS) CPRA_NSWITCH(c.n, (rows_n<NN, SPACE_POS, LAST>(st, c, j0, cnt, a.beta)))
        else CPRA_NSWITCH(c.n, (rows_n<NN, SPACE_CPLX, LAST>(st, c, j0, cnt, a.beta)))
    }

    void iteration(hipStream_t st, bool last, KernelTimer* kt)
    {
        if (mode == CPRA_HIP_MIXED_FUSED) {
            if (kt) kt->begin(h, st, KT_COL);
            hipLaunchKernelGGL(k_mixed_cols, dim3(col_grid), dim3(512), col_lds, st, a);
            check_launch("k_mixed_cols");
            if (kt) kt->end(h, st);
            if (kt) kt->begin(h, st, KT_ROW);
            if (last) rows_fused<true>(st);
            else rows_fused<false>(st);
            check_launch("k_mixed_dm_rows");
            if (kt) kt->end(h, st);
            return;
        }
        if (kt) kt->begin(h, st, KT_MIXED);
        for (int k = 0; k < a.ncls; ++k) {
            const MixedClass& c = a.c[k];
            if (mode == CPRA_HIP_MIXED_PER_CLASS) {
                CPRA_NSWITCH(c.n, (cols_n<NN>(st, c, 0, c.count, a.skip_zero, a.beta)))
                if (last) rows_split<true>(st, c, 0, c.count);
                else rows_split<false>(st, c, 0, c.count);
            } else {
                for (int j = 0; j < c.count; ++j) {
                    CPRA_NSWITCH(c.n, (cols_n<NN>(st, c, j, 1, a.skip_zero, a.beta)))
                    if (last) rows_split<true>(st, c, j, 1);
                    else rows_split<false>(st, c, j, 1);
                }
            }
        }
        check_launch("mixed iteration (split launches)");
        if (kt) kt->end(h, st);
    }
};

template <typename F> static int xguard(cpra_hip_ctx* h, F&& f)
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

int cpra_hip_dm_mixed_run(cpra_hip_handle h, const cpra_hip_size_class* classes, int nclasses,
                          const cpra_hip_dm_params* p, int launch_mode)
{
    return xguard(h, [&]() -> int {
        CPRA_ARG(h && classes && p, "cpra_hip_dm_mixed_run: NULL argument");
        CPRA_ARG(nclasses >= 1 && nclasses <= kMaxClasses, "cpra_hip_dm_mixed_run: 1..%d size classes", kMaxClasses);
        CPRA_ARG(p->iters >= 0, "cpra_hip_dm_mixed_run: iters < 0");
        CPRA_ARG(p->beta != 0.f, "cpra_hip_dm_mixed_run: beta == 0");
        CPRA_ARG(p->space_mode >= 0 && p->space_mode <= 2, "cpra_hip_dm_mixed_run: bad space_mode");
        CPRA_ARG(launch_mode >= CPRA_HIP_MIXED_FUSED && launch_mode <= CPRA_HIP_MIXED_PER_PATTERN,
                 "cpra_hip_dm_mixed_run: bad launch mode");
        if (h->dtype != CPRA_HIP_F32) fail(CPRA_HIP_ERR_UNSUPPORTED, "cpra_hip_dm_mixed_run: FP32 handles only");
        if (!p->textbook)
            fail(CPRA_HIP_ERR_UNSUPPORTED, "cpra_hip_dm_mixed_run: the mixed batch runs the textbook map");
        MixedRun r;
        r.h = h;
        r.space = p->space_mode;
        r.mode = launch_mode;
        r.a.ncls = nclasses;
        r.a.skip_zero = p->dc_skip_zero;
        r.a.beta = p->beta;
        size_t total = 0;
        for (int k = 0; k < nclasses; ++k) {
            const cpra_hip_size_class& c = classes[k];
            CPRA_ARG(c.psi && c.amp && c.support, "cpra_hip_dm_mixed_run: NULL pointer in class %d", k);
            CPRA_ARG(c.count >= 1, "cpra_hip_dm_mixed_run: empty class %d", k);
            if (!size_geo((int)c.n).C)
                fail(CPRA_HIP_ERR_UNSUPPORTED, "cpra_hip_dm_mixed_run: class %d size %u not in {64..512}", k, c.n);
            total += 2 * (size_t)c.count * c.n * c.n;
        }
        if (p->iters == 0) return CPRA_HIP_OK;
        h->ws_dm.ensure(sizeof(c2<float>) * total);
        h->ws_z.ensure(sizeof(c2<float>) * total);
        size_t off = 0;
        uint64_t cb = 0, rb = 0;
        for (int k = 0; k < nclasses; ++k) {
            const cpra_hip_size_class& in = classes[k];
            MixedClass& c = r.a.c[k];
            const SizeGeo sg = size_geo((int)in.n);
            c.Y = static_cast<c2<float>*>(h->ws_dm.ptr) + off;
            c.Z = static_cast<c2<float>*>(h->ws_z.ptr) + off;
            c.psi = static_cast<c2<float>*>(in.psi);
            c.A = in.amp;
            c.mask = in.support;
            c.n = (int)in.n;
            c.count = (int)in.count;
            c.scale = (float)(1.0 / ((double)in.n * in.n));
            c.cb0 = (unsigned)cb;
            c.rb0 = (unsigned)rb;
            cb += (uint64_t)in.count * in.n / sg.C;
            rb += ((uint64_t)in.count * in.n + sg.LPB - 1) / sg.LPB;
            r.col_lds = sg.col_lds > r.col_lds ? sg.col_lds : r.col_lds;
            r.row_lds = sg.row_lds > r.row_lds ? sg.row_lds : r.row_lds;
            off += 2 * (size_t)in.count * in.n * in.n;
        }
        CPRA_ARG(cb < (1ull << 31) && rb < (1ull << 31), "cpra_hip_dm_mixed_run: batch too large");
        r.col_grid = (unsigned)cb;
        r.row_grid = (unsigned)rb;
        hipStream_t st = h->stream;
        r.entry(st);
        const int body = p->iters - 1;
        KernelTimer kt;
        if (p->use_graph && !h->timing && !h->capturing && !debug_launches() && body >= 2) {
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
            CPRA_CHECK_HIP(hipStreamBeginCapture(gst, hipStreamCaptureModeThreadLocal));
            r.iteration(gst, false, nullptr);
            CPRA_CHECK_HIP(hipStreamEndCapture(gst, &graph));
            KeptGraph& kg = h->kept[KG_MIXED];
            try {
                exec = kept_graph_exec(h, kg, graph);
            } catch (...) {
                (void)hipGraphDestroy(graph);
                throw;
            }
            (void)hipGraphDestroy(graph);
            graph = nullptr;
            for (int i = 0; i < body; ++i) CPRA_CHECK_HIP(hipGraphLaunch(exec, gst));
            kept_graph_launched(kg, gst);
            if (ev0) {
                CPRA_CHECK_HIP(hipEventRecord(ev1, gst));
                CPRA_CHECK_HIP(hipStreamWaitEvent(st, ev1, 0));
                CPRA_CHECK_HIP(hipEventSynchronize(ev1));
                (void)hipEventDestroy(ev0);
                (void)hipEventDestroy(ev1);
            }
        } else {
            for (int i = 0; i < body; ++i) r.iteration(st, false, h->timing ? &kt : nullptr);
        }
        r.iteration(st, true, h->timing ? &kt : nullptr);
        kt.collect(h);
        return CPRA_HIP_OK;
    });
}

}  // extern "C"
