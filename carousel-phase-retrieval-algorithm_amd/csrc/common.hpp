// common.hpp -- handle state, error plumbing and launch helpers shared by the
// translation units of libcpra_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cpra_hip.h"
#include "../../include/cpra_hip_engine.h"
#include "fft_core.hpp"

namespace cpra {

// Error carrying a C-ABI status; thrown inside the library, converted to a
// status code at the extern "C" boundary.
struct Status {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Status{code, buf};
}

#define CPRA_CHECK_HIP(expr)                                                                 \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            ::cpra::fail(e_ == hipErrorOutOfMemory ? CPRA_HIP_ERR_NOMEM : CPRA_HIP_ERR_HIP, \
                         "%s failed: %s", #expr, hipGetErrorString(e_));                     \
    } while (0)

#define CPRA_ARG(cond, ...)                                   \
    do {                                                      \
        if (!(cond)) ::cpra::fail(CPRA_HIP_ERR_ARG, __VA_ARGS__); \
    } while (0)

// Device workspace owned by a handle; grows on demand, freed at destroy.
struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t want)
    {
        if (want <= bytes) return;
        if (ptr) CPRA_CHECK_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
        CPRA_CHECK_HIP(hipMalloc(&ptr, want));
        bytes = want;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

struct EngineGraph;  // engine2d.hip

// An engine's iteration graph kept by its handle across calls.  Each call
// captures its body again (host-side recording) and re-points the kept
// executable at it with hipGraphExecUpdate; only a changed topology (another
// launch sequence) is instantiated anew.  `done` marks the last replay
// queued, which an update waits for, so that no replay in flight sees its
// arguments change.
// `key`, when an engine keeps one, lists everything its captured body
// depends on (buffers, parameters, workspaces, tuning); a call with the same
// key replays the kept graph without capturing again.
struct KeptGraph {
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    std::vector<uint64_t> key;
    void release()
    {
        if (done) (void)hipEventSynchronize(done);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (done) (void)hipEventDestroy(done);
        exec = nullptr;
        done = nullptr;
        key.clear();
    }
};
enum { KG_CAROUSEL = 0, KG_MIXED, KG_DM, KG_DM3D, KG_COUNT };

}  // namespace cpra

struct cpra_hip_ctx {
    int dtype = CPRA_HIP_F32;
    uint64_t m = 1, n = 1, l = 1, batch = 1;
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipStream_t aux_stream = nullptr;  // graph capture / replay
    std::string last_error;
    // twiddle tables keyed by (n << 1) | is_double
    std::map<uint64_t, void*> twiddles;
    // engine workspaces
    cpra::Buffer ws_z, ws_blur, ws_keys, ws_part, ws_part2, ws_misc, ws_backup, ws_rows, ws_dm, ws_trig, ws_mcount, ws_mlist, ws_mheavy;
    cpra::Buffer ws_zsup;  // the carousel's 3D support packed one bit per voxel
    // tile-major HIO state and amplitudes (engine2d.hip)
    cpra::Buffer ws_state, ws_amp_t;
    // one sub-batch's support while the HIO run goes sub-batch by sub-batch
    cpra::Buffer ws_sup;
    // the HIO support packed one bit per pixel (k_pack_support)
    cpra::Buffer ws_sbits;
    // gather list of the carousel merge (merge3d.hip)
    int mplan_n = 0, mplan_P = 0, mplan_heavy = 0;
    int64_t mplan_entries = 0;
    // tuning choice per (kind, n) set by the autotuner (0 = default)
    std::map<uint64_t, int> tuning;
    // captured HIO chunk bodies: [0] dense rows, [1] sparse rows
    cpra::EngineGraph* graph[2] = {nullptr, nullptr};
    // the other engines' iteration graphs, kept across calls (KG_*)
    cpra::KeptGraph kept[cpra::KG_COUNT];
    // sparse rows for engine_hio_iteration (decided by engine_hio_prepare)
    bool hio_sparse = false;
    // the support rows have been listed (the adaptive mode's flag is valid)
    bool rows_listed = false;
    // graphs captured from the handle's stream by the caller
    // (cpra_hip_graph_begin / _end); engines run eagerly while capturing
    bool capturing = false;
    uint64_t next_graph = 1;
    std::map<uint64_t, std::pair<hipGraph_t, hipGraphExec_t>> user_graphs;
    // per-kernel event timing of the engines (eager launches only)
    bool timing = false;
    // iteration graphs the engines have captured on this handle (reported by
    // cpra_hip_get_tuning, so that a caller can tell a replay from a capture)
    int64_t graph_captures = 0;
    double kt_ms[16] = {};
    int64_t kt_n[16] = {};
};

namespace cpra {

size_t dtype_size(int dtype);
// device twiddle table exp(-2 pi i k / n), k < n, in the handle's precision
// (or the requested one), cached per handle.
const void* twiddle_table(cpra_hip_ctx* h, int n, bool dbl);

inline unsigned grid_1d(uint64_t work, unsigned block, unsigned cap = 1u << 20)
{
    uint64_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// The executable of `kg` updated to (or instantiated from) the captured
// graph g, which the caller still owns and destroys.  engine2d.hip.
hipGraphExec_t kept_graph_exec(cpra_hip_ctx* h, KeptGraph& kg, hipGraph_t g);
// Records kg.done on st after the replays the caller has queued.
void kept_graph_launched(KeptGraph& kg, hipStream_t st);
// Appends to `key` the handle state a captured engine body can depend on:
// every workspace address, the tuning choices, the merge's gather-list
// geometry and the HIO row mode.
void engine_state_key(const cpra_hip_ctx* h, std::vector<uint64_t>& key);

// Stream on which an engine captures and replays its iteration graph: the
// handle's own stream when it has one, so that the replays queue behind the
// run's other work with no cross-stream events (replays on a second stream
// measured 2 % slower on the carousel); a non-blocking stream of the handle
// when the handle runs on the null stream, which cannot be captured.
hipStream_t graph_stream(cpra_hip_ctx* h);

// CPRA_HIP_DEBUG=1 in the environment: name every launch on stderr and
// synchronise after it, so a failing kernel is identified by name.
bool debug_launches();

inline void check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(CPRA_HIP_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    if (debug_launches()) {
        std::fprintf(stderr, "[cpra_hip] %s\n", what);
        std::fflush(stderr);
        e = hipDeviceSynchronize();
        if (e != hipSuccess) fail(CPRA_HIP_ERR_HIP, "%s failed at run time: %s", what, hipGetErrorString(e));
    }
}

// FFT pass entry points (fft.hip).  `lines` transforms of length n; contig:
// line j starts at j*n; strided: line j starts at (j / G)*GS + (j % G), element
// stride es.  shift: fold the centring shift S into the index maps.
void fft_pass_contig(cpra_hip_ctx* h, void* data, int n, int64_t lines, int dir, bool shift,
                     double scale, bool dbl);
void fft_pass_strided(cpra_hip_ctx* h, void* data, int n, int64_t lines, int64_t es, int64_t G,
                      int64_t GS, int dir, bool shift, double scale, bool dbl);
// Forward centred column transform of complex64 lines (as fft_pass_strided
// with dir -1, shift on, no scale) written line-major to out[line * n + u].
// Returns false, launching nothing, for sizes it has no kernel for (n other
// than 64, 128, 192, 256, 384, 512, or lines / G not whole tiles).
bool fft_fast_size(int n);
// Event pair around one engine kernel when the handle's timing is on; the
// elapsed time is read back after the run (kernel_timer_collect).
struct KernelTimer {
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    void begin(cpra_hip_ctx* h, hipStream_t st, int kind)
    {
        if (!h->timing) return;
        Rec r{kind, nullptr, nullptr};
        CPRA_CHECK_HIP(hipEventCreate(&r.a));
        CPRA_CHECK_HIP(hipEventCreate(&r.b));
        CPRA_CHECK_HIP(hipEventRecord(r.a, st));
        recs.push_back(r);
    }
    void end(cpra_hip_ctx* h, hipStream_t st)
    {
        if (!h->timing) return;
        CPRA_CHECK_HIP(hipEventRecord(recs.back().b, st));
    }
    void collect(cpra_hip_ctx* h)
    {
        for (auto& r : recs) {
            float ms = 0.f;
            CPRA_CHECK_HIP(hipEventSynchronize(r.b));
            CPRA_CHECK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            h->kt_ms[r.kind] += ms;
            h->kt_n[r.kind] += 1;
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        recs.clear();
    }
};

// kinds of cpra_hip_get_kernel_times
enum KernelKind {
    KT_COL = 0, KT_ROW = 1, KT_SW = 2, KT_GD = 3, KT_MERGE = 4, KT_VOL = 5, KT_EXTRACT = 6, KT_PTRANS = 7,
    KT_MIXED = 8, KT_PENCIL = 9
};

// engine2d.hip: tile-major layout of square patterns (Tiled<N> in
// engine_kernels.hpp): whether the handle uses it, the row pass between the
// row-major x and the tiled state (dir -1: x -> DFT_r into the tiled state;
// +1: back, scaled 1/n), and the tiled copy of amplitudes
bool engine_tiled_state(const cpra_hip_ctx* h);
// whole-plane x/y transforms in the carousel's volume constraint (tuning plane=)
bool engine_plane_fft(const cpra_hip_ctx* h);
bool engine_merge_lines(const cpra_hip_ctx* h);
int engine_plane_groups(const cpra_hip_ctx* h);
void engine_retile(hipStream_t st, int n, int64_t patterns, int dir, const void* in, void* out);
void engine_amp_retile(hipStream_t st, int n, int64_t patterns, const float* in, float* out);
// engine2d.hip: one HIO iteration on the row-spectral state (see there)
void engine_hio_prepare(cpra_hip_ctx* h, hipStream_t st, const uint8_t* support);
bool engine_hio_iteration(cpra_hip_ctx* h, hipStream_t st, void* Y, const float* amp,
                          int64_t amp_batch, const uint8_t* support, const cpra_hip_hio_params* p,
                          KernelTimer* kt, void* lines_out = nullptr);
// merge3d.hip: complex64 gather merge R^T(s)/R^T(1) and re-projection R
void merge_c64(cpra_hip_ctx* h, const void* slices, void* vol, const float* angles, int n, int P);
void extract_c64(cpra_hip_ctx* h, const void* vol, void* slices, const float* angles, int n, int P);
// merge through a gather list built once per geometry (synchronises)
void merge_plan_c64(cpra_hip_ctx* h, const float* angles, int n, int P);
void* merge_lines_ws(cpra_hip_ctx* h, int n, int P);
void merge_transpose_c64(cpra_hip_ctx* h, const void* slices, void* sT, int n, int P);
void merge_apply_lines_c64(cpra_hip_ctx* h, const void* sT, void* vol, int n, int P);

// frees the captured engine graphs of a handle (engine2d.hip)
void cpra_engine_release(cpra_hip_ctx* h);

}  // namespace cpra
