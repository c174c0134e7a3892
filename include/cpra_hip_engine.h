/*
 * cpra_hip_engine.h -- fused iteration engines of libcpra_hip.so.
 *
 * These entry points have no single counterpart in the reference's
 * CpraImpl<T> interface: the reference spells its iterations out as
 * sequences of primitive calls in the drivers (src/CPRA_MKL.cpp:37-55,
 * 94-110, 126-140; src/benchmark_mkl.cpp:45-63).  The engines run the same
 * mathematics (or the additions beyond the reference: HIO, shrinkwrap,
 * R-factor) as a few fused HBM passes per iteration.  FP32 (complex64) handles only; the
 * pattern size {m, n} of the handle must be in {64,128,192,256,384,512}.
 */
#ifndef CPRA_HIP_ENGINE_H
#define CPRA_HIP_ENGINE_H
#include "cpra_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* HIO (Fienup 1982) with a shrinkwrap schedule. */
typedef struct {
    int32_t iters;            /* HIO iterations                                      */
    float beta;               /* feedback parameter                                  */
    int32_t real_positive;    /* 1: inside the support keep Re x' >= 0, drop Im      */
    int32_t dc_skip_zero;     /* 1: A == 0 leaves the Fourier pixel unconstrained    */
                              /*    (MklImpl::DataConstraint, MKL_Impl.inl:95-101)   */
    int32_t error_reduction;  /* 1: error reduction instead of HIO (outside -> 0)    */
    int32_t sw_every;         /* shrinkwrap after every sw_every iterations (0: off) */
    float sw_sigma0;          /* sigma of the first update (pixels)                  */
    float sw_sigma1;          /* sigma of the last update (linear in between)        */
    float sw_frac;            /* threshold fraction of the per-pattern blur maximum  */
    int32_t use_graph;        /* capture sw_every iterations in a hipGraph           */
} cpra_hip_hio_params;

void cpra_hip_hio_default_params(cpra_hip_hio_params* p);

/* Runs p->iters HIO iterations on x [batch][m][n] (complex64, in place) with
 * amplitudes amp [amp_batch][m][n] (float; pattern b uses amp[b % amp_batch])
 * and the u8 support [batch][m][n] (updated in place by shrinkwrap).
 * After iteration i (1-based) with i % sw_every == 0, update k = i/sw_every-1
 * of K = iters/sw_every uses sigma_k = sigma0 + (sigma1 - sigma0) k/(K-1).
 * If rfactor_host != NULL it receives, per pattern, the R-factor
 * sum_{A>0} | |Forward2D(x)| - A | / sum_{A>0} A of the final x.
 * The call is asynchronous unless rfactor_host is given; the row mode
 * (tuning rows_sparse, adaptive by default) is decided on the device. */
int cpra_hip_hio_run(cpra_hip_handle h, void* x, const float* amp, uint64_t amp_batch,
                     uint8_t* support, const cpra_hip_hio_params* p, float* rfactor_host);

/* Difference map.  The reference variant (textbook = 0) is the iteration
 * body of src/CPRA_MKL.cpp:94-110 with t1, t2 carried across iterations:
 *   t1 <- P_S(-psi/beta + (1+1/beta) P_M t1)
 *   t2 <- P_M(psi/beta + (1-1/beta) P_S t2)
 *   psi <- psi + beta t1 - beta t2
 * textbook = 1 is Elser's map (t1 = t2 = psi at every iteration).  P_M is
 * Forward2D, DataConstraint (dc_skip_zero as in MklImpl), Backward2D; P_S uses
 * the u8 support: SPACE_MKL = the reference SpaceConstraint on the mask
 * equivalent of its float constraint (cpra_hip_support_from_constraint),
 * SPACE_POSITIVE additionally clamps Re < 0, SPACE_COMPLEX keeps complex
 * values inside the support.  psi (and t1, t2 for the reference variant) are
 * [batch][m][n] complex64 in pattern space, updated in place.  The handle's
 * l is ignored (the reference's drivers give their 2D handles the volume's
 * depth, src/CPRA_MKL.cpp:24). */
#define CPRA_HIP_DM_SPACE_MKL 0
#define CPRA_HIP_DM_SPACE_POSITIVE 1
#define CPRA_HIP_DM_SPACE_COMPLEX 2
typedef struct {
    int32_t iters;
    float beta;
    int32_t textbook;
    int32_t space_mode;
    int32_t dc_skip_zero;
    int32_t use_graph;
} cpra_hip_dm_params;

void cpra_hip_dm_default_params(cpra_hip_dm_params* p);
int cpra_hip_dm_run(cpra_hip_handle h, void* psi, void* t1, void* t2, const float* amp,
                    uint64_t amp_batch, const uint8_t* support, const cpra_hip_dm_params* p);
/* cpra_hip_dm_run without any allocation, for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  No counterpart in the reference.  Same
 * arguments and results (bit for bit) as cpra_hip_dm_run; the iteration
 * state [Y1; Y2; Z1; Z2], 32 bytes per pixel of the batch, lives in a
 * static device array of CPRA_HIP_DM_NOALLOC_MAX_BYTES that every GPU the
 * library runs on holds until the process ends, so batches up to
 * CPRA_HIP_DM_NOALLOC_MAX_BYTES / 32 pixels (e.g. 8 patterns of 512^2 or 32
 * of 256^2); larger ones are refused with CPRA_HIP_ERR_ARG.  Launches are
 * eager (p->use_graph is ignored) and nothing is allocated, created or
 * synchronised, so a caller may capture the call.  The handle's device must
 * be current.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other threads, other handles) race and
 * must be serialised by the caller. */
#define CPRA_HIP_DM_NOALLOC_MAX_BYTES (64u << 20)
int cpra_hip_dm_run_noalloc(cpra_hip_handle h, void* psi, void* t1, void* t2, const float* amp,
                            uint64_t amp_batch, const uint8_t* support, const cpra_hip_dm_params* p);
/* The same map on volumes [batch][l][m][n] of the handle's 3D geometry (l, m,
 * n in {64,128,192,256,384,512}), with Forward3D/Backward3D in P_M: the
 * iteration body of the reference's conventional reconstruction,
 * src/CONV_MKL.cpp:47-63.  amp [amp_batch][l][m][n], support u8
 * [batch][l][m][n].  Four HBM passes per iteration: z lines with the modulus
 * projection, y lines, x rows with the update, y lines. */
int cpra_hip_dm3d_run(cpra_hip_handle h, void* psi, void* t1, void* t2, const float* amp,
                      uint64_t amp_batch, const uint8_t* support, const cpra_hip_dm_params* p);

/* Shrinkwrap of x [batch][m][n]: b = Re Backward2D(Forward2D(|x|) * G_sigma),
 * G_sigma(kr, kc) = exp(-2 pi^2 sigma^2 [((kr-m/2)/m)^2 + ((kc-n/2)/n)^2]);
 * support = (b >= frac * max_pattern(b)).  blur_out (device, [batch][m][n]
 * float) receives b when not NULL. */
int cpra_hip_shrinkwrap(cpra_hip_handle h, const void* x, uint8_t* support, float sigma,
                        float frac, float* blur_out);

/* R-factor per pattern of x (see cpra_hip_hio_run); r_host has batch floats. */
int cpra_hip_rfactor(cpra_hip_handle h, const void* x, const float* amp, uint64_t amp_batch,
                     float* r_host);

/* u8 support equivalent to the reference SpaceConstraint's float mask:
 * mask[i] = !((double)S[i] < 0.9)  (MKL_Impl.inl:80). */
int cpra_hip_support_from_constraint(cpra_hip_handle h, const float* S, uint8_t* mask,
                                     uint64_t num);

/* Autotuner (no reference counterpart): times each column-kernel variant
 * (thread mapping: 16-column blocks, per-wave columns or 32-column blocks of
 * 512 threads, x XCD-aware tile order) for `iters` eager HIO iterations on
 * the caller's data, then on whole iterations the state layout, the row
 * kernel's radix plan where the size has two (256, 512: 16 or 8 values per
 * thread), dense against sparse row passes (rows without a support pixel
 * finished in the column kernel) and the patterns per launch (whole batch or
 * Infinity-Cache-sized chunks) and the run in 2 or 4 sub-batches (split);
 * keeps the fastest for this handle, restores x and the support, and writes
 * "col_<m>=<v>;row_plan_<n>=<v>;rows_sparse=<0|1>;chunk=<c>;split=<s>;
 * tiled=<0|1>;...;col_v<k>_us=..;iter_*_us=..;" into report (may
 * be NULL).  With a tuning cache set (cpra_hip_set_tuning_cache), a pick
 * stored for the same device architecture, pattern size, batch and amplitude
 * batch is applied instead of timing ("...;cached=1;" in the report), and a
 * new pick is stored. */
int cpra_hip_autotune_2d(cpra_hip_handle h, void* x, const float* amp, uint64_t amp_batch,
                         uint8_t* support, int iters, char* report, uint64_t report_len);
/* Re-applies the "col_<m>=<v>;", "rows_sparse=<0|1|2>;", "chunk=<patterns>;"
 * and "tiled=<0|1>;" choices of an autotune report (rows_sparse: dense rows,
 * sparse rows, or adaptive, the default: sparse while at most 80 % of the rows
 * hold a support pixel, decided on the device after every shrinkwrap update;
 * chunk: patterns
 * per launch of the iteration kernels, 0 = whole batch; tiled: tile-major
 * state for square patterns), "split=<s>;" (cpra_hip_hio_run takes the batch
 * through the whole schedule in s sub-batches of batch/s patterns, one after
 * another, with the same results; applies when s divides the batch and the
 * amplitudes are per pattern or one for all, outside a capture; 1 = whole
 * batch), "plane=<0|1>;" (n <= 128: the carousel's 3D constraint and
 * re-projection as one kernel per z-plane in LDS; 0 = x passes around a
 * y-line constraint kernel, and the re-projection on its own), "merge_lines=<0|1>;"
 * (the carousel's last HIO row pass of an outer iteration stores the
 * projections straight into the merge's [P][v][u] layout where it can;
 * 0 = row-major and a transpose before the merge), "plane_groups=<g>;"
 * (angle groups per plane of the carousel's plane kernel, 1, 2, 4 or 8;
 * 0 = enough to fill the GPU, the default) and "row_plan_<n>=<0|1>;" (the
 * HIO row kernel's radix plan for width n; 1 exists for n = 256 and 512). */
int cpra_hip_set_tuning(cpra_hip_handle h, const char* spec);
/* The autotuner's persistent cache for the whole process: a JSON object
 * mapping "<gcn arch>/hio2d/<m>x<n>/b<batch>/a<amp_batch>" to the tuning
 * spec cpra_hip_set_tuning takes.  path NULL or "" turns it off (the
 * default); refresh != 0 times every workload again and overwrites its
 * entry. */
int cpra_hip_set_tuning_cache(const char* path, int refresh);
/* Current choices as "col_<m>=<v>;rows_sparse=<0|1|2>;chunk=<patterns>;
 * split=<s>;tiled=<0|1>;plane=<0|1>;merge_lines=<0|1>;plane_groups=<g>;row_plan_<n>=<0|1>;
 * rows_sparse_used=<0|1>;
 * graph_captures=<k>;" (rows_sparse_used: the row mode of the last iterations
 * run; in the adaptive mode it is read from the device, which synchronises
 * the handle's stream; graph_captures: how many iteration graphs the engines
 * have captured on this handle, so that a caller can check that a run
 * replayed a graph it had already captured). */
int cpra_hip_get_tuning(cpra_hip_handle h, char* out, uint64_t out_len);

/* Per-kernel timing of the engines: when enabled, engine runs launch
 * eagerly (no graph) with a hipEvent pair around every fused kernel on the
 * stream it runs on; cpra_hip_get_kernel_times returns, per kernel kind
 * (0 = column pass, 1 = row pass, 2 = one shrinkwrap update, 3 = fused
 * gradient-descent pass, 4 = 2D->3D merge, 5 = 3D volume constraint (its
 * five FFT passes, three kernels for n <= 128), 6 = slice extraction, 7 = column transforms of the
 * carousel's projections outside its HIO passes (the carousel's kind 0 includes its column pass that
 * starts from the re-projected spectra), 8 = one whole mixed-batch iteration of the per-class and
 * per-pattern modes; the fused mode records 0 and 1; 9 = the y-line pencil
 * passes of the 3D difference map), the summed milliseconds
 * and the launch count since the last enable (n <= 16 kinds). */
int cpra_hip_set_kernel_timing(cpra_hip_handle h, int enable);
int cpra_hip_get_kernel_times(cpra_hip_handle h, double* ms_total, int64_t* counts, int n);

/* ---- amplitude loss and its adjoint (the reference has no loss or
 * gradient) -----------------------------------------------------------------
 * L_b = sum_k (|F x_b|_k - A_k)^2 with F the centred, unnormalised DFT of the
 * handle's 2D (m x n) or 3D (l x m x n) geometry, A [amp_batch][...] real,
 * broadcast over the batch by b % amp_batch.  The gradient follows the
 * convention dL/d(re x) + i dL/d(im x):
 *   grad = 2 F^H (F x - A (F x)/|F x|),  residual 0 where F x = 0,
 * computed from x at every call (nothing of F x is kept between calls, also
 * for 3D volumes).  loss_host: batch doubles (may be NULL for the gradient).
 * F32 and F64 handles; m, n (and l) even. grad must not alias x. */
int cpra_hip_amp_loss(cpra_hip_handle h, const void* x, const void* amp, uint64_t amp_batch,
                      double* loss_host);
int cpra_hip_amp_grad(cpra_hip_handle h, const void* x, const void* amp, uint64_t amp_batch,
                      void* grad, double* loss_host);
int cpra_hip_amp_loss_3d(cpra_hip_handle h, const void* x, const void* amp, uint64_t amp_batch,
                         double* loss_host);
int cpra_hip_amp_grad_3d(cpra_hip_handle h, const void* x, const void* amp, uint64_t amp_batch,
                         void* grad, double* loss_host);

/* `iters` steps of gradient descent x <- x - mu * grad on the amplitude loss,
 * in place (F32 handles).  One fused pass per step along the outer axis (m in
 * 2D, l in 3D, which must be 64/128/192/256/384/512) over the state kept as
 * the DFT over the inner axes.  loss_host (may be NULL) receives the loss of
 * the final iterate, batch doubles. */
int cpra_hip_amp_gd_run(cpra_hip_handle h, void* x, const float* amp, uint64_t amp_batch,
                        int iters, float mu, double* loss_host);
int cpra_hip_amp_gd_run_3d(cpra_hip_handle h, void* x, const float* amp, uint64_t amp_batch,
                           int iters, float mu, double* loss_host);

/* ---- mixed-size batch (no reference counterpart) ----------------------------
 * Difference map (textbook map, p->textbook must be 1; beta, space_mode,
 * dc_skip_zero, iters, use_graph as in cpra_hip_dm_run) over size classes of
 * square patterns.  Class k: count patterns of n x n (n in
 * {64,128,192,256,384,512}), psi [count][n][n] complex64 in/out, amp and the
 * u8 support [count][n][n] (one amplitude per pattern).  launch_mode:
 *   CPRA_HIP_MIXED_FUSED        one column and one row launch per iteration
 *                               for the whole batch (blocks partitioned by
 *                               class, LDS sized for the largest class);
 *   CPRA_HIP_MIXED_PER_CLASS    one launch pair per class per iteration;
 *   CPRA_HIP_MIXED_PER_PATTERN  one launch triple per pattern per iteration.
 * The modes give identical results.  h: any F32 handle (its stream and
 * workspaces are used). */
#define CPRA_HIP_MAX_CLASSES 8
#define CPRA_HIP_MIXED_FUSED 0
#define CPRA_HIP_MIXED_PER_CLASS 1
#define CPRA_HIP_MIXED_PER_PATTERN 2
typedef struct {
    uint32_t n;              /* pattern side                        */
    uint32_t count;          /* patterns of this size               */
    void* psi;               /* [count][n][n] complex64, in/out     */
    const float* amp;        /* [count][n][n]                       */
    const uint8_t* support;  /* [count][n][n]                       */
} cpra_hip_size_class;

int cpra_hip_dm_mixed_run(cpra_hip_handle h, const cpra_hip_size_class* classes, int nclasses,
                          const cpra_hip_dm_params* p, int launch_mode);

/* ---- 3D carousel (no reference counterpart) ---------------------------------
 * Projections of one real object at angles th_p about the z axis, their
 * diffraction amplitudes measured.  One outer iteration: inner_iters HIO
 * iterations on every projection (real-positive inside support2d); F_p =
 * Forward2D(x_p); V = gather-trilinear merge of the F_p as central slices of
 * the object's centred 3D DFT (cpra_hip_complex2d_to_3d geometry: slice rows
 * along z, columns along (cos th, sin th)); v = Backward3D(V), v <- (Re v)^+
 * inside support3d and 0 elsewhere; V = Forward3D(v); F_p = slices of V
 * (cpra_hip_slice_extract); x_p = Backward2D(F_p).  All projections advance
 * together (the Jacobi form of the reference's sequential sweep,
 * src/CPRA_MKL.cpp:58-124).
 * h: F32 handle with m == n in {64,128,192,256,384,512} and batch = P.
 * x [P][n][n] complex64 (in/out), amp [amp_batch][n][n], support2d u8
 * [P][n][n], angles [P] float radians, support3d u8 [n][n][n] (device
 * pointers).  vol [n][n][n] complex64 receives the constrained real-space
 * object of the last iteration (the run also uses it as its volume workspace,
 * so its prior contents are lost).  rfactor_host (NULL or P floats): R-factor
 * of the final projections. */
typedef struct {
    int32_t outer_iters;   /* outer iterations                                  */
    int32_t inner_iters;   /* HIO iterations on the projections per outer one   */
    float beta;            /* HIO feedback                                      */
    int32_t dc_skip_zero;  /* 1: A == 0 leaves the Fourier pixel unconstrained  */
    int32_t use_graph;     /* capture ten outer iterations per hipGraph, replay */
} cpra_hip_carousel_params;

void cpra_hip_carousel_default_params(cpra_hip_carousel_params* p);
int cpra_hip_carousel3d_run(cpra_hip_handle h, void* x, const float* amp, uint64_t amp_batch,
                            const uint8_t* support2d, const float* angles, const uint8_t* support3d,
                            void* vol, const cpra_hip_carousel_params* p, float* rfactor_host);

#ifdef __cplusplus
}
#endif
#endif
