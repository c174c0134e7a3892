/*
 * cpra_hip.h -- C ABI of libcpra_hip.so, the MI355X (gfx950) backend of the
 * Carousel Phase Retrieval Algorithm.
 *
 * Plain pointers and sizes only.  Every entry point validates its arguments
 * before any HIP call, so argument errors are testable without a GPU.
 *
 * Return values: CPRA_HIP_OK (0) = the reference's `true`; CPRA_HIP_FALSE (1)
 * = the reference's `false` for a well-formed no-op (num == 0, alpha == 0,
 * norm == 0, as in include/detail/MKL_Impl.inl:11,34,55,73,89,112);
 * negative = error class (see below), with a message in
 * cpra_hip_last_error().
 *
 * Data layout follows the reference (include/CPRA_Impl.hpp, Appendix of the
 * reference's MKL backend): complex<T> interleaved (re, im); 2D batches are
 * row-major [batch][m][n]; 3D volumes are [batch][l][m][n] with n fastest.
 * T is float (CPRA_HIP_F32) or double (CPRA_HIP_F64), fixed per handle like
 * the reference's CpraImpl<T>.  Device pointers are caller-owned (allocate
 * with cpra_hip_malloc or any HIP allocator); the handle owns its stream,
 * twiddle tables and engine workspaces.
 */
#ifndef CPRA_HIP_H
#define CPRA_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPRA_HIP_OK 0
#define CPRA_HIP_FALSE 1
#define CPRA_HIP_ERR_ARG (-1)         /* invalid argument                       */
#define CPRA_HIP_ERR_UNSUPPORTED (-2) /* well-formed but unsupported geometry    */
#define CPRA_HIP_ERR_NOMEM (-3)       /* device allocation failed                */
#define CPRA_HIP_ERR_HIP (-4)         /* HIP runtime / launch error              */
#define CPRA_HIP_ERR_NODEV (-5)       /* no usable GPU                           */

#define CPRA_HIP_F32 0
#define CPRA_HIP_F64 1

typedef struct cpra_hip_ctx* cpra_hip_handle;

/* ---- lifetime, errors, streams, memory ------------------------------------ */

/* Replaces NewCpraImpl<T>(type, m, n, l, batch_size) / the CudaImpl
 * constructor (include/Impl_Factory.hpp:13-31, include/CUDA_Impl.cuh:16-52):
 * fixes the 2D plan geometry {m, n} x batch and the 3D geometry {l, m, n} x
 * batch.  m, n, l, batch >= 1. */
int cpra_hip_create(cpra_hip_handle* out, int dtype, uint64_t m, uint64_t n, uint64_t l,
                    uint64_t batch);
/* Replaces ~CudaImpl (include/CUDA_Impl.cuh:105-112). */
int cpra_hip_destroy(cpra_hip_handle h);
/* Message of the last error on this handle (h == NULL: last error of the
 * calling thread, e.g. a failed create).  Never NULL. */
const char* cpra_hip_last_error(cpra_hip_handle h);
/* Library version string. */
const char* cpra_hip_version(void);
/* Number of visible GPUs (0 when none); no handle needed. */
int cpra_hip_device_count(void);

/* Use `stream` (a hipStream_t; NULL = the null stream) for all later work of
 * this handle instead of the handle's own stream. */
int cpra_hip_set_stream(cpra_hip_handle h, void* stream);
/* The stream the handle currently launches on. */
void* cpra_hip_get_stream(cpra_hip_handle h);

/* Replaces CudaAllocator / CudaManagedAllocator::allocate / deallocate
 * (include/Allocator.hpp:18-52). managed != 0 -> hipMallocManaged.  h may be
 * NULL (errors then go to the thread's last error). */
int cpra_hip_malloc(cpra_hip_handle h, void** ptr, uint64_t bytes, int managed);
int cpra_hip_free(cpra_hip_handle h, void* ptr);
/* Replaces CudaImpl::Memcpy (include/CUDA_Impl.cuh:93-98): asynchronous,
 * direction inferred (hipMemcpyDefault), on the handle's stream; with
 * h == NULL a synchronous copy on the null stream. */
int cpra_hip_memcpy(cpra_hip_handle h, void* dst, const void* src, uint64_t bytes);
/* Replaces CudaImpl::Sync (include/CUDA_Impl.cuh:99-103). */
int cpra_hip_sync(cpra_hip_handle h);
/* Graph capture of the handle's stream (no reference counterpart;
 * graph_capture / graph_launch): every call made on the handle between
 * cpra_hip_graph_begin and cpra_hip_graph_end is recorded, not run, and the
 * instantiated graph replays it with cpra_hip_graph_launch on the handle's
 * stream.  The engines launch eagerly inside a capture (no nested graphs);
 * calls that need a host synchronisation (cpra_hip_sync, ConvergeError, the
 * R-factor copy-out, the autotuner, the carousel's gather-list build) are
 * refused while a capture is open.  Pointers and sizes are fixed at capture;
 * workspaces should be sized by one eager run first.  Replaces nothing in the
 * reference, whose backends launch every call eagerly. */
int cpra_hip_graph_begin(cpra_hip_handle h);
int cpra_hip_graph_end(cpra_hip_handle h, uint64_t* graph_id);
int cpra_hip_graph_launch(cpra_hip_handle h, uint64_t graph_id);
int cpra_hip_graph_destroy(cpra_hip_handle h, uint64_t graph_id);

/* Replaces CpraImpl::Initialize (include/CPRA_Impl.hpp:23; CUDA
 * include/detail/CUDA_Impl.inl:8-18): fills num values of T with uniform
 * [0, 1) numbers from a counter-based generator keyed by `seed` (the
 * reference seeds cuRAND with time(NULL) and MKL with a default
 * thread_local mt19937, so neither is reproducible across backends). */
int cpra_hip_initialize(cpra_hip_handle h, void* data, uint64_t num, uint64_t seed);

/* ---- FFT and shift (CpraImpl: include/CPRA_Impl.hpp:25-35) ----------------- */

/* Centred batched 2D DFT in place on [batch][m][n]:
 * X = S(DFT2(S(x))) (MklImpl::Forward2D, include/detail/MKL_Impl.inl:128-139).
 * Backward: x = S((1/(m n)) IDFT2(S(X))) (MKL_Impl.inl:141-152). m, n even. */
int cpra_hip_forward2d(cpra_hip_handle h, void* x);
int cpra_hip_backward2d(cpra_hip_handle h, void* x);
/* Centred batched 3D DFT on [batch][l][m][n] (MKL_Impl.inl:154-178 with the
 * intended {l, m, n} geometry); backward scaled by 1/(l m n).  l, m, n even. */
int cpra_hip_forward3d(cpra_hip_handle h, void* x);
int cpra_hip_backward3d(cpra_hip_handle h, void* x);
/* Quadrant swap of FFTShift2D (MKL_Impl.inl:180-209), bit-exact.  M, N even. */
int cpra_hip_fftshift2d(cpra_hip_handle h, void* x, uint64_t M, uint64_t N, uint64_t batch);
/* Octant swap of FFTShift3D (MKL_Impl.inl:212-251), bit-exact. Even sizes. */
int cpra_hip_fftshift3d(cpra_hip_handle h, void* x, uint64_t M, uint64_t N, uint64_t L,
                        uint64_t batch);

/* ---- pointwise constraints (include/CPRA_Impl.hpp:37-48) ------------------- */

#define CPRA_HIP_SPACE_MKL 0      /* im <- 0; re <- 0 where (double)S < 0.9      */
#define CPRA_HIP_SPACE_POSITIVE 1 /* im <- 0; re <- (re < 0 ? 0 : re * S)        */

/* SpaceConstraint: S has num/batch values broadcast as S[i % (num/batch)].
 * CPRA_HIP_SPACE_MKL follows MKL_Impl.inl:70-84, CPRA_HIP_SPACE_POSITIVE the
 * CUDA kernel (include/detail/CUDA_Impl_Kernel.cuh:16-28). */
int cpra_hip_space_constraint(cpra_hip_handle h, void* x, const void* S, uint64_t num,
                              uint64_t batch, int mode);
/* DataConstraint(real): x <- A * x/|x| where |x| > 0.  skip_zero != 0 leaves
 * pixels with A == 0 untouched (MKL_Impl.inl:86-106); skip_zero == 0 is the
 * CUDA kernel (CUDA_Impl_Kernel.cuh:30-38). */
int cpra_hip_data_constraint_real(cpra_hip_handle h, void* x, const void* A, uint64_t num,
                                  uint64_t batch, int skip_zero);
/* DataConstraint(complex): x <- C where |C|^2 != 0 (MKL_Impl.inl:108-126). */
int cpra_hip_data_constraint_complex(cpra_hip_handle h, void* x, const void* C,
                                     uint64_t num, uint64_t batch);
/* MergeAddData: dst <- alpha*src + beta*dst over num complex values, src may
 * alias dst; alpha == 0 -> CPRA_HIP_FALSE and no change (MKL_Impl.inl:31-50). */
int cpra_hip_merge_add(cpra_hip_handle h, const void* src, void* dst, double alpha,
                       double beta, uint64_t num);
/* Normalization: x <- x * (T)(1/norm) (MKL_Impl.inl:52-68). */
int cpra_hip_normalize(cpra_hip_handle h, void* x, double norm, uint64_t num);

/* ConvergeError (include/CPRA_Impl.hpp:57; CUDA_Impl.inl:145-172):
 * old <- new - old; host_err[b] = ||old_b|| / ||new_b|| over the m*n*l
 * values of batch member b (the reference's CUDA path sums m*n only).
 * host_err holds `batch` entries of T in host-visible memory (host or
 * managed). Synchronises. */
int cpra_hip_converge_error(cpra_hip_handle h, void* old_, const void* new_, void* host_err,
                            uint64_t m, uint64_t n, uint64_t l, uint64_t batch);

/* ---- carousel 2D -> 3D merge (include/CPRA_Impl.hpp:54-55) --------------- */

/* Gather-trilinear merge of p slices [p][m][n] (rows u along the rotation
 * axis, columns v) into the volume [l][m][n] = [z][y][x] of a cube
 * (m == n == l, even, <= 512): slice p at angle angles[p] (device array of T,
 * radians) samples the point x = c + (v-c) cos th, y = c + (v-c) sin th,
 * z = u with c = n/2; the volume is R^T(slices) / R^T(1) where R^T(1) > 0 and
 * 0 elsewhere, R being trilinear sampling.  Replaces Real2DTo3DInterpolation
 * (CUDA include/detail/CUDA_Impl.inl:174-197, a racy nearest-voxel scatter;
 * a stub in the MKL backend, include/MKL_Impl.hpp:151). */
int cpra_hip_real2d_to_3d(cpra_hip_handle h, const void* slices, void* vol, const void* angles,
                          uint64_t m, uint64_t n, uint64_t p, uint64_t l);
/* The same for complex<T> slices and volume (a stub returning true in both
 * reference backends, include/MKL_Impl.hpp:152, include/CUDA_Impl.cuh:89). */
int cpra_hip_complex2d_to_3d(cpra_hip_handle h, const void* slices, void* vol, const void* angles,
                             uint64_t m, uint64_t n, uint64_t p, uint64_t l);
/* Re-projection R: slices[p][u][v] = trilinear sample of the cube volume
 * [n][n][n] at the slice points above (0 outside).  complex_data selects
 * complex<T> or T elements.  No reference counterpart. */
int cpra_hip_slice_extract(cpra_hip_handle h, const void* vol, void* slices, const void* angles,
                           uint64_t n, uint64_t p, int complex_data);

#ifdef __cplusplus
}
#endif
#endif
