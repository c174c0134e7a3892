/*
 * cpra_oracle.h -- CPU oracle for the CPRA hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is used by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and never by the product path (the HIP library does not link or load it).
 *
 * It is a plain-C restatement of the reference's MKL backend semantics
 * (include/detail/MKL_Impl.inl) and of the reference driver compositions
 * (src/CPRA_MKL.cpp, src/CONV_MKL.cpp), plus the additions that
 * the reference does not contain (HIO, shrinkwrap, R-factor, slice
 * merge/extraction, amplitude-loss gradient), all in FP64 unless a function
 * name ends in _f32 (those mirror the reference's single-precision pointwise
 * kernels so that masks and permutations can be compared bit for bit).
 *
 * Complex arrays are interleaved (re, im) like std::complex<T>.
 * Every function returns 1 for "true" and 0 for "false", mirroring the
 * reference's bool returns, unless documented otherwise.
 *
 * Parity pins: see the header comment of each function and tests/test_oracle.py.
 */
#ifndef CPRA_ORACLE_H
#define CPRA_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FFT (FP64) ---------------------------------------------------------- */

/* Batched 1D DFT along lines: element k of line j is at data[j*dist + k*stride]
 * (complex units).  sign = -1 forward (e^{-2 pi i nk/N}), +1 inverse
 * (unnormalised); the result is multiplied by `scale`.  Mixed-radix Stockham
 * with radices 4, 2, 3, 5 and a direct DFT for any remaining prime factor. */
int orc_fft_lines(double* data, int64_t n, int64_t stride, int64_t dist,
                  int64_t count, int sign, double scale);

/* Un-centred, unnormalised forward / 1/(MN)-scaled backward 2D DFT over a
 * batch of row-major [B][M][N] arrays (the DFTI descriptor of
 * include/MKL_Impl.hpp:19-45: forward scale 1, backward scale 1/(m*n)). */
int orc_dft2(double* x, int64_t M, int64_t N, int64_t B, int sign);
/* Same for [B][L][M][N] volumes, backward scale 1/(L*M*N). */
int orc_dft3(double* x, int64_t M, int64_t N, int64_t L, int64_t B, int sign);

/* Centred transforms, literally shift -> DFT -> shift like
 * MklImpl::Forward2D / Backward2D (include/detail/MKL_Impl.inl:128-152) and
 * Forward3D / Backward3D (MKL_Impl.inl:154-178). */
int orc_forward2d(double* x, int64_t M, int64_t N, int64_t B);
int orc_backward2d(double* x, int64_t M, int64_t N, int64_t B);
int orc_forward3d(double* x, int64_t M, int64_t N, int64_t L, int64_t B);
int orc_backward3d(double* x, int64_t M, int64_t N, int64_t L, int64_t B);

/* Quadrant / octant swaps of MklImpl::FFTShift2D (MKL_Impl.inl:180-209) and
 * FFTShift3D (MKL_Impl.inl:212-251) on elements of `elem` bytes. */
int orc_fftshift2d(void* x, size_t elem, int64_t M, int64_t N, int64_t B);
int orc_fftshift3d(void* x, size_t elem, int64_t M, int64_t N, int64_t L, int64_t B);

/* ---- pointwise ops, reference single-precision semantics ------------------ */

/* MklImpl::SpaceConstraint (MKL_Impl.inl:70-84): im <- 0; re <- 0 where
 * (double)S[i % (num/batch)] < 0.9. */
int orc_space_constraint_f32(float* x, const float* S, uint64_t num, uint64_t batch);
/* CUDA ker_SpaceConstraint (CUDA_Impl_Kernel.cuh:16-28): im <- 0;
 * re <- 0 if re < 0 else re * S. */
int orc_space_constraint_pos_f32(float* x, const float* S, uint64_t num, uint64_t batch);
/* MklImpl::DataConstraint(real) (MKL_Impl.inl:86-106) when skip_zero != 0,
 * CUDA ker_DataConstraint (CUDA_Impl_Kernel.cuh:30-38) when skip_zero == 0.
 * Computed in double and rounded to float. */
int orc_data_constraint_f32(float* x, const float* A, uint64_t num, uint64_t batch,
                            int skip_zero);
/* MklImpl::DataConstraint(complex) (MKL_Impl.inl:108-126). */
int orc_data_constraint_cplx_f32(float* x, const float* C, uint64_t num, uint64_t batch);
/* MklImpl::MergeAddData (MKL_Impl.inl:31-50): dst <- alpha*src + beta*dst;
 * alpha == 0 or num == 0 -> no-op, returns 0. */
int orc_merge_add_f32(const float* src, float* dst, float alpha, float beta, uint64_t num);
/* MklImpl::Normalization (MKL_Impl.inl:52-68): x <- x * (T)(1/norm). */
int orc_normalize_f32(float* x, float norm, uint64_t num);

/* FP64 versions of the same (used by the algorithm-level functions). */
int orc_space_constraint_f64(double* x, const double* S, uint64_t num, uint64_t batch);
int orc_data_constraint_f64(double* x, const double* A, uint64_t num, uint64_t batch,
                            int skip_zero);
int orc_merge_add_f64(const double* src, double* dst, double alpha, double beta, uint64_t num);

/* ---- compositions -------------------------------------------------------- */

/* Reference difference-map variant, src/CPRA_MKL.cpp:94-110 (also
 * src/benchmark_mkl.cpp:47-62): t1, t2 carried across iterations.
 * A, S are [amp_batch][M][N] and broadcast over the batch with index
 * b % amp_batch (the reference passes batch_size so that i % (num/batch)
 * broadcasts one pattern over all B starts).  FP64 throughout. */
int orc_dm_ref(double* psi, double* t1, double* t2, const double* A, const double* S,
               int64_t M, int64_t N, int64_t B, int64_t amp_batch, double beta,
               int64_t iters);

/* Final projection, src/CPRA_MKL.cpp:126-140:
 * psi <- P_S(-psi/beta + (1+1/beta) P_M psi). */
int orc_final_projection(double* psi, const double* A, const double* S, int64_t M,
                         int64_t N, int64_t B, int64_t amp_batch, double beta);
/* src/CONV_MKL.cpp:47-63: the same difference-map composition on L x M x N
 * volumes (Forward3D/Backward3D), t1 and t2 carried; and the final
 * projection of src/CONV_MKL.cpp:78-88. */
int orc_dm_ref3d(double* psi, double* t1, double* t2, const double* A, const double* S,
                 int64_t M, int64_t N, int64_t L, int64_t B, int64_t amp_batch, double beta,
                 int64_t iters);
int orc_final_projection3d(double* psi, const double* A, const double* S, int64_t M,
                           int64_t N, int64_t L, int64_t B, int64_t amp_batch, double beta);

/* Textbook difference map (Elser 2003) with gamma_S = -1/beta,
 * gamma_M = 1/beta:  psi <- psi + beta*[P_S((1+1/beta)P_M psi - psi/beta)
 *                                     - P_M((1-1/beta)P_S psi + psi/beta)].
 * space_mode: 0 = reference SpaceConstraint on float S (S < 0.9 -> 0),
 *             1 = real-positive on a u8 mask, 2 = complex u8 mask. */
int orc_dm_textbook(double* psi, const double* A, const double* S, const uint8_t* mask,
                    int64_t M, int64_t N, int64_t B, int64_t amp_batch, double beta,
                    int space_mode, int64_t iters);

/* HIO parameters (Fienup 1982) with the shrinkwrap schedule. */
typedef struct {
    int64_t iters;       /* number of HIO iterations                          */
    double beta;         /* feedback parameter                                */
    int real_positive;   /* 1: object real and non-negative inside support     */
    int dc_skip_zero;    /* 1: A == 0 leaves the Fourier pixel unconstrained   */
    int64_t sw_every;    /* shrinkwrap period (0 = no shrinkwrap)             */
    double sw_sigma0;    /* sigma at the first shrinkwrap update (pixels)     */
    double sw_sigma1;    /* sigma at the last update (linear in between)      */
    double sw_frac;      /* threshold = sw_frac * max(blurred |x|)            */
} orc_hio_params;

/* HIO + shrinkwrap on x [B][M][N] with amplitudes A [amp_batch][M][N] and a
 * u8 support [B][M][N] (updated in place by shrinkwrap). */
int orc_hio(double* x, const double* A, uint8_t* support, int64_t M, int64_t N,
            int64_t B, int64_t amp_batch, const orc_hio_params* p);

/* One HIO step (no shrinkwrap). */
int orc_hio_step(double* x, const double* A, const uint8_t* support, int64_t M,
                 int64_t N, int64_t B, int64_t amp_batch, double beta,
                 int real_positive, int dc_skip_zero);

/* Shrinkwrap: b = Re Backward2D(Forward2D(|x|) * G_sigma), with the centred
 * analytic Gaussian transfer function
 *   G(kr,kc) = exp(-2 pi^2 sigma^2 [((kr-M/2)/M)^2 + ((kc-N/2)/N)^2]);
 * support = (b >= frac * max_pattern(b)).  If blur_out != NULL the blurred
 * real image is written there ([B][M][N] doubles). */
int orc_shrinkwrap(const double* x, uint8_t* support, int64_t M, int64_t N, int64_t B,
                   double sigma, double frac, double* blur_out);

/* R-factor per pattern: sum_{A>0} | |Forward2D(x)| - A | / sum_{A>0} A. */
int orc_rfactor(const double* x, const double* A, int64_t M, int64_t N, int64_t B,
                int64_t amp_batch, double* r_out);

/* Convergence error per batch member over m*n*l elements (the intent of
 * CudaImpl::ConvergeError, include/detail/CUDA_Impl.inl:145-172, which
 * normalises over m*n only): old <- new - old; err_b = ||old_b|| / ||new_b||. */
int orc_converge_error(double* old_, const double* new_, double* err, int64_t m, int64_t n,
                       int64_t l, int64_t batch);

/* ---- carousel slice geometry (gather-trilinear merge) ---------------------
 * Volume [L][M][N] (index z, y, x), cube of side n = L = M = N, centre c = n/2.
 * Slice p (angle th_p) is an [M][N] image with rows u (along the rotation
 * axis z) and columns v; its sample (u, v) sits at the volume point
 *   x = c + (v - c) cos th,  y = c + (v - c) sin th,  z = u.
 * orc_slice_extract (R) samples the volume trilinearly at those points
 * (zero outside); orc_slice_scatter (R^T) scatters slice values with the same
 * trilinear weights.  orc_slice_merge = R^T(s) / R^T(1) where R^T(1) > 0.
 * The pair is pinned by <R v, s> = <v, R^T s>. */
int orc_slice_extract(const double* vol, int64_t n, const double* angles, int64_t P,
                      double* slices, int complex_data);
int orc_slice_scatter(const double* slices, int64_t n, const double* angles, int64_t P,
                      double* vol, double* weight, int complex_data);
int orc_slice_merge(const double* slices, int64_t n, const double* angles, int64_t P,
                    double* vol, int complex_data);

/* ---- amplitude loss --------------------------------------------------------
 * L_b = sum_k (|F x_b|_k - A_k)^2 with F = Forward2D (centred, unnormalised).
 * grad = dL/d(re) + i dL/d(im) = 2*M*N*(x - Ptilde_M x) where Ptilde_M is
 * the modulus projection without the A == 0 skip.  For volumes use *_3d. */
int orc_amp_loss(const double* x, const double* A, int64_t M, int64_t N, int64_t B,
                 int64_t amp_batch, double* loss);
int orc_amp_grad(const double* x, const double* A, int64_t M, int64_t N, int64_t B,
                 int64_t amp_batch, double* grad);
int orc_amp_loss_3d(const double* x, const double* A, int64_t n, double* loss);
int orc_amp_grad_3d(const double* x, const double* A, int64_t n, double* grad);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
