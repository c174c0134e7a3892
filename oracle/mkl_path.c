/* The reference's MKL backend in single precision, over Intel MKL loaded at
 * run time (TEST INFRASTRUCTURE AND CPU BASELINE ONLY).
 *
 * This is the CPU path that the benchmark times beside the GPU
 * (`cpu_baseline` of bench.py), and nothing else may call it: the HIP
 * library never links or loads it.  It restates, in FP32 with MKL DFTI and
 * CBLAS doing the arithmetic, what `CPRA::MklImpl<float>` does:
 *
 *   descriptor     include/MKL_Impl.hpp:19-45 (2D, DFTI_SINGLE, DFTI_COMPLEX,
 *                  NUMBER_OF_TRANSFORMS = batch, INPUT_DISTANCE = m*n,
 *                  BACKWARD_SCALE = 1/(m*n), in place)
 *   Forward2D      include/detail/MKL_Impl.inl:128-139 (shift, DFTI forward, shift)
 *   Backward2D     include/detail/MKL_Impl.inl:141-152
 *   FFTShift2D     include/detail/MKL_Impl.inl:180-209 (OpenMP swap loop)
 *   DataConstraint include/detail/MKL_Impl.inl:86-106 (A == 0 skip, polar(A, arg x))
 *   MergeAddData   include/detail/MKL_Impl.inl:31-50 (cblas_caxpby, alpha 0 no-op)
 *
 * and composes the workload the GPU runs from those calls, one pass per
 * primitive as the reference's drivers do (src/CPRA_MKL.cpp:94-110):
 * HIO (x' = Backward2D(DataConstraint(Forward2D(x))); inside the support
 * with Re x' >= 0, x = Re x'; elsewhere x = x - beta x') with shrinkwrap
 * (|x| -> Forward2D -> centred Gaussian transfer function -> Backward2D ->
 * per-pattern maximum -> threshold) every sw_every iterations, sigma annealed
 * linearly over the updates.
 *
 * MKL has no headers in this image, so the few DFTI/CBLAS entry points are
 * declared here from Intel's published interface and resolved with dlsym
 * from libmkl_rt.  A wrong enumeration value makes DftiCommitDescriptor or
 * the transforms fail with an MKL status, and the known-answer tests
 * (tests/test_mkl_path.py, against numpy's FFT and the FP64 oracle) catch
 * any that would not. */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef long mkl_long;                 /* MKL_LONG on LP64 Linux */
typedef void* dfti_handle;

/* DFTI configuration parameters and values (mkl_dfti.h) */
enum {
    DFTI_BACKWARD_SCALE_ = 5,
    DFTI_NUMBER_OF_TRANSFORMS_ = 7,
    DFTI_INPUT_DISTANCE_ = 14,
    DFTI_COMPLEX_ = 32,
    DFTI_SINGLE_ = 35,
};

static struct {
    void* so;
    mkl_long (*create)(dfti_handle*, int, int, mkl_long, ...);
    mkl_long (*set_value)(dfti_handle, int, ...);
    mkl_long (*commit)(dfti_handle);
    mkl_long (*forward)(dfti_handle, void*, ...);
    mkl_long (*backward)(dfti_handle, void*, ...);
    mkl_long (*free_desc)(dfti_handle*);
    char* (*error_message)(mkl_long);
    void (*caxpby)(int, const void*, const void*, int, const void*, void*, int);
    void (*set_num_threads)(int);
    int (*get_max_threads)(void);
    char path[512];
} M;

static char g_err[1024];

const char* mklp_last_error(void) { return g_err; }

/* Loads libmkl_rt (the given path, else the usual names and /opt/conda/lib).
 * MKL runs on the GNU OpenMP layer, the same runtime as this file's loops,
 * so the two share one thread pool.  Returns 1 when MKL is available. */
int mklp_load(const char* path)
{
    if (M.so) return 1;
    setenv("MKL_THREADING_LAYER", "GNU", 0);
    const char* cands[] = {path, "libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so",
                           "/opt/conda/lib/libmkl_rt.so.1"};
    for (size_t i = 0; i < sizeof(cands) / sizeof(cands[0]) && !M.so; ++i) {
        if (!cands[i]) continue;
        M.so = dlopen(cands[i], RTLD_NOW | RTLD_LOCAL);
        if (M.so) snprintf(M.path, sizeof(M.path), "%s", cands[i]);
    }
    if (!M.so) {
        snprintf(g_err, sizeof(g_err), "libmkl_rt not found: %s", dlerror());
        return 0;
    }
#define SYM(field, name)                                                          \
    do {                                                                          \
        *(void**)&M.field = dlsym(M.so, name);                                    \
        if (!M.field) {                                                           \
            snprintf(g_err, sizeof(g_err), "%s: missing symbol %s", M.path, name); \
            dlclose(M.so);                                                        \
            M.so = NULL;                                                          \
            return 0;                                                             \
        }                                                                         \
    } while (0)
    SYM(create, "DftiCreateDescriptor");
    SYM(set_value, "DftiSetValue");
    SYM(commit, "DftiCommitDescriptor");
    SYM(forward, "DftiComputeForward");
    SYM(backward, "DftiComputeBackward");
    SYM(free_desc, "DftiFreeDescriptor");
    SYM(error_message, "DftiErrorMessage");
    SYM(caxpby, "cblas_caxpby");
    SYM(set_num_threads, "MKL_Set_Num_Threads");
    SYM(get_max_threads, "MKL_Get_Max_Threads");
#undef SYM
    return 1;
}

const char* mklp_library_path(void) { return M.so ? M.path : ""; }

/* MKL and OpenMP threads for the calls that follow. */
void mklp_set_threads(int n)
{
    if (M.so) M.set_num_threads(n);
#ifdef _OPENMP
    omp_set_num_threads(n);
#endif
}

int mklp_max_threads(void) { return M.so ? M.get_max_threads() : 0; }

typedef struct {
    dfti_handle d2;
    int64_t m, n, batch;
} mklp_ctx;

static int check(mkl_long st, const char* what)
{
    if (st == 0) return 1;
    snprintf(g_err, sizeof(g_err), "%s: %s", what, M.error_message(st));
    return 0;
}

/* The 2D descriptor of include/MKL_Impl.hpp:19-45 (single precision). */
void* mklp_create(int64_t m, int64_t n, int64_t batch)
{
    if (!M.so || m < 2 || n < 2 || batch < 1 || (m & 1) || (n & 1)) {
        snprintf(g_err, sizeof(g_err), "mklp_create: MKL not loaded or bad geometry");
        return NULL;
    }
    mklp_ctx* c = (mklp_ctx*)calloc(1, sizeof(mklp_ctx));
    c->m = m;
    c->n = n;
    c->batch = batch;
    mkl_long dims[2] = {(mkl_long)m, (mkl_long)n};
    if (!check(M.create(&c->d2, DFTI_SINGLE_, DFTI_COMPLEX_, 2, dims), "DftiCreateDescriptor") ||
        !check(M.set_value(c->d2, DFTI_NUMBER_OF_TRANSFORMS_, (mkl_long)batch), "NUMBER_OF_TRANSFORMS") ||
        !check(M.set_value(c->d2, DFTI_INPUT_DISTANCE_, (mkl_long)(m * n)), "INPUT_DISTANCE") ||
        !check(M.set_value(c->d2, DFTI_BACKWARD_SCALE_, 1.0 / (double)(m * n)), "BACKWARD_SCALE") ||
        !check(M.commit(c->d2), "DftiCommitDescriptor")) {
        if (c->d2) M.free_desc(&c->d2);
        free(c);
        return NULL;
    }
    return c;
}

void mklp_destroy(void* h)
{
    mklp_ctx* c = (mklp_ctx*)h;
    if (!c) return;
    if (c->d2) M.free_desc(&c->d2);
    free(c);
}

/* include/detail/MKL_Impl.inl:180-209: swap (r, c) with (r + M/2,
 * (c + N/2) mod N) for the top half of every pattern. */
void mklp_fftshift2d(float* x, int64_t M_, int64_t N, int64_t B)
{
    const int64_t half = M_ * N / 2;
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t i = 0; i < half * B; ++i) {
        const int64_t b = i / half;
        const int64_t idx = i - b * half;
        const int64_t col = idx % N, row = idx / N;
        const int64_t ccol = col < N / 2 ? col + N / 2 : col - N / 2;
        float* p = x + 2 * (b * M_ * N + row * N + col);
        float* q = x + 2 * (b * M_ * N + (row + M_ / 2) * N + ccol);
        const float t0 = p[0], t1 = p[1];
        p[0] = q[0];
        p[1] = q[1];
        q[0] = t0;
        q[1] = t1;
    }
}

int mklp_forward2d(void* h, float* x)
{
    mklp_ctx* c = (mklp_ctx*)h;
    mklp_fftshift2d(x, c->m, c->n, c->batch);
    if (!check(M.forward(c->d2, x), "DftiComputeForward")) return 0;
    mklp_fftshift2d(x, c->m, c->n, c->batch);
    return 1;
}

int mklp_backward2d(void* h, float* x)
{
    mklp_ctx* c = (mklp_ctx*)h;
    mklp_fftshift2d(x, c->m, c->n, c->batch);
    if (!check(M.backward(c->d2, x), "DftiComputeBackward")) return 0;
    mklp_fftshift2d(x, c->m, c->n, c->batch);
    return 1;
}

/* include/detail/MKL_Impl.inl:86-106, with the constraint broadcast
 * i % (num / batch_size). */
void mklp_data_constraint(float* x, const float* A, uint64_t num, uint64_t batch_size)
{
    const uint64_t per = num / batch_size;
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (uint64_t i = 0; i < num; ++i) {
        const float a = A[i % per];
        if (a == 0.0f) continue;
        const float re = x[2 * i], im = x[2 * i + 1];
        if (re * re + im * im > 0.0f) {
            const float th = atan2f(im, re);
            x[2 * i] = a * cosf(th);
            x[2 * i + 1] = a * sinf(th);
        }
    }
}

/* include/detail/MKL_Impl.inl:31-50 */
int mklp_merge_add(const float* src, float* dst, float alpha, float beta, uint64_t num)
{
    if (alpha == 0.0f || num == 0) return 0;
    const float a[2] = {alpha, 0.0f}, b[2] = {beta, 0.0f};
    M.caxpby((int)num, a, src, 1, b, dst, 1);
    return 1;
}

static float gauss_transfer(int64_t kr, int64_t kc, int64_t M_, int64_t N, float sigma)
{
    const float fr = (float)(kr - M_ / 2) / (float)M_;
    const float fc = (float)(kc - N / 2) / (float)N;
    return expf(-2.0f * 9.8696044f * sigma * sigma * (fr * fr + fc * fc));
}

/* One shrinkwrap update from the reference's primitives: w = |x| (real),
 * Forward2D, times the centred Gaussian transfer function, Backward2D,
 * support = Re w >= frac * max(Re w) per pattern.  w: the batch's size. */
static int shrinkwrap(mklp_ctx* c, const float* x, uint8_t* s, float* w, float sigma, float frac)
{
    const int64_t px = c->m * c->n, tot = px * c->batch;
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t i = 0; i < tot; ++i) {
        w[2 * i] = hypotf(x[2 * i], x[2 * i + 1]);
        w[2 * i + 1] = 0.0f;
    }
    if (!mklp_forward2d(c, w)) return 0;
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t i = 0; i < tot; ++i) {
        const int64_t r = (i % px) / c->n, col = i % c->n;
        const float g = gauss_transfer(r, col, c->m, c->n, sigma);
        w[2 * i] *= g;
        w[2 * i + 1] *= g;
    }
    if (!mklp_backward2d(c, w)) return 0;
#ifdef _OPENMP
#pragma omp parallel for
#endif
    for (int64_t b = 0; b < c->batch; ++b) {
        const float* wb = w + 2 * b * px;
        float mx = -INFINITY;
        for (int64_t i = 0; i < px; ++i)
            if (wb[2 * i] > mx) mx = wb[2 * i];
        const float thr = frac * mx;
        for (int64_t i = 0; i < px; ++i) s[b * px + i] = wb[2 * i] >= thr ? 1 : 0;
    }
    return 1;
}

/* HIO with shrinkwrap on the handle's batch (x: complex64 [batch][m][n],
 * A: [amp_batch][m][n], support u8 [batch][m][n], updated in place).
 * `work`: 2 * batch * m * n floats.  Same schedule as the GPU engine:
 * an update after every sw_every-th iteration, the k-th of nsw = iters /
 * sw_every updates at sigma0 + (sigma1 - sigma0) k / (nsw - 1). */
int mklp_hio(void* h, float* x, const float* A, int64_t amp_batch, uint8_t* support, int64_t iters,
             float beta, int real_positive, int64_t sw_every, float sigma0, float sigma1, float frac,
             float* work)
{
    mklp_ctx* c = (mklp_ctx*)h;
    const int64_t px = c->m * c->n, tot = px * c->batch;
    if (amp_batch < 1 || c->batch % amp_batch) {
        snprintf(g_err, sizeof(g_err), "mklp_hio: amp_batch must divide the batch");
        return 0;
    }
    const int64_t nsw = sw_every > 0 ? iters / sw_every : 0;
    for (int64_t it = 1; it <= iters; ++it) {
        /* x' = P_M x: Memcpy, Forward2D, DataConstraint, Backward2D */
        memcpy(work, x, sizeof(float) * 2 * (size_t)tot);
        if (!mklp_forward2d(c, work)) return 0;
        mklp_data_constraint(work, A, (uint64_t)tot, (uint64_t)(c->batch / amp_batch));
        if (!mklp_backward2d(c, work)) return 0;
#ifdef _OPENMP
#pragma omp parallel for
#endif
        for (int64_t i = 0; i < tot; ++i) {
            const float pr = work[2 * i], pi = work[2 * i + 1];
            if (support[i] && (!real_positive || pr >= 0.0f)) {
                x[2 * i] = pr;
                x[2 * i + 1] = real_positive ? 0.0f : pi;
            } else {
                x[2 * i] -= beta * pr;
                x[2 * i + 1] -= beta * pi;
            }
        }
        if (sw_every > 0 && it % sw_every == 0) {
            const int64_t k = it / sw_every - 1;
            const float sigma =
                nsw > 1 ? sigma0 + (sigma1 - sigma0) * (float)k / (float)(nsw - 1) : sigma0;
            if (!shrinkwrap(c, x, support, work, sigma, frac)) return 0;
        }
    }
    return 1;
}
