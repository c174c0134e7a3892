/*
 * cpra_oracle.c -- CPU oracle for the CPRA hot path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library.  It restates the reference's MKL backend
 * (include/detail/MKL_Impl.inl) and driver compositions (src/CPRA_MKL.cpp)
 * in plain C; FFTs are an independent FP64 mixed-radix Stockham transform.
 * Each function cites the reference lines it follows; the additions beyond
 * the reference cite the textbook definition they implement.
 */
#include "cpra_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO_PI 6.283185307179586476925286766559

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* FFT                                                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t n;
    int nfac;
    int64_t fac[64];
    double* tw; /* tw[2k], tw[2k+1] = cos, sin(2 pi k / n) */
} orc_plan;

static void plan_init(orc_plan* p, int64_t n)
{
    int64_t m = n;
    p->n = n;
    p->nfac = 0;
    while (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
    while (m % 3 == 0) { p->fac[p->nfac++] = 3; m /= 3; }
    while (m % 5 == 0) { p->fac[p->nfac++] = 5; m /= 5; }
    for (int64_t f = 7; m > 1; f += 2) {
        while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
    }
    p->tw = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        /* reduce the angle exactly before calling cos/sin */
        double a = TWO_PI * (double)k / (double)n;
        p->tw[2 * k] = cos(a);
        p->tw[2 * k + 1] = sin(a);
    }
}

static void plan_free(orc_plan* p) { free(p->tw); }

/* e^{sign * 2 pi i * num / den} for 0 <= num, using the n-point table when
 * den divides n. */
static inline void twid(const orc_plan* p, int64_t num, int64_t den, int sign,
                        double* c, double* s)
{
    int64_t idx = (num * (p->n / den)) % p->n;
    *c = p->tw[2 * idx];
    *s = sign * p->tw[2 * idx + 1];
}

/* One contiguous line of length n, Stockham autosort (Govindaraju et al.,
 * "High performance discrete Fourier transforms on graphics processors",
 * SC'08, Fig. 2), radix r butterflies done as direct r-point DFTs. */
static void fft_line(const orc_plan* p, double* a, double* b, int sign)
{
    const int64_t n = p->n;
    double* in = a;
    double* out = b;
    int64_t ns = 1;
    double v[2 * 64], w[2 * 64];
    for (int f = 0; f < p->nfac; ++f) {
        const int64_t r = p->fac[f];
        const int64_t nr = n / r;
        double* vv = v;
        double* ww = w;
        double* heapv = NULL;
        if (r > 64) {
            heapv = (double*)malloc(sizeof(double) * 4 * (size_t)r);
            vv = heapv;
            ww = heapv + 2 * r;
        }
        for (int64_t j = 0; j < nr; ++j) {
            const int64_t k = j % ns;
            for (int64_t q = 0; q < r; ++q) {
                double xr = in[2 * (j + q * nr)], xi = in[2 * (j + q * nr) + 1];
                double c, s;
                twid(p, q * k, ns * r, sign, &c, &s);
                vv[2 * q] = xr * c - xi * s;
                vv[2 * q + 1] = xr * s + xi * c;
            }
            /* direct r-point DFT */
            for (int64_t kk = 0; kk < r; ++kk) {
                double sr = 0.0, si = 0.0;
                for (int64_t q = 0; q < r; ++q) {
                    double c, s;
                    twid(p, (q * kk) % r, r, sign, &c, &s);
                    sr += vv[2 * q] * c - vv[2 * q + 1] * s;
                    si += vv[2 * q] * s + vv[2 * q + 1] * c;
                }
                ww[2 * kk] = sr;
                ww[2 * kk + 1] = si;
            }
            const int64_t idxd = (j / ns) * ns * r + k;
            for (int64_t q = 0; q < r; ++q) {
                out[2 * (idxd + q * ns)] = ww[2 * q];
                out[2 * (idxd + q * ns) + 1] = ww[2 * q + 1];
            }
        }
        free(heapv);
        double* t = in;
        in = out;
        out = t;
        ns *= r;
    }
    if (in != a) memcpy(a, in, sizeof(double) * 2 * (size_t)n);
}

int orc_fft_lines(double* data, int64_t n, int64_t stride, int64_t dist, int64_t count,
                  int sign, double scale)
{
    if (n < 1 || count < 0) return 0;
    orc_plan p;
    plan_init(&p, n);
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        double* a = (double*)malloc(sizeof(double) * 4 * (size_t)n);
        double* b = a + 2 * n;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t j = 0; j < count; ++j) {
            double* base = data + 2 * j * dist;
            for (int64_t k = 0; k < n; ++k) {
                a[2 * k] = base[2 * k * stride];
                a[2 * k + 1] = base[2 * k * stride + 1];
            }
            fft_line(&p, a, b, sign);
            for (int64_t k = 0; k < n; ++k) {
                base[2 * k * stride] = a[2 * k] * scale;
                base[2 * k * stride + 1] = a[2 * k + 1] * scale;
            }
        }
        free(a);
    }
    plan_free(&p);
    return 1;
}

int orc_dft2(double* x, int64_t M, int64_t N, int64_t B, int sign)
{
    const double scale = sign > 0 ? 1.0 / ((double)M * (double)N) : 1.0;
    for (int64_t b = 0; b < B; ++b) {
        double* xb = x + 2 * b * M * N;
        orc_fft_lines(xb, N, 1, N, M, sign, 1.0);      /* rows    */
        orc_fft_lines(xb, M, N, 1, N, sign, scale);    /* columns */
    }
    return 1;
}

int orc_dft3(double* x, int64_t M, int64_t N, int64_t L, int64_t B, int sign)
{
    const double scale = sign > 0 ? 1.0 / ((double)M * (double)N * (double)L) : 1.0;
    for (int64_t b = 0; b < B; ++b) {
        double* xb = x + 2 * b * L * M * N;
        orc_fft_lines(xb, N, 1, N, L * M, sign, 1.0); /* x: contiguous rows */
        for (int64_t l = 0; l < L; ++l)                /* y: stride N        */
            orc_fft_lines(xb + 2 * l * M * N, M, N, 1, N, sign, 1.0);
        orc_fft_lines(xb, L, M * N, 1, M * N, sign, scale); /* z: stride MN */
    }
    return 1;
}

/* MklImpl::FFTShift2D, include/detail/MKL_Impl.inl:180-209 */
int orc_fftshift2d(void* x, size_t elem, int64_t M, int64_t N, int64_t B)
{
    unsigned char* d = (unsigned char*)x;
    unsigned char tmp[64];
    const uint64_t total = (uint64_t)M * N * B / 2;
    for (uint64_t i = 0; i < total; ++i) {
        uint64_t bidx = i / ((uint64_t)M * N / 2);
        uint64_t idx = i - bidx * M * N / 2;
        uint64_t col = idx % N;
        uint64_t row = idx / N;
        uint64_t off = bidx * M * N;
        uint64_t crow = row + M / 2;
        uint64_t ccol = col < (uint64_t)N / 2 ? col + N / 2 : col - N / 2;
        unsigned char* a = d + elem * (crow * N + ccol + off);
        unsigned char* c = d + elem * (row * N + col + off);
        memcpy(tmp, a, elem);
        memcpy(a, c, elem);
        memcpy(c, tmp, elem);
    }
    return 1;
}

/* MklImpl::FFTShift3D, include/detail/MKL_Impl.inl:212-251 */
int orc_fftshift3d(void* x, size_t elem, int64_t M, int64_t N, int64_t L, int64_t B)
{
    unsigned char* d = (unsigned char*)x;
    unsigned char tmp[64];
    const uint64_t total = (uint64_t)M * N * L * B / 2;
    for (uint64_t i = 0; i < total; ++i) {
        uint64_t bidx = i / ((uint64_t)M * N * L / 2);
        uint64_t idx = i - bidx * M * N * L / 2;
        uint64_t dep = idx / ((uint64_t)M * N);
        idx -= dep * M * N;
        uint64_t col = idx % N;
        uint64_t row = idx / N;
        uint64_t off = bidx * M * N * L;
        uint64_t cdep = dep + L / 2;
        uint64_t crow = row < (uint64_t)M / 2 ? row + M / 2 : row - M / 2;
        uint64_t ccol = col < (uint64_t)N / 2 ? col + N / 2 : col - N / 2;
        unsigned char* a = d + elem * (cdep * M * N + crow * N + ccol + off);
        unsigned char* c = d + elem * (dep * M * N + row * N + col + off);
        memcpy(tmp, a, elem);
        memcpy(a, c, elem);
        memcpy(c, tmp, elem);
    }
    return 1;
}

/* MklImpl::Forward2D / Backward2D, MKL_Impl.inl:128-152 */
int orc_forward2d(double* x, int64_t M, int64_t N, int64_t B)
{
    orc_fftshift2d(x, 16, M, N, B);
    orc_dft2(x, M, N, B, -1);
    orc_fftshift2d(x, 16, M, N, B);
    return 1;
}

int orc_backward2d(double* x, int64_t M, int64_t N, int64_t B)
{
    orc_fftshift2d(x, 16, M, N, B);
    orc_dft2(x, M, N, B, +1);
    orc_fftshift2d(x, 16, M, N, B);
    return 1;
}

/* MklImpl::Forward3D / Backward3D, MKL_Impl.inl:154-178, with the intended
 * {m, n, l} geometry (the reference builds the 3D descriptor from dim2D). */
int orc_forward3d(double* x, int64_t M, int64_t N, int64_t L, int64_t B)
{
    orc_fftshift3d(x, 16, M, N, L, B);
    orc_dft3(x, M, N, L, B, -1);
    orc_fftshift3d(x, 16, M, N, L, B);
    return 1;
}

int orc_backward3d(double* x, int64_t M, int64_t N, int64_t L, int64_t B)
{
    orc_fftshift3d(x, 16, M, N, L, B);
    orc_dft3(x, M, N, L, B, +1);
    orc_fftshift3d(x, 16, M, N, L, B);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* pointwise ops                                                             */
/* ------------------------------------------------------------------------ */

/* MKL_Impl.inl:70-84 */
int orc_space_constraint_f32(float* x, const float* S, uint64_t num, uint64_t batch)
{
    if (num == 0 || batch == 0) return 0;
    const uint64_t per = num / batch;
    for (uint64_t i = 0; i < num; ++i) {
        x[2 * i + 1] = 0.0f;
        if ((double)S[i % per] < 0.9) x[2 * i] = 0.0f;
    }
    return 1;
}

/* CUDA_Impl_Kernel.cuh:16-28 */
int orc_space_constraint_pos_f32(float* x, const float* S, uint64_t num, uint64_t batch)
{
    if (num == 0 || batch == 0) return 0;
    const uint64_t per = num / batch;
    for (uint64_t i = 0; i < num; ++i) {
        x[2 * i + 1] = 0.0f;
        if (x[2 * i] < 0.0f)
            x[2 * i] = 0.0f;
        else
            x[2 * i] = x[2 * i] * S[i % per];
    }
    return 1;
}

/* MKL_Impl.inl:86-106 (skip_zero) and CUDA_Impl_Kernel.cuh:30-38 */
int orc_data_constraint_f32(float* x, const float* A, uint64_t num, uint64_t batch,
                            int skip_zero)
{
    if (num == 0 || batch == 0) return 0;
    const uint64_t per = num / batch;
    for (uint64_t i = 0; i < num; ++i) {
        const float a = A[i % per];
        if (skip_zero && a == 0.0f) continue;
        const float re = x[2 * i], im = x[2 * i + 1];
        if (re * re + im * im > 0.0f) {
            const double th = atan2((double)im, (double)re);
            x[2 * i] = (float)((double)a * cos(th));
            x[2 * i + 1] = (float)((double)a * sin(th));
        }
    }
    return 1;
}

/* MKL_Impl.inl:108-126 */
int orc_data_constraint_cplx_f32(float* x, const float* C, uint64_t num, uint64_t batch)
{
    if (num == 0 || batch == 0) return 0;
    const uint64_t per = num / batch;
    for (uint64_t i = 0; i < num; ++i) {
        const float cr = C[2 * (i % per)], ci = C[2 * (i % per) + 1];
        if (cr * cr + ci * ci != 0.0f) {
            x[2 * i] = cr;
            x[2 * i + 1] = ci;
        }
    }
    return 1;
}

/* MKL_Impl.inl:31-50 (cblas_caxpby with real alpha, beta) */
int orc_merge_add_f32(const float* src, float* dst, float alpha, float beta, uint64_t num)
{
    if (alpha == 0.0f || num == 0) return 0;
    for (uint64_t i = 0; i < 2 * num; ++i) dst[i] = alpha * src[i] + beta * dst[i];
    return 1;
}

/* MKL_Impl.inl:52-68 */
int orc_normalize_f32(float* x, float norm, uint64_t num)
{
    if (norm == 0.0f || num == 0) return 0;
    const float a = (float)(1.0 / norm);
    for (uint64_t i = 0; i < 2 * num; ++i) x[i] = x[i] * a;
    return 1;
}

int orc_space_constraint_f64(double* x, const double* S, uint64_t num, uint64_t batch)
{
    if (num == 0 || batch == 0) return 0;
    const uint64_t per = num / batch;
    for (uint64_t i = 0; i < num; ++i) {
        x[2 * i + 1] = 0.0;
        if (S[i % per] < 0.9) x[2 * i] = 0.0;
    }
    return 1;
}

int orc_data_constraint_f64(double* x, const double* A, uint64_t num, uint64_t batch,
                            int skip_zero)
{
    if (num == 0 || batch == 0) return 0;
    const uint64_t per = num / batch;
    for (uint64_t i = 0; i < num; ++i) {
        const double a = A[i % per];
        if (skip_zero && a == 0.0) continue;
        const double re = x[2 * i], im = x[2 * i + 1];
        const double m2 = re * re + im * im;
        if (m2 > 0.0) {
            const double th = atan2(im, re);
            x[2 * i] = a * cos(th);
            x[2 * i + 1] = a * sin(th);
        }
    }
    return 1;
}

int orc_merge_add_f64(const double* src, double* dst, double alpha, double beta, uint64_t num)
{
    if (alpha == 0.0 || num == 0) return 0;
    for (uint64_t i = 0; i < 2 * num; ++i) dst[i] = alpha * src[i] + beta * dst[i];
    return 1;
}

/* ------------------------------------------------------------------------ */
/* compositions                                                              */
/* ------------------------------------------------------------------------ */

/* P_M on one batch member: Forward2D, DataConstraint, Backward2D (L == 1),
 * or Forward3D, DataConstraint, Backward3D on an L x M x N volume
 * (src/CONV_MKL.cpp:48-50). */
static void pm_one(double* z, const double* A, int64_t M, int64_t N, int64_t L, int skip_zero)
{
    if (L == 1) orc_forward2d(z, M, N, 1);
    else orc_forward3d(z, M, N, L, 1);
    orc_data_constraint_f64(z, A, (uint64_t)(M * N * L), 1, skip_zero);
    if (L == 1) orc_backward2d(z, M, N, 1);
    else orc_backward3d(z, M, N, L, 1);
}

/* src/CPRA_MKL.cpp:94-110 in 2D, src/CONV_MKL.cpp:47-63 in 3D (the same
 * composition over Forward3D/Backward3D); one batch member at a time (the
 * reference's batched calls are independent per member). */
static int dm_ref_geom(double* psi, double* t1, double* t2, const double* A, const double* S,
                       int64_t M, int64_t N, int64_t L, int64_t B, int64_t amp_batch,
                       double beta, int64_t iters)
{
    const int64_t px = M * N * L;
    if (amp_batch < 1 || B % amp_batch) return 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t b = 0; b < B; ++b) {
        double* p = psi + 2 * b * px;
        double* a1 = t1 + 2 * b * px;
        double* a2 = t2 + 2 * b * px;
        const double* Ab = A + (b % amp_batch) * px;
        const double* Sb = S + (b % amp_batch) * px;
        for (int64_t it = 0; it < iters; ++it) {
            /* Part 1: CPRA_MKL.cpp:95-99 (CONV_MKL.cpp:48-52) */
            pm_one(a1, Ab, M, N, L, 1);
            orc_merge_add_f64(p, a1, -1.0 / beta, 1.0 + 1.0 / beta, (uint64_t)px);
            orc_space_constraint_f64(a1, Sb, (uint64_t)px, 1);
            /* Part 2: CPRA_MKL.cpp:102-106 (CONV_MKL.cpp:55-59) */
            orc_space_constraint_f64(a2, Sb, (uint64_t)px, 1);
            orc_merge_add_f64(p, a2, 1.0 / beta, 1.0 - 1.0 / beta, (uint64_t)px);
            pm_one(a2, Ab, M, N, L, 1);
            /* Merge: CPRA_MKL.cpp:109-110 (CONV_MKL.cpp:62-63) */
            orc_merge_add_f64(a1, p, beta, 1.0, (uint64_t)px);
            orc_merge_add_f64(a2, p, -1.0 * beta, 1.0, (uint64_t)px);
        }
    }
    return 1;
}

int orc_dm_ref(double* psi, double* t1, double* t2, const double* A, const double* S,
               int64_t M, int64_t N, int64_t B, int64_t amp_batch, double beta,
               int64_t iters)
{
    return dm_ref_geom(psi, t1, t2, A, S, M, N, 1, B, amp_batch, beta, iters);
}

int orc_dm_ref3d(double* psi, double* t1, double* t2, const double* A, const double* S,
                 int64_t M, int64_t N, int64_t L, int64_t B, int64_t amp_batch, double beta,
                 int64_t iters)
{
    return dm_ref_geom(psi, t1, t2, A, S, M, N, L, B, amp_batch, beta, iters);
}

/* src/CPRA_MKL.cpp:126-140 in 2D, src/CONV_MKL.cpp:78-88 in 3D */
static int final_projection_geom(double* psi, const double* A, const double* S, int64_t M,
                                 int64_t N, int64_t L, int64_t B, int64_t amp_batch,
                                 double beta)
{
    const int64_t px = M * N * L;
    if (amp_batch < 1 || B % amp_batch) return 0;
    double* t = (double*)malloc(sizeof(double) * 2 * (size_t)px);
    for (int64_t b = 0; b < B; ++b) {
        double* p = psi + 2 * b * px;
        memcpy(t, p, sizeof(double) * 2 * (size_t)px);
        pm_one(t, A + (b % amp_batch) * px, M, N, L, 1);
        orc_merge_add_f64(p, t, -1.0 / beta, 1.0 + 1.0 / beta, (uint64_t)px);
        orc_space_constraint_f64(t, S + (b % amp_batch) * px, (uint64_t)px, 1);
        memcpy(p, t, sizeof(double) * 2 * (size_t)px);
    }
    free(t);
    return 1;
}

int orc_final_projection(double* psi, const double* A, const double* S, int64_t M,
                         int64_t N, int64_t B, int64_t amp_batch, double beta)
{
    return final_projection_geom(psi, A, S, M, N, 1, B, amp_batch, beta);
}

int orc_final_projection3d(double* psi, const double* A, const double* S, int64_t M,
                           int64_t N, int64_t L, int64_t B, int64_t amp_batch, double beta)
{
    return final_projection_geom(psi, A, S, M, N, L, B, amp_batch, beta);
}

static void ps_apply(double* z, const double* S, const uint8_t* mask, int64_t px,
                     int space_mode)
{
    if (space_mode == 0) {
        orc_space_constraint_f64(z, S, (uint64_t)px, 1);
        return;
    }
    for (int64_t i = 0; i < px; ++i) {
        if (!mask[i]) {
            z[2 * i] = 0.0;
            z[2 * i + 1] = 0.0;
        } else if (space_mode == 1) {
            z[2 * i + 1] = 0.0;
            if (z[2 * i] < 0.0) z[2 * i] = 0.0;
        }
    }
}

/* Elser, "Phase retrieval by iterated projections", JOSA A 20 (2003) 40. */
int orc_dm_textbook(double* psi, const double* A, const double* S, const uint8_t* mask,
                    int64_t M, int64_t N, int64_t B, int64_t amp_batch, double beta,
                    int space_mode, int64_t iters)
{
    const int64_t px = M * N;
    if (amp_batch < 1 || B % amp_batch) return 0;
    if (space_mode == 0 && !S) return 0;
    if (space_mode != 0 && !mask) return 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t b = 0; b < B; ++b) {
        double* p = psi + 2 * b * px;
        const double* Ab = A + (b % amp_batch) * px;
        const double* Sb = S ? S + (b % amp_batch) * px : NULL;
        const uint8_t* mb = mask ? mask + b * px : NULL;
        double* u = (double*)malloc(sizeof(double) * 4 * (size_t)px);
        double* w = u + 2 * px;
        for (int64_t it = 0; it < iters; ++it) {
            /* u = P_S((1+1/beta) P_M psi - psi/beta) */
            memcpy(u, p, sizeof(double) * 2 * (size_t)px);
            pm_one(u, Ab, M, N, 1, 1);
            for (int64_t i = 0; i < 2 * px; ++i)
                u[i] = (1.0 + 1.0 / beta) * u[i] - p[i] / beta;
            ps_apply(u, Sb, mb, px, space_mode);
            /* w = P_M((1-1/beta) P_S psi + psi/beta) */
            memcpy(w, p, sizeof(double) * 2 * (size_t)px);
            ps_apply(w, Sb, mb, px, space_mode);
            for (int64_t i = 0; i < 2 * px; ++i)
                w[i] = (1.0 - 1.0 / beta) * w[i] + p[i] / beta;
            pm_one(w, Ab, M, N, 1, 1);
            for (int64_t i = 0; i < 2 * px; ++i) p[i] += beta * (u[i] - w[i]);
        }
        free(u);
    }
    return 1;
}

/* One HIO step on one pattern (Fienup, Appl. Opt. 21 (1982) 2758):
 * x' = P_M x; where the pixel satisfies the object-domain constraints
 * (inside the support and, in real-positive mode, Re x' >= 0) the new value
 * is x' (real part only in real-positive mode), elsewhere x - beta x'. */
static void hio_one(double* x, double* xp, const double* A, const uint8_t* s, int64_t M,
                    int64_t N, double beta, int real_positive, int skip_zero)
{
    const int64_t px = M * N;
    memcpy(xp, x, sizeof(double) * 2 * (size_t)px);
    pm_one(xp, A, M, N, 1, skip_zero);
    for (int64_t i = 0; i < px; ++i) {
        const double pr = xp[2 * i], pi = xp[2 * i + 1];
        const int ok = s[i] && (!real_positive || pr >= 0.0);
        if (ok) {
            x[2 * i] = pr;
            x[2 * i + 1] = real_positive ? 0.0 : pi;
        } else {
            x[2 * i] = x[2 * i] - beta * pr;
            x[2 * i + 1] = x[2 * i + 1] - beta * pi;
        }
    }
}

int orc_hio_step(double* x, const double* A, const uint8_t* support, int64_t M, int64_t N,
                 int64_t B, int64_t amp_batch, double beta, int real_positive,
                 int dc_skip_zero)
{
    const int64_t px = M * N;
    if (amp_batch < 1 || B % amp_batch) return 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t b = 0; b < B; ++b) {
        double* xp = (double*)malloc(sizeof(double) * 2 * (size_t)px);
        hio_one(x + 2 * b * px, xp, A + (b % amp_batch) * px, support + b * px, M, N, beta,
                real_positive, dc_skip_zero);
        free(xp);
    }
    return 1;
}

static double gauss_transfer(int64_t kr, int64_t kc, int64_t M, int64_t N, double sigma)
{
    const double fr = (double)(kr - M / 2) / (double)M;
    const double fc = (double)(kc - N / 2) / (double)N;
    return exp(-2.0 * M_PI * M_PI * sigma * sigma * (fr * fr + fc * fc));
}

static void shrinkwrap_one(const double* x, uint8_t* s, int64_t M, int64_t N, double sigma,
                           double frac, double* work, double* blur_out)
{
    const int64_t px = M * N;
    for (int64_t i = 0; i < px; ++i) {
        work[2 * i] = sqrt(x[2 * i] * x[2 * i] + x[2 * i + 1] * x[2 * i + 1]);
        work[2 * i + 1] = 0.0;
    }
    orc_forward2d(work, M, N, 1);
    for (int64_t r = 0; r < M; ++r)
        for (int64_t c = 0; c < N; ++c) {
            const double g = gauss_transfer(r, c, M, N, sigma);
            work[2 * (r * N + c)] *= g;
            work[2 * (r * N + c) + 1] *= g;
        }
    orc_backward2d(work, M, N, 1);
    double mx = -INFINITY;
    for (int64_t i = 0; i < px; ++i)
        if (work[2 * i] > mx) mx = work[2 * i];
    const double thr = frac * mx;
    for (int64_t i = 0; i < px; ++i) s[i] = work[2 * i] >= thr ? 1 : 0;
    if (blur_out)
        for (int64_t i = 0; i < px; ++i) blur_out[i] = work[2 * i];
}

int orc_shrinkwrap(const double* x, uint8_t* support, int64_t M, int64_t N, int64_t B,
                   double sigma, double frac, double* blur_out)
{
    const int64_t px = M * N;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t b = 0; b < B; ++b) {
        double* work = (double*)malloc(sizeof(double) * 2 * (size_t)px);
        shrinkwrap_one(x + 2 * b * px, support + b * px, M, N, sigma, frac, work,
                       blur_out ? blur_out + b * px : NULL);
        free(work);
    }
    return 1;
}

int orc_hio(double* x, const double* A, uint8_t* support, int64_t M, int64_t N, int64_t B,
            int64_t amp_batch, const orc_hio_params* p)
{
    const int64_t px = M * N;
    if (amp_batch < 1 || B % amp_batch || !p) return 0;
    const int64_t nsw = p->sw_every > 0 ? p->iters / p->sw_every : 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t b = 0; b < B; ++b) {
        double* xp = (double*)malloc(sizeof(double) * 2 * (size_t)px);
        double* xb = x + 2 * b * px;
        uint8_t* sb = support + b * px;
        const double* Ab = A + (b % amp_batch) * px;
        for (int64_t it = 1; it <= p->iters; ++it) {
            hio_one(xb, xp, Ab, sb, M, N, p->beta, p->real_positive, p->dc_skip_zero);
            if (p->sw_every > 0 && it % p->sw_every == 0) {
                const int64_t k = it / p->sw_every - 1;
                const double sigma =
                    nsw > 1 ? p->sw_sigma0 + (p->sw_sigma1 - p->sw_sigma0) * (double)k /
                                                 (double)(nsw - 1)
                            : p->sw_sigma0;
                shrinkwrap_one(xb, sb, M, N, sigma, p->sw_frac, xp, NULL);
            }
        }
        free(xp);
    }
    return 1;
}

int orc_rfactor(const double* x, const double* A, int64_t M, int64_t N, int64_t B,
                int64_t amp_batch, double* r_out)
{
    const int64_t px = M * N;
    if (amp_batch < 1 || B % amp_batch) return 0;
    double* w = (double*)malloc(sizeof(double) * 2 * (size_t)px);
    for (int64_t b = 0; b < B; ++b) {
        memcpy(w, x + 2 * b * px, sizeof(double) * 2 * (size_t)px);
        orc_forward2d(w, M, N, 1);
        const double* Ab = A + (b % amp_batch) * px;
        double num = 0.0, den = 0.0;
        for (int64_t i = 0; i < px; ++i) {
            if (Ab[i] > 0.0) {
                num += fabs(hypot(w[2 * i], w[2 * i + 1]) - Ab[i]);
                den += Ab[i];
            }
        }
        r_out[b] = den > 0.0 ? num / den : 0.0;
    }
    free(w);
    return 1;
}

int orc_converge_error(double* old_, const double* new_, double* err, int64_t m, int64_t n,
                       int64_t l, int64_t batch)
{
    const int64_t per = m * n * l;
    for (int64_t i = 0; i < 2 * per * batch; ++i) old_[i] = new_[i] - old_[i];
    for (int64_t b = 0; b < batch; ++b) {
        double so = 0.0, sn = 0.0;
        for (int64_t i = 0; i < 2 * per; ++i) {
            so += old_[2 * b * per + i] * old_[2 * b * per + i];
            sn += new_[2 * b * per + i] * new_[2 * b * per + i];
        }
        err[b] = sqrt(so) / sqrt(sn);
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* slice geometry                                                            */
/* ------------------------------------------------------------------------ */

/* Visit the (up to 4) voxels that the trilinear footprint of slice sample
 * (p, u, v) touches; z = u exactly, so the weights are bilinear in (x, y). */
#define SLICE_FOOTPRINT(BODY)                                                   \
    do {                                                                        \
        const double c0 = (double)(n / 2);                                      \
        const double ct = cos(angles[p]), st = sin(angles[p]);                  \
        const double px_ = c0 + ((double)v - c0) * ct;                          \
        const double py_ = c0 + ((double)v - c0) * st;                          \
        const double fx = floor(px_), fy = floor(py_);                          \
        const double ax = px_ - fx, ay = py_ - fy;                              \
        for (int dy = 0; dy < 2; ++dy)                                          \
            for (int dx = 0; dx < 2; ++dx) {                                    \
                const int64_t xi = (int64_t)fx + dx, yi = (int64_t)fy + dy;     \
                const double wgt = (dx ? ax : 1.0 - ax) * (dy ? ay : 1.0 - ay); \
                if (xi < 0 || xi >= n || yi < 0 || yi >= n || wgt == 0.0)      \
                    continue;                                                   \
                const int64_t vox = (u * n + yi) * n + xi;                      \
                BODY                                                            \
            }                                                                   \
    } while (0)

int orc_slice_extract(const double* vol, int64_t n, const double* angles, int64_t P,
                      double* slices, int complex_data)
{
    const int cw = complex_data ? 2 : 1;
    for (int64_t p = 0; p < P; ++p)
        for (int64_t u = 0; u < n; ++u)
            for (int64_t v = 0; v < n; ++v) {
                double acc[2] = {0.0, 0.0};
                SLICE_FOOTPRINT({
                    for (int q = 0; q < cw; ++q) acc[q] += wgt * vol[cw * vox + q];
                });
                for (int q = 0; q < cw; ++q) slices[cw * ((p * n + u) * n + v) + q] = acc[q];
            }
    return 1;
}

int orc_slice_scatter(const double* slices, int64_t n, const double* angles, int64_t P,
                      double* vol, double* weight, int complex_data)
{
    const int cw = complex_data ? 2 : 1;
    memset(vol, 0, sizeof(double) * (size_t)(cw * n * n * n));
    if (weight) memset(weight, 0, sizeof(double) * (size_t)(n * n * n));
    for (int64_t p = 0; p < P; ++p)
        for (int64_t u = 0; u < n; ++u)
            for (int64_t v = 0; v < n; ++v) {
                const double* s = slices + cw * ((p * n + u) * n + v);
                SLICE_FOOTPRINT({
                    for (int q = 0; q < cw; ++q) vol[cw * vox + q] += wgt * s[q];
                    if (weight) weight[vox] += wgt;
                });
            }
    return 1;
}

int orc_slice_merge(const double* slices, int64_t n, const double* angles, int64_t P,
                    double* vol, int complex_data)
{
    const int cw = complex_data ? 2 : 1;
    double* w = (double*)malloc(sizeof(double) * (size_t)(n * n * n));
    orc_slice_scatter(slices, n, angles, P, vol, w, complex_data);
    for (int64_t i = 0; i < n * n * n; ++i)
        if (w[i] > 0.0)
            for (int q = 0; q < cw; ++q) vol[cw * i + q] /= w[i];
    free(w);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* amplitude loss                                                            */
/* ------------------------------------------------------------------------ */

int orc_amp_loss(const double* x, const double* A, int64_t M, int64_t N, int64_t B,
                 int64_t amp_batch, double* loss)
{
    const int64_t px = M * N;
    if (amp_batch < 1 || B % amp_batch) return 0;
    double* w = (double*)malloc(sizeof(double) * 2 * (size_t)px);
    for (int64_t b = 0; b < B; ++b) {
        memcpy(w, x + 2 * b * px, sizeof(double) * 2 * (size_t)px);
        orc_forward2d(w, M, N, 1);
        const double* Ab = A + (b % amp_batch) * px;
        double s = 0.0;
        for (int64_t i = 0; i < px; ++i) {
            const double d = hypot(w[2 * i], w[2 * i + 1]) - Ab[i];
            s += d * d;
        }
        loss[b] = s;
    }
    free(w);
    return 1;
}

int orc_amp_grad(const double* x, const double* A, int64_t M, int64_t N, int64_t B,
                 int64_t amp_batch, double* grad)
{
    const int64_t px = M * N;
    if (amp_batch < 1 || B % amp_batch) return 0;
    for (int64_t b = 0; b < B; ++b) {
        const double* xb = x + 2 * b * px;
        double* g = grad + 2 * b * px;
        memcpy(g, xb, sizeof(double) * 2 * (size_t)px);
        pm_one(g, A + (b % amp_batch) * px, M, N, 1, 0);
        for (int64_t i = 0; i < 2 * px; ++i) g[i] = 2.0 * (double)px * (xb[i] - g[i]);
    }
    return 1;
}

int orc_amp_loss_3d(const double* x, const double* A, int64_t n, double* loss)
{
    const int64_t px = n * n * n;
    double* w = (double*)malloc(sizeof(double) * 2 * (size_t)px);
    memcpy(w, x, sizeof(double) * 2 * (size_t)px);
    orc_forward3d(w, n, n, n, 1);
    double s = 0.0;
    for (int64_t i = 0; i < px; ++i) {
        const double d = hypot(w[2 * i], w[2 * i + 1]) - A[i];
        s += d * d;
    }
    *loss = s;
    free(w);
    return 1;
}

int orc_amp_grad_3d(const double* x, const double* A, int64_t n, double* grad)
{
    const int64_t px = n * n * n;
    memcpy(grad, x, sizeof(double) * 2 * (size_t)px);
    orc_forward3d(grad, n, n, n, 1);
    orc_data_constraint_f64(grad, A, (uint64_t)px, 1, 0);
    orc_backward3d(grad, n, n, n, 1);
    for (int64_t i = 0; i < 2 * px; ++i) grad[i] = 2.0 * (double)px * (x[i] - grad[i]);
    return 1;
}
