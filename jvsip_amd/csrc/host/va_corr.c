/*
 * va_corr.c -- 1-D correlation: vsip_corr1d (real) and vsip_ccorr1d
 * (complex) objects and vsip_[c]correlate1d_[fd]; 1-D convolution: vsip_conv1d
 * objects and vsip_convolve1d_[fd] (at the end of the file).  Both are
 * callers of the FFT path.
 *
 * The reference (c_VSIP_src/vsip_ccorrelate1d_f.c:123-196,
 * vsip_correlate1d_f.c:82-151, creation in vsip_ccorr1d_create_f.c:17-54)
 * pads the data x (length n) and the reference h (length m) into two
 * complex vectors of length N, the smallest power of two above
 * mn = m + n - 1 (x at the end after N - n zeros, h at the start), takes
 * their forward FFTs, forms conj(X conj(H)) / N, takes one more forward FFT
 * and copies lags N - mn + s .. of it to y (s = 0 FULL, (m - 1)/2 SAME,
 * m - 1 MIN), divided by the support's unbiasing weights on request; the
 * real form uses zero imaginary parts and keeps the real part.
 *
 * Here the same steps run on the device, with N the smallest 2^a 3^b 5^c
 * >= mn instead (the lags kept are the same for any N >= mn, and the
 * reference's power of two can be nearly 2 mn): the two padded rows of one
 * 2 x N transform buffer are filled by one element pass, transformed by
 * one multiple-FFT launch (two ROW vectors), multiplied in one pass,
 * transformed again (row 0) and windowed into y in one pass (corr_ops.hip);
 * the FFTs are this library's own plans, so every length class they serve
 * applies.  create returns NULL for m < 1 or m > n (the reference's SAME
 * and MIN windows assume m <= n) and for an unknown support.
 *   create   vsip.h:1784-1785, 1797-1798
 *   destroy  vsip.h:379-380, 407-408 (returns 0)
 *   getattr  vsip_ccorr1d_getattr_f.c:17-26 (ref_len m, data_len n,
 *            lag_len, support)
 *   execute  vsip.h:724-725, 822-823
 *
 * C89.
 */
#include <stdlib.h>
#include "va_internal.h"

typedef struct {
    int markings;
    int prec;                  /* 4 or 8 */
    int cplx;                  /* 1 ccorr1d, 0 corr1d */
    vsip_length m, n, mn, N, lag_len;
    vsip_support_region support;
    VA_block *buf;             /* 2 x N complex: row 0 the data, row 1 the reference */
    VA_mview *rows, *row0;
    VA_fftm fft2, fft1;        /* forward, scale 1: both rows; row 0 */
    int have2, have1;
} VA_corr;

struct vsip_corr1dattributes_f { VA_corr c; };
struct vsip_corr1dattributes_d { VA_corr c; };
struct vsip_ccorr1dattributes_f { VA_corr c; };
struct vsip_ccorr1dattributes_d { VA_corr c; };

static void va_corr_free(VA_corr *c)
{
    if (!c) return;
    if (c->have2) va_fftm_release(&c->fft2);
    if (c->have1) va_fftm_release(&c->fft1);
    if (c->rows) va_mview_destroy(c->rows);
    if (c->row0) va_mview_destroy(c->row0);
    if (c->buf) va_block_destroy(c->buf);
    c->markings = VA_FREED;
    free(c);
}

/* transform length for mn = m + n - 1 lags: any length >= mn gives the same
 * circular product with no wrap into the lags kept.  The reference takes the
 * smallest power of two above mn (up to 2x mn).  This takes the smallest
 * 2^a 3^b 5^c >= mn unless that power of two is within 1.15x of it (a
 * power-of-two transform costs about 9% less per element here; MI355X,
 * m x n = 1024 x 2^20: 107 -> 70.5 us, 1000 x 10^6 keeps 2^20: 66-68 us
 * against 70.8).  VSIP_AMD_CORR_POW2=1 keeps the reference's length. */
static vsip_length va_corr_length(vsip_length mn)
{
    const char *e = getenv("VSIP_AMD_CORR_POW2");
    vsip_length pow2 = 2, best, p2, p3, p5;
    while (pow2 <= mn) pow2 *= 2;
    if (e && atoi(e) == 1) return pow2;
    best = pow2;
    for (p5 = 1; p5 < pow2; p5 *= 5)
        for (p3 = p5; p3 < pow2; p3 *= 3)
            for (p2 = p3; p2 < pow2; p2 *= 2)
                if (p2 >= mn && p2 < best) best = p2;
    return (double)pow2 <= 1.15 * (double)best ? pow2 : best;
}

static VA_corr *va_corr_new(const char *fn, int prec, int cplx, vsip_length m, vsip_length n,
                            vsip_support_region support, unsigned int ntimes, vsip_alg_hint hint)
{
    VA_corr *c;
    vsip_length N;
    if (va_ensure_init()) return NULL;
    if (m < 1 || n < m || n > ((vsip_length)1 << 40)) {
        va_error(fn, "needs 1 <= M <= N", -1);
        return NULL;
    }
    if (support != VSIP_SUPPORT_FULL && support != VSIP_SUPPORT_SAME && support != VSIP_SUPPORT_MIN) {
        va_error(fn, "unknown support region", (int)support);
        return NULL;
    }
    c = (VA_corr *)calloc(1, sizeof(VA_corr));
    if (!c) return NULL;
    c->prec = prec;
    c->cplx = cplx;
    c->m = m;
    c->n = n;
    c->mn = m + n - 1;
    c->support = support;
    c->lag_len = support == VSIP_SUPPORT_MIN ? n - m + 1 : support == VSIP_SUPPORT_SAME ? n : c->mn;
    N = c->N = va_corr_length(c->mn);
    c->buf = va_block_create(prec, 1, 2 * N, VSIP_MEM_NONE);
    if (!c->buf) goto fail;
    c->rows = va_mview_bind(c->buf, 0, (vsip_stride)N, 2, 1, N);
    c->row0 = va_mview_bind(c->buf, 0, (vsip_stride)N, 1, 1, N);
    if (!c->rows || !c->row0) goto fail;
    /* a plan whose creation failed is still released */
    c->have2 = 1;
    if (va_fftm_init(&c->fft2, prec, VA_CCFFTIP, 2, N, 1.0, VSIP_FFT_FWD, VSIP_ROW, ntimes, hint)) goto fail;
    c->have1 = 1;
    if (va_fftm_init(&c->fft1, prec, VA_CCFFTIP, 1, N, 1.0, VSIP_FFT_FWD, VSIP_ROW, ntimes, hint)) goto fail;
    c->markings = VA_VALID;
    return c;
fail:
    va_error(fn, "allocation or plan creation failed", -1);
    va_corr_free(c);
    return NULL;
}

static int va_corr_view(const char *fn, const VA_mview *v, vsip_length len)
{
    if (!v || !v->block || !v->block->admitted || !v->block->dev_r) {
        va_error(fn, "view of a released block", -1);
        return -1;
    }
    if (v->row_length != len || v->col_length != 1) {
        va_error(fn, "view length does not match the correlation object", (int)v->row_length);
        return -1;
    }
    return 0;
}

static void va_correlate(const char *fn, const VA_corr *c, int prec, int cplx, vsip_bias bias, const VA_mview *h,
                         const VA_mview *x, const VA_mview *y)
{
    vhip_mat mh, mx, my;
    void *st, *buf;
    long o;
    int rc;
    if (va_ensure_init()) return;
    if (!c || c->markings != VA_VALID || c->prec != prec || c->cplx != cplx) {
        va_error(fn, "not a correlation object of this type", -1);
        return;
    }
    if (va_corr_view(fn, h, c->m) || va_corr_view(fn, x, c->n) || va_corr_view(fn, y, c->lag_len)) return;
    va_view_mat(h, 1, &mh);
    va_view_mat(x, 1, &mx);
    va_view_mat(y, 1, &my);
    st = vhip_stream();
    buf = c->buf->dev_r;
    o = (long)(c->N - c->mn);
    if (c->support == VSIP_SUPPORT_SAME) o += (long)(c->m - 1) / 2;
    else if (c->support == VSIP_SUPPORT_MIN) o += (long)c->m - 1;
    rc = vhip_corr_pack(prec, &mx, cplx, (long)c->n, &mh, cplx, (long)c->m, buf, (long)c->N, st);
    if (!rc) rc = va_fftm_exec(&c->fft2, c->rows, c->rows, st);
    if (!rc) rc = vhip_corr_product(prec, buf, (long)c->N, st);
    if (!rc) rc = va_fftm_exec(&c->fft1, c->row0, c->row0, st);
    if (!rc)
        rc = vhip_corr_window(prec, buf, (long)c->N, o, &my, cplx, (long)c->lag_len,
                              c->support == VSIP_SUPPORT_FULL ? VHIP_CORR_FULL
                              : c->support == VSIP_SUPPORT_SAME ? VHIP_CORR_SAME : VHIP_CORR_MIN,
                              bias == VSIP_UNBIASED, (long)c->m, (long)c->n, st);
    if (rc) va_error(fn, vhip_error_string(rc), rc);
}

static void va_corr_attr(const VA_corr *c, vsip_corr1d_attr *a)
{
    if (!c || !a) return;
    a->ref_len = c->m;
    a->data_len = c->n;
    a->support = c->support;
    a->lag_len = c->lag_len;
}

/* ---- create ------------------------------------------------------------- */

vsip_corr1d_f *vsip_corr1d_create_f(vsip_length M, vsip_length N, vsip_support_region support,
                                    unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_corr1d_f *)va_corr_new("vsip_corr1d_create_f", 4, 0, M, N, support, ntimes, hint);
}

vsip_corr1d_d *vsip_corr1d_create_d(vsip_length M, vsip_length N, vsip_support_region support,
                                    unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_corr1d_d *)va_corr_new("vsip_corr1d_create_d", 8, 0, M, N, support, ntimes, hint);
}

vsip_ccorr1d_f *vsip_ccorr1d_create_f(vsip_length M, vsip_length N, vsip_support_region support,
                                      unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_ccorr1d_f *)va_corr_new("vsip_ccorr1d_create_f", 4, 1, M, N, support, ntimes, hint);
}

vsip_ccorr1d_d *vsip_ccorr1d_create_d(vsip_length M, vsip_length N, vsip_support_region support,
                                      unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_ccorr1d_d *)va_corr_new("vsip_ccorr1d_create_d", 8, 1, M, N, support, ntimes, hint);
}

/* ---- destroy ------------------------------------------------------------ */

int vsip_corr1d_destroy_f(vsip_corr1d_f *c) { va_corr_free(c ? &c->c : NULL); return 0; }
int vsip_corr1d_destroy_d(vsip_corr1d_d *c) { va_corr_free(c ? &c->c : NULL); return 0; }
int vsip_ccorr1d_destroy_f(vsip_ccorr1d_f *c) { va_corr_free(c ? &c->c : NULL); return 0; }
int vsip_ccorr1d_destroy_d(vsip_ccorr1d_d *c) { va_corr_free(c ? &c->c : NULL); return 0; }

/* ---- getattr ------------------------------------------------------------ */

void vsip_corr1d_getattr_f(const vsip_corr1d_f *c, vsip_corr1d_attr *a) { va_corr_attr(c ? &c->c : NULL, a); }
void vsip_corr1d_getattr_d(const vsip_corr1d_d *c, vsip_corr1d_attr *a) { va_corr_attr(c ? &c->c : NULL, a); }
void vsip_ccorr1d_getattr_f(const vsip_ccorr1d_f *c, vsip_corr1d_attr *a) { va_corr_attr(c ? &c->c : NULL, a); }
void vsip_ccorr1d_getattr_d(const vsip_ccorr1d_d *c, vsip_corr1d_attr *a) { va_corr_attr(c ? &c->c : NULL, a); }

/* ---- execute ------------------------------------------------------------ */

void vsip_correlate1d_f(const vsip_corr1d_f *c, vsip_bias bias, const vsip_vview_f *h, const vsip_vview_f *x,
                        const vsip_vview_f *y)
{
    va_correlate("vsip_correlate1d_f", c ? &c->c : NULL, 4, 0, bias, h ? &h->v : NULL, x ? &x->v : NULL,
                 y ? &y->v : NULL);
}

void vsip_correlate1d_d(const vsip_corr1d_d *c, vsip_bias bias, const vsip_vview_d *h, const vsip_vview_d *x,
                        const vsip_vview_d *y)
{
    va_correlate("vsip_correlate1d_d", c ? &c->c : NULL, 8, 0, bias, h ? &h->v : NULL, x ? &x->v : NULL,
                 y ? &y->v : NULL);
}

void vsip_ccorrelate1d_f(const vsip_ccorr1d_f *c, vsip_bias bias, const vsip_cvview_f *h, const vsip_cvview_f *x,
                         const vsip_cvview_f *y)
{
    va_correlate("vsip_ccorrelate1d_f", c ? &c->c : NULL, 4, 1, bias, h ? &h->v : NULL, x ? &x->v : NULL,
                 y ? &y->v : NULL);
}

void vsip_ccorrelate1d_d(const vsip_ccorr1d_d *c, vsip_bias bias, const vsip_cvview_d *h, const vsip_cvview_d *x,
                         const vsip_cvview_d *y)
{
    va_correlate("vsip_ccorrelate1d_d", c ? &c->c : NULL, 8, 1, bias, h ? &h->v : NULL, x ? &x->v : NULL,
                 y ? &y->v : NULL);
}

/* ---- convolution -----------------------------------------------------------
 * The reference (vsip_conv1d_create_f.c, vsip_convolve1d_f.c) expands the
 * kernel by its symmetry (VSIP_NONSYM: h; VSIP_SYM_EVEN_LEN_ODD: h then h
 * reversed without its last element; VSIP_SYM_EVEN_LEN_EVEN: h then h
 * reversed), transforms it at creation and convolves by FFT, in sections for
 * short kernels, keeping y[k] = z[k D + s] of z = kernel * x with s = 0
 * (FULL, (M + N - 2)/D + 1 outputs), M/2 (SAME, (N - 1)/D + 1) or M - 1
 * (MIN, (N - 1)/D - (M - 1)/D + 1).  Its sectioned form loses the lags of z
 * at and past (number of sections) x (section data length), the tail past
 * the last full section; this computes every lag.
 * Here, with real transforms of an even length L >= N + M - 1 (twice the
 * correlation's length rule for half of it): at creation the kernel,
 * zero-padded, goes through a real-to-complex transform to its half
 * spectrum K (L/2 + 1 bins).  A call packs x into a real row, transforms it
 * to X, forms X K / L on the half spectrum, and its complex-to-real
 * transform (the inverse sign, scale 1) is z, whose decimated lags go to y:
 * about half the transform work of complex rows.  create returns NULL for an empty kernel, D < 1, an unknown
 * symmetry or support, or a MIN support with no output.
 *   create   vsip.h:1795-1796      destroy  vsip.h:405-406 (returns 0)
 *   getattr  vsip_conv1d_getattr_f.c:17-28     execute  vsip.h:818-819 */

typedef struct {
    int markings;
    int prec;
    vsip_length m, n, L, out_len;   /* kernel length (expanded), data length, transform (even), outputs */
    int D;
    vsip_symmetry symm;
    vsip_support_region support;
    VA_block *rbuf;                 /* L reals: the padded data, then the convolution */
    VA_block *sbuf;                 /* 2 x (L/2 + 1) complex: the data's half spectrum, the kernel's */
    VA_mview *rrow, *s0, *s1;
    VA_fftm rc, cr;                 /* real to complex, complex to real, scale 1 */
    int have_rc, have_cr;
} VA_conv;

struct vsip_conv1dattributes_f { VA_conv c; };
struct vsip_conv1dattributes_d { VA_conv c; };

static void va_conv_free(VA_conv *c)
{
    if (!c) return;
    if (c->have_rc) va_fftm_release(&c->rc);
    if (c->have_cr) va_fftm_release(&c->cr);
    if (c->rrow) va_mview_destroy(c->rrow);
    if (c->s0) va_mview_destroy(c->s0);
    if (c->s1) va_mview_destroy(c->s1);
    if (c->rbuf) va_block_destroy(c->rbuf);
    if (c->sbuf) va_block_destroy(c->sbuf);
    c->markings = VA_FREED;
    free(c);
}

static VA_conv *va_conv_new(const char *fn, int prec, const VA_mview *h, vsip_symmetry symm, vsip_length n, int D,
                            vsip_support_region support, unsigned int ntimes, vsip_alg_hint hint)
{
    VA_conv *c;
    vsip_length hl, m, L, nb, k;
    double *kern = NULL;
    void *up = NULL;
    int rc;
    if (va_ensure_init()) return NULL;
    if (!h || !h->block || h->col_length != 1 || h->row_length < 1 || n < 1 || D < 1 ||
        n > ((vsip_length)1 << 40)) {
        va_error(fn, "needs a kernel view, N >= 1 and D >= 1", -1);
        return NULL;
    }
    if (symm != VSIP_NONSYM && symm != VSIP_SYM_EVEN_LEN_ODD && symm != VSIP_SYM_EVEN_LEN_EVEN) {
        va_error(fn, "unknown symmetry", (int)symm);
        return NULL;
    }
    if (support != VSIP_SUPPORT_FULL && support != VSIP_SUPPORT_SAME && support != VSIP_SUPPORT_MIN) {
        va_error(fn, "unknown support region", (int)support);
        return NULL;
    }
    hl = h->row_length;
    m = symm == VSIP_SYM_EVEN_LEN_ODD ? 2 * hl - 1 : symm == VSIP_SYM_EVEN_LEN_EVEN ? 2 * hl : hl;
    if (support == VSIP_SUPPORT_MIN && (m > n || (n - 1) / (vsip_length)D < (m - 1) / (vsip_length)D)) {
        va_error(fn, "MIN support with no output (kernel longer than the data)", -1);
        return NULL;
    }
    c = (VA_conv *)calloc(1, sizeof(VA_conv));
    if (!c) return NULL;
    c->prec = prec;
    c->m = m;
    c->n = n;
    c->D = D;
    c->symm = symm;
    c->support = support;
    c->out_len = support == VSIP_SUPPORT_FULL ? (m + n - 2) / (vsip_length)D + 1
                 : support == VSIP_SUPPORT_SAME ? (n - 1) / (vsip_length)D + 1
                 : (n - 1) / (vsip_length)D - (m - 1) / (vsip_length)D + 1;
    /* an even length >= m + n - 1 for the real transforms */
    L = c->L = 2 * va_corr_length((m + n) / 2);
    nb = L / 2 + 1;
    c->rbuf = va_block_create(prec, 0, L, VSIP_MEM_NONE);
    c->sbuf = va_block_create(prec, 1, 2 * nb, VSIP_MEM_NONE);
    if (!c->rbuf || !c->sbuf) goto fail;
    c->rrow = va_mview_bind(c->rbuf, 0, (vsip_stride)L, 1, 1, L);
    c->s0 = va_mview_bind(c->sbuf, 0, (vsip_stride)nb, 1, 1, nb);
    c->s1 = va_mview_bind(c->sbuf, nb, (vsip_stride)nb, 1, 1, nb);
    if (!c->rrow || !c->s0 || !c->s1) goto fail;
    /* a plan whose creation failed is still released */
    c->have_rc = 1;
    if (va_fftm_init(&c->rc, prec, VA_RCFFTOP, 1, L, 1.0, VSIP_FFT_FWD, VSIP_ROW, ntimes, hint)) goto fail;
    c->have_cr = 1;
    if (va_fftm_init(&c->cr, prec, VA_CRFFTOP, 1, L, 1.0, VSIP_FFT_INV, VSIP_ROW, ntimes, hint)) goto fail;
    /* the kernel, expanded on the host and zero-padded to L, then its half
     * spectrum into row 1 of the spectrum buffer */
    kern = (double *)calloc((size_t)L, sizeof(double));
    up = malloc((size_t)prec * (size_t)L);
    if (!kern || !up) goto fail;
    {
        void *hv = malloc((size_t)h->block->prec * (size_t)hl);
        if (!hv) goto fail;
        if (va_mview_copy_host(h, 0, hv, NULL, VSIP_ROW)) { free(hv); goto fail; }
        for (k = 0; k < hl; k++)
            kern[k] = h->block->prec == 4 ? (double)((float *)hv)[k] : ((double *)hv)[k];
        free(hv);
    }
    for (k = hl; k < m; k++) kern[k] = kern[symm == VSIP_SYM_EVEN_LEN_ODD ? 2 * hl - 2 - k : 2 * hl - 1 - k];
    for (k = 0; k < L; k++) {
        if (prec == 4) ((float *)up)[k] = (float)kern[k];
        else ((double *)up)[k] = kern[k];
    }
    rc = vhip_memcpy_h2d(c->rbuf->dev_r, up, (size_t)prec * (size_t)L);
    if (rc || va_fftm_exec(&c->rc, c->rrow, c->s1, vhip_stream())) goto fail;
    free(kern);
    free(up);
    c->markings = VA_VALID;
    return c;
fail:
    va_error(fn, "allocation, kernel copy or plan creation failed", -1);
    free(kern);
    free(up);
    va_conv_free(c);
    return NULL;
}

static void va_convolve(const char *fn, const VA_conv *c, int prec, const VA_mview *x, const VA_mview *y)
{
    vhip_mat mx, my;
    void *st, *r, *spec;
    long s0, nb;
    int rc;
    if (va_ensure_init()) return;
    if (!c || c->markings != VA_VALID || c->prec != prec) {
        va_error(fn, "not a convolution object of this type", -1);
        return;
    }
    if (va_corr_view(fn, x, c->n) || va_corr_view(fn, y, c->out_len)) return;
    va_view_mat(x, 1, &mx);
    va_view_mat(y, 1, &my);
    st = vhip_stream();
    r = c->rbuf->dev_r;
    spec = c->sbuf->dev_r;
    nb = (long)c->L / 2 + 1;
    s0 = c->support == VSIP_SUPPORT_FULL ? 0 : c->support == VSIP_SUPPORT_SAME ? (long)(c->m / 2) : (long)c->m - 1;
    rc = vhip_conv_pack_r(prec, &mx, (long)c->n, r, (long)c->L, st);
    if (!rc) rc = va_fftm_exec(&c->rc, c->rrow, c->s0, st);
    if (!rc) rc = vhip_conv_mul_half(prec, spec, (char *)spec + (size_t)nb * 2 * (size_t)prec, (long)c->L, st);
    if (!rc) rc = va_fftm_exec(&c->cr, c->s0, c->rrow, st);
    if (!rc) rc = vhip_conv_window_r(prec, r, (long)c->L, s0, (long)c->D, &my, (long)c->out_len, st);
    if (rc) va_error(fn, vhip_error_string(rc), rc);
}

static void va_conv_attr(const VA_conv *c, vsip_conv1d_attr *a)
{
    if (!c || !a) return;
    a->kernel_len = c->m;
    a->symm = c->symm;
    a->data_len = c->n;
    a->support = c->support;
    a->out_len = c->out_len;
    a->decimation = c->D;
}

vsip_conv1d_f *vsip_conv1d_create_f(const vsip_vview_f *h, vsip_symmetry symm, vsip_length N, int D,
                                    vsip_support_region support, unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_conv1d_f *)va_conv_new("vsip_conv1d_create_f", 4, h ? &h->v : NULL, symm, N, D, support, ntimes,
                                        hint);
}

vsip_conv1d_d *vsip_conv1d_create_d(const vsip_vview_d *h, vsip_symmetry symm, vsip_length N, int D,
                                    vsip_support_region support, unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_conv1d_d *)va_conv_new("vsip_conv1d_create_d", 8, h ? &h->v : NULL, symm, N, D, support, ntimes,
                                        hint);
}

int vsip_conv1d_destroy_f(vsip_conv1d_f *c) { va_conv_free(c ? &c->c : NULL); return 0; }
int vsip_conv1d_destroy_d(vsip_conv1d_d *c) { va_conv_free(c ? &c->c : NULL); return 0; }
void vsip_conv1d_getattr_f(const vsip_conv1d_f *c, vsip_conv1d_attr *a) { va_conv_attr(c ? &c->c : NULL, a); }
void vsip_conv1d_getattr_d(const vsip_conv1d_d *c, vsip_conv1d_attr *a) { va_conv_attr(c ? &c->c : NULL, a); }

void vsip_convolve1d_f(const vsip_conv1d_f *c, const vsip_vview_f *x, const vsip_vview_f *y)
{
    va_convolve("vsip_convolve1d_f", c ? &c->c : NULL, 4, x ? &x->v : NULL, y ? &y->v : NULL);
}

void vsip_convolve1d_d(const vsip_conv1d_d *c, const vsip_vview_d *x, const vsip_vview_d *y)
{
    va_convolve("vsip_convolve1d_d", c ? &c->c : NULL, 8, x ? &x->v : NULL, y ? &y->v : NULL);
}
