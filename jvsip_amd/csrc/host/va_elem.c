/*
 * va_elem.c -- element operations on FFT operands: real-scalar times complex
 * view, complex difference and complex mean square value (the operations
 * the reference's own tests of this path apply to transform results,
 * c_VSIP_testing/ccfftmip_f.h and cblock_admit_release_f.h).
 * Reference definitions (rrjudd/jvsip c_VSIP_src/):
 *   rscvmul / rscmmul   vsip.h:1207-1208, 1215-1216: r = alpha * a
 *   cvsub / cmsub       vsip.h:804-805, 940-941: c = a - b
 *   c[mv]meansqval      vsip_cmmeansqval_f.c:18-55: sum |a|^2 / elements
 * and the real matrix operations of the reference's k-omega beamformer
 * (examples/komegaExamples/kw_c/kw.c:7-25):
 *   vmmul               vsip_vmmul_f.c: R = a (.) B, a along each row (ROW)
 *                       or down each column (COL)
 *   mcmagsq             vsip_mcmagsq_f.c: r = re^2 + im^2
 *   smmul, madd, mfill  vsip_smmul_f.c, vsip_madd_f.c, vsip_mfill_f.c
 *   vcreate_hanning     vsip_vcreate_hanning_f.c: w[n-1] = 0.5 (1 - cos(n t)),
 *                       t = 2 pi / (N + 1), n = 1..N, in the reference's
 *                       precision steps (float t and product, double cos)
 * Operands must be views of admitted blocks (their data is in HBM); the
 * operations run on the library stream, and meansqval returns after the
 * device reduction.  The real operations round every product and sum
 * separately, like the reference's scalar loops, so their results are the
 * reference's bit for bit.
 *
 * C89.
 */
#include <math.h>
#include <stdlib.h>
#include "va_internal.h"

static int va_dev_views(const char *fn, const VA_mview *a, const VA_mview *b, const VA_mview *c)
{
    const VA_mview *v[3];
    int k;
    v[0] = a; v[1] = b; v[2] = c;
    for (k = 0; k < 3; k++) {
        if (!v[k]) continue;
        if (!v[k]->block || !v[k]->block->admitted || !v[k]->block->dev_r) {
            va_error(fn, "view of a released block", -1);
            return -1;
        }
        if (v[k]->row_length != a->row_length || v[k]->col_length != a->col_length) {
            va_error(fn, "view shapes differ", -1);
            return -1;
        }
    }
    return 0;
}

static void va_scale(const char *fn, double alpha, const VA_mview *a, const VA_mview *r)
{
    vhip_mat ma, mr;
    int rc;
    if (va_ensure_init() || va_dev_views(fn, a, r, NULL)) return;
    va_view_mat(a, 1, &ma);
    va_view_mat(r, 1, &mr);
    rc = vhip_elem_scale(a->block->prec, &ma, &mr, (long)a->col_length, (long)a->row_length, alpha, vhip_stream());
    if (rc) va_error(fn, vhip_error_string(rc), rc);
}

static void va_sub(const char *fn, const VA_mview *a, const VA_mview *b, const VA_mview *c)
{
    vhip_mat ma, mb, mc;
    int rc;
    if (va_ensure_init() || va_dev_views(fn, a, b, c)) return;
    va_view_mat(a, 1, &ma);
    va_view_mat(b, 1, &mb);
    va_view_mat(c, 1, &mc);
    rc = vhip_elem_sub(a->block->prec, &ma, &mb, &mc, (long)a->col_length, (long)a->row_length, vhip_stream());
    if (rc) va_error(fn, vhip_error_string(rc), rc);
}

static double va_meansq(const char *fn, const VA_mview *a)
{
    vhip_mat ma;
    double s = 0.0;
    long n;
    int rc;
    if (va_ensure_init() || va_dev_views(fn, a, NULL, NULL)) return 0.0;
    n = (long)a->col_length * (long)a->row_length;
    if (n == 0) return 0.0;
    va_view_mat(a, 1, &ma);
    rc = vhip_elem_sumsq(a->block->prec, &ma, (long)a->col_length, (long)a->row_length, &s, vhip_stream());
    if (rc) { va_error(fn, vhip_error_string(rc), rc); return 0.0; }
    return s / (double)n;
}

void vsip_rscvmul_f(vsip_scalar_f alpha, const vsip_cvview_f *a, const vsip_cvview_f *r)
{
    va_scale("vsip_rscvmul_f", alpha, &a->v, &r->v);
}

void vsip_rscvmul_d(vsip_scalar_d alpha, const vsip_cvview_d *a, const vsip_cvview_d *r)
{
    va_scale("vsip_rscvmul_d", alpha, &a->v, &r->v);
}

void vsip_rscmmul_f(vsip_scalar_f alpha, const vsip_cmview_f *a, const vsip_cmview_f *r)
{
    va_scale("vsip_rscmmul_f", alpha, &a->v, &r->v);
}

void vsip_rscmmul_d(vsip_scalar_d alpha, const vsip_cmview_d *a, const vsip_cmview_d *r)
{
    va_scale("vsip_rscmmul_d", alpha, &a->v, &r->v);
}

void vsip_cvsub_f(const vsip_cvview_f *a, const vsip_cvview_f *b, const vsip_cvview_f *c)
{
    va_sub("vsip_cvsub_f", &a->v, &b->v, &c->v);
}

void vsip_cvsub_d(const vsip_cvview_d *a, const vsip_cvview_d *b, const vsip_cvview_d *c)
{
    va_sub("vsip_cvsub_d", &a->v, &b->v, &c->v);
}

void vsip_cmsub_f(const vsip_cmview_f *a, const vsip_cmview_f *b, const vsip_cmview_f *c)
{
    va_sub("vsip_cmsub_f", &a->v, &b->v, &c->v);
}

void vsip_cmsub_d(const vsip_cmview_d *a, const vsip_cmview_d *b, const vsip_cmview_d *c)
{
    va_sub("vsip_cmsub_d", &a->v, &b->v, &c->v);
}

vsip_scalar_f vsip_cvmeansqval_f(const vsip_cvview_f *a)
{
    return (vsip_scalar_f)va_meansq("vsip_cvmeansqval_f", &a->v);
}

vsip_scalar_d vsip_cvmeansqval_d(const vsip_cvview_d *a)
{
    return va_meansq("vsip_cvmeansqval_d", &a->v);
}

vsip_scalar_f vsip_cmmeansqval_f(const vsip_cmview_f *a)
{
    return (vsip_scalar_f)va_meansq("vsip_cmmeansqval_f", &a->v);
}

vsip_scalar_d vsip_cmmeansqval_d(const vsip_cmview_d *a)
{
    return va_meansq("vsip_cmmeansqval_d", &a->v);
}

/* ---- real matrix operations -------------------------------------------- */

static void va_real_op(const char *fn, int op, const VA_mview *a, const VA_mview *b, const VA_mview *r,
                       const VA_mview *vec, double alpha)
{
    vhip_mat ma, mb, mr, mv;
    int rc;
    if (va_ensure_init() || va_dev_views(fn, r, a, b)) return;
    if (a) va_view_mat(a, 1, &ma);
    if (b) va_view_mat(b, 1, &mb);
    va_view_mat(r, 1, &mr);
    if (vec) {
        vsip_length want = (op == VHIP_EOP_VMUL_ROW) ? r->row_length : r->col_length;
        if (!vec->block || !vec->block->admitted || !vec->block->dev_r) {
            va_error(fn, "view of a released block", -1);
            return;
        }
        if (vec->row_length != want) {
            va_error(fn, "vector length does not match the matrix", -1);
            return;
        }
        va_view_mat(vec, 1, &mv);
    }
    rc = vhip_elem_real(r->block->prec, op, a ? &ma : NULL, b ? &mb : NULL, &mr, vec ? mv.re : NULL,
                        vec ? mv.rs : 0, alpha, (long)r->col_length, (long)r->row_length, vhip_stream());
    if (rc) va_error(fn, vhip_error_string(rc), rc);
}

void vsip_vmmul_f(const vsip_vview_f *a, const vsip_mview_f *B, vsip_major major, const vsip_mview_f *R)
{
    va_real_op("vsip_vmmul_f", major == VSIP_ROW ? VHIP_EOP_VMUL_ROW : VHIP_EOP_VMUL_COL, &B->v, NULL, &R->v,
               &a->v, 0.0);
}

void vsip_vmmul_d(const vsip_vview_d *a, const vsip_mview_d *B, vsip_major major, const vsip_mview_d *R)
{
    va_real_op("vsip_vmmul_d", major == VSIP_ROW ? VHIP_EOP_VMUL_ROW : VHIP_EOP_VMUL_COL, &B->v, NULL, &R->v,
               &a->v, 0.0);
}

void vsip_mcmagsq_f(const vsip_cmview_f *a, const vsip_mview_f *r)
{
    va_real_op("vsip_mcmagsq_f", VHIP_EOP_MAGSQ, &a->v, NULL, &r->v, NULL, 0.0);
}

void vsip_mcmagsq_d(const vsip_cmview_d *a, const vsip_mview_d *r)
{
    va_real_op("vsip_mcmagsq_d", VHIP_EOP_MAGSQ, &a->v, NULL, &r->v, NULL, 0.0);
}

void vsip_smmul_f(vsip_scalar_f alpha, const vsip_mview_f *a, const vsip_mview_f *r)
{
    va_real_op("vsip_smmul_f", VHIP_EOP_SMUL, &a->v, NULL, &r->v, NULL, alpha);
}

void vsip_smmul_d(vsip_scalar_d alpha, const vsip_mview_d *a, const vsip_mview_d *r)
{
    va_real_op("vsip_smmul_d", VHIP_EOP_SMUL, &a->v, NULL, &r->v, NULL, alpha);
}

void vsip_madd_f(const vsip_mview_f *a, const vsip_mview_f *b, const vsip_mview_f *r)
{
    va_real_op("vsip_madd_f", VHIP_EOP_ADD, &a->v, &b->v, &r->v, NULL, 0.0);
}

void vsip_madd_d(const vsip_mview_d *a, const vsip_mview_d *b, const vsip_mview_d *r)
{
    va_real_op("vsip_madd_d", VHIP_EOP_ADD, &a->v, &b->v, &r->v, NULL, 0.0);
}

void vsip_mfill_f(vsip_scalar_f alpha, const vsip_mview_f *r)
{
    va_real_op("vsip_mfill_f", VHIP_EOP_FILL, NULL, NULL, &r->v, NULL, alpha);
}

void vsip_mfill_d(vsip_scalar_d alpha, const vsip_mview_d *r)
{
    va_real_op("vsip_mfill_d", VHIP_EOP_FILL, NULL, NULL, &r->v, NULL, alpha);
}

/* 2 pi as the reference writes it (vsip_vcreate_hanning_f.c:21) */
#define VA_HANN_TWOPI 6.2831853071796

vsip_vview_f *vsip_vcreate_hanning_f(vsip_length N, vsip_memory_hint h)
{
    vsip_vview_f *v = vsip_vcreate_f(N, h);
    float *w, t = (float)VA_HANN_TWOPI / (float)(N + 1);
    vsip_length n;
    if (!v || N == 0) return v;
    w = (float *)malloc(sizeof(float) * N);
    if (!w) { vsip_valldestroy_f(v); return NULL; }
    for (n = 1; n <= N; n++) {
        volatile float x = t * (float)n;      /* rounded to float before cos, as in the reference */
        w[n - 1] = 0.5f * (1.0f - (float)cos((double)x));
    }
    vsip_vcopyfrom_user_f(w, v);
    free(w);
    return v;
}

vsip_vview_d *vsip_vcreate_hanning_d(vsip_length N, vsip_memory_hint h)
{
    vsip_vview_d *v = vsip_vcreate_d(N, h);
    double *w, t = VA_HANN_TWOPI / (double)(N + 1);
    vsip_length n;
    if (!v || N == 0) return v;
    w = (double *)malloc(sizeof(double) * N);
    if (!w) { vsip_valldestroy_d(v); return NULL; }
    for (n = 1; n <= N; n++) w[n - 1] = 0.5 * (1.0 - cos(t * (double)n));
    vsip_vcopyfrom_user_d(w, v);
    free(w);
    return v;
}
