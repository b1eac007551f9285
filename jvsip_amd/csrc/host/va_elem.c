/*
 * va_elem.c -- element operations on FFT operands: real-scalar times complex
 * view, complex difference and complex mean square value (the operations
 * the reference's own tests of this path apply to transform results,
 * c_VSIP_testing/ccfftmip_f.h and cblock_admit_release_f.h).
 * Reference definitions (rrjudd/jvsip c_VSIP_src/):
 *   rscvmul / rscmmul   vsip.h:1207-1208, 1215-1216: r = alpha * a
 *   cvsub / cmsub       vsip.h:804-805, 940-941: c = a - b
 *   c[mv]meansqval      vsip_cmmeansqval_f.c:18-55: sum |a|^2 / elements
 * Operands must be views of admitted blocks (their data is in HBM); the
 * operations run on the library stream, and meansqval returns after the
 * device reduction.
 *
 * C89.
 */
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
