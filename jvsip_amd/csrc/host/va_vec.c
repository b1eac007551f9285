/*
 * va_vec.c -- VSIPL vector views (the 1-D FFT's operands).
 *
 * A vector view (offset, stride, length) over a block is held as the
 * 1 x length matrix view (offset, col_stride length*stride, 1, row_stride
 * stride, length), so every storage path of the matrix views (device data
 * of admitted blocks, host data of released user blocks, gets and puts,
 * user copies, the transforms' view arithmetic) serves vectors unchanged.
 * Reference behaviour (file:line in rrjudd/jvsip c_VSIP_src/):
 *   bind / create   vsip.h:491, 503, 488, 501, 514, 517, 512, 515 (create:
 *                   unit stride over a new block of the view's length)
 *   destroy         vsip.h:304, 296, 360, 352 (returns the block)
 *   alldestroy      vsip.h:1292, 1291, 865, 864 (destroys the block too)
 *   get / put       vsip.h:638, 605, 1556, 1555, 568, 539, 915, 914
 *   getattrib       vsip.h:1422, 1421, 889, 888; attributes vsip.h:231-234
 *   copy from/to user  vsip.h:1378, 1377, 1384, 1383, 875, 874, 877, 876
 *                   (complex: interleaved when the imaginary pointer is NULL)
 *
 * C89.
 */
#include <stdlib.h>
#include "va_internal.h"

static VA_mview *va_vbind(VA_block *b, vsip_offset o, vsip_stride s, vsip_length n)
{
    if (!b) return NULL;
    return va_mview_bind(b, o, (vsip_stride)n * s, 1, s, n);
}

static VA_mview *va_vcreate(int prec, int cplx, vsip_length n, vsip_memory_hint h)
{
    VA_block *b = va_block_create(prec, cplx, n, h);
    VA_mview *v;
    if (!b) return NULL;
    v = va_vbind(b, 0, 1, n);
    if (!v) va_block_destroy(b);
    return v;
}

static void va_vattr(const VA_mview *v, vsip_offset *o, vsip_stride *s, vsip_length *n)
{
    *o = v->offset;
    *s = v->row_stride;
    *n = v->row_length;
}

/* ---- real vectors --------------------------------------------------------- */

vsip_vview_f *vsip_vbind_f(const vsip_block_f *b, vsip_offset o, vsip_stride s, vsip_length n)
{
    return (vsip_vview_f *)va_vbind(b ? (VA_block *)&b->b : NULL, o, s, n);
}

vsip_vview_d *vsip_vbind_d(const vsip_block_d *b, vsip_offset o, vsip_stride s, vsip_length n)
{
    return (vsip_vview_d *)va_vbind(b ? (VA_block *)&b->b : NULL, o, s, n);
}

vsip_vview_f *vsip_vcreate_f(vsip_length n, vsip_memory_hint h)
{
    return (vsip_vview_f *)va_vcreate(4, 0, n, h);
}

vsip_vview_d *vsip_vcreate_d(vsip_length n, vsip_memory_hint h)
{
    return (vsip_vview_d *)va_vcreate(8, 0, n, h);
}

vsip_block_f *vsip_vdestroy_f(vsip_vview_f *v)
{
    vsip_block_f *b = v ? (vsip_block_f *)v->v.block : NULL;
    va_mview_destroy(v ? &v->v : NULL);
    return b;
}

vsip_block_d *vsip_vdestroy_d(vsip_vview_d *v)
{
    vsip_block_d *b = v ? (vsip_block_d *)v->v.block : NULL;
    va_mview_destroy(v ? &v->v : NULL);
    return b;
}

void vsip_valldestroy_f(vsip_vview_f *v)
{
    if (v) va_block_destroy((VA_block *)vsip_vdestroy_f(v));
}

void vsip_valldestroy_d(vsip_vview_d *v)
{
    if (v) va_block_destroy((VA_block *)vsip_vdestroy_d(v));
}

vsip_scalar_f vsip_vget_f(const vsip_vview_f *v, vsip_index i)
{
    vsip_scalar_f x = 0;
    va_get(&v->v, 0, i, &x, NULL);
    return x;
}

vsip_scalar_d vsip_vget_d(const vsip_vview_d *v, vsip_index i)
{
    vsip_scalar_d x = 0;
    va_get(&v->v, 0, i, &x, NULL);
    return x;
}

void vsip_vput_f(const vsip_vview_f *v, vsip_index i, vsip_scalar_f x)
{
    va_put(&v->v, 0, i, &x, NULL);
}

void vsip_vput_d(const vsip_vview_d *v, vsip_index i, vsip_scalar_d x)
{
    va_put(&v->v, 0, i, &x, NULL);
}

void vsip_vgetattrib_f(const vsip_vview_f *v, vsip_vattr_f *a)
{
    va_vattr(&v->v, &a->offset, &a->stride, &a->length);
    a->block = (vsip_block_f *)v->v.block;
}

void vsip_vgetattrib_d(const vsip_vview_d *v, vsip_vattr_d *a)
{
    va_vattr(&v->v, &a->offset, &a->stride, &a->length);
    a->block = (vsip_block_d *)v->v.block;
}

void vsip_vcopyfrom_user_f(vsip_scalar_f *const r, const vsip_vview_f *v)
{
    va_mview_copy_host(&v->v, 1, (void *)r, NULL, VSIP_ROW);
}

void vsip_vcopyfrom_user_d(vsip_scalar_d *const r, const vsip_vview_d *v)
{
    va_mview_copy_host(&v->v, 1, (void *)r, NULL, VSIP_ROW);
}

void vsip_vcopyto_user_f(const vsip_vview_f *v, vsip_scalar_f *const r)
{
    va_mview_copy_host(&v->v, 0, (void *)r, NULL, VSIP_ROW);
}

void vsip_vcopyto_user_d(const vsip_vview_d *v, vsip_scalar_d *const r)
{
    va_mview_copy_host(&v->v, 0, (void *)r, NULL, VSIP_ROW);
}

/* ---- complex vectors ------------------------------------------------------ */

vsip_cvview_f *vsip_cvbind_f(const vsip_cblock_f *b, vsip_offset o, vsip_stride s, vsip_length n)
{
    return (vsip_cvview_f *)va_vbind(b ? (VA_block *)&b->b : NULL, o, s, n);
}

vsip_cvview_d *vsip_cvbind_d(const vsip_cblock_d *b, vsip_offset o, vsip_stride s, vsip_length n)
{
    return (vsip_cvview_d *)va_vbind(b ? (VA_block *)&b->b : NULL, o, s, n);
}

vsip_cvview_f *vsip_cvcreate_f(vsip_length n, vsip_memory_hint h)
{
    return (vsip_cvview_f *)va_vcreate(4, 1, n, h);
}

vsip_cvview_d *vsip_cvcreate_d(vsip_length n, vsip_memory_hint h)
{
    return (vsip_cvview_d *)va_vcreate(8, 1, n, h);
}

vsip_cblock_f *vsip_cvdestroy_f(vsip_cvview_f *v)
{
    vsip_cblock_f *b = v ? (vsip_cblock_f *)v->v.block : NULL;
    va_mview_destroy(v ? &v->v : NULL);
    return b;
}

vsip_cblock_d *vsip_cvdestroy_d(vsip_cvview_d *v)
{
    vsip_cblock_d *b = v ? (vsip_cblock_d *)v->v.block : NULL;
    va_mview_destroy(v ? &v->v : NULL);
    return b;
}

void vsip_cvalldestroy_f(vsip_cvview_f *v)
{
    if (v) va_block_destroy((VA_block *)vsip_cvdestroy_f(v));
}

void vsip_cvalldestroy_d(vsip_cvview_d *v)
{
    if (v) va_block_destroy((VA_block *)vsip_cvdestroy_d(v));
}

vsip_cscalar_f vsip_cvget_f(const vsip_cvview_f *v, vsip_index i)
{
    vsip_cscalar_f z;
    z.r = 0; z.i = 0;
    va_get(&v->v, 0, i, &z.r, &z.i);
    return z;
}

vsip_cscalar_d vsip_cvget_d(const vsip_cvview_d *v, vsip_index i)
{
    vsip_cscalar_d z;
    z.r = 0; z.i = 0;
    va_get(&v->v, 0, i, &z.r, &z.i);
    return z;
}

void vsip_cvput_f(const vsip_cvview_f *v, vsip_index i, vsip_cscalar_f z)
{
    va_put(&v->v, 0, i, &z.r, &z.i);
}

void vsip_cvput_d(const vsip_cvview_d *v, vsip_index i, vsip_cscalar_d z)
{
    va_put(&v->v, 0, i, &z.r, &z.i);
}

void vsip_cvgetattrib_f(const vsip_cvview_f *v, vsip_cvattr_f *a)
{
    va_vattr(&v->v, &a->offset, &a->stride, &a->length);
    a->block = (vsip_cblock_f *)v->v.block;
}

void vsip_cvgetattrib_d(const vsip_cvview_d *v, vsip_cvattr_d *a)
{
    va_vattr(&v->v, &a->offset, &a->stride, &a->length);
    a->block = (vsip_cblock_d *)v->v.block;
}

void vsip_cvcopyfrom_user_f(vsip_scalar_f *const r, vsip_scalar_f *const i, const vsip_cvview_f *v)
{
    va_mview_copy_host(&v->v, 1, (void *)r, (void *)i, VSIP_ROW);
}

void vsip_cvcopyfrom_user_d(vsip_scalar_d *const r, vsip_scalar_d *const i, const vsip_cvview_d *v)
{
    va_mview_copy_host(&v->v, 1, (void *)r, (void *)i, VSIP_ROW);
}

void vsip_cvcopyto_user_f(const vsip_cvview_f *v, vsip_scalar_f *const r, vsip_scalar_f *const i)
{
    va_mview_copy_host(&v->v, 0, (void *)r, (void *)i, VSIP_ROW);
}

void vsip_cvcopyto_user_d(const vsip_cvview_d *v, vsip_scalar_d *const r, vsip_scalar_d *const i)
{
    va_mview_copy_host(&v->v, 0, (void *)r, (void *)i, VSIP_ROW);
}
