/*
 * va_view.c -- VSIPL matrix views (real and complex, float and double).
 *
 * A view is the reference's stride descriptor (vsip_cmviewattributes_f.h:47-55):
 * element (i, j) is block element offset + i * col_stride + j * row_stride.
 *   bind      vsip_cmbind_f.c:18-36
 *   create    vsip_cmcreate_f.c:19-39 (ROW: col_stride n, row_stride 1;
 *                                     COL: col_stride 1, row_stride m)
 *   subview   vsip_cmsubview_f.c:18-30
 *   copyfrom/copyto_user  vsip_cmcopyfrom_user_f.c:18-90 (user data dense in
 *             ROW or COL order, interleaved when the imaginary pointer is NULL)
 * Data of an admitted block is in HBM; gets, puts and user copies move it
 * with synchronous device copies.  A released user block's data is the
 * user's own memory, and gets/puts then touch that memory directly.
 *
 * C89.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "va_internal.h"

VA_mview *va_mview_bind(VA_block *b, vsip_offset o, vsip_stride cs, vsip_length cl,
                        vsip_stride rs, vsip_length rl)
{
    VA_mview *v = (VA_mview *)malloc(sizeof(VA_mview));
    if (!v) return NULL;
    v->block = b;
    v->offset = o;
    v->col_stride = cs;
    v->col_length = cl;
    v->row_stride = rs;
    v->row_length = rl;
    v->markings = VA_VALID;
    return v;
}

void va_mview_destroy(VA_mview *v)
{
    if (v) {
        v->markings = VA_FREED;
        free(v);
    }
}

static VA_mview *va_mview_create(int prec, int cplx, vsip_length m, vsip_length n,
                                 vsip_major major, vsip_memory_hint h)
{
    VA_block *b = va_block_create(prec, cplx, m * n, h);
    VA_mview *v;
    if (!b) return NULL;
    if (major == VSIP_ROW) v = va_mview_bind(b, 0, (vsip_stride)n, m, 1, n);
    else v = va_mview_bind(b, 0, 1, m, (vsip_stride)m, n);
    if (!v) va_block_destroy(b);
    return v;
}

static long va_elem_index(const VA_mview *v, vsip_index i, vsip_index j)
{
    return (long)v->offset + (long)i * v->col_stride + (long)j * v->row_stride;
}

/* device (admitted) or host (released user block) address of element (i, j) */
static int va_elem_addr(const VA_mview *v, vsip_index i, vsip_index j, void **r, void **im,
                        int *on_device)
{
    VA_block *b = v->block;
    long k = va_elem_index(v, i, j);
    char *base_r, *base_i;
    size_t p = (size_t)b->prec;
    *on_device = b->admitted;
    base_r = (char *)(b->admitted ? b->dev_r : b->user_r);
    base_i = (char *)(b->admitted ? b->dev_i : b->user_i);
    if (!base_r) return 1;
    if (b->cplx && !b->split) {
        *r = base_r + (size_t)k * 2 * p;
        *im = base_r + (size_t)k * 2 * p + p;
    } else {
        *r = base_r + (size_t)k * (size_t)VA_RS(b) * p;
        *im = b->cplx ? (void *)(base_i + (size_t)k * p) : NULL;
    }
    return 0;
}

int va_elem_dev(const VA_mview *v, vsip_index i, vsip_index j, void **r, void **im)
{
    int dev;
    return va_elem_addr(v, i, j, r, im, &dev);
}

void va_get(const VA_mview *v, vsip_index i, vsip_index j, void *re, void *im)
{
    void *pr, *pi;
    int dev, rc = 0;
    size_t p = (size_t)v->block->prec;
    if (va_elem_addr(v, i, j, &pr, &pi, &dev)) { va_error("get", "block has no data", 0); return; }
    if (dev) {
        if (v->block->cplx && !v->block->split) {
            char buf[16];
            rc = vhip_memcpy_d2h(buf, pr, 2 * p);
            memcpy(re, buf, p);
            if (im) memcpy(im, buf + p, p);
        } else {
            rc = vhip_memcpy_d2h(re, pr, p);
            if (!rc && im && pi) rc = vhip_memcpy_d2h(im, pi, p);
        }
        if (rc) va_error("get", "device-to-host copy failed", rc);
    } else {
        memcpy(re, pr, p);
        if (im && pi) memcpy(im, pi, p);
    }
}

void va_put(const VA_mview *v, vsip_index i, vsip_index j, const void *re, const void *im)
{
    void *pr, *pi;
    int dev, rc = 0;
    size_t p = (size_t)v->block->prec;
    if (va_elem_addr(v, i, j, &pr, &pi, &dev)) { va_error("put", "block has no data", 0); return; }
    if (dev) {
        if (v->block->cplx && !v->block->split) {
            char buf[16];
            memcpy(buf, re, p);
            memcpy(buf + p, im, p);
            rc = vhip_memcpy_h2d(pr, buf, 2 * p);
        } else {
            rc = vhip_memcpy_h2d(pr, re, p);
            if (!rc && im && pi) rc = vhip_memcpy_h2d(pi, im, p);
        }
        if (rc) va_error("put", "host-to-device copy failed", rc);
    } else {
        memcpy(pr, re, p);
        if (im && pi) memcpy(pi, im, p);
    }
}

/* matrix descriptor of a view over its block's current storage */
void va_view_mat(const VA_mview *v, int dev, vhip_mat *m)
{
    VA_block *b = v->block;
    size_t p = (size_t)b->prec;
    char *br = (char *)(dev ? b->dev_r : b->user_r);
    char *bi = (char *)(dev ? b->dev_i : b->user_i);
    m->rs = v->row_stride * VA_RS(b);
    m->cs = v->col_stride * VA_RS(b);
    m->split = b->cplx && b->split;
    m->pad_ = 0;
    if (b->cplx && !b->split) {
        m->re = br + (size_t)v->offset * 2 * p;
        m->im = NULL;
    } else {
        m->re = br + (size_t)v->offset * (size_t)VA_RS(b) * p;
        m->im = (b->cplx && bi) ? (void *)(bi + (size_t)v->offset * p) : NULL;
    }
}

/* host-side strided copy between a view in host memory and dense user data */
static void va_copy_host_view(const VA_mview *v, int into_view, void *r, void *im,
                              vsip_major major)
{
    vhip_mat m;
    vsip_length M = v->col_length, N = v->row_length, i, j;
    size_t p = (size_t)v->block->prec;
    int cplx = v->block->cplx;
    va_view_mat(v, 0, &m);
    for (i = 0; i < M; i++) {
        for (j = 0; j < N; j++) {
            long so = (long)i * m.cs + (long)j * m.rs;
            long u = (major == VSIP_ROW) ? (long)(i * N + j) : (long)(j * M + i);
            char *vr, *vi = NULL, *ur, *ui = NULL;
            if (cplx && !m.split) { vr = (char *)m.re + so * 2 * (long)p; vi = vr + p; }
            else { vr = (char *)m.re + so * (long)p; if (cplx) vi = (char *)m.im + so * (long)p; }
            if (cplx && im == NULL) { ur = (char *)r + u * 2 * (long)p; ui = ur + p; }
            else { ur = (char *)r + u * (long)p; if (cplx) ui = (char *)im + u * (long)p; }
            if (into_view) { memcpy(vr, ur, p); if (cplx) memcpy(vi, ui, p); }
            else { memcpy(ur, vr, p); if (cplx) memcpy(ui, vi, p); }
        }
    }
}

int va_mview_copy_host(const VA_mview *v, int to_view, void *r, void *im, vsip_major major)
{
    VA_block *b = v->block;
    vsip_length M = v->col_length, N = v->row_length;
    size_t p = (size_t)b->prec, count = (size_t)M * (size_t)N, esz;
    void *st_r = NULL, *st_i = NULL;
    vhip_mat vm, sm;
    int cplx = b->cplx, rc = 0, user_split = (cplx && im != NULL);
    if (count == 0) return 0;
    if (!b->admitted) {
        va_copy_host_view(v, to_view, r, im, major);
        return 0;
    }
    esz = (cplx && !user_split) ? 2 * p : p;
    rc = vhip_malloc(&st_r, esz * count);
    if (!rc && user_split) rc = vhip_malloc(&st_i, esz * count);
    if (rc) { va_error("copy_user", "staging allocation failed", rc); goto done; }
    sm.re = st_r;
    sm.im = st_i;
    sm.split = user_split;
    sm.pad_ = 0;
    if (major == VSIP_ROW) { sm.rs = 1; sm.cs = (long)N; }
    else { sm.rs = (long)M; sm.cs = 1; }
    va_view_mat(v, 1, &vm);
    if (to_view) {
        rc = vhip_memcpy_h2d(st_r, r, esz * count);
        if (!rc && user_split) rc = vhip_memcpy_h2d(st_i, im, esz * count);
        if (!rc) rc = vhip_mat_copy((int)p, cplx, &sm, &vm, (long)M, (long)N, vhip_stream());
        if (!rc) rc = vhip_sync();
    } else {
        rc = vhip_mat_copy((int)p, cplx, &vm, &sm, (long)M, (long)N, vhip_stream());
        if (!rc) rc = vhip_memcpy_d2h(r, st_r, esz * count);
        if (!rc && user_split) rc = vhip_memcpy_d2h(im, st_i, esz * count);
    }
    if (rc) va_error("copy_user", "device copy failed", rc);
done:
    vhip_sync();
    if (st_r) vhip_free(st_r);
    if (st_i) vhip_free(st_i);
    return rc;
}

/* ---- public API: complex matrix views ------------------------------------ */


vsip_cmview_f *vsip_cmbind_f(const vsip_cblock_f *b, vsip_offset o, vsip_stride cs,
                             vsip_length cl, vsip_stride rs, vsip_length rl)
{
    return (vsip_cmview_f *)va_mview_bind((VA_block *)&b->b, o, cs, cl, rs, rl);
}

vsip_cmview_d *vsip_cmbind_d(const vsip_cblock_d *b, vsip_offset o, vsip_stride cs,
                             vsip_length cl, vsip_stride rs, vsip_length rl)
{
    return (vsip_cmview_d *)va_mview_bind((VA_block *)&b->b, o, cs, cl, rs, rl);
}

vsip_cmview_f *vsip_cmcreate_f(vsip_length m, vsip_length n, vsip_major major, vsip_memory_hint h)
{
    return (vsip_cmview_f *)va_mview_create(4, 1, m, n, major, h);
}

vsip_cmview_d *vsip_cmcreate_d(vsip_length m, vsip_length n, vsip_major major, vsip_memory_hint h)
{
    return (vsip_cmview_d *)va_mview_create(8, 1, m, n, major, h);
}

vsip_cmview_f *vsip_cmsubview_f(const vsip_cmview_f *v, vsip_index i, vsip_index j,
                                vsip_length m, vsip_length n)
{
    return (vsip_cmview_f *)va_mview_bind(v->v.block, v->v.offset + i * v->v.col_stride
                                          + j * v->v.row_stride, v->v.col_stride, m,
                                          v->v.row_stride, n);
}

vsip_cmview_d *vsip_cmsubview_d(const vsip_cmview_d *v, vsip_index i, vsip_index j,
                                vsip_length m, vsip_length n)
{
    return (vsip_cmview_d *)va_mview_bind(v->v.block, v->v.offset + i * v->v.col_stride
                                          + j * v->v.row_stride, v->v.col_stride, m,
                                          v->v.row_stride, n);
}

vsip_cmview_f *vsip_cmtransview_f(const vsip_cmview_f *v)
{
    return (vsip_cmview_f *)va_mview_bind(v->v.block, v->v.offset, v->v.row_stride,
                                          v->v.row_length, v->v.col_stride, v->v.col_length);
}

vsip_cmview_d *vsip_cmtransview_d(const vsip_cmview_d *v)
{
    return (vsip_cmview_d *)va_mview_bind(v->v.block, v->v.offset, v->v.row_stride,
                                          v->v.row_length, v->v.col_stride, v->v.col_length);
}

/* real / imaginary part of a complex matrix view: the same offset, strides
 * and lengths over the complex block's derived real block
 * (VI_mrealview_f.h:20-33, vsip_mimagview_f.c) */
static VA_mview *va_part_view(const VA_mview *v, int imag)
{
    VA_block *pb = v ? va_block_part(v->block, imag) : NULL;
    if (!pb) return NULL;
    return va_mview_bind(pb, v->offset, v->col_stride, v->col_length, v->row_stride, v->row_length);
}

vsip_mview_f *vsip_mrealview_f(const vsip_cmview_f *v) { return (vsip_mview_f *)va_part_view(v ? &v->v : NULL, 0); }
vsip_mview_d *vsip_mrealview_d(const vsip_cmview_d *v) { return (vsip_mview_d *)va_part_view(v ? &v->v : NULL, 0); }
vsip_mview_f *vsip_mimagview_f(const vsip_cmview_f *v) { return (vsip_mview_f *)va_part_view(v ? &v->v : NULL, 1); }
vsip_mview_d *vsip_mimagview_d(const vsip_cmview_d *v) { return (vsip_mview_d *)va_part_view(v ? &v->v : NULL, 1); }

vsip_cblock_f *vsip_cmdestroy_f(vsip_cmview_f *v)
{
    vsip_cblock_f *b = v ? (vsip_cblock_f *)v->v.block : NULL;
    va_mview_destroy(v ? &v->v : NULL);
    return b;
}

vsip_cblock_d *vsip_cmdestroy_d(vsip_cmview_d *v)
{
    vsip_cblock_d *b = v ? (vsip_cblock_d *)v->v.block : NULL;
    va_mview_destroy(v ? &v->v : NULL);
    return b;
}

void vsip_cmalldestroy_f(vsip_cmview_f *v)
{
    if (v) va_block_destroy((VA_block *)vsip_cmdestroy_f(v));
}

void vsip_cmalldestroy_d(vsip_cmview_d *v)
{
    if (v) va_block_destroy((VA_block *)vsip_cmdestroy_d(v));
}

vsip_cscalar_f vsip_cmget_f(const vsip_cmview_f *v, vsip_index i, vsip_index j)
{
    vsip_cscalar_f z;
    z.r = 0; z.i = 0;
    va_get(&v->v, i, j, &z.r, &z.i);
    return z;
}

vsip_cscalar_d vsip_cmget_d(const vsip_cmview_d *v, vsip_index i, vsip_index j)
{
    vsip_cscalar_d z;
    z.r = 0; z.i = 0;
    va_get(&v->v, i, j, &z.r, &z.i);
    return z;
}

void vsip_cmput_f(const vsip_cmview_f *v, vsip_index i, vsip_index j, vsip_cscalar_f z)
{
    va_put(&v->v, i, j, &z.r, &z.i);
}

void vsip_cmput_d(const vsip_cmview_d *v, vsip_index i, vsip_index j, vsip_cscalar_d z)
{
    va_put(&v->v, i, j, &z.r, &z.i);
}

void vsip_cmgetattrib_f(const vsip_cmview_f *v, vsip_cmattr_f *a)
{
    a->offset = v->v.offset;
    a->row_stride = v->v.row_stride;
    a->row_length = v->v.row_length;
    a->col_stride = v->v.col_stride;
    a->col_length = v->v.col_length;
    a->block = (vsip_cblock_f *)v->v.block;
}

void vsip_cmgetattrib_d(const vsip_cmview_d *v, vsip_cmattr_d *a)
{
    a->offset = v->v.offset;
    a->row_stride = v->v.row_stride;
    a->row_length = v->v.row_length;
    a->col_stride = v->v.col_stride;
    a->col_length = v->v.col_length;
    a->block = (vsip_cblock_d *)v->v.block;
}

vsip_cblock_f *vsip_cmgetblock_f(const vsip_cmview_f *v) { return (vsip_cblock_f *)v->v.block; }
vsip_cblock_d *vsip_cmgetblock_d(const vsip_cmview_d *v) { return (vsip_cblock_d *)v->v.block; }

void vsip_cmcopyfrom_user_f(vsip_scalar_f *const r, vsip_scalar_f *const i, vsip_major major,
                            const vsip_cmview_f *v)
{
    va_mview_copy_host(&v->v, 1, (void *)r, (void *)i, major);
}

void vsip_cmcopyfrom_user_d(vsip_scalar_d *const r, vsip_scalar_d *const i, vsip_major major,
                            const vsip_cmview_d *v)
{
    va_mview_copy_host(&v->v, 1, (void *)r, (void *)i, major);
}

void vsip_cmcopyto_user_f(const vsip_cmview_f *v, vsip_major major, vsip_scalar_f *const r,
                          vsip_scalar_f *const i)
{
    va_mview_copy_host(&v->v, 0, (void *)r, (void *)i, major);
}

void vsip_cmcopyto_user_d(const vsip_cmview_d *v, vsip_major major, vsip_scalar_d *const r,
                          vsip_scalar_d *const i)
{
    va_mview_copy_host(&v->v, 0, (void *)r, (void *)i, major);
}

/* ---- public API: real matrix views --------------------------------------- */

vsip_mview_f *vsip_mbind_f(const vsip_block_f *b, vsip_offset o, vsip_stride cs,
                           vsip_length cl, vsip_stride rs, vsip_length rl)
{
    return (vsip_mview_f *)va_mview_bind((VA_block *)&b->b, o, cs, cl, rs, rl);
}

vsip_mview_d *vsip_mbind_d(const vsip_block_d *b, vsip_offset o, vsip_stride cs,
                           vsip_length cl, vsip_stride rs, vsip_length rl)
{
    return (vsip_mview_d *)va_mview_bind((VA_block *)&b->b, o, cs, cl, rs, rl);
}

vsip_mview_f *vsip_mcreate_f(vsip_length m, vsip_length n, vsip_major major, vsip_memory_hint h)
{
    return (vsip_mview_f *)va_mview_create(4, 0, m, n, major, h);
}

vsip_mview_d *vsip_mcreate_d(vsip_length m, vsip_length n, vsip_major major, vsip_memory_hint h)
{
    return (vsip_mview_d *)va_mview_create(8, 0, m, n, major, h);
}

vsip_mview_f *vsip_msubview_f(const vsip_mview_f *v, vsip_index i, vsip_index j,
                              vsip_length m, vsip_length n)
{
    return (vsip_mview_f *)va_mview_bind(v->v.block, v->v.offset + i * v->v.col_stride
                                         + j * v->v.row_stride, v->v.col_stride, m,
                                         v->v.row_stride, n);
}

vsip_mview_d *vsip_msubview_d(const vsip_mview_d *v, vsip_index i, vsip_index j,
                              vsip_length m, vsip_length n)
{
    return (vsip_mview_d *)va_mview_bind(v->v.block, v->v.offset + i * v->v.col_stride
                                         + j * v->v.row_stride, v->v.col_stride, m,
                                         v->v.row_stride, n);
}

vsip_block_f *vsip_mdestroy_f(vsip_mview_f *v)
{
    vsip_block_f *b = v ? (vsip_block_f *)v->v.block : NULL;
    va_mview_destroy(v ? &v->v : NULL);
    return b;
}

vsip_block_d *vsip_mdestroy_d(vsip_mview_d *v)
{
    vsip_block_d *b = v ? (vsip_block_d *)v->v.block : NULL;
    va_mview_destroy(v ? &v->v : NULL);
    return b;
}

void vsip_malldestroy_f(vsip_mview_f *v)
{
    if (v) va_block_destroy((VA_block *)vsip_mdestroy_f(v));
}

void vsip_malldestroy_d(vsip_mview_d *v)
{
    if (v) va_block_destroy((VA_block *)vsip_mdestroy_d(v));
}

vsip_scalar_f vsip_mget_f(const vsip_mview_f *v, vsip_index i, vsip_index j)
{
    vsip_scalar_f x = 0;
    va_get(&v->v, i, j, &x, NULL);
    return x;
}

vsip_scalar_d vsip_mget_d(const vsip_mview_d *v, vsip_index i, vsip_index j)
{
    vsip_scalar_d x = 0;
    va_get(&v->v, i, j, &x, NULL);
    return x;
}

void vsip_mput_f(const vsip_mview_f *v, vsip_index i, vsip_index j, vsip_scalar_f x)
{
    va_put(&v->v, i, j, &x, NULL);
}

void vsip_mput_d(const vsip_mview_d *v, vsip_index i, vsip_index j, vsip_scalar_d x)
{
    va_put(&v->v, i, j, &x, NULL);
}

void vsip_mgetattrib_f(const vsip_mview_f *v, vsip_mattr_f *a)
{
    a->offset = v->v.offset;
    a->row_stride = v->v.row_stride;
    a->row_length = v->v.row_length;
    a->col_stride = v->v.col_stride;
    a->col_length = v->v.col_length;
    a->block = (vsip_block_f *)v->v.block;
}

void vsip_mgetattrib_d(const vsip_mview_d *v, vsip_mattr_d *a)
{
    a->offset = v->v.offset;
    a->row_stride = v->v.row_stride;
    a->row_length = v->v.row_length;
    a->col_stride = v->v.col_stride;
    a->col_length = v->v.col_length;
    a->block = (vsip_block_d *)v->v.block;
}

void vsip_mcopyfrom_user_f(vsip_scalar_f *const r, vsip_major major, const vsip_mview_f *v)
{
    va_mview_copy_host(&v->v, 1, (void *)r, NULL, major);
}

void vsip_mcopyfrom_user_d(vsip_scalar_d *const r, vsip_major major, const vsip_mview_d *v)
{
    va_mview_copy_host(&v->v, 1, (void *)r, NULL, major);
}

void vsip_mcopyto_user_f(const vsip_mview_f *v, vsip_major major, vsip_scalar_f *const r)
{
    va_mview_copy_host(&v->v, 0, (void *)r, NULL, major);
}

void vsip_mcopyto_user_d(const vsip_mview_d *v, vsip_major major, vsip_scalar_d *const r)
{
    va_mview_copy_host(&v->v, 0, (void *)r, NULL, major);
}
