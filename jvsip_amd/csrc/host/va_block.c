/*
 * va_block.c -- VSIPL blocks whose admitted data lives in HBM.
 *
 * Behaviour follows the reference's default-interleaved block family:
 *   bind      VI_cblockbind_f_di.h:17-57   (user block, released; Ip == NULL
 *                                           -> interleaved, else split)
 *   admit     VI_cblockadmit_f_di.h:16-29  (0 on success; 1 if the data
 *                                           pointer is NULL or the block is
 *                                           not a user block)
 *   release   VI_cblockrelease_f_di.h:16-34 (returns the user pointers, Ip is
 *                                           NULL for interleaved; NULL pointers
 *                                           for a non-user block)
 *   create    vsip_cblockcreate_f.c:20-86  (library block, born admitted)
 *   destroy   VI_cblockdestroy_f_di.h:16-24 (user arrays are never freed)
 * and adds the copies of the copy-on-admit model (VI_cblockadmit_f_ai.h:28-47,
 * VI_cblockrelease_f_ai.h:24-35): admit(update) uploads the user arrays to
 * the device, release(update) downloads them.
 *
 * C89.
 */
#include <stdio.h>
#include <stdlib.h>
#include "va_internal.h"

static size_t va_elem_bytes(const VA_block *b)
{
    /* bytes per element of one device array */
    if (b->cplx && !b->split) return (size_t)(2 * b->prec);
    return (size_t)b->prec;
}

static int va_block_alloc_dev(VA_block *b)
{
    size_t bytes = va_elem_bytes(b) * (size_t)b->size;
    int rc;
    if (b->dev_r) return 0;
    rc = vhip_malloc(&b->dev_r, bytes);
    if (rc) { b->dev_r = NULL; return rc; }
    if (b->cplx && b->split) {
        rc = vhip_malloc(&b->dev_i, bytes);
        if (rc) { vhip_free(b->dev_r); b->dev_r = NULL; b->dev_i = NULL; return rc; }
    }
    return 0;
}

/* keep the derived real/imaginary blocks of complex block b pointing at b's
 * current data and admission state */
static void va_block_sync_parts(VA_block *b)
{
    int k;
    for (k = 0; k < 2; k++) {
        VA_block *s = b->part[k];
        size_t off = (size_t)k * (size_t)b->prec;
        if (!s) continue;
        s->admitted = b->admitted;
        s->size = b->size;
        if (b->split) {
            s->rstride = 1;
            s->user_r = k ? b->user_i : b->user_r;
            s->dev_r = k ? b->dev_i : b->dev_r;
        } else {
            s->rstride = 2;
            s->user_r = b->user_r ? (void *)((char *)b->user_r + off) : NULL;
            s->dev_r = b->dev_r ? (void *)((char *)b->dev_r + off) : NULL;
        }
    }
}

/* the real (imag = 0) or imaginary part of complex block b as a derived real
 * block, made on first use and owned by b */
VA_block *va_block_part(VA_block *b, int imag)
{
    VA_block *s;
    if (!b || !b->cplx) return NULL;
    if (b->part[imag]) return b->part[imag];
    s = (VA_block *)calloc(1, sizeof(VA_block));
    if (!s) return NULL;
    s->prec = b->prec;
    s->cplx = 0;
    s->kind = VA_DERIVED_BLOCK;
    s->markings = VA_VALID;
    s->hint = b->hint;
    s->parent = b;
    b->part[imag] = s;
    va_block_sync_parts(b);
    return s;
}

VA_block *va_block_bind(int prec, int cplx, void *r, void *i, vsip_length n, vsip_memory_hint h)
{
    VA_block *b;
    if (va_ensure_init()) return NULL;
    b = (VA_block *)calloc(1, sizeof(VA_block));
    if (!b) return NULL;
    b->prec = prec;
    b->cplx = cplx;
    b->kind = VA_USER_BLOCK;
    b->admitted = 0;
    b->split = (cplx && i != NULL) ? 1 : 0;
    b->markings = VA_VALID;
    b->size = n;
    b->user_r = r;
    b->user_i = cplx ? i : NULL;
    b->hint = h;
    return b;
}

VA_block *va_block_create(int prec, int cplx, vsip_length n, vsip_memory_hint h)
{
    VA_block *b;
    if (va_ensure_init()) return NULL;
    b = (VA_block *)calloc(1, sizeof(VA_block));
    if (!b) return NULL;
    b->prec = prec;
    b->cplx = cplx;
    b->kind = VA_LIB_BLOCK;
    b->admitted = 1;
    b->split = 0;
    b->markings = VA_VALID;
    b->size = n;
    b->hint = h;
    if (va_block_alloc_dev(b)) {
        free(b);
        return NULL;
    }
    return b;
}

int va_block_admit(VA_block *b, int update)
{
    size_t bytes;
    int rc;
    if (!b || b->user_r == NULL || b->kind != VA_USER_BLOCK) return 1;
    if (b->admitted) return 0;
    if (va_block_alloc_dev(b)) return 1;
    if (update) {
        bytes = va_elem_bytes(b) * (size_t)b->size;
        rc = vhip_memcpy_h2d(b->dev_r, b->user_r, bytes);
        if (!rc && b->split) rc = vhip_memcpy_h2d(b->dev_i, b->user_i, bytes);
        if (rc) { va_error("blockadmit", "host-to-device copy failed", rc); return 1; }
    }
    b->admitted = 1;
    va_block_sync_parts(b);
    return 0;
}

int va_block_release(VA_block *b, int update, void **r, void **i)
{
    size_t bytes;
    int rc;
    if (r) *r = NULL;
    if (i) *i = NULL;
    if (!b || b->kind != VA_USER_BLOCK) return 1;
    if (b->admitted && update && b->dev_r) {
        bytes = va_elem_bytes(b) * (size_t)b->size;
        rc = vhip_memcpy_d2h(b->user_r, b->dev_r, bytes);
        if (!rc && b->split) rc = vhip_memcpy_d2h(b->user_i, b->dev_i, bytes);
        if (rc) va_error("blockrelease", "device-to-host copy failed", rc);
    } else if (b->admitted) {
        vhip_sync();
    }
    b->admitted = 0;
    va_block_sync_parts(b);
    if (r) *r = b->user_r;
    if (i) *i = b->split ? b->user_i : NULL;
    return 0;
}

void va_block_destroy(VA_block *b)
{
    int k;
    if (!b || b->kind == VA_DERIVED_BLOCK) return;   /* parts go with their complex block */
    for (k = 0; k < 2; k++) {
        if (b->part[k]) {
            b->part[k]->markings = VA_FREED;
            free(b->part[k]);
        }
    }
    if (b->dev_r || b->dev_i) vhip_sync();
    if (b->dev_r) vhip_free(b->dev_r);
    if (b->dev_i) vhip_free(b->dev_i);
    b->markings = VA_FREED;
    free(b);
}

/* ---- public API: real blocks ------------------------------------------- */

vsip_block_f *vsip_blockbind_f(vsip_scalar_f *const p, vsip_length n, vsip_memory_hint h)
{
    return (vsip_block_f *)va_block_bind(4, 0, (void *)p, NULL, n, h);
}

vsip_block_d *vsip_blockbind_d(vsip_scalar_d *const p, vsip_length n, vsip_memory_hint h)
{
    return (vsip_block_d *)va_block_bind(8, 0, (void *)p, NULL, n, h);
}

vsip_block_f *vsip_blockcreate_f(vsip_length n, vsip_memory_hint h)
{
    return (vsip_block_f *)va_block_create(4, 0, n, h);
}

vsip_block_d *vsip_blockcreate_d(vsip_length n, vsip_memory_hint h)
{
    return (vsip_block_d *)va_block_create(8, 0, n, h);
}

int vsip_blockadmit_f(vsip_block_f *b, vsip_scalar_bl update)
{
    return va_block_admit(b ? &b->b : NULL, update);
}

int vsip_blockadmit_d(vsip_block_d *b, vsip_scalar_bl update)
{
    return va_block_admit(b ? &b->b : NULL, update);
}

vsip_scalar_f *vsip_blockrelease_f(vsip_block_f *b, vsip_scalar_bl update)
{
    void *r = NULL;
    va_block_release(b ? &b->b : NULL, update, &r, NULL);
    return (vsip_scalar_f *)r;
}

vsip_scalar_d *vsip_blockrelease_d(vsip_block_d *b, vsip_scalar_bl update)
{
    void *r = NULL;
    va_block_release(b ? &b->b : NULL, update, &r, NULL);
    return (vsip_scalar_d *)r;
}

void vsip_blockdestroy_f(vsip_block_f *b) { va_block_destroy(b ? &b->b : NULL); }
void vsip_blockdestroy_d(vsip_block_d *b) { va_block_destroy(b ? &b->b : NULL); }

vsip_scalar_f *vsip_blockfind_f(const vsip_block_f *b)
{
    return b ? (vsip_scalar_f *)b->b.user_r : NULL;
}

vsip_scalar_d *vsip_blockfind_d(const vsip_block_d *b)
{
    return b ? (vsip_scalar_d *)b->b.user_r : NULL;
}

/* rebind: only a released user block may take new arrays; returns the old one */
static void *va_rebind(VA_block *b, void *r, void *i, void **old_i)
{
    void *old = NULL;
    if (old_i) *old_i = NULL;
    if (!b || b->kind != VA_USER_BLOCK || b->admitted) return NULL;
    old = b->user_r;
    if (old_i) *old_i = b->split ? b->user_i : NULL;
    if (b->cplx && ((i != NULL) != (b->split != 0))) {
        /* changing between split and interleaved changes the device layout */
        if (b->dev_r) { vhip_free(b->dev_r); b->dev_r = NULL; }
        if (b->dev_i) { vhip_free(b->dev_i); b->dev_i = NULL; }
        b->split = (i != NULL);
    }
    b->user_r = r;
    b->user_i = b->cplx ? i : NULL;
    va_block_sync_parts(b);
    return old;
}

vsip_scalar_f *vsip_blockrebind_f(vsip_block_f *b, vsip_scalar_f *const p)
{
    return (vsip_scalar_f *)va_rebind(b ? &b->b : NULL, (void *)p, NULL, NULL);
}

vsip_scalar_d *vsip_blockrebind_d(vsip_block_d *b, vsip_scalar_d *const p)
{
    return (vsip_scalar_d *)va_rebind(b ? &b->b : NULL, (void *)p, NULL, NULL);
}

/* ---- public API: complex blocks ---------------------------------------- */

vsip_cblock_f *vsip_cblockbind_f(vsip_scalar_f *const r, vsip_scalar_f *const i,
                                 vsip_length n, vsip_memory_hint h)
{
    return (vsip_cblock_f *)va_block_bind(4, 1, (void *)r, (void *)i, n, h);
}

vsip_cblock_d *vsip_cblockbind_d(vsip_scalar_d *const r, vsip_scalar_d *const i,
                                 vsip_length n, vsip_memory_hint h)
{
    return (vsip_cblock_d *)va_block_bind(8, 1, (void *)r, (void *)i, n, h);
}

vsip_cblock_f *vsip_cblockcreate_f(vsip_length n, vsip_memory_hint h)
{
    return (vsip_cblock_f *)va_block_create(4, 1, n, h);
}

vsip_cblock_d *vsip_cblockcreate_d(vsip_length n, vsip_memory_hint h)
{
    return (vsip_cblock_d *)va_block_create(8, 1, n, h);
}

int vsip_cblockadmit_f(vsip_cblock_f *b, vsip_scalar_bl update)
{
    return va_block_admit(b ? &b->b : NULL, update);
}

int vsip_cblockadmit_d(vsip_cblock_d *b, vsip_scalar_bl update)
{
    return va_block_admit(b ? &b->b : NULL, update);
}

void vsip_cblockrelease_f(vsip_cblock_f *b, vsip_scalar_bl update,
                          vsip_scalar_f **r, vsip_scalar_f **i)
{
    void *pr = NULL, *pi = NULL;
    va_block_release(b ? &b->b : NULL, update, &pr, &pi);
    if (r) *r = (vsip_scalar_f *)pr;
    if (i) *i = (vsip_scalar_f *)pi;
}

void vsip_cblockrelease_d(vsip_cblock_d *b, vsip_scalar_bl update,
                          vsip_scalar_d **r, vsip_scalar_d **i)
{
    void *pr = NULL, *pi = NULL;
    va_block_release(b ? &b->b : NULL, update, &pr, &pi);
    if (r) *r = (vsip_scalar_d *)pr;
    if (i) *i = (vsip_scalar_d *)pi;
}

void vsip_cblockdestroy_f(vsip_cblock_f *b) { va_block_destroy(b ? &b->b : NULL); }
void vsip_cblockdestroy_d(vsip_cblock_d *b) { va_block_destroy(b ? &b->b : NULL); }

void vsip_cblockfind_f(const vsip_cblock_f *b, vsip_scalar_f **r, vsip_scalar_f **i)
{
    *r = b ? (vsip_scalar_f *)b->b.user_r : NULL;
    *i = (b && b->b.split) ? (vsip_scalar_f *)b->b.user_i : NULL;
}

void vsip_cblockfind_d(const vsip_cblock_d *b, vsip_scalar_d **r, vsip_scalar_d **i)
{
    *r = b ? (vsip_scalar_d *)b->b.user_r : NULL;
    *i = (b && b->b.split) ? (vsip_scalar_d *)b->b.user_i : NULL;
}

void vsip_cblockrebind_f(vsip_cblock_f *b, vsip_scalar_f *const r, vsip_scalar_f *const i,
                         vsip_scalar_f **old_r, vsip_scalar_f **old_i)
{
    void *oi = NULL;
    void *orr = va_rebind(b ? &b->b : NULL, (void *)r, (void *)i, &oi);
    if (old_r) *old_r = (vsip_scalar_f *)orr;
    if (old_i) *old_i = (vsip_scalar_f *)oi;
}

void vsip_cblockrebind_d(vsip_cblock_d *b, vsip_scalar_d *const r, vsip_scalar_d *const i,
                         vsip_scalar_d **old_r, vsip_scalar_d **old_i)
{
    void *oi = NULL;
    void *orr = va_rebind(b ? &b->b : NULL, (void *)r, (void *)i, &oi);
    if (old_r) *old_r = (vsip_scalar_d *)orr;
    if (old_i) *old_i = (vsip_scalar_d *)oi;
}
