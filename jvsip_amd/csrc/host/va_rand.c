/*
 * va_rand.c -- the VSIPL random number generators (vsip_randcreate and the
 * randu/randn family), with the vector and matrix fills on the device.
 *
 * Reference behaviour (rrjudd/jvsip c_VSIP_src/):
 *   state and creation   vsip_randobject.h, vsip_randcreate.c:17-19, 20-104
 *     nonportable (VSIP_NPRNG): X <- a X + c with a = 1664525^numseqs and
 *       c = 1013904223 (1 + a0 + ... + a0^(numseqs-1)) (the leapfrog of
 *       stream id), X started `id` base steps after the seed;
 *     portable (VSIP_PRNG): X <- 1664525 X + 1013904223 started
 *       floor((2^32-1)/numseqs) (id-1) steps after the seed, and
 *       X1 <- 69069 X1 + c1 with c1 the id-th odd prime, X1 = X2 = 1; a
 *       draw is X - X1, and X1 steps past its marker X2 (both + 1) when it
 *       meets it (vsip_randu_f.c:30-36)
 *   draws to values      vsip_randu_f.c, vsip_randn_f.c, vsip_crandu_f.c,
 *                        vsip_crandn_f.c (and _d), vsip_vrandu_f.c,
 *                        vsip_vrandn_f.c, vsip_cvrandu_f.c, vsip_cvrandn_f.c
 *                        (and _d), vsip_[c]mrand[un]_[fd].c (by column when
 *                        col_stride < row_stride, else by row)
 * The values keep every rounding step of the reference, including where
 * its functions differ in how a portable draw t = X - X1 becomes a value:
 *   (t >> 8 | 1) / 2^24   scalar _f draws; vrandu_f, vrandn_f, cvrandu_f,
 *                         cvrandn_f, mrandn_f, cmrandn_f, mrandn_d, cmrandn_d
 *   t / 2^32              scalar randn_d, crandn_d, crandu_d (real part);
 *                         mrandu_f, cmrandu_f (t rounded to float first) and
 *                         the other _d fills
 *   (t + 0.5) / 2^32      vsip_randu_d
 *   t / 2^24              vsip_crandu_d's imaginary part
 *
 * The state struct has the reference's field order.  C89.
 */
#include <stdlib.h>
#include <string.h>
#include "va_internal.h"

#define VA_A0 1664525u
#define VA_C0 1013904223u
#define VA_A1 69069u

struct vsip_randobject {
    vsip_scalar_ue32 a, c, a1, c1;
    vsip_scalar_ue32 X, X1, X2;
    int type;
};

/* x after k steps of x <- a x + c (k < 2^33) */
static vsip_scalar_ue32 va_lcg_jump(vsip_scalar_ue32 a, vsip_scalar_ue32 c, unsigned long k, vsip_scalar_ue32 x)
{
    while (k) {
        if (k & 1) x = a * x + c;
        c = (a + 1u) * c;
        a = a * a;
        k >>= 1;
    }
    return x;
}

/* maps of 2^i steps */
static void va_lcg_tables(vsip_scalar_ue32 a, vsip_scalar_ue32 c, unsigned *ja, unsigned *jc)
{
    int i;
    for (i = 0; i < 32; i++) {
        ja[i] = a;
        jc[i] = c;
        c = (a + 1u) * c;
        a = a * a;
    }
}

/* the map of k steps */
static void va_lcg_compose(vsip_scalar_ue32 a, vsip_scalar_ue32 c, unsigned long k, unsigned *ak, unsigned *ck)
{
    vsip_scalar_ue32 A = 1u, C = 0u;
    while (k) {
        if (k & 1) { C = a * C + c; A = a * A; }
        c = (a + 1u) * c;
        a = a * a;
        k >>= 1;
    }
    *ak = A;
    *ck = C;
}

/* smallest t >= 1 with (t steps of X1 from x1) == x2, for the full-period
 * sequence X1 <- a1 X1 + c1 (a1 = 1 mod 4, c1 odd): its value mod 2^j
 * repeats with period 2^j, so t is found one bit at a time */
static unsigned long va_lcg_meet(vsip_scalar_ue32 a1, vsip_scalar_ue32 c1, vsip_scalar_ue32 x1, vsip_scalar_ue32 x2)
{
    unsigned long t = 0;
    int j;
    for (j = 1; j <= 32; j++) {
        vsip_scalar_ue32 mask = (j == 32) ? 0xffffffffu : ((1u << j) - 1u);
        if (((va_lcg_jump(a1, c1, t, x1) ^ x2) & mask) != 0) t += 1UL << (j - 1);
    }
    return t ? t : (1UL << 32);
}

vsip_randstate *vsip_randcreate(vsip_index seed, vsip_index numseqs, vsip_index id, vsip_rng type)
{
    vsip_randstate *s;
    vsip_scalar_ue32 k = (vsip_scalar_ue32)numseqs, x0 = (vsip_scalar_ue32)seed;
    /* numseqs 0 and (portable) id 0 leave the reference undefined (it loops
     * forever or reads before its prime table); they are refused here, and
     * so are portable ids past its 100 primes */
    if (k < 1) return NULL;
    if (type == VSIP_PRNG && (id < 1 || id > 100 || id > numseqs)) return NULL;
    s = (vsip_randstate *)calloc(1, sizeof(*s));
    if (!s) return NULL;
    s->type = (int)type;
    if (type != VSIP_PRNG) {
        /* every numseqs-th value of the base sequence: the map of numseqs
         * steps, a = A0^k, c = C0 (1 + A0 + ... + A0^(k-1)) */
        unsigned a, c;
        va_lcg_compose(VA_A0, VA_C0, (unsigned long)k, &a, &c);
        s->X = va_lcg_jump(VA_A0, VA_C0, (unsigned long)id, x0);
        s->a = a;
        s->c = c;
    } else {
        vsip_scalar_ue32 p = 1u, found = 0u;
        while (found < (vsip_scalar_ue32)id) {            /* the id-th odd prime */
            vsip_scalar_ue32 q, prime = 1u;
            p += 2u;
            for (q = 3u; q * q <= p; q += 2u)
                if (p % q == 0u) { prime = 0u; break; }
            if (prime) found++;
        }
        if (id > 1) {
            vsip_scalar_ue32 skip = (vsip_scalar_ue32)((0xffffffffu / k) * (vsip_scalar_ue32)(id - 1));
            x0 = va_lcg_jump(VA_A0, VA_C0, (unsigned long)skip, x0);
        }
        s->X = x0;
        s->X1 = 1u;
        s->X2 = 1u;
        s->a = VA_A0;
        s->c = VA_C0;
        s->a1 = VA_A1;
        s->c1 = p;
    }
    return s;
}

int vsip_randdestroy(vsip_randstate *s)
{
    free(s);
    return 0;
}

/* one draw: the new X (nonportable) or X - X1 (portable), state advanced */
static vsip_scalar_ue32 va_draw(vsip_randstate *s)
{
    vsip_scalar_ue32 t;
    s->X = s->a * s->X + s->c;
    if (s->type) return s->X;
    s->X1 = s->a1 * s->X1 + s->c1;
    t = s->X - s->X1;
    if (s->X1 == s->X2) { s->X1++; s->X2++; }
    return t;
}

static float va_uf(vsip_randstate *s)
{
    vsip_scalar_ue32 t = va_draw(s);
    if (s->type) return (float)t / 4294967296.0f;
    return (float)((t >> 8) | 1u) / 16777216.0f;
}

static double va_ud(vsip_randstate *s)
{
    vsip_scalar_ue32 t = va_draw(s);
    return (double)t / 4294967296.0;
}

vsip_scalar_f vsip_randu_f(vsip_randstate *s) { return va_uf(s); }

vsip_scalar_d vsip_randu_d(vsip_randstate *s)
{
    vsip_scalar_ue32 t = va_draw(s);
    if (s->type) return (double)t / 4294967296.0;
    return ((double)t + 0.5) / 4294967296.0;
}

vsip_scalar_f vsip_randn_f(vsip_randstate *s)
{
    float r = 0.0f;
    int i;
    for (i = 0; i < 12; i++) r += va_uf(s);
    return s->type ? r - 6.0f : 6.0f - r;
}

vsip_scalar_d vsip_randn_d(vsip_randstate *s)
{
    double r = 0.0;
    int i;
    for (i = 0; i < 12; i++) r += va_ud(s);
    return s->type ? r - 6.0 : 6.0 - r;
}

vsip_cscalar_f vsip_crandu_f(vsip_randstate *s)
{
    vsip_cscalar_f z;
    z.r = va_uf(s);
    z.i = va_uf(s);
    return z;
}

vsip_cscalar_d vsip_crandu_d(vsip_randstate *s)
{
    vsip_cscalar_d z;
    z.r = va_ud(s);
    if (s->type) {
        z.i = va_ud(s);
    } else {
        vsip_scalar_ue32 t = va_draw(s);
        z.i = (double)t / 16777216.0;
    }
    return z;
}

vsip_cscalar_f vsip_crandn_f(vsip_randstate *s)
{
    vsip_cscalar_f z;
    float t1, t2;
    t1 = va_uf(s);
    t1 += va_uf(s);
    t1 += va_uf(s);
    t2 = va_uf(s);
    t2 += va_uf(s);
    t2 += va_uf(s);
    z.i = t1 - t2;
    z.r = 3.0f - t2 - t1;
    return z;
}

vsip_cscalar_d vsip_crandn_d(vsip_randstate *s)
{
    vsip_cscalar_d z;
    double t1, t2;
    t1 = va_ud(s);
    t1 += va_ud(s);
    t1 += va_ud(s);
    t2 = va_ud(s);
    t2 += va_ud(s);
    t2 += va_ud(s);
    z.i = t1 - t2;
    z.r = 3.0 - t2 - t1;
    return z;
}

/* ---- device fills ----------------------------------------------------- */

/* one launch over view v (every element; draws of the launch < 2^32) */
static int va_fill_launch(vsip_randstate *s, const VA_mview *v, int dist, int cplx, int p24, int by_col)
{
    vhip_rng_desc d;
    const long D = dist == VHIP_RNG_UNIFORM ? (cplx ? 2 : 1) : (cplx ? 6 : 12);
    long total = (long)v->col_length * (long)v->row_length;
    unsigned long draws = (unsigned long)(D * total);
    unsigned long meet = 0;
    int rc;
    if (total <= 0) return 0;
    memset(&d, 0, sizeof(d));
    d.prec = v->block->prec;
    d.dist = dist;
    d.cplx = cplx;
    d.portable = (s->type == 0);
    d.by_col = by_col;
    d.p24 = p24;
    d.m = (long)v->col_length;
    d.n = (long)v->row_length;
    d.lanes = vhip_rng_lanes(total);
    d.a = s->a; d.c = s->c; d.a1 = s->a1; d.c1 = s->c1;
    d.x0 = s->X; d.x10 = s->X1; d.x2 = s->X2;
    va_lcg_tables(s->a, s->c, d.ja, d.jc);
    va_lcg_compose(s->a, s->c, (unsigned long)(D * d.lanes), &d.stride_a, &d.stride_c);
    if (d.portable) {
        va_lcg_tables(s->a1, s->c1, d.ja1, d.jc1);
        va_lcg_compose(s->a1, s->c1, (unsigned long)(D * d.lanes), &d.stride_a1, &d.stride_c1);
        meet = va_lcg_meet(s->a1, s->c1, s->X1, s->X2);
        if (meet <= draws) d.kstar = (long)meet;
    }
    va_view_mat(v, 1, &d.r);
    rc = vhip_rng_fill(&d, vhip_stream());
    if (rc) return rc;
    /* the state after the launch's draws */
    s->X = va_lcg_jump(s->a, s->c, draws, s->X);
    if (d.portable) {
        if (d.kstar) {
            s->X1 = va_lcg_jump(s->a1, s->c1, draws - meet, s->X2 + 1u);
            s->X2 += 1u;
        } else {
            s->X1 = va_lcg_jump(s->a1, s->c1, draws, s->X1);
        }
    }
    return 0;
}

/* fill view v in the reference's order, in launches of fewer than 2^32 draws */
static void va_fill(const char *fn, vsip_randstate *s, const VA_mview *v, int dist, int cplx, int p24)
{
    const long D = dist == VHIP_RNG_UNIFORM ? (cplx ? 2 : 1) : (cplx ? 6 : 12);
    int by_col;
    long outer, inner, per, k0, lim;
    if (!s || !v || !v->block) { va_error(fn, "null argument", -1); return; }
    if (va_ensure_init()) return;
    if (!v->block->admitted || !v->block->dev_r) { va_error(fn, "view of a released block", -1); return; }
    /* VSIP_AMD_RNG_LAUNCH_DRAWS lowers the draws per launch (tests of the
     * split into several launches) */
    lim = (1L << 32) - 1;
    if (getenv("VSIP_AMD_RNG_LAUNCH_DRAWS")) {
        long l = atol(getenv("VSIP_AMD_RNG_LAUNCH_DRAWS"));
        if (l > 0 && l < lim) lim = l;
    }
    by_col = v->col_stride < v->row_stride;
    inner = by_col ? (long)v->col_length : (long)v->row_length;
    outer = by_col ? (long)v->row_length : (long)v->col_length;
    if (inner * D > lim) { va_error(fn, "a column or row of the view holds too many draws for one launch", -1); return; }
    per = lim / (inner * D > 0 ? inner * D : 1);
    for (k0 = 0; k0 < outer; k0 += per) {
        VA_mview part = *v;
        long cnt = outer - k0 < per ? outer - k0 : per;
        int rc;
        if (by_col) {
            part.offset = (vsip_offset)((long)v->offset + k0 * v->row_stride);
            part.row_length = (vsip_length)cnt;
        } else {
            part.offset = (vsip_offset)((long)v->offset + k0 * v->col_stride);
            part.col_length = (vsip_length)cnt;
        }
        rc = va_fill_launch(s, &part, dist, cplx, p24, by_col);
        if (rc) { va_error(fn, vhip_error_string(rc), rc); return; }
    }
}

#define U VHIP_RNG_UNIFORM
#define N VHIP_RNG_NORMAL
void vsip_vrandu_f(vsip_randstate *s, const vsip_vview_f *r) { va_fill("vsip_vrandu_f", s, &r->v, U, 0, 1); }
void vsip_vrandu_d(vsip_randstate *s, const vsip_vview_d *r) { va_fill("vsip_vrandu_d", s, &r->v, U, 0, 0); }
void vsip_vrandn_f(vsip_randstate *s, const vsip_vview_f *r) { va_fill("vsip_vrandn_f", s, &r->v, N, 0, 1); }
void vsip_vrandn_d(vsip_randstate *s, const vsip_vview_d *r) { va_fill("vsip_vrandn_d", s, &r->v, N, 0, 0); }
void vsip_cvrandu_f(vsip_randstate *s, const vsip_cvview_f *r) { va_fill("vsip_cvrandu_f", s, &r->v, U, 1, 1); }
void vsip_cvrandu_d(vsip_randstate *s, const vsip_cvview_d *r) { va_fill("vsip_cvrandu_d", s, &r->v, U, 1, 0); }
void vsip_cvrandn_f(vsip_randstate *s, const vsip_cvview_f *r) { va_fill("vsip_cvrandn_f", s, &r->v, N, 1, 1); }
void vsip_cvrandn_d(vsip_randstate *s, const vsip_cvview_d *r) { va_fill("vsip_cvrandn_d", s, &r->v, N, 1, 0); }
void vsip_mrandu_f(vsip_randstate *s, const vsip_mview_f *r) { va_fill("vsip_mrandu_f", s, &r->v, U, 0, 0); }
void vsip_mrandu_d(vsip_randstate *s, const vsip_mview_d *r) { va_fill("vsip_mrandu_d", s, &r->v, U, 0, 0); }
void vsip_mrandn_f(vsip_randstate *s, const vsip_mview_f *r) { va_fill("vsip_mrandn_f", s, &r->v, N, 0, 1); }
void vsip_mrandn_d(vsip_randstate *s, const vsip_mview_d *r) { va_fill("vsip_mrandn_d", s, &r->v, N, 0, 1); }
void vsip_cmrandu_f(vsip_randstate *s, const vsip_cmview_f *r) { va_fill("vsip_cmrandu_f", s, &r->v, U, 1, 0); }
void vsip_cmrandu_d(vsip_randstate *s, const vsip_cmview_d *r) { va_fill("vsip_cmrandu_d", s, &r->v, U, 1, 0); }
void vsip_cmrandn_f(vsip_randstate *s, const vsip_cmview_f *r) { va_fill("vsip_cmrandn_f", s, &r->v, N, 1, 1); }
void vsip_cmrandn_d(vsip_randstate *s, const vsip_cmview_d *r) { va_fill("vsip_cmrandn_d", s, &r->v, N, 1, 1); }
