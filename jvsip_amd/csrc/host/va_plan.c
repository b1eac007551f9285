/*
 * va_plan.c -- transform schedules: factorisation, launch shapes, twiddles.
 *
 * The reference factors the length in the fixed order 3,5,7,8,4,2 into at
 * most six radix groups plus a remainder handled by an O(L^2) DFT
 * (VI_nuV.h:22-136) and builds an FP32 twiddle table from a float angle ramp
 * (VI_ccfftmip_create_f_def.h:63-73, vsip_vramp_f.c:31-34).  Only the result
 * of the transform has to agree, so the schedule here is chosen for the
 * hardware instead:
 *   - radices 16, 8, 4, 2 for the power-of-two part, split as evenly as
 *     possible into ceil(log2(p)/4) stages, then one stage per factor
 *     3, 5, 7, 11, 13;
 *   - per stage s with Ns = R_0 ... R_{s-1}, twiddles W_{Ns R}^{r k},
 *     r = 1..R-1, k = 0..Ns-1, computed in double precision from the exact
 *     integer angle and stored in the working precision;
 *   - launch shapes (vectors per workgroup, threads, register class) for the
 *     two lane mappings, bounded by the per-workgroup LDS of the device.
 * Lengths with a prime factor the kernels have no butterfly for are
 * rejected (create returns NULL); lengths beyond one workgroup's LDS use the
 * four-step split planned in va_fftm.c.
 *
 * C89.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "va_internal.h"

#define VA_PI 3.14159265358979323846264338327950288

static int va_pad(int prec, int i)
{
    return i + (i >> (prec == 4 ? 5 : 4));
}

/*
 * Factor n into stage radices: the power-of-two part split as evenly as
 * possible into stages of at most `maxbits` bits, then one stage per factor
 * 3, 5, 7, 11, 13 (odd stages first when odd_first).  Returns the stage count,
 * or -1 for a prime factor above 13.
 */
int va_factor_variant(int n, int prec, int maxbits, int odd_first, int *radix)
{
    int m = n, p2 = 0, nst = 0, s, stages2, base, extra, i, nodd = 0;
    int odds[VHIP_MAX_STAGES], pw[VHIP_MAX_STAGES], npw = 0;
    static const int odd[5] = {3, 5, 7, 11, 13};
    if (prec != 4 && maxbits > 3) maxbits = 3;   /* radix 16 only in the float kernels */
    if (maxbits < 1) maxbits = 1;
    if (n < 1) return -1;
    while (m % 2 == 0) { m /= 2; p2++; }
    if (p2 > 0) {
        stages2 = (p2 + maxbits - 1) / maxbits;
        base = p2 / stages2;
        extra = p2 % stages2;
        for (s = 0; s < stages2; s++) {
            if (npw >= VHIP_MAX_STAGES) return -1;
            pw[npw++] = 1 << (base + (s < extra ? 1 : 0));
        }
    }
    for (i = 0; i < 5; i++) {
        while (m % odd[i] == 0) {
            if (nodd >= VHIP_MAX_STAGES) return -1;
            odds[nodd++] = odd[i];
            m /= odd[i];
        }
    }
    if (m != 1 || npw + nodd > VHIP_MAX_STAGES) return -1;
    if (odd_first) {
        for (i = 0; i < nodd; i++) radix[nst++] = odds[i];
        for (i = 0; i < npw; i++) radix[nst++] = pw[i];
    } else {
        for (i = 0; i < npw; i++) radix[nst++] = pw[i];
        for (i = 0; i < nodd; i++) radix[nst++] = odds[i];
    }
    if (n == 1) radix[nst++] = 1;   /* length 1: one identity stage */
    return nst;
}

/*
 * Pack the prime factors of n (all <= 13) into as few stages as possible with
 * every radix <= 16 (float kernels: 2..16 including the composite radices
 * 6, 9, 10, 12, 14, 15): first-fit over the factors in decreasing order.
 * Returns the stage count or -1.
 */
int va_factor_pack(int n, int *radix)
{
    int f[64], nf = 0, m = n, p, i, j, nst = 0;
    static const int primes[6] = {13, 11, 7, 5, 3, 2};
    if (n < 1) return -1;
    for (i = 0; i < 6; i++)
        while (m % primes[i] == 0 && nf < 64) { f[nf++] = primes[i]; m /= primes[i]; }
    if (m != 1) return -1;
    for (i = 0; i < nf; i++) {
        p = f[i];
        for (j = 0; j < nst; j++)
            if (radix[j] * p <= 16) { radix[j] *= p; break; }
        if (j == nst) {
            if (nst >= VHIP_MAX_STAGES) return -1;
            radix[nst++] = p;
        }
    }
    if (n == 1) radix[nst++] = 1;
    /* larger radices first */
    for (i = 1; i < nst; i++)
        for (j = i; j > 0 && radix[j] > radix[j - 1]; j--) { p = radix[j]; radix[j] = radix[j - 1]; radix[j - 1] = p; }
    return nst;
}

/* a radix sequence for length n from VSIP_AMD_RADICES ("n:r1,r2,...;..."),
 * used only when n is not a power of two (the power-of-two kernels take the
 * even split), the radices multiply to n and the kernels have each of them
 * (2..16 for float, 2, 3, 4, 5, 7, 8, 11, 13 for double); else 0 */
static int va_factor_env(int n, int prec, int *radix)
{
    const char *e = getenv("VSIP_AMD_RADICES");
    if (n < 2 || (n & (n - 1)) == 0) return 0;
    while (e && *e) {
        char *end;
        long len = strtol(e, &end, 10);
        int nst = 0, ok = 1;
        long prod = 1;
        if (end == e || *end != ':') return 0;
        e = end + 1;
        for (;;) {
            long r = strtol(e, &end, 10);
            if (end == e) { ok = 0; break; }
            e = end;
            if (nst >= VHIP_MAX_STAGES || r < 2 || r > 16 || (prec != 4 && r != 2 && r != 3 && r != 4 && r != 5 &&
                                                              r != 7 && r != 8 && r != 11 && r != 13))
                ok = 0;
            else radix[nst++] = (int)r;
            prod *= r;
            if (*e != ',') break;
            e++;
        }
        if (ok && len == n && prod == n && nst > 0) return nst;
        while (*e && *e != ';') e++;
        if (*e == ';') e++;
    }
    return 0;
}

/* the default schedule's factorisation: powers of two split evenly (the
 * specialised power-of-two kernels expect that split), other float lengths
 * packed into radices <= 16, double lengths in radix <= 8 and prime stages;
 * VSIP_AMD_RADICES replaces it for the lengths it names (tuning) */
static int va_factor(int n, int prec, int *radix)
{
    int nst = va_factor_env(n, prec, radix);
    if (nst > 0) return nst;
    if (prec == 4 && (n & (n - 1)) != 0) return va_factor_pack(n, radix);
    return va_factor_variant(n, prec, prec == 4 ? 4 : 3, 0, radix);
}

/* butterflies one thread holds in a stage of radix r (the kernels' rule) */
static int va_maxb(int r, int maxe)
{
    if ((r == 6 || r >= 9) && r != 16) return 1;
    return (maxe / r) > 0 ? (maxe / r) : 1;
}

/* does (vpb, threads, maxe) keep every stage within the register budget? */
static int va_shape_ok(const VA_sched *s, int vpb, int thr, int maxe)
{
    int i;
    for (i = 0; i < s->nst; i++) {
        long bt = (long)vpb * (s->n / s->radix[i]);
        long nb = (bt + thr - 1) / thr;
        if (nb > va_maxb(s->radix[i], maxe)) return 0;
    }
    return 1;
}

static long va_lds_bytes(const VA_sched *s, int prec, int vpb)
{
    return (long)vpb * s->lds_stride * 2 * prec + (long)vpb * 20;
}

/* smallest thread count (multiple of 64, <= 1024; <= 512 for the 32-element
 * register class), at least thr_min, that fits; 0 if none */
static int va_fit_threads(const VA_sched *s, int vpb, int maxe, int thr_min)
{
    long want = ((long)vpb * s->n + maxe - 1) / maxe;
    int thr = (int)(((want + 63) / 64) * 64), tmax = maxe > 16 ? 512 : 1024;
    if (thr < thr_min) thr = ((thr_min + 63) / 64) * 64;
    if (thr < 64) thr = 64;
    for (; thr <= tmax; thr += 64)
        if (va_shape_ok(s, vpb, thr, maxe)) return thr;
    return 0;
}

/* min(v, the positive integer in environment variable `name`), a tuning cap */
static int va_env_cap(const char *name, int v)
{
    const char *e = getenv(name);
    int c = e ? atoi(e) : 0;
    return (c > 0 && c < v) ? c : v;
}

/* a tuning floor on lanes per workgroup from the environment (0: none) */
static int va_env_floor(const char *name)
{
    const char *e = getenv(name);
    int c = e ? atoi(e) : 0;
    return (c > 0 && c <= 1024) ? c : 0;
}

static int va_pick_launch(const VA_sched *s, int prec, int vpb_want, int lds_soft, int lds_max,
                          int thr_min, VA_launch *out)
{
    int vpb, thr, m;
    static const int maxes_f[2] = {16, 32};
    static const int maxes_d[1] = {8};
    const int *maxes = (prec == 4) ? maxes_f : maxes_d;
    int nmax = (prec == 4) ? 2 : 1;
    for (vpb = vpb_want; vpb >= 1; vpb--) {
        long lds = va_lds_bytes(s, prec, vpb);
        if (lds > lds_max) continue;
        if (vpb > 1 && lds > lds_soft) continue;
        for (m = 0; m < nmax; m++) {
            thr = va_fit_threads(s, vpb, maxes[m], thr_min);
            if (thr) {
                out->vpb = vpb;
                out->threads = thr;
                out->maxe = maxes[m];
                return 0;
            }
        }
    }
    return -1;
}

/*
 * Shape of a single-pass schedule of length n for nvec vectors: radices,
 * twiddle offsets, LDS stride and launch shapes, with no device work.
 * Returns 0, or -1 when n has an unsupported prime factor, or -2 when one
 * vector does not fit a workgroup (the caller then splits n).
 */
int va_sched_shape(VA_sched *s, int prec, int n, vsip_length nvec, int lds_max)
{
    int i, off, Ns, vpb_along, vpb_across;
    long want;
    memset(s, 0, sizeof(*s));
    s->n = n;
    s->nst = va_factor(n, prec, s->radix);
    if (s->nst < 0 || n < 1) return -1;
    s->lds_stride = va_pad(prec, n);
    if (s->lds_stride % 2 == 0) s->lds_stride += 1;

    /* along: enough elements per workgroup to keep ~256 lanes busy, but no
     * more vectors per workgroup than leave every CU two workgroups */
    want = (prec == 4) ? 4096 : 2048;
    vpb_along = (int)(want / n);
    if (nvec > 0) {
        int cus = 256;
        long fill;
        vhip_cu_count(&cus);
        if (cus < 1) cus = 256;
        fill = ((long)nvec + 2L * cus - 1) / (2L * cus);
        if (vpb_along > fill) vpb_along = (int)fill;
    }
    vpb_along = va_env_cap("VSIP_AMD_VPB_ALONG", vpb_along);
    if (vpb_along < 1) vpb_along = 1;
    if ((vsip_length)vpb_along > nvec && nvec > 0) vpb_along = (int)nvec;
    {
        /* A launch of at most two workgroups per CU runs each workgroup once
         * and is bound by one workgroup's latency: FP32 rows then take enough
         * lanes for one butterfly per lane in every stage (ROW 512 x 1001,
         * [13, 11, 7]: 128 -> 192 lanes, 6.53 -> 5.86 us; no gain for
         * larger grids, and FP64 loses) */
        int thr_min = va_env_floor("VSIP_AMD_THREADS_ALONG");
        if (prec == 4 && nvec > 0 && !thr_min) {
            int cus = 256, k;
            long most = 0;
            vhip_cu_count(&cus);
            if (cus < 1) cus = 256;
            if (((long)nvec + vpb_along - 1) / vpb_along <= 2L * cus) {
                for (k = 0; k < s->nst; k++)
                    if ((long)vpb_along * (n / s->radix[k]) > most) most = (long)vpb_along * (n / s->radix[k]);
                thr_min = most <= 1024 ? (int)most : 0;
            }
        }
        if (va_pick_launch(s, prec, vpb_along, 65536, lds_max, thr_min, &s->along)) return -2;
    }

    /* across: 16 vectors give 128-byte (float) row segments per load when
     * two such workgroups fit a CU; otherwise 4, whose neighbouring groups
     * share the remaining parts of each line through the XCD's L2 (the
     * kernel places consecutive groups on one XCD; measured faster than 8
     * vectors at one workgroup per CU) */
    vpb_across = (va_lds_bytes(s, prec, 16) <= lds_max / 2) ? 16 : 4;
    /* few vectors: fewer per workgroup, until at least half the CUs get one
     * (k-omega's 128-row real transform: 26.6 -> 19.2 us per step; 1000
     * columns keep 4, which beats 2 there) */
    if (nvec > 0) {
        int cus = 256;
        long half;
        vhip_cu_count(&cus);
        half = (cus > 1 ? cus : 256) / 2;
        if ((long)nvec < (long)vpb_across * half) {
            long v = (long)nvec / half;
            vpb_across = v < 1 ? 1 : (int)v;
        }
    }
    vpb_across = va_env_cap("VSIP_AMD_VPB_ACROSS", vpb_across);
    if ((vsip_length)vpb_across > nvec && nvec > 0) vpb_across = (int)nvec;
    {
        /* the rows' rule for launches of at most two workgroups per CU, for
         * FP32 columns too: one butterfly per lane in every stage */
        int thr_min = va_env_floor("VSIP_AMD_THREADS_ACROSS");
        if (prec == 4 && nvec > 0 && !thr_min) {
            int cus = 256, k;
            long most = 0;
            vhip_cu_count(&cus);
            if (cus < 1) cus = 256;
            if (((long)nvec + vpb_across - 1) / vpb_across <= 2L * cus) {
                for (k = 0; k < s->nst; k++)
                    if ((long)vpb_across * (n / s->radix[k]) > most) most = (long)vpb_across * (n / s->radix[k]);
                thr_min = most <= 1024 ? (int)most : 0;
            }
        }
        if (va_pick_launch(s, prec, vpb_across, lds_max, lds_max, thr_min, &s->across)) return -2;
    }

    /* stage twiddle offsets */
    off = 0;
    Ns = 1;
    for (i = 0; i < s->nst; i++) {
        s->tw_off[i] = off;
        if (Ns > 1) off += (s->radix[i] - 1) * Ns;
        Ns *= s->radix[i];
    }
    s->tw_len = off;
    return 0;
}

/*
 * Shape for a fixed workgroup size (the batched kernel): the given radix
 * sequence (or the default factorisation when radix == NULL), and the largest
 * vectors-per-workgroup up to vpb_max that fits `threads` lanes with register
 * class maxe and `lds_max` bytes of LDS.  Returns 0, -1 (unsupported length)
 * or -2 (does not fit).
 */
int va_sched_shape_fixed(VA_sched *s, int prec, int n, const int *radix, int nst,
                         int threads, int maxe, int lds_max, int vpb_max)
{
    int i, off, Ns, vpb;
    memset(s, 0, sizeof(*s));
    s->n = n;
    if (radix) {
        long prod = 1;
        if (nst < 1 || nst > VHIP_MAX_STAGES) return -1;
        for (i = 0; i < nst; i++) { s->radix[i] = radix[i]; prod *= radix[i]; }
        if (prod != n) return -1;
        s->nst = nst;
    } else {
        s->nst = va_factor(n, prec, s->radix);
        if (s->nst < 0) return -1;
    }
    s->lds_stride = va_pad(prec, n);
    if (s->lds_stride % 2 == 0) s->lds_stride += 1;
    for (vpb = vpb_max; vpb >= 1; vpb--) {
        if (va_lds_bytes(s, prec, vpb) > lds_max) continue;
        if (va_shape_ok(s, vpb, threads, maxe)) break;
    }
    if (vpb < 1) return -2;
    s->along.vpb = s->across.vpb = vpb;
    s->along.threads = s->across.threads = threads;
    s->along.maxe = s->across.maxe = maxe;
    off = 0;
    Ns = 1;
    for (i = 0; i < s->nst; i++) {
        s->tw_off[i] = off;
        if (Ns > 1) off += (s->radix[i] - 1) * Ns;
        Ns *= s->radix[i];
    }
    s->tw_len = off;
    return 0;
}

/* lg of the batched kernel's two-level twiddle table of length n:
 * ceil(ceil(log2 n) / 2) (the kernel derives the same value from n) */
static int va_btw_lg(int n)
{
    int cl = 0;
    while ((1L << cl) < (long)n) cl++;
    return (cl + 1) / 2;
}

/* the stage twiddles of a shaped schedule, uploaded to the device; float
 * schedules of two or more stages also carry the batched kernel's table of
 * the powers of W_n = e^{-2 pi i / n} in two levels (after the stage
 * twiddles, at btw_off): lo[t] = W_n^t for t < 2^lg, then
 * hi[m] = W_n^{m 2^lg} for m < ceil(n / 2^lg) */
int va_sched_upload(VA_sched *s, int prec)
{
    int i, r, k, Ns, off;
    int blg = 0, bl = 0, bh = 0;
    off = s->tw_len;
    s->btw_off = -1;
    if (prec == 4 && s->nst > 1) {
        blg = va_btw_lg(s->n);
        bl = 1 << blg;
        bh = (s->n + bl - 1) >> blg;
        s->btw_off = off;
        off += bl + bh;
    }
    if (off > 0) {
        size_t esz = (size_t)(2 * prec);
        char *host = (char *)malloc(esz * (size_t)off);
        int rc;
        if (!host) return -3;
        Ns = 1;
        for (i = 0; i < s->nst; i++) {
            int R = s->radix[i];
            if (Ns > 1) {
                long L = (long)Ns * R;
                for (r = 1; r < R; r++) {
                    for (k = 0; k < Ns; k++) {
                        long m = ((long)r * k) % L;
                        double a = -2.0 * VA_PI * (double)m / (double)L;
                        double c = cos(a), sn = sin(a);
                        long idx = s->tw_off[i] + (long)(r - 1) * Ns + k;
                        if (prec == 4) {
                            float *f = (float *)host;
                            f[2 * idx] = (float)c;
                            f[2 * idx + 1] = (float)sn;
                        } else {
                            double *d = (double *)host;
                            d[2 * idx] = c;
                            d[2 * idx + 1] = sn;
                        }
                    }
                }
            }
            Ns *= R;
        }
        for (i = 0; i < bl + bh; i++) {
            long e = i < bl ? (long)i : (long)(i - bl) << blg;
            double a = -2.0 * VA_PI * (double)e / (double)s->n;
            float *f = (float *)host + 2L * (s->btw_off + i);
            f[0] = (float)cos(a);
            f[1] = (float)sin(a);
        }
        rc = vhip_malloc(&s->dev_tw, esz * (size_t)off);
        if (!rc) rc = vhip_memcpy_h2d(s->dev_tw, host, esz * (size_t)off);
        free(host);
        if (rc) {
            if (s->dev_tw) vhip_free(s->dev_tw);
            s->dev_tw = NULL;
            return -3;
        }
    }
    return 0;
}

/* shape plus the stage twiddles uploaded to the device */
int va_sched_plan(VA_sched *s, int prec, int n, vsip_length nvec)
{
    int lds_max = 65536, rc;
    vhip_lds_max(&lds_max);
    rc = va_sched_shape(s, prec, n, nvec, lds_max);
    if (rc) return rc;
    return va_sched_upload(s, prec);
}

void va_sched_free(VA_sched *s)
{
    if (s->dev_tw) vhip_free(s->dev_tw);
    s->dev_tw = NULL;
}
