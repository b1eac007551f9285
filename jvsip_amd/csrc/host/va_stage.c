/*
 * va_stage.c -- staged transforms: the lengths the single-pass kernels and
 * the two-pass split do not serve.
 *
 * The reference transforms every length: its factoriser leaves a remainder
 * L whose prime factors are above 8 to a direct O(L^2) DFT
 * (VI_nuV.h:129-134, VI_fftm_building_blocks_f.h:95-155), and its real
 * transforms wrap a complex FFT of any length (VI_rcfftmop_f_def.h:53-115,
 * VI_crfftmop_f_def.h:74-124).  Here such plans run in stages over
 * contiguous scratch rows, `chunk` vectors at a time:
 *
 *   PRE   view -> scratch (real packing or Hermitian pre-combine, chirp,
 *         zero padding)
 *   FFT   forward c2c of length sm on the scratch rows (an ordinary
 *         ccfftmip plan: one pass, or the two-pass split)
 *   MUL   Bluestein only: times the chirp filter's spectrum, then FFT again
 *   POST  scratch -> view (chirp and index reversal, Euler combine or real
 *         unpacking, conjugation, scale)
 *
 * Bluestein (chirp-z): with a_j = e^{-i pi j^2 / n},
 *   y_k = sum_j x_j e^{-2 pi i jk/n} = a_k sum_j (x_j a_j) conj(a_{k-j}),
 * a linear convolution, computed as a cyclic one of length sm >= 2n - 1:
 *   conv = IDFT_sm(DFT_sm(u) .* B),  B = DFT_sm(b),  b_l = conj(a_l) for
 *   0 <= l < n and b_{sm-l} = conj(a_l) for 0 < l < n, zero elsewhere,
 * and IDFT_sm(W)_k = DFT_sm(W)_{(sm-k) mod sm} / sm, so both transforms are
 * forward and B / sm is the table of the MUL pass.  The chirp angles use
 * j^2 mod 2n (exact integers), and B is computed on the host in double
 * precision with the mixed-radix FFT below, then rounded to the plan's
 * precision.  Inverse transforms conjugate on the way in and out.
 *
 * C89.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "va_internal.h"

#define VA_PI 3.14159265358979323846264338327950288
#define VA_STAGE_BYTES ((size_t)256 << 20)   /* scratch per slab */

/* Bluestein's convolution length for an n-point DFT: the smallest
 * 2^a 3^b 5^c >= 2n - 1, or the power of two when it is within 25% of that */
long va_bluestein_length(long n)
{
    long best = -1, p2 = 1, a, b, c, t = 2 * n - 1;
    if (t < 1) t = 1;
    while (p2 < t) p2 *= 2;
    for (a = 1; a <= p2; a *= 2)
        for (b = a; b <= p2; b *= 3)
            for (c = b; c <= p2; c *= 5)
                if (c >= t && (best < 0 || c < best)) best = c;
    if (best < 0 || 4 * p2 <= 5 * best) return p2;
    return best;
}

/* scratch bytes per slab; VSIP_AMD_STAGE_BYTES overrides the default */
static size_t va_stage_bytes(void)
{
    const char *e = getenv("VSIP_AMD_STAGE_BYTES");
    if (e) {
        long v = atol(e);
        if (v > 0) return (size_t)v;
    }
    return VA_STAGE_BYTES;
}

static long va_smallest_factor(long n)
{
    long p;
    if (n % 2 == 0) return 2;
    for (p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/*
 * Decimation in time: the n-point DFT of x[0], x[xs], ... is assembled from
 * the p DFTs of length n/p of the decimated sequences (p the smallest prime
 * factor of n), written to y[0 .. n).  w holds e^{-2 pi i j / N}, j < N, for
 * the top length N (n divides N); tmp has room for n complex values.
 */
static void va_fft_rec(long n, const double *x, long xs, double *y, const double *w, long N,
                       double *tmp)
{
    long p, m, r, k, q;
    if (n == 1) {
        y[0] = x[0];
        y[1] = x[1];
        return;
    }
    p = va_smallest_factor(n);
    m = n / p;
    for (r = 0; r < p; r++) va_fft_rec(m, x + 2 * r * xs, xs * p, y + 2 * r * m, w, N, tmp);
    for (k = 0; k < m; k++)
        for (q = 0; q < p; q++) {
            long kk = k + m * q;
            double sr = 0.0, si = 0.0;
            for (r = 0; r < p; r++) {
                long e = ((r * kk) % n) * (N / n);
                double cr = w[2 * e], ci = w[2 * e + 1];
                double ar = y[2 * (r * m + k)], ai = y[2 * (r * m + k) + 1];
                sr += ar * cr - ai * ci;
                si += ar * ci + ai * cr;
            }
            tmp[2 * kk] = sr;
            tmp[2 * kk + 1] = si;
        }
    memcpy(y, tmp, sizeof(double) * 2 * (size_t)n);
}

/* in-place forward DFT of n complex doubles (interleaved) on the host */
void va_host_fft(long n, double *x)
{
    double *w, *y, *tmp;
    long j;
    if (n <= 1) return;
    w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    y = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    tmp = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    if (!w || !y || !tmp) {
        free(w); free(y); free(tmp);
        return;
    }
    for (j = 0; j < n; j++) {
        double a = -2.0 * VA_PI * (double)j / (double)n;
        w[2 * j] = cos(a);
        w[2 * j + 1] = sin(a);
    }
    va_fft_rec(n, x, 1, y, w, n, tmp);
    memcpy(x, y, sizeof(double) * 2 * (size_t)n);
    free(w); free(y); free(tmp);
}

static int va_upload_cplx(void **dev, int prec, const double *v, long count)
{
    size_t esz = (size_t)(2 * prec);
    void *host = malloc(esz * (size_t)count);
    long k;
    int rc;
    if (!host) return -1;
    for (k = 0; k < 2 * count; k++) {
        if (prec == 4) ((float *)host)[k] = (float)v[k];
        else ((double *)host)[k] = v[k];
    }
    rc = vhip_malloc(dev, esz * (size_t)count);
    if (!rc) rc = vhip_memcpy_h2d(*dev, host, esz * (size_t)count);
    free(host);
    if (rc && *dev) { vhip_free(*dev); *dev = NULL; }
    return rc;
}

/* chirp a_j = e^{-i pi j^2 / n}, j < n, and the MUL table DFT_sm(b) / sm */
static int va_bluestein_tables(VA_fftm *p)
{
    long n = (long)p->MN, sm = (long)p->sm, j;
    double *a = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *b = (double *)calloc(2 * (size_t)sm, sizeof(double));
    int rc = -1;
    if (!a || !b) goto done;
    for (j = 0; j < n; j++) {
        unsigned long q = ((unsigned long)j * (unsigned long)j) % (2UL * (unsigned long)n);
        double ang = -VA_PI * (double)q / (double)n;
        a[2 * j] = cos(ang);
        a[2 * j + 1] = sin(ang);
    }
    for (j = 0; j < n; j++) {
        b[2 * j] = a[2 * j];
        b[2 * j + 1] = -a[2 * j + 1];
        if (j) {
            b[2 * (sm - j)] = a[2 * j];
            b[2 * (sm - j) + 1] = -a[2 * j + 1];
        }
    }
    va_host_fft(sm, b);
    for (j = 0; j < 2 * sm; j++) b[j] /= (double)sm;
    rc = va_upload_cplx(&p->dev_chirp, p->prec, a, n);
    if (!rc) rc = va_upload_cplx(&p->dev_filt, p->prec, b, sm);
done:
    free(a);
    free(b);
    return rc;
}

int va_staged_init(VA_fftm *p, int bluestein)
{
    size_t row;
    long sm = bluestein ? va_bluestein_length((long)p->MN) : (long)p->MN;
    if (sm < 1 || sm > 0x7fffffffL) return -1;
    p->staged = bluestein ? 2 : 1;
    p->sm = (vsip_length)sm;
    row = (size_t)sm * 2 * (size_t)p->prec;
    p->chunk = (vsip_length)(va_stage_bytes() / row);
    if (p->chunk < 1) p->chunk = 1;
    if (p->chunk > p->mN && p->mN > 0) p->chunk = p->mN;
    p->inner = (VA_fftm *)calloc(1, sizeof(VA_fftm));
    if (!p->inner) return -1;
    if (va_fftm_init_inner(p->inner, p->prec, p->chunk, p->sm)) {
        va_fftm_release(p->inner);
        free(p->inner);
        p->inner = NULL;
        return -1;
    }
    if (vhip_malloc(&p->dev_stage, (size_t)p->chunk * row)) return -1;
    if (bluestein && va_bluestein_tables(p)) return -1;
    return 0;
}

void va_staged_release(VA_fftm *p)
{
    if (p->inner) {
        va_fftm_release(p->inner);
        free(p->inner);
        p->inner = NULL;
    }
    if (p->dev_chirp) vhip_free(p->dev_chirp);
    if (p->dev_filt) vhip_free(p->dev_filt);
    if (p->dev_stage) vhip_free(p->dev_stage);
    p->dev_chirp = p->dev_filt = p->dev_stage = NULL;
    p->staged = 0;
}

static long va_labs(long x) { return x < 0 ? -x : x; }

/* the inner forward FFT over the first `cnt` scratch rows */
static int va_inner_fft(const VA_fftm *p, long cnt, void *stream)
{
    VA_block blk;
    VA_mview sv;
    vhip_fftm_desc d[VA_MAX_DESC];
    int nd, k, rc = 0;
    memset(&blk, 0, sizeof(blk));
    blk.prec = p->prec;
    blk.cplx = 1;
    blk.kind = VA_LIB_BLOCK;
    blk.admitted = 1;
    blk.markings = VA_VALID;
    blk.size = p->chunk * p->sm;
    blk.dev_r = p->dev_stage;
    memset(&sv, 0, sizeof(sv));
    sv.block = &blk;
    sv.row_stride = 1;
    sv.row_length = p->sm;
    sv.col_stride = (vsip_stride)p->sm;
    sv.col_length = p->chunk;
    sv.markings = VA_VALID;
    nd = va_fftm_descs(p->inner, &sv, &sv, NULL, NULL, NULL, NULL, d);
    if (nd < 0) return -1;
    for (k = 0; k < nd && !rc; k++) {
        /* every launch covers whole rows: vectors per row times rows */
        d[k].nvec = d[k].nvec / (long)p->chunk * cnt;
        rc = vhip_fftm_launch(&d[k], stream);
    }
    return rc;
}

int va_staged_exec(const VA_fftm *p, const VA_mview *in, const VA_mview *out, void *stream)
{
    long es_i, vs_i, es_o, vs_o, v0, mN = (long)p->mN;
    vhip_view gi, go;
    vhip_xform_desc x;
    int kind, rc = 0;
    kind = va_exec_views(p, in, out, &gi, &go, &es_i, &vs_i, &es_o, &vs_o);
    if (kind < 0) return -1;
    gi.es = es_i; gi.vs0 = vs_i; gi.nv0 = mN; gi.vs1 = 0; gi.contig = 0;
    go.es = es_o; go.vs0 = vs_o; go.nv0 = mN; go.vs1 = 0; go.contig = 0;
    memset(&x, 0, sizeof(x));
    x.prec = p->prec;
    x.kind = kind;
    x.conj = (kind == VHIP_C2C && p->dir > 0);
    x.n = (long)p->MN;
    x.m = (long)p->sm;
    x.scale = p->scale;
    x.euler = p->dev_euler;
    x.work = p->dev_stage;
    for (v0 = 0; v0 < mN && !rc; v0 += (long)p->chunk) {
        long cnt = mN - v0 < (long)p->chunk ? mN - v0 : (long)p->chunk;
        x.nvec = cnt;
        x.vbase = v0;
        /* PRE */
        x.step = VHIP_XF_PRE;
        x.view = gi;
        x.across = (cnt > 1 && va_labs(vs_i) < va_labs(es_i));
        x.chirp = (p->staged == 2) ? p->dev_chirp : NULL;
        x.reverse = 0;
        rc = vhip_xform_launch(&x, stream);
        if (!rc) rc = va_inner_fft(p, cnt, stream);
        if (!rc && p->staged == 2) {
            x.step = VHIP_XF_MUL;
            x.table = p->dev_filt;
            x.across = 0;
            rc = vhip_xform_launch(&x, stream);
            if (!rc) rc = va_inner_fft(p, cnt, stream);
        }
        /* POST */
        if (!rc) {
            x.step = VHIP_XF_POST;
            x.view = go;
            x.across = (cnt > 1 && va_labs(vs_o) < va_labs(es_o));
            x.reverse = (p->staged == 2);
            rc = vhip_xform_launch(&x, stream);
        }
    }
    if (rc) va_error("fftm", vhip_error_string(rc), rc);
    return rc;
}
