/*
 * fftm_oracle.c -- CPU oracle for the VSIPL multiple-FFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this file's library, and only to check
 * results.  The product library (jvsip_amd/lib/libvsip_amd.so) never calls
 * it.
 *
 * It restates, in double precision (long double for the brute-force DFT),
 * what rrjudd/jvsip computes on this path:
 *
 *  - view geometry (c_VSIP_src/vsip_cmviewattributes_f.h:47-55,
 *    VI_fftm_building_blocks_f.h:1044-1060): element (i, j) of a matrix view is
 *    block element offset + i*col_stride + j*row_stride; VSIP_ROW transforms
 *    each of the col_length rows (length row_length), VSIP_COL each of the
 *    row_length columns (length col_length).
 *  - ccfftm{ip,op} (VI_fftm_building_blocks_f.h:1064-1088,
 *    VI_ccfftmop_f_def.h:20-29): y_k = scale * sum_j x_j e^{dir 2 pi i jk/L}
 *    with dir = -1 forward, +1 inverse (vsip.h:151), no normalisation.
 *  - rcfftmop (VI_rcfftmop_f_def.h:24-120): per vector of real length L = 2n,
 *    the first n+1 bins of scale * sum_j x_j e^{-2 pi i jk/L}.  The reference
 *    gets them by packing x[2m] + i x[2m+1], an n-point FFT and the combine
 *    Z_k = 1/2[(Y_k + conj Y_{n-k}) - i e^{-i pi k/n}(Y_k - conj Y_{n-k})];
 *    for real input that equals the direct real DFT used here.
 *  - crfftmop (VI_crfftmop_f_def.h:25-126): per vector of n+1 complex inputs,
 *    Y_k = (Z_k + conj Z_{n-k}) + i e^{+i pi k/n}(Z_k - conj Z_{n-k}),
 *    k = 0..n-1 (no 1/2), y = scale * sum_k Y_k e^{+2 pi i km/n},
 *    x[2m] = Re y_m, x[2m+1] = Im y_m.  This is restated literally because it
 *    defines the result also for inputs that are not Hermitian.
 *
 * The transforms are a recursive mixed-radix decimation-in-time FFT (any
 * length; a prime length p costs O(p^2)) with twiddles from a long double
 * table, and a brute-force long double DFT for pinning small cases.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    long offset;
    long col_stride;
    long col_length;
    long row_stride;
    long row_length;
} oview;

#define O_PI 3.141592653589793238462643383279502884L

/* y[k] = sum_j x[j] exp(sign * 2 pi i j k / n), long double accumulation */
void oracle_dft_ld(long n, const double *x, int sign, double *y)
{
    long k, j;
    for (k = 0; k < n; k++) {
        long double sr = 0.0L, si = 0.0L;
        for (j = 0; j < n; j++) {
            long m = (long)(((long long)j * k) % n);
            long double a = (long double)sign * 2.0L * O_PI * (long double)m / (long double)n;
            long double c = cosl(a), s = sinl(a);
            sr += (long double)x[2 * j] * c - (long double)x[2 * j + 1] * s;
            si += (long double)x[2 * j] * s + (long double)x[2 * j + 1] * c;
        }
        y[2 * k] = (double)sr;
        y[2 * k + 1] = (double)si;
    }
}

static long o_smallest_factor(long n)
{
    long p;
    if (n % 2 == 0) return 2;
    for (p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* tw[t] = exp(sign 2 pi i t / n), t = 0..n-1 */
static double *o_twiddles(long n, int sign)
{
    double *tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    long t;
    if (!tw) return NULL;
    for (t = 0; t < n; t++) {
        long double a = (long double)sign * 2.0L * O_PI * (long double)t / (long double)n;
        tw[2 * t] = (double)cosl(a);
        tw[2 * t + 1] = (double)sinl(a);
    }
    return tw;
}

/*
 * out[k*os] = sum_j in[j*is] W^{jk}, length n, twiddle table tw of the full
 * length N with W_n = tw[N/n]; scratch holds 2*n doubles.
 */
static void o_fft_rec(long n, const double *in, long is, double *out, long os,
                      const double *tw, long N, double *scratch)
{
    long p, m, r, k, q;
    if (n == 1) {
        out[0] = in[0];
        out[1] = in[1];
        return;
    }
    p = o_smallest_factor(n);
    m = n / p;
    /* sub-transforms of the p decimated sequences into scratch[r*m + k] */
    for (r = 0; r < p; r++)
        o_fft_rec(m, in + 2 * r * is, is * p, scratch + 2 * r * m, 1, tw, N, scratch + 2 * n);
    /* combine: X[k + q m] = sum_r W_n^{r (k + q m)} Y_r[k] */
    for (k = 0; k < m; k++) {
        for (q = 0; q < p; q++) {
            long kk = k + q * m;
            double sr = 0.0, si = 0.0;
            for (r = 0; r < p; r++) {
                long t = ((r * kk) % n) * (N / n);
                double wr = tw[2 * t], wi = tw[2 * t + 1];
                double yr = scratch[2 * (r * m + k)], yi = scratch[2 * (r * m + k) + 1];
                sr += yr * wr - yi * wi;
                si += yr * wi + yi * wr;
            }
            out[2 * kk * os] = sr;
            out[2 * kk * os + 1] = si;
        }
    }
}

/* scratch requirement of o_fft_rec for length n */
static long o_scratch_len(long n)
{
    long total = 0, m = n;
    while (m > 1) {
        total += 2 * m;
        m /= o_smallest_factor(m);
    }
    return total + 2;
}

/* y = sum_j x_j exp(sign 2 pi i jk/n); x, y interleaved double, may not alias */
int oracle_fft_d(long n, const double *x, int sign, double *y)
{
    double *tw, *scratch;
    if (n <= 0) return 0;
    tw = o_twiddles(n, sign);
    scratch = (double *)malloc(sizeof(double) * (size_t)o_scratch_len(n));
    if (!tw || !scratch) { free(tw); free(scratch); return -1; }
    o_fft_rec(n, x, 1, y, 1, tw, n, scratch);
    free(tw);
    free(scratch);
    return 0;
}

/* axis geometry of a view for a given major */
static void o_axis(const oview *v, int major, long *es, long *vs, long *len, long *cnt)
{
    if (major == 0) { *es = v->row_stride; *vs = v->col_stride; *len = v->row_length; *cnt = v->col_length; }
    else { *es = v->col_stride; *vs = v->row_stride; *len = v->col_length; *cnt = v->row_length; }
}

/*
 * ccfftmip / ccfftmop.  xblk, yblk: interleaved complex blocks (double);
 * for the in-place form pass the same block and view twice.
 */
int oracle_ccfftm(const double *xblk, oview xv, double *yblk, oview yv, int major, int dir,
                  double scale)
{
    long es, vs, n, cnt, eso, vso, no, cnto, v, j;
    double *a, *b, *tw, *scratch;
    o_axis(&xv, major, &es, &vs, &n, &cnt);
    o_axis(&yv, major, &eso, &vso, &no, &cnto);
    if (n != no || cnt != cnto) return -1;
    if (n <= 0 || cnt <= 0) return 0;
    a = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    b = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    tw = o_twiddles(n, dir < 0 ? -1 : 1);
    scratch = (double *)malloc(sizeof(double) * (size_t)o_scratch_len(n));
    if (!a || !b || !tw || !scratch) { free(a); free(b); free(tw); free(scratch); return -2; }
    for (v = 0; v < cnt; v++) {
        for (j = 0; j < n; j++) {
            long e = xv.offset + v * vs + j * es;
            a[2 * j] = xblk[2 * e];
            a[2 * j + 1] = xblk[2 * e + 1];
        }
        o_fft_rec(n, a, 1, b, 1, tw, n, scratch);
        for (j = 0; j < n; j++) {
            long e = yv.offset + v * vso + j * eso;
            yblk[2 * e] = scale * b[2 * j];
            yblk[2 * e + 1] = scale * b[2 * j + 1];
        }
    }
    free(a); free(b); free(tw); free(scratch);
    return 0;
}

/* rcfftmop: real block x (length 2n per vector) -> complex block z (n+1) */
int oracle_rcfftm(const double *xblk, oview xv, double *zblk, oview zv, int major, double scale)
{
    long es, vs, L, cnt, eso, vso, no, cnto, v, j, n;
    double *a, *b, *tw, *scratch;
    o_axis(&xv, major, &es, &vs, &L, &cnt);
    o_axis(&zv, major, &eso, &vso, &no, &cnto);
    n = L / 2;
    if (L % 2 || no != n + 1 || cnt != cnto) return -1;
    if (cnt <= 0) return 0;
    a = (double *)malloc(sizeof(double) * 2 * (size_t)L);
    b = (double *)malloc(sizeof(double) * 2 * (size_t)L);
    tw = o_twiddles(L, -1);
    scratch = (double *)malloc(sizeof(double) * (size_t)o_scratch_len(L));
    if (!a || !b || !tw || !scratch) { free(a); free(b); free(tw); free(scratch); return -2; }
    for (v = 0; v < cnt; v++) {
        for (j = 0; j < L; j++) {
            a[2 * j] = xblk[xv.offset + v * vs + j * es];
            a[2 * j + 1] = 0.0;
        }
        o_fft_rec(L, a, 1, b, 1, tw, L, scratch);
        for (j = 0; j <= n; j++) {
            long e = zv.offset + v * vso + j * eso;
            zblk[2 * e] = scale * b[2 * j];
            zblk[2 * e + 1] = scale * b[2 * j + 1];
        }
    }
    free(a); free(b); free(tw); free(scratch);
    return 0;
}

/* crfftmop: complex block z (n+1 per vector) -> real block x (2n) */
int oracle_crfftm(const double *zblk, oview zv, double *xblk, oview xv, int major, double scale)
{
    long es, vs, n1, cnt, eso, vso, L, cnto, v, k, n;
    double *a, *b, *tw, *scratch;
    o_axis(&zv, major, &es, &vs, &n1, &cnt);
    o_axis(&xv, major, &eso, &vso, &L, &cnto);
    n = n1 - 1;
    if (n < 1 || L != 2 * n || cnt != cnto) return -1;
    if (cnt <= 0) return 0;
    a = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    b = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    tw = o_twiddles(n, +1);
    scratch = (double *)malloc(sizeof(double) * (size_t)o_scratch_len(n));
    if (!a || !b || !tw || !scratch) { free(a); free(b); free(tw); free(scratch); return -2; }
    for (v = 0; v < cnt; v++) {
        for (k = 0; k < n; k++) {
            long ef = zv.offset + v * vs + k * es;
            long eb = zv.offset + v * vs + (n - k) * es;
            long double fr = zblk[2 * ef], fi = zblk[2 * ef + 1];
            long double br = zblk[2 * eb], bi = -zblk[2 * eb + 1];      /* conj Z_{n-k} */
            long double ang = O_PI * (long double)k / (long double)n;
            long double wr = cosl(ang), wi = sinl(ang);
            long double ar = fr + br, ai = fi + bi;                     /* Z_k + conj Z_{n-k} */
            long double dr = fr - br, di = fi - bi;                     /* Z_k - conj Z_{n-k} */
            long double pr = wr * dr - wi * di, pi = wr * di + wi * dr; /* W D */
            a[2 * k] = (double)(ar - pi);                               /* A + i W D */
            a[2 * k + 1] = (double)(ai + pr);
        }
        o_fft_rec(n, a, 1, b, 1, tw, n, scratch);
        for (k = 0; k < n; k++) {
            xblk[xv.offset + v * vso + (2 * k) * eso] = scale * b[2 * k];
            xblk[xv.offset + v * vso + (2 * k + 1) * eso] = scale * b[2 * k + 1];
        }
    }
    free(a); free(b); free(tw); free(scratch);
    return 0;
}
