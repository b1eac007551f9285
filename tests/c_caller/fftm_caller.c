/*
 * fftm_caller.c -- a plain C89 VSIPL program, written the way a jvsip user
 * writes one (bind a user array, admit, bind a view, create a plan,
 * execute, release), built against include/vsip.h and linked against
 * libvsip_amd.so.  It runs ccfftmip (ROW, forward), ccfftmop (COL,
 * inverse, scale 0.5) and an rcfftmop/crfftmop round trip, and checks each
 * against a direct DFT in double precision.  Prints one line per case and
 * "ALL OK" at the end; exit status 0 on success.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "vsip.h"

#define PI 3.14159265358979323846

static unsigned long state = 12345UL;
static float frand(void)
{
    state = state * 6364136223846793005UL + 1442695040888963407UL;
    return (float)((double)(state >> 11) / 9007199254740992.0 * 2.0 - 1.0);
}

/* y_k = scale * sum_j x_j e^{sign 2 pi i jk/n}, x at stride xs, y dense */
static void dft(int n, const float *x, long xs, int sign, double scale, double *y)
{
    int k, j;
    for (k = 0; k < n; k++) {
        double sr = 0.0, si = 0.0;
        for (j = 0; j < n; j++) {
            double a = sign * 2.0 * PI * (double)((long)j * k % n) / n;
            double xr = x[2 * j * xs], xi = x[2 * j * xs + 1];
            sr += xr * cos(a) - xi * sin(a);
            si += xr * sin(a) + xi * cos(a);
        }
        y[2 * k] = scale * sr;
        y[2 * k + 1] = scale * si;
    }
}

static int report(const char *name, double num, double den)
{
    double e = sqrt(num / den);
    printf("%-44s rel L2 %.3e %s\n", name, e, e < 1e-5 ? "ok" : "FAIL");
    return e < 1e-5 ? 0 : 1;
}

int main(void)
{
    const int M = 24, N = 60;           /* 60 = 4 * 3 * 5 */
    float *a = (float *)malloc(sizeof(float) * 2 * M * N);
    float *a0 = (float *)malloc(sizeof(float) * 2 * M * N);
    float *c = (float *)malloc(sizeof(float) * 2 * M * N);
    float *r = (float *)malloc(sizeof(float) * M * N);
    float *r0 = (float *)malloc(sizeof(float) * M * N);
    float *z = (float *)malloc(sizeof(float) * 2 * M * (N / 2 + 1));
    double *y = (double *)malloc(sizeof(double) * 2 * (M > N ? M : N));
    vsip_scalar_f *pr, *pi;
    vsip_cblock_f *ba, *bc, *bz;
    vsip_block_f *br;
    vsip_cmview_f *va, *vc, *vz;
    vsip_mview_f *vr;
    vsip_fftm_f *f;
    int i, k, fails = 0;
    double num, den;

    if (vsip_init(NULL)) { printf("vsip_init failed\n"); return 2; }
    for (i = 0; i < 2 * M * N; i++) a[i] = a0[i] = frand();
    for (i = 0; i < M * N; i++) r[i] = r0[i] = frand();

    /* 1. ccfftmip ROW forward, scale 1, on a user block */
    ba = vsip_cblockbind_f(a, NULL, (vsip_length)(M * N), VSIP_MEM_NONE);
    if (!ba || vsip_cblockadmit_f(ba, VSIP_TRUE)) { printf("bind/admit failed\n"); return 2; }
    va = vsip_cmbind_f(ba, 0, N, M, 1, N);
    f = vsip_ccfftmip_create_f(M, N, 1.0f, VSIP_FFT_FWD, VSIP_ROW, 0, VSIP_ALG_TIME);
    if (!va || !f) { printf("view/plan failed\n"); return 2; }
    vsip_ccfftmip_f(f, va);
    vsip_cblockrelease_f(ba, VSIP_TRUE, &pr, &pi);
    if (pr != a || pi != NULL) fails++;
    num = den = 0.0;
    for (i = 0; i < M; i++) {
        dft(N, a0 + 2 * i * N, 1, -1, 1.0, y);
        for (k = 0; k < 2 * N; k++) {
            double d = a[2 * i * N + k] - y[k];
            num += d * d;
            den += y[k] * y[k];
        }
    }
    fails += report("ccfftmip_f ROW 24x60 forward", num, den);
    vsip_fftm_destroy_f(f);
    vsip_cmalldestroy_f(va);

    /* 2. ccfftmop COL inverse, scale 0.5, into a library block */
    for (i = 0; i < 2 * M * N; i++) a[i] = a0[i];
    ba = vsip_cblockbind_f(a, NULL, (vsip_length)(M * N), VSIP_MEM_NONE);
    vsip_cblockadmit_f(ba, VSIP_TRUE);
    va = vsip_cmbind_f(ba, 0, N, M, 1, N);
    bc = vsip_cblockbind_f(c, NULL, (vsip_length)(M * N), VSIP_MEM_NONE);
    vsip_cblockadmit_f(bc, VSIP_FALSE);
    vc = vsip_cmbind_f(bc, 0, N, M, 1, N);
    f = vsip_ccfftmop_create_f(M, N, 0.5f, VSIP_FFT_INV, VSIP_COL, 0, VSIP_ALG_TIME);
    vsip_ccfftmop_f(f, va, vc);
    vsip_cblockrelease_f(bc, VSIP_TRUE, &pr, &pi);
    vsip_cblockrelease_f(ba, VSIP_TRUE, &pr, &pi);
    num = den = 0.0;
    for (i = 0; i < N; i++) {
        dft(M, a0 + 2 * i, N, +1, 0.5, y);
        for (k = 0; k < M; k++) {
            double dr = c[2 * (k * N + i)] - y[2 * k], di = c[2 * (k * N + i) + 1] - y[2 * k + 1];
            num += dr * dr + di * di;
            den += y[2 * k] * y[2 * k] + y[2 * k + 1] * y[2 * k + 1];
        }
    }
    for (i = 0; i < 2 * M * N; i++)
        if (a[i] != a0[i]) { fails++; break; }       /* ccfftmop keeps its input */
    fails += report("ccfftmop_f COL 24x60 inverse scale 0.5", num, den);
    vsip_fftm_destroy_f(f);
    vsip_cmalldestroy_f(va);
    vsip_cmalldestroy_f(vc);

    /* 3. rcfftmop ROW then crfftmop ROW with scale 1/N: the identity */
    br = vsip_blockbind_f(r, (vsip_length)(M * N), VSIP_MEM_NONE);
    vsip_blockadmit_f(br, VSIP_TRUE);
    vr = vsip_mbind_f(br, 0, N, M, 1, N);
    bz = vsip_cblockbind_f(z, NULL, (vsip_length)(M * (N / 2 + 1)), VSIP_MEM_NONE);
    vsip_cblockadmit_f(bz, VSIP_FALSE);
    vz = vsip_cmbind_f(bz, 0, N / 2 + 1, M, 1, N / 2 + 1);
    f = vsip_rcfftmop_create_f(M, N, 1.0f, VSIP_ROW, 0, VSIP_ALG_TIME);
    vsip_rcfftmop_f(f, vr, vz);
    vsip_fftm_destroy_f(f);
    f = vsip_crfftmop_create_f(M, N, 1.0f / N, VSIP_ROW, 0, VSIP_ALG_TIME);
    vsip_crfftmop_f(f, vz, vr);
    vsip_fftm_destroy_f(f);
    vsip_blockrelease_f(br, VSIP_TRUE);
    num = den = 0.0;
    for (i = 0; i < M * N; i++) {
        double d = r[i] - r0[i];
        num += d * d;
        den += (double)r0[i] * r0[i];
    }
    fails += report("rcfftmop_f -> crfftmop_f(1/N) ROW 24x60", num, den);
    vsip_mdestroy_f(vr);
    vsip_blockdestroy_f(br);
    vsip_cmalldestroy_f(vz);

    vsip_finalize(NULL);
    free(a); free(a0); free(c); free(r); free(r0); free(z); free(y);
    printf(fails ? "FAILED %d\n" : "ALL OK\n", fails);
    return fails ? 1 : 0;
}
