/*
 * vsip_amd.h -- extensions of the MI355X VSIPL FFTM library.
 *
 * Nothing here has a counterpart in the reference's vsip.h.  These entry
 * points serve callers that need to see the device side: the library stream
 * (to time kernels with HIP events or to order their own work after an FFT),
 * device-resident block data, error text, and the HBM budget that bounds the
 * library's allocations.
 *
 * C89.
 */
#ifndef _vsip_amd_h
#define _vsip_amd_h 1

#include <stddef.h>
#include "vsip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t of the library stream all plain vsip_* calls run on */
void *vsip_amd_stream(void);
/* wait for all queued work; returns 0 or a HIP error code */
int vsip_amd_sync(void);
/* text of the most recent error reported by the library ("" if none) */
const char *vsip_amd_last_error(void);
void vsip_amd_clear_error(void);
/* device bytes currently held by the library */
size_t vsip_amd_bytes_in_use(void);
/* cap the library's device allocations at `bytes` (0 = no cap); allocations
 * beyond the cap fail as if HBM were exhausted (create returns NULL, admit 1) */
void vsip_amd_set_hbm_budget(size_t bytes);

/* device address of a block's data (NULL unless the block holds device data);
 * for split complex blocks `im` receives the imaginary array */
void *vsip_amd_cblock_device_f(const vsip_cblock_f *b, void **im);
void *vsip_amd_cblock_device_d(const vsip_cblock_d *b, void **im);
void *vsip_amd_block_device_f(const vsip_block_f *b);
void *vsip_amd_block_device_d(const vsip_block_d *b);

/* HBM reference rates on this device: four buffers of `bytes` each (rounded
 * down to whole 32 KiB rows) visited in turn (so caches do not serve the
 * data), `reps` visits timed with stream events; *rmw_gbs gets the
 * in-place read-modify-write rate (bytes read + written per second / 1e9)
 * of the in-place transforms' pattern -- each workgroup reads a 32 KiB row,
 * synchronises and writes it back, with non-temporal loads and stores --
 * and *read_gbs the read-only rate.  Returns 0 or an error code. */
int vsip_amd_hbm_probe(size_t bytes, int reps, double *rmw_gbs, double *read_gbs);

/* The column-mapped transforms' access pattern with no transform in it:
 * for a rows x cols complex FP32 view whose rows are in_row_stride elements
 * apart, each workgroup reads every row of cols_per_group (1, 2, 4, 8 or 16;
 * rows x cols_per_group <= 16384) adjacent columns into registers,
 * synchronises and writes them to a view whose rows are out_row_stride
 * apart; consecutive groups run on one XCD, as in the column kernels.  Four
 * input/output pairs visited in turn, `reps` visits timed with stream
 * events; *gbs gets 2 x 8 x rows x cols bytes per visit / seconds / 1e9: the
 * ceiling of that pattern at that size.  Returns 0, -1 for arguments out of
 * range, or an error code. */
int vsip_amd_col_probe(long rows, long cols, long in_row_stride, long out_row_stride, int cols_per_group,
                       int reps, double *gbs);

/* Single-pass schedule the planner picks for `nvec` transforms of length n
 * (prec 4 or 8) on a device with `lds_max` bytes of LDS per workgroup, as
 * JSON text; no device work.  Returns 0, -1 for an unsupported length, -2 if
 * one vector does not fit a workgroup (such lengths run as two passes). */
int vsip_amd_schedule(int prec, long n, long nvec, int lds_max, char *buf, size_t len);

/* Schedule of a plan as text: passes, radices, launch shapes. */
int vsip_amd_fftm_describe_f(const vsip_fftm_f *plan, char *buf, size_t len);
int vsip_amd_fftm_describe_d(const vsip_fftm_d *plan, char *buf, size_t len);

/* ---- batched multiple FFTs: one persistent launch in a HIP graph --------
 * njobs (plan, input view, output view) triples; ccfftmip_f plans run in
 * place on ins[i] (outs may be NULL), ccfftmop_f plans write outs[i].  Every
 * transform runs with the schedule the autotuner picked for its length and
 * the batched workgroup shape.  run() replays the captured graph on the
 * library stream (asynchronous, ordered with the plain calls).  The views'
 * blocks must stay admitted for the life of the batch.  Returns NULL on
 * failure (vsip_amd_last_error says why). */
typedef struct vsip_amd_batch_f vsip_amd_batch_f;
vsip_amd_batch_f *vsip_amd_batch_create_f(vsip_length njobs, vsip_fftm_f *const *plans,
                                          const vsip_cmview_f *const *ins,
                                          const vsip_cmview_f *const *outs);
int vsip_amd_batch_run_f(vsip_amd_batch_f *b);
void vsip_amd_batch_destroy_f(vsip_amd_batch_f *b);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the transforms of vsip_amd_batch_create_f +
 * vsip_amd_batch_run_f in one call that allocates nothing.  The launch
 * descriptors, work items and counters live in a static device scratch of
 * 4 MiB per GPU (held until the process ends), so a call takes at most
 * (4 MiB - 256 B) / (464 B per job + 16 B per work item) of them.  Every
 * job must be a complex FP32 ccfftmip/ccfftmop plan whose length is single
 * pass in the batched shape and already tuned (vsip_amd_tune, or a
 * vsip_amd_batch_create_f over it, tunes and allocates once beforehand).
 * Runs asynchronously on the library stream.  Calls share the scratch and
 * are ordered by the library stream; calls from several host threads at once
 * race.  Returns 0, or -1 with the reason in vsip_amd_last_error (sizes
 * beyond the scratch, an untuned or two-pass length) before any launch. */
int vsip_amd_batch_run_noalloc_f(vsip_length njobs, vsip_fftm_f *const *plans,
                                 const vsip_cmview_f *const *ins, const vsip_cmview_f *const *outs);
/* JSON: jobs, descriptors, work items, grid, LDS bytes, elements */
int vsip_amd_batch_info_f(const vsip_amd_batch_f *b, char *buf, size_t len);
/* the batched schedule the autotuner picks for length n (tunes on first use;
 * VSIP_AMD_TUNE_CACHE=<file> keeps results across processes), as JSON */
int vsip_amd_tune(int prec, long n, char *buf, size_t len);

/* ---- serving: continuous batching of arriving FFT requests ---------------
 * A request is a rows x n complex FP32 transform (ROW, dense rows), forward
 * (dir -1) or inverse (+1), times scale, read from the input pool block at
 * complex element in_offset and written to the output pool block at
 * out_offset, or (out_offset < 0) to memory the server takes from its HBM
 * budget and returns when the request completes.  Requests must be sorted by
 * t_arrive (seconds after the run starts).  Arrived requests are admitted
 * while their output (and scratch) fits the budget, on one of two paths:
 *   * resident: requests of at most VSIP_AMD_SERVE_RESIDENT elements (default
 *     131072; 0 turns the path off) whose length runs in one pass of the
 *     batched workgroup join one persistent kernel that stays on the device
 *     for the whole stream (VSIP_AMD_SERVE_RESIDENT_WGS workgroups, default
 *     32): admission writes the request's descriptor to pinned host memory
 *     and publishes its work items, the running kernel picks them up with no
 *     launch, and a per-request flag in host memory reports completion;
 *   * lanes: the other requests, up to max_jobs per group, launch on a free
 *     lane (a stream; lanes overlap on the device and with the resident
 *     kernel): requests of at least VSIP_AMD_SERVE_DIRECT elements (default
 *     32768) launch their own length-specialised kernel, and the smaller ones
 *     of the group share one persistent batched kernel when there are more
 *     than two of them.
 * With lanes == 1 and max_jobs == 1 every request runs alone, in arrival
 * order.  Finished requests are retired as they complete; res[i] gets the
 * admission and completion times (res[i].stream == lanes for the resident
 * path).  stats receives a JSON summary (with the counts of resident,
 * directly launched and batched requests).  Returns 0 or an error code. */
typedef struct {
    double t_arrive;
    long rows;
    long n;
    int dir;
    float scale;
    long in_offset;
    long out_offset;
} vsip_amd_request;

typedef struct {
    double t_admit;
    double t_done;
    int batch;
    int stream;
} vsip_amd_result;

int vsip_amd_serve_f(const vsip_amd_request *reqs, long nreq, const vsip_cblock_f *in_pool,
                     const vsip_cblock_f *out_pool, size_t hbm_budget, int lanes, int max_jobs,
                     vsip_amd_result *res, char *stats, size_t stats_len);
/* the same stream served by plain calls in arrival order on the library
 * stream, each launched when its request arrives: with sync_each != 0 every
 * call is followed by a synchronisation (a caller that waits for each
 * result), otherwise the calls go back to back and res[i].t_done is when the
 * event recorded after request i is seen complete */
int vsip_amd_serve_plain_f(const vsip_amd_request *reqs, long nreq, const vsip_cblock_f *in_pool,
                           const vsip_cblock_f *out_pool, size_t hbm_budget, int sync_each,
                           vsip_amd_result *res, char *stats, size_t stats_len);

/* One k-omega beamformer step, the komega() of the reference's example
 * (examples/komegaExamples/kw_c/kw.c:7-25), in two fused launches:
 *     gram += gain * |ccfftm( rcfftm( data .* ts_taper (rows) .* array_taper (columns) ) )|^2
 * rcfftm: an rcfftmop ROW plan of data's shape (M x N, N even); ccfftm: a
 * ccfftmip or ccfftmop COL plan of freq's shape (M x N/2+1), both single
 * pass; freq: interleaved complex; gram: real, with freq's strides.  data is
 * not modified (the reference tapers it in place); freq holds the row
 * transforms of the tapered data afterwards.  Returns 0, or -1 with the
 * reason in vsip_amd_last_error. */
int vsip_amd_komega_f(const vsip_fftm_f *rcfftm, const vsip_fftm_f *ccfftm, const vsip_vview_f *ts_taper,
                      const vsip_vview_f *array_taper, const vsip_mview_f *data, const vsip_cmview_f *freq,
                      const vsip_mview_f *gram, vsip_scalar_f gain);

#ifdef __cplusplus
}
#endif
#endif /* _vsip_amd_h */
