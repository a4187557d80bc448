/*
 * wavemoth.h -- public C ABI of libwavemoth_hip.so, the MI355X (gfx950)
 * a_lm -> HEALPix-map synthesis engine.
 *
 * The first block keeps the reference's stable driver API unchanged, so a
 * caller of the reference links against this library instead:
 *   reference: src/wavemoth.h:20-46 (declarations), src/wavemoth.c (impl).
 * Layouts are the reference's:
 *   input   complex128 a_lm, m-major triangle, idx(l,m) = m(2 lmax - m + 3)/2
 *           + (l - m), each entry holding nmaps interleaved complex values
 *           (src/wavemoth.c:1016-1018, wavemoth/lib.pyx:119-135);
 *   output  float64 map (npix, nmaps) C order, HEALPix RING order from the
 *           north pole (src/wavemoth.c:1182-1187, 1199-1232).
 * Ownership follows the reference: the plan stores the caller's host
 * pointers (src/wavemoth.c:396-397); they must outlive the plan and are read
 * / written at execute time.
 *
 * Differences, all deliberate and documented:
 *   - errors never abort(): plan creation returns NULL and every other call
 *     returns/records an error; wavemoth_last_error() says why;
 *   - nthreads must be > 0 (as src/wavemoth.c:411-414) but is otherwise
 *     unused: the work runs on the current HIP device;
 *   - Nside is limited to 8192 (the reference: 2^15, src/wavemoth.c:92).
 *     Ring FFTs of up to 8192 points stay in one workgroup's LDS (Nside <=
 *     4096); longer ones run as four-step FFTs through a device scratch
 *     buffer.  Batch plans are limited to Nside <= 4096, the adjoint to
 *     Nside <= 2048;
 *   - no resource file is needed: Lambda_l^m is generated on the GPU in FP64.
 *     A resource file passed to wavemoth_plan_to_healpix is validated against
 *     (Nside, lmax, mmax) through its header and otherwise unused.
 */
#ifndef WAVEMOTH_H_
#define WAVEMOTH_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _wavemoth_plan *wavemoth_plan;

/* src/wavemoth.h:22-27 */
#define WAVEMOTH_MMAJOR 0x0
#define WAVEMOTH_ESTIMATE 0x0
#define WAVEMOTH_MEASURE 0x1
#define WAVEMOTH_NO_RESOURCE_COPY 0x10
/* Extension flag: generate the per-(m, ring) Lambda seeds inside every
   execute instead of once at plan time. */
#define WAVEMOTH_SEEDS_ON_THE_FLY 0x100
/* Extension: run the Legendre stage from the plan's resource file (given,
   or found under the wavemoth_configure directory) instead of generating
   Lambda on the device: the strips' stored start rows, recurrence
   coefficients and dense blocks, summed as src/wavemoth.c:958-1009 and
   src/legendre_transform.c.in:680-716 do, and for butterfly-compressed
   matrices the transposed interpolations of their trees
   (src/butterfly.c.in:284-432), on the GPU (FP64).  Excludes
   WAVEMOTH_MIXED_*. */
#define WAVEMOTH_RESOURCE_APPLY 0x400
/* Extension flags (at most one): the Legendre contraction on the matrix
   cores, the Lambda recurrence and its seeds still in FP64.  Sums are
   accumulated in FP32 within l-chunks of 128 and reduced across chunks in
   FP64.  Meant for nmaps >= 16 (any nmaps works).
     WAVEMOTH_MIXED_F32     FP32 inputs, FP32 MFMA;
     WAVEMOTH_MIXED_BF16X3  Lambda and a split into two bf16 parts each,
                            three products (hi*hi + hi*lo + lo*hi);
     WAVEMOTH_MIXED_BF16X6  three bf16 parts each, six products;
     WAVEMOTH_MIXED_F16X3   two fp16 parts each, three products, with
                            power-of-two scales keeping fp16's range.
   These plans keep the Legendre-stage output in FP32 and run the ring FFTs
   in FP32 (the map is FP64); wavemoth_get_work widens it to complex128. */
#define WAVEMOTH_MIXED_F32 0x1000
#define WAVEMOTH_MIXED_BF16X3 0x2000
#define WAVEMOTH_MIXED_BF16X6 0x4000
#define WAVEMOTH_MIXED_F16X3 0x8000

/* ---- stable driver API (replaces src/wavemoth.h:33-46) ---------------- */

/* replaces src/wavemoth.h:33 (impl src/wavemoth.c:129-135) */
void wavemoth_configure(char *resource_dir);

/* replaces src/wavemoth.h:35-39 (impl src/wavemoth.c:383-615) */
wavemoth_plan wavemoth_plan_to_healpix(int Nside, int lmax, int mmax, int nmaps,
                                       int nthreads, double *input, double *output,
                                       int ordering, unsigned flags,
                                       char *resource_filename);

/* replaces src/wavemoth.h:41 (impl src/wavemoth.c:837-874); frees everything */
void wavemoth_destroy_plan(wavemoth_plan plan);

/* replaces src/wavemoth.h:42 (impl src/wavemoth.c:1271-1283): synchronous
   host input -> device -> host output */
void wavemoth_execute(wavemoth_plan plan);

/* replaces src/wavemoth.h:44 (impl src/wavemoth.c:876-885): multiply-adds
   counted as 2 flops for nvecs = 2 over the (l, ring) elements the plan
   actually evaluates for (m, odd) (4 * element_count, as the reference) */
int64_t wavemoth_get_legendre_flops(wavemoth_plan plan, int m, int odd);

/* replaces src/wavemoth.h:46 (impl src/wavemoth.c:144-155) */
int wavemoth_query_resourcefile(char *filename, int *out_Nside, int *out_lmax);

/* Reads and validates a rev1 resource file the way plan creation does
   (the reader of src/wavemoth.c:203-264 and the walk of
   src/butterfly.c.in:284-432 / src/wavemoth.c:958-1009, every offset and
   size checked).  Extension: the reference has no validator.  Fills up to
   ninfo of: 0 lmax, 1 mmax, 2 Nside, 3 matrices present, 4 matrices
   missing (start 0), 5 compressed trees (heap of more than one node),
   6 recurrence strips, 7 dense blocks and strips, 8 stored matrix elements
   (strip rows x columns, interpolation blocks), 9 deepest tree level,
   10 largest |x^2 - cos^2 theta| of the strips' x^2 against the HEALPix
   rings, in units of 1e-18 (-1 without strips).  Returns 0, or -1 with
   wavemoth_last_error set for a missing, truncated or malformed file.
   wavemoth_plan_to_healpix reads resources the same way: the file given,
   or <dir>/rev1/<Nside>.dat when wavemoth_configure set a directory and
   the file exists (a missing one is not an error here: the device path
   generates Lambda itself), with Nside < 0 taken from the file. */
int wavemoth_resource_info(const char *filename, int64_t *info, int ninfo);

/* ---- stage-wise hooks (replace src/wavemoth_private.h:132-138, used by
        wavemoth/lib.pyx:55-57,80-81,151-152,187-193) --------------------- */

/* replaces wavemoth_perform_legendre_transforms (src/wavemoth.c:933-941):
   copies the host input to the device and runs the Legendre stage only */
void wavemoth_perform_legendre_transforms(wavemoth_plan plan);
/* replaces wavemoth_perform_backward_ffts (src/wavemoth.c:1193-1195): runs
   the ring stage on the current work buffer and copies the map to host */
void wavemoth_perform_backward_ffts(wavemoth_plan plan);
/* replaces wavemoth_disable_phase_shifting (src/wavemoth.c:1240-1248);
   waits for all work queued on the current device before it changes the
   plan's ring table (executes issued later use phi0 = 0) */
void wavemoth_disable_phase_shifting(wavemoth_plan plan);

/* ---- per-strip Legendre transform (replaces src/legendre_transform.h:14-25,
        the calls wavemoth/lib.pyx:220-292 wraps and
        wavemoth/test/test_legendre_transforms.py:22-72 drives) ------------ */

/* replaces wavemoth_legendre_transform_auxdata (src/legendre_transform.h:14-15,
   impl src/legendre_transform.c.in:723-747): for rows l = lmin + 2k,
   auxdata[3(k-2) + {0,1,2}] = {alpha, beta, gamma} of
   P_k = (x^2 + alpha) beta P_{k-1} + gamma P_{k-2}; 3 (nk - 2) entries.
   Host arithmetic (plan-time data, as in the reference). */
void wavemoth_legendre_transform_auxdata(size_t m, size_t lmin, size_t nk, double *auxdata);
/* replaces wavemoth_legendre_transform (src/legendre_transform.h:17-23, impl
   src/legendre_transform.c.in:680-716; its SSE twin :640-678 computes the
   same sums): y[i][j] = sum_{k<nk} P_k(x_i) a[2 k nvecs + j] for nx rings,
   P_0 = P[i], P_1 = Pp1[i].  Host arrays; the sums run on the current HIP
   device, synchronously.  nk >= 2 and even nvecs as the reference asserts;
   errors go to wavemoth_last_error() instead of assert(). */
void wavemoth_legendre_transform(size_t nx, size_t nk, size_t nvecs, double *a, double *y,
                                 double *x_squared, double *auxdata, double *P, double *Pp1);
/* Extension: the same on device arrays, asynchronous on `stream` (a
   hipStream_t, NULL = default).  Returns 0 on success. */
int wavemoth_legendre_transform_device(size_t nx, size_t nk, size_t nvecs, const double *a,
                                       double *y, const double *x_squared,
                                       const double *auxdata, const double *P,
                                       const double *Pp1, void *stream);

/* ---- extensions (no counterpart in the reference) ---------------------- */

/* Last error of the calling thread ("" if none). */
const char *wavemoth_last_error(void);

/* Copy the Legendre-stage output to host in the reference's get_work layout
   (wavemoth/lib.pyx:158-179): complex128 [mmax+1][2 (odd)][2 Nside][nmaps],
   rings ordered from the equator to the pole.  Returns 0 on success. */
int wavemoth_get_work(wavemoth_plan plan, double *out);

/* Asynchronous execute on device buffers: d_input (same layout as input)
   and d_output (same layout as output) are device pointers, `stream` a
   hipStream_t (NULL = default stream).  Returns 0 on success. */
int wavemoth_execute_device(wavemoth_plan plan, const double *d_input, double *d_output,
                            void *stream);

/* Adjoint of the synthesis, b = Y^T g (the unweighted analysis):
   b_lm = c_m sum_p g_p Lambda_l^m(theta_p) e^{-i m phi_p}, c_0 = 1, c_m = 2.
   It is the gradient of sum_p g_p f_p with respect to (Re a, Im a) and the
   backward of the synthesis.  The direction of the reference's only GPU
   kernel (wavemoth/cuda/legendre_transform.cu.in:589-679, Lambda^T q).
   _device: d_map (npix, nmaps) float64 and d_alm (nalm, nmaps) complex128
   on the device, asynchronous on `stream`; returns 0 on success.
   wavemoth_execute_adjoint: synchronous, reads the plan's host *output*
   array and overwrites its host *input* array.  Plans of Nside > 2048 have
   no adjoint (an error). */
int wavemoth_execute_adjoint_device(wavemoth_plan plan, const double *d_map, double *d_alm,
                                    void *stream);
void wavemoth_execute_adjoint(wavemoth_plan plan);

/* The adjoint without any device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Same arguments, kernels and result as
   wavemoth_execute_adjoint_device, but the partial sums over ring groups
   live in a fixed 256 MiB device array of the library (held by each GPU it
   runs on until the process ends) instead of a buffer the plan allocates
   on first use.  The ring groups are the plan's (wavemoth_plan_info key 12)
   reduced by halving until G * nalm_padded * nmaps * 16 bytes fit (key 13);
   with the same count the result is bitwise that of the other entry,
   otherwise it differs only by the summation order.  Returns -1 with
   wavemoth_last_error set when one ring group does not fit (it fits while
   nmaps * lmax^2 <= ~3.3e7 for mmax = lmax: lmax 5780 at one map, 2040 at
   eight), when the plan's device
   is not current, or for NULL / misaligned pointers.  Every call on a GPU
   shares the array: calls that overlap on one GPU (other streams, other
   threads, other plans) race and must be serialised by the caller. */
int wavemoth_execute_adjoint_device_noalloc(wavemoth_plan plan, const double *d_map,
                                            double *d_alm, void *stream);

/* Stage timing: with profiling on, every execute records HIP events on its
   stream around (0) a_lm prescale (plus the seed kernels with
   WAVEMOTH_SEEDS_ON_THE_FLY), (1) the Legendre sweep, (2) the ring stage.
   wavemoth_stage_times waits for them and returns the average milliseconds
   of each stage over the executes recorded since the last call, then
   forgets them.  Both return 0 on success. */
int wavemoth_set_profiling(wavemoth_plan plan, int on);
int wavemoth_stage_times(wavemoth_plan plan, double *avg_ms, int *count);

/* Plan facts: 0 Nside, 1 lmax, 2 mmax, 3 nmaps, 4 npix, 5 dense triples T
   (sum over m of (lmax-m+1) * 2 Nside), 6 trimmed triples T_eps evaluated by
   the plan, 7 flags, 8 legendre-kernel launches per execute, 9 device bytes
   held by the plan, 10 (l, ring-pair) slots the Legendre sweep evaluates
   including its wave/chunk padding, 11 contraction scheme (-1 FP64,
   0 MIXED_F32, 1 MIXED_BF16X3, 2 MIXED_BF16X6, 3 MIXED_F16X3), 12 ring groups
   of the adjoint's partial sums (0: the plan has no adjoint), 13 ring
   groups of the allocation-free adjoint (0: it cannot run this plan),
   14..17 ring pairs transformed by the general LDS ring kernel, the
   register-FFT belt, polar-cap and wrapping-ring kernels, 18 the longest
   register-FFT length among them, 19 ring pairs of the four-step path
   (transforms longer than 8192 points).  Returns -1 for an unknown key
   (key 11 of an FP64 plan is also -1). */
int64_t wavemoth_plan_info(wavemoth_plan plan, int key);

/* ---- batched synthesis (extension) -------------------------------------
   Many independent syntheses of uneven (Nside, lmax, mmax, nmaps) in one
   execute.  The reference has no batch call; its nearest equivalent is one
   wavemoth_plan_to_healpix + wavemoth_execute per request, and every
   request's map equals (bit for bit) what a single plan of the same
   request produces here.  The prescale, Legendre sweep (one persistent
   work-queue kernel per sweep shape) and ring stage (one workgroup per item,
   per ring class) each run over all requests' work items, costliest request
   first; requests of equal (Nside, lmax, mmax) share
   their seed and ring tables.  The launches are captured once into a
   HIP graph (unless WAVEMOTH_BATCH_NO_GRAPH) and replayed by every execute.
   With WAVEMOTH_MEASURE the sweep shape of each distinct (Nside, lmax,
   mmax, nmaps) -- rings per lane and maps per pass -- and the kernel of its
   wrapping rings (general LDS kernel or register-FFT kernel) are chosen by
   timing the candidates on the device at plan time (results are kept for
   the process); otherwise a fixed heuristic chooses, the single plan's.  A
   measured ring kernel other than the single plan's changes the map in the
   last bits only.  Executes of one batch must not overlap. */
typedef struct _wavemoth_batch *wavemoth_batch;
#define WAVEMOTH_BATCH_NO_GRAPH 0x200

/* Returns NULL on error (wavemoth_last_error). */
wavemoth_batch wavemoth_plan_batch(int nreq, const int *Nside, const int *lmax,
                                   const int *mmax, const int *nmaps, unsigned flags);
void wavemoth_destroy_batch(wavemoth_batch batch);
/* d_inputs[r] / d_outputs[r]: device buffers of request r in the layouts
   of wavemoth_execute_device, 16-byte aligned; asynchronous on `stream`
   (a hipStream_t, NULL = default).  Returns 0 on success. */
int wavemoth_execute_batch_device(wavemoth_batch batch, const double *const *d_inputs,
                                  double *const *d_outputs, void *stream);
/* Synchronous host-buffer variant (stages through batch-owned device
   buffers allocated on first use). */
int wavemoth_execute_batch(wavemoth_batch batch, const double *const *inputs,
                           double *const *outputs);
/* Stage timing as wavemoth_set_profiling / wavemoth_stage_times, for
   (0) prescale, (1) Legendre sweep, (2) ring stage.  While profiling is on
   executes run the launches directly instead of replaying the graph. */
int wavemoth_batch_set_profiling(wavemoth_batch batch, int on);
int wavemoth_batch_stage_times(wavemoth_batch batch, double *avg_ms, int *count);
/* req = -1, batch facts: 0 requests, 1 total maps, 2 sum of T_eps, 3 sum of
   dense T, 4 algorithmic flops per execute (sum of (5 + 4 nmaps) T_eps),
   5 device bytes held, 6 sweep work items, 7 ring work items, 8 graph in
   use, 9 sum of executed (l, ring-pair, map-pass) slots, 10 persistent
   sweep workgroups.  req >= 0, request facts: 0 Nside, 1 lmax, 2 mmax,
   3 nmaps, 4 rings per lane, 5 maps per pass, 6 T_eps, 7 dense T, 8 npix,
   9 executed slots, 10 shape chosen by measurement (1) or heuristic (0),
   11 wrapping rings in the register-FFT kernels (1) or the general one (0).
   Returns -1 for an unknown key or request. */
int64_t wavemoth_batch_info(wavemoth_batch batch, int req, int key);

#ifdef __cplusplus
}
#endif

#endif /* WAVEMOTH_H_ */
