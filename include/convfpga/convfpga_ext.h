/*
 * convfpga_ext.h -- extensions of the drop-in API (not needed by callers of
 * convfpga.h).  Plain C ABI: opaque handles, pointers and sizes; device
 * pointers are HIP device addresses, `stream` is a hipStream_t passed as
 * void* (NULL = HIP's null stream, so the call is ordered with other work
 * on it, e.g. torch's default stream).  All calls require a successful
 * fpga_initialize() and return 0 or a negative CF_ERR_* code; the message of
 * the last error on the calling thread is in cf_last_error().
 *
 * Data layout is the drop-in one: N^3 interleaved complex float cubes, index
 * z*N*N + y*N + x, consecutive cubes of a batch N^3 elements apart.
 */
#ifndef CONVFPGA_EXT_H
#define CONVFPGA_EXT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CF_OK 0
#define CF_ERR_ARG (-1)      /* NULL pointer, zero count, bad enum          */
#define CF_ERR_SIZE (-2)     /* N not in {8,16,32,64,128,256,512}           */
#define CF_ERR_STATE (-3)    /* fpga_initialize() has not succeeded         */
#define CF_ERR_RUNTIME (-4)  /* HIP runtime error                           */
#define CF_ERR_NOMEM (-5)    /* device or host allocation failed            */

/* scale applied in the last inverse pass */
#define CF_SCALE_NONE 0     /* reference semantics: N^3 * circular convolution */
#define CF_SCALE_INV_N3 1   /* 1/N^3: the circular convolution itself          */

/* arithmetic of the radix-16 DFT stages */
#define CF_PREC_FP32 0        /* FP32 FMA butterflies                        */
#define CF_PREC_SPLIT_FP16 1  /* 16-point DFTs as complex GEMMs on the matrix cores:
                               * hi+lo FP16 operands, three products, FP32
                               * accumulation, one power-of-two scale per tile.
                               * N = 256 with a spectrum of a four-step plan
                               * whose middle tiles are 2 x wide (split3_g22,
                               * g42, g24m2, g28m2 and their o / h forms), and
                               * N = 512 with a spectrum of CF_PLAN_AXIS5 or
                               * CF_PLAN_AXIS5Z (512 = 16 * 16 * 2: two matrix
                               * stages and an FP32 radix-2 stage per axis);
                               * other spectra: CF_ERR_ARG */

/* execution plans (which passes and tiles a convolution of edge N runs) */
#define CF_PLAN_AUTO 0       /* the autotuner's choice (default)                              */
#define CF_PLAN_AXIS5 1      /* x, y, z*H*z, y^-1, x^-1: 11 volume transfers, every N          */
#define CF_PLAN_PLANE3 2     /* xy-plane, z*H*z, xy-plane^-1: 7 transfers, N in {32,64,128}    */
#define CF_PLAN_SPLIT3_G22 3 /* four-step y split, 7 transfers, N = 128, 256, tile groups 2x2  */
#define CF_PLAN_SPLIT3_G24 4 /*   ... tile groups 2x4                                          */
#define CF_PLAN_SPLIT3_G42 5 /*   ... tile groups 4x2                                          */
#define CF_PLAN_SPLIT3_G44 6 /*   ... tile groups 4x4                                          */
/* the same with middle-pass tiles narrower than the intermediate's x group:
 * G<GA><GB>M<GM> = outer tiles of GA y2 values, runs of GA*GB points, middle tiles of GM x */
#define CF_PLAN_SPLIT3_G24M2 7
#define CF_PLAN_SPLIT3_G24M1 8
#define CF_PLAN_SPLIT3_G28M4 9
#define CF_PLAN_SPLIT3_G28M2 10
#define CF_PLAN_SPLIT3_G28M1 11
#define CF_PLAN_AXIS5Z 12 /* the five passes with a z-blocked intermediate in the workspace: the
                             z pass reads and writes contiguous blocks; N in {64,...,512}   */
/* four-step plans whose middle pass writes a second intermediate, laid out so
 * that every middle tile (2 x wide) stores one contiguous block and the last
 * pass reads runs of 4 points: the first pass writes runs of 2*4 (G24M2O) or
 * 2*8 (G28M2O) points.  The second intermediate needs one more cube of
 * workspace when the output is the signal (in place); otherwise the output
 * buffer holds the first one.  Spectrum layout of the other GM = 2 plans. */
#define CF_PLAN_SPLIT3_G24M2O 13
#define CF_PLAN_SPLIT3_G28M2O 14
/* the same, with the second intermediate in the first one's x groups and
 * pairs of y2 interleaved: the middle pass writes runs of 4 points, the last
 * pass reads runs of 8 (G24M2H) or 16 (G28M2H) */
#define CF_PLAN_SPLIT3_G24M2H 15
#define CF_PLAN_SPLIT3_G28M2H 16
/* the first pass writes whole-line runs (x groups of 8, as G28M2H), the middle
 * pass writes x groups of 4 with y2 pairs interleaved (runs of 4 points, two
 * tiles per 64-byte run), the last pass reads runs of 8 (as G24M2H) */
#define CF_PLAN_SPLIT3_G28M2H4 17
/* G28M2O with last-pass tiles of 4 y2 values (1024 threads), which read runs
 * of 4*2 points of the middle tiles' blocks, and those blocks ordered with k1
 * inside the x group */
#define CF_PLAN_SPLIT3_G28M2OC4 18
/* G28M2OC4 with first-pass tiles of one y2 value (16 rows; 256 threads at
 * N = 256) writing the same layout, two tiles per 128-byte line */
#define CF_PLAN_SPLIT3_G18M2OC4 19
#define CF_NUM_PLANS 20

const char *cf_last_error(void);
/* 1 if N is a supported cube edge, else 0 */
int cf_supported_size(unsigned N);

/* 1 if `plan` can run edge N */
int cf_plan_supported(unsigned N, int plan);
/* Force the plan for edge N (CF_PLAN_AUTO returns the choice to the autotuner).
 * Spectra already created keep the plan they were created for. */
int cf_set_plan(unsigned N, int plan);
/* The plan new spectra of edge N get; runs the autotuner on first use. */
int cf_get_plan(unsigned N);
/* Time every supported plan for edge N on scratch data (one cube, FP32) and
 * make the fastest the automatic choice.  ms[CF_NUM_PLANS] (nullable) receives
 * the mean device time per convolution of each plan, 0 for unsupported ones.
 * The choice is also appended to the file named by $CONVFPGA_TUNE_FILE, if
 * set, and read back from it by later processes. Returns the chosen plan. */
int cf_autotune(unsigned N, float *ms);

/* ---- precomputed filter spectra ("filter setup") ---------------------------
 * The reference recomputes the filter transform in every call
 * (host/src/convfpga.c:87-172); a spectrum handle computes it once.
 * `count` filters of edge N, contiguous, on the host or the device. */
typedef struct cf_spectrum_s *cf_spectrum_t;
int cf_spectrum_create(unsigned N, unsigned count, const void *filters, int on_device, void *stream,
                       cf_spectrum_t *out);
void cf_spectrum_destroy(cf_spectrum_t h);
/* device address of the spectra: count * N^3 complex floats, natural order for
 * CF_PLAN_AXIS5 / CF_PLAN_PLANE3, an internal blocked order for CF_PLAN_SPLIT3_*
 * and CF_PLAN_AXIS5Z */
const void *cf_spectrum_data(cf_spectrum_t h);
unsigned cf_spectrum_edge(cf_spectrum_t h);
unsigned cf_spectrum_count(cf_spectrum_t h);
/* the plan the spectrum was laid out for (its layout is internal to the plan) */
int cf_spectrum_plan(cf_spectrum_t h);
/* A handle of count 1 for filter `index` of `parent`, sharing its memory (no
 * copy); valid while the parent lives, destroyed with cf_spectrum_destroy.
 * Instances of a graph batch whose spectra are consecutive views of one
 * parent (and whose signals and outputs are consecutive) run as one batched
 * pass sequence. */
int cf_spectrum_view(cf_spectrum_t parent, unsigned index, cf_spectrum_t *out);

/* ---- device-resident convolution --------------------------------------------
 * how_many cubes d_sig -> d_out (may be the same buffer).  With a spectrum of
 * count 1 every cube uses it; with count == how_many cube b uses spectrum b. */
int cf_conv3d_device(cf_spectrum_t h, const void *d_sig, void *d_out, unsigned how_many, int scale_mode,
                     int precision, void *stream);

/* ---- host-resident batches --------------------------------------------------
 * how_many cubes h_sig -> h_out in ordinary (pageable) host memory; the two
 * may be the same buffer.  The cubes stream through the device in groups of
 * whole cubes so that the upload of one group, the passes of the group before
 * it and the download of the one before that run at the same time (three
 * streams of the library, pinned chunk buffers, three device slots): the
 * overlapped counterpart of the reference's pairwise batch pipeline
 * (host/src/convfpga_svm_batch.c:241-662), with any spectrum of
 * cf_conv3d_device (shared, or one filter per cube), scale mode and
 * precision.  Blocking.  times (nullable): summed device time of the groups'
 * passes and the wall time of the whole call, in milliseconds. */
typedef struct {
  double pass_ms, wall_ms;
} cf_host_times_t;
int cf_conv3d_host(cf_spectrum_t h, const void *h_sig, void *h_out, unsigned how_many, int scale_mode,
                   int precision, cf_host_times_t *times);

/* Per-pass device times (milliseconds) of one cf_conv3d_device call, measured
 * with HIP events on the launch stream; blocks until the call has finished. */
#define CF_MAX_PASSES 8
typedef struct {
  unsigned npass;
  float ms[CF_MAX_PASSES];
  char name[CF_MAX_PASSES][24];
  double bytes[CF_MAX_PASSES]; /* algorithmic HBM bytes of each pass */
} cf_pass_times_t;
int cf_conv3d_device_timed(cf_spectrum_t h, const void *d_sig, void *d_out, unsigned how_many, int scale_mode,
                           int precision, void *stream, cf_pass_times_t *times);

/* One GPU's share of a slab-decomposed convolution (design for several GPUs:
 * each of `shares` GPUs holds N/shares z-planes for the x pass and the outer
 * passes of the four-step plans, and after an all-to-all a 1/shares range of
 * pencils for the passes along z).  Runs every pass of the spectrum's plan on
 * one cube over its first 1/shares of tiles -- that share, at its real shape
 * -- without the exchanges between the passes, and returns each pass's time
 * and algorithmic bytes.  shares = 1, 2, 4, 8 or 16; with shares > 1 d_out and
 * the workspace receive partial data, not a convolution; with shares = 1 it is
 * cf_conv3d_device_timed of one cube with CF_SCALE_INV_N3.  Blocks. */
int cf_conv3d_share_timed(cf_spectrum_t h, const void *d_sig, void *d_out, unsigned shares, int precision,
                          void *stream, cf_pass_times_t *times);

/* Event profiling of the next cf_conv3d_device calls without a host sync per
 * call: cf_profile_begin arms it for up to max_calls recorded calls, taking
 * the last call of every group of `every` (1: each one; the others run without
 * events, whose packets between the passes cost the GPU a few microseconds
 * each);
 * cf_profile_end waits for them and returns each pass's mean device time over
 * the recorded calls (times->ms, from events recorded on the launch stream
 * between the passes) and the number of calls recorded. */
int cf_profile_begin(unsigned max_calls, unsigned every);
int cf_profile_end(cf_pass_times_t *avg, unsigned *ncalls);

/* ---- the same without any allocation ---------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same arguments and result as cf_conv3d_device, but the
 * intermediate cube of the four-step and z-blocked plans lives in a fixed
 * device scratch array of the code object (allocated when fpga_initialize()
 * loads it, held until fpga_final()), so the call allocates and frees
 * nothing: it only launches the passes on `stream`.  Limit: an intermediate
 * of at most CF_NOALLOC_SCRATCH_BYTES (N^3 * 8 * how_many bytes: one 256^3
 * cube, 8 at 128^3, 64 at 64^3; the plans with two intermediates, the o and
 * h four-step forms, need twice that when d_sig == d_out) -- larger calls
 * fail with CF_ERR_SIZE; plans without an intermediate (CF_PLAN_AXIS5,
 * CF_PLAN_PLANE3) have no limit.
 * Every call shares the one scratch array: calls that overlap in time on the
 * GPU (on different streams, or from different threads) race.  The spectrum
 * must have been created in the current session (its edge's kernels and
 * twiddles are set up then). */
#define CF_NOALLOC_SCRATCH_BYTES 134217728ull
int cf_conv3d_device_noalloc(cf_spectrum_t h, const void *d_sig, void *d_out, unsigned how_many, int scale_mode,
                             int precision, void *stream);

/* ---- batches of mixed sizes ------------------------------------------------
 * A fixed list of convolutions (each: device signal, device output, spectrum
 * of count 1 -- instances may share a spectrum) prepared once and replayed
 * with one graph launch per call.
 *   CF_BATCH_PERSISTENT: one persistent kernel runs every pass tile of every
 *     instance, stage by stage (all first passes, then all second passes,
 *     ...), each instance on the plan of its spectrum: CF_PLAN_AXIS5
 *     (N = 16..256), CF_PLAN_PLANE3 (N = 32, 64), CF_PLAN_SPLIT3_G24M2 or
 *     CF_PLAN_SPLIT3_G24M2H (N = 128, 256); other plans: CF_ERR_ARG.  The
 *     batch holds the intermediates of its four-step instances.  Its
 *     work-groups wait for each other, so they must all be resident at once:
 *     with other work filling the GPU at the same time, the kernel's bounded
 *     waits expire and cf_batch_check reports the error.
 *   CF_BATCH_GRAPH: the passes of each instance's own plan, recorded in order
 *     into one graph (any edge and plan).
 * An output may alias its own signal.  In CF_BATCH_PERSISTENT mode the
 * instances run concurrently, so each output must be disjoint from every
 * other instance's signal and output (CF_ERR_ARG otherwise); graph mode keeps
 * the given order, so chained instances (one reads another's output) and
 * shared outputs (the last instance wins) are allowed there.  Concurrent
 * launches of one batch race. */
typedef struct {
  const void *d_sig;
  void *d_out;
  cf_spectrum_t h;
} cf_instance_t;
typedef struct cf_batch_s *cf_batch_t;
#define CF_BATCH_PERSISTENT 0
#define CF_BATCH_GRAPH 1
int cf_batch_create(const cf_instance_t *inst, unsigned count, int scale_mode, int mode, cf_batch_t *out);
int cf_batch_launch(cf_batch_t b, void *stream);
/* waits for the last launch; CF_ERR_RUNTIME if the persistent kernel reported a stalled dependency */
int cf_batch_check(cf_batch_t b);
void cf_batch_destroy(cf_batch_t b);
/* Test support for the persistent kernel's bounded wait: a CF_BATCH_PERSISTENT
 * batch whose last stage waits for one tile more than the stage before it
 * completes, polling at most spin_limit times; every launch therefore ends
 * with the error cf_batch_check reports, and with the batch's last pass not
 * run (its outputs are incomplete).  For tests of the error path only. */
int cf_batch_create_stall_test(const cf_instance_t *inst, unsigned count, int scale_mode,
                               unsigned long long spin_limit, cf_batch_t *out);

/* ---- plain 3D transforms (device) ------------------------------------------ */
/* forward (inverse = 0) or unnormalised inverse (inverse = 1) 3D FFT */
int cf_fft3d_device(unsigned N, const void *d_in, void *d_out, unsigned how_many, int inverse, void *stream);

/* ---- FP64 error norm on the device -----------------------------------------
 * rel_l2 = sqrt(sum|test-ref|^2 / sum|ref|^2) accumulated in double; ref is
 * complex double (ref_is_f64 = 1) or complex float. Blocks until done. */
int cf_rel_l2_device(const void *d_test, const void *d_ref, int ref_is_f64, size_t npts, void *stream,
                     double *rel_l2);

/* ---- synthetic inputs --------------------------------------------------------
 * nfloats floats uniform on [-1, 1) from the portable splitmix64 generator
 * (value i = k/2^23 - 1, k = top 24 bits of splitmix64(seed + (i+1)*0x9E3779B97F4A7C15)),
 * written on the device; asynchronous on `stream`. */
int cf_fill_uniform(void *d_out, size_t nfloats, unsigned long long seed, void *stream);
/* npts complex points with log-uniform magnitudes 10^(6 u1 - 3) in [1e-3, 1e3)
 * and uniform phase 2 pi u2 (u1, u2: the 24-bit splitmix64 variates 2p and
 * 2p+1 of the same generator), computed in double on the device. */
int cf_fill_loguniform(void *d_out, size_t npts, unsigned long long seed, void *stream);

#ifdef __cplusplus
}
#endif
#endif
