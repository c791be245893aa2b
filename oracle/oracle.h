/*
 * oracle.h -- CPU checker for the FFT-based 3D convolution path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for the baseline, as the CPU path being timed beside the GPU).  The
 * product library (libconvfpga) never links or calls it.
 *
 * What it restates (reference file:line):
 *   - orc_conv3d_f32      : fft_conv3D_cpu_verify, src/fft_conv3D.cpp:186-233
 *                           (rank-3 c2c forward of every signal and of the one
 *                           filter, fp32 pointwise product in the order of
 *                           fft_conv3D.cpp:225-226 / kernels/fft/fft_8.cl:180-185,
 *                           unnormalised backward transform).  FFTW3f is not
 *                           present in this image, so the transforms are this
 *                           file's own fp32 radix-2 FFT.
 *   - orc_snr_db          : the SNR test of src/fft_conv3D.cpp:235-258
 *                           (double accumulators, 10*log10(sum|ref|^2/sum|ref-out|^2)).
 *   - orc_fill_ref01      : create_data, src/helper.cpp:148-172 (glibc rand()
 *                           with its default seed, rand()/RAND_MAX in fp32, batch
 *                           instances copied from instance 0).
 *   - orc_fft3d_f64 / orc_conv3d_f64 : the FP64 oracle; same transform
 *                           convention as the reference (forward sign -, backward
 *                           unnormalised: kernels/fft/fft_8.cl:242 and 280-282,
 *                           346-348; src/fft_conv3D.cpp:201-203, 233).
 *   - orc_conv3d_brute_f64: O(P^2) direct circular convolution (definition).
 *
 * Pinning (tests/test_oracle.py), every function by a check that a plausible
 * mistake fails:
 *   orc_fft3d_f64         numpy pocketfft golden fixtures and live numpy,
 *                         single-tone KAT (sign -, amplitude P), Parseval,
 *                         linearity, round trip
 *   orc_conv3d_f64        O(P^2) brute force (N <= 16), delta / shifted-delta
 *                         closed forms, golden fixtures
 *   orc_conv3d_brute_f64  numpy index-arithmetic restatement of the definition
 *   orc_fft3d_f32         FP64 transform and numpy (1e-9 < rel-L2 < 1e-6, both
 *                         signs), single-tone KAT in both directions, round
 *                         trip, thread-count independence
 *   orc_conv3d_f32        FP64 oracle (FP32-rounding-sized error), batch
 *                         instances equal single calls
 *   orc_conv3d_f32_spec   bit-equal to orc_conv3d_f32 with H = orc_fft3d_f32(filter)
 *   orc_snr_db, orc_rel_l2_*  numpy norms and closed forms
 *   orc_fill_*            numpy port (bit-exact / one ulp), glibc rand()
 *   orc_fftw_available, orc_max_threads  direct dlopen probe, OpenMP default
 * FFTW's own fp32 output is not available (FFTW3 is absent): parity at the
 * FFTW boundary is pinned only statistically.
 *
 * Data layout everywhere: interleaved complex {re, im}, linear index
 * z*N*N + y*N + x with x fastest (host/include/convfpga/convfpga.h:16-19).
 */
#ifndef CONVFPGA_ORACLE_H
#define CONVFPGA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- portable seeded inputs (identical in C, numpy and on the GPU) ---- */
/* splitmix64 of (seed, index); 24 high bits -> k/2^23 - 1 in [-1, 1). */
void orc_fill_uniform(float *out, size_t nfloats, uint64_t seed);
/* magnitude 10^U(-3,3), phase U[0,2pi); computed in double, rounded to float. */
void orc_fill_loguniform(float *out, size_t npts, uint64_t seed);
/* the reference's create_data (helper.cpp:148-172): srand(1) then rand()/RAND_MAX. */
void orc_fill_ref01(float *out, size_t npts, unsigned batch);

/* ---- FP64 transforms ---- */
/* in-place 3D DFT of an N^3 interleaved complex double cube; sign -1 forward,
 * +1 backward (unnormalised). N must be a power of two, 1 <= N <= 1024. */
int orc_fft3d_f64(int n, double *data, int sign);
/* out = scale * IFFT3(FFT3(sig) .* FFT3(filt)) in double; out has 2*N^3 doubles. */
int orc_conv3d_f64(int n, const float *sig, const float *filt, double *out, double scale);
/* the same by the O(P^2) definition: out[p] = scale*P * sum_q sig[q] filt[(p-q) mod N]. */
int orc_conv3d_brute_f64(int n, const float *sig, const float *filt, double *out, double scale);

/* ---- FP32 restatement of the reference verifier's transform ---- */
/* out[b] = IFFT3_unnorm(FFT3(sig[b]) .* FFT3(filt)) in fp32; batch cubes share
 * one filter.  nthreads <= 0 means all OpenMP threads. */
int orc_conv3d_f32(int n, int batch, const float *sig, const float *filt, float *out, int nthreads);
int orc_fft3d_f32(int n, float *data, int sign, int nthreads);
/* CPU baseline step of fft_conv3D_cpu (src/fft_conv3D.cpp:142-165): with the
 * filter spectrum H precomputed, copy sig into out, forward FFT, fp32 product
 * with H, unnormalised backward FFT -- all in place in out. */
int orc_conv3d_f32_spec(int n, const float *sig, const float *H, float *out, int nthreads);

/* ---- metrics (double accumulators) ---- */
double orc_snr_db(const float *ref, const float *test, size_t npts);
double orc_rel_l2_f32_f64(const float *test, const double *ref, size_t npts);
double orc_rel_l2_f32_f32(const float *test, const float *ref, size_t npts);

/* ---- FFTW3f probe (dlopen at run time; absent in this image) ---- */
/* 1 when libfftw3f.so.3 can be opened and exports the plan/execute/destroy
 * entry points: the CPU-baseline label says whether FFTW was present. */
int orc_fftw_available(void);

int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
