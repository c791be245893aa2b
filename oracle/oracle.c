/*
 * oracle.c -- CPU checker for the FFT-based 3D convolution path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h for who may call it and what each
 * function restates, with reference file:line).  Never linked into libconvfpga.
 *
 * Transforms: textbook iterative radix-2 decimation-in-time FFT (bit-reversal
 * permutation, then log2(N) butterfly stages), applied along x, y and z in
 * turn.  Twiddles exp(-2*pi*i*k/len) are computed from the exact rational
 * angle k/len in double (libm cos/sin, < 1 ulp) and, for the fp32 path,
 * rounded once to float.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double ORC_PI = 3.14159265358979323846264338327950288;

/* ------------------------------------------------------------------ PRNG */
static inline uint64_t orc_mix(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1u) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void orc_fill_uniform(float *out, size_t nfloats, uint64_t seed) {
  for (size_t i = 0; i < nfloats; ++i) {
    int64_t k = (int64_t)(orc_mix(seed, i) >> 40); /* 24 bits */
    out[i] = (float)(k - (1 << 23)) / 8388608.0f;   /* exact: k/2^23 - 1 */
  }
}

void orc_fill_loguniform(float *out, size_t npts, uint64_t seed) {
  for (size_t p = 0; p < npts; ++p) {
    double u1 = (double)(orc_mix(seed, 2 * p) >> 40) / 16777216.0;
    double u2 = (double)(orc_mix(seed, 2 * p + 1) >> 40) / 16777216.0;
    double mag = pow(10.0, 6.0 * u1 - 3.0);
    double ph = 2.0 * ORC_PI * u2;
    out[2 * p] = (float)(mag * cos(ph));
    out[2 * p + 1] = (float)(mag * sin(ph));
  }
}

/* helper.cpp:148-172: rand()/RAND_MAX in fp32 for re then im, glibc default
 * seed (no srand in the reference, i.e. seed 1); instances 1.. copy instance 0. */
void orc_fill_ref01(float *out, size_t npts, unsigned batch) {
  srand(1);
  for (size_t j = 0; j < npts; ++j) {
    out[2 * j] = (float)((float)rand() / (float)RAND_MAX);
    out[2 * j + 1] = (float)((float)rand() / (float)RAND_MAX);
  }
  for (unsigned b = 1; b < batch; ++b)
    memcpy(out + 2 * npts * b, out, 2 * npts * sizeof(float));
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static int orc_is_pow2(int n) { return n >= 1 && (n & (n - 1)) == 0; }

/* ------------------------------------------------------- FP64 line FFT */
typedef struct { double re, im; } cd;

static void bitrev_perm_d(cd *a, int n) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { cd t = a[i]; a[i] = a[j]; a[j] = t; }
  }
}

/* tw[k] = exp(-2 pi i k / n), k < n/2 */
static void twiddles_d(cd *tw, int n) {
  for (int k = 0; k < n / 2; ++k) {
    double ang = 2.0 * ORC_PI * ((double)k / (double)n);
    tw[k].re = cos(ang);
    tw[k].im = -sin(ang);
  }
}

static void fft_line_d(cd *a, int n, const cd *tw, int sign) {
  bitrev_perm_d(a, n);
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1, step = n / len;
    for (int i = 0; i < n; i += len) {
      for (int j = 0; j < half; ++j) {
        cd w = tw[j * step];
        if (sign > 0) w.im = -w.im;
        cd u = a[i + j], v = a[i + j + half];
        cd t = {v.re * w.re - v.im * w.im, v.re * w.im + v.im * w.re};
        a[i + j].re = u.re + t.re; a[i + j].im = u.im + t.im;
        a[i + j + half].re = u.re - t.re; a[i + j + half].im = u.im - t.im;
      }
    }
  }
}

/* apply a line transform along axis ax (0=x,1=y,2=z) of an n^3 cube, gathering
 * B adjacent lines at a time so strided axes read whole cache lines. */
#define ORC_B 8
int orc_fft3d_f64(int n, double *data, int sign) {
  if (!orc_is_pow2(n) || n > 1024 || !data) return -1;
  if (n == 1) return 0;
  cd *a = (cd *)data;
  cd *tw = (cd *)malloc(sizeof(cd) * (size_t)(n / 2));
  twiddles_d(tw, n);
  const size_t nn = (size_t)n * n;
  for (int ax = 0; ax < 3; ++ax) {
    const size_t stride = ax == 0 ? 1 : (ax == 1 ? (size_t)n : nn);
    #pragma omp parallel
    {
      cd *buf = (cd *)malloc(sizeof(cd) * (size_t)n * ORC_B);
      #pragma omp for schedule(static)
      for (int blk = 0; blk < (int)((nn + ORC_B - 1) / ORC_B); ++blk) {
        for (int l = 0; l < ORC_B; ++l) {
          size_t line = (size_t)blk * ORC_B + l;
          if (line >= nn) break;
          size_t base;
          if (ax == 0) base = line * n;
          else if (ax == 1) base = (line % n) + (line / n) * nn;     /* (x, z) */
          else base = line;                                          /* (x, y) */
          for (int j = 0; j < n; ++j) buf[(size_t)l * n + j] = a[base + j * stride];
          fft_line_d(buf + (size_t)l * n, n, tw, sign);
          for (int j = 0; j < n; ++j) a[base + j * stride] = buf[(size_t)l * n + j];
        }
      }
      free(buf);
    }
  }
  free(tw);
  return 0;
}

int orc_conv3d_f64(int n, const float *sig, const float *filt, double *out, double scale) {
  if (!orc_is_pow2(n) || !sig || !filt || !out) return -1;
  const size_t p = (size_t)n * n * n;
  double *h = (double *)malloc(sizeof(double) * 2 * p);
  if (!h) return -2;
  for (size_t i = 0; i < 2 * p; ++i) { out[i] = sig[i]; h[i] = filt[i]; }
  orc_fft3d_f64(n, out, -1);
  orc_fft3d_f64(n, h, -1);
  for (size_t i = 0; i < p; ++i) {
    double ar = out[2 * i], ai = out[2 * i + 1], br = h[2 * i], bi = h[2 * i + 1];
    out[2 * i] = ar * br - ai * bi;
    out[2 * i + 1] = ar * bi + ai * br;
  }
  orc_fft3d_f64(n, out, +1);
  for (size_t i = 0; i < 2 * p; ++i) out[i] *= scale;
  free(h);
  return 0;
}

/* Unnormalised IFFT(FFT(s) .* FFT(f)) equals P times the circular convolution
 * sum_q s[q] f[(p - q) mod N] (per axis).  O(P^2): meant for N <= 16. */
int orc_conv3d_brute_f64(int n, const float *sig, const float *filt, double *out, double scale) {
  if (!orc_is_pow2(n) || n > 32 || !sig || !filt || !out) return -1;
  const int m = n - 1;
  const double pts = (double)n * n * n;
  #pragma omp parallel for collapse(3)
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        double re = 0.0, im = 0.0;
        for (int qz = 0; qz < n; ++qz)
          for (int qy = 0; qy < n; ++qy)
            for (int qx = 0; qx < n; ++qx) {
              size_t iq = ((size_t)qz * n + qy) * n + qx;
              size_t id = ((size_t)((z - qz) & m) * n + ((y - qy) & m)) * n + ((x - qx) & m);
              double sr = sig[2 * iq], si = sig[2 * iq + 1];
              double fr = filt[2 * id], fi = filt[2 * id + 1];
              re += sr * fr - si * fi;
              im += sr * fi + si * fr;
            }
        size_t ip = ((size_t)z * n + y) * n + x;
        out[2 * ip] = re * pts * scale;
        out[2 * ip + 1] = im * pts * scale;
      }
  return 0;
}

/* ------------------------------------------------------- FP32 line FFT */
typedef struct { float re, im; } cf;

static void bitrev_perm_f(cf *a, int n) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { cf t = a[i]; a[i] = a[j]; a[j] = t; }
  }
}

static void fft_line_f(cf *a, int n, const cf *tw, int sign) {
  bitrev_perm_f(a, n);
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1, step = n / len;
    for (int i = 0; i < n; i += len) {
      for (int j = 0; j < half; ++j) {
        cf w = tw[j * step];
        if (sign > 0) w.im = -w.im;
        cf u = a[i + j], v = a[i + j + half];
        cf t = {v.re * w.re - v.im * w.im, v.re * w.im + v.im * w.re};
        a[i + j].re = u.re + t.re; a[i + j].im = u.im + t.im;
        a[i + j + half].re = u.re - t.re; a[i + j + half].im = u.im - t.im;
      }
    }
  }
}

int orc_fft3d_f32(int n, float *data, int sign, int nthreads) {
  if (!orc_is_pow2(n) || n > 1024 || !data) return -1;
  if (n == 1) return 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  cf *a = (cf *)data;
  cf *tw = (cf *)malloc(sizeof(cf) * (size_t)(n / 2));
  for (int k = 0; k < n / 2; ++k) {
    double ang = 2.0 * ORC_PI * ((double)k / (double)n);
    tw[k].re = (float)cos(ang);
    tw[k].im = (float)(-sin(ang));
  }
  const size_t nn = (size_t)n * n;
  /* blocks of nb lines adjacent in x: nb divides n, so a block never wraps */
  const int nb = n < ORC_B ? n : ORC_B;
  const int nblk = (int)(nn / nb);
  for (int ax = 0; ax < 3; ++ax) {
    const size_t stride = ax == 0 ? 1 : (ax == 1 ? (size_t)n : nn);
    #pragma omp parallel num_threads(nthreads)
    {
      cf *buf = (cf *)malloc(sizeof(cf) * (size_t)n * ORC_B);
      #pragma omp for schedule(static)
      for (int blk = 0; blk < nblk; ++blk) {
        size_t l0 = (size_t)blk * nb;
        int nl = nb;
        if (ax == 0) {
          for (int l = 0; l < nl; ++l) fft_line_f(a + (l0 + l) * n, n, tw, sign);
        } else {
          /* the nl lines are adjacent in x (nl divides n or equals n) */
          size_t base0 = ax == 1 ? (l0 % n) + (l0 / n) * nn : l0;
          for (int j = 0; j < n; ++j)
            for (int l = 0; l < nl; ++l) buf[(size_t)l * n + j] = a[base0 + l + j * stride];
          for (int l = 0; l < nl; ++l) fft_line_f(buf + (size_t)l * n, n, tw, sign);
          for (int j = 0; j < n; ++j)
            for (int l = 0; l < nl; ++l) a[base0 + l + j * stride] = buf[(size_t)l * n + j];
        }
      }
      free(buf);
    }
  }
  free(tw);
  return 0;
}

/* fft_conv3D.cpp:186-233 restated: forward FFT of every signal and of the
 * filter, fp32 product (225-226), unnormalised backward FFT. */
int orc_conv3d_f32(int n, int batch, const float *sig, const float *filt, float *out, int nthreads) {
  if (!orc_is_pow2(n) || batch < 1 || !sig || !filt || !out) return -1;
  const size_t p = (size_t)n * n * n;
  float *h = (float *)malloc(sizeof(float) * 2 * p);
  if (!h) return -2;
  memcpy(h, filt, sizeof(float) * 2 * p);
  memcpy(out, sig, sizeof(float) * 2 * p * (size_t)batch);
  orc_fft3d_f32(n, h, -1, nthreads);
  for (int b = 0; b < batch; ++b) {
    float *o = out + 2 * p * (size_t)b;
    orc_fft3d_f32(n, o, -1, nthreads);
    #pragma omp parallel for num_threads(nthreads > 0 ? nthreads : orc_max_threads())
    for (long long i = 0; i < (long long)p; ++i) {
      float sr = o[2 * i], si = o[2 * i + 1], fr = h[2 * i], fi = h[2 * i + 1];
      float tr = (sr * fr) - (si * fi);
      float ti = (sr * fi) + (si * fr);
      o[2 * i] = tr;
      o[2 * i + 1] = ti;
    }
    orc_fft3d_f32(n, o, +1, nthreads);
  }
  free(h);
  return 0;
}

int orc_conv3d_f32_spec(int n, const float *sig, const float *H, float *out, int nthreads) {
  if (!orc_is_pow2(n) || !sig || !H || !out) return -1;
  const size_t p = (size_t)n * n * n;
  memcpy(out, sig, sizeof(float) * 2 * p);
  orc_fft3d_f32(n, out, -1, nthreads);
  #pragma omp parallel for num_threads(nthreads > 0 ? nthreads : orc_max_threads())
  for (long long i = 0; i < (long long)p; ++i) {
    float sr = out[2 * i], si = out[2 * i + 1], fr = H[2 * i], fi = H[2 * i + 1];
    float tr = (sr * fr) - (si * fi);
    float ti = (sr * fi) + (si * fr);
    out[2 * i] = tr;
    out[2 * i + 1] = ti;
  }
  orc_fft3d_f32(n, out, +1, nthreads);
  return 0;
}

/* ------------------------------------------------------------- metrics */
/* fft_conv3D.cpp:236-250: per-point squares formed in fp32, sums in double. */
double orc_snr_db(const float *ref, const float *test, size_t npts) {
  double mag_sum = 0.0, noise_sum = 0.0;
  for (size_t i = 0; i < npts; ++i) {
    float rr = ref[2 * i], ri = ref[2 * i + 1];
    float dr = rr - test[2 * i], di = ri - test[2 * i + 1];
    double magnitude = rr * rr + ri * ri;
    double noise = dr * dr + di * di;
    mag_sum += magnitude;
    noise_sum += noise;
  }
  if (noise_sum == 0.0) return INFINITY;
  return 10.0 * log(mag_sum / noise_sum) / log(10.0);
}

double orc_rel_l2_f32_f64(const float *test, const double *ref, size_t npts) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < 2 * npts; ++i) {
    double d = (double)test[i] - ref[i];
    num += d * d;
    den += ref[i] * ref[i];
  }
  return den > 0.0 ? sqrt(num / den) : sqrt(num);
}

double orc_rel_l2_f32_f32(const float *test, const float *ref, size_t npts) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < 2 * npts; ++i) {
    double d = (double)test[i] - (double)ref[i];
    num += d * d;
    den += (double)ref[i] * (double)ref[i];
  }
  return den > 0.0 ? sqrt(num / den) : sqrt(num);
}

/* ------------------------------------------------ FFTW availability probe */
/* The CPU baseline's label names whether FFTW3f was present; the checker never
 * routes through it (its single-precision output is not reproducible here). */
int orc_fftw_available(void) {
  static int tried = 0, ok = 0;
  if (!tried) {
    tried = 1;
    void *lib = dlopen("libfftw3f.so.3", RTLD_NOW | RTLD_LOCAL);
    if (lib) {
      ok = dlsym(lib, "fftwf_plan_many_dft") && dlsym(lib, "fftwf_execute") && dlsym(lib, "fftwf_destroy_plan");
    }
  }
  return ok;
}
