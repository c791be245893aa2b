// convolution -- command-line harness of the library, with the flags, the
// dispatch rules and the report of the reference's sample application
// (src/main.cpp:11-106, src/helper.cpp:59-110 and 180-270), driving the
// drop-in C API (include/convfpga/convfpga.h) on the GPU.
//
//   -p, --path        code object (default: convfpga_gfx950.hsaco next to this binary)
//   -w, --wisdomfile  FFTW wisdom file of the CPU paths (default a.out)
//   -n, --num         cube edge N (default 64)
//   -i, --iter        iterations (default 1)
//   -t, --threads     CPU threads of the FFTW paths (default 1)
//   -y, --noverify    skip the check against the CPU FFTW convolution
//   -b, --batch       number of signals (default 1; > 1 needs --usesvm)
//   -c, --cpu-only    time the CPU FFTW convolution only
//   -s, --usesvm      use the host-mapped (_svm) calls
//   -e, --emulate     the reference's FPGA emulator: not available for the GPU
//   -h, --help
//   --fftw-plan       estimate | measure | patient | exhaustive: FFTW planning
//                     rigour of the CPU-only mode (the reference fixes it when
//                     it is built, FFTW_PLAN in src/CMakeLists.txt:5-6;
//                     default estimate as there)
//   --dump-out <file> after the last iteration, write the filter, the signals
//                     and the outputs (float32 re/im interleaved, in that
//                     order: N^3, batch*N^3 and batch*N^3 points) to <file>,
//                     so that a checker can verify the numbers without FFTW
//
// The CPU check and the CPU-only mode follow src/fft_conv3D.cpp:50-266 on
// FFTW single precision, loaded at run time (libfftw3f.so.3 and its threads
// library); without it they report "FFTW not found" as the reference does
// when it was built without FFTW (src/main.cpp:17-32).
#include <dlfcn.h>
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "convfpga/convfpga.h"

namespace {

struct Config {
  std::string path, wisdomfile = "a.out", fftw_plan = "estimate", dump_out;
  unsigned num = 64, iter = 1, threads = 1, batch = 1;
  bool noverify = false, cpuonly = false, usesvm = false, emulate = false;
};

double now_ms() {
  timespec a;
  clock_gettime(CLOCK_MONOTONIC, &a);
  return (double)a.tv_nsec * 1.0e-6 + (double)a.tv_sec * 1.0e3;
}

// ---- FFTW single precision, resolved at run time --------------------------
// (public API of fftw3.h: names, signatures and flag values)
typedef float fftwf_cpx[2];
typedef void* fftwf_plan_t;
constexpr int kFftwForward = -1, kFftwBackward = 1;
constexpr unsigned kFftwEstimate = 1u << 6, kFftwMeasure = 0u, kFftwPatient = 1u << 5, kFftwExhaustive = 1u << 3;
constexpr unsigned kFftwWisdomOnly = 1u << 21;

// planner flag of a --fftw-plan name and the label the reference prints
// (src/fft_conv3D.cpp:72-83); false for an unknown name
bool fftw_plan_flag(const std::string& name, unsigned* flag, const char** label) {
  static const struct {
    const char* name;
    unsigned flag;
    const char* label;
  } modes[] = {{"estimate", kFftwEstimate, "Estimate"},
               {"measure", kFftwMeasure, "Measure"},
               {"patient", kFftwPatient, "Patient"},
               {"exhaustive", kFftwExhaustive, "Exhaustive"}};
  for (const auto& m : modes)
    if (name == m.name) {
      *flag = m.flag;
      *label = m.label;
      return true;
    }
  return false;
}

struct Fftw {
  void* lib = nullptr;
  void* thr = nullptr;
  fftwf_plan_t (*plan_many_dft)(int, const int*, int, fftwf_cpx*, const int*, int, int, fftwf_cpx*, const int*, int,
                                int, int, unsigned) = nullptr;
  void (*execute)(fftwf_plan_t) = nullptr;
  void (*destroy_plan)(fftwf_plan_t) = nullptr;
  void* (*malloc_)(size_t) = nullptr;
  void (*free_)(void*) = nullptr;
  int (*init_threads)() = nullptr;
  void (*plan_with_nthreads)(int) = nullptr;
  void (*cleanup_threads)() = nullptr;
  int (*import_wisdom)(const char*) = nullptr;
  int (*export_wisdom)(const char*) = nullptr;

  bool load() {
    lib = dlopen("libfftw3f.so.3", RTLD_NOW | RTLD_GLOBAL);
    if (!lib) return false;
    thr = dlopen("libfftw3f_threads.so.3", RTLD_NOW | RTLD_GLOBAL);
    auto sym = [&](const char* n) -> void* {
      void* p = dlsym(lib, n);
      if (!p && thr) p = dlsym(thr, n);
      return p;
    };
    plan_many_dft = (decltype(plan_many_dft))sym("fftwf_plan_many_dft");
    execute = (decltype(execute))sym("fftwf_execute");
    destroy_plan = (decltype(destroy_plan))sym("fftwf_destroy_plan");
    malloc_ = (decltype(malloc_))sym("fftwf_malloc");
    free_ = (decltype(free_))sym("fftwf_free");
    init_threads = (decltype(init_threads))sym("fftwf_init_threads");
    plan_with_nthreads = (decltype(plan_with_nthreads))sym("fftwf_plan_with_nthreads");
    cleanup_threads = (decltype(cleanup_threads))sym("fftwf_cleanup_threads");
    import_wisdom = (decltype(import_wisdom))sym("fftwf_import_wisdom_from_filename");
    export_wisdom = (decltype(export_wisdom))sym("fftwf_export_wisdom_to_filename");
    return plan_many_dft && execute && destroy_plan && malloc_ && free_;
  }
};

fftwf_plan_t plan3(Fftw& f, unsigned num, int howmany, fftwf_cpx* data, int sign, unsigned flags) {
  const int n[3] = {(int)num, (int)num, (int)num};
  const int dist = (int)(num * num * num);
  return f.plan_many_dft(3, n, howmany, data, nullptr, 1, dist, data, nullptr, 1, dist, sign, flags);
}

// N^3-unnormalised FFT convolution of the batch with the shared filter on the
// CPU and its SNR against `out` (src/fft_conv3D.cpp:186-266: in-place
// FFTW_ESTIMATE plans, product re = sr fr - si fi, im = sr fi + si fr,
// double sums, pass at > 120 dB).
bool cpu_verify(Fftw& f, const Config& c, const float2* sig, const float2* filt, const float2* out) {
  const size_t P = (size_t)c.num * c.num * c.num;
  fftwf_cpx* fs = (fftwf_cpx*)f.malloc_(sizeof(fftwf_cpx) * P * c.batch);
  fftwf_cpx* ff = (fftwf_cpx*)f.malloc_(sizeof(fftwf_cpx) * P);
  fftwf_plan_t ps = plan3(f, c.num, (int)c.batch, fs, kFftwForward, kFftwEstimate);
  fftwf_plan_t pi = plan3(f, c.num, (int)c.batch, fs, kFftwBackward, kFftwEstimate);
  fftwf_plan_t pf = plan3(f, c.num, 1, ff, kFftwForward, kFftwEstimate);
  for (size_t i = 0; i < P * c.batch; ++i) {
    fs[i][0] = sig[i].x;
    fs[i][1] = sig[i].y;
  }
  for (size_t i = 0; i < P; ++i) {
    ff[i][0] = filt[i].x;
    ff[i][1] = filt[i].y;
  }
  f.execute(pf);
  f.execute(ps);
  for (unsigned b = 0; b < c.batch; ++b)
    for (size_t i = 0; i < P; ++i) {
      fftwf_cpx& s = fs[b * P + i];
      const float re = s[0] * ff[i][0] - s[1] * ff[i][1];
      const float im = s[0] * ff[i][1] + s[1] * ff[i][0];
      s[0] = re;
      s[1] = im;
    }
  f.execute(pi);
  double mag = 0.0, noise = 0.0;
  for (size_t i = 0; i < P * c.batch; ++i) {
    const float dr = fs[i][0] - out[i].x, di = fs[i][1] - out[i].y;
    mag += (double)(fs[i][0] * fs[i][0] + fs[i][1] * fs[i][1]);
    noise += (double)(dr * dr + di * di);
  }
  f.destroy_plan(ps);
  f.destroy_plan(pi);
  f.destroy_plan(pf);
  f.free_(fs);
  f.free_(ff);
  const double db = 10.0 * log10(mag / noise);
  if (db > 120.0) return true;
  printf("Signal to noise ratio on output sample: %g --> FAILED\n\n", db);
  return false;
}

// rand()/RAND_MAX in [0, 1] for re then im from glibc's default seed, every
// instance of the batch a copy of the first (src/helper.cpp:148-172)
void create_data(float2* inp, unsigned npts, unsigned batch) {
  for (unsigned j = 0; j < npts; ++j) {
    inp[j].x = (float)rand() / (float)RAND_MAX;
    inp[j].y = (float)rand() / (float)RAND_MAX;
  }
  for (unsigned i = 1; i < batch; ++i) memcpy(inp + (size_t)i * npts, inp, sizeof(float2) * npts);
}

// FFTW-only timing: filter transform once, then iter x (transform, product,
// inverse) (src/fft_conv3D.cpp:50-177)
int cpu_only(Fftw& f, const Config& c) {
  const size_t P = (size_t)c.num * c.num * c.num;
  if (f.init_threads && f.plan_with_nthreads) {
    if (f.init_threads() == 0) {
      fprintf(stderr, "Something went wrong with Multithreaded FFTW! Exiting...\n");
      return EXIT_FAILURE;
    }
    f.plan_with_nthreads((int)c.threads);
  }
  fftwf_cpx* fs = (fftwf_cpx*)f.malloc_(sizeof(fftwf_cpx) * P);
  fftwf_cpx* ff = (fftwf_cpx*)f.malloc_(sizeof(fftwf_cpx) * P);
  unsigned flags = kFftwEstimate;
  const char* label = "Estimate";
  fftw_plan_flag(c.fftw_plan, &flags, &label);  // validated by parse_args
  printf("-- Creating Plan\n-- FFTW Plan: %s\n", label);
  const bool wis = f.import_wisdom && f.import_wisdom(c.wisdomfile.c_str()) != 0;
  printf("-- %s wisdom from %s\n", wis ? "Importing" : "Cannot import", c.wisdomfile.c_str());
  if (wis) flags = kFftwWisdomOnly | kFftwEstimate;  // plans come from the wisdom (fft_conv3D.cpp:85-92)
  const double t0 = now_ms();
  fftwf_plan_t pf = plan3(f, c.num, 1, ff, kFftwForward, flags);
  fftwf_plan_t ps = plan3(f, c.num, 1, fs, kFftwForward, flags);
  fftwf_plan_t pi = plan3(f, c.num, 1, fs, kFftwBackward, flags);
  printf("-- Time to Plan: %g\n", now_ms() - t0);
  if (!wis && f.export_wisdom) f.export_wisdom(c.wisdomfile.c_str());
  create_data((float2*)ff, (unsigned)P, 1);
  create_data((float2*)fs, (unsigned)P, 1);
  double tf = now_ms();
  f.execute(pf);
  tf = now_ms() - tf;
  double tc = 0.0;
  for (unsigned it = 0; it < c.iter; ++it) {
    const double a = now_ms();
    f.execute(ps);
#pragma omp parallel for num_threads(c.threads)
    for (size_t i = 0; i < P; ++i) {
      const float re = fs[i][0] * ff[i][0] - fs[i][1] * ff[i][1];
      const float im = fs[i][0] * ff[i][1] + fs[i][1] * ff[i][0];
      fs[i][0] = re;
      fs[i][1] = im;
    }
    f.execute(pi);
    tc += now_ms() - a;
  }
  f.destroy_plan(pf);
  f.destroy_plan(ps);
  f.destroy_plan(pi);
  if (f.cleanup_threads) f.cleanup_threads();
  f.free_(fs);
  f.free_(ff);
  printf("\n\nMEASUREMENTS \n--------------\n");
  printf("Points           : %u^3\nThreads          : %u\nIterations       : %u\n\n", c.num, c.threads, c.iter);
  printf("CPU:\n----\nFilter Runtime   : %gms\nConv3D Runtime   : %gms\n\n", tf, tc / c.iter);
  return EXIT_SUCCESS;
}

// mean and population standard deviation of every field (src/helper.cpp:180-270)
void disp_results(const Config& c, const std::vector<fpga_t>& rt) {
  double m[6] = {0}, v[6] = {0};
  auto fields = [](const fpga_t& t, double* o) {
    o[0] = t.filter_pcie_wr_t;
    o[1] = t.filter_exec_t;
    o[2] = t.sig_pcie_wr_t;
    o[3] = t.sig_pcie_rd_t;
    o[4] = t.sig_exec_t;
    o[5] = t.siginv_exec_t;
  };
  for (const auto& t : rt) {
    double o[6];
    fields(t, o);
    for (int k = 0; k < 6; ++k) m[k] += o[k];
  }
  for (double& x : m) x /= (double)c.iter;
  if (c.batch > 1 && c.usesvm) m[5] = m[3] = m[2] = 0.0;
  for (const auto& t : rt) {
    double o[6];
    fields(t, o);
    for (int k = 0; k < 6; ++k) v[k] += (o[k] - m[k]) * (o[k] - m[k]);
  }
  double sd[6];
  for (int k = 0; k < 6; ++k) sd[k] = sqrt(v[k] / (double)c.iter);
  printf("\n\n------------------------------------------\nMeasurements in ms\n");
  printf("--------------------------------------------\n%s", c.iter > 1 ? "Average Measurements of iterations\n" : "");
  printf("GPU:\n-----\n- Filter:\n");
  if (!c.usesvm) printf("  PCIe Host to Device : %g\n", m[0]);
  printf("  Execution           : %g\n\n- Signal Convolution:\n", m[1]);
  if (!c.usesvm) printf("  PCIe Host to Device : %g\n", m[2]);
  printf("  FFT + Conv          : %g\n  Inverse FFT         : %g\n  Total Computation   : %g\n", m[4], m[5],
         m[4] + m[5]);
  if (c.batch > 1) printf("  Total Comp per batch: %g\n", (m[4] + m[5]) / c.batch);
  if (!c.usesvm) printf("  PCIe Dev to Host    : %g\n", m[3]);
  if (c.iter > 1) {
    printf("\nDeviation of runtimes among iterations\n- Filter:\n");
    if (!c.usesvm) printf("PCIe Host to Dev    = %.4lfms\n", sd[0]);
    printf("Execution    = %.4lfms\n- Signal:\n", sd[1]);
    printf("FFT + Conv          = %.4lfms\nInverse FFT         = %.4lfms\nTotal Computation   = %.4lfms\n", sd[4],
           sd[5], sd[4] + sd[5]);
    if (!c.usesvm) printf("PCIe Host to Dev      = %.4lfms\nPCIe Dev to Host      = %.4lfms\n", sd[2], sd[3]);
    printf("\n");
  }
}

// filter (nf points), signals and outputs (ns points each), float32 re/im
bool dump(const std::string& file, const float2* filter, size_t nf, const float2* sig, const float2* out, size_t ns) {
  FILE* fp = fopen(file.c_str(), "wb");
  if (!fp) return false;
  bool ok = fwrite(filter, sizeof(float2), nf, fp) == nf && fwrite(sig, sizeof(float2), ns, fp) == ns &&
            fwrite(out, sizeof(float2), ns, fp) == ns;
  ok = fclose(fp) == 0 && ok;
  return ok;
}

void usage(const char* argv0) {
  printf("Usage: %s [options]\n"
         "  -p, --path <file>        code object (default: convfpga_gfx950.hsaco next to this binary)\n"
         "  -w, --wisdomfile <file>  FFTW wisdom file (default a.out)\n"
         "  -n, --num <N>            cube edge (default 64)\n"
         "  -i, --iter <k>           iterations (default 1)\n"
         "  -t, --threads <k>        CPU threads of the FFTW paths (default 1)\n"
         "  -y, --noverify           no verification\n"
         "  -b, --batch <k>          number of signals (default 1)\n"
         "  -c, --cpu-only           CPU FFTW only\n"
         "  -s, --usesvm             host-mapped calls\n"
         "  -e, --emulate            FPGA emulation (not available)\n"
         "  -h, --help               this text\n"
         "      --fftw-plan <mode>   estimate | measure | patient | exhaustive (CPU-only mode; default estimate)\n"
         "      --dump-out <file>    write filter, signals and outputs (float32 re/im) after the last iteration\n",
         argv0);
}

bool parse_unsigned(const char* s, unsigned* out) {
  char* end = nullptr;
  const unsigned long v = strtoul(s, &end, 10);
  if (!s[0] || *end || v > 0xffffffffUL) return false;
  *out = (unsigned)v;
  return true;
}

// 0 = run, 1 = exit with failure, 2 = exit with success (help)
int parse_args(int argc, char** argv, Config& c) {
  static const option opts[] = {{"path", required_argument, nullptr, 'p'},
                                {"wisdomfile", required_argument, nullptr, 'w'},
                                {"num", required_argument, nullptr, 'n'},
                                {"iter", required_argument, nullptr, 'i'},
                                {"threads", required_argument, nullptr, 't'},
                                {"noverify", no_argument, nullptr, 'y'},
                                {"batch", required_argument, nullptr, 'b'},
                                {"cpu-only", no_argument, nullptr, 'c'},
                                {"usesvm", no_argument, nullptr, 's'},
                                {"emulate", no_argument, nullptr, 'e'},
                                {"help", no_argument, nullptr, 'h'},
                                {"fftw-plan", required_argument, nullptr, 'P'},
                                {"dump-out", required_argument, nullptr, 'D'},
                                {nullptr, 0, nullptr, 0}};
  int o;
  while ((o = getopt_long(argc, argv, "p:w:n:i:t:yb:cseh", opts, nullptr)) != -1) {
    bool ok = true;
    switch (o) {
      case 'p': c.path = optarg; break;
      case 'w': c.wisdomfile = optarg; break;
      case 'n': ok = parse_unsigned(optarg, &c.num); break;
      case 'i': ok = parse_unsigned(optarg, &c.iter) && c.iter > 0; break;
      case 't': ok = parse_unsigned(optarg, &c.threads) && c.threads > 0; break;
      case 'y': c.noverify = true; break;
      case 'b': ok = parse_unsigned(optarg, &c.batch) && c.batch > 0; break;
      case 'c': c.cpuonly = true; break;
      case 's': c.usesvm = true; break;
      case 'e': c.emulate = true; break;
      case 'D': c.dump_out = optarg; break;
      case 'h': usage(argv[0]); return 2;
      case 'P': {
        unsigned flag;
        const char* label;
        c.fftw_plan = optarg;
        ok = fftw_plan_flag(c.fftw_plan, &flag, &label);
        break;
      }
      default: return 1;
    }
    if (!ok) {
      if (o == 'P')
        fprintf(stderr, "Error parsing options: bad value '%s' for --fftw-plan\n", optarg);
      else
        fprintf(stderr, "Error parsing options: bad value '%s' for -%c\n", optarg, o);
      return 1;
    }
  }
  if (optind < argc) {
    fprintf(stderr, "Error parsing options: unexpected argument '%s'\n", argv[optind]);
    return 1;
  }
  if (!c.cpuonly && c.path.empty()) {  // the code object installed beside this binary
    char self[4096];
    const ssize_t k = readlink("/proc/self/exe", self, sizeof self - 1);
    if (k > 0) {
      self[k] = 0;
      std::string dir(self);
      dir = dir.substr(0, dir.rfind('/'));
      c.path = dir + "/convfpga_gfx950.hsaco";
    }
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Config c;
  const int pr = parse_args(argc, argv, c);
  if (pr == 2) return EXIT_SUCCESS;
  if (pr != 0) return EXIT_FAILURE;
  printf("\nCONFIGURATION: \n---------------\n");
  unsigned plan_flag;
  const char* plan_label = "Estimate";
  fftw_plan_flag(c.fftw_plan, &plan_flag, &plan_label);
  printf("Code object  = %s\nPoints       = {%u, %u, %u}\nWisdom Path  = %s\nFFTW Plan    = %s\n",
         c.path.c_str(), c.num, c.num, c.num, c.wisdomfile.c_str(), plan_label);
  printf("Threads      = %u\nIterations   = %u\nBatch        = %u\n----------------\n\n", c.threads, c.iter,
         c.batch);

  Fftw fftw;
  const bool have_fftw = fftw.load();
  if (c.cpuonly) {
    if (!have_fftw) {
      fprintf(stderr, "FFTW not found\n");
      return EXIT_FAILURE;
    }
    return cpu_only(fftw, c);
  }
  if (c.emulate) {
    fprintf(stderr, "no emulation platform: the GPU build runs on the device only\n");
    return EXIT_FAILURE;
  }
  if (!c.usesvm && c.batch > 1) {  // dispatch rules of src/main.cpp:67-74
    fprintf(stderr, "Non-SVM Batch not implemented for Convolution 3D\n");
    return EXIT_FAILURE;
  }
  if (!c.noverify && !have_fftw) printf("-- FFTW not found: results are not verified\n");

  if (fpga_initialize("gfx950", c.path.c_str(), c.usesvm) != 0) {
    fprintf(stderr, "GPU initialization error\n");
    return EXIT_FAILURE;
  }
  const size_t P = (size_t)c.num * c.num * c.num;
  float2* filter = (float2*)fpgaf_complex_malloc(sizeof(float2) * P);
  float2* sig = (float2*)fpgaf_complex_malloc(sizeof(float2) * P * c.batch);
  float2* out = (float2*)fpgaf_complex_malloc(sizeof(float2) * P * c.batch);
  std::vector<fpga_t> rt(c.iter);
  int rc = EXIT_SUCCESS;
  if (!filter || !sig || !out) {
    fprintf(stderr, "host allocation failed\n");
    rc = EXIT_FAILURE;
  } else {
    create_data(filter, (unsigned)P, 1);
    create_data(sig, (unsigned)P, c.batch);
    for (unsigned i = 0; i < c.iter && rc == EXIT_SUCCESS; ++i) {
      printf("\n%u: Calculating Conv3D\n", i);
      if (c.usesvm && c.batch > 1)
        rt[i] = fpgaf_conv3D_svm_batch(c.num, sig, filter, out, c.batch);
      else if (c.usesvm)
        rt[i] = fpgaf_conv3D_svm(c.num, sig, filter, out);
      else
        rt[i] = fpgaf_conv3D(c.num, sig, filter, out);
      if (!rt[i].valid) {
        fprintf(stderr, "GPU execution found invalid\n");
        rc = EXIT_FAILURE;
      } else if (!c.noverify && have_fftw && !cpu_verify(fftw, c, sig, filter, out)) {
        fprintf(stderr, "Iter %u: GPU result incorrect in comparison to CPU\n", i);
        rc = EXIT_FAILURE;
      }
    }
  }
  fpga_final();
  if (rc == EXIT_SUCCESS && !c.dump_out.empty() && !dump(c.dump_out, filter, P, sig, out, P * c.batch)) {
    fprintf(stderr, "cannot write %s\n", c.dump_out.c_str());
    rc = EXIT_FAILURE;
  }
  if (rc == EXIT_SUCCESS) disp_results(c, rt);
  free(sig);
  free(filter);
  free(out);
  return rc;
}
