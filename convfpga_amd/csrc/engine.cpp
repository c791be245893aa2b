// engine.cpp -- HIP host runtime of libconvfpga: device selection, the kernel
// code object, per-size twiddle tables, workspaces, pass scheduling and timing.
//
// Replaces the reference's OpenCL host layer (host/src/fpga_init.c,
// host/src/convfpga*.c, host/src/opencl_utils.c).  Differences by design:
//  * one in-order HIP stream instead of eight profiling queues per call
//    (fpga_init.c:142-183): the passes are dependent kernels, not a dataflow
//    pipeline of concurrently running tasks;
//  * device buffers, pinned staging and twiddle tables are cached across calls
//    (the reference creates and frees 4 buffers and 8 kernels per call,
//    convfpga.c:40-72, 377-405);
//  * errors come back as return codes (the reference exit()s,
//    opencl_utils.c:421-439);
//  * every entry point holds one process-wide lock (the reference keeps global
//    OpenCL state and is not thread-safe, fpga_state.h:8-17).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch_desc.h"
#include "batch_spans.h"
#include "convfpga/convfpga_ext.h"
#include "engine.h"
#include "geometry.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return fail(CF_ERR_RUNTIME, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define CFCHK(expr)          \
  do {                       \
    int r_ = (expr);         \
    if (r_ != CF_OK) return r_; \
  } while (0)

struct SizeFns {
  hipFunction_t row_fwd = nullptr, row_inv = nullptr, coly_fwd = nullptr, coly_inv = nullptr;
  hipFunction_t colz_fwd = nullptr, colz_inv = nullptr, colz_conv = nullptr;
  hipFunction_t plane_fwd = nullptr, plane_inv = nullptr;                 // N in {32, 64, 128}
  hipFunction_t coly_fwd_zb = nullptr, coly_inv_zb = nullptr, colz_fwd_zb = nullptr, colz_conv_zb = nullptr;  // N >= 64
  hipFunction_t xy1_fwd[cfgeo::kSplitVariants] = {}, xy1_inv[cfgeo::kSplitVariants] = {};  // N = 256
  hipFunction_t mid[cfgeo::kSplitVariants] = {}, mspec[cfgeo::kSplitVariants] = {};
  hipFunction_t mx_xy1_fwd = nullptr, mx_xy1_inv = nullptr, mx_mid = nullptr;  // N = 256, split-FP16 MFMA
  // N = 512, split-FP16 MFMA axis passes: natural layout, then z-blocked (AXIS5Z)
  hipFunction_t mx_row_fwd = nullptr, mx_row_inv = nullptr, mx_coly_fwd = nullptr, mx_coly_inv = nullptr;
  hipFunction_t mx_colz_conv = nullptr, mx_coly_fwd_zb = nullptr, mx_colz_conv_zb = nullptr, mx_coly_inv_zb = nullptr;
  void* tw = nullptr;  // N float2 twiddles W_N^j
  int forced = CF_PLAN_AUTO;  // cf_set_plan
  int tuned = CF_PLAN_AUTO;   // autotuner's choice, CF_PLAN_AUTO until it has run
  float tune_ms[CF_NUM_PLANS] = {};
};

// four-step plans: CF_PLAN_SPLIT3_G22 .. G28M1 are variants 0..8, then
// CF_PLAN_SPLIT3_G24M2O .. G18M2OC4 variants 9..15 (geometry.h)
bool is_split(int plan) {
  return (plan >= CF_PLAN_SPLIT3_G22 && plan <= CF_PLAN_SPLIT3_G28M1) ||
         (plan >= CF_PLAN_SPLIT3_G24M2O && plan <= CF_PLAN_SPLIT3_G18M2OC4);
}
int split_variant(int plan) {
  return plan <= CF_PLAN_SPLIT3_G28M1 ? plan - CF_PLAN_SPLIT3_G22 : 9 + (plan - CF_PLAN_SPLIT3_G24M2O);
}
static_assert(cfgeo::kSplitVariants == 16 && CF_PLAN_SPLIT3_G28M1 - CF_PLAN_SPLIT3_G22 == 8 &&
                  CF_PLAN_SPLIT3_G28M2OC4 == CF_PLAN_SPLIT3_G24M2O + 5 &&
                  CF_PLAN_SPLIT3_G18M2OC4 == CF_PLAN_SPLIT3_G28M2OC4 + 1 && CF_NUM_PLANS == CF_PLAN_SPLIT3_G18M2OC4 + 1,
              "plan ids of the four-step variants");
// plans whose passes go through the workspace (one cube per convolution in flight)
bool needs_work(int plan) { return is_split(plan) || plan == CF_PLAN_AXIS5Z; }
// four-step plans whose middle pass writes a second intermediate
bool two_intermediates(int plan) { return is_split(plan) && cfgeo::split_out_of_place(split_variant(plan)); }

bool plan_ok(unsigned n, int plan) {
  if (!cfgeo::supported((int)n)) return false;
  if (plan == CF_PLAN_AXIS5) return true;
  if (plan == CF_PLAN_PLANE3) return cfgeo::plane_supported((int)n);
  if (is_split(plan)) return cfgeo::split_supported((int)n, split_variant(plan));
  if (plan == CF_PLAN_AXIS5Z) return n >= 64 && cfgeo::supported((int)n);
  return false;
}

struct Engine {
  std::recursive_mutex mu;
  bool ready = false;
  bool svm = false;
  int device = 0;
  hipModule_t mod = nullptr;
  hipStream_t stream = nullptr;
  hipFunction_t twgen = nullptr, dft16gen = nullptr, errp = nullptr, errf = nullptr, fillu = nullptr, filll = nullptr;
  hipFunction_t batch_run = nullptr, batch_run_3p = nullptr, batch_run_3p_small = nullptr;
  hipStream_t capture = nullptr;  // private stream for graph capture
  std::map<unsigned, SizeFns> sizes;
  // workspace of the host-buffer (drop-in) calls
  void* d_data = nullptr;
  size_t d_data_bytes = 0;
  void* d_filt = nullptr;
  size_t d_filt_bytes = 0;
  void* h_sig = nullptr;   // pinned, host-mapped
  size_t h_sig_bytes = 0;
  void* h_filt = nullptr;
  size_t h_filt_bytes = 0;
  void* h_out = nullptr;
  // pinned staging of the copying drop-in call: two chunk buffers, each with
  // the event of the last transfer that used it
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  bool stage_used[2] = {false, false};
  size_t h_out_bytes = 0;
  // workspace of the 3-pass four-step plans, ordered across streams by work_ev
  void* d_work = nullptr;
  size_t d_work_bytes = 0;
  hipStream_t work_stream = nullptr;
  bool work_used = false;
  hipEvent_t work_ev = nullptr;
  std::string arch;
  // error-norm scratch
  void* d_err = nullptr;
  size_t d_err_bytes = 0;
  hipEvent_t ev[CF_MAX_PASSES + 8] = {};
  // device address of the code object's cf_scratch array on g.device (the
  // only device this engine drives), looked up once when the module is loaded
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // cf_profile_begin/end: events of up to prof_max calls, (CF_MAX_PASSES+1) each
  std::vector<hipEvent_t> prof_ev;
  unsigned prof_max = 0, prof_calls = 0, prof_npass = 0;
  unsigned prof_every = 1, prof_seen = 0;  // record the last call of every group of prof_every while armed
  std::vector<std::string> prof_names;
  std::vector<double> prof_bytes;
};

Engine g;
constexpr int kErrBlocks = 1024;
constexpr int kNumEvents = CF_MAX_PASSES + 8;

size_t cube_elems(unsigned n) { return (size_t)n * n * n; }
size_t cube_bytes(unsigned n) { return cube_elems(n) * 8; }

std::string lower(const std::string& s) {
  std::string r = s;
  for (auto& ch : r) ch = (char)std::tolower((unsigned char)ch);
  return r;
}

void pipe_release();

void release_all() {
  if (g.ready || g.stream) (void)hipSetDevice(g.device);
  if (g.stream) (void)hipStreamSynchronize(g.stream);
  pipe_release();
  for (auto& kv : g.sizes)
    if (kv.second.tw) (void)hipFree(kv.second.tw);
  g.sizes.clear();
  if (g.d_data) (void)hipFree(g.d_data);
  if (g.d_filt) (void)hipFree(g.d_filt);
  if (g.d_err) (void)hipFree(g.d_err);
  if (g.d_work) (void)hipFree(g.d_work);
  g.d_work = nullptr;
  g.d_work_bytes = 0;
  g.work_used = false;
  g.work_stream = nullptr;
  if (g.work_ev) (void)hipEventDestroy(g.work_ev);
  g.work_ev = nullptr;
  if (g.h_sig) (void)hipHostFree(g.h_sig);
  if (g.h_filt) (void)hipHostFree(g.h_filt);
  if (g.h_out) (void)hipHostFree(g.h_out);
  for (int b = 0; b < 2; ++b) {
    if (g.stage[b]) (void)hipHostFree(g.stage[b]);
    if (g.stage_ev[b]) (void)hipEventDestroy(g.stage_ev[b]);
    g.stage[b] = nullptr;
    g.stage_ev[b] = nullptr;
    g.stage_used[b] = false;
  }
  g.d_data = g.d_filt = g.d_err = g.h_sig = g.h_filt = g.h_out = nullptr;
  g.d_data_bytes = g.d_filt_bytes = g.d_err_bytes = g.h_sig_bytes = g.h_filt_bytes = g.h_out_bytes = 0;
  for (auto& e : g.ev)
    if (e) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
  for (auto e : g.prof_ev) (void)hipEventDestroy(e);
  g.prof_ev.clear();
  g.prof_max = g.prof_calls = 0;
  if (g.stream) (void)hipStreamDestroy(g.stream);
  g.stream = nullptr;
  if (g.mod) (void)hipModuleUnload(g.mod);
  g.mod = nullptr;
  g.twgen = g.dft16gen = g.errp = g.errf = g.fillu = g.filll = g.batch_run = g.batch_run_3p = g.batch_run_3p_small = nullptr;
  g.scratch = nullptr;  // part of the unloaded module
  g.scratch_bytes = 0;
  if (g.capture) (void)hipStreamDestroy(g.capture);
  g.capture = nullptr;
  g.ready = false;
  g.svm = false;
}

int get_fn(hipFunction_t* f, const char* name) {
  hipError_t e = hipModuleGetFunction(f, g.mod, name);
  if (e != hipSuccess) return fail(CF_ERR_RUNTIME, "kernel %s not in code object: %s", name, hipGetErrorString(e));
  return CF_OK;
}

int require_ready() {
  if (!g.ready) return fail(CF_ERR_STATE, "fpga_initialize() has not succeeded");
  hipError_t e = hipSetDevice(g.device);
  if (e != hipSuccess) return fail(CF_ERR_RUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
  return CF_OK;
}

// Autotuner cache: lines "<arch> <N> <plan>" in $CONVFPGA_TUNE_FILE (last one wins).
void read_tune_file(unsigned n, SizeFns* s) {
  const char* path = getenv("CONVFPGA_TUNE_FILE");
  if (!path || !*path) return;
  FILE* fp = fopen(path, "r");
  if (!fp) return;
  char arch[128];
  unsigned nn = 0;
  int plan = 0;
  char line[512];
  while (fgets(line, sizeof line, fp)) {
    if (sscanf(line, "%127s %u %d", arch, &nn, &plan) == 3 && nn == n && g.arch == arch && plan_ok(n, plan))
      s->tuned = plan;
  }
  fclose(fp);
}

void write_tune_file(unsigned n, const SizeFns& s) {
  const char* path = getenv("CONVFPGA_TUNE_FILE");
  if (!path || !*path) return;
  FILE* fp = fopen(path, "a");
  if (!fp) return;
  fprintf(fp, "%s %u %d", g.arch.c_str(), n, s.tuned);
  for (int p = 1; p < CF_NUM_PLANS; ++p) fprintf(fp, " %.4f", s.tune_ms[p]);
  fprintf(fp, "\n");
  fclose(fp);
}

// Kernels and FP64-generated twiddle table of edge n (created on first use).
int size_res(unsigned n, SizeFns** out) {
  if (!cfgeo::supported((int)n)) return fail(CF_ERR_SIZE, "N=%u is not one of 8,16,32,64,128,256,512", n);
  auto it = g.sizes.find(n);
  if (it != g.sizes.end()) {
    *out = &it->second;
    return CF_OK;
  }
  SizeFns s;
  char name[64];
  struct {
    hipFunction_t* f;
    const char* fmt;
  } tab[] = {{&s.row_fwd, "cf_row_fwd_%u"},   {&s.row_inv, "cf_row_inv_%u"},   {&s.coly_fwd, "cf_coly_fwd_%u"},
             {&s.coly_inv, "cf_coly_inv_%u"}, {&s.colz_fwd, "cf_colz_fwd_%u"}, {&s.colz_inv, "cf_colz_inv_%u"},
             {&s.colz_conv, "cf_colz_conv_%u"}};
  for (auto& t : tab) {
    snprintf(name, sizeof name, t.fmt, n);
    CFCHK(get_fn(t.f, name));
  }
  if (plan_ok(n, CF_PLAN_AXIS5Z)) {
    struct {
      hipFunction_t* f;
      const char* fmt;
    } zb[] = {{&s.coly_fwd_zb, "cf_coly_fwd_zb_%u"},
              {&s.coly_inv_zb, "cf_coly_inv_zb_%u"},
              {&s.colz_fwd_zb, "cf_colz_fwd_zb_%u"},
              {&s.colz_conv_zb, "cf_colz_conv_zb_%u"}};
    for (auto& t : zb) {
      snprintf(name, sizeof name, t.fmt, n);
      CFCHK(get_fn(t.f, name));
    }
  }
  if (cfgeo::plane_supported((int)n)) {
    snprintf(name, sizeof name, "cf_plane_fwd_%u", n);
    CFCHK(get_fn(&s.plane_fwd, name));
    snprintf(name, sizeof name, "cf_plane_inv_%u", n);
    CFCHK(get_fn(&s.plane_inv, name));
  }
  for (int vi = 0; vi < cfgeo::kSplitVariants; ++vi) {
    if (!cfgeo::split_supported((int)n, vi)) continue;
    const int ga = cfgeo::split_ga(vi), gb = cfgeo::split_gb(vi), gm = cfgeo::split_gm(vi);
    const int gbo = cfgeo::split_gbo(vi), gao = cfgeo::split_gao(vi);
    // the middle pass (and its filter setup) depends on the first pass's
    // layout (x group gb), not on its y2 group: the first pass of one y2
    // value (ga = 1) shares the kernels of ga = 2
    const int gam = ga < 2 ? 2 : ga;
    char sfx[16];  // the in-place middle pass of the same input layout and tile (its filter setup too)
    if (gm == gb)
      snprintf(sfx, sizeof sfx, "g%d%d", gam, gb);
    else
      snprintf(sfx, sizeof sfx, "g%d%dm%d", gam, gb, gm);
    snprintf(name, sizeof name, "cf_xy1_fwd_%u_g%d%d", n, ga, gb);
    CFCHK(get_fn(&s.xy1_fwd[vi], name));
    // the last pass reads the middle pass's output layout: x group gbo, y2 pairs interleaved for "h"
    const bool ybo2 = cfgeo::split_ybo(vi) == 2;
    const bool k1o = cfgeo::split_k1o(vi);  // second intermediate with k1 inside the x group
    snprintf(name, sizeof name, ybo2 ? "cf_xy1_inv_%u_g%d%dh" : k1o ? "cf_xy1_inv_%u_g%d%dk" : "cf_xy1_inv_%u_g%d%d", n,
             gao, gbo);
    CFCHK(get_fn(&s.xy1_inv[vi], name));
    if (ybo2 && gbo != gb)
      snprintf(name, sizeof name, "cf_mid_%u_%sh%d", n, sfx, gbo);
    else
      snprintf(name, sizeof name, ybo2 ? "cf_mid_%u_%sh" : gbo == gb ? "cf_mid_%u_%s" : k1o ? "cf_mid_%u_%sok" : "cf_mid_%u_%so",
               n, sfx);
    CFCHK(get_fn(&s.mid[vi], name));
    snprintf(name, sizeof name, "cf_midspec_%u_%s", n, sfx);
    CFCHK(get_fn(&s.mspec[vi], name));
  }
  if (n == 256) {
    CFCHK(get_fn(&s.mx_xy1_fwd, "cf_xy1_fwd_256_s24"));
    CFCHK(get_fn(&s.mx_xy1_inv, "cf_xy1_inv_256_s24"));
    CFCHK(get_fn(&s.mx_mid, "cf_mid_256_s24m2"));
  }
  if (n == 512) {
    CFCHK(get_fn(&s.mx_row_fwd, "cf_row_fwd_512_s"));
    CFCHK(get_fn(&s.mx_row_inv, "cf_row_inv_512_s"));
    CFCHK(get_fn(&s.mx_coly_fwd, "cf_coly_fwd_512_s"));
    CFCHK(get_fn(&s.mx_coly_inv, "cf_coly_inv_512_s"));
    CFCHK(get_fn(&s.mx_colz_conv, "cf_colz_conv_512_s"));
    CFCHK(get_fn(&s.mx_coly_fwd_zb, "cf_coly_fwd_zb_512_s"));
    CFCHK(get_fn(&s.mx_colz_conv_zb, "cf_colz_conv_zb_512_s"));
    CFCHK(get_fn(&s.mx_coly_inv_zb, "cf_coly_inv_zb_512_s"));
  }
  read_tune_file(n, &s);
  // twiddles W_N^j; for N = 256 and 512 followed by the DFT16 operand table
  // of the split-FP16 kernels (2 x 64 lanes x 4 x 16 bytes)
  const size_t tw_bytes = (size_t)n * 8, table_bytes = n == 256 || n == 512 ? 8192 : 0;
  HIPCHK(hipMalloc(&s.tw, tw_bytes + table_bytes));
  void* tw = s.tw;
  int nn = (int)n;
  void* args[] = {&tw, &nn};
  HIPCHK(hipModuleLaunchKernel(g.twgen, (n + 255) / 256, 1, 1, 256, 1, 1, 0, g.stream, args, nullptr));
  if (table_bytes) {
    void* table = (char*)s.tw + tw_bytes;
    void* targs[] = {&table};
    HIPCHK(hipModuleLaunchKernel(g.dft16gen, 1, 1, 1, 128, 1, 1, 0, g.stream, targs, nullptr));
  }
  HIPCHK(hipStreamSynchronize(g.stream));
  *out = &(g.sizes[n] = s);
  return CF_OK;
}

// ---- passes -------------------------------------------------------------------
struct Pass {
  hipFunction_t f;
  unsigned tiles, threads;  // work-groups (tiles) per cube, threads per work-group
  const void* in;
  void* out;
  const void* H;
  long long hstride;
  float scale;
  const char* name;
  double bytes;  // algorithmic HBM bytes for the whole launch
};

// The passes of one call: at most CF_MAX_PASSES, held in place (building a
// schedule touches no heap, so the allocation-free entry point allocates
// nothing at all).
struct PassList {
  Pass p[CF_MAX_PASSES];
  size_t n = 0;
  void push_back(const Pass& x) {
    if (n < CF_MAX_PASSES) p[n++] = x;  // every schedule here has at most 5 passes
  }
  size_t size() const { return n; }
  Pass& operator[](size_t i) { return p[i]; }
  const Pass& operator[](size_t i) const { return p[i]; }
  Pass& back() { return p[n - 1]; }
  const Pass* begin() const { return p; }
  const Pass* end() const { return p + n; }
};

int launch_pass(const Pass& p, unsigned n, unsigned nb, hipStream_t s, const void* tw) {
  const size_t ce = cube_elems(n);
  // one 1-D grid over the whole batch (one work-group per tile, tile indices
  // in int), split into launches of whole cubes below 2^31 work-items each
  const unsigned long long per_cube = (unsigned long long)std::max(1u, p.tiles) * p.threads;
  const unsigned max_cubes = (unsigned)std::max(1ull, 0x7fffffffull / per_cube);
  for (unsigned b0 = 0; b0 < nb; b0 += max_cubes) {
    const unsigned cnt = std::min(max_cubes, nb - b0);
    const void* in = (const char*)p.in + b0 * ce * 8;
    void* out = (char*)p.out + b0 * ce * 8;
    const void* H = p.H ? (const void*)((const char*)p.H + (size_t)b0 * p.hstride * 8) : nullptr;
    long long hs = p.hstride;
    float sc = p.scale;
    const void* twp = tw;
    int ntotal = (int)(p.tiles * cnt);
    void* args[] = {(void*)&in, (void*)&out, (void*)&H, (void*)&twp, (void*)&hs, (void*)&sc, (void*)&ntotal};
    HIPCHK(hipModuleLaunchKernel(p.f, (unsigned)ntotal, 1, 1, p.threads, 1, 1, 0, s, args, nullptr));
  }
  return CF_OK;
}

// forward 3D FFT: x, y, z passes (in may equal out)
void fwd_passes(const SizeFns& f, unsigned n, unsigned nb, const void* in, void* out, PassList& v) {
  const double io = 2.0 * cube_bytes(n) * nb;
  const unsigned rt = cfgeo::row_threads(n), rT = cfgeo::row_tiles(n);
  const unsigned ct = cfgeo::col_threads(n), cT = cfgeo::col_tiles(n);
  v.push_back({f.row_fwd, rT, rt, in, out, nullptr, 0, 1.0f, "x_fwd", io});
  v.push_back({f.coly_fwd, cT, ct, out, out, nullptr, 0, 1.0f, "y_fwd", io});
  v.push_back({f.colz_fwd, cT, ct, out, out, nullptr, 0, 1.0f, "z_fwd", io});
}

void inv_passes(const SizeFns& f, unsigned n, unsigned nb, const void* in, void* out, float scale, PassList& v) {
  const double io = 2.0 * cube_bytes(n) * nb;
  const unsigned rt = cfgeo::row_threads(n), rT = cfgeo::row_tiles(n);
  const unsigned ct = cfgeo::col_threads(n), cT = cfgeo::col_tiles(n);
  v.push_back({f.colz_inv, cT, ct, in, out, nullptr, 0, 1.0f, "z_inv", io});
  v.push_back({f.coly_inv, cT, ct, out, out, nullptr, 0, 1.0f, "y_inv", io});
  v.push_back({f.row_inv, rT, rt, out, out, nullptr, 0, scale, "x_inv", io});
}

// Convolution passes of `plan`.  AXIS5: x, y, z*H*z, y^-1, x^-1(+scale), in
// place in out.  PLANE3: xy-plane, z*H*z, xy-plane^-1(+scale), in place.
// SPLIT3: sig -> work (pass A), work in place (middle), work -> out (pass C).
// Returns the index of the first pass after the product (the fpga_t split).
size_t conv_passes(const SizeFns& f, int plan, unsigned n, unsigned nb, const void* sig, void* out, void* work,
                   const void* H, long long hstride, float scale, PassList& v,
                   int precision = CF_PREC_FP32) {
  const double V = (double)cube_bytes(n);
  const double io = 2.0 * V * nb;
  const double hbytes = hstride ? V * nb : V;
  if (precision == CF_PREC_SPLIT_FP16 && n == 512) {  // AXIS5 or AXIS5Z spectrum (checked by the caller)
    const unsigned rt = 256, rT = 512 * 512 / 8, ct = 512, cT = 32 * 512;
    v.push_back({f.mx_row_fwd, rT, rt, sig, out, nullptr, 0, 1.0f, "x_fwd_mfma", io});
    if (plan == CF_PLAN_AXIS5Z) {
      v.push_back({f.mx_coly_fwd_zb, cT, ct, out, work, nullptr, 0, 1.0f, "y_fwd_mfma", io});
      v.push_back({f.mx_colz_conv_zb, cT, ct, work, work, H, hstride, 1.0f, "z_mfma_mul_inv", io + hbytes});
      const size_t split = v.size();
      v.push_back({f.mx_coly_inv_zb, cT, ct, work, out, nullptr, 0, 1.0f, "y_inv_mfma", io});
      v.push_back({f.mx_row_inv, rT, rt, out, out, nullptr, 0, scale, "x_inv_mfma", io});
      return split;
    }
    v.push_back({f.mx_coly_fwd, cT, ct, out, out, nullptr, 0, 1.0f, "y_fwd_mfma", io});
    v.push_back({f.mx_colz_conv, cT, ct, out, out, H, hstride, 1.0f, "z_mfma_mul_inv", io + hbytes});
    const size_t split = v.size();
    v.push_back({f.mx_coly_inv, cT, ct, out, out, nullptr, 0, 1.0f, "y_inv_mfma", io});
    v.push_back({f.mx_row_inv, rT, rt, out, out, nullptr, 0, scale, "x_inv_mfma", io});
    return split;
  }
  if (precision == CF_PREC_SPLIT_FP16) {  // four-step plan with middle tiles of 2 x (checked by the caller)
    const unsigned at = 512, aT = 256 * 8, bt = 512, bT = 16 * 128;
    v.push_back({f.mx_xy1_fwd, aT, at, sig, work, nullptr, 0, 1.0f, "xy1_fwd_mfma", io});
    v.push_back({f.mx_mid, bT, bt, work, work, H, hstride, 1.0f, "y2z_mfma_mul_inv", io + hbytes});
    const size_t split = v.size();
    v.push_back({f.mx_xy1_inv, aT, at, work, out, nullptr, 0, scale, "xy1_inv_mfma", io});
    return split;
  }
  const unsigned rt = cfgeo::row_threads(n), rT = cfgeo::row_tiles(n);
  const unsigned ct = cfgeo::col_threads(n), cT = cfgeo::col_tiles(n);
  if (plan == CF_PLAN_PLANE3) {
    const unsigned pt = cfgeo::plane_threads(n), pT = cfgeo::plane_tiles(n);
    v.push_back({f.plane_fwd, pT, pt, sig, out, nullptr, 0, 1.0f, "xy_fwd", io});
    v.push_back({f.colz_conv, cT, ct, out, out, H, hstride, 1.0f, "z_fwd_mul_z_inv", io + hbytes});
    const size_t split = v.size();
    v.push_back({f.plane_inv, pT, pt, out, out, nullptr, 0, scale, "xy_inv", io});
    return split;
  }
  if (is_split(plan)) {
    const int vi = split_variant(plan);
    const unsigned at = cfgeo::split_a_threads((int)n, vi), aT = cfgeo::split_a_tiles((int)n, vi);
    const unsigned bt = cfgeo::split_b_threads((int)n, vi), bT = cfgeo::split_b_tiles((int)n, vi);
    const unsigned ct3 = cfgeo::split_c_threads((int)n, vi), cT3 = cfgeo::split_c_tiles((int)n, vi);
    void* w1 = work;  // first intermediate (the middle pass's input)
    void* w2 = work;  // second one (its output): the same cube unless the plan writes out of place
    if (two_intermediates(plan)) {
      if (sig != out) {
        w1 = out;  // overwritten by the last pass anyway
      } else {
        w2 = (char*)work + (size_t)nb * cube_bytes(n);  // in place: the workspace holds both
      }
    }
    v.push_back({f.xy1_fwd[vi], aT, at, sig, w1, nullptr, 0, 1.0f, "xy1_fwd", io});
    v.push_back({f.mid[vi], bT, bt, w1, w2, H, hstride, 1.0f, "y2z_fwd_mul_inv", io + hbytes});
    const size_t split = v.size();
    v.push_back({f.xy1_inv[vi], cT3, ct3, w2, out, nullptr, 0, scale, "xy1_inv", io});
    return split;
  }
  if (plan == CF_PLAN_AXIS5Z) {  // natural -> z-blocked in the workspace -> natural
    v.push_back({f.row_fwd, rT, rt, sig, out, nullptr, 0, 1.0f, "x_fwd", io});
    v.push_back({f.coly_fwd_zb, cT, ct, out, work, nullptr, 0, 1.0f, "y_fwd", io});
    v.push_back({f.colz_conv_zb, cT, ct, work, work, H, hstride, 1.0f, "z_fwd_mul_z_inv", io + hbytes});
    const size_t split = v.size();
    v.push_back({f.coly_inv_zb, cT, ct, work, out, nullptr, 0, 1.0f, "y_inv", io});
    v.push_back({f.row_inv, rT, rt, out, out, nullptr, 0, scale, "x_inv", io});
    return split;
  }
  v.push_back({f.row_fwd, rT, rt, sig, out, nullptr, 0, 1.0f, "x_fwd", io});
  v.push_back({f.coly_fwd, cT, ct, out, out, nullptr, 0, 1.0f, "y_fwd", io});
  v.push_back({f.colz_conv, cT, ct, out, out, H, hstride, 1.0f, "z_fwd_mul_z_inv", io + hbytes});
  const size_t split = v.size();
  v.push_back({f.coly_inv, cT, ct, out, out, nullptr, 0, 1.0f, "y_inv", io});
  v.push_back({f.row_inv, rT, rt, out, out, nullptr, 0, scale, "x_inv", io});
  return split;
}

// Workspace bytes conv_passes needs for nb cubes of edge n under `plan`: one
// intermediate cube per convolution, two for a plan with two intermediates
// run in place (out of place, the output buffer holds the first).
size_t work_bytes(int plan, unsigned n, unsigned nb, const void* sig, const void* out) {
  if (!needs_work(plan)) return 0;
  const size_t one = cube_bytes(n) * nb;
  return two_intermediates(plan) && sig == out ? 2 * one : one;
}

// Filter setup of `plan`: the forward FFT3 of `count` filters into dst, in the
// layout the plan's product pass reads (in place when src == dst).  Natural
// order by the axis passes (PLANE3: the plane pass and z); z-blocked for
// AXIS5Z; for the four-step plans the
// first pass of the convolution into the workspace, then the forward half of
// the middle pass straight into the middle pass's spectrum order (4 volume
// transfers instead of 6 for the axis passes and 2 for a re-layout).
void spectrum_passes(const SizeFns& f, int plan, unsigned n, unsigned count, const void* src, void* dst, void* work,
                     PassList& v) {
  const double io = 2.0 * cube_bytes(n) * count;
  if (plan == CF_PLAN_AXIS5Z) {  // x into the workspace, y into dst (z-blocked), z in place
    const unsigned rt = cfgeo::row_threads(n), rT = cfgeo::row_tiles(n);
    const unsigned ct = cfgeo::col_threads(n), cT = cfgeo::col_tiles(n);
    v.push_back({f.row_fwd, rT, rt, src, work, nullptr, 0, 1.0f, "x_fwd", io});
    v.push_back({f.coly_fwd_zb, cT, ct, work, dst, nullptr, 0, 1.0f, "y_fwd", io});
    v.push_back({f.colz_fwd_zb, cT, ct, dst, dst, nullptr, 0, 1.0f, "z_fwd", io});
    return;
  }
  if (plan == CF_PLAN_PLANE3) {  // x and y in one pass per plane, then z
    const unsigned pt = cfgeo::plane_threads(n), pT = cfgeo::plane_tiles(n);
    const unsigned ct = cfgeo::col_threads(n), cT = cfgeo::col_tiles(n);
    v.push_back({f.plane_fwd, pT, pt, src, dst, nullptr, 0, 1.0f, "xy_fwd", io});
    v.push_back({f.colz_fwd, cT, ct, dst, dst, nullptr, 0, 1.0f, "z_fwd", io});
    return;
  }
  if (!is_split(plan)) {
    fwd_passes(f, n, count, src, dst, v);
    return;
  }
  const int vi = split_variant(plan);
  const unsigned at = cfgeo::split_a_threads((int)n, vi), aT = cfgeo::split_a_tiles((int)n, vi);
  const unsigned bt = cfgeo::split_b_threads((int)n, vi), bT = cfgeo::split_b_tiles((int)n, vi);
  v.push_back({f.xy1_fwd[vi], aT, at, src, work, nullptr, 0, 1.0f, "xy1_fwd", io});
  v.push_back({f.mspec[vi], bT, bt, work, dst, nullptr, 0, 1.0f, "y2z_fwd", io});
}

int run_passes(const PassList& v, unsigned n, unsigned nb, hipStream_t s, const void* tw,
               hipEvent_t* ev_before /* nullable, v.size()+1 events */) {
  if (ev_before) HIPCHK(hipEventRecord(ev_before[0], s));
  for (size_t i = 0; i < v.size(); ++i) {
    CFCHK(launch_pass(v[i], n, nb, s, tw));
    if (ev_before) HIPCHK(hipEventRecord(ev_before[i + 1], s));
  }
  return CF_OK;
}

int ensure_device(void** p, size_t* have, size_t need) {
  if (*have >= need) return CF_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  hipError_t e = hipMalloc(p, need);
  if (e != hipSuccess) return fail(CF_ERR_NOMEM, "hipMalloc(%zu): %s", need, hipGetErrorString(e));
  *have = need;
  return CF_OK;
}

int ensure_pinned(void** p, size_t* have, size_t need) {
  if (*have >= need) return CF_OK;
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  *have = 0;
  hipError_t e = hipHostMalloc(p, need, hipHostMallocMapped);
  if (e != hipSuccess) return fail(CF_ERR_NOMEM, "hipHostMalloc(%zu): %s", need, hipGetErrorString(e));
  *have = need;
  return CF_OK;
}

// ---- host transfers of the copying drop-in call ------------------------------
// Pageable host memory (what the reference's callers pass) moves at about half
// the bus rate through the runtime's own path (measured 29 GB/s H2D, 18 GB/s
// D2H against 57 GB/s from pinned memory).  The call stages it through two
// pinned chunk buffers instead: the host copy of one chunk overlaps the DMA
// of the other, and the host copies use a few threads.
constexpr size_t kStageChunk = size_t(8) << 20;

// threads per host copy: 4, or $CONVFPGA_COPY_THREADS (1..32)
int copy_threads() {
  static const int t = [] {
    const char* e = getenv("CONVFPGA_COPY_THREADS");
    const int v = e ? atoi(e) : 0;
    return v >= 1 && v <= 32 ? v : 4;
  }();
  return t;
}

void host_copy(void* dst, const void* src, size_t n) {
  const int kThreads = copy_threads();
  if (n < (size_t(1) << 20)) {
    memcpy(dst, src, n);
    return;
  }
  const size_t part = ((n + kThreads - 1) / kThreads + 63) & ~size_t(63);
#pragma omp parallel for num_threads(kThreads) schedule(static)
  for (int i = 0; i < kThreads; ++i) {
    const size_t o = (size_t)i * part;
    if (o < n) memcpy((char*)dst + o, (const char*)src + o, std::min(part, n - o));
  }
}

int stage_ready() {
  for (int b = 0; b < 2; ++b) {
    if (!g.stage[b]) {
      const hipError_t e = hipHostMalloc(&g.stage[b], kStageChunk, hipHostMallocDefault);
      if (e != hipSuccess) return fail(CF_ERR_NOMEM, "hipHostMalloc(%zu): %s", kStageChunk, hipGetErrorString(e));
    }
    if (!g.stage_ev[b]) HIPCHK(hipEventCreateWithFlags(&g.stage_ev[b], hipEventDisableTiming));
  }
  return CF_OK;
}

// waits until the last transfer that used staging buffer b has finished
int stage_wait(int b) {
  if (g.stage_used[b]) HIPCHK(hipEventSynchronize(g.stage_ev[b]));
  g.stage_used[b] = false;
  return CF_OK;
}

// host -> device through the staging buffers; returns once the last DMA has been queued on s
int staged_h2d(void* d, const void* h, size_t bytes, hipStream_t s) {
  CFCHK(stage_ready());
  for (size_t off = 0, k = 0; off < bytes; off += kStageChunk, ++k) {
    const int b = (int)(k & 1);
    const size_t n = std::min(kStageChunk, bytes - off);
    CFCHK(stage_wait(b));
    host_copy(g.stage[b], (const char*)h + off, n);
    HIPCHK(hipMemcpyAsync((char*)d + off, g.stage[b], n, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(g.stage_ev[b], s));
    g.stage_used[b] = true;
  }
  return CF_OK;
}

// device -> host through the staging buffers, after the work queued on s; blocks until done
int staged_d2h(void* h, const void* d, size_t bytes, hipStream_t s) {
  CFCHK(stage_ready());
  const size_t nchunk = (bytes + kStageChunk - 1) / kStageChunk;
  auto issue = [&](size_t k) -> int {
    const int b = (int)(k & 1);
    CFCHK(stage_wait(b));
    HIPCHK(hipMemcpyAsync(g.stage[b], (const char*)d + k * kStageChunk, std::min(kStageChunk, bytes - k * kStageChunk),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(g.stage_ev[b], s));
    g.stage_used[b] = true;
    return CF_OK;
  };
  if (nchunk) CFCHK(issue(0));
  for (size_t k = 0; k < nchunk; ++k) {
    if (k + 1 < nchunk) CFCHK(issue(k + 1));
    const int b = (int)(k & 1);
    CFCHK(stage_wait(b));
    host_copy((char*)h + k * kStageChunk, g.stage[b], std::min(kStageChunk, bytes - k * kStageChunk));
  }
  return CF_OK;
}

double host_ms(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// Workspace of `bytes` for a call on stream s.  Calls on different streams are
// ordered through work_ev so that they never overlap on the shared workspace.
int acquire_work(size_t bytes, hipStream_t s, void** out) {
  if (!g.work_ev) HIPCHK(hipEventCreateWithFlags(&g.work_ev, hipEventDisableTiming));
  if (g.d_work_bytes < bytes) {
    if (g.d_work) {
      HIPCHK(hipDeviceSynchronize());  // earlier users on any stream are done
      (void)hipFree(g.d_work);
    }
    g.d_work = nullptr;
    g.d_work_bytes = 0;
    g.work_used = false;
    hipError_t e = hipMalloc(&g.d_work, bytes);
    if (e != hipSuccess) return fail(CF_ERR_NOMEM, "hipMalloc(%zu) for the workspace: %s", bytes, hipGetErrorString(e));
    g.d_work_bytes = bytes;
  }
  if (g.work_used && g.work_stream != s) {
    // Everything submitted so far to the previous user's stream includes its
    // use of the workspace, so an event recorded there now orders s after it.
    // (Recording lazily, only on a change of stream, keeps event packets --
    // each a gap of several microseconds on the GPU -- out of the
    // single-stream sequence.)  A stream destroyed meanwhile has finished its
    // work; should the record fail, wait for the whole device instead.
    if (hipEventRecord(g.work_ev, g.work_stream) == hipSuccess) {
      HIPCHK(hipStreamWaitEvent(s, g.work_ev, 0));
    } else {
      (void)hipGetLastError();
      HIPCHK(hipDeviceSynchronize());
    }
  }
  *out = g.d_work;
  return CF_OK;
}

int release_work(hipStream_t s) {
  g.work_stream = s;
  g.work_used = true;
  return CF_OK;
}

int ensure_events() {
  for (auto& e : g.ev)
    if (!e) HIPCHK(hipEventCreate(&e));  // system-scope fence: host reads (SVM output) follow these
  return CF_OK;
}

double ms_between(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
  return (double)ms;
}

void* mapped(void* host) {
  void* d = nullptr;
  if (hipHostGetDevicePointer(&d, host, 0) != hipSuccess) return host;
  return d;
}

// Times every supported plan on one scratch cube (two rounds of 2 warm-up and
// 5 timed convolutions each, HIP events on the library stream) and keeps the
// fastest.
int autotune(unsigned n, SizeFns* f) {
  std::vector<int> cand;
  for (int p = 1; p < CF_NUM_PLANS; ++p)
    if (plan_ok(n, p)) cand.push_back(p);
  for (auto& m : f->tune_ms) m = 0.0f;
  if (cand.size() == 1) {
    f->tuned = cand[0];
    return CF_OK;
  }
  const size_t V = cube_bytes(n);
  void *sig = nullptr, *out = nullptr, *H = nullptr;
  void* work = nullptr;
  hipStream_t s = g.stream;
  int rc = CF_OK;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    if (sig) (void)hipFree(sig);
    if (out) (void)hipFree(out);
    if (H) (void)hipFree(H);
  };
  if (hipMalloc(&sig, V) != hipSuccess || hipMalloc(&out, V) != hipSuccess || hipMalloc(&H, V) != hipSuccess) {
    cleanup();
    return fail(CF_ERR_NOMEM, "autotune scratch allocation failed");
  }
  size_t wneed = 0;  // every candidate runs out of place on the scratch cubes
  for (int p : cand) wneed = std::max(wneed, work_bytes(p, n, 1, sig, out));
  rc = acquire_work(std::max(wneed, V), s, &work);
  long long nf = (long long)(V / 4);
  unsigned long long seed = 12345;
  void* fa[] = {&sig, &nf, &seed};
  void* fb[] = {&H, &nf, &seed};
  if (rc == CF_OK && (hipModuleLaunchKernel(g.fillu, 1024, 1, 1, 256, 1, 1, 0, s, fa, nullptr) != hipSuccess ||
                      hipModuleLaunchKernel(g.fillu, 1024, 1, 1, 256, 1, 1, 0, s, fb, nullptr) != hipSuccess))
    rc = fail(CF_ERR_RUNTIME, "autotune fill failed");
  // two interleaved rounds over the candidates, each plan keeping its better
  // round: a plan is not lost to one disturbed measurement
  for (int round = 0; round < 2 && rc == CF_OK; ++round) {
    for (int p : cand) {
      if (rc != CF_OK) break;
      PassList v;
      conv_passes(*f, p, n, 1, sig, out, work, H, 0, 1.0f, v);
      for (int w = 0; w < 2 && rc == CF_OK; ++w) rc = run_passes(v, n, 1, s, f->tw, nullptr);
      if (rc == CF_OK && hipEventRecord(g.ev[0], s) != hipSuccess) rc = fail(CF_ERR_RUNTIME, "event");
      for (int r = 0; r < 5 && rc == CF_OK; ++r) rc = run_passes(v, n, 1, s, f->tw, nullptr);
      if (rc == CF_OK && (hipEventRecord(g.ev[1], s) != hipSuccess || hipEventSynchronize(g.ev[1]) != hipSuccess))
        rc = fail(CF_ERR_RUNTIME, "event");
      if (rc != CF_OK) break;
      const float ms = (float)(ms_between(g.ev[0], g.ev[1]) / 5.0);
      if (round == 0 || ms < f->tune_ms[p]) f->tune_ms[p] = ms;
    }
  }
  float best = 0.0f;
  int best_plan = cand[0];
  for (int p : cand) {
    if (best == 0.0f || f->tune_ms[p] < best) {
      best = f->tune_ms[p];
      best_plan = p;
    }
  }
  // The plans within 3% of the best are timed again, three interleaved rounds
  // of 10 convolutions each, and ranked by their mean: near-equal plans are
  // not decided by one short measurement.
  std::vector<int> close;
  for (int p : cand)
    if (f->tune_ms[p] <= best * 1.03f) close.push_back(p);
  if (rc == CF_OK && close.size() > 1) {
    float acc[CF_NUM_PLANS] = {};
    for (int round = 0; round < 3 && rc == CF_OK; ++round) {
      for (int p : close) {
        if (rc != CF_OK) break;
        PassList v;
        conv_passes(*f, p, n, 1, sig, out, work, H, 0, 1.0f, v);
        for (int w = 0; w < 2 && rc == CF_OK; ++w) rc = run_passes(v, n, 1, s, f->tw, nullptr);
        if (rc == CF_OK && hipEventRecord(g.ev[0], s) != hipSuccess) rc = fail(CF_ERR_RUNTIME, "event");
        for (int r = 0; r < 10 && rc == CF_OK; ++r) rc = run_passes(v, n, 1, s, f->tw, nullptr);
        if (rc == CF_OK && (hipEventRecord(g.ev[1], s) != hipSuccess || hipEventSynchronize(g.ev[1]) != hipSuccess))
          rc = fail(CF_ERR_RUNTIME, "event");
        if (rc == CF_OK) acc[p] += (float)(ms_between(g.ev[0], g.ev[1]) / 10.0);
      }
    }
    if (rc == CF_OK) {
      best = 0.0f;
      for (int p : close) {
        f->tune_ms[p] = acc[p] / 3.0f;
        if (best == 0.0f || f->tune_ms[p] < best) {
          best = f->tune_ms[p];
          best_plan = p;
        }
      }
    }
  }
  if (rc == CF_OK) rc = release_work(s);
  cleanup();
  if (rc != CF_OK) return rc;
  f->tuned = best_plan;
  write_tune_file(n, *f);
  return CF_OK;
}

int choose_plan(unsigned n, SizeFns* f, int* plan) {
  if (f->forced != CF_PLAN_AUTO) {
    *plan = f->forced;
    return CF_OK;
  }
  if (f->tuned == CF_PLAN_AUTO) CFCHK(autotune(n, f));
  *plan = f->tuned;
  return CF_OK;
}

// ---- host-resident batches: transfers overlapped with the passes -------------
// Cubes move in groups of whole cubes (about kPipeGroupBytes each) through
// kPipeSlots device slots.  The calling thread copies group k from pageable
// memory into two pinned chunk buffers and uploads it on s_in, then queues the
// group's passes (in place in its slot) on the compute stream; a second host
// thread downloads finished groups on s_out through two more pinned chunk
// buffers and copies them out.  Upload of group k+1, the passes of group k and
// the download of group k-1 therefore run at the same time, the idiomatic
// analogue of the reference's pairwise pipelining of consecutive instances
// (convfpga_svm_batch.c:241-662).
//
// Ordering, all on the device through events:
//   slot s is refilled (group k+S) only after the download of group k has
//   read it (ev_freed[s], recorded on s_out);
//   the passes of group k wait for its upload (ev_loaded[s], on s_in);
//   the download of group k waits for its passes (ev_done[s], compute stream).
// Each event is re-recorded only after every wait on its previous record has
// been queued: the two host threads hand over through `computed` (groups
// whose passes are queued) and `drained` (groups whose downloads are queued).
constexpr int kPipeSlots = 3;
constexpr size_t kPipeGroupBytes = size_t(16) << 20;

struct PipeState {
  hipStream_t s_in = nullptr, s_out = nullptr;
  void* stage_in[2] = {nullptr, nullptr};
  void* stage_out[2] = {nullptr, nullptr};
  hipEvent_t ev_in[2] = {}, ev_out[2] = {};
  hipEvent_t ev_loaded[kPipeSlots] = {}, ev_done[kPipeSlots] = {}, ev_freed[kPipeSlots] = {};
  void* slots = nullptr;
  size_t slots_bytes = 0;
};
PipeState gp;

void pipe_release() {
  if (gp.s_in) (void)hipStreamSynchronize(gp.s_in);
  if (gp.s_out) (void)hipStreamSynchronize(gp.s_out);
  for (int b = 0; b < 2; ++b) {
    if (gp.stage_in[b]) (void)hipHostFree(gp.stage_in[b]);
    if (gp.stage_out[b]) (void)hipHostFree(gp.stage_out[b]);
    if (gp.ev_in[b]) (void)hipEventDestroy(gp.ev_in[b]);
    if (gp.ev_out[b]) (void)hipEventDestroy(gp.ev_out[b]);
  }
  for (int i = 0; i < kPipeSlots; ++i) {
    if (gp.ev_loaded[i]) (void)hipEventDestroy(gp.ev_loaded[i]);
    if (gp.ev_done[i]) (void)hipEventDestroy(gp.ev_done[i]);
    if (gp.ev_freed[i]) (void)hipEventDestroy(gp.ev_freed[i]);
  }
  if (gp.slots) (void)hipFree(gp.slots);
  if (gp.s_in) (void)hipStreamDestroy(gp.s_in);
  if (gp.s_out) (void)hipStreamDestroy(gp.s_out);
  gp = PipeState();
}

int pipe_ready(size_t slot_bytes) {
  if (!gp.s_in) HIPCHK(hipStreamCreateWithFlags(&gp.s_in, hipStreamNonBlocking));
  if (!gp.s_out) HIPCHK(hipStreamCreateWithFlags(&gp.s_out, hipStreamNonBlocking));
  for (int b = 0; b < 2; ++b) {
    if (!gp.stage_in[b]) HIPCHK(hipHostMalloc(&gp.stage_in[b], kStageChunk, hipHostMallocDefault));
    if (!gp.stage_out[b]) HIPCHK(hipHostMalloc(&gp.stage_out[b], kStageChunk, hipHostMallocDefault));
    if (!gp.ev_in[b]) HIPCHK(hipEventCreateWithFlags(&gp.ev_in[b], hipEventDisableTiming));
    if (!gp.ev_out[b]) HIPCHK(hipEventCreateWithFlags(&gp.ev_out[b], hipEventDisableTiming));
  }
  for (int i = 0; i < kPipeSlots; ++i) {
    if (!gp.ev_loaded[i]) HIPCHK(hipEventCreateWithFlags(&gp.ev_loaded[i], hipEventDisableTiming));
    if (!gp.ev_done[i]) HIPCHK(hipEventCreateWithFlags(&gp.ev_done[i], hipEventDisableTiming));
    if (!gp.ev_freed[i]) HIPCHK(hipEventCreateWithFlags(&gp.ev_freed[i], hipEventDisableTiming));
  }
  const size_t need = slot_bytes * kPipeSlots;
  if (gp.slots_bytes < need) {
    if (gp.slots) {
      HIPCHK(hipDeviceSynchronize());
      (void)hipFree(gp.slots);
    }
    gp.slots = nullptr;
    gp.slots_bytes = 0;
    const hipError_t e = hipMalloc(&gp.slots, need);
    if (e != hipSuccess) return fail(CF_ERR_NOMEM, "hipMalloc(%zu) for the batch slots: %s", need, hipGetErrorString(e));
    gp.slots_bytes = need;
  }
  return CF_OK;
}

struct PipeJob {
  const SizeFns* f;
  int plan, precision;
  unsigned n, how_many, group;  // cubes per group (the last one may hold fewer)
  const void* H;                // spectrum; hstride 0 (shared) or N^3 (one per cube)
  long long hstride;
  float scale;
  const char* sig;
  char* out;
  hipStream_t s;  // compute stream
};

struct PipeSync {
  std::mutex mu;
  std::condition_variable cv;
  unsigned computed = 0, drained = 0;
  bool abort = false;
  int rc = CF_OK;
  std::string err;
};

// download side: runs on its own host thread
void pipe_drain(const PipeJob& j, PipeSync& ps) {
  auto stop = [&](int rc) {
    std::lock_guard<std::mutex> lk(ps.mu);
    if (ps.rc == CF_OK) {
      ps.rc = rc;
      ps.err = g_err;  // thread-local in this thread
    }
    ps.abort = true;
    ps.cv.notify_all();
  };
  if (hipSetDevice(g.device) != hipSuccess) return stop(fail(CF_ERR_RUNTIME, "hipSetDevice failed"));
  const size_t V = cube_bytes(j.n);
  const unsigned ngroups = (j.how_many + j.group - 1) / j.group;
  struct Chunk {
    size_t dst_off, bytes;
    int b;
  };
  Chunk pend{0, 0, -1};  // downloaded chunk not yet copied out
  unsigned nchunk = 0;
  auto finish = [&](const Chunk& c) -> int {
    HIPCHK(hipEventSynchronize(gp.ev_out[c.b]));
    host_copy(j.out + c.dst_off, gp.stage_out[c.b], c.bytes);
    return CF_OK;
  };
  for (unsigned k = 0; k < ngroups; ++k) {
    {
      std::unique_lock<std::mutex> lk(ps.mu);
      ps.cv.wait(lk, [&] { return ps.abort || ps.computed > k; });
      if (ps.abort) return;
    }
    const int sl = (int)(k % kPipeSlots);
    const size_t k0 = (size_t)k * j.group;
    const size_t bytes = std::min<size_t>(j.group, j.how_many - k0) * V;
    const char* src = (const char*)gp.slots + sl * (size_t)j.group * V;
    if (hipStreamWaitEvent(gp.s_out, gp.ev_done[sl], 0) != hipSuccess)
      return stop(fail(CF_ERR_RUNTIME, "hipStreamWaitEvent (download)"));
    for (size_t off = 0; off < bytes; off += kStageChunk) {
      const Chunk c{k0 * V + off, std::min(kStageChunk, bytes - off), (int)(nchunk++ & 1)};
      // buffer c.b last held the chunk before `pend`, already copied out
      if (hipMemcpyAsync(gp.stage_out[c.b], src + off, c.bytes, hipMemcpyDeviceToHost, gp.s_out) != hipSuccess ||
          hipEventRecord(gp.ev_out[c.b], gp.s_out) != hipSuccess)
        return stop(fail(CF_ERR_RUNTIME, "download of group %u failed", k));
      if (off + kStageChunk >= bytes) {  // slot sl fully read once this is done
        if (hipEventRecord(gp.ev_freed[sl], gp.s_out) != hipSuccess)
          return stop(fail(CF_ERR_RUNTIME, "hipEventRecord (slot freed)"));
        std::lock_guard<std::mutex> lk(ps.mu);
        ps.drained = k + 1;
        ps.cv.notify_all();
      }
      if (pend.b >= 0) {
        const int rc = finish(pend);
        if (rc != CF_OK) return stop(rc);
      }
      pend = c;
    }
  }
  if (pend.b >= 0) {
    const int rc = finish(pend);
    if (rc != CF_OK) return stop(rc);
  }
}

// upload side and passes: the calling thread.  Returns the summed span of the
// groups' passes (host clock of the compute events is not needed: spans come
// from events recorded around each group).
int pipe_feed(const PipeJob& j, PipeSync& ps, std::vector<hipEvent_t>& tev) {
  const size_t V = cube_bytes(j.n);
  const unsigned ngroups = (j.how_many + j.group - 1) / j.group;
  const size_t slot_bytes = (size_t)j.group * V;
  unsigned nchunk = 0;
  bool used[2] = {false, false};
  auto stop = [&](int rc) {
    std::lock_guard<std::mutex> lk(ps.mu);
    if (ps.rc == CF_OK) {
      ps.rc = rc;
      ps.err = g_err;
    }
    ps.abort = true;
    ps.cv.notify_all();
    return rc;
  };
  for (unsigned k = 0; k < ngroups; ++k) {
    const int sl = (int)(k % kPipeSlots);
    const size_t k0 = (size_t)k * j.group;
    const unsigned nb = (unsigned)std::min<size_t>(j.group, j.how_many - k0);
    const size_t bytes = nb * V;
    char* slot = (char*)gp.slots + sl * slot_bytes;
    if (k >= (unsigned)kPipeSlots) {  // the download of group k-S has to be queued before its event is waited on
      std::unique_lock<std::mutex> lk(ps.mu);
      ps.cv.wait(lk, [&] { return ps.abort || ps.drained > k - kPipeSlots; });
      if (ps.abort) return ps.rc;
      lk.unlock();
      if (hipStreamWaitEvent(gp.s_in, gp.ev_freed[sl], 0) != hipSuccess)
        return stop(fail(CF_ERR_RUNTIME, "hipStreamWaitEvent (slot reuse)"));
    }
    for (size_t off = 0; off < bytes; off += kStageChunk) {
      const int b = (int)(nchunk++ & 1);
      const size_t cb = std::min(kStageChunk, bytes - off);
      if (used[b] && hipEventSynchronize(gp.ev_in[b]) != hipSuccess)
        return stop(fail(CF_ERR_RUNTIME, "upload of group %u failed", k));
      host_copy(gp.stage_in[b], j.sig + k0 * V + off, cb);
      if (hipMemcpyAsync(slot + off, gp.stage_in[b], cb, hipMemcpyHostToDevice, gp.s_in) != hipSuccess ||
          hipEventRecord(gp.ev_in[b], gp.s_in) != hipSuccess)
        return stop(fail(CF_ERR_RUNTIME, "upload of group %u failed", k));
      used[b] = true;
    }
    if (hipEventRecord(gp.ev_loaded[sl], gp.s_in) != hipSuccess ||
        hipStreamWaitEvent(j.s, gp.ev_loaded[sl], 0) != hipSuccess)
      return stop(fail(CF_ERR_RUNTIME, "upload event of group %u", k));
    PassList p;
    const void* H = (const char*)j.H + (size_t)k0 * j.hstride * 8;
    void* work = nullptr;
    if (needs_work(j.plan)) {
      const int rc = acquire_work(work_bytes(j.plan, j.n, j.group, slot, slot), j.s, &work);
      if (rc != CF_OK) return stop(rc);
    }
    conv_passes(*j.f, j.plan, j.n, nb, slot, slot, work, H, j.hstride, j.scale, p, j.precision);
    if (hipEventRecord(tev[2 * k], j.s) != hipSuccess) return stop(fail(CF_ERR_RUNTIME, "hipEventRecord"));
    int rc = run_passes(p, j.n, nb, j.s, j.f->tw, nullptr);
    if (work) release_work(j.s);
    if (rc != CF_OK) return stop(rc);
    if (hipEventRecord(tev[2 * k + 1], j.s) != hipSuccess || hipEventRecord(gp.ev_done[sl], j.s) != hipSuccess)
      return stop(fail(CF_ERR_RUNTIME, "hipEventRecord (passes done)"));
    std::lock_guard<std::mutex> lk(ps.mu);
    ps.computed = k + 1;
    ps.cv.notify_all();
  }
  return CF_OK;
}

// Runs a host-resident batch through the pipeline; blocks until the last
// output byte is in `out`.  pass_ms: summed span of the groups' passes;
// wall_ms: first upload to last copy-out, host clock.
int pipe_run(PipeJob j, double* pass_ms, double* wall_ms) {
  const size_t V = cube_bytes(j.n);
  j.group = (unsigned)std::max<size_t>(1, std::min<size_t>(j.how_many, kPipeGroupBytes / V));
  const unsigned ngroups = (j.how_many + j.group - 1) / j.group;
  CFCHK(pipe_ready((size_t)j.group * V));
  std::vector<hipEvent_t> tev(2 * (size_t)ngroups, nullptr);
  struct Events {
    std::vector<hipEvent_t>& v;
    ~Events() {
      for (auto e : v)
        if (e) (void)hipEventDestroy(e);
    }
  } hold{tev};
  for (auto& e : tev) HIPCHK(hipEventCreate(&e));
  // earlier work on the compute stream (e.g. the filter transform) stays first
  const auto t0 = std::chrono::steady_clock::now();
  PipeSync ps;
  std::thread drain(pipe_drain, std::cref(j), std::ref(ps));
  pipe_feed(j, ps, tev);
  drain.join();
  const auto t1 = std::chrono::steady_clock::now();
  const hipError_t e1 = hipStreamSynchronize(j.s), e2 = hipStreamSynchronize(gp.s_in),
                   e3 = hipStreamSynchronize(gp.s_out);
  if (ps.rc != CF_OK) {
    g_err = ps.err;
    return ps.rc;
  }
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
    return fail(CF_ERR_RUNTIME, "host batch pipeline: %s",
                hipGetErrorString(e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3));
  double sum = 0.0;
  for (unsigned k = 0; k < ngroups; ++k) sum += ms_between(tev[2 * k], tev[2 * k + 1]);
  if (pass_ms) *pass_ms = sum;
  if (wall_ms) *wall_ms = host_ms(t0, t1);
  return CF_OK;
}

}  // namespace

// =============================================================================
// drop-in back end
// =============================================================================
extern "C" int cfe_initialize(const char* platform_name, const char* path, int use_svm) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (g.ready || g.stream || g.mod) release_all();
  // same order of checks as the reference (fpga_init.c:65-112)
  if (path == nullptr || path[0] == '\0') return -1;
  if (platform_name == nullptr || platform_name[0] == '\0') return -2;
  int ndev = 0;
  const bool have_rt = hipGetDeviceCount(&ndev) == hipSuccess;
  if (!have_rt) ndev = 0;
  std::vector<std::string> names = {"AMD HIP", "ROCm"};
  hipDeviceProp_t prop;
  if (ndev > 0 && hipGetDeviceProperties(&prop, 0) == hipSuccess) {
    names.push_back(prop.name);
    names.push_back(prop.gcnArchName);
    g.arch = std::string(prop.gcnArchName).substr(0, std::string(prop.gcnArchName).find(':'));
  }
  const std::string want = lower(platform_name);
  bool match = false;
  for (auto& nm : names) match = match || lower(nm).find(want) != std::string::npos;
  if (!match) return -2;
  if (ndev < 1) return -3;
  g.device = 0;
  if (hipSetDevice(g.device) != hipSuccess) return -3;
  if (use_svm) {
    int can_map = 0;
    if (hipDeviceGetAttribute(&can_map, hipDeviceAttributeCanMapHostMemory, g.device) != hipSuccess || !can_map)
      return -5;
    g.svm = true;
  }
  if (hipStreamCreate(&g.stream) != hipSuccess) {
    release_all();
    return -3;
  }
  if (hipModuleLoad(&g.mod, path) != hipSuccess) {
    g.mod = nullptr;
    fprintf(stderr, "convfpga: cannot load code object '%s'\n", path);
    release_all();
    return -4;
  }
  if (get_fn(&g.twgen, "cf_twiddle_gen") != CF_OK || get_fn(&g.dft16gen, "cf_dft16_table_gen") != CF_OK || get_fn(&g.errp, "cf_errnorm_partial") != CF_OK ||
      get_fn(&g.errf, "cf_errnorm_final") != CF_OK || get_fn(&g.fillu, "cf_fill_uniform") != CF_OK ||
      get_fn(&g.filll, "cf_fill_loguniform") != CF_OK ||
      get_fn(&g.batch_run, "cf_batch_run") != CF_OK || get_fn(&g.batch_run_3p, "cf_batch_run_3p") != CF_OK ||
      get_fn(&g.batch_run_3p_small, "cf_batch_run_3p_small") != CF_OK ||
      ensure_events() != CF_OK) {
    fprintf(stderr, "convfpga: '%s' is not a convfpga code object (%s)\n", path, g_err.c_str());
    release_all();
    return -4;
  }
  // The allocation-free entry point's scratch: a device array of the module,
  // allocated by the module load.  Its address on this device comes from the
  // runtime (the host never names the array itself).
  {
    hipDeviceptr_t p = nullptr;
    size_t bytes = 0;
    const hipError_t e = hipModuleGetGlobal(&p, &bytes, g.mod, "cf_scratch");
    if (e != hipSuccess || !p || bytes != (size_t)cfgeo::noalloc_scratch_bytes) {
      fprintf(stderr, "convfpga: '%s' has no scratch array of %llu bytes (%s)\n", path,
              (unsigned long long)cfgeo::noalloc_scratch_bytes, hipGetErrorString(e));
      release_all();
      return -4;
    }
    g.scratch = (void*)p;
    g.scratch_bytes = bytes;
  }
  g.ready = true;
  return 0;
}

extern "C" void cfe_final(void) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  release_all();
}

extern "C" int cfe_conv_host(unsigned n, const void* sig, const void* filter, void* out, unsigned how_many, int mode,
                             cfe_times_t* t) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!sig || !filter || !out || how_many == 0 || !t) return fail(CF_ERR_ARG, "NULL pointer or zero count");
  if (mode != CFE_MODE_COPY && mode != CFE_MODE_SVM && mode != CFE_MODE_SVM_BATCH)
    return fail(CF_ERR_ARG, "unknown mode %d", mode);
  CFCHK(require_ready());
  if (mode != CFE_MODE_COPY && !g.svm) return fail(CF_ERR_STATE, "fpga_initialize() was called with use_svm = false");
  SizeFns* f = nullptr;
  CFCHK(size_res(n, &f));
  int plan = 0;
  CFCHK(choose_plan(n, f, &plan));
  const size_t V = cube_bytes(n);
  const size_t SV = V * how_many;
  memset(t, 0, sizeof *t);
  hipEvent_t* ev = g.ev;
  hipStream_t s = g.stream;
  CFCHK(ensure_device(&g.d_filt, &g.d_filt_bytes, V));
  PassList filt_p, conv_p;
  if (mode == CFE_MODE_SVM_BATCH) {
    // filter upload and transform, then the cubes stream through the
    // overlapped pipeline; times as the reference's batch call reports them
    // (convfpga_svm_batch.c:664-665): summed pass spans and the wall span
    void* fwork = nullptr;
    if (needs_work(plan)) CFCHK(acquire_work(V, s, &fwork));
    CFCHK(staged_h2d(g.d_filt, filter, V, s));
    spectrum_passes(*f, plan, n, 1, g.d_filt, g.d_filt, fwork, filt_p);
    const size_t nf = filt_p.size();
    CFCHK(ensure_events());
    const int rc = run_passes(filt_p, n, 1, s, f->tw, ev);  // ev[0 .. nf]
    if (fwork) release_work(s);
    CFCHK(rc);
    PipeJob j{f, plan, CF_PREC_FP32, n, how_many, 1, g.d_filt, 0, 1.0f, (const char*)sig, (char*)out, s};
    double pass_ms = 0.0, wall_ms = 0.0;
    CFCHK(pipe_run(j, &pass_ms, &wall_ms));
    t->filter_exec_t = ms_between(ev[0], ev[nf]);
    t->sig_exec_t = pass_ms;
    t->siginv_exec_t = wall_ms;
    return CF_OK;
  }
  CFCHK(ensure_device(&g.d_data, &g.d_data_bytes, SV));
  void* work = nullptr;
  // the copying call runs in place in d_data; the _svm call reads the mapped
  // signal and writes d_data, then the last pass writes the mapped output
  if (needs_work(plan))
    CFCHK(acquire_work(std::max(V, work_bytes(plan, n, how_many, g.d_data, mode == CFE_MODE_COPY ? g.d_data : nullptr)),
                       s, &work));

  if (mode == CFE_MODE_COPY) {
    // Blocking phases as in the reference (convfpga.c:75-375); each host
    // transfer goes through the pinned staging buffers and is timed on the
    // host clock (it includes the host copies), the passes with events.
    using clk = std::chrono::steady_clock;
    HIPCHK(hipStreamSynchronize(s));  // earlier work on the stream is not part of the timing
    // step 1-2: filter H2D and transform (convfpga.c:75-172)
    auto t0 = clk::now();
    CFCHK(staged_h2d(g.d_filt, filter, V, s));
    HIPCHK(hipStreamSynchronize(s));
    auto t1 = clk::now();
    spectrum_passes(*f, plan, n, 1, g.d_filt, g.d_filt, work, filt_p);
    const size_t nf = filt_p.size();
    CFCHK(run_passes(filt_p, n, 1, s, f->tw, ev));  // ev[0 .. nf]
    // step 3: signal H2D, forward + product (convfpga.c:174-279)
    auto t2 = clk::now();
    CFCHK(staged_h2d(g.d_data, sig, SV, s));
    HIPCHK(hipStreamSynchronize(s));
    auto t3 = clk::now();
    const size_t split = conv_passes(*f, plan, n, how_many, g.d_data, g.d_data, work, g.d_filt, 0, 1.0f, conv_p);
    const size_t nc = conv_p.size();
    hipEvent_t* ec = ev + 1 + nf;
    CFCHK(run_passes(conv_p, n, how_many, s, f->tw, ec));  // ec[0 .. nc]
    if (work) CFCHK(release_work(s));
    HIPCHK(hipStreamSynchronize(s));
    // step 4 ran in the last passes; D2H (convfpga.c:365-375)
    auto t4 = clk::now();
    CFCHK(staged_d2h(out, g.d_data, SV, s));
    auto t5 = clk::now();
    t->filter_pcie_wr_t = host_ms(t0, t1);
    t->filter_exec_t = ms_between(ev[0], ev[nf]);
    t->sig_pcie_wr_t = host_ms(t2, t3);
    t->sig_exec_t = ms_between(ec[0], ec[split]);
    t->siginv_exec_t = ms_between(ec[split], ec[nc]);
    t->sig_pcie_rd_t = host_ms(t4, t5);
    return CF_OK;
  }
  CFCHK(ensure_pinned(&g.h_filt, &g.h_filt_bytes, V));
  CFCHK(ensure_pinned(&g.h_sig, &g.h_sig_bytes, SV));
  CFCHK(ensure_pinned(&g.h_out, &g.h_out_bytes, SV));
  // host copies into the mapped buffers (the reference's SVM map + memcpy,
  // convfpga_svm.c:86-103); the first and last passes then use them over the bus.
  host_copy(g.h_filt, filter, V);
  host_copy(g.h_sig, sig, SV);
  spectrum_passes(*f, plan, n, 1, mapped(g.h_filt), g.d_filt, work, filt_p);
  const size_t nf = filt_p.size();
  CFCHK(run_passes(filt_p, n, 1, s, f->tw, ev));  // ev[0 .. nf]
  const size_t split = conv_passes(*f, plan, n, how_many, mapped(g.h_sig), g.d_data, work, g.d_filt, 0, 1.0f, conv_p);
  conv_p.back().out = mapped(g.h_out);  // the last pass writes host memory directly
  const size_t nc = conv_p.size();
  hipEvent_t* ec = ev + nf + 1;
  CFCHK(run_passes(conv_p, n, how_many, s, f->tw, ec));  // ec[0 .. nc]
  if (work) CFCHK(release_work(s));
  HIPCHK(hipStreamSynchronize(s));
  host_copy(out, g.h_out, SV);
  t->filter_exec_t = ms_between(ev[0], ev[nf]);
  if (how_many == 1) {
    t->sig_exec_t = ms_between(ec[0], ec[split]);
    t->siginv_exec_t = ms_between(ec[split], ec[nc]);
  } else {
    // batch: summed stage time and first-to-last wall span (convfpga_svm_batch.c:664-665)
    t->sig_exec_t = ms_between(ec[0], ec[split]) + ms_between(ec[split], ec[nc]);
    t->siginv_exec_t = ms_between(ec[0], ec[nc]);
  }
  return CF_OK;
}

// =============================================================================
// extension API (convfpga_ext.h) -- the only symbols of this file exported
// from the shared library (it is built with -fvisibility=hidden)
// =============================================================================
#pragma GCC visibility push(default)
struct cf_spectrum_s {
  unsigned n = 0, count = 0;
  int plan = CF_PLAN_AXIS5;
  void* d = nullptr;
  bool view = false;  // aliases one filter of another handle's memory
};

extern "C" const char* cf_last_error(void) { return g_err.c_str(); }

extern "C" int cf_supported_size(unsigned n) { return cfgeo::supported((int)n) ? 1 : 0; }

extern "C" int cf_spectrum_create(unsigned n, unsigned count, const void* filters, int on_device, void* stream,
                                  cf_spectrum_t* out) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!filters || !out || count == 0) return fail(CF_ERR_ARG, "NULL pointer or zero count");
  *out = nullptr;
  CFCHK(require_ready());
  SizeFns* f = nullptr;
  CFCHK(size_res(n, &f));
  hipStream_t s = (hipStream_t)stream;
  auto* h = new cf_spectrum_s;
  h->n = n;
  h->count = count;
  const size_t bytes = cube_bytes(n) * count;
  hipError_t e = hipMalloc(&h->d, bytes);
  if (e != hipSuccess) {
    delete h;
    return fail(CF_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  const void* src = filters;
  if (!on_device) {
    e = hipMemcpyAsync(h->d, filters, bytes, hipMemcpyHostToDevice, s);
    src = h->d;
  }
  int rc = e == hipSuccess ? choose_plan(n, f, &h->plan)
                           : fail(CF_ERR_RUNTIME, "hipMemcpyAsync: %s", hipGetErrorString(e));
  void* work = nullptr;
  if (rc == CF_OK && needs_work(h->plan)) rc = acquire_work(bytes, s, &work);
  if (rc == CF_OK) {
    PassList p;
    spectrum_passes(*f, h->plan, n, count, src, h->d, work, p);
    rc = run_passes(p, n, count, s, f->tw, nullptr);
    if (rc == CF_OK && work) rc = release_work(s);
  }
  if (rc == CF_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(CF_ERR_RUNTIME, "spectrum setup failed");
  if (rc != CF_OK) {
    (void)hipFree(h->d);
    delete h;
    return rc;
  }
  *out = h;
  return CF_OK;
}

extern "C" void cf_spectrum_destroy(cf_spectrum_t h) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!h) return;
  if (h->d && !h->view) {
    (void)hipSetDevice(g.device);
    (void)hipFree(h->d);
  }
  delete h;
}

extern "C" int cf_spectrum_view(cf_spectrum_t parent, unsigned index, cf_spectrum_t* out) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!parent || !out) return fail(CF_ERR_ARG, "NULL pointer");
  if (index >= parent->count) return fail(CF_ERR_ARG, "index %u of a spectrum of %u filters", index, parent->count);
  auto* h = new cf_spectrum_s;
  h->n = parent->n;
  h->count = 1;
  h->plan = parent->plan;
  h->d = (char*)parent->d + (size_t)index * cube_bytes(parent->n);
  h->view = true;
  *out = h;
  return CF_OK;
}

extern "C" const void* cf_spectrum_data(cf_spectrum_t h) { return h ? h->d : nullptr; }
extern "C" int cf_spectrum_plan(cf_spectrum_t h) { return h ? h->plan : CF_ERR_ARG; }

extern "C" int cf_plan_supported(unsigned n, int plan) { return plan_ok(n, plan) ? 1 : 0; }

extern "C" int cf_set_plan(unsigned n, int plan) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (plan != CF_PLAN_AUTO && !plan_ok(n, plan)) return fail(CF_ERR_ARG, "plan %d cannot run N=%u", plan, n);
  CFCHK(require_ready());
  SizeFns* f = nullptr;
  CFCHK(size_res(n, &f));
  f->forced = plan;
  return CF_OK;
}

extern "C" int cf_get_plan(unsigned n) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  CFCHK(require_ready());
  SizeFns* f = nullptr;
  CFCHK(size_res(n, &f));
  int plan = 0;
  CFCHK(choose_plan(n, f, &plan));
  return plan;
}

extern "C" int cf_autotune(unsigned n, float* ms) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  CFCHK(require_ready());
  SizeFns* f = nullptr;
  CFCHK(size_res(n, &f));
  CFCHK(autotune(n, f));
  if (ms)
    for (int p = 0; p < CF_NUM_PLANS; ++p) ms[p] = f->tune_ms[p];
  return f->tuned;
}
extern "C" unsigned cf_spectrum_edge(cf_spectrum_t h) { return h ? h->n : 0; }
extern "C" unsigned cf_spectrum_count(cf_spectrum_t h) { return h ? h->count : 0; }

static int check_conv_args(cf_spectrum_t h, unsigned how_many, int scale_mode, int precision) {
  if (h->count != 1 && h->count != how_many)
    return fail(CF_ERR_ARG, "spectrum holds %u filters; need 1 or how_many (%u)", h->count, how_many);
  if (scale_mode != CF_SCALE_NONE && scale_mode != CF_SCALE_INV_N3) return fail(CF_ERR_ARG, "bad scale_mode");
  if (precision != CF_PREC_FP32 && precision != CF_PREC_SPLIT_FP16)
    return fail(CF_ERR_ARG, "unknown precision %d", precision);
  if (precision == CF_PREC_SPLIT_FP16 &&
      !(h->n == 256 && is_split(h->plan) && cfgeo::split_gm(split_variant(h->plan)) == 2) &&
      !(h->n == 512 && (h->plan == CF_PLAN_AXIS5 || h->plan == CF_PLAN_AXIS5Z)))
    return fail(CF_ERR_ARG,
                "split-FP16 precision runs N = 256 with a spectrum of a four-step plan with middle tiles of 2 x "
                "(split3_g22, g42, g24m2, g28m2 and their o/h forms) and N = 512 with a spectrum of axis5 or "
                "axis5z; this spectrum is N = %u, plan %d",
                h->n, h->plan);
  return CF_OK;
}

static int conv_device_impl(cf_spectrum_t h, const void* d_sig, void* d_out, unsigned how_many, int scale_mode,
                            int precision, void* stream, cf_pass_times_t* times) {
  if (!h || !d_sig || !d_out || how_many == 0) return fail(CF_ERR_ARG, "NULL pointer or zero count");
  CFCHK(check_conv_args(h, how_many, scale_mode, precision));
  CFCHK(require_ready());
  SizeFns* f = nullptr;
  CFCHK(size_res(h->n, &f));
  const unsigned n = h->n;
  const float scale = scale_mode == CF_SCALE_INV_N3 ? (float)(1.0 / (double)cube_elems(n)) : 1.0f;
  const long long hstride = h->count == 1 ? 0 : (long long)cube_elems(n);
  hipStream_t s = (hipStream_t)stream;
  void* work = nullptr;
  if (needs_work(h->plan)) CFCHK(acquire_work(work_bytes(h->plan, n, how_many, d_sig, d_out), s, &work));
  PassList p;
  conv_passes(*f, h->plan, n, how_many, d_sig, d_out, work, h->d, hstride, scale, p, precision);
  struct WorkRelease {  // the workspace is marked busy on s whatever path returns
    void* w;
    hipStream_t s;
    ~WorkRelease() {
      if (w) release_work(s);
    }
  } wr{work, s};
  if (!times) {
    // armed by cf_profile_begin: the last call of every group of prof_every (a
    // call that follows others in the stream, not one that starts from an idle GPU)
    if (g.prof_max > g.prof_calls && ++g.prof_seen % g.prof_every == 0) {
      hipEvent_t* ev = &g.prof_ev[(size_t)g.prof_calls * (CF_MAX_PASSES + 1)];
      CFCHK(run_passes(p, n, how_many, s, f->tw, ev));
      if (g.prof_calls == 0) {
        g.prof_npass = (unsigned)p.size();
        g.prof_names.clear();
        g.prof_bytes.clear();
        for (auto& q : p) {
          g.prof_names.push_back(q.name);
          g.prof_bytes.push_back(q.bytes);
        }
      }
      ++g.prof_calls;
      return CF_OK;
    }
    return run_passes(p, n, how_many, s, f->tw, nullptr);
  }
  CFCHK(ensure_events());
  CFCHK(run_passes(p, n, how_many, s, f->tw, g.ev));
  HIPCHK(hipEventSynchronize(g.ev[p.size()]));
  memset(times, 0, sizeof *times);
  times->npass = (unsigned)p.size();
  for (size_t i = 0; i < p.size(); ++i) {
    times->ms[i] = (float)ms_between(g.ev[i], g.ev[i + 1]);
    snprintf(times->name[i], sizeof times->name[i], "%s", p[i].name);
    times->bytes[i] = p[i].bytes;
  }
  return CF_OK;
}

extern "C" int cf_conv3d_device(cf_spectrum_t h, const void* d_sig, void* d_out, unsigned how_many, int scale_mode,
                                int precision, void* stream) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  return conv_device_impl(h, d_sig, d_out, how_many, scale_mode, precision, stream, nullptr);
}

static_assert(CF_NOALLOC_SCRATCH_BYTES == cfgeo::noalloc_scratch_bytes, "scratch size of the header and the kernels");

extern "C" int cf_conv3d_device_noalloc(cf_spectrum_t h, const void* d_sig, void* d_out, unsigned how_many,
                                        int scale_mode, int precision, void* stream) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!h || !d_sig || !d_out || how_many == 0) return fail(CF_ERR_ARG, "NULL pointer or zero count");
  CFCHK(check_conv_args(h, how_many, scale_mode, precision));
  CFCHK(require_ready());
  // the edge's kernels and twiddles were set up when the spectrum was made;
  // nothing is created here
  auto it = g.sizes.find(h->n);
  if (it == g.sizes.end())
    return fail(CF_ERR_STATE, "N = %u has not been set up in this session: create its spectrum after fpga_initialize()",
                h->n);
  const SizeFns& f = it->second;
  const unsigned n = h->n;
  void* work = nullptr;
  if (needs_work(h->plan)) {
    const size_t need = work_bytes(h->plan, n, how_many, d_sig, d_out);
    if (need > g.scratch_bytes)
      return fail(CF_ERR_SIZE,
                  "plan %d of N = %u needs a %zu-byte intermediate for %u cubes; the allocation-free call has %zu "
                  "bytes of scratch (fewer cubes per call, or cf_conv3d_device)",
                  h->plan, n, need, how_many, g.scratch_bytes);
    work = g.scratch;
  }
  const float scale = scale_mode == CF_SCALE_INV_N3 ? (float)(1.0 / (double)cube_elems(n)) : 1.0f;
  const long long hstride = h->count == 1 ? 0 : (long long)cube_elems(n);
  PassList p;
  conv_passes(f, h->plan, n, how_many, d_sig, d_out, work, h->d, hstride, scale, p, precision);
  return run_passes(p, n, how_many, (hipStream_t)stream, f.tw, nullptr);
}

extern "C" int cf_conv3d_host(cf_spectrum_t h, const void* h_sig, void* h_out, unsigned how_many, int scale_mode,
                              int precision, cf_host_times_t* times) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!h || !h_sig || !h_out || how_many == 0) return fail(CF_ERR_ARG, "NULL pointer or zero count");
  CFCHK(check_conv_args(h, how_many, scale_mode, precision));
  CFCHK(require_ready());
  SizeFns* f = nullptr;
  CFCHK(size_res(h->n, &f));
  const unsigned n = h->n;
  const float scale = scale_mode == CF_SCALE_INV_N3 ? (float)(1.0 / (double)cube_elems(n)) : 1.0f;
  const long long hstride = h->count == 1 ? 0 : (long long)cube_elems(n);
  PipeJob j{f, h->plan, precision, n, how_many, 1, h->d, hstride, scale, (const char*)h_sig, (char*)h_out, g.stream};
  double pass_ms = 0.0, wall_ms = 0.0;
  CFCHK(pipe_run(j, &pass_ms, &wall_ms));
  if (times) {
    times->pass_ms = pass_ms;
    times->wall_ms = wall_ms;
  }
  return CF_OK;
}

extern "C" int cf_conv3d_device_timed(cf_spectrum_t h, const void* d_sig, void* d_out, unsigned how_many,
                                      int scale_mode, int precision, void* stream, cf_pass_times_t* times) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!times) return fail(CF_ERR_ARG, "times is NULL");
  return conv_device_impl(h, d_sig, d_out, how_many, scale_mode, precision, stream, times);
}

extern "C" int cf_conv3d_share_timed(cf_spectrum_t h, const void* d_sig, void* d_out, unsigned shares,
                                     int precision, void* stream, cf_pass_times_t* times) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!h || !d_sig || !d_out || !times) return fail(CF_ERR_ARG, "NULL pointer");
  if (shares == 0 || shares > 16 || (shares & (shares - 1)))
    return fail(CF_ERR_ARG, "shares must be 1, 2, 4, 8 or 16, not %u", shares);
  CFCHK(check_conv_args(h, 1, CF_SCALE_INV_N3, precision));
  CFCHK(require_ready());
  SizeFns* f = nullptr;
  CFCHK(size_res(h->n, &f));
  const unsigned n = h->n;
  hipStream_t s = (hipStream_t)stream;
  void* work = nullptr;
  if (needs_work(h->plan)) CFCHK(acquire_work(work_bytes(h->plan, n, 1, d_sig, d_out), s, &work));
  struct WorkRelease {
    void* w;
    hipStream_t s;
    ~WorkRelease() {
      if (w) release_work(s);
    }
  } wr{work, s};
  PassList p;
  conv_passes(*f, h->plan, n, 1, d_sig, d_out, work, h->d, 0, (float)(1.0 / (double)cube_elems(n)), p, precision);
  // every work-group count stays a multiple of 8 x the widest sibling group (8)
  for (const Pass& q : p)
    if (shares > 1 && q.tiles % (shares * 64) != 0)
      return fail(CF_ERR_ARG, "pass %s of N = %u has %u tiles: too few for %u shares", q.name, n, q.tiles, shares);
  CFCHK(ensure_events());
  HIPCHK(hipEventRecord(g.ev[0], s));
  for (size_t i = 0; i < p.size(); ++i) {
    const Pass& q = p[i];
    const void* in = q.in;
    void* out = q.out;
    const void* H = q.H;
    long long hs = q.hstride;
    float sc = q.scale;
    const void* twp = f->tw;
    int ntotal = (int)(q.tiles / shares);
    void* args[] = {(void*)&in, (void*)&out, (void*)&H, (void*)&twp, (void*)&hs, (void*)&sc, (void*)&ntotal};
    HIPCHK(hipModuleLaunchKernel(q.f, (unsigned)ntotal, 1, 1, q.threads, 1, 1, 0, s, args, nullptr));
    HIPCHK(hipEventRecord(g.ev[i + 1], s));
  }
  HIPCHK(hipEventSynchronize(g.ev[p.size()]));
  memset(times, 0, sizeof *times);
  times->npass = (unsigned)p.size();
  for (size_t i = 0; i < p.size(); ++i) {
    times->ms[i] = (float)ms_between(g.ev[i], g.ev[i + 1]);
    snprintf(times->name[i], sizeof times->name[i], "%s", p[i].name);
    times->bytes[i] = p[i].bytes / shares;
  }
  return CF_OK;
}

extern "C" int cf_profile_begin(unsigned max_calls, unsigned every) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (max_calls == 0 || max_calls > 100000) return fail(CF_ERR_ARG, "max_calls must be in 1..100000");
  if (every == 0 || every > 1000) return fail(CF_ERR_ARG, "every must be in 1..1000");
  CFCHK(require_ready());
  for (auto e : g.prof_ev) (void)hipEventDestroy(e);
  g.prof_ev.assign((size_t)max_calls * (CF_MAX_PASSES + 1), nullptr);
  // timing-only events: no system-scope fence (cache writeback) when they
  // complete, which would otherwise stall the next pass for microseconds
  for (auto& e : g.prof_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  // each event recorded once here: the runtime sets an event up on its first
  // record, which would otherwise stall the host inside the profiled calls
  // (and let the GPU run dry behind it)
  for (auto e : g.prof_ev) HIPCHK(hipEventRecord(e, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  g.prof_max = max_calls;
  g.prof_calls = 0;
  g.prof_every = every;
  g.prof_seen = 0;
  return CF_OK;
}

extern "C" int cf_profile_end(cf_pass_times_t* avg, unsigned* ncalls) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!avg || !ncalls) return fail(CF_ERR_ARG, "NULL pointer");
  memset(avg, 0, sizeof *avg);
  *ncalls = g.prof_calls;
  int rc = CF_OK;
  if (g.prof_calls > 0) {
    const size_t stride = CF_MAX_PASSES + 1;
    hipEvent_t last = g.prof_ev[(size_t)(g.prof_calls - 1) * stride + g.prof_npass];
    if (hipEventSynchronize(last) != hipSuccess) rc = fail(CF_ERR_RUNTIME, "event synchronize failed");
    avg->npass = g.prof_npass;
    for (unsigned i = 0; i < g.prof_npass && rc == CF_OK; ++i) {
      double sum = 0.0;
      for (unsigned c = 0; c < g.prof_calls; ++c)
        sum += ms_between(g.prof_ev[c * stride + i], g.prof_ev[c * stride + i + 1]);
      avg->ms[i] = (float)(sum / g.prof_calls);
      snprintf(avg->name[i], sizeof avg->name[i], "%s", g.prof_names[i].c_str());
      avg->bytes[i] = g.prof_bytes[i];
    }
  }
  for (auto e : g.prof_ev) (void)hipEventDestroy(e);
  g.prof_ev.clear();
  g.prof_max = g.prof_calls = 0;
  return rc;
}

extern "C" int cf_fft3d_device(unsigned n, const void* d_in, void* d_out, unsigned how_many, int inverse,
                               void* stream) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!d_in || !d_out || how_many == 0) return fail(CF_ERR_ARG, "NULL pointer or zero count");
  CFCHK(require_ready());
  SizeFns* f = nullptr;
  CFCHK(size_res(n, &f));
  PassList p;
  if (inverse)
    inv_passes(*f, n, how_many, d_in, d_out, 1.0f, p);
  else
    fwd_passes(*f, n, how_many, d_in, d_out, p);
  return run_passes(p, n, how_many, (hipStream_t)stream, f->tw, nullptr);
}

extern "C" int cf_rel_l2_device(const void* d_test, const void* d_ref, int ref_is_f64, size_t npts, void* stream,
                                double* rel_l2) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!d_test || !d_ref || !rel_l2 || npts == 0) return fail(CF_ERR_ARG, "NULL pointer or zero size");
  CFCHK(require_ready());
  CFCHK(ensure_device(&g.d_err, &g.d_err_bytes, (size_t)(2 * kErrBlocks + 2) * sizeof(double)));
  hipStream_t s = (hipStream_t)stream;
  int nblk = (int)std::min<size_t>(kErrBlocks, (npts + 255) / 256);
  const void* t = d_test;
  const void* r = d_ref;
  int f64 = ref_is_f64 ? 1 : 0;
  long long np = (long long)npts;
  void* part = g.d_err;
  void* a1[] = {(void*)&t, (void*)&r, (void*)&f64, (void*)&np, (void*)&part};
  HIPCHK(hipModuleLaunchKernel(g.errp, nblk, 1, 1, 256, 1, 1, 0, s, a1, nullptr));
  void* res = (char*)g.d_err + (size_t)2 * kErrBlocks * sizeof(double);
  void* a2[] = {(void*)&part, (void*)&nblk, (void*)&res};
  HIPCHK(hipModuleLaunchKernel(g.errf, 1, 1, 1, 256, 1, 1, 0, s, a2, nullptr));
  double host[2] = {0.0, 0.0};
  HIPCHK(hipMemcpyAsync(host, res, sizeof host, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *rel_l2 = host[1] > 0.0 ? std::sqrt(host[0] / host[1]) : std::sqrt(host[0]);
  return CF_OK;
}
extern "C" int cf_fill_uniform(void* d_out, size_t nfloats, unsigned long long seed, void* stream) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!d_out || nfloats == 0) return fail(CF_ERR_ARG, "NULL pointer or zero size");
  CFCHK(require_ready());
  long long nf = (long long)nfloats;
  unsigned long long sd = seed;
  const unsigned blocks = (unsigned)std::min<size_t>(4096, (nfloats + 255) / 256);
  void* args[] = {&d_out, &nf, &sd};
  HIPCHK(hipModuleLaunchKernel(g.fillu, blocks, 1, 1, 256, 1, 1, 0, (hipStream_t)stream, args, nullptr));
  return CF_OK;
}
extern "C" int cf_fill_loguniform(void* d_out, size_t npts, unsigned long long seed, void* stream) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!d_out || npts == 0) return fail(CF_ERR_ARG, "NULL pointer or zero size");
  CFCHK(require_ready());
  long long np = (long long)npts;
  unsigned long long sd = seed;
  const unsigned blocks = (unsigned)std::min<size_t>(4096, (npts + 255) / 256);
  void* args[] = {&d_out, &np, &sd};
  HIPCHK(hipModuleLaunchKernel(g.filll, blocks, 1, 1, 256, 1, 1, 0, (hipStream_t)stream, args, nullptr));
  return CF_OK;
}

// ---- batches -------------------------------------------------------------------
struct cf_batch_s {
  int mode = CF_BATCH_PERSISTENT;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* d_segs = nullptr;
  unsigned* d_ctr = nullptr;
  void* d_work = nullptr;  // own workspace of the graph mode (four-step plans)
  hipStream_t last = nullptr;
  bool launched = false;
};

static void batch_free(cf_batch_s* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_segs) (void)hipFree(b->d_segs);
  if (b->d_ctr) (void)hipFree(b->d_ctr);
  if (b->d_work) (void)hipFree(b->d_work);
  delete b;
}

static int batch_capture_end(cf_batch_s* b, int rc) {
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(g.capture, &graph);
  if (rc != CF_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(CF_ERR_RUNTIME, "hipStreamEndCapture: %s", hipGetErrorString(e));
  b->graph = graph;
  HIPCHK(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0));
  return CF_OK;
}

// Persistent-mode schedule of one instance: its passes (kind, tiles, input,
// output) on the plan of its spectrum, or CF_ERR_ARG when the persistent kernel
// does not run that plan.
struct BatchPass {
  int kind, variant, ntiles, tile_threads, siblings;
  const void* in;
  void* out;
  const void* H;
  float scale;
};

static int batch_schedule(const cf_instance_t& x, unsigned i, void* w1, void* w2, float last_scale,
                          std::vector<BatchPass>& v) {
  const unsigned n = x.h->n;
  const int plan = x.h->plan;
  const void* sig = x.d_sig;
  void* out = x.d_out;
  const void* H = x.h->d;
  const int rt = cfgeo::row_tiles((int)n), ct = cfgeo::col_tiles((int)n);
  if (plan == CF_PLAN_AXIS5 && n >= 16 && n <= 256) {
    v.push_back({CF_BK_ROW_FWD, 0, rt, 256, 1, sig, out, nullptr, 1.0f});
    v.push_back({CF_BK_COLY_FWD, 0, ct, 256, 1, out, out, nullptr, 1.0f});
    v.push_back({CF_BK_COLZ_CONV, 0, ct, 256, 1, out, out, H, 1.0f});
    v.push_back({CF_BK_COLY_INV, 0, ct, 256, 1, out, out, nullptr, 1.0f});
    v.push_back({CF_BK_ROW_INV, 0, rt, 256, 1, out, out, nullptr, last_scale});
    return CF_OK;
  }
  if (plan == CF_PLAN_PLANE3 && (n == 32 || n == 64)) {
    const int pt = cfgeo::plane_tiles((int)n), pth = cfgeo::plane_threads((int)n);
    v.push_back({CF_BK_PLANE_FWD, 0, pt, pth, 1, sig, out, nullptr, 1.0f});
    v.push_back({CF_BK_COLZ_CONV, 0, ct, 256, 1, out, out, H, 1.0f});
    v.push_back({CF_BK_PLANE_INV, 0, pt, pth, 1, out, out, nullptr, last_scale});
    return CF_OK;
  }
  if (is_split(plan) && cfgeo::batch_split_supported(split_variant(plan)) && (n == 128 || n == 256)) {
    const int vi = split_variant(plan), nn = (int)n;
    const int ga = cfgeo::split_ga(vi), gb = cfgeo::split_gb(vi), gm = cfgeo::split_gm(vi), gbo = cfgeo::split_gbo(vi);
    const int gao = cfgeo::split_gao(vi);
    const int sa = ga * gb >= 16 ? 1 : 16 / (ga * gb), sc = gao * gbo >= 16 ? 1 : 16 / (gao * gbo);
    v.push_back({CF_BK_XY1_FWD, vi, cfgeo::split_a_tiles(nn, vi), cfgeo::split_a_threads(nn, vi), sa, sig, w1, nullptr,
                 1.0f});
    v.push_back({CF_BK_MID, vi, cfgeo::split_b_tiles(nn, vi), cfgeo::split_b_threads(nn, vi), gb / gm, w1, w2, H, 1.0f});
    v.push_back({CF_BK_XY1_INV, vi, cfgeo::split_c_tiles(nn, vi), cfgeo::split_c_threads(nn, vi), sc, w2, out, nullptr,
                 last_scale});
    return CF_OK;
  }
  return fail(CF_ERR_ARG,
              "instance %u: the persistent batch runs the axis plan (N = 16..256), the plane plan (N = 32, 64) and "
              "the four-step plans split3_g24m2 / split3_g24m2h (N = 128, 256); this spectrum is N = %u, plan %d",
              i, n, plan);
}

static int batch_create(const cf_instance_t* inst, unsigned count, int scale_mode, int mode, bool stall_test,
                        unsigned long long spin_limit, cf_batch_t* out) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!inst || !out || count == 0) return fail(CF_ERR_ARG, "NULL pointer or zero count");
  *out = nullptr;
  if (scale_mode != CF_SCALE_NONE && scale_mode != CF_SCALE_INV_N3) return fail(CF_ERR_ARG, "bad scale_mode");
  if (mode != CF_BATCH_PERSISTENT && mode != CF_BATCH_GRAPH) return fail(CF_ERR_ARG, "bad batch mode %d", mode);
  CFCHK(require_ready());
  std::vector<SizeFns*> fns(count);
  size_t work_need = 0;
  for (unsigned i = 0; i < count; ++i) {
    const cf_instance_t& x = inst[i];
    if (!x.d_sig || !x.d_out || !x.h) return fail(CF_ERR_ARG, "instance %u: NULL pointer", i);
    if (x.h->count != 1) return fail(CF_ERR_ARG, "instance %u: spectrum must hold one filter", i);
    CFCHK(size_res(x.h->n, &fns[i]));
    if (mode == CF_BATCH_PERSISTENT) {
      if (x.h->n < 16 || x.h->n > 256)
        return fail(CF_ERR_SIZE, "instance %u: the persistent batch runs edges 16..256, not %u", i, x.h->n);
      std::vector<BatchPass> probe;
      CFCHK(batch_schedule(x, i, nullptr, nullptr, 1.0f, probe));
    }
  }
  if (mode == CF_BATCH_PERSISTENT) {
    // The passes of different instances run at the same time in one launch,
    // so no instance may write memory that another one reads or writes: every
    // output must be disjoint from every other instance's signal and output
    // (an output may be its own signal; a chained pipeline needs graph mode,
    // which keeps the instance order).
    struct Span {
      const char* a;
      const char* b;
    };
    auto span = [&](const void* p, unsigned i) {
      return Span{(const char*)p, (const char*)p + cube_bytes(inst[i].h->n)};
    };
    auto meet = [](Span x, Span y) { return x.a < y.b && y.a < x.b; };
    std::vector<cfspan::Span> all;
    all.reserve(2 * (size_t)count);
    for (unsigned i = 0; i < count; ++i) {
      const Span si = span(inst[i].d_sig, i), oi = span(inst[i].d_out, i);
      if (meet(si, oi) && si.a != oi.a)
        return fail(CF_ERR_ARG, "instance %u: output partly overlaps its own signal", i);
      all.push_back({(uintptr_t)oi.a, (uintptr_t)oi.b, i, true});
      if (si.a != oi.a) all.push_back({(uintptr_t)si.a, (uintptr_t)si.b, i, false});
    }
    unsigned i = 0, j = 0;
    if (cfspan::find_conflict(std::move(all), &i, &j))
      return fail(CF_ERR_ARG,
                  "instances %u and %u: the output of %u overlaps the buffers of %u; instances of a "
                  "persistent batch run concurrently (use CF_BATCH_GRAPH for chained convolutions)",
                  i, j, i, j);
  }
  if (!g.capture) HIPCHK(hipStreamCreateWithFlags(&g.capture, hipStreamNonBlocking));
  auto* b = new cf_batch_s;
  b->mode = mode;
  int rc = CF_OK;
  if (mode == CF_BATCH_PERSISTENT) {
    // workspace: the intermediates of the four-step instances, one after the other
    std::vector<size_t> woff(count, 0);
    for (unsigned i = 0; i < count; ++i) {
      woff[i] = work_need;
      work_need += work_bytes(inst[i].h->plan, inst[i].h->n, 1, inst[i].d_sig, inst[i].d_out);
    }
    if (work_need) {
      const hipError_t e = hipMalloc(&b->d_work, work_need);
      if (e != hipSuccess) {
        batch_free(b);
        return fail(CF_ERR_NOMEM, "batch workspace: %s", hipGetErrorString(e));
      }
    }
    // every instance's passes; segment = one pass of one instance, stage-major
    std::vector<std::vector<BatchPass>> sched(count);
    int nstages = 0;
    for (unsigned i = 0; i < count; ++i) {
      const unsigned n = inst[i].h->n;
      void* w1 = (char*)b->d_work + woff[i];
      void* w2 = w1;  // as conv_passes places the intermediates
      if (two_intermediates(inst[i].h->plan)) {
        if (inst[i].d_sig != inst[i].d_out)
          w1 = inst[i].d_out;
        else
          w2 = (char*)w1 + cube_bytes(n);
      }
      const float last = scale_mode == CF_SCALE_INV_N3 ? (float)(1.0 / (double)cube_elems(n)) : 1.0f;
      rc = batch_schedule(inst[i], i, w1, w2, last, sched[i]);
      if (rc != CF_OK) {
        batch_free(b);
        return rc;
      }
      nstages = std::max(nstages, (int)sched[i].size());
    }
    // the kernel specialised to the three-pass set when every instance is in
    // it, with half the LDS (more work-groups per CU) when every edge is 32 or 64
    bool three_pass = true, small = true;
    for (unsigned i = 0; i < count; ++i) {
      const unsigned n = inst[i].h->n;
      const int p = inst[i].h->plan;
      three_pass = three_pass && ((p == CF_PLAN_PLANE3 && (n == 32 || n == 64)) ||
                                  (is_split(p) && split_variant(p) == cfgeo::kBatchSplit[1] && (n == 128 || n == 256)));
      small = small && n <= 64;
    }
    hipFunction_t kfn = !three_pass ? g.batch_run : small ? g.batch_run_3p_small : g.batch_run_3p;
    // The small-edge kernel's 41 KiB of LDS would let three work-groups share
    // a CU; two run c3 (64 x 64^3) faster (0.185 against 0.202 ms; four: 0.222),
    // so its launch reserves LDS it does not use to hold it to two.
    const unsigned dyn_lds = kfn == g.batch_run_3p_small ? kBatchSmallReserve : 0u;
    int per_cu = 1, cus = 1;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kBatchThreads, dyn_lds) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g.device) != hipSuccess || cus < 1) cus = 1;
    const int full_grid = per_cu * cus;  // every work-group resident
    std::vector<CfBatchSeg> segs;
    std::vector<int> seg_stage;
    int stage_need[kBatchMaxStages] = {};
    int stage_first[kBatchMaxStages + 1] = {};  // first work item of each stage (stage-major numbering)
    int item = 0;
    // 256^3 instances whose spectrum no other instance reads stream their
    // signal, spectrum and output past the memory-side cache (256 MB), which
    // then keeps more of their intermediates, as the single-cube passes do;
    // smaller cubes fit it and lose by it (measured)
    std::vector<short> stream_of(count, 0);
    {
      std::vector<const void*> hs(count);
      for (unsigned i = 0; i < count; ++i) hs[i] = inst[i].h->d;
      std::sort(hs.begin(), hs.end());
      for (unsigned i = 0; i < count; ++i) {
        const auto r = std::equal_range(hs.begin(), hs.end(), (const void*)inst[i].h->d);
        stream_of[i] = (short)(inst[i].h->n == 256 && r.second - r.first == 1);
      }
    }
    // within a stage, the segments by edge, largest first: a work-group's
    // items (G apart) then change edge, and restage its twiddles, rarely
    std::vector<unsigned> by_edge(count);
    for (unsigned i = 0; i < count; ++i) by_edge[i] = i;
    std::stable_sort(by_edge.begin(), by_edge.end(),
                     [&](unsigned a, unsigned c) { return inst[a].h->n > inst[c].h->n; });
    for (int st = 0; st < nstages; ++st) {
      stage_first[st] = item;
      for (const unsigned i : by_edge) {
        if (st >= (int)sched[i].size()) continue;
        const BatchPass& q = sched[i][st];
        CfBatchSeg sg{};
        sg.in = q.in;
        sg.out = q.out;
        sg.H = q.H;
        sg.tw = fns[i]->tw;
        sg.scale = q.scale;
        sg.n = (int)inst[i].h->n;
        sg.kind = q.kind;
        sg.variant = (short)q.variant;
        sg.stream = stream_of[i];
        sg.first_item = item;
        sg.ntiles = q.ntiles;
        // A work item: the tiles a work-group runs side by side.  Sibling tiles
        // that share lines of an intermediate must meet in one L2: when there
        // are more of them than that, their items are placed 8 apart (one XCD,
        // run at the same time) if the grid is a multiple of 8 and the
        // segment's items fill whole groups of 8 siblings; otherwise one item
        // holds them all and one work-group runs them one after another.
        const int side = kBatchThreads / q.tile_threads;
        sg.per_item = std::max(side, q.siblings);
        sg.xcd_group = 1;
        if (q.siblings > side && q.siblings % side == 0 && full_grid % 8 == 0 && q.ntiles % side == 0 &&
            (q.ntiles / side) % (8 * (q.siblings / side)) == 0) {
          sg.per_item = side;
          sg.xcd_group = q.siblings / side;
        }
        seg_stage.push_back(st);
        if (sg.n >= 32 && sg.ntiles % (kBatchThreads / q.tile_threads) != 0) {  // tiles run in whole groups
          batch_free(b);
          return fail(CF_ERR_RUNTIME, "pass kind %d of N = %d: %d tiles do not fill work-groups", q.kind, sg.n,
                      sg.ntiles);
        }
        item += (sg.ntiles + sg.per_item - 1) / sg.per_item;
        stage_need[st] += sg.ntiles;
        segs.push_back(sg);
      }
    }
    for (int st = nstages; st <= kBatchMaxStages; ++st) stage_first[st] = item;
    // per work item: its segment and stage (segment | stage << 24)
    if (segs.size() >= (1u << 24)) {
      batch_free(b);
      return fail(CF_ERR_ARG, "too many segments in one persistent batch");
    }
    std::vector<int> item_tab(item);
    for (size_t k = 0; k < segs.size(); ++k) {
      const CfBatchSeg& sg = segs[k];
      const int items = (sg.ntiles + sg.per_item - 1) / sg.per_item;
      for (int j = 0; j < items; ++j) item_tab[sg.first_item + j] = (int)k | (seg_stage[k] << 24);
    }
    if (stall_test) {
      // test support: the last stage waits for one tile more than the one
      // before it has, so its wait runs into the (short) spin limit
      if (nstages < 2) {
        batch_free(b);
        return fail(CF_ERR_ARG, "a stall test needs a schedule of at least two passes");
      }
      stage_need[nstages - 2] += 1;
    }
    const int nitems = item;
    const size_t seg_bytes = sizeof(CfBatchSeg) * segs.size();
    const size_t tab_bytes = sizeof(int) * item_tab.size();
    hipError_t e = hipMalloc(&b->d_segs, seg_bytes + sizeof stage_need + sizeof stage_first + tab_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_ctr, sizeof(unsigned) * kBatchCounters * kBatchCtrStride);
    if (e == hipSuccess) e = hipMemcpy(b->d_segs, segs.data(), seg_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy((char*)b->d_segs + seg_bytes, stage_need, sizeof stage_need, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy((char*)b->d_segs + seg_bytes + sizeof stage_need, stage_first, sizeof stage_first,
                    hipMemcpyHostToDevice);
    if (e == hipSuccess && tab_bytes)
      e = hipMemcpy((char*)b->d_segs + seg_bytes + sizeof stage_need + sizeof stage_first, item_tab.data(), tab_bytes,
                    hipMemcpyHostToDevice);
    // the counters start at zero once; the kernel leaves them at zero for the next launch
    if (e == hipSuccess) e = hipMemset(b->d_ctr, 0, sizeof(unsigned) * kBatchCounters * kBatchCtrStride);
    if (e != hipSuccess) {
      batch_free(b);
      return fail(CF_ERR_NOMEM, "batch descriptor allocation: %s", hipGetErrorString(e));
    }
    // as many work-groups as the largest stage has items (rounded up to whole
    // groups of the 8 XCDs): more would sit idle in every stage yet take part
    // in its counting (one 128^3 convolution: 0.066 against 0.078 ms)
    int most = 1;
    for (int st = 0; st < nstages; ++st) most = std::max(most, stage_first[st + 1] - stage_first[st]);
    const unsigned grid = (unsigned)std::min(full_grid, (most + 7) / 8 * 8);
    HIPCHK(hipStreamBeginCapture(g.capture, hipStreamCaptureModeRelaxed));
    unsigned* ctr = b->d_ctr;
    const void* sp = b->d_segs;
    const void* need = (const char*)b->d_segs + seg_bytes;
    int ni = nitems;
    unsigned long long sl = spin_limit;
    void* ba[] = {(void*)&sp, (void*)&need, &ni, &ctr, &sl};  // the kernels' parameters, in order
    if (hipModuleLaunchKernel(kfn, grid, 1, 1, kBatchThreads, 1, 1, dyn_lds, g.capture, ba, nullptr) != hipSuccess)
      rc = fail(CF_ERR_RUNTIME, "batch kernel capture failed");
    rc = batch_capture_end(b, rc);
  } else {
    // runs of consecutive instances with one edge and plan and contiguous
    // signals / outputs, and either one spectrum or spectra that follow each
    // other in memory (views of one multi-filter spectrum), are recorded as
    // one batched pass sequence; plans that go through the workspace get
    // room for their longest run
    struct Run {
      unsigned i, run;
      bool shared;
    };
    std::vector<Run> runs;
    work_need = 0;
    for (unsigned i = 0; i < count;) {
      const unsigned n = inst[i].h->n;
      const size_t V = cube_bytes(n);
      unsigned run = 1;
      const bool shared = i + 1 < count && inst[i + 1].h == inst[i].h;
      while (i + run < count && inst[i + run].h->n == n && inst[i + run].h->plan == inst[i].h->plan &&
             (shared ? inst[i + run].h == inst[i].h
                     : (const char*)inst[i + run].h->d == (const char*)inst[i].h->d + run * V) &&
             (const char*)inst[i + run].d_sig == (const char*)inst[i].d_sig + run * V &&
             (char*)inst[i + run].d_out == (char*)inst[i].d_out + run * V)
        ++run;
      work_need = std::max(work_need, work_bytes(inst[i].h->plan, n, run, inst[i].d_sig, inst[i].d_out));
      runs.push_back({i, run, shared});
      i += run;
    }
    if (work_need) {
      const hipError_t e = hipMalloc(&b->d_work, work_need);
      if (e != hipSuccess) {
        batch_free(b);
        return fail(CF_ERR_NOMEM, "batch workspace: %s", hipGetErrorString(e));
      }
    }
    HIPCHK(hipStreamBeginCapture(g.capture, hipStreamCaptureModeRelaxed));
    for (const Run& r : runs) {
      if (rc != CF_OK) break;
      const unsigned i = r.i, n = inst[i].h->n;
      const long long hstride = (r.run > 1 && !r.shared) ? (long long)cube_elems(n) : 0;
      const float scale = scale_mode == CF_SCALE_INV_N3 ? (float)(1.0 / (double)cube_elems(n)) : 1.0f;
      PassList p;
      conv_passes(*fns[i], inst[i].h->plan, n, r.run, inst[i].d_sig, inst[i].d_out, b->d_work, inst[i].h->d, hstride,
                  scale, p);
      rc = run_passes(p, n, r.run, g.capture, fns[i]->tw, nullptr);
    }
    rc = batch_capture_end(b, rc);
  }
  if (rc != CF_OK) {
    batch_free(b);
    return rc;
  }
  *out = b;
  return CF_OK;
}

extern "C" int cf_batch_create(const cf_instance_t* inst, unsigned count, int scale_mode, int mode, cf_batch_t* out) {
  return batch_create(inst, count, scale_mode, mode, false, 1ull << 21, out);
}

extern "C" int cf_batch_create_stall_test(const cf_instance_t* inst, unsigned count, int scale_mode,
                                          unsigned long long spin_limit, cf_batch_t* out) {
  if (spin_limit == 0) return fail(CF_ERR_ARG, "spin_limit must be at least 1");
  return batch_create(inst, count, scale_mode, CF_BATCH_PERSISTENT, true, spin_limit, out);
}

extern "C" int cf_batch_launch(cf_batch_t b, void* stream) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!b || !b->exec) return fail(CF_ERR_ARG, "NULL batch");
  CFCHK(require_ready());
  HIPCHK(hipGraphLaunch(b->exec, (hipStream_t)stream));
  b->last = (hipStream_t)stream;
  b->launched = true;
  return CF_OK;
}

extern "C" int cf_batch_check(cf_batch_t b) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!b) return fail(CF_ERR_ARG, "NULL batch");
  CFCHK(require_ready());
  if (!b->launched) return CF_OK;
  HIPCHK(hipStreamSynchronize(b->last));
  if (b->mode == CF_BATCH_PERSISTENT) {
    unsigned flag = 0;
    HIPCHK(hipMemcpy(&flag, b->d_ctr + kBatchCtrStride, sizeof flag, hipMemcpyDeviceToHost));  // counter 1
    if (flag) {
      const unsigned zero = 0;  // the next launch starts clean
      HIPCHK(hipMemcpy(b->d_ctr + kBatchCtrStride, &zero, sizeof zero, hipMemcpyHostToDevice));
      return fail(CF_ERR_RUNTIME, "persistent batch: a stage's dependency was not met in time");
    }
  }
  return CF_OK;
}

extern "C" void cf_batch_destroy(cf_batch_t b) {
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!b) return;
  (void)hipSetDevice(g.device);
  if (b->launched) (void)hipStreamSynchronize(b->last);
  batch_free(b);
}
#pragma GCC visibility pop
