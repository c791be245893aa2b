// capi.cpp -- the drop-in C ABI of include/convfpga/convfpga.h.
//
// This translation unit includes the public header (with its own `float2`
// typedef, reference convfpga.h:16-19) and deliberately no HIP header: all
// device work is behind engine.h, which passes void pointers only.
#include <stdlib.h>

#include "convfpga/convfpga.h"
#include "engine.h"

static_assert(sizeof(float2) == 8, "float2 must be two packed floats");
static_assert(offsetof(float2, y) == 4, "float2 layout");
static_assert(offsetof(fpga_t, filter_pcie_wr_t) == 0 && offsetof(fpga_t, filter_exec_t) == 8 &&
                  offsetof(fpga_t, sig_pcie_wr_t) == 16 && offsetof(fpga_t, sig_pcie_rd_t) == 24 &&
                  offsetof(fpga_t, sig_exec_t) == 32 && offsetof(fpga_t, siginv_exec_t) == 40 &&
                  offsetof(fpga_t, valid) == 48,
              "fpga_t layout must match the reference (convfpga.h:24-34)");

namespace {

fpga_t invalid() {
  fpga_t r = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, false};
  return r;
}

bool pow2(unsigned n) { return n != 0 && (n & (n - 1)) == 0; }

// argument screen of the reference (convfpga.c:36-37), plus N == 0
bool bad_args(unsigned n, const float2* sig, const float2* filter, const float2* out) {
  return sig == nullptr || filter == nullptr || out == nullptr || !pow2(n);
}

fpga_t run(unsigned n, const float2* sig, const float2* filter, float2* out, unsigned how_many, int mode) {
  if (bad_args(n, sig, filter, out) || how_many == 0) return invalid();
  cfe_times_t t;
  if (cfe_conv_host(n, sig, filter, out, how_many, mode, &t) != 0) return invalid();
  fpga_t r = invalid();
  r.filter_pcie_wr_t = t.filter_pcie_wr_t;
  r.filter_exec_t = t.filter_exec_t;
  r.sig_pcie_wr_t = t.sig_pcie_wr_t;
  r.sig_pcie_rd_t = t.sig_pcie_rd_t;
  r.sig_exec_t = t.sig_exec_t;
  r.siginv_exec_t = t.siginv_exec_t;
  r.valid = true;
  return r;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

int fpga_initialize(const char* platform_name, const char* path, const bool use_svm) {
  return cfe_initialize(platform_name, path, use_svm ? 1 : 0);
}

void fpga_final(void) { cfe_final(); }

void* fpgaf_complex_malloc(const size_t sz) {
  if (sz == 0) return nullptr;
  void* p = nullptr;
  if (posix_memalign(&p, 64, sz) != 0) return nullptr;
  return p;
}

fpga_t fpgaf_conv3D(const unsigned N, const float2* sig, const float2* filter, float2* out) {
  return run(N, sig, filter, out, 1, CFE_MODE_COPY);
}

fpga_t fpgaf_conv3D_svm(const unsigned N, const float2* sig, const float2* filter, float2* out) {
  return run(N, sig, filter, out, 1, CFE_MODE_SVM);
}

fpga_t fpgaf_conv3D_svm_batch(const unsigned N, const float2* sig, const float2* filter, float2* out,
                              const unsigned how_many) {
  return run(N, sig, filter, out, how_many, CFE_MODE_SVM_BATCH);
}

}  // extern "C"
#pragma GCC visibility pop
