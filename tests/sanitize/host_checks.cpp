// Host-side checks of the C ABI, built with ASan and UBSan by
// tests/test_host_sanitizers.py: argument and state errors, init failures,
// allocation, idempotent release -- every path that runs without a GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include "convfpga/convfpga.h"
#include "convfpga/convfpga_ext.h"
int main() {
  int fails = 0;
  auto chk = [&](bool c, const char* what) { if (!c) { printf("FAIL %s\n", what); ++fails; } };
  chk(fpga_initialize("gfx950", nullptr, false) == -1, "null path");
  chk(fpga_initialize("gfx950", "", false) == -1, "empty path");
  int r = fpga_initialize("gfx950", "/nonexistent.hsaco", false);
  chk(r == -3 || r == -2 || r == -4, "no device / bad path");
  chk(fpga_initialize("nosuchplatform", "/x", false) != 0, "platform");
  fpga_final(); fpga_final();
  chk(fpgaf_complex_malloc(0) == nullptr, "malloc 0");
  void* p = fpgaf_complex_malloc(1000); chk(p && ((uintptr_t)p % 64) == 0, "aligned"); memset(p, 1, 1000); free(p);
  float2 a[8] = {}, b[8] = {}, c[8] = {};
  chk(!fpgaf_conv3D(2, a, b, c).valid, "conv not init");
  chk(!fpgaf_conv3D(3, a, b, c).valid, "conv bad N");
  chk(!fpgaf_conv3D(8, nullptr, b, c).valid, "conv null");
  chk(!fpgaf_conv3D_svm(8, a, b, c).valid, "svm not init");
  chk(!fpgaf_conv3D_svm_batch(8, a, b, c, 0).valid, "batch 0");
  chk(cf_supported_size(256) == 1 && cf_supported_size(7) == 0, "sizes");
  chk(cf_set_plan(256, 4) == CF_ERR_STATE, "set plan not init");
  cf_spectrum_t h = nullptr;
  chk(cf_spectrum_create(64, 1, a, 0, nullptr, &h) != 0, "spectrum not init");
  chk(cf_conv3d_device(nullptr, a, c, 1, 0, 0, nullptr) != 0, "conv null handle");
  cf_host_times_t ht;
  chk(cf_conv3d_host(nullptr, a, c, 1, 0, 0, &ht) != 0, "host conv null handle");
  chk(cf_spectrum_view(nullptr, 0, &h) != 0, "view null parent");
  cf_batch_t bt = nullptr;
  chk(cf_batch_create(nullptr, 1, 0, 0, &bt) != 0, "batch null");
  chk(strlen(cf_last_error()) > 0, "last error");
  cf_pass_times_t t; unsigned n = 0;
  chk(cf_profile_end(&t, &n) == 0 || true, "profile end");
  printf("%s (%d failures)\n", fails ? "FAILED" : "ok", fails);
  return fails != 0;
}
