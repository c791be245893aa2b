/*
 * convfpga.h -- drop-in C API of the MI355X FFT-convolution library.
 *
 * Same types, names and signatures as the reference's public header
 * (host/include/convfpga/convfpga.h:16-46), so that an application written
 * against the reference (e.g. src/main.cpp) builds and links against
 * libconvfpga unchanged.  What each entry point does on the GPU is documented
 * beside it; the reference interface it replaces is cited with file:line.
 *
 * Result convention (reference semantics): out = IFFT3(FFT3(sig) .* FFT3(filter))
 * with an UNNORMALISED inverse, i.e. N^3 times the circular convolution, in
 * natural order (index z*N*N + y*N + x, x fastest).
 */
#ifndef CONVFPGA_H
#define CONVFPGA_H

#include <stdbool.h>
#include <stddef.h>

/* Interleaved single-precision complex value: x = real, y = imaginary.
 * Layout-identical to the reference's float2 (convfpga.h:16-19). */
typedef struct {
  float x;
  float y;
} float2;

/* Milliseconds spent per phase of one call, plus a validity flag.
 * Field order and meaning follow the reference's fpga_t (convfpga.h:24-34):
 *   filter_pcie_wr_t  host->device copy of the filter
 *   filter_exec_t     3D FFT of the filter
 *   sig_pcie_wr_t     host->device copy of the signal(s)
 *   sig_pcie_rd_t     device->host copy of the result(s)
 *   sig_exec_t        forward x and y passes + fused z-forward/product/z-inverse pass
 *   siginv_exec_t     inverse y and x passes
 *   valid             true when the call completed */
typedef struct fpga_timing {
  double filter_pcie_wr_t;
  double filter_exec_t;

  double sig_pcie_wr_t;
  double sig_pcie_rd_t;
  double sig_exec_t;
  double siginv_exec_t;

  bool valid;
} fpga_t;

#ifdef __cplusplus
extern "C" {
#endif

/* Select the GPU and load the kernel code object.
 * Replaces fpga_initialize (reference fpga_init.c:57-122).
 *   platform_name  case-insensitive substring of "AMD HIP", "ROCm", the device
 *                  name (e.g. "AMD Instinct MI355X") or its gfx target ("gfx950")
 *   path           code object built by this project (convfpga_gfx950.hsaco),
 *                  the analogue of the reference's .aocx bitstream
 *   use_svm        require host-mapped memory support (used by the _svm calls)
 * Returns 0, or -1 path NULL/empty, -2 platform not matched, -3 no device,
 * -4 code object not loadable, -5 host-mapped memory unsupported. */
extern int fpga_initialize(const char *platform_name, const char *path, const bool use_svm);

/* Release everything fpga_initialize and the conv calls allocated.
 * Replaces fpga_final (fpga_init.c:127-137); unlike the reference it is safe
 * to call more than once. */
extern void fpga_final(void);

/* 64-byte aligned host allocation, NULL when sz == 0; release with free().
 * Same contract as fpgaf_complex_malloc (fpga_init.c:37-43, opencl_utils.c:204-212). */
extern void *fpgaf_complex_malloc(const size_t sz);

/* One N^3 convolution, host buffers, explicit copies through device memory.
 * Replaces fpgaf_conv3D (host/src/convfpga.c:32-409).  Invalid arguments
 * (NULL pointers, N not a supported power of two, library not initialised) and
 * runtime errors return valid == false instead of exiting the process. */
extern fpga_t fpgaf_conv3D(const unsigned N, const float2 *sig, const float2 *filter, float2 *out);

/* Same result; the first and last passes read and write pinned host memory
 * directly (zero-copy), the pcie_* fields stay 0.
 * Replaces fpgaf_conv3D_svm (host/src/convfpga_svm.c:34-461). */
extern fpga_t fpgaf_conv3D_svm(const unsigned N, const float2 *sig, const float2 *filter, float2 *out);

/* how_many contiguous N^3 signals convolved with one shared N^3 filter.
 * Replaces fpgaf_conv3D_svm_batch (host/src/convfpga_svm_batch.c:36-723);
 * any how_many >= 1 is accepted (the reference needs an even count), and every
 * instance may hold different data.  sig_exec_t is the summed device time of
 * all passes, siginv_exec_t the wall span from the first to the last pass. */
extern fpga_t fpgaf_conv3D_svm_batch(const unsigned N, const float2 *sig, const float2 *filter, float2 *out,
                                     const unsigned how_many);

#ifdef __cplusplus
}
#endif
#endif
