// engine.h -- internal boundary between the drop-in C-ABI translation unit
// (capi.cpp, which sees the public header's own `float2` typedef) and the HIP
// runtime translation unit (engine.cpp, which must not see that typedef
// because HIP defines a vector type of the same name).  Only void pointers and
// plain structs cross it.
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double filter_pcie_wr_t, filter_exec_t, sig_pcie_wr_t, sig_pcie_rd_t, sig_exec_t, siginv_exec_t;
} cfe_times_t;

#define CFE_MODE_COPY 0 /* explicit host<->device copies (fpgaf_conv3D)            */
#define CFE_MODE_SVM 1  /* kernels read/write pinned host memory (fpgaf_conv3D_svm)  */
#define CFE_MODE_SVM_BATCH 2 /* cubes streamed through overlapped transfers and passes
                                (fpgaf_conv3D_svm_batch)                                */

int cfe_initialize(const char *platform_name, const char *path, int use_svm);
void cfe_final(void);
/* 0 on success; negative CF_ERR_* otherwise (message in cf_last_error()). */
int cfe_conv_host(unsigned N, const void *sig, const void *filter, void *out, unsigned how_many, int mode,
                  cfe_times_t *times);

#ifdef __cplusplus
}
#endif
