"""convfpga_amd -- FFT-based 3D convolution on AMD Instinct MI355X (gfx950).

The product is the native library under lib/ (C ABI: include/convfpga/*.h);
this package only builds it (convfpga_amd.build) and binds it from Python
(convfpga_amd.api).
"""
from . import api  # noqa: F401
from .api import (CODE_OBJECT, LIB_PATH, SUPPORTED_SIZES, ConvError, Spectrum, conv3d_device,  # noqa: F401
                  conv3d_device_timed, conv3d_host, fft3d_device, fpga_final, fpga_initialize, fpgaf_complex_malloc,
                  fpgaf_conv3D, fpgaf_conv3D_svm, fpgaf_conv3D_svm_batch, rel_l2_device)
