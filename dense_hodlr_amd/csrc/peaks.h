/*
 * Sustained throughput of the current device (see peaks.hip).
 */
#ifndef DENSE_HODLR_PEAKS_H
#define DENSE_HODLR_PEAKS_H

namespace hodlr {

struct DevicePeaks {
  double hbm_read_gbs;      // GB/s read
  double hbm_copy_gbs;      // GB/s read + written
  double fp64_mfma_tflops;  // TFLOP/s
  double fp64_valu_tflops;
  double fp32_mfma_tflops;
};

/* Runs the probes on the current device (about half a second; allocates and
 * frees 1 GiB). */
DevicePeaks measure_device_peaks();

}  // namespace hodlr

#endif
