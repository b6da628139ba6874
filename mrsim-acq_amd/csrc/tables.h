// tables.h -- host-side protocol tables shared by the plan (plan.cpp), the 3D train (ge3d.cpp) and
// their C entry points (tables.cpp).  No HIP dependency.  Not part of the public ABI.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/mrsim_gpu.h"

namespace mrs {

// PE lines a Scanner-mode partial Fourier acquisition skips (their readout starts that much earlier).
size_t pf_skipped_lines(size_t ny, const mrs_acquisition& a);
// LineTiming::for_acquisition (src/kspace.rs:310-326) from SingleShotEpi (src/readout.rs:21-69).
void timing_table(const mrs_acquisition& a, size_t nx, size_t ny, double* t, double* trf, double* tread,
                  int8_t* pol);
// Acquired PE rows (sampling_mask rows, src/kspace.rs:546-588).
std::vector<uint8_t> acquired_rows(size_t ny, const mrs_acquisition& a);
// Round-to-nearest-even bf16 of a float, its value, and the packed (hi, lo) split of an FP64 complex.
uint16_t bf16_rne(float f);
float bf16_to_float(uint16_t h);
void split_c(double re, double im, uint32_t& hi, uint32_t& lo);
// coil_sensitivity (src/kspace.rs:226-237).
double ring_sensitivity(size_t coil, size_t n_coils, double x, double y, size_t nx, size_t ny);
// readout::grase_block (src/readout.rs:206-221) and readout::kz_order (:170-189).
int grase_block(size_t ny, size_t kseg, double t_line, bool reverse, std::vector<double>& t, std::vector<int8_t>& pol,
                char* err, size_t errlen);
std::vector<size_t> kz_order(size_t nz, bool centric);

}  // namespace mrs
