// tables_check.cpp -- drives the library's host-side protocol tables (mrsim-acq_amd/csrc/tables.cpp)
// under AddressSanitizer and UndefinedBehaviorSanitizer (built and run by tests/test_sanitize.py):
// the EPI timing table and sampling rows over odd and even matrices, every partial-Fourier mode,
// reversal and acceleration; the ring coil maps; the bf16 split on ordinary and special values; the
// 3D gradient-echo block timing, line table and timing checks, including the refusals.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "../../mrsim-acq_amd/csrc/tables.h"

int main() {
    double sum = 0.0;
    char err[512];
    const size_t mats[][2] = {{1, 1}, {2, 3}, {8, 9}, {16, 40}, {17, 70}, {64, 64}, {96, 97}, {128, 128}};
    const int pf_modes[] = {MRS_PF_FIBERFOX, MRS_PF_CONTIGUOUS, MRS_PF_SCANNER};
    for (const auto& m : mats) {
        const size_t nx = m[0], ny = m[1];
        for (int pf_mode : pf_modes)
            for (double pf : {1.0, 0.875, 0.75, 0.625, 0.5})
                for (int rev = 0; rev < 2; rev++)
                    for (uint64_t accel : {1u, 2u, 3u, 4u}) {
                        mrs_acquisition a;
                        mrs_acquisition_default(&a);
                        a.pf_mode = pf_mode;
                        a.partial_fourier = pf;
                        a.reverse_phase = rev;
                        a.accel = accel;
                        std::vector<double> t(ny), trf(ny), tread(ny);
                        std::vector<int8_t> pol(ny);
                        if (mrs_line_timing_for_acquisition(&a, nx, ny, t.data(), trf.data(), tread.data(), pol.data()))
                            return 2;
                        std::vector<uint8_t> mask(nx * ny);
                        if (mrs_sampling_mask(nx, ny, &a, mask.data())) return 2;
                        for (size_t k = 0; k < ny; k++) sum += t[k] + trf[k] + tread[k] + pol[k];
                        for (uint8_t v : mask) sum += v;
                    }
        const size_t o = 2, snx = nx * o, sny = ny * o;
        for (size_t nc : {1u, 3u, 8u}) {
            std::vector<mrs_c32> maps(nc * snx * sny);
            if (mrs_ring_coil_maps(0, nc, nc, snx, sny, nx, ny, 100.0, maps.data())) return 2;
            for (const auto& v : maps) sum += v.re;
        }
    }
    const double special[] = {0.0, -0.0, 1.0, -3.5e-41, 1e-45, 3.4e38, std::numeric_limits<double>::infinity(),
                              -std::numeric_limits<double>::infinity(), std::nan(""), 1.0 / 3.0};
    for (double re : special)
        for (double im : special) {
            uint32_t hi, lo;
            mrs::split_c(re, im, hi, lo);
            sum += (double)(hi ^ lo);
        }
    for (size_t ny : {4u, 6u, 8u, 12u, 64u, 96u})
        for (size_t kseg : {1u, 2u, 3u, 4u, 5u})
            for (int rev = 0; rev < 2; rev++) {
                std::vector<double> t(ny);
                std::vector<int8_t> pol(ny);
                const int rc = mrs_grase_block(ny, kseg, 0.4, rev, t.data(), pol.data(), err, sizeof err);
                if (rc == 0)
                    for (size_t k = 0; k < ny; k++) sum += t[k] + pol[k];
                else
                    sum += (double)std::strlen(err);
                for (size_t nz : {1u, 2u, 6u, 24u, 25u})
                    for (size_t kzs : {1u, 2u, 3u})
                        for (int centric = 0; centric < 2; centric++) {
                            std::vector<int32_t> shot(ny * nz), exc(ny * nz);
                            size_t nshots = 0, nexc = 0, ec = 0;
                            const int rl = mrs_ge3d_lines(ny, nz, kseg, 0.4, rev, kzs, centric, shot.data(), exc.data(),
                                                          &nshots, &nexc, &ec, err, sizeof err);
                            sum += rl == 0 ? (double)(nshots + nexc + ec + shot[0] + exc[ny * nz - 1])
                                           : (double)std::strlen(err);
                        }
                const double tes_ok[] = {12.0, 24.0, 36.0, 48.0};
                const double tes_bad[] = {1.0, 1.5};
                sum += mrs_check_ge3d_timing(tes_ok, 4, 60.0, 2.0, ny, kseg, 0.4, rev, 8, 0.0, 600.0, err, sizeof err);
                sum += mrs_check_ge3d_timing(tes_bad, 2, 60.0, 2.0, ny, kseg, 0.4, rev, 8, 0.0, 600.0, err, sizeof err);
                sum += mrs_check_ge3d_timing(tes_ok, 0, 60.0, 2.0, ny, kseg, 0.4, rev, 8, 0.0, 600.0, err, sizeof err);
            }
    std::printf("tables_check ok %.6e\n", sum);
    return 0;
}
